// GPU.h -- drop-in for the reference's src/GPU.cuh (:1-11): the GPU entry
// points with the same signatures, backed by the gfx950 library.
#pragma once

#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "config.h"

// Windowed encoding (src/encode.cu:197-409).  For each sequence of length
// L >= SEQUENCE_LENGTH, the (L - SEQUENCE_LENGTH) / SIM_STEP + 1 windows
// starting at 0, SIM_STEP, ...; each window's HV_16 equals HVCache::compute
// of its substring (src/encoder.cpp:106-132).  Shorter sequences give no
// windows.  kmer_map is accepted for source compatibility and not read: the
// device binds every k-mer from the base HVs, as HVCache's own table does.
std::vector<std::vector<HV_16>> encode(const std::vector<int16_t>& kmer_map,
                                       const std::vector<std::string>& sequences);

// Top-k cosine search (src/cosine.cu:502-564): for each input, the k (clamped
// to db.size()) most similar db entries as (cosine, db pointer), cosine
// descending, ties by db position; cosine = dot / (|a| |b|) (the GPU scale,
// not the 2c - 1 of cosine() in HV.h), 0 for a zero vector.  Any k: past
// the candidate-list length the library selects from the exact key matrix.
std::unordered_map<HV_16*, std::vector<std::pair<float, HV_16*>>> cosine_search(
    const std::vector<HV_16*>& inputs, const std::vector<HV_16*>& db, int k = 1);
