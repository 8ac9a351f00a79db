// HV.h -- drop-in for the reference's src/HV.h (:1-108): binary hypervector
// primitives on the host.  Same names, signatures and results.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "config.h"

// All-ones HV of HV_DIM bits (src/HV.cpp:9-18).
HV empty_hv();
// Base HV of a nucleotide; anything but A/C/G/T gives empty_hv()
// (src/HV.cpp:20-51).
const HV& get_base_hv(char symbol);
// out = l XOR rotR(r, k_bits) on the 64*HV_WORDS-bit ring, tail masked
// (src/HV.cpp:98-147).  out may alias l.
void mult(const HV& l, const HV& r, HV& out, size_t k_bits);

HV_16 make_accumulator();
// acc[d] += 2 * #{b : bit d of binders[b]} - batch_size (src/HV.cpp:156-216).
void superpose_batch(HV_16& acc, const HV* binders, size_t batch_size);
// acc[d] += 2 * (bit d of *news[b]) - 2 * (bit d of *olds[b]) (src/HV.h:21-106).
void replace_batch(HV_16& acc, const HV* const* olds, const HV* const* news, size_t batch_size);
// NORM_COSINE (default, as install.sh selects): float(2c - 1), -1 on a zero
// norm; with HVK_PLAIN_DOT_COSINE defined: (float) dot (src/HV.cpp:219-252).
float cosine(const HV_16& l, const HV_16& r);
