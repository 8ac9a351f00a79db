// config.h -- drop-in for the reference's src/config.h (:1-20).
//
// Same macro names and types.  Each macro can be set on the compiler command
// line (the reference fixes them in the header); the GPU library takes D and
// k at run time, so one build of libhdc_amd.so serves every configuration.
#pragma once
#include <stdint.h>

#include <vector>

#ifndef KMER_SIZE
#define KMER_SIZE 5
#endif
#ifndef HV_DIM
#define HV_DIM 10'000
#endif
#ifndef SEQUENCE_LENGTH
#define SEQUENCE_LENGTH 125
#endif
#ifndef SIM_STEP
#define SIM_STEP 10
#endif

#define HV_WORDS ((HV_DIM + 63) / 64)
#define HV_BLOCK_SIZE 1024
#define HV_BLOCKS ((HV_DIM + HV_BLOCK_SIZE - 1) / HV_BLOCK_SIZE)
#define KMER_BITS (2 * KMER_SIZE)
#define KMER_COUNT (1 << KMER_BITS)
#define KMER_MASK (KMER_COUNT - 1)
#define K (SEQUENCE_LENGTH - KMER_SIZE)

using HV = std::vector<uint64_t>;
using HV_16 = std::vector<int16_t>;
