// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" shim, written for this repository, over the reference's
// own CPU sources (src/HV.cpp, src/encoder.cpp), which oracle/Makefile
// compiles in place from the reference checkout into oracle/_ref/libref.so.
// It lets tests/ and tests/golden/make_golden.py pin the C oracle
// (hdc_oracle.c) against the real reference at the reference's compile-time
// configuration (src/config.h:5-8: HV_DIM 10000, KMER_SIZE 5, window 125,
// step 10), built with NORM_COSINE as install.sh:3 selects.
#include <cstring>
#include <string>
#include <vector>

#include "HV.h"
#include "encoder.h"

namespace {
// A cache used only through get() on fresh strings: get == compute.
HVCache& compute_cache() {
    static HVCache c(1u << 30);
    return c;
}
}  // namespace

extern "C" {

void ref_config(int* dim, int* kmer, int* window, int* step) {
    *dim = HV_DIM;
    *kmer = KMER_SIZE;
    *window = SEQUENCE_LENGTH;
    *step = SIM_STEP;
}

void ref_base_hv(char symbol, uint64_t* out) {
    const HV& h = get_base_hv(symbol);
    std::memcpy(out, h.data(), sizeof(uint64_t) * HV_WORDS);
}

void ref_empty_hv(uint64_t* out) {
    HV h = empty_hv();
    std::memcpy(out, h.data(), sizeof(uint64_t) * HV_WORDS);
}

void ref_mult(const uint64_t* l, const uint64_t* r, uint64_t* out, size_t k_bits) {
    HV a(l, l + HV_WORDS), b(r, r + HV_WORDS), o(HV_WORDS);
    mult(a, b, o, k_bits);
    std::memcpy(out, o.data(), sizeof(uint64_t) * HV_WORDS);
}

// 1 and the HV when the k-mer is in the reference's table, else 0.
int ref_kmer_hv(const char* kmer, uint64_t* out) {
    auto& m = compute_cache().kmers;
    auto it = m.find(std::string(kmer, KMER_SIZE));
    if (it == m.end()) return 0;
    std::memcpy(out, it->second.data(), sizeof(uint64_t) * HV_WORDS);
    return 1;
}

// The whole k-mer table of a fresh HVCache: its entry count, then the HV of
// every k-mer over ACGT in lexicographic order (4^KMER_SIZE rows; a k-mer the
// table lacks gets an all-zero row and is not counted).
int ref_kmers_all(uint64_t* out) {
    HVCache c(16);
    const char acgt[4] = {'A', 'C', 'G', 'T'};
    const size_t n = size_t(1) << (2 * KMER_SIZE);
    std::string s(KMER_SIZE, 'A');
    for (size_t i = 0; i < n; ++i) {
        for (size_t p = 0; p < KMER_SIZE; ++p) s[p] = acgt[(i >> (2 * (KMER_SIZE - 1 - p))) & 3];
        auto it = c.kmers.find(s);
        if (it == c.kmers.end())
            std::memset(out + i * HV_WORDS, 0, sizeof(uint64_t) * HV_WORDS);
        else
            std::memcpy(out + i * HV_WORDS, it->second.data(), sizeof(uint64_t) * HV_WORDS);
    }
    return (int)c.kmers.size();
}

void ref_compute(const char* seq, size_t len, int16_t* out) {
    const HV_16& h = compute_cache().get(std::string(seq, len));
    std::memcpy(out, h.data(), sizeof(int16_t) * HV_DIM);
}

void ref_superpose_batch(int16_t* acc, const uint64_t* hvs, size_t n) {
    HV_16 a(acc, acc + HV_DIM);
    std::vector<HV> b(n);
    for (size_t i = 0; i < n; ++i) b[i].assign(hvs + i * HV_WORDS, hvs + (i + 1) * HV_WORDS);
    superpose_batch(a, b.data(), n);
    std::memcpy(acc, a.data(), sizeof(int16_t) * HV_DIM);
}

void ref_replace_batch(int16_t* acc, const uint64_t* olds, const uint64_t* news, size_t n) {
    HV_16 a(acc, acc + HV_DIM);
    std::vector<HV> o(n), w(n);
    std::vector<const HV*> op(n), np(n);
    for (size_t i = 0; i < n; ++i) {
        o[i].assign(olds + i * HV_WORDS, olds + (i + 1) * HV_WORDS);
        w[i].assign(news + i * HV_WORDS, news + (i + 1) * HV_WORDS);
        op[i] = &o[i];
        np[i] = &w[i];
    }
    replace_batch(a, op.data(), np.data(), n);
    std::memcpy(acc, a.data(), sizeof(int16_t) * HV_DIM);
}

float ref_cosine(const int16_t* l, const int16_t* r) {
    HV_16 a(l, l + HV_DIM), b(r, r + HV_DIM);
    return cosine(a, b);
}

// Fresh cache: precompute(ref) then similarity(query, ref).
float ref_precompute_similarity(const char* q, size_t ql, const char* r, size_t rl, size_t* pos) {
    HVCache c(1u << 20);
    std::string rs(r, rl), qs(q, ql);
    c.precompute(rs);
    SIM s = c.similarity(qs, rs);
    if (pos) *pos = s.pos;
    return s.score;
}

// Fresh cache: precompute(ref), then read each window back through get().
int ref_precompute_windows(const char* r, size_t rl, int16_t* out, int max_windows) {
    HVCache c(1u << 20);
    std::string rs(r, rl);
    c.precompute(rs);
    int n = 0;
    if (rl < SEQUENCE_LENGTH) return 0;
    for (size_t x = 0; (x == 0 || x + SEQUENCE_LENGTH < rl) && n < max_windows; x += SIM_STEP) {
        const HV_16& h = c.get(rs.substr(x, SEQUENCE_LENGTH));
        std::memcpy(out + (size_t)n * HV_DIM, h.data(), sizeof(int16_t) * HV_DIM);
        ++n;
    }
    return n;
}

}  // extern "C"
