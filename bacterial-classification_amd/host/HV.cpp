// HV.cpp -- host HV primitives with the reference's semantics
// (src/HV.cpp:9-252, src/HV.h:21-106).
#include "HV.h"

#include <cmath>
#include <random>

namespace {

constexpr size_t kExcess = HV_WORDS * 64 - HV_DIM;
constexpr uint64_t kTail = kExcess ? (~uint64_t(0)) >> kExcess : ~uint64_t(0);

HV draw(uint64_t seed) {
    std::mt19937_64 g(seed);
    HV h(HV_WORDS);
    for (auto& w : h) w = g();
    h.back() &= kTail;
    return h;
}

inline int bit(const HV& h, size_t d) { return (int)((h[d >> 6] >> (d & 63)) & 1); }

}  // namespace

HV empty_hv() {
    HV h(HV_WORDS, ~uint64_t(0));
    h.back() &= kTail;
    return h;
}

const HV& get_base_hv(char symbol) {
    static const HV a = draw(0), c = draw(1), t = draw(2), g = draw(3), n = empty_hv();
    switch (symbol) {
        case 'A': return a;
        case 'C': return c;
        case 'T': return t;
        case 'G': return g;
        default: return n;
    }
}

void mult(const HV& l, const HV& r, HV& out, size_t k_bits) {
    const size_t W = HV_WORDS;
    const size_t ws = (k_bits >> 6) % W;
    const unsigned s = (unsigned)(k_bits & 63);
    HV tmp(W);
    for (size_t i = 0; i < W; ++i) {
        const size_t a = (ws + i) % W, b = (a + 1) % W;
        const uint64_t rw = s ? ((r[a] >> s) | (r[b] << (64 - s))) : r[a];
        tmp[i] = l[i] ^ rw;
    }
    tmp[W - 1] &= kTail;
    out.swap(tmp);
}

HV_16 make_accumulator() { return HV_16(HV_DIM, 0); }

void superpose_batch(HV_16& acc, const HV* binders, size_t batch_size) {
    // per 64-dim word: count set bits over the batch, then one update
    for (size_t w = 0; w < HV_WORDS; ++w) {
        uint16_t cnt[64] = {0};
        for (size_t b = 0; b < batch_size; ++b) {
            uint64_t x = binders[b][w];
            while (x) {
                cnt[__builtin_ctzll(x)]++;
                x &= x - 1;
            }
        }
        const size_t lim = (w + 1) * 64 <= HV_DIM ? 64 : HV_DIM - w * 64;
        for (size_t i = 0; i < lim; ++i)
            acc[w * 64 + i] = (int16_t)(acc[w * 64 + i] - (int16_t)batch_size + (int16_t)(2 * cnt[i]));
    }
}

void replace_batch(HV_16& acc, const HV* const* olds, const HV* const* news, size_t batch_size) {
    for (size_t b = 0; b < batch_size; ++b)
        for (size_t d = 0; d < HV_DIM; ++d)
            acc[d] = (int16_t)(acc[d] + 2 * bit(*news[b], d) - 2 * bit(*olds[b], d));
}

float cosine(const HV_16& l, const HV_16& r) {
    int64_t dot = 0, nl = 0, nr = 0;
    const size_t n = l.size();
    for (size_t i = 0; i < n; ++i) {
        const int64_t a = l[i], b = r[i];
        dot += a * b;
        nl += a * a;
        nr += b * b;
    }
#ifdef HVK_PLAIN_DOT_COSINE
    (void)nl;
    (void)nr;
    return (float)dot;
#else
    if (nl == 0 || nr == 0) return -1.0f;
    return (float)(2.0 * ((double)dot / (std::sqrt((double)nl) * std::sqrt((double)nr)) - 0.5));
#endif
}
