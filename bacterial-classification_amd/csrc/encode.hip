// encode.hip -- k-mer hypervector encoding on gfx950.
//
// Semantics: HVCache::compute (src/encoder.cpp:106-132) with the k-mer HVs of
// the HVCache constructor (src/encoder.cpp:47-74):
//   HV(kmer)[d] = NOT( XOR_i rotR(base[c_i], i)[d] ),  c = lexmin(kmer, revcomp)
//   count[d]    = sum over all-ACGT k-mers of (2*HV[d] - 1)
//               = n_valid - 2 * #{k-mers whose XOR word has bit d set}.
//
// Design (one workgroup per sequence, persistent over sequences):
//   * the table rot[b][i][w] (4 x k x W32 words, 61 KB at D=10240, k=12) is
//     staged in LDS once per workgroup -- or, when it fits, the table of base
//     pairs (encode_pair_kernel, below), which halves the lookups;
//   * phase 1: lanes compute the canonical code of each k-mer position of a
//     1024-position chunk from the 2-bit packed sequence (rolling-free: one
//     64-bit window per position), written to LDS as a "digit i = c_i" code;
//   * phase 2: lane w owns 32 dimensions (word w).  The k-mer code is wave
//     uniform (readfirstlane -> SGPR), so the k table loads of a k-mer hit k
//     conflict-free LDS rows; 8 XOR words are reduced by a carry-save adder
//     tree and rippled into P bit-sliced counter planes;
//   * phase 3: planes are transposed to per-dimension counts, written as
//     int16 counts and/or int8 limbs, with the exact sum of squares and the
//     max |count| reduced over the workgroup.
#include "hvk_internal.h"

namespace hvk {
namespace {

constexpr uint32_t kInvalidCode = 0xFFFFFFFFu;
constexpr int kChunk = 1024;

__device__ __forceinline__ uint32_t rev_digits(uint32_t v, int k) {
    uint32_t y = __builtin_bitreverse32(v);
    y = ((y >> 1) & 0x55555555u) | ((y & 0x55555555u) << 1);
    return y >> (32 - 2 * k);
}

__device__ __forceinline__ void full_add(uint32_t a, uint32_t b, uint32_t c, uint32_t& s,
                                         uint32_t& cy) {
    const uint32_t t = a ^ b;
    s = t ^ c;
    cy = (a & b) | (t & c);
}

// planes += popcount-per-lane of x[0..7] (bit-sliced, lane = bit position)
template <int P>
__device__ __forceinline__ void add8(uint32_t (&planes)[P], const uint32_t (&x)[8]) {
    uint32_t a1, a2, b1, b2, c1, c2, d1, d2, e2, e4, f2, f4, g4, g8;
    full_add(x[0], x[1], x[2], a1, a2);
    full_add(x[3], x[4], x[5], b1, b2);
    full_add(x[6], x[7], a1, c1, c2);
    d1 = b1 ^ c1;  // half adder
    d2 = b1 & c1;
    full_add(a2, b2, c2, e2, e4);
    f2 = e2 ^ d2;
    f4 = e2 & d2;
    g4 = e4 ^ f4;
    g8 = e4 & f4;
    // ripple (d1 + 2 f2 + 4 g4 + 8 g8) into the planes
    uint32_t cy = planes[0] & d1;
    planes[0] ^= d1;
    uint32_t s;
    full_add(planes[1], f2, cy, s, cy);
    planes[1] = s;
    full_add(planes[2], g4, cy, s, cy);
    planes[2] = s;
    full_add(planes[3], g8, cy, s, cy);
    planes[3] = s;
#pragma unroll
    for (int p = 4; p < P; ++p) {
        const uint32_t t = planes[p] & cy;
        planes[p] ^= cy;
        cy = t;
    }
}

// a ^ b ^ c and the majority of a, b, c: one v_bitop3 each.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8);
}

// Carry-save adder: a + b + c = l + 2 h, bitwise.
__device__ __forceinline__ void csa(uint32_t& h, uint32_t& l, uint32_t a, uint32_t b, uint32_t c) {
    h = maj3(a, b, c);
    l = xor3(a, b, c);
}

typedef __attribute__((address_space(3))) const uint32_t lds_u32;

// LDS word at byte address (digit << (kPairShift + 2)) + base.  The digit is
// wave uniform: one s_bfe extracts it and one v_lshl_add forms the address
// (the compiler would shift and add on the scalar unit first).
__device__ __forceinline__ uint32_t pair_row(uint32_t digit, uint32_t base);

// planes += popcount-per-lane of x[0..15], Harley-Seal form: planes[0..3]
// are the ones, twos, fours and eights of a carry-save accumulator and
// planes[4..] a binary counter of sixteens.  The value held is
// sum_p planes[p] 2^p in both readings, so the planes stay an ordinary
// bit-sliced counter for the part merge and phase 3.  Per 16 words: 15
// carry-save adders and one ripple of a single word into planes[4..].
template <int P>
__device__ __forceinline__ void add16(uint32_t (&planes)[P], const uint32_t (&x)[16]) {
    static_assert(P >= 5, "add16 needs at least 5 planes");
    uint32_t tA, tB, fA, fB, eA, eB, cy;
    csa(tA, planes[0], planes[0], x[0], x[1]);
    csa(tB, planes[0], planes[0], x[2], x[3]);
    csa(fA, planes[1], planes[1], tA, tB);
    csa(tA, planes[0], planes[0], x[4], x[5]);
    csa(tB, planes[0], planes[0], x[6], x[7]);
    csa(fB, planes[1], planes[1], tA, tB);
    csa(eA, planes[2], planes[2], fA, fB);
    csa(tA, planes[0], planes[0], x[8], x[9]);
    csa(tB, planes[0], planes[0], x[10], x[11]);
    csa(fA, planes[1], planes[1], tA, tB);
    csa(tA, planes[0], planes[0], x[12], x[13]);
    csa(tB, planes[0], planes[0], x[14], x[15]);
    csa(fB, planes[1], planes[1], tA, tB);
    csa(eB, planes[2], planes[2], fA, fB);
    csa(cy, planes[3], planes[3], eA, eB);
#pragma unroll
    for (int p = 4; p < P; ++p) {
        const uint32_t t = planes[p] & cy;
        planes[p] ^= cy;
        cy = t;
    }
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int lo = __shfl_xor((int)(uint32_t)v, o, 64);
        int hi = __shfl_xor((int)(v >> 32), o, 64);
        v += (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
    return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Codes of chunk positions j .. j+7 (j a multiple of 8): two 16-byte reads at
// a wave-uniform address.  Phase 1 fills the positions past the chunk up to
// the next multiple of 8 with kInvalidCode.
__device__ __forceinline__ void load_codes8(const uint32_t* codes, int j, uint32_t (&c)[8]) {
    const uint4 a = *reinterpret_cast<const uint4*>(codes + j);
    const uint4 b = *reinterpret_cast<const uint4*>(codes + j + 4);
    c[0] = a.x; c[1] = a.y; c[2] = a.z; c[3] = a.w;
    c[4] = b.x; c[5] = b.y; c[6] = b.z; c[7] = b.w;
}

// Phase 1: canonical code of chunk position j (kInvalidCode when a base of
// the k-mer is not A/C/G/T or j >= cn).
// The 64-bit windows of the 2-bit packed bases and the invalid-base bits that
// hold position p.
__device__ __forceinline__ uint64_t base_window(const uint32_t* __restrict__ packed, int64_t p) {
    return (uint64_t)packed[p >> 4] | ((uint64_t)packed[(p >> 4) + 1] << 32);
}
__device__ __forceinline__ uint64_t invalid_window(const uint32_t* __restrict__ invalid,
                                                   int64_t p) {
    return (uint64_t)invalid[p >> 5] | ((uint64_t)invalid[(p >> 5) + 1] << 32);
}

// Canonical code of the k-mer at position p from its windows.
__device__ __forceinline__ uint32_t code_of_windows(uint64_t w2, uint64_t iv, int64_t p, int k,
                                                    uint32_t mask2k, uint32_t maskk) {
    const uint32_t v = (uint32_t)(w2 >> (2 * (p & 15))) & mask2k;
    const uint32_t bad = (uint32_t)(iv >> (p & 31)) & maskk;
    if (bad) return kInvalidCode;
    const uint32_t f = rev_digits(v, k);
    const uint32_t r = mask2k - v;
    return (r < f) ? (mask2k - f) : v;
}

__device__ __forceinline__ uint32_t kmer_code(const uint32_t* __restrict__ packed,
                                              const uint32_t* __restrict__ invalid, int64_t p,
                                              int j, int cn, int k, uint32_t mask2k,
                                              uint32_t maskk) {
    if (j >= cn) return kInvalidCode;
    const uint64_t w2 = base_window(packed, p);
    const uint32_t v = (uint32_t)(w2 >> (2 * (p & 15))) & mask2k;
    const uint64_t iv = invalid_window(invalid, p);
    const uint32_t bad = (uint32_t)(iv >> (p & 31)) & maskk;
    if (bad) return kInvalidCode;
    // v: digit i = base i; forward code f has base 0 most significant;
    // revcomp code = mask - v.  Stored: rev(canonical), digit i = c_i.
    const uint32_t f = rev_digits(v, k);
    const uint32_t r = mask2k - v;
    return (r < f) ? (mask2k - f) : v;
}

typedef short i16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Even bits of x to bits 0..15, odd bits to 16..31 (inverse outer perfect
// shuffle, four delta swaps).
__device__ __forceinline__ uint32_t unzip_bits(uint32_t x) {
    uint32_t t;
    t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
    t = (x ^ (x >> 2)) & 0x0C0C0C0Cu; x ^= t ^ (t << 2);
    t = (x ^ (x >> 4)) & 0x00F000F0u; x ^= t ^ (t << 4);
    t = (x ^ (x >> 8)) & 0x0000FF00u; x ^= t ^ (t << 8);
    return x;
}

// One stage of the transposes below: swap the SH x SH sub-blocks.  A stage
// per template instance, so every register index is a constant (a loop over
// the stages stays rolled and indexes registers at run time).
template <int SH>
__device__ __forceinline__ void transpose_stage(uint32_t (&w)[16], uint32_t m) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (j & SH) continue;
        const uint32_t t = ((w[j] >> SH) ^ w[j + SH]) & m;
        w[j + SH] ^= t;
        w[j] ^= t << SH;
    }
}

// Two 16 x 16 bit-matrix transposes at once (one per 16-bit half): on
// return bit p of half h of w[i] is bit i of half h of the input w[p].
__device__ __forceinline__ void transpose16x2(uint32_t (&w)[16]) {
    transpose_stage<8>(w, 0x00FF00FFu);
    transpose_stage<4>(w, 0x0F0F0F0Fu);
    transpose_stage<2>(w, 0x33333333u);
    transpose_stage<1>(w, 0x55555555u);
}

__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ i16x2 as_i16x2(uint32_t v) { return __builtin_bit_cast(i16x2, v); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t as_u32(i16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// Phase 3 for one word: the thread owns word w of row s and holds its
// bit-sliced counts; it writes the word's counts / limbs / clamped values and
// adds the word's sum of squares, max |count| and clamping residual to sq,
// mx and res.
template <int P>
__device__ __forceinline__ void finish_word(const uint32_t (&planes)[P], int w, int nvalid, int D,
                                            int64_t s, const EncodeOut& out, int64_t& sq, int& mx,
                                            int& res) {
    const int64_t ld = out.ld;
    {
        uint32_t packed16[16];
        uint32_t plo[8], phi[8], pcl[8];
        if constexpr (P <= 16) {
            // Word i of the transposed planes holds the popcounts of dims 2i
            // (low half) and 2i+1 (high half); the rest is packed 16-bit
            // arithmetic, modulo 2^16 as the int16 counts are.
            uint32_t t[16];
#pragma unroll
            for (int p = 0; p < 16; ++p) t[p] = p < P ? unzip_bits(planes[p]) : 0u;
            transpose16x2(t);
            const int lim = D - 32 * w;  // dims >= lim of this word are padding
            const bool tail = (D & 31) != 0 && lim < 32;  // uniform test first
            const u16x2 nv2 = {(unsigned short)nvalid, (unsigned short)nvalid};
            const u16x2 c127 = {127, 127}, c128 = {128, 128}, ones = {1, 1};
            const i16x2 p127 = {127, 127}, m127 = {-127, -127};
            u16x2 mx2 = {0, 0};
            uint32_t res_u = 0;
            int sq32 = 0;
            int64_t sq64 = 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const u16x2 pc = as_u16x2(t[i]);
                u16x2 c = nv2 - (pc + pc);
                if (tail) {
                    if (2 * i >= lim) c.x = 0;
                    if (2 * i + 1 >= lim) c.y = 0;
                }
                const i16x2 cs = as_i16x2(as_u32(c));
                if constexpr (P <= 11) sq32 = __builtin_amdgcn_sdot2(cs, cs, sq32, false);
                else sq64 += (int64_t)(uint32_t)__builtin_amdgcn_sdot2(cs, cs, 0, false);
                const u16x2 zero = {0, 0};
                const u16x2 a = as_u16x2(as_u32(__builtin_elementwise_max(cs, as_i16x2(as_u32(zero - c)))));
                mx2 = __builtin_elementwise_max(mx2, a);
                res_u = __builtin_amdgcn_udot2(__builtin_elementwise_sub_sat(a, c127), ones, res_u,
                                               false);
                packed16[i] = as_u32(c);
                t[i] = as_u32(c + c128);  // high bytes: the hi limb
                const i16x2 cl = __builtin_elementwise_max(__builtin_elementwise_min(cs, p127), m127);
                if (i & 1) {
                    plo[i >> 1] = __builtin_amdgcn_perm(packed16[i], packed16[i - 1], 0x06040200u);
                    phi[i >> 1] = __builtin_amdgcn_perm(t[i], t[i - 1], 0x07050301u);
                    pcl[i >> 1] = __builtin_amdgcn_perm(as_u32(cl), pcl[i >> 1], 0x06040200u);
                } else {
                    pcl[i >> 1] = as_u32(cl);
                }
            }
            if constexpr (P <= 11) sq += sq32;
            else sq += sq64;
            mx = max(mx, max((int)mx2.x, (int)mx2.y));
            res += (int)res_u;
        } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) { plo[q] = 0; phi[q] = 0; pcl[q] = 0; }
#pragma unroll
        for (int b = 0; b < 32; ++b) {
            uint32_t pc = 0;
#pragma unroll
            for (int p = 0; p < P; ++p) pc |= ((planes[p] >> b) & 1u) << p;
            int cnt = (32 * w + b < D) ? (int)(int16_t)(nvalid - 2 * (int)pc) : 0;
            sq += (int64_t)cnt * cnt;
            const int a = cnt < 0 ? -cnt : cnt;
            mx = max(mx, a);
            res += a > 127 ? a - 127 : 0;
            const uint32_t u16 = (uint32_t)(uint16_t)(int16_t)cnt;
            if (b & 1) packed16[b >> 1] |= u16 << 16;
            else packed16[b >> 1] = u16;
            const int l8 = (int)(int8_t)(cnt & 0xFF);
            const int h8 = (cnt - l8) >> 8;
            const int c8 = cnt > 127 ? 127 : (cnt < -127 ? -127 : cnt);
            plo[b >> 2] |= ((uint32_t)(uint8_t)l8) << (8 * (b & 3));
            phi[b >> 2] |= ((uint32_t)(uint8_t)h8) << (8 * (b & 3));
            pcl[b >> 2] |= ((uint32_t)(uint8_t)c8) << (8 * (b & 3));
        }
        }
        const int64_t row = s * ld + 32 * (int64_t)w;
        if (out.counts) {
            uint4* dst = reinterpret_cast<uint4*>(out.counts + row);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                dst[q] = make_uint4(packed16[4 * q], packed16[4 * q + 1], packed16[4 * q + 2],
                                    packed16[4 * q + 3]);
        }
        if (out.lo) {
            uint4* dst = reinterpret_cast<uint4*>(out.lo + row);
            dst[0] = make_uint4(plo[0], plo[1], plo[2], plo[3]);
            dst[1] = make_uint4(plo[4], plo[5], plo[6], plo[7]);
        }
        if (out.hi) {
            uint4* dst = reinterpret_cast<uint4*>(out.hi + row);
            dst[0] = make_uint4(phi[0], phi[1], phi[2], phi[3]);
            dst[1] = make_uint4(phi[4], phi[5], phi[6], phi[7]);
        }
        if (out.clamp8) {
            uint4* dst = reinterpret_cast<uint4*>(out.clamp8 + row);
            dst[0] = make_uint4(pcl[0], pcl[1], pcl[2], pcl[3]);
            dst[1] = make_uint4(pcl[4], pcl[5], pcl[6], pcl[7]);
        }
    }
}

// Phase 3, called by every thread of the workgroup: the threads with
// `active` set own word w of the row and hold its bit-sliced counts; they
// write counts / limbs / clamped rows, and the workgroup reduces the exact
// sum of squares, max |count| and clamping residual of row s.
template <int P>
__device__ __forceinline__ void finish_row(const uint32_t (&planes)[P], bool active, int w,
                                           int nvalid, int D, int W32, int64_t s,
                                           const EncodeOut& out, int64_t* red_sq, int* red_mx,
                                           int* red_res) {
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nwaves = nthr >> 6;
    int64_t sq = 0;
    int mx = 0, res = 0;
    const int64_t ld = out.ld;
    if (active) finish_word<P>(planes, w, nvalid, D, s, out, sq, mx, res);
    // zero padding columns beyond the last full word
    for (int64_t d = 32 * (int64_t)W32 + tid; d < ld; d += nthr) {
        const int64_t o = s * ld + d;
        if (out.counts) out.counts[o] = 0;
        if (out.lo) out.lo[o] = 0;
        if (out.hi) out.hi[o] = 0;
        if (out.clamp8) out.clamp8[o] = 0;
    }
    sq = wave_sum64(sq);
    mx = wave_max(mx);
    res = wave_sum(res);
    if (lane == 0) { red_sq[wave] = sq; red_mx[wave] = mx; red_res[wave] = res; }
    __syncthreads();
    if (tid == 0) {
        int64_t t = 0;
        int m = 0, r = 0;
        for (int q = 0; q < nwaves; ++q) {
            t += red_sq[q];
            m = max(m, red_mx[q]);
            r += red_res[q];
        }
        if (out.sqnorm) out.sqnorm[s] = t;
        if (out.maxabs) out.maxabs[s] = m;
        if (out.nkmers) out.nkmers[s] = nvalid;
        if (out.resid1) out.resid1[s] = r;
    }
    __syncthreads();
}

template <int K, int P>
__global__ void __launch_bounds__(1024)
encode_kernel(const uint32_t* __restrict__ gtable, int k_rt, int D, int W32,
              const uint32_t* __restrict__ packed, const uint32_t* __restrict__ invalid,
              const int64_t* __restrict__ offsets, const int64_t* __restrict__ starts,
              const int32_t* __restrict__ lens, int window_len, int n_seqs, EncodeOut out) {
    const int k = K > 0 ? K : k_rt;
    extern __shared__ uint32_t smem[];
    uint32_t* tab = smem;
    uint32_t* codes = smem + 4 * k * W32;
    __shared__ int64_t red_sq[16];
    __shared__ int red_mx[16];
    __shared__ int red_res[16];

    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    for (int i = tid; i < 4 * k * W32; i += nthr) tab[i] = gtable[i];

    const uint32_t mask2k = (k == 16) ? 0xFFFFFFFFu : ((1u << (2 * k)) - 1u);
    const uint32_t maskk = (1u << k) - 1u;
    const int w = tid;
    const int kW = k * W32;

    for (int s = blockIdx.x; s < n_seqs; s += gridDim.x) {
        // sequence s: [offsets[s], offsets[s+1]), or lens[s] (else window_len)
        // bases from starts[s]
        const int64_t off = starts ? starts[s] : offsets[s];
        const int64_t len = starts ? (lens ? (int64_t)lens[s] : (int64_t)window_len)
                                   : offsets[s + 1] - off;
        const int64_t npos = len - k + 1;
        uint32_t planes[P];
#pragma unroll
        for (int p = 0; p < P; ++p) planes[p] = 0;
        int nvalid = 0;

        for (int64_t c0 = 0; c0 < npos; c0 += kChunk) {
            const int cn = (int)((npos - c0) < kChunk ? (npos - c0) : kChunk);
            __syncthreads();
            const int cn8 = (cn + 7) & ~7;
            for (int j = tid; j < cn8; j += nthr)
                codes[j] = kmer_code(packed, invalid, off + c0 + j, j, cn, k, mask2k, maskk);
            __syncthreads();
            if (w < W32) {
                for (int j = 0; j < cn; j += 8) {
                    // branch-free: an invalid code (all digits 3) reads valid
                    // table rows and its word is dropped afterwards
                    uint32_t c[8], x[8];
                    load_codes8(codes, j, c);
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const uint32_t code = __builtin_amdgcn_readfirstlane(c[u]);
                        uint32_t y = 0;
                        if (K > 0) {
#pragma unroll
                            for (int i = 0; i < (K > 0 ? K : 1); ++i)
                                y ^= tab[((code >> (2 * i)) & 3u) * kW + i * W32 + w];
                        } else {
                            for (int i = 0; i < k; ++i)
                                y ^= tab[((code >> (2 * i)) & 3u) * kW + i * W32 + w];
                        }
                        const bool valid = code != kInvalidCode;
                        nvalid += valid ? 1 : 0;
                        x[u] = valid ? y : 0u;
                    }
                    add8<P>(planes, x);
                }
            }
        }

        finish_row<P>(planes, w < W32, w, nvalid, D, W32, s, out, red_sq, red_mx, red_res);
    }
}

// Pair-table encoder: the table of base PAIRS T2[c_2j + 4 c_2j+1][j][w] =
// rot(base[c_2j], 2j) ^ rot(base[c_2j+1], 2j+1) halves the LDS lookups and
// address adds per k-mer (6 instead of 12 at k = 12).  The 16 entries sit
// kPairStride words apart (128 KB), so an entry's offset is a shift of its
// digit folded into the lane's address add (no scalar multiply).  One
// workgroup fills a CU; to keep more waves per CU, NP parts of the workgroup
// (as many as 1024 threads hold, at most 4) share each sequence: they take
// alternate blocks of 16 k-mers into their own Harley-Seal counters, which
// are added (bit-sliced) into part 0's before phase 3.
constexpr int kPairShift = 11;
constexpr int kPairStride = 1 << kPairShift;  // words per pair-table entry
constexpr int kPairBlock = 16;                // k-mers per block of a part

__device__ __forceinline__ uint32_t pair_row(uint32_t digit, uint32_t base) {
    uint32_t a;
    asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(a) : "s"(digit), "i"(kPairShift + 2), "v"(base));
    return *(lds_u32*)(size_t)a;
}

template <int K, int P>
__global__ void __launch_bounds__(1024)
encode_pair_kernel(const uint32_t* __restrict__ gtable, int k_rt, int D, int W32, int TW, int NP,
                   const uint32_t* __restrict__ packed, const uint32_t* __restrict__ invalid,
                   const int64_t* __restrict__ offsets, const int64_t* __restrict__ starts,
                   const int32_t* __restrict__ lens, int window_len, int n_seqs,
                   EncodeOut out) {
    const int k = K > 0 ? K : k_rt;
    const int KH = k / 2;
    extern __shared__ uint32_t smem[];
    uint32_t* tab2 = smem;                                  // [16][kPairStride]: [KH][W32]
    uint32_t* tab1 = tab2 + 16 * kPairStride;               // [4][W32], odd k: last base
    uint32_t* codes = tab1 + ((k & 1) ? 4 * W32 : 0);      // [kChunk]
    uint32_t* xplanes = codes + kChunk;                     // [P][W32]
    __shared__ int64_t red_sq[16];
    __shared__ int red_mx[16];
    __shared__ int red_res[16];
    __shared__ int s_nvalid;

    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int part = tid / TW;
    const int w = tid - part * TW;
    const int kW = k * W32;
    const int KHW = KH * W32;
    const uint32_t lds_w = (uint32_t)(size_t)(lds_u32*)tab2 + 4u * w;  // LDS byte address
    for (int i = tid; i < 16 * KHW; i += nthr) {
        const int v = i / KHW, rem = i - v * KHW, j = rem / W32, ww = rem - j * W32;
        tab2[v * kPairStride + rem] = gtable[(v & 3) * kW + (2 * j) * W32 + ww] ^
                                      gtable[(v >> 2) * kW + (2 * j + 1) * W32 + ww];
    }
    if (k & 1)
        for (int i = tid; i < 4 * W32; i += nthr) {
            const int b = i / W32, ww = i - b * W32;
            tab1[i] = gtable[b * kW + (k - 1) * W32 + ww];
        }

    const uint32_t mask2k = (k == 16) ? 0xFFFFFFFFu : ((1u << (2 * k)) - 1u);
    const uint32_t maskk = (1u << k) - 1u;

    // The next sequence's bounds are loaded at the start of a sequence, and
    // the windows of its first chunk's position tid after this sequence's
    // chunks, so that both loads overlap this sequence's work.
    auto bounds = [&](int q, int64_t& o, int64_t& l) {
        o = starts ? starts[q] : offsets[q];
        l = starts ? (lens ? (int64_t)lens[q] : (int64_t)window_len) : offsets[q + 1] - o;
    };
    int64_t off_n = 0, len_n = 0;
    uint64_t pw_n = 0, iv_n = 0;  // windows of the next sequence's position tid
    bool pre_n = false;           // ... when that position is a k-mer of it
    if (blockIdx.x < n_seqs) bounds(blockIdx.x, off_n, len_n);
    for (int s = blockIdx.x; s < n_seqs; s += gridDim.x) {
        const int64_t off = off_n, len = len_n;
        const uint64_t pw = pw_n, iv = iv_n;
        const bool pre = pre_n;
        const int s2 = s + gridDim.x;
        if (s2 < n_seqs) bounds(s2, off_n, len_n);
        const int64_t npos = len - k + 1;
        uint32_t planes[P];
#pragma unroll
        for (int p = 0; p < P; ++p) planes[p] = 0;
        int nvalid = 0;

        for (int64_t c0 = 0; c0 < npos; c0 += kChunk) {
            const int cn = (int)((npos - c0) < kChunk ? (npos - c0) : kChunk);
            __syncthreads();
            const int cnb = (cn + kPairBlock - 1) & ~(kPairBlock - 1);
            for (int j = tid; j < cnb; j += nthr)
                codes[j] = (c0 == 0 && j == tid && pre)
                               ? code_of_windows(pw, iv, off + j, k, mask2k, maskk)
                               : kmer_code(packed, invalid, off + c0 + j, j, cn, k, mask2k, maskk);
            __syncthreads();
            if (w < W32) {
                // alternate 16-blocks per part
                for (int j = kPairBlock * part; j < cn; j += kPairBlock * NP) {
                    // branch-free: an invalid code (all digits 3) reads valid
                    // table rows and its word is dropped afterwards
                    uint32_t c[16], x[16];
                    load_codes8(codes, j, *reinterpret_cast<uint32_t(*)[8]>(&c[0]));
                    load_codes8(codes, j + 8, *reinterpret_cast<uint32_t(*)[8]>(&c[8]));
                    if constexpr (K == 12 || K == 10) {
                        // software pipeline: the rows of k-mer u + 1 are
                        // requested before the rows of k-mer u are combined
                        // (the scheduling fences keep the extracted digits
                        // from piling up in SGPRs)
                        constexpr int KH_ = K / 2;
                        uint32_t r[2][KH_];
                        auto issue = [&](int u, uint32_t (&rr)[KH_]) {
                            const uint32_t code = __builtin_amdgcn_readfirstlane(c[u]);
#pragma unroll
                            for (int jj = 0; jj < KH_; ++jj)
                                rr[jj] = pair_row((code >> (4 * jj)) & 15u, lds_w + 4u * jj * W32);
                        };
                        issue(0, r[0]);
#pragma unroll
                        for (int u = 0; u < kPairBlock; ++u) {
                            if (u + 1 < kPairBlock) issue(u + 1, r[(u + 1) & 1]);
                            __builtin_amdgcn_sched_barrier(0);
                            const uint32_t(&q)[KH_] = r[u & 1];
                            uint32_t y;
                            if constexpr (K == 12) y = xor3(xor3(q[0], q[1], q[2]), q[3], q[4]) ^ q[5];
                            else y = xor3(xor3(q[0], q[1], q[2]), q[3], q[4]);
                            const bool valid = __builtin_amdgcn_readfirstlane(c[u]) != kInvalidCode;
                            nvalid += valid ? 1 : 0;
                            x[u] = valid ? y : 0u;
                        }
                    } else {
#pragma unroll
                    for (int u = 0; u < kPairBlock; ++u) {
                        const uint32_t code = __builtin_amdgcn_readfirstlane(c[u]);
                        auto row = [&](int jj) {
                            return pair_row((code >> (4 * jj)) & 15u, lds_w + 4u * jj * W32);
                        };
                        uint32_t y = 0;
                        if constexpr (K > 0) {
#pragma unroll
                            for (int jj = 0; jj < (K > 0 ? K / 2 : 1); ++jj) y ^= row(jj);
                            if (K & 1) y ^= tab1[((code >> (2 * (k - 1))) & 3u) * W32 + w];
                        } else {
                            for (int jj = 0; jj < KH; ++jj) y ^= row(jj);
                            if (k & 1) y ^= tab1[((code >> (2 * (k - 1))) & 3u) * W32 + w];
                        }
                        const bool valid = code != kInvalidCode;
                        nvalid += valid ? 1 : 0;
                        x[u] = valid ? y : 0u;
                    }
                    }
                    add16<P>(planes, x);
                }
            }
        }
        pre_n = false;
        if (s2 < n_seqs && tid < kChunk && tid < len_n - k + 1) {
            pw_n = base_window(packed, off_n + tid);
            iv_n = invalid_window(invalid, off_n + tid);
            pre_n = true;
        }
        // add the other parts' counters into part 0's (bit-sliced ripple add)
        for (int q = 1; q < NP; ++q) {
            __syncthreads();
            if (part == q && w < W32) {
#pragma unroll
                for (int p = 0; p < P; ++p) xplanes[p * W32 + w] = planes[p];
                if (w == 0) s_nvalid = nvalid;
            }
            __syncthreads();
            if (part == 0 && w < W32) {
                uint32_t cy = 0;
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const uint32_t a = planes[p], b = xplanes[p * W32 + w];
                    const uint32_t t = a ^ b;
                    planes[p] = t ^ cy;
                    cy = (a & b) | (t & cy);
                }
                nvalid += s_nvalid;
            }
        }
        finish_row<P>(planes, part == 0 && w < W32, w, nvalid, D, W32, s, out, red_sq, red_mx,
                      red_res);
    }
}

// One wave per sequence (encode_wave_kernel), for even k in {10, 12} and
// 257..320 or up to 256 words of 32 dimensions.
//
// Each wave64 LDS read costs about the same number of LDS cycles whatever
// its width (measured on gfx950: 3.9 clocks for 4 bytes per lane, 6.2 for
// 16), so the 4-byte pair-table reads of the kernel above keep the LDS busy
// while the VALU idles.  Here lane l owns the five words 4l .. 4l+3 (one
// 16-byte read per lookup) and 256 + l (one 4-byte read), and a wave holds a
// whole sequence: 12 LDS reads per k-mer instead of 30, and no workgroup
// barrier after the pair table is staged.  The k-mer words go into the
// Harley-Seal counters as they arrive (ones, twos, fours, eights in
// planes[0..3], sixteens rippled into planes[4..]), so only the carries of a
// 16-k-mer block are live.  Phase 1 writes the pair digits in 5-bit fields;
// a k-mer with a base other than A/C/G/T (and the padding past a chunk)
// points every field at entry 16, an all-zero row, so it adds nothing and
// needs no select.  Phase 1 also counts the valid k-mers.
constexpr int kWaveChunk = 512;  // k-mer positions per chunk (codes per wave)
constexpr int kWaveWaves = 8;    // waves (sequences in flight) per workgroup
constexpr int kZeroEntry = 16;   // pair-table entry of zeros

// The 2-bit digits of a canonical code (digit i = c_i) as KH 5-bit pair
// fields (field j = c_2j + 4 c_2j+1); kInvalidCode maps to kZeroEntry in
// every field.
template <int KH>
__device__ __forceinline__ uint32_t pair_fields(uint32_t code) {
    uint32_t f = 0;
#pragma unroll
    for (int j = 0; j < KH; ++j) f |= ((code >> (4 * j)) & 15u) << (5 * j);
    constexpr uint32_t zero = [] {
        uint32_t z = 0;
        for (int j = 0; j < KH; ++j) z |= (uint32_t)kZeroEntry << (5 * j);
        return z;
    }();
    return code == kInvalidCode ? zero : f;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ u32x4 pair_row4(uint32_t digit, uint32_t base) {
    uint32_t a;
    asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(a) : "s"(digit), "i"(kPairShift + 2), "v"(base));
    return *(const __attribute__((address_space(3))) u32x4*)(size_t)a;
}

template <int K, int P, int NB32>
__global__ void __launch_bounds__(64 * kWaveWaves)
encode_wave_kernel(const uint32_t* __restrict__ gtable, int D, int W32,
                   const uint32_t* __restrict__ packed, const uint32_t* __restrict__ invalid,
                   const int64_t* __restrict__ offsets, const int64_t* __restrict__ starts,
                   const int32_t* __restrict__ lens, int window_len, int n_seqs,
                   EncodeOut out) {
    static_assert(K == 10 || K == 12, "even k with a fixed pair count");
    constexpr int KH = K / 2;
    constexpr int NW = 4 + NB32;             // words per lane
    constexpr int W32P = 256 + 64 * NB32;    // table row length (words)
    static_assert(KH * W32P <= kPairStride, "pair-table entry overflows its stride");
    extern __shared__ uint32_t smem[];
    uint32_t* tab2 = smem;  // [17][kPairStride]: [KH][W32P], words >= W32 and entry 16 zero
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t* codes = smem + (kZeroEntry + 1) * kPairStride + wave * kWaveChunk;
    const int kW = K * W32;
    for (int i = tid; i < (kZeroEntry + 1) * KH * W32P; i += blockDim.x) {
        const int v = i / (KH * W32P), rem = i - v * (KH * W32P), j = rem / W32P,
                  ww = rem - j * W32P;
        tab2[v * kPairStride + rem] =
            (ww < W32 && v < kZeroEntry) ? gtable[(v & 3) * kW + (2 * j) * W32 + ww] ^
                                               gtable[(v >> 2) * kW + (2 * j + 1) * W32 + ww]
                                         : 0u;
    }
    __syncthreads();

    constexpr uint32_t mask2k = (1u << (2 * K)) - 1u, maskk = (1u << K) - 1u;
    const uint32_t tab_lds = (uint32_t)(size_t)(lds_u32*)tab2;
    const uint32_t a128 = tab_lds + 16u * lane;          // words 4 lane .. 4 lane + 3
    const uint32_t a32 = tab_lds + 4u * (256 + lane);    // word 256 + lane
    const int64_t ld = out.ld;
    const int stride = gridDim.x * kWaveWaves;
    for (int s = blockIdx.x * kWaveWaves + wave; s < n_seqs; s += stride) {
        const int64_t off = starts ? starts[s] : offsets[s];
        const int64_t len = starts ? (lens ? (int64_t)lens[s] : (int64_t)window_len)
                                   : offsets[s + 1] - off;
        const int64_t npos = len - K + 1;
        uint32_t planes[P][NW];
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int g = 0; g < NW; ++g) planes[p][g] = 0;
        int nvalid = 0;

        for (int64_t c0 = 0; c0 < npos; c0 += kWaveChunk) {
            const int cn = (int)((npos - c0) < kWaveChunk ? (npos - c0) : kWaveChunk);
            const int cnb = (cn + 15) & ~15;
            // the previous chunk's code reads are done before they are overwritten
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int j0 = 0; j0 < cnb; j0 += 64) {  // same trip count on every lane
                const int j = j0 + lane;
                const uint32_t code = kmer_code(packed, invalid, off + c0 + j, j, cn, K, mask2k,
                                                maskk);
                if (j < cnb) codes[j] = pair_fields<KH>(code);
                nvalid += __builtin_popcountll(__ballot(code != kInvalidCode));
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int j = 0; j < cn; j += 16) {
                uint32_t c[16];
                load_codes8(codes, j, *reinterpret_cast<uint32_t(*)[8]>(&c[0]));
                load_codes8(codes, j + 8, *reinterpret_cast<uint32_t(*)[8]>(&c[8]));
                uint32_t r[2][KH][NW];
                auto issue = [&](int u, uint32_t (&rr)[KH][NW]) {
                    const uint32_t code = __builtin_amdgcn_readfirstlane(c[u]);
#pragma unroll
                    for (int jj = 0; jj < KH; ++jj) {
                        const uint32_t e = (code >> (5 * jj)) & 31u;
                        const u32x4 v = pair_row4(e, a128 + 4u * jj * W32P);
                        rr[jj][0] = v.x; rr[jj][1] = v.y; rr[jj][2] = v.z; rr[jj][3] = v.w;
                        if constexpr (NB32) rr[jj][4] = pair_row(e, a32 + 4u * jj * W32P);
                    }
                };
                uint32_t y0[NW], tA[NW], fA[NW], eA[NW];
                issue(0, r[0]);
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    if (u + 1 < 16) issue(u + 1, r[(u + 1) & 1]);
                    __builtin_amdgcn_sched_barrier(0);
                    const uint32_t(&q)[KH][NW] = r[u & 1];
                    uint32_t y[NW];
#pragma unroll
                    for (int g = 0; g < NW; ++g) {
                        y[g] = xor3(xor3(q[0][g], q[1][g], q[2][g]), q[3][g], q[4][g]);
                        if constexpr (K == 12) y[g] ^= q[5][g];
                    }
                    if ((u & 1) == 0) {
#pragma unroll
                        for (int g = 0; g < NW; ++g) y0[g] = y[g];
                        continue;
                    }
                    // Harley-Seal step on the pair (u - 1, u)
#pragma unroll
                    for (int g = 0; g < NW; ++g) {
                        uint32_t t, f, e8, cy;
                        csa(t, planes[0][g], planes[0][g], y0[g], y[g]);
                        if ((u & 3) == 1) { tA[g] = t; continue; }
                        csa(f, planes[1][g], planes[1][g], tA[g], t);
                        if ((u & 7) == 3) { fA[g] = f; continue; }
                        csa(e8, planes[2][g], planes[2][g], fA[g], f);
                        if (u == 7) { eA[g] = e8; continue; }
                        csa(cy, planes[3][g], planes[3][g], eA[g], e8);
#pragma unroll
                        for (int p = 4; p < P; ++p) {
                            const uint32_t tt = planes[p][g] & cy;
                            planes[p][g] ^= cy;
                            cy = tt;
                        }
                    }
                }
            }
        }

        // phase 3: the lane's five (or four) words, then the row reductions
        int64_t sq = 0;
        int mx = 0, res = 0;
#pragma unroll
        for (int g = 0; g < NW; ++g) {
            const int w = g < 4 ? 4 * lane + g : 256 + lane;
            if (w < W32) {
                uint32_t pl[P];
#pragma unroll
                for (int p = 0; p < P; ++p) pl[p] = planes[p][g];
                finish_word<P>(pl, w, nvalid, D, s, out, sq, mx, res);
            }
        }
        for (int64_t d = 32 * (int64_t)W32 + lane; d < ld; d += 64) {
            const int64_t o = s * ld + d;
            if (out.counts) out.counts[o] = 0;
            if (out.lo) out.lo[o] = 0;
            if (out.hi) out.hi[o] = 0;
            if (out.clamp8) out.clamp8[o] = 0;
        }
        sq = wave_sum64(sq);
        mx = wave_max(mx);
        res = wave_sum(res);
        if (lane == 0) {
            if (out.sqnorm) out.sqnorm[s] = sq;
            if (out.maxabs) out.maxabs[s] = mx;
            if (out.nkmers) out.nkmers[s] = nvalid;
            if (out.resid1) out.resid1[s] = res;
        }
    }
}

template <int K, int P>
int launch_kp(hvk_ctx* ctx, const uint32_t* packed, const uint32_t* invalid,
              const int64_t* offsets, const int64_t* starts, const int32_t* lens, int window_len,
              int n, const EncodeOut& out, hipStream_t st) {
    const int W32 = ctx->w32;
    const int threads = (W32 + 63) / 64 * 64;
    const int k = ctx->k;
    // pair table when an entry's k/2 rows fit its kPairStride words and the
    // table, the codes and (with more than one part) one part's counters fit
    // one workgroup; as many parts (<= 4) as 1024 threads hold, else one
    const size_t lds_one = sizeof(uint32_t) * ((size_t)16 * kPairStride +
                                               ((k & 1) ? 4 * (size_t)W32 : 0) + kChunk);
    const size_t lds_parts = lds_one + sizeof(uint32_t) * (size_t)P * W32;
    int np = std::min(4, 1024 / threads);
    if (np > 1 && lds_parts + 512 > (size_t)ctx->max_lds) np = 1;
    const size_t lds_pair = np > 1 ? lds_parts : lds_one;
    const bool pair = k >= 2 && np >= 1 && (size_t)(k / 2) * W32 <= (size_t)kPairStride &&
                      lds_pair + 512 <= (size_t)ctx->max_lds;
    int grid = n < 4 * ctx->n_cu ? n : 4 * ctx->n_cu;
    if constexpr ((K == 10 || K == 12) && P <= 16) {
        // one wave per sequence when a lane's four (or five) words cover the row
        const size_t lds_wave = sizeof(uint32_t) * ((size_t)(kZeroEntry + 1) * kPairStride +
                                                    (size_t)kWaveWaves * kWaveChunk);
        const int nb32 = W32 <= 256 ? 0 : (W32 <= 320 ? 1 : -1);
        if (nb32 >= 0 && W32 > 128 && lds_wave + 512 <= (size_t)ctx->max_lds &&
            !ctx->encode_legacy) {
            const int wgrid = std::max(1, std::min(ctx->n_cu, (n + kWaveWaves - 1) / kWaveWaves));
            timing_mark(ctx, 0, false, st);
            if (nb32)
                hipLaunchKernelGGL((encode_wave_kernel<K, P, 1>), dim3(wgrid),
                                   dim3(64 * kWaveWaves), lds_wave, st, ctx->d_table, ctx->dim,
                                   W32, packed, invalid, offsets, starts, lens, window_len, n, out);
            else
                hipLaunchKernelGGL((encode_wave_kernel<K, P, 0>), dim3(wgrid),
                                   dim3(64 * kWaveWaves), lds_wave, st, ctx->d_table, ctx->dim,
                                   W32, packed, invalid, offsets, starts, lens, window_len, n, out);
            timing_mark(ctx, 0, true, st);
            HVK_HIP(hipGetLastError());
            return HVK_OK;
        }
    }
    timing_mark(ctx, 0, false, st);
    if (pair) {
        hipLaunchKernelGGL((encode_pair_kernel<K, P>), dim3(grid), dim3(np * threads), lds_pair,
                           st, ctx->d_table, k, ctx->dim, W32, threads, np, packed, invalid,
                           offsets, starts, lens, window_len, n, out);
    } else {
        const size_t lds = sizeof(uint32_t) * ((size_t)4 * k * W32 + kChunk);
        if (lds + 256 > (size_t)ctx->max_lds) return HVK_ERR_ARG;
        hipLaunchKernelGGL((encode_kernel<K, P>), dim3(grid), dim3(threads), lds, st, ctx->d_table,
                           k, ctx->dim, W32, packed, invalid, offsets, starts, lens, window_len,
                           n, out);
    }
    timing_mark(ctx, 0, true, st);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

template <int K>
int launch_k(hvk_ctx* ctx, int max_len, const uint32_t* packed, const uint32_t* invalid,
             const int64_t* offsets, const int64_t* starts, const int32_t* lens, int window_len,
             int n, const EncodeOut& out, hipStream_t st) {
    const int64_t nmax = (int64_t)max_len - ctx->k + 1;
    if (nmax < (1 << 11))
        return launch_kp<K, 11>(ctx, packed, invalid, offsets, starts, lens, window_len, n, out, st);
    if (nmax < (1 << 16))
        return launch_kp<K, 16>(ctx, packed, invalid, offsets, starts, lens, window_len, n, out, st);
    return launch_kp<K, 24>(ctx, packed, invalid, offsets, starts, lens, window_len, n, out, st);
}

}  // namespace

int launch_encode(hvk_ctx* ctx, const uint32_t* packed, const uint32_t* invalid,
                  const int64_t* offsets, const int64_t* starts, int window_len, int n,
                  int max_len, const EncodeOut& out, hipStream_t st, const int32_t* lens) {
    switch (ctx->k) {
        case 5: return launch_k<5>(ctx, max_len, packed, invalid, offsets, starts, lens, window_len, n, out, st);
        case 10: return launch_k<10>(ctx, max_len, packed, invalid, offsets, starts, lens, window_len, n, out, st);
        case 12: return launch_k<12>(ctx, max_len, packed, invalid, offsets, starts, lens, window_len, n, out, st);
        default: return launch_k<0>(ctx, max_len, packed, invalid, offsets, starts, lens, window_len, n, out, st);
    }
}

}  // namespace hvk

namespace {
int check_out(const hvk_ctx* ctx, hvk::EncodeOut& o) {
    if (o.ld == 0) o.ld = ctx->ld;
    if (o.ld < 32 * (int64_t)ctx->w32 || (o.ld % 16) != 0) return HVK_ERR_ARG;
    return HVK_OK;
}
hvk::EncodeOut to_out(const hvk_encode_out* o) {
    hvk::EncodeOut e{};
    if (!o) return e;
    e.counts = o->counts;
    e.lo = o->lo;
    e.hi = o->hi;
    e.clamp8 = o->clamp8;
    e.ld = o->ld;
    e.sqnorm = o->sqnorm;
    e.maxabs = o->maxabs;
    e.nkmers = o->nkmers;
    e.resid1 = o->resid1;
    return e;
}
}  // namespace

extern "C" int hvk_encode_ex(hvk_ctx* ctx, const uint32_t* d_packed, const uint32_t* d_invalid,
                             const int64_t* d_offsets, int n_seqs, int max_len,
                             const hvk_encode_out* out, hvk_stream stream) {
    if (!ctx || !d_packed || !d_invalid || !d_offsets || !out || n_seqs < 0 || max_len < 0)
        return HVK_ERR_ARG;
    if (n_seqs == 0) return HVK_OK;
    hvk::EncodeOut o = to_out(out);
    if (int st = check_out(ctx, o)) return st;
    if ((int64_t)max_len - ctx->k + 1 >= (1 << 24)) return HVK_ERR_RANGE;
    return hvk::launch_encode(ctx, d_packed, d_invalid, d_offsets, nullptr, 0, n_seqs, max_len, o,
                              (hipStream_t)stream);
}

extern "C" int hvk_encode_windows(hvk_ctx* ctx, const uint32_t* d_packed,
                                  const uint32_t* d_invalid, const int64_t* d_starts,
                                  int n_windows, int window_len, const hvk_encode_out* out,
                                  hvk_stream stream) {
    if (!ctx || !d_packed || !d_invalid || !d_starts || !out || n_windows < 0 || window_len < 0)
        return HVK_ERR_ARG;
    if (n_windows == 0) return HVK_OK;
    hvk::EncodeOut o = to_out(out);
    if (int st = check_out(ctx, o)) return st;
    if ((int64_t)window_len - ctx->k + 1 >= (1 << 24)) return HVK_ERR_RANGE;
    return hvk::launch_encode(ctx, d_packed, d_invalid, nullptr, d_starts, window_len, n_windows,
                              window_len, o, (hipStream_t)stream);
}

extern "C" int hvk_encode(hvk_ctx* ctx, const uint32_t* d_packed, const uint32_t* d_invalid,
                          const int64_t* d_offsets, int n_seqs, int max_len, int16_t* d_counts,
                          int8_t* d_lo, int8_t* d_hi, int64_t ld, int64_t* d_sqnorm,
                          int32_t* d_maxabs, int32_t* d_nkmers, hvk_stream stream) {
    hvk_encode_out o{};
    o.counts = d_counts;
    o.lo = d_lo;
    o.hi = d_hi;
    o.ld = ld;
    o.sqnorm = d_sqnorm;
    o.maxabs = d_maxabs;
    o.nkmers = d_nkmers;
    return hvk_encode_ex(ctx, d_packed, d_invalid, d_offsets, n_seqs, max_len, &o, stream);
}
