// descend.hip -- hierarchical descent on the GPU: identify's rank-by-rank
// beam (targets/identify.cpp:89-202) with one integer prototype bundle per
// category, scored by the reference's cosine (src/HV.cpp:219-242) from exact
// int64 dots.
//
// One level is seven launches on the caller's stream, with no host round trip:
//   1. bucket count: one key per live (read, beam entry) = its category, on
//      one of two paths: int8 (the read's counts all fit int8 and every
//      candidate of the category splits into two int8 digits) or int16
//   2. bucket scan (tile sums, one-workgroup scan of the sums, tile scans)
//   3. scatter: (read, entry) pairs grouped by path, then category
//   4. score, int8 path: v_dot4_i32_i8 on the read's int8 row and the
//      candidates' two digit planes (p = d0 + 256 d1), eight pairs per wave
//   5. score, int16 path: every other pair against the candidate list of its
//      category.  A wave takes 4 consecutive pairs of the grouped list
//      (mostly one category) and splits K across its lanes, so every load
//      reads 1 KB of one row; each child slab is loaded once for all of the
//      wave's pairs of that category.  v_dot2_i32_i16 multiplies int16
//      counts by the int16 limbs (p = 65536 hi + lo) into int32 partial
//      sums, widened to int64 every E elements per lane, E chosen so no
//      partial can overflow; the lane sums are reduced across the wave.
//   6. select: per read, gather its candidates in identify's order (beam
//      order, then each entry's candidate list), sort them exactly as
//      std::sort does (lsort.h), then exit or keep the beam.
// Level 0, where every read scores the roots, is one dense matrix-core
// product that also writes the int8 read rows.  Both paths compute the same
// exact integer dots, so the path taken never changes a score.  A final
// launch writes the surviving reads' best paths.
#include <algorithm>
#include <vector>

#include "hvk_internal.h"
#include "lsort.h"

namespace hvk {
namespace {

// candidates scored per pass over a read row: 5 when no candidate list is
// longer (the registers of 5 fit five waves per SIMD), else 8 (four waves)
constexpr int kDescGSmall = 5, kDescG = 8;
constexpr int kDescThreads = 256;
constexpr int kMaxBeam = 16;
constexpr int kMaxLevels = 32;
// bucket counters are kept in `copies` copies per key (workgroup b uses copy
// b % copies) so that few hot keys -- 40 phyla for 3M pairs at level 1 --
// do not serialise on one atomic; the scan runs over (key, copy) in key-major
// order, so the pairs stay grouped by key.  64 copies while the counters stay
// within 2^24 (a 76,040-category taxonomy, 3M pairs over 40 phyla: the level-1
// count and scatter take 0.23 ms each, 1.3 ms with 8 copies), halved down to 8
// for larger trees.
constexpr int kDescCopiesMax = 64, kDescCopiesMin = 8;
constexpr int64_t kDescMaxCounters = int64_t(1) << 24;

typedef short short2_t __attribute__((ext_vector_type(2)));

// Beam state of one level, structure of arrays over reads.
struct Beams {
    int32_t* nb;    // [n] live entries (0: the read has exited)
    int32_t* cat;   // [n][B] last category of each entry (-1: the root)
    int32_t* path;  // [n][B][L]
    float* pscore;  // [n][B][L]
};

// src/HV.cpp:236-241 on the exact integer dot and norms, as a candidate's
// score: identify starts every candidate at -1.0f and keeps the max of that
// and its similarity (targets/identify.cpp:122-128,156), so a negative
// cosine scores -1, not below it
__device__ __forceinline__ float ref_score(int64_t dot, int64_t nq, int64_t np) {
    if (nq == 0 || np == 0) return -1.0f;
    return fmaxf(-1.0f, (float)(2.0 * ((double)dot / (sqrt((double)nq) * sqrt((double)np)) - 0.5)));
}

// candidate list that follows category c (c < 0: the root)
__device__ __forceinline__ void span_of(const TreeDev& t, int c, const int32_t*& list, int& n) {
    if (c < 0) {
        list = t.roots;
        n = t.n_roots;
    } else {
        const int a = t.child_off[c];
        list = t.child_idx + a;
        n = t.child_off[c + 1] - a;
    }
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ctr[key] += 1 for every lane with act set, one atomic per distinct key in
// the wave; returns each such lane's slot (the counter value it took).  All
// lanes of the wave must call it.
__device__ __forceinline__ int wave_agg_add(int32_t* ctr, int key, bool act) {
    int slot = 0;
    const int lane = lane_id();
    uint64_t todo = __ballot(act);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int k = __shfl(key, leader);
        const uint64_t m = __ballot(act && key == k) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(&ctr[k], __popcll(m));
        base = __shfl(base, leader);
        if ((m >> lane) & 1) slot = base + __popcll(m & ((1ull << lane) - 1));
        todo &= ~m;
    }
    return slot;
}

__global__ void desc_fill(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void desc_init(Beams b, int n, int B) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    b.nb[r] = 1;
    b.cat[(int64_t)r * B] = -1;
}

// key of an entry: its category, n_cat for the root.  The bucket of a pair is
// (path, key, copy), path-major: pairs whose read has an int8 row (qmode 0)
// and whose candidates all have digit planes (cflag) come first, scored by
// desc_score8; the others follow, scored by desc_score from the int16 rows.
// Without the int8 rows (qmode null) every pair takes path 1.
__device__ __forceinline__ int pair_bucket(int key, int r, int n_cat, const uint8_t* qmode,
                                           const uint8_t* cflag, int copies) {
    const int path = (qmode && qmode[r] == 0 && cflag[key]) ? 0 : 1;
    return (path * (n_cat + 1) + key) * copies + (int)(blockIdx.x % (unsigned)copies);
}

__global__ void desc_hist(Beams b, int n, int B, int n_cat, const uint8_t* __restrict__ qmode,
                          const uint8_t* __restrict__ cflag, int copies, int32_t* hist) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = r < n ? b.nb[r] : 0;
    for (int j = 0; j < B; ++j) {
        const bool act = j < nb;
        int bucket = 0;
        if (act) {
            const int c = b.cat[(int64_t)r * B + j];
            bucket = pair_bucket(c < 0 ? n_cat : c, r, n_cat, qmode, cflag, copies);
        }
        (void)wave_agg_add(hist, bucket, act);
    }
}

// Exclusive scan of the bucket counters in three launches: per-tile sums,
// a one-workgroup scan of the tile sums (and the total), then each tile's
// scan plus its offset.  A tile is 1024 threads x 4 consecutive counters.
constexpr int kScanTile = 4096;

// inclusive scan of v over the workgroup's 1024 threads (16 waves)
__device__ __forceinline__ int block_scan_incl(int v, int32_t* s16) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) s16[wv] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        int x = threadIdx.x < 16 ? s16[threadIdx.x] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int u = __shfl_up(x, o, 64);
            if (lane >= o) x += u;
        }
        if (threadIdx.x < 16) s16[threadIdx.x] = x;  // inclusive wave totals
    }
    __syncthreads();
    const int before = wv > 0 ? s16[wv - 1] : 0;
    __syncthreads();  // s16 may be reused by the caller
    return v + before;
}

__global__ void __launch_bounds__(1024)
desc_scan_sums(const int32_t* __restrict__ hist, int n_keys, int32_t* __restrict__ part) {
    __shared__ int32_t s16[16];
    const int64_t j = (int64_t)blockIdx.x * kScanTile + 4 * threadIdx.x;
    int v = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) v += j + q < n_keys ? hist[j + q] : 0;
    v = block_scan_incl(v, s16);
    if (threadIdx.x == 1023) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(1024)
desc_scan_top(int32_t* part, int nb, int32_t* total) {
    __shared__ int32_t s16[16];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int b = b0 + threadIdx.x;
        const int v = b < nb ? part[b] : 0;
        const int incl = block_scan_incl(v, s16);
        const int c = carry;
        if (b < nb) part[b] = c + incl - v;  // exclusive
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(1024)
desc_scan_apply(const int32_t* __restrict__ hist, int n_keys, const int32_t* __restrict__ part,
                int32_t* __restrict__ cursor) {
    __shared__ int32_t s16[16];
    const int64_t j = (int64_t)blockIdx.x * kScanTile + 4 * threadIdx.x;
    int h[4], v = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        h[q] = j + q < n_keys ? hist[j + q] : 0;
        v += h[q];
    }
    int run = part[blockIdx.x] + block_scan_incl(v, s16) - v;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (j + q < n_keys) cursor[j + q] = run;
        run += h[q];
    }
}

__global__ void desc_scatter(Beams b, int n, int B, int n_cat, const uint8_t* __restrict__ qmode,
                             const uint8_t* __restrict__ cflag, int copies, int32_t* cursor,
                             int32_t* pairs, int32_t* pkey) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = r < n ? b.nb[r] : 0;
    for (int j = 0; j < B; ++j) {
        const bool act = j < nb;
        int key = 0, bucket = 0;
        if (act) {
            const int c = b.cat[(int64_t)r * B + j];
            key = c < 0 ? n_cat : c;
            bucket = pair_bucket(key, r, n_cat, qmode, cflag, copies);
        }
        const int pos = wave_agg_add(cursor, bucket, act);
        if (act) {
            pairs[pos] = r * B + j;
            pkey[pos] = key;
        }
    }
}

// Read rows as int8: q8[r] = count[r] where every count is in [-128, 127]
// (then the row is exact in int8), qmode[r] = 0; else qmode[r] = 1 and the
// int16 row is used.  One wave per read; columns [dim, K) are zero in every
// valid row.
__global__ void desc_q8(const int16_t* __restrict__ q, int64_t qld, int n, int64_t K,
                        int8_t* __restrict__ q8, uint8_t* __restrict__ qmode) {
    const int r = blockIdx.x * (kDescThreads / 64) + (threadIdx.x >> 6);
    if (r >= n) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    const int16_t* row = q + (int64_t)r * qld;
    int8_t* out = q8 + (int64_t)r * K;
    bool bad = false;
    for (int64_t k = 16 * lane; k < K; k += 1024) {
        const int4 a = *reinterpret_cast<const int4*>(row + k);
        const int4 b = *reinterpret_cast<const int4*>(row + k + 8);
        const int w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint32_t o[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int lo0 = (int)(int16_t)w[2 * d], hi0 = w[2 * d] >> 16;
            const int lo1 = (int)(int16_t)w[2 * d + 1], hi1 = w[2 * d + 1] >> 16;
            bad |= ((uint32_t)(lo0 + 128) | (uint32_t)(hi0 + 128) | (uint32_t)(lo1 + 128) |
                    (uint32_t)(hi1 + 128)) > 255u;
            o[d] = (uint32_t)(lo0 & 0xFF) | ((uint32_t)(hi0 & 0xFF) << 8) |
                   ((uint32_t)(lo1 & 0xFF) << 16) | ((uint32_t)(hi1 & 0xFF) << 24);
        }
        *reinterpret_cast<uint4*>(out + k) = make_uint4(o[0], o[1], o[2], o[3]);
    }
    const bool any_bad = __ballot(bad) != 0;  // every lane of the wave votes
    if (lane == 0) qmode[r] = any_bad ? 1 : 0;
}

// Prototype rows as two signed base-256 int8 digits, p = d0 + 256 d1, where
// every |p| <= 32639 (dflag[c] = 1), for the int8 path of desc_score; rows
// with a larger entry keep dflag[c] = 0 and the int16 limbs.  One wave per
// category.
__global__ void desc_digits(TreeDev t, int64_t K, int8_t* __restrict__ d0, int8_t* __restrict__ d1,
                            uint8_t* __restrict__ dflag) {
    const int c = blockIdx.x * (kDescThreads / 64) + (threadIdx.x >> 6);
    if (c >= t.n_cat) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    const int16_t* lo = t.lo + (int64_t)c * t.ld;
    const int16_t* hi = t.hi + (int64_t)c * t.ld;
    int8_t* o0 = d0 + (int64_t)c * K;
    int8_t* o1 = d1 + (int64_t)c * K;
    bool ok = true;
    for (int64_t k = 8 * lane; k < K; k += 512) {
        const int4 a = *reinterpret_cast<const int4*>(lo + k);
        const int4 b = *reinterpret_cast<const int4*>(hi + k);
        const int la[4] = {a.x, a.y, a.z, a.w}, hb[4] = {b.x, b.y, b.z, b.w};
        uint32_t x0[2] = {0, 0}, x1[2] = {0, 0};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int l = (int)(int16_t)(la[e >> 1] >> (16 * (e & 1)));
            const int h = (int)(int16_t)(hb[e >> 1] >> (16 * (e & 1)));
            const int64_t p = (int64_t)l + 65536 * (int64_t)h;
            ok = ok && p >= -32639 && p <= 32639;
            const int v0 = (int)(int8_t)(p & 0xFF);
            const int v1 = (int)((p - v0) >> 8);
            x0[e >> 2] |= (uint32_t)(v0 & 0xFF) << (8 * (e & 3));
            x1[e >> 2] |= (uint32_t)(v1 & 0xFF) << (8 * (e & 3));
        }
        *reinterpret_cast<uint2*>(o0 + k) = make_uint2(x0[0], x0[1]);
        *reinterpret_cast<uint2*>(o1 + k) = make_uint2(x1[0], x1[1]);
    }
    const bool all = __ballot(!ok) == 0;
    if (lane == 0) dflag[c] = all ? 1 : 0;
}

// The int8 rows of desc_score8: read rows q8 (qmode 0) and prototype digit
// planes d0, d1 (dflag 1), row stride K; cflag per candidate list.
struct DescI8 {
    int8_t* q8;
    uint8_t* qmode;
    int8_t* d0;
    int8_t* d1;
    uint8_t* dflag;
    uint8_t* cflag;
};

__device__ __forceinline__ int32_t sdot2(int32_t a, int32_t b, int32_t c) {
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, a), __builtin_bit_cast(short2_t, b),
                                  c, false);
}

// Scores pairs [*begin, *total) of the grouped list.  E: elements per lane
// between widenings of the int32 partial sums (8 or 2; a 32-element chunk
// would keep 2*P*G more partials live and halve the waves per SIMD).
//
// One wave takes kDescP consecutive pairs of the grouped list and splits K
// across its lanes: lane l holds elements [8l, 8l+8) of every 512-element
// step, so each load instruction reads 1 KB of one row (one page, fully
// coalesced) -- a lane-per-pair layout would touch 64 rows, and at the deep
// levels 64 pages, per instruction.  The wave walks the distinct categories
// of its pairs (wave-uniform), G children at a time; each child slab is
// loaded once for all the wave's pairs of that category.  The per-lane int64
// sums are reduced across the wave at the end of each pass.
// pairs per wave: 4 leaves ~120 VGPRs (G = 8) and four waves per SIMD, which
// beat 8 pairs (one child load shared by more pairs, two waves per SIMD) and 2
constexpr int kDescP = 4;
constexpr int kDescStep = 64 * 8;         // elements per wave step

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int E, int GW>
__global__ void __launch_bounds__(kDescThreads, GW <= kDescGSmall ? 5 : 4)
desc_score(TreeDev t, const int16_t* __restrict__ q, int64_t qld, const int64_t* __restrict__ qn,
           int B, int64_t K, const int32_t* __restrict__ pairs, const int32_t* __restrict__ pkey,
           const int32_t* __restrict__ begin, const int32_t* __restrict__ total, int Fmax,
           float* __restrict__ sc) {
    static_assert(E == 8 || E == 2, "partial-sum chunk");
    const int n_pairs = *total;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int base = *begin + (blockIdx.x * (kDescThreads / 64) + wave) * kDescP;
    if (base >= n_pairs) return;  // uniform per wave
    const int lane = lane_id();
    int key[kDescP], pr[kDescP];
#pragma unroll
    for (int p = 0; p < kDescP; ++p) {
        const bool ok = base + p < n_pairs;
        key[p] = ok ? pkey[base + p] : -1;
        pr[p] = ok ? pairs[base + p] : 0;
    }
    uint32_t todo = 0;
#pragma unroll
    for (int p = 0; p < kDescP; ++p) todo |= (key[p] >= 0 ? 1u : 0u) << p;
    while (todo) {
        const int pk = key[__builtin_ctz(todo)];
        uint32_t mine = 0;
#pragma unroll
        for (int p = 0; p < kDescP; ++p) mine |= (key[p] == pk ? 1u : 0u) << p;
        todo &= ~mine;
        const int32_t* list;
        int F;
        span_of(t, pk == t.n_cat ? -1 : pk, list, F);
        const int16_t* qrow[kDescP];
#pragma unroll
        for (int p = 0; p < kDescP; ++p) qrow[p] = q + (int64_t)(pr[p] / B) * qld;
        for (int g0 = 0; g0 < F; g0 += GW) {
            const int G = min(GW, F - g0);
            const int16_t* lo[GW];
            const int16_t* hi[GW];
#pragma unroll
            for (int g = 0; g < GW; ++g) {
                const int c = list[g0 + (g < G ? g : G - 1)];
                lo[g] = t.lo + (int64_t)c * t.ld;
                hi[g] = t.hi + (int64_t)c * t.ld;
            }
            // A child with every |p| < 2^15 has a zero hi limb row (p = lo),
            // which is then neither loaded nor multiplied: by its row maximum
            // when the tree carries them, else when its squared norm is below
            // 2^30 (which bounds every |p| below 2^15).
            bool lo_only = true;
#pragma unroll
            for (int g = 0; g < GW; ++g)
                if (g < G) {
                    const int c = list[g0 + g];
                    lo_only = lo_only && (t.row_maxabs ? t.row_maxabs[c] < 32768
                                                       : t.sqnorm[c] < (int64_t(1) << 30));
                }
            int64_t acc[kDescP][GW];
#pragma unroll
            for (int p = 0; p < kDescP; ++p)
#pragma unroll
                for (int g = 0; g < GW; ++g) acc[p][g] = 0;
            auto pass = [&](auto HI) {
                constexpr bool kHi = decltype(HI)::value;
                for (uint32_t k = 8 * lane; k < (uint32_t)K; k += kDescStep) {
                    int4 qv[kDescP];
#pragma unroll
                    for (int p = 0; p < kDescP; ++p)
                        qv[p] = ((mine >> p) & 1) ? *reinterpret_cast<const int4*>(qrow[p] + k)
                                                   : make_int4(0, 0, 0, 0);
#pragma unroll
                    for (int g = 0; g < GW; ++g) {
                        if (g >= G) continue;  // uniform
                        const int4 lv = *reinterpret_cast<const int4*>(lo[g] + k);
                        const int4 hv = kHi ? *reinterpret_cast<const int4*>(hi[g] + k)
                                            : make_int4(0, 0, 0, 0);
                        const int32_t l4[4] = {lv.x, lv.y, lv.z, lv.w};
                        const int32_t h4[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
                        for (int p = 0; p < kDescP; ++p) {
                            const int32_t q4[4] = {qv[p].x, qv[p].y, qv[p].z, qv[p].w};
                            int32_t pl = 0, ph = 0;
#pragma unroll
                            for (int d = 0; d < 4; ++d) {
                                pl = sdot2(q4[d], l4[d], pl);
                                if (kHi) ph = sdot2(q4[d], h4[d], ph);
                                if (E == 2) {
                                    acc[p][g] += kHi ? (int64_t)pl + ((int64_t)ph << 16) : (int64_t)pl;
                                    pl = 0;
                                    ph = 0;
                                }
                            }
                            if (E == 8)
                                acc[p][g] += kHi ? (int64_t)pl + ((int64_t)ph << 16) : (int64_t)pl;
                        }
                    }
                }
            };
            if (lo_only) pass(std::false_type{});
            else pass(std::true_type{});
#pragma unroll
            for (int p = 0; p < kDescP; ++p)
#pragma unroll
                for (int g = 0; g < GW; ++g) {
                    if (g < G && ((mine >> p) & 1)) {
                        const int64_t dot = wave_sum_i64(acc[p][g]);
                        if (lane == 0) {
                            const int r = pr[p] / B;
                            const int c = list[g0 + g];
                            sc[(int64_t)pr[p] * Fmax + g0 + g] = ref_score(dot, qn[r], t.sqnorm[c]);
                        }
                    }
                }
        }
    }
}

// cflag[key] = 1 when every candidate of key (the children of category key,
// the roots for key n_cat) has two digit planes.
__global__ void desc_cflag(TreeDev t, const uint8_t* __restrict__ dflag, uint8_t* __restrict__ cflag) {
    const int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key > t.n_cat) return;
    const int32_t* list;
    int F;
    span_of(t, key == t.n_cat ? -1 : key, list, F);
    bool ok = true;
    for (int i = 0; i < F; ++i) ok = ok && dflag[list[i]] != 0;
    cflag[key] = ok ? 1 : 0;
}

// Scores pairs [0, *n_i8) of the grouped list: reads with int8 rows against
// candidates with two int8 digit planes.  One wave takes P consecutive pairs
// (mostly one category) and walks their candidates GW at a time; lane l holds
// elements [4 NW l, 4 NW l + 4 NW) of every 256 NW-element step, so each load
// reads 256 NW bytes of one row.  Each lane sums q d0 + 256 q d1 (v_dot4 on
// int8 operands) in one int32 per (pair, candidate): with |q| <= 128 and
// |p| <= 32639, at most K / 64 * 128 * 32639 < 2^31 for K <= 32512 (checked by
// the host), and every partial sum of a step stays within 2^31 as well.
// Eight pairs per wave and 8-byte lane slices: one candidate slab serves eight
// read rows, and the 40 accumulators (G = 5), 16 read and 20 candidate
// registers fit four waves per SIMD with every load of a step in flight
// (measured on the whole descent: four pairs per wave 7% slower, 16-byte
// slices 1% slower; skipping the slots past the list end made no difference
// there and saves work on short lists).
constexpr int kDescP8 = 8;

template <int N>
struct IVec {
    typedef int32_t type __attribute__((ext_vector_type(N)));
};

template <int P, int NW, int GW, int OCC>
__global__ void __launch_bounds__(kDescThreads, OCC)
desc_score8(TreeDev t, const int8_t* __restrict__ q8, const int64_t* __restrict__ qn, int B,
            int64_t K, const int8_t* __restrict__ d0, const int8_t* __restrict__ d1,
            const int32_t* __restrict__ pairs, const int32_t* __restrict__ pkey,
            const int32_t* __restrict__ n_i8, int Fmax, float* __restrict__ sc) {
    typedef typename IVec<NW>::type vec;
    const int n_pairs = *n_i8;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int base = (blockIdx.x * (kDescThreads / 64) + wave) * P;
    if (base >= n_pairs) return;  // uniform per wave
    const int lane = lane_id();
    int key[P], pr[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const bool ok = base + p < n_pairs;
        key[p] = ok ? pkey[base + p] : -1;
        pr[p] = ok ? pairs[base + p] : 0;
    }
    uint32_t todo = 0;
#pragma unroll
    for (int p = 0; p < P; ++p) todo |= (key[p] >= 0 ? 1u : 0u) << p;
    while (todo) {
        const int pk = key[__builtin_ctz(todo)];
        uint32_t mine = 0;
#pragma unroll
        for (int p = 0; p < P; ++p) mine |= (key[p] == pk ? 1u : 0u) << p;
        todo &= ~mine;
        const int32_t* list;
        int F;
        span_of(t, pk == t.n_cat ? -1 : pk, list, F);
        // every load of a step is issued unconditionally, so that one wait
        // covers them all: a pair of another category reads the row of this
        // pass's first pair (same cache lines), a slot past the list end
        // reads the last candidate, and their sums are dropped
        int rfirst = 0;
#pragma unroll
        for (int p = P - 1; p >= 0; --p)
            if ((mine >> p) & 1) rfirst = pr[p];
        const int8_t* qrow[P];
#pragma unroll
        for (int p = 0; p < P; ++p)
            qrow[p] = q8 + (int64_t)((((mine >> p) & 1) ? pr[p] : rfirst) / B) * K;
        for (int g0 = 0; g0 < F; g0 += GW) {
            const int G = min(GW, F - g0);
            const int8_t* u_row[GW];
            const int8_t* v_row[GW];
#pragma unroll
            for (int g = 0; g < GW; ++g) {
                const int c = list[g0 + min(g, G - 1)];
                u_row[g] = d0 + (int64_t)c * K;
                v_row[g] = d1 + (int64_t)c * K;
            }
            int32_t a[P][GW];
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int g = 0; g < GW; ++g) a[p][g] = 0;
            // 32-bit lane offsets: each load is a uniform row base plus one
            // VGPR (K < 2^31 bytes)
#pragma nounroll
            for (uint32_t k = 4 * NW * lane; k < (uint32_t)K; k += 256 * NW) {
                vec qv[P];
#pragma unroll
                for (int p = 0; p < P; ++p) qv[p] = *reinterpret_cast<const vec*>(qrow[p] + k);
                vec uv[GW], vv[GW];
#pragma unroll
                for (int g = 0; g < GW; ++g) {
                    uv[g] = *reinterpret_cast<const vec*>(u_row[g] + k);
                    vv[g] = *reinterpret_cast<const vec*>(v_row[g] + k);
                }
#pragma unroll
                for (int g = 0; g < GW; ++g) {
                    if (g >= G) continue;  // uniform
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        int32_t t1 = 0;
#pragma unroll
                        for (int w = 0; w < NW; ++w)
                            t1 = __builtin_amdgcn_sdot4(qv[p][w], vv[g][w], t1, false);
                        int32_t s0 = a[p][g] + (t1 << 8);
#pragma unroll
                        for (int w = 0; w < NW; ++w)
                            s0 = __builtin_amdgcn_sdot4(qv[p][w], uv[g][w], s0, false);
                        a[p][g] = s0;
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int g = 0; g < GW; ++g) {
                    if (g < G && ((mine >> p) & 1)) {
                        const int64_t dot = wave_sum_i64((int64_t)a[p][g]);
                        if (lane == 0) {
                            const int r = pr[p] / B;
                            const int c = list[g0 + g];
                            sc[(int64_t)pr[p] * Fmax + g0 + g] = ref_score(dot, qn[r], t.sqnorm[c]);
                        }
                    }
                }
        }
    }
}

// Level 0 as a limb GEMM on the matrix cores.  At level 0 every read scores
// the same candidate list (the roots), so the scores are a dense product
// reads x roots x D.  Operands are split into signed base-256 digits: a read
// count q = q0 + 256 q1 (|q| <= 32639), a prototype entry
// p = d0 + 256 d1 + ... over ND digits, all in [-128, 127]: ND = 2 covers
// |p| <= 32639, 3 covers |p| <= 8355711, 4 every |p| < 2^30 (the host picks
// the fewest for the roots' largest |p|).  v_mfma_i32_32x32x32_i8 multiplies
// digit pairs; the 2 ND products fall into ND + 1 weights 2^(8w) (w = a + b),
// each accumulated in int32 -- at most 2 products x 10240 x 128 x 128 < 2^31
// -- and combined in int64 at the end: the exact dot.  A wave owns 32 reads;
// the workgroup (4 waves) stages each 32-element K step of the roots' digits
// (split once per launch by desc_root_digits) in LDS, double-buffered with one
// barrier a step, and every step's loads go out a step ahead.  Lane l holds
// row / column l % 32 and K bytes [16 (l / 32), +16) of both operands, as the
// MFMA pairs them.  (Staging the read rows through LDS too, so that each load
// reads whole lines: 8.4 -> 11.4 ms per 1M reads, not kept.)
typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef int v16i_t __attribute__((ext_vector_type(16)));
constexpr int kDenseN = 64;     // roots (two 32-wide blocks)
constexpr int kDenseRow = 48;   // LDS bytes per (digit, root) row: 32 + 16 pad, conflict-free
constexpr int kDenseStage = 144;  // LDS bytes per staged int8 read row: 128 + 16 pad

__device__ __forceinline__ uint32_t pack4(int a, int b, int c, int d) {
    return (uint32_t)(a & 0xFF) | ((uint32_t)(b & 0xFF) << 8) | ((uint32_t)(c & 0xFF) << 16) |
           ((uint32_t)(d & 0xFF) << 24);
}

// The roots' prototype rows as ND signed base-256 digits for desc_dense, in
// the order its workgroups stage them: for each 32-element K step s, digit g
// and root slot c, the 32 bytes of digit g of p[roots[c]][32 s, 32 s + 32)
// (p = d0 + 256 d1 + ...), zero for c >= n_roots up to kDenseN.  A step's
// staging then reads ND contiguous 2 KB blocks.  One thread per 8 elements.
template <int ND>
__global__ void desc_root_digits(TreeDev t, int64_t K, int8_t* __restrict__ droot) {
    const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (j >= (int64_t)kDenseN * K) return;
    const int c = (int)(j / K);
    const int64_t k = j - (int64_t)c * K;
    const int cat = c < t.n_roots ? t.roots[c] : -1;
    int4 lv = make_int4(0, 0, 0, 0), hv = make_int4(0, 0, 0, 0);
    if (cat >= 0) {
        lv = *reinterpret_cast<const int4*>(t.lo + (int64_t)cat * t.ld + k);
        if (ND > 2) hv = *reinterpret_cast<const int4*>(t.hi + (int64_t)cat * t.ld + k);
    }  // ND = 2: every |p| <= 32639, so the hi limb is zero
    const int lw[4] = {lv.x, lv.y, lv.z, lv.w}, hw[4] = {hv.x, hv.y, hv.z, hv.w};
    int d[ND][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int lo = (int)(int16_t)(lw[e >> 1] >> (16 * (e & 1)));
        const int hi = (int)(int16_t)(hw[e >> 1] >> (16 * (e & 1)));
        int p = lo + hi * 65536;
#pragma unroll
        for (int g = 0; g + 1 < ND; ++g) {
            d[g][e] = (int)(int8_t)p;
            p = (p - d[g][e]) >> 8;
        }
        d[ND - 1][e] = p;
    }
    const int64_t step = k >> 5;
#pragma unroll
    for (int g = 0; g < ND; ++g)
        *reinterpret_cast<uint2*>(droot + ((step * ND + g) * kDenseN + c) * 32 + (k & 31)) =
            make_uint2(pack4(d[g][0], d[g][1], d[g][2], d[g][3]),
                       pack4(d[g][4], d[g][5], d[g][6], d[g][7]));
}

// (Two workgroups per CU: the accumulator registers -- 160 at ND = 4 -- and
// the rest fit in 256, and the second wave per SIMD hides the read-row loads.)
template <int ND>
__global__ void __launch_bounds__(256, 2)
desc_dense(TreeDev t, const int16_t* __restrict__ q, int64_t qld, const int64_t* __restrict__ qn,
           int n, int B, int64_t K, int Fmax, float* __restrict__ sc,
           const int8_t* __restrict__ droot, int8_t* __restrict__ q8, uint8_t* __restrict__ qmode) {
    static_assert(ND >= 2 && ND <= 4, "prototype digits");
    // the roots' digits of one 32-element K step, double-buffered: step s reads
    // buffer s & 1 while it stages step s + 1 in the other, one barrier a step
    __shared__ __attribute__((aligned(16))) int8_t bl[2][ND][kDenseN][kDenseRow];
    // the int8 rows pass through LDS, 128 K bytes of the wave's 32 reads at a
    // time, so that each store writes whole 128-byte lines (8 rows x 128 B)
    __shared__ __attribute__((aligned(16))) int8_t qst[4][32][kDenseStage];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = (blockIdx.x * 4 + wave) * 32;
    const int i = lane & 31, kg = lane >> 5;
    const int16_t* __restrict__ qrow = q + (int64_t)min(r0 + i, n - 1) * qld;  // rows past n: clamped
    // staging: thread -> (root bc, 8 elements at bk) of the K step, from the
    // roots' digits in step order (desc_root_digits; slots past the roots are
    // zero): 8 bytes at tid * 8 of each digit's 2 KB block
    const int bc = tid >> 2, bk = (tid & 3) * 8;
    auto load_b = [&](int64_t k0, uint2 (&dv)[ND]) {
        const int8_t* blk = droot + (k0 >> 5) * ND * (kDenseN * 32) + tid * 8;
#pragma unroll
        for (int g = 0; g < ND; ++g)
            dv[g] = *reinterpret_cast<const uint2*>(blk + g * (kDenseN * 32));
    };
    auto stage_b = [&](int buf, const uint2 (&dv)[ND]) {
#pragma unroll
        for (int g = 0; g < ND; ++g) *reinterpret_cast<uint2*>(&bl[buf][g][bc][bk]) = dv[g];
    };
    v16i_t acc[ND + 1][2];
#pragma unroll
    for (int w = 0; w < ND + 1; ++w)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) acc[w][nb] = v16i_t{};
    int qbad = 0;  // nonzero when some q of this lane's part of the row is outside int8
    // the read's K bytes [16 kg, +16) of the current step, and the roots' slabs
    int4 qa = *reinterpret_cast<const int4*>(qrow + 16 * kg);
    int4 qb = *reinterpret_cast<const int4*>(qrow + 16 * kg + 8);
    {
        uint2 dv[ND];
        load_b(0, dv);
        stage_b(0, dv);
    }
    __syncthreads();
    for (int64_t k0 = 0; k0 < K; k0 += 32) {
        const int buf = (int)(k0 >> 5) & 1;
        // the next step's loads go out before this step's work (the last step
        // reloads its own slabs: in bounds, unused)
        const int64_t kn = k0 + 32 < K ? k0 + 32 : k0;
        const int4 qan = *reinterpret_cast<const int4*>(qrow + kn + 16 * kg);
        const int4 qbn = *reinterpret_cast<const int4*>(qrow + kn + 16 * kg + 8);
        uint2 dvn[ND];
        load_b(kn, dvn);
        const int qw[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
        int q0[16], q1[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int v = (int)(int16_t)(qw[e >> 1] >> (16 * (e & 1)));
            q0[e] = (int)(int8_t)v;
            q1[e] = (v - q0[e]) >> 8;
        }
        v4i_t a[2];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            a[0][c] = (int)pack4(q0[4 * c], q0[4 * c + 1], q0[4 * c + 2], q0[4 * c + 3]);
            a[1][c] = (int)pack4(q1[4 * c], q1[4 * c + 1], q1[4 * c + 2], q1[4 * c + 3]);
        }
        if (q8) {  // the int8 rows of the later levels: the low digit, exact where q1 is 0
            qbad |= a[1][0] | a[1][1] | a[1][2] | a[1][3];
            const int s4 = (int)(k0 >> 5) & 3;  // K is a multiple of 128
            *reinterpret_cast<v4i_t*>(&qst[wave][i][32 * s4 + 16 * kg]) = a[0];
            if (s4 == 3) {  // wave-local: this wave's own LDS rows, in program order
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = 8 * j + (lane >> 3), col = 16 * (lane & 7);
                    const v4i_t v = *reinterpret_cast<const v4i_t*>(&qst[wave][row][col]);
                    if (r0 + row < n)
                        *reinterpret_cast<v4i_t*>(q8 + (int64_t)(r0 + row) * K + (k0 - 96) + col) = v;
                }
            }
        }
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
#pragma unroll
            for (int g = 0; g < ND; ++g) {
                const v4i_t b = *reinterpret_cast<const v4i_t*>(&bl[buf][g][nb * 32 + i][16 * kg]);
#pragma unroll
                for (int qa_i = 0; qa_i < 2; ++qa_i)
                    acc[qa_i + g][nb] =
                        __builtin_amdgcn_mfma_i32_32x32x32_i8(a[qa_i], b, acc[qa_i + g][nb], 0, 0, 0);
            }
        // every wave is past step s - 1, so the other buffer is free
        stage_b(buf ^ 1, dvn);
        __syncthreads();  // step s + 1 staged; this step's buffer consumed
        qa = qan;
        qb = qbn;
    }
    if (q8) {
        qbad |= __shfl_xor(qbad, 32);  // lanes i and i + 32 hold one read
        if (kg == 0 && r0 + i < n) qmode[r0 + i] = qbad ? 1 : 0;
    }
    // lane l holds column l % 32, rows 8 (e / 4) + 4 (l / 32) + e % 4 of each block
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int r = r0 + 8 * (e >> 2) + 4 * kg + (e & 3);
            const int col = nb * 32 + i;
            if (r < n && col < t.n_roots) {
                int64_t dot = 0;
#pragma unroll
                for (int w = 0; w < ND + 1; ++w) dot += (int64_t)acc[w][nb][e] << (8 * w);
                sc[(int64_t)r * B * Fmax + col] = ref_score(dot, qn[r], t.sqnorm[t.roots[col]]);
            }
        }
}

struct DescOut {
    int32_t* path;   // [n][L]
    float* pscore;   // [n][L]
    int32_t* len;    // [n]
    int32_t* exit;   // [n]
};

__global__ void desc_select(TreeDev t, Beams cur, Beams nxt, int n, int B, int L, int lvl,
                            float thr, const float* __restrict__ sc, int Fmax,
                            float* __restrict__ ws_s, int32_t* __restrict__ ws_c, int maxN,
                            DescOut out, int in_lds) {
    // candidate lists: columns of LDS (element i of thread t at [i][t]) when
    // in_lds, else a row of the global workspace
    extern __shared__ float sel_lds[];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int nb = cur.nb[r];
    if (nb == 0) {  // exited at an earlier level
        nxt.nb[r] = 0;
        return;
    }
    float* s = in_lds ? sel_lds + threadIdx.x : ws_s + (int64_t)r * maxN;
    int32_t* c = in_lds ? reinterpret_cast<int32_t*>(sel_lds + (size_t)maxN * blockDim.x) + threadIdx.x
                        : ws_c + (int64_t)r * maxN;
    const int st = in_lds ? (int)blockDim.x : 1;
    int m = 0;
    for (int j = 0; j < nb; ++j) {
        const int32_t* list;
        int F;
        span_of(t, cur.cat[(int64_t)r * B + j], list, F);
        const float* row = sc + ((int64_t)r * B + j) * Fmax;
        for (int f = 0; f < F; ++f) {
            s[m * st] = row[f];
            c[m * st] = j * Fmax + f;
            ++m;
        }
    }
    lsort::sort_desc(s, c, m, st);
    // candidate id -> (beam entry, category)
    auto resolve = [&](int id, int& j, int& cat) {
        j = id / Fmax;
        const int32_t* list;
        int F;
        span_of(t, cur.cat[(int64_t)r * B + j], list, F);
        cat = list[id - j * Fmax];
    };
    if (m == 0 || s[0] < thr) {  // early exit (targets/identify.cpp:186-197)
        int len = 0;
        if (m > 0) {
            int j, cat;
            resolve(c[0], j, cat);
            const int64_t src = ((int64_t)r * B + j) * L;
            for (int x = 0; x < lvl; ++x) {
                out.path[(int64_t)r * L + x] = cur.path[src + x];
                out.pscore[(int64_t)r * L + x] = cur.pscore[src + x];
            }
            out.path[(int64_t)r * L + lvl] = cat;
            out.pscore[(int64_t)r * L + lvl] = s[0];
            len = lvl + 1;
        }
        out.len[r] = len;
        out.exit[r] = lvl;
        nxt.nb[r] = 0;
        return;
    }
    int keep = 0;
    while (keep < m && keep < B && s[keep * st] >= thr) ++keep;
    for (int i = 0; i < keep; ++i) {
        int j, cat;
        resolve(c[i * st], j, cat);
        const int64_t src = ((int64_t)r * B + j) * L;
        const int64_t dst = ((int64_t)r * B + i) * L;
        for (int x = 0; x < lvl; ++x) {
            nxt.path[dst + x] = cur.path[src + x];
            nxt.pscore[dst + x] = cur.pscore[src + x];
        }
        nxt.path[dst + lvl] = cat;
        nxt.pscore[dst + lvl] = s[i * st];
        nxt.cat[(int64_t)r * B + i] = cat;
    }
    nxt.nb[r] = keep;
}

// reads that ran every level: the first beam entry (targets/identify.cpp:205-211)
__global__ void desc_final(Beams cur, int n, int B, int L, int levels, DescOut out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || cur.nb[r] == 0) return;
    const int64_t src = (int64_t)r * B * L;
    for (int x = 0; x < levels; ++x) {
        out.path[(int64_t)r * L + x] = cur.path[src + x];
        out.pscore[(int64_t)r * L + x] = cur.pscore[src + x];
    }
    out.len[r] = levels;
    out.exit[r] = -1;
}

// ---- prototype construction -------------------------------------------

__global__ void tree_place_kernel(const int16_t* __restrict__ counts, int64_t cld, int64_t ld,
                                  const int32_t* __restrict__ cat, int32_t* __restrict__ proto) {
    const int i = blockIdx.x;
    const int16_t* src = counts + (int64_t)i * cld;
    int32_t* dst = proto + (int64_t)cat[i] * ld;
    const int64_t K = cld < ld ? cld : ld;
    for (int64_t k = threadIdx.x; k < ld; k += blockDim.x) dst[k] = k < K ? (int32_t)src[k] : 0;
}

__global__ void tree_sum_kernel(int32_t* __restrict__ proto, int64_t ld,
                                const int32_t* __restrict__ child_off,
                                const int32_t* __restrict__ child_idx,
                                const int32_t* __restrict__ parents, int32_t* overflow) {
    const int p = parents[blockIdx.x];
    const int a = child_off[p], e = child_off[p + 1];
    bool bad = false;
    for (int64_t k = threadIdx.x; k < ld; k += blockDim.x) {
        int64_t s = 0;
        for (int j = a; j < e; ++j) s += proto[(int64_t)child_idx[j] * ld + k];
        bad |= s > INT32_MAX || s < INT32_MIN;
        proto[(int64_t)p * ld + k] = (int32_t)s;
    }
    if (bad && overflow) atomicOr(overflow, 1);
}

// p = 65536 hi + lo, lo in [-32768, 32767]; a row whose hi would not fit in
// int16 (p >= 2^31 - 32768) reports 65536 as its max |limb|
__global__ void tree_limbs_kernel(const int32_t* __restrict__ proto, int64_t ld,
                                  int16_t* __restrict__ lo, int16_t* __restrict__ hi,
                                  int64_t* __restrict__ sqnorm, int32_t* __restrict__ maxabs,
                                  int32_t* __restrict__ maxabs_p) {
    __shared__ int64_t s_n[kDescThreads];
    __shared__ int32_t s_m[kDescThreads];
    __shared__ uint32_t s_p[kDescThreads];
    const int i = blockIdx.x;
    int64_t nrm = 0;
    int32_t mx = 0;
    uint32_t mp = 0;  // max |p| (2^31 for INT32_MIN)
    for (int64_t k = threadIdx.x; k < ld; k += blockDim.x) {
        const int32_t p = proto[(int64_t)i * ld + k];
        const int32_t l = (int32_t)(int16_t)(p & 0xFFFF);
        const int64_t h = ((int64_t)p - l) >> 16;
        lo[(int64_t)i * ld + k] = (int16_t)l;
        hi[(int64_t)i * ld + k] = (int16_t)h;
        nrm += (int64_t)p * p;
        const int32_t al = l < 0 ? -l : l;
        const int32_t ah = h > 32767 ? 65536 : (int32_t)(h < 0 ? -h : h);
        mx = max(mx, max(al, ah));
        mp = max(mp, p < 0 ? 0u - (uint32_t)p : (uint32_t)p);
    }
    s_n[threadIdx.x] = nrm;
    s_m[threadIdx.x] = mx;
    s_p[threadIdx.x] = mp;
    __syncthreads();
    for (int o = kDescThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_n[threadIdx.x] += s_n[threadIdx.x + o];
            s_m[threadIdx.x] = max(s_m[threadIdx.x], s_m[threadIdx.x + o]);
            s_p[threadIdx.x] = max(s_p[threadIdx.x], s_p[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (sqnorm) sqnorm[i] = s_n[0];
        if (maxabs) maxabs[i] = s_m[0];
        if (maxabs_p) maxabs_p[i] = (int32_t)min(s_p[0], 0x7FFFFFFFu);
    }
}

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// copies host int32 arrays into one device workspace (in order); returns
// the device pointers
int upload_i32(hvk_ctx* ctx, Workspace& ws, std::initializer_list<std::pair<const int32_t*, size_t>> parts,
               std::vector<const int32_t*>& dev, hipStream_t st) {
    size_t total = 0;
    for (auto& p : parts) total += align_up(p.second * sizeof(int32_t));
    if (int e = ws.ensure(total ? total : 256)) return e;
    char* base = (char*)ws.ptr;
    size_t off = 0;
    dev.clear();
    for (auto& p : parts) {
        dev.push_back((const int32_t*)(base + off));
        if (p.second)
            HVK_HIP(hipMemcpyAsync(base + off, p.first, p.second * sizeof(int32_t),
                                   hipMemcpyHostToDevice, st));
        off += align_up(p.second * sizeof(int32_t));
    }
    // the host arrays may be pageable and are the caller's: finish the copies
    HVK_HIP(hipStreamSynchronize(st));
    return HVK_OK;
}

}  // namespace
}  // namespace hvk

using namespace hvk;

namespace hvk {

// Checks the tree and the read bound, picks the partial-sum chunk E and the
// level-0 form, and uploads the candidate lists to the context (the copies
// come from the caller's host arrays and are waited for).  Not capturable.
int desc_prepare(hvk_ctx* ctx, const hvk_tree* tree, int q_maxabs, int beam, int levels,
                 DescPlan& pl, hipStream_t st) {
    if (!ctx || !tree || beam < 1 || beam > kMaxBeam || levels < 1 || levels > kMaxLevels ||
        q_maxabs < 0)
        return HVK_ERR_ARG;
    const int n_cat = tree->n_cat;
    if (n_cat < 0 || tree->n_roots < 0 || !tree->child_off || (tree->n_roots && !tree->roots))
        return HVK_ERR_ARG;
    if (n_cat > 0 && (!tree->d_lo || !tree->d_hi || !tree->d_sqnorm)) return HVK_ERR_ARG;
    const int64_t tld = tree->ld == 0 ? ctx->ld : tree->ld;
    if (tld < ctx->dim || tld % 128) return HVK_ERR_ARG;
    // candidate lists: offsets monotone from 0, indices in range
    if (tree->child_off[0] != 0) return HVK_ERR_ARG;
    int max_span = 0;
    for (int c = 0; c < n_cat; ++c) {
        const int w = tree->child_off[c + 1] - tree->child_off[c];
        if (w < 0) return HVK_ERR_ARG;
        max_span = std::max(max_span, w);
    }
    const int nnz = tree->child_off[n_cat];
    if (nnz > 0 && !tree->child_idx) return HVK_ERR_ARG;
    for (int i = 0; i < nnz; ++i)
        if (tree->child_idx[i] < 0 || tree->child_idx[i] >= n_cat) return HVK_ERR_ARG;
    for (int i = 0; i < tree->n_roots; ++i)
        if (tree->roots[i] < 0 || tree->roots[i] >= n_cat) return HVK_ERR_ARG;
    // int32 partial sums: E elements of |q| * |limb| each
    if (tree->max_abs_limb < 0 || tree->max_abs_limb > 32768) return HVK_ERR_RANGE;
    const int64_t prod = (int64_t)q_maxabs * tree->max_abs_limb;
    const int64_t lim = int64_t(1) << 31;
    if (8 * prod < lim) pl.E = 8;
    else if (2 * prod < lim) pl.E = 2;
    else return HVK_ERR_RANGE;
    pl.B = beam;
    pl.L = levels;
    pl.max_span = max_span;
    pl.Fmax = std::max(1, std::max(tree->n_roots, max_span));
    pl.maxN = std::max(1, std::max(tree->n_roots, beam * max_span));
    pl.n_cat = n_cat;
    pl.n_roots = tree->n_roots;
    pl.max_abs_proto = tree->max_abs_proto;
    // level 0 on the matrix cores when its operands fit the digit split; each
    // digit-weight class sums < 2 K 2^14 in int32
    pl.dense = tree->n_roots > 0 && tree->n_roots <= kDenseN && tree->max_abs_proto > 0 &&
               tree->max_abs_proto < (int64_t(1) << 30) && q_maxabs <= 32639 && ctx->ld < 65536;
    std::vector<const int32_t*> dev;
    if (int e = upload_i32(ctx, ctx->ws_desc_tree,
                           {{tree->child_off, (size_t)n_cat + 1},
                            {tree->child_idx, (size_t)nnz},
                            {tree->roots, (size_t)tree->n_roots}},
                           dev, st))
        return e;
    pl.t = TreeDev{tree->d_lo, tree->d_hi, tree->d_sqnorm, tld, n_cat,
                   dev[0], dev[1], dev[2], tree->n_roots, tree->d_row_maxabs};
    return HVK_OK;
}

// The launches of a descent on a prepared plan: the workspace (grown here
// when it is too small) and a fixed sequence of launches per level on st,
// with no host synchronisation.  A stream capture of it records launches only when the
// workspace already holds n_q reads.
int desc_launch(hvk_ctx* ctx, const DescPlan& pl, const int16_t* d_q_counts,
                const int64_t* d_q_sqnorm, int n_q, int64_t q_ld, float threshold,
                int32_t* d_path, float* d_path_score, int32_t* d_path_len,
                int32_t* d_exit_level, hipStream_t st) {
    if (n_q == 0) return HVK_OK;
    const TreeDev& t = pl.t;
    const int n_cat = pl.n_cat;
    const int B = pl.B, L = pl.L, E = pl.E;
    const int Fmax = pl.Fmax, maxN = pl.maxN, max_span = pl.max_span;
    const int64_t n = n_q;
    const bool dense = pl.dense;
    // workspace: two beam states, scores, grouped pairs, buckets, sort rows
    const size_t beam_bytes = align_up(n * 4) + align_up(n * B * 4) + 2 * align_up(n * B * L * 4);
    const size_t sc_bytes = align_up((size_t)n * B * Fmax * 4);
    const size_t pair_bytes = 2 * align_up((size_t)n * B * 4);
    int copies = kDescCopiesMax;
    while (copies > kDescCopiesMin && int64_t(2) * (n_cat + 1) * copies > kDescMaxCounters)
        copies /= 2;
    const int n_keys = 2 * (n_cat + 1) * copies;  // counter slots: (path, key, copy)
    const int n_scan = (n_keys + kScanTile - 1) / kScanTile;  // scan tiles
    const size_t key_bytes = 2 * align_up((size_t)n_keys * 4) + 256 + align_up((size_t)n_scan * 4);
    const size_t sort_bytes = 2 * align_up((size_t)n * maxN * 4);
    // the int8 rows of the scoring levels: reads (K bytes each) and the
    // prototypes' two digit planes, with their flags
    const int64_t K = ctx->ld;
    const bool use_i8 = !ctx->tune_desc_i16 && L > (dense ? 1 : 0) && K <= 32512;
    const size_t i8_bytes = use_i8 ? align_up((size_t)n * K) + align_up((size_t)n) +
                                         2 * align_up((size_t)n_cat * K) + align_up((size_t)n_cat) +
                                         align_up((size_t)n_cat + 1)
                                   : 0;
    // the roots' digit planes of the dense level 0 (at most 4 digits)
    const size_t root_bytes = dense ? align_up((size_t)4 * kDenseN * K) : 0;
    if (int e = ctx->ws_desc.ensure(2 * beam_bytes + sc_bytes + pair_bytes + key_bytes + sort_bytes +
                                    i8_bytes + root_bytes))
        return e;
    char* w = (char*)ctx->ws_desc.ptr;
    auto take = [&](size_t bytes) { char* p = w; w += align_up(bytes); return p; };
    Beams bm[2];
    for (int s = 0; s < 2; ++s) {
        bm[s].nb = (int32_t*)take(n * 4);
        bm[s].cat = (int32_t*)take(n * B * 4);
        bm[s].path = (int32_t*)take(n * B * L * 4);
        bm[s].pscore = (float*)take(n * B * L * 4);
    }
    float* sc = (float*)take((size_t)n * B * Fmax * 4);
    int32_t* pairs = (int32_t*)take((size_t)n * B * 4);
    int32_t* pkey = (int32_t*)take((size_t)n * B * 4);
    int32_t* hist = (int32_t*)take((size_t)n_keys * 4);
    int32_t* cursor = (int32_t*)take((size_t)n_keys * 4);
    int32_t* total = (int32_t*)take(256);
    int32_t* scan_part = (int32_t*)take((size_t)n_scan * 4);
    float* ws_s = (float*)take((size_t)n * maxN * 4);
    int32_t* ws_c = (int32_t*)take((size_t)n * maxN * 4);
    int8_t* droot = dense ? (int8_t*)take((size_t)4 * kDenseN * K) : nullptr;
    DescI8 i8{};
    if (use_i8) {
        i8.q8 = (int8_t*)take((size_t)n * K);
        i8.qmode = (uint8_t*)take((size_t)n);
        i8.d0 = (int8_t*)take((size_t)n_cat * K);
        i8.d1 = (int8_t*)take((size_t)n_cat * K);
        i8.dflag = (uint8_t*)take((size_t)n_cat);
        i8.cflag = (uint8_t*)take((size_t)n_cat + 1);
    }

    const DescOut out{d_path, d_path_score, d_path_len, d_exit_level};
    const int rb = (int)((n + kDescThreads - 1) / kDescThreads);
    const int pb = (int)((n * B + kDescP * (kDescThreads / 64) - 1) / (kDescP * (kDescThreads / 64)));
    const int pb8 =
        (int)((n * B + kDescP8 * (kDescThreads / 64) - 1) / (kDescP8 * (kDescThreads / 64)));
    // columns [dim, ctx->ld) of every row are zero, and any valid stride is at
    // least ctx->ld: the dots run over the context's ld (K), whatever the strides

    // candidate lists of up to kSelMaxLds entries are sorted in LDS
    constexpr int kSelThreads = 128, kSelMaxLds = 64;
    const bool sel_lds = maxN <= kSelMaxLds;
    const size_t sel_bytes = sel_lds ? 2 * sizeof(float) * (size_t)maxN * kSelThreads : 0;
    // level 0 on the matrix cores when its operands fit the digit split
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL(desc_init, dim3(rb), dim3(kDescThreads), 0, st, bm[0], n_q, B);
    if (use_i8) {
        const int wpb = kDescThreads / 64;
        if (!dense)  // else level 0 writes them
            hipLaunchKernelGGL(desc_q8, dim3((unsigned)((n + wpb - 1) / wpb)), dim3(kDescThreads),
                               0, st, d_q_counts, q_ld, n_q, K, i8.q8, i8.qmode);
        if (n_cat > 0)
            hipLaunchKernelGGL(desc_digits, dim3((unsigned)((n_cat + wpb - 1) / wpb)),
                               dim3(kDescThreads), 0, st, t, K, i8.d0, i8.d1, i8.dflag);
        hipLaunchKernelGGL(desc_cflag, dim3((unsigned)((n_cat + kDescThreads) / kDescThreads)),
                           dim3(kDescThreads), 0, st, t, i8.dflag, i8.cflag);
    }
    // after the scatter, the cursor of the last path-0 bucket holds the
    // number of path-0 pairs: the first pair of desc_score
    const int32_t* n_i8 = cursor + (n_cat + 1) * copies - 1;
    HVK_HIP(hipGetLastError());
    int cur = 0;
    for (int lvl = 0; lvl < L; ++lvl) {
        if (lvl == 0 && dense) {  // every read has one entry, the root
            const int nd = pl.max_abs_proto <= 32639 ? 2 : pl.max_abs_proto <= 8355711 ? 3 : 4;
            auto digits_k = nd == 2 ? desc_root_digits<2>
                            : nd == 3 ? desc_root_digits<3>
                                      : desc_root_digits<4>;
            auto dense_k = nd == 2 ? desc_dense<2> : nd == 3 ? desc_dense<3> : desc_dense<4>;
            const int64_t n8 = (int64_t)kDenseN * K / 8;
            hipLaunchKernelGGL(digits_k, dim3((unsigned)((n8 + 255) / 256)), dim3(256), 0, st, t, K,
                               droot);
            hipLaunchKernelGGL(dense_k, dim3((n_q + 127) / 128), dim3(256), 0, st, t, d_q_counts,
                               q_ld, d_q_sqnorm, n_q, B, K, Fmax, sc, droot, i8.q8, i8.qmode);
        } else {
            hipLaunchKernelGGL(desc_fill, dim3((n_keys + kDescThreads - 1) / kDescThreads),
                               dim3(kDescThreads), 0, st, hist, (int64_t)n_keys, 0);
            hipLaunchKernelGGL(desc_hist, dim3(rb), dim3(kDescThreads), 0, st, bm[cur], n_q, B,
                               n_cat, i8.qmode, i8.cflag, copies, hist);
            hipLaunchKernelGGL(desc_scan_sums, dim3(n_scan), dim3(1024), 0, st, hist, n_keys,
                               scan_part);
            hipLaunchKernelGGL(desc_scan_top, dim3(1), dim3(1024), 0, st, scan_part, n_scan, total);
            hipLaunchKernelGGL(desc_scan_apply, dim3(n_scan), dim3(1024), 0, st, hist, n_keys,
                               scan_part, cursor);
            hipLaunchKernelGGL(desc_scatter, dim3(rb), dim3(kDescThreads), 0, st, bm[cur], n_q, B,
                               n_cat, i8.qmode, i8.cflag, copies, cursor, pairs, pkey);
            // the longest candidate list of this level: the roots at level 0
            const bool narrow = (lvl == 0 ? pl.n_roots : max_span) <= kDescGSmall;
            if (use_i8) {
                auto score8 = [&](auto kernel) {
                    hipLaunchKernelGGL(kernel, dim3(pb8), dim3(kDescThreads), 0, st, t, i8.q8,
                                       d_q_sqnorm, B, K, i8.d0, i8.d1, pairs, pkey, n_i8, Fmax, sc);
                };
                if (narrow) score8(desc_score8<kDescP8, 2, kDescGSmall, 4>);
                else score8(desc_score8<kDescP8, 2, kDescG, 2>);
            }
            auto score = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, dim3(pb), dim3(kDescThreads), 0, st, t, d_q_counts, q_ld,
                                   d_q_sqnorm, B, K, pairs, pkey, n_i8, total, Fmax, sc);
            };
            if (E == 8) {
                if (narrow) score(desc_score<8, kDescGSmall>);
                else score(desc_score<8, kDescG>);
            } else {
                if (narrow) score(desc_score<2, kDescGSmall>);
                else score(desc_score<2, kDescG>);
            }
        }
        if (sel_lds)
            hipLaunchKernelGGL(desc_select, dim3((n_q + kSelThreads - 1) / kSelThreads),
                               dim3(kSelThreads), sel_bytes, st, t, bm[cur], bm[cur ^ 1], n_q, B, L,
                               lvl, threshold, sc, Fmax, ws_s, ws_c, maxN, out, 1);
        else
            hipLaunchKernelGGL(desc_select, dim3(rb), dim3(kDescThreads), 0, st, t, bm[cur],
                               bm[cur ^ 1], n_q, B, L, lvl, threshold, sc, Fmax, ws_s, ws_c, maxN,
                               out, 0);
        HVK_HIP(hipGetLastError());
        cur ^= 1;
    }
    hipLaunchKernelGGL(desc_final, dim3(rb), dim3(kDescThreads), 0, st, bm[cur], n_q, B, L, L, out);
    HVK_HIP(hipGetLastError());
    timing_mark(ctx, 1, true, st);
    return HVK_OK;
}

}  // namespace hvk

extern "C" int hvk_descend(hvk_ctx* ctx, const hvk_tree* tree, const int16_t* d_q_counts,
                           const int64_t* d_q_sqnorm, int n_q, int64_t q_ld, int q_maxabs,
                           int beam, float threshold, int levels, int32_t* d_path,
                           float* d_path_score, int32_t* d_path_len, int32_t* d_exit_level,
                           hvk_stream stream) {
    if (!ctx || !tree || n_q < 0 || beam < 1 || beam > kMaxBeam || levels < 1 ||
        levels > kMaxLevels || q_maxabs < 0)
        return HVK_ERR_ARG;
    if (n_q == 0) return HVK_OK;
    if (!d_q_counts || !d_q_sqnorm || !d_path || !d_path_score || !d_path_len || !d_exit_level)
        return HVK_ERR_ARG;
    if (q_ld == 0) q_ld = ctx->ld;
    if (q_ld < ctx->dim || q_ld % 128) return HVK_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = (hipStream_t)stream;
    DescPlan pl;
    if (int e = desc_prepare(ctx, tree, q_maxabs, beam, levels, pl, st)) return e;
    return desc_launch(ctx, pl, d_q_counts, d_q_sqnorm, n_q, q_ld, threshold, d_path, d_path_score,
                       d_path_len, d_exit_level, st);
}

extern "C" int hvk_tree_place(hvk_ctx* ctx, const int16_t* d_counts, int64_t counts_ld, int n,
                              const int32_t* h_cat, int32_t* d_proto, int64_t ld, int n_cat,
                              hvk_stream stream) {
    if (!ctx || n < 0) return HVK_ERR_ARG;
    if (n == 0) return HVK_OK;
    if (!d_counts || !h_cat || !d_proto) return HVK_ERR_ARG;
    if (counts_ld == 0) counts_ld = ctx->ld;
    if (ld == 0) ld = ctx->ld;
    if (counts_ld < ctx->dim || ld < ctx->dim) return HVK_ERR_ARG;
    for (int i = 0; i < n; ++i)
        if (h_cat[i] < 0 || h_cat[i] >= n_cat) return HVK_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = (hipStream_t)stream;
    std::vector<const int32_t*> dev;
    if (int e = upload_i32(ctx, ctx->ws_desc_tree, {{h_cat, (size_t)n}}, dev, st)) return e;
    hipLaunchKernelGGL(tree_place_kernel, dim3(n), dim3(kDescThreads), 0, st, d_counts, counts_ld,
                       ld, dev[0], d_proto);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

extern "C" int hvk_tree_sum(hvk_ctx* ctx, int32_t* d_proto, int64_t ld, int n_cat,
                            const int32_t* h_child_off, const int32_t* h_child_idx,
                            const int32_t* h_parents, int n_parents, int32_t* d_overflow,
                            hvk_stream stream) {
    if (!ctx || n_cat < 0 || n_parents < 0) return HVK_ERR_ARG;
    if (n_parents == 0) return HVK_OK;
    if (!d_proto || !h_child_off || !h_parents) return HVK_ERR_ARG;
    if (ld == 0) ld = ctx->ld;
    if (ld < ctx->dim || h_child_off[0] != 0) return HVK_ERR_ARG;
    for (int c = 0; c < n_cat; ++c)
        if (h_child_off[c + 1] < h_child_off[c]) return HVK_ERR_ARG;
    const int nnz = h_child_off[n_cat];
    if (nnz > 0 && !h_child_idx) return HVK_ERR_ARG;
    for (int i = 0; i < nnz; ++i)
        if (h_child_idx[i] < 0 || h_child_idx[i] >= n_cat) return HVK_ERR_ARG;
    // the listed parents are written while their children are read: no
    // listed parent may be a child of a listed parent (one rank per call)
    std::vector<char> listed(n_cat, 0);
    for (int i = 0; i < n_parents; ++i) {
        const int p = h_parents[i];
        if (p < 0 || p >= n_cat || listed[p]) return HVK_ERR_ARG;
        listed[p] = 1;
    }
    for (int i = 0; i < n_parents; ++i) {
        const int p = h_parents[i];
        for (int j = h_child_off[p]; j < h_child_off[p + 1]; ++j)
            if (listed[h_child_idx[j]]) return HVK_ERR_ARG;
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = (hipStream_t)stream;
    std::vector<const int32_t*> dev;
    if (int e = upload_i32(ctx, ctx->ws_desc_tree,
                           {{h_child_off, (size_t)n_cat + 1}, {h_child_idx, (size_t)nnz},
                            {h_parents, (size_t)n_parents}},
                           dev, st))
        return e;
    hipLaunchKernelGGL(tree_sum_kernel, dim3(n_parents), dim3(kDescThreads), 0, st, d_proto, ld,
                       dev[0], dev[1], dev[2], d_overflow);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

extern "C" int hvk_tree_limbs(hvk_ctx* ctx, const int32_t* d_proto, int n, int64_t ld,
                              int16_t* d_lo, int16_t* d_hi, int64_t* d_sqnorm,
                              int32_t* d_maxabs_limb, int32_t* d_maxabs_proto,
                              hvk_stream stream) {
    if (!ctx || n < 0) return HVK_ERR_ARG;
    if (n == 0) return HVK_OK;
    if (!d_proto || !d_lo || !d_hi) return HVK_ERR_ARG;
    if (ld == 0) ld = ctx->ld;
    if (ld < ctx->dim) return HVK_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(tree_limbs_kernel, dim3(n), dim3(kDescThreads), 0, st, d_proto, ld, d_lo,
                       d_hi, d_sqnorm, d_maxabs_limb, d_maxabs_proto);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}
