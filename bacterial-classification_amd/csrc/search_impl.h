// search_impl.h -- kernels and launch code of the exact cosine top-k search,
// shared by the translation units search.hip (entry points, norms, re-scan)
// and search_kp{8,16,32}.hip (one candidate-list length each).
//
// Order: (c desc, reference index asc), c = dot / (sqrt(|q|^2) sqrt(|r|^2))
// in FP64 from the exact integer dot (the reference's cosine, src/HV.cpp:
// 219-242, before its 2c-1 map).  Pipeline per call:
//   1. inv_norm_kernel  FP32 inverse norms 1/sqrt(|v|^2), correctly rounded.
//   2. plain_kernel | mfma_kernel
//        MT queries x NT references per tile, K-slabs staged through LDS.
//        Each tile's scores s = dot * iq * ir (FP32) are offered to per-query
//        LDS top lists keyed by (s, -index).  A workgroup sweeps one slice
//        ("split") of the references and writes its per-query top-KP keys.
//   3. finalize_kernel  one wave per query: merges the split lists,
//        recomputes the exact int64 dot of each of the KP survivors from the
//        stored HVs, re-scores them in FP64, sorts, and checks the margin:
//        every excluded reference has s <= s_KP and |s - c| < kMarginEps, so
//        the top-k is exact when c_k > s_KP + kMarginEps.  Queries failing
//        the check are queued for
//   4. rescan_kernel    an exact FP64 brute-force pass over all references.
#pragma once

#include <cmath>
#include <type_traits>
#include <utility>

#include "hvk_internal.h"

namespace hvk {


typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef short v2s __attribute__((ext_vector_type(2)));

// Bound on |s - c| for the FP32 score: below 16 units in the last place of
// 1.0 (2^-24) for the worst limb combination (two limbs x two limbs).
constexpr float kMarginEps = 2e-6f;

__device__ __forceinline__ uint32_t ord_f32(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float f32_of_ord(uint32_t o) {
    const uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(b);
}
__device__ __forceinline__ uint64_t make_key(float s, int n) {
    return ((uint64_t)ord_f32(s) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)n);
}
__device__ __forceinline__ int idx_of_key(uint64_t key) {
    return (int)(0xFFFFFFFFu - (uint32_t)key);
}

__device__ __forceinline__ double cos64(int64_t dot, int64_t nq, int64_t nr) {
    if (nq == 0 || nr == 0) return 0.0;
    return (double)dot / (sqrt((double)nq) * sqrt((double)nr));
}

// FP64 value <-> unsigned key of the same order (no NaN occurs here).
__device__ __forceinline__ uint64_t ord_f64(double c) {
    const uint64_t b = (uint64_t)__double_as_longlong(c);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double f64_of_ord(uint64_t o) {
    const uint64_t b = (o >> 63) ? (o & 0x7FFFFFFFFFFFFFFFull) : ~o;
    return __longlong_as_double((long long)b);
}

template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
// f(integral_constant<int, i>) for i = 0..N-1, expanded at compile time so
// register arrays indexed by i stay in registers.
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// ------------------------------------------------------------------------
// per-workgroup top lists in LDS (keys only; 0 = empty).  Query-minor
// layout ([j][m]): the merges, one thread per query, touch consecutive
// words.  The candidate buffer (CAP keys per query, layout [slot][m]) is
// passed in: the MFMA kernel keeps it in LDS stages idle during its
// epilogue.
// ------------------------------------------------------------------------
template <int MT, int KP>
struct TopLds {
    uint64_t top[KP][MT];  // top[j][m] descending in j
    uint64_t thr[MT];      // top[KP-1][m] (0 while the list is not full)
    float thr_f[MT];       // its FP32 score, -inf while thr is 0
    int cnt[MT];
    int ovf[2];
    int zero;  // always 0; an operand the compiler cannot fold

    __device__ void init(int tid, int nthr) {
        for (int i = tid; i < MT * KP; i += nthr) top[i / MT][i % MT] = 0;
        for (int i = tid; i < MT; i += nthr) { thr[i] = 0; thr_f[i] = -INFINITY; cnt[i] = 0; }
        if (tid == 0) { ovf[0] = 0; ovf[1] = 0; zero = 0; }
    }

    // Offers one (query m, reference) key.  Returns true when it beat the
    // threshold but the buffer was full: the caller keeps it for the next
    // pass, which runs after the merge has raised the threshold.  A key is
    // therefore offered at most once until accepted or fallen below.
    template <int CAP>
    __device__ __forceinline__ bool offer(int m, uint64_t key, int pass, uint64_t* buf) {
        if (key <= thr[m]) return false;
        const int slot = atomicAdd(&cnt[m], 1);
        if (slot < CAP) {
            buf[slot * MT + m] = key;
            return false;
        }
        ovf[pass & 1] = 1;
        return true;
    }

    // Merges query m's buffered keys into its list, held in registers while
    // the candidates are inserted (r[p] = max(t[p], min(t[p-1], x))).
    template <int CAP>
    __device__ void merge(int m, const uint64_t* buf) {
        const int c = cnt[m];
        const int n = c < CAP ? c : CAP;
        if (n > 0) {
            uint64_t t[KP];
#pragma unroll
            for (int j = 0; j < KP; ++j) t[j] = top[j][m];
            for (int q = 0; q < n; ++q) {
                const uint64_t x = buf[q * MT + m];
                if (x <= t[KP - 1]) continue;
#pragma unroll
                for (int p = KP - 1; p > 0; --p) {
                    const uint64_t lo = t[p - 1] < x ? t[p - 1] : x;
                    t[p] = t[p] > lo ? t[p] : lo;
                }
                t[0] = t[0] > x ? t[0] : x;
            }
#pragma unroll
            for (int j = 0; j < KP; ++j) top[j][m] = t[j];
            if (t[KP - 1] > thr[m]) {
                thr[m] = t[KP - 1];
                thr_f[m] = f32_of_ord((uint32_t)(t[KP - 1] >> 32));
            }
        }
        cnt[m] = 0;
    }
};

// End of one offer pass: merge the buffers; true when a key was dropped
// (the caller then runs another pass over its dropped keys).
template <int CAP, int MT, int KP>
__device__ __forceinline__ bool topk_pass_end(TopLds<MT, KP>& tl, const uint64_t* buf, int tid,
                                              int pass) {
    __syncthreads();
    const int ov = tl.ovf[pass & 1];
    if (tid == 0) tl.ovf[(pass + 1) & 1] = 0;
    if (tid < MT) tl.template merge<CAP>(tid, buf);
    __syncthreads();
    return ov != 0;
}

// ------------------------------------------------------------------------
// 1. inverse norms; also clears the re-scan queue counter
// ------------------------------------------------------------------------
struct BoundRows {  // nullable inputs of the bounded form
    const int32_t* q_maxabs;
    const int32_t* q_resid1;
    const int32_t* r_maxabs;
    const int32_t* r_resid1;
};

// bq[m] = (min(max|q|,127) iq, resid1(q) iq), br[n] = (resid1(r) ir, max|r| ir):
// ------------------------------------------------------------------------
// 2a. plain form: VALU int16 dot (v_dot2_i32_i16), 4 queries x 8 references
//     per thread, K-slab of 64 dims; int32 per slab, int64 across slabs.
// ------------------------------------------------------------------------
constexpr int kPlainMT = 32, kPlainNT = 256, kPlainKS = 64;

// MAT: instead of top lists, every pair's exact FP64 cosine, as an ord_f64
// key, is written to mat[m][n] (row stride nr); qn64 / rn64 are the exact
// squared norms.
template <int KP, bool MAT = false>
__global__ void __launch_bounds__(256)
plain_kernel(const int16_t* __restrict__ q, const float* __restrict__ iq, int nq,
             const int16_t* __restrict__ r, const float* __restrict__ ir, int nr, int64_t ld,
             int split_tiles, uint64_t* __restrict__ part, const int64_t* __restrict__ qn64 = nullptr,
             const int64_t* __restrict__ rn64 = nullptr, uint64_t* __restrict__ mat = nullptr) {
    constexpr int MT = kPlainMT, NT = kPlainNT, KS = kPlainKS, KW = KS / 2;
    constexpr int CAP = 16;
    __shared__ uint32_t sQ[KW][MT];
    __shared__ uint32_t sR[KW][NT + 1];
    __shared__ TopLds<MT, KP> tl;
    __shared__ uint64_t cbuf[CAP * MT];
    __shared__ float s_iq[MT];

    const int tid = threadIdx.x;
    const int tx = tid & 31, ty = tid >> 5;
    const int m0 = blockIdx.x * MT;
    const int split = blockIdx.y;
    if constexpr (!MAT) {
        tl.init(tid, 256);
        if (tid < MT) s_iq[tid] = (m0 + tid < nq) ? iq[m0 + tid] : 0.0f;
    }
    const int n_begin = split * split_tiles * NT;
    int n_end = n_begin + split_tiles * NT;
    if (n_end > nr) n_end = nr;
    __syncthreads();

    for (int n0 = n_begin; n0 < n_end; n0 += NT) {
        int64_t acc[4][8];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = 0;

        for (int64_t k0 = 0; k0 < ld; k0 += KS) {
            {   // Q slab: 32 rows x 8 chunks of 16 B
                const int row = tid >> 3, ch = tid & 7;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (m0 + row < nq)
                    v = *reinterpret_cast<const uint4*>(q + (int64_t)(m0 + row) * ld + k0 + ch * 8);
                sQ[ch * 4 + 0][row] = v.x;
                sQ[ch * 4 + 1][row] = v.y;
                sQ[ch * 4 + 2][row] = v.z;
                sQ[ch * 4 + 3][row] = v.w;
            }
#pragma unroll
            for (int it = 0; it < 8; ++it) {  // R slab: 256 rows x 8 chunks
                const int c = it * 256 + tid;
                const int row = c >> 3, ch = c & 7;
                uint4 v = make_uint4(0, 0, 0, 0);
                if (n0 + row < nr)
                    v = *reinterpret_cast<const uint4*>(r + (int64_t)(n0 + row) * ld + k0 + ch * 8);
                sR[ch * 4 + 0][row] = v.x;
                sR[ch * 4 + 1][row] = v.y;
                sR[ch * 4 + 2][row] = v.z;
                sR[ch * 4 + 3][row] = v.w;
            }
            __syncthreads();
            int a32[4][8];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) a32[i][j] = 0;
#pragma unroll 4
            for (int c = 0; c < KW; ++c) {
                const uint4 qv = *reinterpret_cast<const uint4*>(&sQ[c][4 * ty]);
                const uint32_t qa[4] = {qv.x, qv.y, qv.z, qv.w};
                uint32_t rv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) rv[j] = sR[c][tx + 32 * j];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        a32[i][j] = __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, qa[i]),
                                                          __builtin_bit_cast(v2s, rv[j]),
                                                          a32[i][j], false);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] += a32[i][j];
            __syncthreads();
        }

        if constexpr (MAT) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + 4 * ty + i;
                if (m >= nq) continue;
                const int64_t qm = qn64[m];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int n = n0 + tx + 32 * j;
                    if (n < nr) mat[(int64_t)m * nr + n] = ord_f64(cos64(acc[i][j], qm, rn64[n]));
                }
            }
            continue;
        }
        float irv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int n = n0 + tx + 32 * j;
            irv[j] = n < nr ? ir[n] : 0.0f;
        }
        uint32_t dropped = 0;  // bit 8*i+j: pair (i, j) waits for another pass
        for (int pass = 0;; ++pass) {
            static_for<32>([&](auto I) {
                constexpr int i = decltype(I)::value / 8, j = decltype(I)::value % 8;
                constexpr uint32_t bit = 1u << decltype(I)::value;
                const int m = 4 * ty + i;
                const int n = n0 + tx + 32 * j;
                if ((pass == 0 || (dropped & bit)) && m0 + m < nq && n < nr) {
                    const float s = __fmul_rn(__fmul_rn((float)acc[i][j], s_iq[m]), irv[j]);
                    if (tl.template offer<CAP>(m, make_key(s, n), pass, cbuf)) dropped |= bit;
                    else dropped &= ~bit;
                }
            });
            if (!topk_pass_end<CAP>(tl, cbuf, tid, pass)) break;
        }
    }
    if constexpr (!MAT) {
        for (int i = tid; i < MT * KP; i += 256) {
            const int j = i / MT, m = i % MT;
            if (m0 + m < nq) part[((int64_t)split * nq + m0 + m) * KP + j] = tl.top[j][m];
        }
    }
}

// ------------------------------------------------------------------------
// 2b. MFMA form: v_mfma_i32_32x32x32_i8.  Workgroup of 4 waves, one per
//     SIMD, on a (64 WX)-query x (64 WY)-reference tile; wave grid 2 x 2,
//     wave tile (32 WX) x (32 WY), i.e. QL RL WX WY = 16 MFMA blocks of
//     32 x 32 per wave, 256 accumulator registers.  K-slabs of 64 bytes are
//     filled by direct global->LDS loads (global_load_lds_dwordx4) into ST
//     LDS stages, ST - 1 slabs ahead of the multiply; one barrier per slab.
//     Queries are the A operand (rows), references the B operand (columns).
//
//     Tiles are handed to workgroups XCD by XCD: workgroup b runs on XCD
//     b mod 8, and the rx workgroups that one XCD holds at a time form a
//     group of gq query tiles x n_splits reference slices, so that each
//     query tile is read by n_splits workgroups and each reference tile by
//     gq workgroups of the same XCD at about the same time, from its L2.
// ------------------------------------------------------------------------
constexpr int kMfmaKS = 64, kXcds = 8;
constexpr int kLdsBytes = 160 * 1024;
// 8 waves, two per SIMD, in a GM x GN = 2 x 4 grid: while one wave of a
// SIMD waits (slab requests, barrier, fragment reads), the other keeps the
// matrix pipe busy.  128 accumulator registers per wave.
constexpr int kMfmaGM = 2, kMfmaGN = 4, kMfmaWaves = kMfmaGM * kMfmaGN;
constexpr int kMfmaThreads = 64 * kMfmaWaves;

// Tile of a launch.  SH = 0: the default, a 2 x 4 wave grid whose wave tiles
// hold QL * RL * WX * WY = 8 blocks of 32 x 32.  SH > 0 (one limb per side
// only, the bounded form's shapes for the batch autotuner): other wave grids
// GM x GN (8 waves) and wave tiles, at most 8 blocks per wave:
//   SH  GM x GN  WX x WY  queries x references per tile
//    1  2 x 4    2 x 2    128 x 256
//    2  2 x 4    4 x 1    256 x 128
//    3  2 x 4    2 x 4    128 x 512
//    4  4 x 2    4 x 2    512 x 128
//    5  4 x 2    2 x 4    256 x 256
// (Fewer than 128 rows or columns would leave waves without a share of each
// slab's 16-row loads, and the floor groups need GN <= KP / 2.)
constexpr int kMfmaShapes = 6;
template <int QL, int RL, int SH = 0>
struct MfmaTile {
    static_assert(SH == 0 || (QL == 1 && RL == 1), "tile shapes are one-limb only");
    static constexpr int GM = SH == 4 || SH == 5 ? 4 : kMfmaGM;
    static constexpr int GN = SH == 4 || SH == 5 ? 2 : kMfmaGN;
    static constexpr int WX = SH == 0 ? (QL == 1 ? 4 : 2)
                            : (SH == 1 || SH == 3 || SH == 5) ? 2 : 4;
    static constexpr int WY = SH == 0 ? (RL == 1 ? 2 : 1)
                            : SH == 1 ? 2 : SH == 2 ? 1 : SH == 4 ? 2 : 4;
    static_assert(GM * GN == kMfmaWaves, "8 waves");
    static constexpr int MT = GM * 32 * WX, NT = GN * 32 * WY;
};

struct TileMap {
    int q_tiles;      // query tiles of MT rows
    int n_splits;     // reference slices
    int split_tiles;  // reference tiles of NT per slice
    int gq;           // query tiles per XCD group
    int sg;           // reference slices per XCD group (sg <= n_splits)
    int n_sgroups;    // slice groups = ceil(n_splits / sg)
    int rx;           // workgroups per XCD group = gq * sg
};

// 16-byte global -> LDS copy (global_load_lds_dwordx4): lane l of the wave
// writes LDS bytes [lds + 16 l, lds + 16 l + 16); lds must be wave-uniform.
// The builtin exists only for the device pass.
__device__ __forceinline__ void load_lds16(const int8_t* g, int8_t* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
#endif
}

// LDS tile rows are 64 bytes; 16-byte chunk c of row `row` is stored at
// chunk c ^ swz_key(row).  A fragment read (ds_read_b128) takes one chunk
// of 32 consecutive rows.  gfx950's LDS serves a ds_read_b128 16 lanes (256
// bytes, 64 banks) at a time; with this key the 16 rows of each group fall
// on 16 distinct 16-byte bank groups (measured: no bank-conflict cycles).
__device__ __forceinline__ int swz_key(int row) { return (row >> 2) & 3; }
__device__ __forceinline__ int swz(int row, int chunk) {
    return row * 64 + ((chunk ^ swz_key(row)) << 4);
}

template <int QL, int RL, int KP, int ST, int SH = 0>
struct MfmaSmem {
    static constexpr int MT = MfmaTile<QL, RL, SH>::MT, NT = MfmaTile<QL, RL, SH>::NT;
    int8_t q[ST][QL][MT * kMfmaKS];
    int8_t r[ST][RL][NT * kMfmaKS];  // also the candidate buffer of the epilogue
    TopLds<MT, KP> tl;
    float iq[MT];
    float2 bq[MT];
    // The epilogue's per-wave score stash (16 floats per lane) sits in the
    // idle query stages and the candidate buffer in the idle reference
    // stages; when the query stages are too small (tiles of 128 rows) the
    // stash starts at the query stages and runs on, and the buffer follows it.
    static constexpr int kStash = kMfmaWaves * 64 * 16 * (int)sizeof(float);
    static constexpr bool kStashInQ = (int)sizeof(q) >= kStash;
    static constexpr int kBufBytes = kStashInQ ? (int)sizeof(r) : (int)(sizeof(q) + sizeof(r)) - kStash;
    // candidates per query and pass that the buffer holds
    static constexpr int kCap = kBufBytes / (MT * 8) < 32 ? kBufBytes / (MT * 8) : 32;
};

// LDS stages: three when they fit beside the top lists, else two.
template <int QL, int RL, int KP, int SH = 0>
constexpr int mfma_stages() {
    return sizeof(MfmaSmem<QL, RL, KP, 3, SH>) <= (size_t)kLdsBytes ? 3 : 2;
}

// s_waitcnt with vmcnt(n), expcnt and lgkmcnt left free (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// s_waitcnt vmcnt(n) lgkmcnt(0).
template <int N>
__device__ __forceinline__ void wait_vmcnt_lgkm0() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
}

// Lane index recomputed where used (v_mbcnt in a volatile asm), so that the
// compiler neither keeps it nor values derived from it live across loops.
__device__ __forceinline__ int opaque_lane() {
    int ln;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(ln));
    return ln;
}

// Workgroup barrier without a memory fence; the caller has waited for the
// memory operations that must complete before it.  The memory clobber keeps
// the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_barrier" ::: "memory"); }

// BND: one-limb bounded form -- both operands hold clamped counts and each
// score is raised by bq[m].x br[n].x + bq[m].y br[n].y, an upper bound on
// |dot - dot_clamped| * iq * ir.
// IND: query m of the launch is stored row q_rows[m], and the launch covers
// m < *q_count (a count on the device, at most nq).
// MAT: no top lists; the exact dot of every pair (limb products combined in
// int64) gives its FP64 cosine, written as an ord_f64 key to mat[m][n] (row
// stride nr); qn64 / rn64 are the exact squared norms.
template <int QL, int RL, int KP, bool BND = false, bool IND = false, bool MAT = false, int SH = 0>
__global__ void __launch_bounds__(kMfmaThreads, 1)
mfma_kernel(const int8_t* __restrict__ qlo, const int8_t* __restrict__ qhi,
            const float* __restrict__ iq, int nq, const int8_t* __restrict__ rlo,
            const int8_t* __restrict__ rhi, const float* __restrict__ ir, int nr, int64_t ld,
            TileMap map, uint64_t* __restrict__ part, const float2* __restrict__ bq,
            const float2* __restrict__ br, const int32_t* __restrict__ q_rows,
            const int32_t* __restrict__ q_count, uint64_t* __restrict__ qthr,
            const int64_t* __restrict__ qn64 = nullptr, const int64_t* __restrict__ rn64 = nullptr,
            uint64_t* __restrict__ mat = nullptr) {
    static_assert(!(MAT && (IND || BND)), "the key matrix takes direct rows and exact limbs");
    // qthr (nullable, zeroed by the caller): per query, the largest
    // threshold key any slice has reached.  Every slice publishes its own
    // after each tile and adopts the largest: a slice's KP-th best key is a
    // lower bound on the KP-th best of the merged lists, so the keys it
    // rejects could not have entered the merged top KP.
    using T = MfmaTile<QL, RL, SH>;
    constexpr int WX = T::WX, WY = T::WY, MT = T::MT, NT = T::NT, KS = kMfmaKS;
    constexpr int GM = T::GM, GN = T::GN;
    constexpr int ST = mfma_stages<QL, RL, KP, SH>();
    constexpr int NTH = kMfmaThreads;
    constexpr int QJ = MT * KS / 1024 / kMfmaWaves;  // wave-loads per wave per Q limb
    constexpr int RJ = NT * KS / 1024 / kMfmaWaves;  // per R limb
    static_assert(QJ >= 1 && RJ >= 1, "every wave requests a share of each operand");
    constexpr int LPS = QL * QJ + RL * RJ;  // wave-loads per wave per slab
    __shared__ __attribute__((aligned(16))) MfmaSmem<QL, RL, KP, ST, SH> sm;
    auto& tl = sm.tl;

    // XCD-grouped tile of this workgroup (all branches below are uniform)
    const int b = blockIdx.x;
    const int slot = b / kXcds, gi = slot / map.rx, w = slot - gi * map.rx;
    const int g = gi * kXcds + b % kXcds;  // group: gq query tiles x sg slices
    const int qg = g / map.n_sgroups, sgi = g - qg * map.n_sgroups;
    const int qt = qg * map.gq + w / map.sg;
    const int split = sgi * map.sg + w % map.sg;
    if (qt >= map.q_tiles || split >= map.n_splits) return;
    const int m0 = qt * MT;
    if constexpr (IND) {
        const int c = *q_count;
        nq = c < nq ? c : nq;
    }
    if (m0 >= nq) return;  // the whole workgroup leaves before any barrier

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int wm = wave % GM, wn = wave / GM;
    auto qrow = [&](int m) { return IND ? q_rows[m] : m; };
    if constexpr (!MAT) {
        tl.init(tid, NTH);
        for (int i = tid; i < MT; i += NTH) {
            const bool in = m0 + i < nq;
            sm.iq[i] = in ? iq[qrow(m0 + i)] : 0.0f;
            sm.bq[i] = (BND && in) ? bq[qrow(m0 + i)] : make_float2(0.0f, 0.0f);
        }
    }
    const int n_begin = split * map.split_tiles * NT;
    int n_end = n_begin + map.split_tiles * NT;
    if (n_end > nr) n_end = nr;
    const int8_t* qsrc[2] = {qlo, qhi};
    const int8_t* rsrc[2] = {rlo, rhi};
    const int nk = (int)(ld / KS);

    // Direct global->LDS loads: one wave instruction fills 1 KB of LDS (16
    // rows of 64 B), lane l writing bytes [16 l, 16 l + 16).  Lane l of
    // wave-load j therefore fetches row 16 j + l / 4, logical chunk
    // (l & 3) ^ swz_key(row), so that LDS holds the layout swz().  Addresses
    // are a wave-uniform tile base plus a 32-bit lane offset (the rows of a
    // tile span less than 2^32 bytes: the host checks 256 ld < 2^32); with
    // IND the query rows are gathered through 64-bit lane pointers.  Rows
    // past the end read a valid row; their results are never offered.
    const int8_t* gq[IND ? QL : 1][QJ];
    if constexpr (IND) {
#pragma unroll
        for (int t = 0; t < QJ; ++t) {
            const int row = 16 * (wave * QJ + t) + (lane >> 2);
            const int off = 16 * ((lane & 3) ^ swz_key(row));
            const int rr = min(m0 + row, nq - 1);
#pragma unroll
            for (int l = 0; l < QL; ++l) gq[l][t] = qsrc[l] + (int64_t)qrow(rr) * ld + off;
        }
    }
    // 32-bit offset of the lane's chunk from the tile base (row0 = first row
    // of the tile, rows clamped to < n_rows); recomputed for each slab from
    // an opaque lane index rather than kept in registers
    auto tile_off = [&](int ln, int t, int per_wave, int row0, int n_rows) -> uint32_t {
        const int row = 16 * (wave * per_wave + t) + (ln >> 2);
        const int off = 16 * ((ln & 3) ^ swz_key(row));
        return (uint32_t)((uint32_t)(min(row0 + row, n_rows - 1) - row0) * (uint32_t)ld + off);
    };
    __syncthreads();

    for (int n0 = n_begin; n0 < n_end; n0 += NT) {
        auto issue_slab = [&](int kk, int buf, int ln) {
            const int64_t k0 = (int64_t)kk * KS;
#pragma unroll
            for (int t = 0; t < QJ; ++t) {
                const uint32_t vo = IND ? 0u : tile_off(ln, t, QJ, m0, nq);
#pragma unroll
                for (int l = 0; l < QL; ++l) {
                    const int8_t* src;
                    if constexpr (IND) src = gq[l][t] + k0;
                    else src = qsrc[l] + (int64_t)m0 * ld + k0 + vo;
                    load_lds16(src, &sm.q[buf][l][1024 * (wave * QJ + t)]);
                }
            }
#pragma unroll
            for (int t = 0; t < RJ; ++t) {
                const uint32_t vo = tile_off(ln, t, RJ, n0, nr);
#pragma unroll
                for (int l = 0; l < RL; ++l)
                    load_lds16(rsrc[l] + (int64_t)n0 * ld + k0 + vo,
                               &sm.r[buf][l][1024 * (wave * RJ + t)]);
            }
        };

        v16i acc[QL][RL][WX][WY];
        static_for<QL * RL * WX * WY>([&](auto I) {
            constexpr int i = decltype(I)::value;
            acc[i / (RL * WX * WY)][(i / (WX * WY)) % RL][(i / WY) % WX][i % WY] = v16i{};
        });

        // Fragments of K-step ks (32 bytes) of the slab in stage st: one
        // lane base per operand, block and limb offsets as immediates.
        // Lane (r, h) = (lane % 32, lane / 32) reads 16 bytes of row r of a
        // 32-row block, logical chunk 2 ks + h.
        auto load_frags = [&](v4i (&fa)[QL][WX], v4i (&fb)[RL][WY], int st, int ks, int ln) {
            const int r = ln & 31;
            const int c = r * 64 + (((2 * ks + (ln >> 5)) ^ swz_key(r)) << 4);
            const int8_t* pa = &sm.q[st][0][0] + wm * 32 * WX * 64 + c;
            const int8_t* pb = &sm.r[st][0][0] + wn * 32 * WY * 64 + c;
#pragma unroll
            for (int l = 0; l < QL; ++l)
#pragma unroll
                for (int x = 0; x < WX; ++x)
                    fa[l][x] = *reinterpret_cast<const v4i*>(pa + l * MT * KS + x * 32 * 64);
#pragma unroll
            for (int l = 0; l < RL; ++l)
#pragma unroll
                for (int y = 0; y < WY; ++y)
                    fb[l][y] = *reinterpret_cast<const v4i*>(pb + l * NT * KS + y * 32 * 64);
        };
        auto mfma_all = [&](const v4i (&fa)[QL][WX], const v4i (&fb)[RL][WY]) {
            static_for<QL * RL * WX * WY>([&](auto I) {
                constexpr int i = decltype(I)::value;
                constexpr int qa = i / (RL * WX * WY), rb = (i / (WX * WY)) % RL;
                constexpr int x = (i / WY) % WX, y = i % WY;
                acc[qa][rb][x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                    fa[qa][x], fb[rb][y], acc[qa][rb][x][y], 0, 0, 0);
            });
        };
        // Single items of the above, for explicit interleaving: MFMA i, slab
        // request i (queries first), fragment read i (A blocks first).
        auto mfma_one = [&](auto I, const v4i (&fa)[QL][WX], const v4i (&fb)[RL][WY]) {
            constexpr int i = decltype(I)::value;
            constexpr int qa = i / (RL * WX * WY), rb = (i / (WX * WY)) % RL;
            constexpr int x = (i / WY) % WX, y = i % WY;
            acc[qa][rb][x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                fa[qa][x], fb[rb][y], acc[qa][rb][x][y], 0, 0, 0);
        };
        auto request_one = [&](auto I, int kk, int buf, int ln) {
            constexpr int i = decltype(I)::value;
            const int64_t k0 = (int64_t)kk * KS;
            if constexpr (i < QL * QJ) {
                constexpr int l = i / QJ, t = i % QJ;
                const int8_t* src;
                if constexpr (IND) src = gq[l][t] + k0;
                else src = qsrc[l] + (int64_t)m0 * ld + k0 + tile_off(ln, t, QJ, m0, nq);
                load_lds16(src, &sm.q[buf][l][1024 * (wave * QJ + t)]);
            } else {
                constexpr int l = (i - QL * QJ) / RJ, t = (i - QL * QJ) % RJ;
                load_lds16(rsrc[l] + (int64_t)n0 * ld + k0 + tile_off(ln, t, RJ, n0, nr),
                           &sm.r[buf][l][1024 * (wave * RJ + t)]);
            }
        };
        auto frag_one = [&](auto I, v4i (&fa)[QL][WX], v4i (&fb)[RL][WY], int st, int ks, int ln) {
            constexpr int i = decltype(I)::value;
            const int r = ln & 31;
            const int c = r * 64 + (((2 * ks + (ln >> 5)) ^ swz_key(r)) << 4);
            if constexpr (i < QL * WX) {
                constexpr int l = i / WX, x = i % WX;
                fa[l][x] = *reinterpret_cast<const v4i*>(&sm.q[st][0][0] + wm * 32 * WX * 64 + c +
                                                         l * MT * KS + x * 32 * 64);
            } else {
                constexpr int l = (i - QL * WX) / WY, y = (i - QL * WX) % WY;
                fb[l][y] = *reinterpret_cast<const v4i*>(&sm.r[st][0][0] + wn * 32 * WY * 64 + c +
                                                         l * NT * KS + y * 32 * 64);
            }
        };

        // Pipeline.  Slab kk + ST - 1 is requested while slab kk is
        // multiplied; the barrier sits between the two K-steps of a slab, so
        // that each batch of fragment reads is issued one batch of 16 MFMAs
        // before its use.  The barrier is a bare s_barrier: a fenced
        // __syncthreads would make the compiler wait for every load in
        // flight, the prefetched slabs included.  Each wave waits for its own
        // loads of the next slab (vmcnt) and its own fragment reads
        // (lgkmcnt) before the barrier; after it the next slab is complete
        // in LDS and no wave still reads the stage that the next request
        // overwrites.  (ld >= 128, so nk >= 2 >= ST - 1.)
        static_assert(KS / 32 == 2, "two K-steps per slab");
        // ln: the lane index, recomputed at the end of each iteration (an
        // opaque asm is a scheduling boundary, so none sits inside a half)
        int ln = opaque_lane();
#pragma unroll
        for (int s = 0; s < ST - 1; ++s) issue_slab(s, s, ln);
        wait_vmcnt<(ST - 2) * LPS>();
        lds_barrier();
        v4i fa0[QL][WX], fb0[RL][WY], fa1[QL][WX], fb1[RL][WY];
        load_frags(fa0, fb0, 0, 0, ln);
        int cur = 0, nxt = ST - 1;  // stage of slab kk, stage for slab kk + ST - 1
        // Issue order inside each half: the slab requests and fragment reads
        // are spread over the MFMA stream in program order (each group
        // fenced by sched_barrier), so that with one wave per SIMD the
        // matrix pipe is not left idle while they issue.
        constexpr int NMF = QL * RL * WX * WY;  // MFMAs per K-step
        constexpr int NFR = QL * WX + RL * WY;  // fragment reads per K-step
        constexpr int NSL = NMF / 2;            // slots of two MFMAs
        // (tiles with more requests or fragment reads than that give the extra
        // ones to the slots again, in turn)
        static_assert(NMF % 2 == 0 && LPS <= 3 * NSL && NFR <= 3 * NSL, "interleave pattern");
        for (int kk = 0; kk + 1 < nk; ++kk) {
            // past the last slab the request repeats slab nk - 1 into the
            // stage of slab kk - 1, which no later read uses: the loop body
            // stays branch-free
            const int kreq = kk + ST - 1 < nk ? kk + ST - 1 : nk - 1;
            // first half: slab requests spread over the slots, fragment
            // reads front-loaded (they have landed by the wait before the
            // barrier)
            static_for<NSL>([&](auto I) {
                constexpr int i = decltype(I)::value;
                mfma_one(std::integral_constant<int, 2 * i>{}, fa0, fb0);
                if constexpr (i < LPS) request_one(I, kreq, nxt, ln);
                if constexpr (i + NSL < LPS)
                    request_one(std::integral_constant<int, i + NSL>{}, kreq, nxt, ln);
                if constexpr (i + 2 * NSL < LPS)
                    request_one(std::integral_constant<int, i + 2 * NSL>{}, kreq, nxt, ln);
                if constexpr (2 * i < NFR)
                    frag_one(std::integral_constant<int, 2 * i>{}, fa1, fb1, cur, 1, ln);
                if constexpr (2 * i + 1 < NFR)
                    frag_one(std::integral_constant<int, 2 * i + 1>{}, fa1, fb1, cur, 1, ln);
                if constexpr (2 * NSL + i < NFR)
                    frag_one(std::integral_constant<int, 2 * NSL + i>{}, fa1, fb1, cur, 1, ln);
                mfma_one(std::integral_constant<int, 2 * i + 1>{}, fa0, fb0);
                __builtin_amdgcn_sched_barrier(0);
            });
            wait_vmcnt_lgkm0<(ST - 2) * LPS>();  // slab kk+1 landed
            lds_barrier();
            cur = cur + 1 == ST ? 0 : cur + 1;
            nxt = nxt + 1 == ST ? 0 : nxt + 1;
            // second half: next slab's first-step fragments spread over it
            static_for<NSL>([&](auto I) {
                constexpr int i = decltype(I)::value;
                if constexpr (i < NFR) frag_one(I, fa0, fb0, cur, 0, ln);
                mfma_one(std::integral_constant<int, 2 * i>{}, fa1, fb1);
                if constexpr (i + NSL < NFR)
                    frag_one(std::integral_constant<int, i + NSL>{}, fa0, fb0, cur, 0, ln);
                if constexpr (i + 2 * NSL < NFR)
                    frag_one(std::integral_constant<int, i + 2 * NSL>{}, fa0, fb0, cur, 0, ln);
                mfma_one(std::integral_constant<int, 2 * i + 1>{}, fa1, fb1);
                __builtin_amdgcn_sched_barrier(0);
            });
            ln = opaque_lane();
        }
        // last slab: already complete in LDS
        load_frags(fa1, fb1, cur, 1, ln);
        mfma_all(fa0, fb0);
        mfma_all(fa1, fb1);
        // the repeated requests land before the stages turn into the
        // epilogue's buffers (and the next tile's stages)
        wait_vmcnt<0>();
        __syncthreads();

        if constexpr (MAT) {
            // pair e of block (x, y): query row mb(x) + 8(e/4) + e%4, reference
            // column n0 + wn*32*WY + y*32 + lane%32 (the 32x32 output layout);
            // the 32 lanes of a row write 256 consecutive bytes
            const int lane_m = opaque_lane();
            static_for<WY>([&](auto Y) {
                constexpr int y = decltype(Y)::value;
                const int n = n0 + wn * 32 * WY + y * 32 + (lane_m & 31);
                if (n < nr) {
                    const int64_t rnn = rn64[n];
                    static_for<WX>([&](auto X) {
                        constexpr int x = decltype(X)::value;
                        const int mb = wm * 32 * WX + x * 32 + 4 * (lane_m >> 5);
                        static_for<16>([&](auto E) {
                            constexpr int e = decltype(E)::value;
                            const int m = m0 + mb + 8 * (e >> 2) + (e & 3);
                            if (m < nq) {
                                int64_t d = 0;
                                static_for<QL * RL>([&](auto J) {
                                    constexpr int qa = decltype(J)::value / RL;
                                    constexpr int rb = decltype(J)::value % RL;
                                    d += (int64_t)acc[qa][rb][x][y][e] << (8 * (qa + rb));
                                });
                                mat[(int64_t)m * nr + n] = ord_f64(cos64(d, qn64[m], rnn));
                            }
                        });
                    });
                }
            });
            continue;
        }

        // The epilogue derives its addresses from an opaque copy of the lane
        // index, so that none of them is computed ahead and kept live
        // through the K loop.
        const int lane_e = opaque_lane();
        float irv[WY];
        float2 brv[WY];
#pragma unroll
        for (int y = 0; y < WY; ++y) {
            const int n = n0 + wn * 32 * WY + y * 32 + (lane_e & 31);
            irv[y] = n < nr ? ir[n] : 0.0f;
            brv[y] = (BND && n < nr) ? br[n] : make_float2(0.0f, 0.0f);
        }
        // Epilogue.  Pair e of block (x, y) is query row
        // mb(x) + 8(e/4) + e%4, mb(x) = wm*32*WX + x*32 + 4(lane/32), and
        // reference column n0 + wn*32*WY + y*32 + lane%32 (the 32x32 MFMA
        // output layout).  A float prefilter against the query's current
        // FP32 threshold rejects almost every pair; a lane with survivors in
        // a block stages that block's 16 scores in LDS (the idle query
        // stages) and offers the survivors one by one.  The blocks of one x are
        // offered together and merged (topk_pass_end); pairs whose offer
        // found the buffer full are offered again in the next pass.
        // FP32 score (upper bound with BND) of pair e of block (x, y)
        auto pair_score = [&](auto X, auto Y, auto E, float iqh, float2 bqh) -> float {
            constexpr int x = decltype(X)::value, y = decltype(Y)::value, e = decltype(E)::value;
            float v = 0.0f;  // sum_{qa,rb} acc * 256^(qa+rb) in FP32
            static_for<QL * RL>([&](auto J) {
                constexpr int qa = decltype(J)::value / RL, rb = decltype(J)::value % RL;
                constexpr float scale = (float)(1u << (8 * (qa + rb)));
                v = __fmaf_rn((float)acc[qa][rb][x][y][e], scale, v);
            });
            float sc = __fmul_rn(__fmul_rn(v, iqh), irv[y]);
            if constexpr (BND) sc = __fmaf_rn(bqh.y, brv[y].y, __fmaf_rn(bqh.x, brv[y].x, sc));
            return sc;
        };
        // First tile of the slice: the lists are empty, and without a
        // threshold nearly every pair would be offered.  Raise each query's
        // threshold to a floor first: the tile's NT references of a query
        // (the kMfmaGN waves of its row group, 32 lanes x WY each) are split
        // into KP / 2 disjoint groups of GL lanes x WY references; each
        // group's two best scores belong to two distinct references, so the
        // minimum over the groups of each group's second-best score is
        // reached by KP distinct references of the tile: it does not exceed
        // the query's KP-th best key, and keys below it could not enter the
        // top KP.  (Groups of 32 references at KP = 16 leave ~33 of 256
        // references per query above the floor on random scores; the maxima
        // of 16 groups of 16 leave ~48.)
        if (n0 == n_begin) {
            // (the reference stages hold the floors; the offers below reuse them)
            uint32_t* floor_o = reinterpret_cast<uint32_t*>(&sm.r[0][0][0]);
            constexpr int GL = 2 * 32 * GN / KP;  // lanes per group
            static_assert(KP % (2 * GN) == 0 && GL >= 1 && GL <= 32, "floor groups");
            for (int i = tid; i < MT; i += NTH) floor_o[i] = 0xFFFFFFFFu;
            __syncthreads();
            static_for<WX>([&](auto X) {
                constexpr int x = decltype(X)::value;
                const int mb = wm * 32 * WX + x * 32 + 4 * (lane_e >> 5);
                static_for<16>([&](auto E) {
                    constexpr int e = decltype(E)::value;
                    const int row = mb + 8 * (e >> 2) + (e & 3);
                    const float iqh = sm.iq[row];
                    const float2 bqh = sm.bq[row];
                    float a = -INFINITY, v = -INFINITY;  // best, second best
                    static_for<WY>([&](auto Y) {
                        constexpr int y = decltype(Y)::value;
                        const int n = n0 + wn * 32 * WY + y * 32 + (lane_e & 31);
                        const float sc = n < nr ? pair_score(X, Y, E, iqh, bqh) : -INFINITY;
                        v = fmaxf(v, fminf(a, sc));
                        a = fmaxf(a, sc);
                    });
                    for (int o = 1; o < GL; o <<= 1) {
                        const float a2 = __shfl_xor(a, o), v2 = __shfl_xor(v, o);
                        v = fmaxf(fminf(a, a2), fmaxf(v, v2));
                        a = fmaxf(a, a2);
                    }
                    for (int o = GL; o < 32; o <<= 1) v = fminf(v, __shfl_xor(v, o));
                    if ((lane_e & 31) == 0) atomicMin(&floor_o[row], ord_f32(v));
                });
            });
            __syncthreads();
            for (int i = tid; i < MT; i += NTH) {
                const uint64_t key = (uint64_t)floor_o[i] << 32;
                if (key > tl.thr[i]) {
                    tl.thr[i] = key;
                    tl.thr_f[i] = f32_of_ord(floor_o[i]);
                }
            }
            __syncthreads();
        }
        using Smem = MfmaSmem<QL, RL, KP, ST, SH>;
        static_assert(Smem::kCap >= 1 && sizeof(sm.q) + sizeof(sm.r) >= (size_t)Smem::kStash,
                      "score stash and candidate buffer");
        float* stage = reinterpret_cast<float*>(&sm.q[0][0][0]) + wave * 1024 + lane_e;
        constexpr int CAP = Smem::kCap;
        uint64_t* cbuf = Smem::kStashInQ
                             ? reinterpret_cast<uint64_t*>(&sm.r[0][0][0])
                             : reinterpret_cast<uint64_t*>(&sm.q[0][0][0] + Smem::kStash);
        static_for<WX>([&](auto X) {
            constexpr int x = decltype(X)::value;
            const int mb = wm * 32 * WX + x * 32 + 4 * (lane_e >> 5);
            uint32_t dropped[WY];  // bit e of block (x, y): pair e awaits another pass
#pragma unroll
            for (int y = 0; y < WY; ++y) dropped[y] = 0;
            for (int pass = 0;; ++pass) {
                static_for<WY>([&](auto Y) {
                    constexpr int y = decltype(Y)::value;
                    const int n = n0 + wn * 32 * WY + y * 32 + (lane_e & 31);
                    float sv[16];
                    uint32_t hm = 0;  // bit e: pair e reaches its query's threshold
                    static_for<4>([&](auto G) {
                        constexpr int gg = 3 - decltype(G)::value;  // rows mb + 8 gg + 0..3
                        const float4 i4 = *reinterpret_cast<const float4*>(&sm.iq[mb + 8 * gg]);
                        const float4 t4 = *reinterpret_cast<const float4*>(&tl.thr_f[mb + 8 * gg]);
                        const float iqv[4] = {i4.x, i4.y, i4.z, i4.w};
                        const float thv[4] = {t4.x, t4.y, t4.z, t4.w};
                        float2 bqv[4];
                        if constexpr (BND) {
                            const float4 b01 = *reinterpret_cast<const float4*>(&sm.bq[mb + 8 * gg]);
                            const float4 b23 = *reinterpret_cast<const float4*>(&sm.bq[mb + 8 * gg + 2]);
                            bqv[0] = make_float2(b01.x, b01.y); bqv[1] = make_float2(b01.z, b01.w);
                            bqv[2] = make_float2(b23.x, b23.y); bqv[3] = make_float2(b23.z, b23.w);
                        }
                        static_for<4>([&](auto H) {
                            constexpr int h = 3 - decltype(H)::value, e = 4 * gg + h;
                            sv[e] = pair_score(X, Y, std::integral_constant<int, e>{}, iqv[h],
                                               BND ? bqv[h] : make_float2(0.0f, 0.0f));
                            hm = (hm << 1) | (sv[e] >= thv[h] ? 1u : 0u);
                        });
                    });
                    if (pass > 0) hm &= dropped[y];
                    if (n >= nr) hm = 0;
                    dropped[y] = 0;
                    if (hm) {
                        static_for<16>([&](auto E) {
                            constexpr int e = decltype(E)::value;
                            stage[64 * e] = sv[e];
                        });
                        uint32_t ov = 0;
                        do {
                            const int e = __builtin_ctz(hm);
                            hm &= hm - 1;
                            const int m = mb + 8 * (e >> 2) + (e & 3);
                            if (m0 + m < nq &&
                                tl.template offer<CAP>(m, make_key(stage[64 * e], n), pass, cbuf))
                                ov |= 1u << e;
                        } while (hm);
                        dropped[y] = ov;
                    }
                });
                if (!topk_pass_end<CAP>(tl, cbuf, tid, pass)) break;
            }
        });
        if (qthr && tid < MT && m0 + tid < nq) {
            const uint64_t mine = tl.thr[tid];
            const uint64_t g = atomicMax((unsigned long long*)&qthr[m0 + tid],
                                         (unsigned long long)mine);
            if (g > mine) {
                tl.thr[tid] = g;
                tl.thr_f[tid] = f32_of_ord((uint32_t)(g >> 32));
            }
        }
    }
    if constexpr (!MAT) {
        for (int i = tid; i < MT * KP; i += NTH) {
            const int j = i / MT, m = i % MT;
            if (m0 + m < nq) part[((int64_t)split * nq + m0 + m) * KP + j] = tl.top[j][m];
        }
    }
}

// ------------------------------------------------------------------------
// Exact int64 dot of two stored rows by one wave: int16 counts (plain form)
// or int8 limbs (value = lo + 256 hi; hi rows may be absent).  Every lane
// returns the full dot.
// ------------------------------------------------------------------------
__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int lo = __shfl_xor((int)(uint32_t)v, o, 64);
        const int hi = __shfl_xor((int)(v >> 32), o, 64);
        v += (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
    return v;
}

struct RowsI16 {
    const int16_t* q;
    const int16_t* r;
    int64_t ld;
    __device__ int64_t dot(int lane, int64_t qrow, int64_t rrow) const {
        const uint4* a = reinterpret_cast<const uint4*>(q + qrow * ld);
        const uint4* b = reinterpret_cast<const uint4*>(r + rrow * ld);
        int64_t s = 0;
        for (int64_t c = lane; c < ld / 8; c += 64) {
            const uint4 x = a[c], y = b[c];
            int t = 0;
            t = __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, x.x), __builtin_bit_cast(v2s, y.x), t, false);
            t = __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, x.y), __builtin_bit_cast(v2s, y.y), t, false);
            t = __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, x.z), __builtin_bit_cast(v2s, y.z), t, false);
            t = __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, x.w), __builtin_bit_cast(v2s, y.w), t, false);
            s += t;
        }
        return wave_sum64(s);
    }
};

__device__ __forceinline__ int dot4x4(const uint4& x, const uint4& y) {
    int t = __builtin_amdgcn_sdot4((int)x.x, (int)y.x, 0, false);
    t = __builtin_amdgcn_sdot4((int)x.y, (int)y.y, t, false);
    t = __builtin_amdgcn_sdot4((int)x.z, (int)y.z, t, false);
    return __builtin_amdgcn_sdot4((int)x.w, (int)y.w, t, false);
}

struct RowsLimb {
    const int8_t *qlo, *qhi, *rlo, *rhi;
    int64_t ld;
    __device__ int64_t dot(int lane, int64_t qrow, int64_t rrow) const {
        int64_t s = 0;
        for (int64_t c = lane; c < ld / 16; c += 64) {
            const int64_t oq = qrow * ld + 16 * c, orr = rrow * ld + 16 * c;
            const uint4 ql = *reinterpret_cast<const uint4*>(qlo + oq);
            const uint4 rl = *reinterpret_cast<const uint4*>(rlo + orr);
            int64_t t = dot4x4(ql, rl);
            if (rhi) t += (int64_t)dot4x4(ql, *reinterpret_cast<const uint4*>(rhi + orr)) << 8;
            if (qhi) {
                const uint4 qh = *reinterpret_cast<const uint4*>(qhi + oq);
                t += (int64_t)dot4x4(qh, rl) << 8;
                if (rhi) t += (int64_t)dot4x4(qh, *reinterpret_cast<const uint4*>(rhi + orr)) << 16;
            }
            s += t;
        }
        return wave_sum64(s);
    }
};

// ------------------------------------------------------------------------
// 3. finalize: one wave per query
// ------------------------------------------------------------------------
template <int KP, typename Rows>
__global__ void __launch_bounds__(256)
finalize_kernel(Rows rows, const uint64_t* __restrict__ part, int n_splits, int nq,
                const int64_t* __restrict__ qn, const int64_t* __restrict__ rn, int nr, int k,
                int32_t* __restrict__ out_idx, float* __restrict__ out_score,
                double* __restrict__ out_cos, int32_t* __restrict__ flag_list,
                int32_t* __restrict__ flag_count, const int32_t* __restrict__ q_rows,
                const int32_t* __restrict__ q_count, int part_rows,
                const int32_t* __restrict__ order) {
    // q_rows (nullable): query m is stored row q_rows[m], m < *q_count; the
    // partial lists hold part_rows queries, later ones are forwarded to the
    // flag list unexamined.  order (nullable): wave i takes query order[i]
    // (queries grouped by their best candidate, so neighbouring waves re-read
    // the same reference rows from L2).
    __shared__ uint64_t s_key[4][KP];
    __shared__ double s_c[4][KP];
    __shared__ int s_i[4][KP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wv;
    int n_all = nq;
    if (q_rows) n_all = *q_count < nq ? *q_count : nq;
    if (i >= n_all) return;  // whole wave exits together
    const int m = order ? order[i] : i;
    const int row = q_rows ? q_rows[m] : m;
    if (m >= part_rows) {
        if (lane == 0) flag_list[atomicAdd(flag_count, 1)] = row;
        return;
    }
    const int np = n_all < part_rows ? n_all : part_rows;  // queries with partial lists
    uint64_t* key = s_key[wv];
    if (lane == 0) {
        for (int j = 0; j < KP; ++j) key[j] = part[(int64_t)m * KP + j];
        for (int s = 1; s < n_splits; ++s)
            for (int j = 0; j < KP; ++j) {
                const uint64_t c = part[((int64_t)s * np + m) * KP + j];
                if (c <= key[KP - 1]) break;  // each list is sorted
                int p = KP - 1;
                while (p > 0 && key[p - 1] < c) { key[p] = key[p - 1]; --p; }
                key[p] = c;
            }
    }
    __builtin_amdgcn_wave_barrier();
    int nv = 0;
    while (nv < KP && key[nv] != 0) ++nv;
    const int64_t qnm = qn[row];
    // Exact scores in list order (upper-bound keys, descending), kept
    // sorted (c desc, index asc).  Once k are scored, a candidate whose key
    // score plus the FP32 margin is below the k-th exact score cannot enter
    // the top k, nor can any after it: the rest of the list is skipped.
    double* cs = s_c[wv];
    int* ix = s_i[wv];
    int nc = 0;  // candidates scored
    for (int j = 0; j < nv; ++j) {
        if (j >= k) {  // every lane reads the same LDS values: uniform
            const float u = f32_of_ord((uint32_t)(key[j] >> 32));
            if ((double)u + (double)kMarginEps < cs[k - 1]) break;
        }
        const int idx = idx_of_key(key[j]);
        const int64_t d = rows.dot(lane, row, idx);
        if (lane == 0) {
            const double c = cos64(d, qnm, rn[idx]);
            int p = j;
            while (p > 0 && (cs[p - 1] < c || (cs[p - 1] == c && ix[p - 1] > idx))) {
                cs[p] = cs[p - 1];
                ix[p] = ix[p - 1];
                --p;
            }
            cs[p] = c;
            ix[p] = idx;
        }
        __builtin_amdgcn_wave_barrier();
        nc = j + 1;
    }
    if (lane != 0) return;
    bool flag = false;
    if (nv == KP && nr > KP) {
        const float s_kp = f32_of_ord((uint32_t)(key[KP - 1] >> 32));
        flag = !(cs[k - 1] > (double)s_kp + (double)kMarginEps);
    }
    for (int j = 0; j < k; ++j) {
        const int64_t o = (int64_t)row * k + j;
        if (j < nc) {
            out_idx[o] = ix[j];
            out_score[o] = (float)cs[j];
            if (out_cos) out_cos[o] = cs[j];
        } else {
            out_idx[o] = -1;
            out_score[o] = -INFINITY;
            if (out_cos) out_cos[o] = -INFINITY;
        }
    }
    if (flag) flag_list[atomicAdd(flag_count, 1)] = row;
}

// ------------------------------------------------------------------------
// 4. exact re-scan of queued queries: FP64 brute force over all references,
//    one workgroup (4 waves) per queued query.
// ------------------------------------------------------------------------
template <typename Rows>
__global__ void __launch_bounds__(256)
rescan_kernel(Rows rows, const int64_t* __restrict__ qn, const int64_t* __restrict__ rn, int nr,
              int k, const int32_t* __restrict__ flag_list, const int32_t* __restrict__ flag_count,
              int32_t* __restrict__ out_idx, float* __restrict__ out_score,
              double* __restrict__ out_cos) {
    __shared__ double s_c[4][kMaxTopK];
    __shared__ int s_i[4][kMaxTopK];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nflag = *flag_count;
    for (int f = blockIdx.x; f < nflag; f += gridDim.x) {
        const int m = flag_list[f];
        const int64_t qnm = qn[m];
        double* bc = s_c[wv];
        int* bi = s_i[wv];
        if (lane == 0)
            for (int j = 0; j < k; ++j) { bc[j] = -INFINITY; bi[j] = 0x7FFFFFFF; }
        for (int n = wv; n < nr; n += 4) {
            const int64_t d = rows.dot(lane, m, n);
            if (lane == 0) {
                const double c = cos64(d, qnm, rn[n]);
                if (c > bc[k - 1] || (c == bc[k - 1] && n < bi[k - 1])) {
                    int p = k - 1;
                    while (p > 0 && (bc[p - 1] < c || (bc[p - 1] == c && bi[p - 1] > n))) {
                        bc[p] = bc[p - 1];
                        bi[p] = bi[p - 1];
                        --p;
                    }
                    bc[p] = c;
                    bi[p] = n;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < 4; ++w)  // merge the wave lists into wave 0's
                for (int j = 0; j < k; ++j) {
                    const double c = s_c[w][j];
                    const int n = s_i[w][j];
                    if (n == 0x7FFFFFFF) break;
                    if (!(c > s_c[0][k - 1] || (c == s_c[0][k - 1] && n < s_i[0][k - 1]))) continue;
                    int p = k - 1;
                    while (p > 0 && (s_c[0][p - 1] < c || (s_c[0][p - 1] == c && s_i[0][p - 1] > n))) {
                        s_c[0][p] = s_c[0][p - 1];
                        s_i[0][p] = s_i[0][p - 1];
                        --p;
                    }
                    s_c[0][p] = c;
                    s_i[0][p] = n;
                }
            for (int j = 0; j < k; ++j) {
                const int64_t o = (int64_t)m * k + j;
                const bool ok = s_i[0][j] != 0x7FFFFFFF;
                out_idx[o] = ok ? s_i[0][j] : -1;
                out_score[o] = ok ? (float)s_c[0][j] : -INFINITY;
                if (out_cos) out_cos[o] = ok ? s_c[0][j] : -INFINITY;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------
inline int pick_kp(int k) { return k <= 5 ? 8 : (k <= 13 ? 16 : 32); }
// Largest k the candidate-list kernels serve (the margin check needs a list
// longer than k); larger k go through the key-matrix selection (select.hip).
constexpr int kListMaxK = 29;

struct SearchBufs {
    float* iq;
    float* ir;
    float2* bq;
    float2* br;
    uint64_t* part;
    int32_t* flag_list;   // queued by the first finalize
    int32_t* okey;        // finalize order: best-candidate index per query
    int32_t* order;       // queries sorted by okey
    int32_t* ohist;       // [nr + 1] bucket counts, then cursors
    int32_t* ocur;
    int32_t* ototal;
    int32_t* flag_count;
    uint64_t* part2;      // second (exact) pass
    int32_t* flag_list2;  // queued by the second finalize
    int32_t* flag_count2;
    uint64_t* qthr;       // per query: the largest threshold key any slice has reached
};

constexpr int kPass2Cap = 16384;  // queries the second pass takes at once
// reference slices of the second pass (4,126 queued reads x 100k references,
// two-limb lists: 21.0 ms with 16 slices, 18.8 with 32, 21.0 with 64)
constexpr int kPass2Splits = 32;
constexpr int kPass2KP = 32;

// Bytes of each search buffer for nq queries, nr references, S slices of
// kp-entry lists; the second pass takes min(nq, kPass2Cap) queries.
struct SearchSizes {
    size_t norms, part, flags, part2;
};
inline SearchSizes search_sizes(int nq, int nr, int n_splits, int kp) {
    const size_t nqp = ((size_t)nq + 63) / 64 * 64, nrp = ((size_t)nr + 63) / 64 * 64;
    const size_t cap2 = (size_t)(nq < kPass2Cap ? nq : kPass2Cap);
    SearchSizes z;
    z.norms = sizeof(float) * (nqp + nrp) + sizeof(float2) * (nqp + nrp) + sizeof(uint64_t) * nqp;
    z.part = sizeof(uint64_t) * (size_t)n_splits * nq * kp;
    z.flags = sizeof(int32_t) * (4 * (size_t)nq + 2 * (size_t)nr + 512);
    z.part2 = sizeof(uint64_t) * (size_t)kPass2Splits * cap2 * kPass2KP;
    return z;
}

// Points b into the four regions (sized by search_sizes).
inline void carve(SearchBufs& b, void* norms, void* part, void* flags, void* part2, int nq,
                  int nr) {
    const size_t nqp = ((size_t)nq + 63) / 64 * 64, nrp = ((size_t)nr + 63) / 64 * 64;
    b.iq = (float*)norms;
    b.ir = b.iq + nqp;
    b.bq = reinterpret_cast<float2*>(b.ir + nrp);
    b.br = b.bq + nqp;
    b.qthr = reinterpret_cast<uint64_t*>(b.br + nrp);
    b.part = (uint64_t*)part;
    b.flag_count = (int32_t*)flags;
    b.flag_count2 = b.flag_count + 32;
    b.flag_list = b.flag_count + 64;
    b.flag_list2 = b.flag_list + nq + 64;
    b.okey = b.flag_list2 + nq + 64;
    b.order = b.okey + nq;
    b.ohist = b.order + nq;
    b.ocur = b.ohist + nr + 1;
    b.ototal = b.ocur + nr + 1;
    b.part2 = (uint64_t*)part2;
}

// The search buffers in the context's workspaces (grown on demand).
inline int prepare(hvk_ctx* ctx, int nq, int nr, int n_splits, int kp, SearchBufs& b) {
    const SearchSizes z = search_sizes(nq, nr, n_splits, kp);
    if (int st = ctx->ws_norms.ensure(z.norms)) return st;
    if (int st = ctx->ws_search.ensure(z.part)) return st;
    if (int st = ctx->ws_flags.ensure(z.flags)) return st;
    if (int st = ctx->ws_search2.ensure(z.part2)) return st;
    carve(b, ctx->ws_norms.ptr, ctx->ws_search.ptr, ctx->ws_flags.ptr, ctx->ws_search2.ptr, nq, nr);
    return HVK_OK;
}

// splits of the reference range so that q_tiles * splits fills the device
inline int choose_splits(hvk_ctx* ctx, int q_tiles, int r_tiles, int wg_per_cu) {
    const int want = ctx->n_cu * wg_per_cu;
    int s = (want + q_tiles - 1) / q_tiles;
    if (s < 1) s = 1;
    if (s > r_tiles) s = r_tiles;
    if (s > 64) s = 64;
    return s;
}

// Reference slices and XCD groups of an MFMA launch (one workgroup per CU).
// Automatic choice: 16 slices from 2048 query tiles up, else 8, or as many as
// fill the device when there are few query tiles (every slice rebuilds its
// candidate lists from an empty start, while few slices leave the last round
// of workgroups part-empty).  want_splits > 0 (the second pass) or the
// context's tuning value forces the slice count.
constexpr int kGroupSlices = 8;  // reference slices per XCD group by default
struct MfmaPlan {
    TileMap map;
    int blocks;
};
inline MfmaPlan plan_mfma(hvk_ctx* ctx, int q_tiles, int r_tiles, int want_splits) {
    const int per_xcd = ctx->n_cu / kXcds > 0 ? ctx->n_cu / kXcds : 1;
    int s = want_splits > 0 ? want_splits : ctx->tune_splits;
    if (s <= 0) {
        // measured on 100k references: 16 slices at 1M reads (3,907 read tiles) and 8 at
        // 262k reads (1,024 tiles) beat 4 by 2-9%; at least enough to fill every CU
        s = q_tiles >= 2048 ? 16 : 8;
        const int fill = (ctx->n_cu + q_tiles - 1) / q_tiles;
        if (s < fill) s = fill;
    }
    if (s > r_tiles) s = r_tiles;
    if (s > 64) s = 64;
    if (s < 1) s = 1;
    MfmaPlan p;
    p.map.q_tiles = q_tiles;
    p.map.split_tiles = (r_tiles + s - 1) / s;
    p.map.n_splits = (r_tiles + p.map.split_tiles - 1) / p.map.split_tiles;
    // XCD groups of gq query tiles x sg slices (gq * sg = CUs per XCD): the
    // XCD's L2 serves each query slab to sg workgroups and each reference
    // slab to gq.  Measured at 1M reads x 100k references with 16 slices:
    // 8 slices per group (4 x 8) 905 ms, 4 (8 x 4) 907 ms, all 16 (2 x 16)
    // 920 ms, 2 (16 x 2) 928-968 ms; identical results.
    int sg = ctx->tune_group_slices > 0 ? ctx->tune_group_slices : kGroupSlices;
    if (sg > p.map.n_splits) sg = p.map.n_splits;
    p.map.sg = sg;
    p.map.n_sgroups = (p.map.n_splits + sg - 1) / sg;
    p.map.gq = per_xcd / sg > 0 ? per_xcd / sg : 1;
    p.map.rx = p.map.gq * sg;
    const int groups = (q_tiles + p.map.gq - 1) / p.map.gq * p.map.n_sgroups;
    p.blocks = (groups + kXcds - 1) / kXcds * kXcds * p.map.rx;
    return p;
}

// Orders the queries by the index of their best candidate (a counting sort
// over the split lists' first keys) into b.order, for finalize.
int launch_finalize_order(const SearchBufs& b, int S, int nq, int kp, int nr, hipStream_t st);

template <typename Rows, int KP>
int finalize(hvk_ctx* ctx, const Rows& rows, const SearchBufs& b, int S, const int64_t* qn, int nq,
             const int64_t* rn, int nr, int k, int32_t* idx, float* score, double* cosv,
             hipStream_t st, bool ordered = false) {
    if (ordered)
        if (int rc = launch_finalize_order(b, S, nq, KP, nr, st)) return rc;
    hipLaunchKernelGGL((finalize_kernel<KP, Rows>), dim3((nq + 3) / 4), dim3(256), 0, st, rows,
                       b.part, S, nq, qn, rn, nr, k, idx, score, cosv, b.flag_list, b.flag_count,
                       nullptr, nullptr, nq, ordered ? b.order : nullptr);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

// Exact FP64 brute force of the queries left in (list, count); reports the
// number of queries that took any exact pass.
int rescan(hvk_ctx* ctx, const RowsLimb& rows, const int32_t* list, const int32_t* count,
           const SearchBufs& b, const int64_t* qn, const int64_t* rn, int nr, int k, int32_t* idx,
           float* score, double* cosv, int32_t* n_res, hipStream_t st);
int rescan(hvk_ctx* ctx, const RowsI16& rows, const int32_t* list, const int32_t* count,
           const SearchBufs& b, const int64_t* qn, const int64_t* rn, int nr, int k, int32_t* idx,
           float* score, double* cosv, int32_t* n_res, hipStream_t st);

// FP32 inverse norms (and the bound terms when bnd holds rows); clears the
// re-scan queue counters.
void launch_inv_norm(const int64_t* qn, int nq, const int64_t* rn, int nr, const SearchBufs& b,
                     const BoundRows& bnd, bool bounds, hipStream_t st);

// Exact top-k for kListMaxK < k (any k; k > nr pads with -1): the key
// matrix of a chunk of queries from the plain kernel (int16 rows) or the
// MFMA kernel (exact limbs; hi rows nullable), then a radix select and sort
// per query.  *n_res (nullable) is set to 0: nothing is re-scored.
int select_topk_plain(hvk_ctx* ctx, const int16_t* q, const int64_t* qn, int nq, const int16_t* r,
                      const int64_t* rn, int nr, int64_t ld, int k, int32_t* idx, float* score,
                      double* cosv, int32_t* n_res, hipStream_t st);
int select_topk_limb(hvk_ctx* ctx, const int8_t* qlo, const int8_t* qhi, const int64_t* qn, int nq,
                     const int8_t* rlo, const int8_t* rhi, const int64_t* rn, int nr, int64_t ld,
                     int k, int32_t* idx, float* score, double* cosv, int32_t* n_res,
                     hipStream_t st);

// Second pass of the bounded form (one limb, 32-entry lists) over the
// queries its first finalize queued, its own finalize, then the FP64 re-scan
// of what remains.
int exact_pass2(hvk_ctx* ctx, const hvk_bundles& q, const hvk_bundles& r, int64_t ld, int k,
                const SearchBufs& b, int32_t* idx, float* score, double* cosv, int32_t* n_res,
                hipStream_t st);

template <typename Rows, int KP>
int finish(hvk_ctx* ctx, const Rows& rows, const SearchBufs& b, int S, const int64_t* qn, int nq,
           const int64_t* rn, int nr, int k, int32_t* idx, float* score, double* cosv,
           int32_t* n_res, hipStream_t st) {
    // the limb rows of the MFMA forms are re-read per candidate: order the
    // queries so neighbouring waves share them in L2
    if (int rc = finalize<Rows, KP>(ctx, rows, b, S, qn, nq, rn, nr, k, idx, score, cosv, st,
                                    std::is_same<Rows, RowsLimb>::value))
        return rc;
    return rescan(ctx, rows, b.flag_list, b.flag_count, b, qn, rn, nr, k, idx, score, cosv, n_res,
                  st);
}

template <int KP>
int run_plain(hvk_ctx* ctx, const int16_t* q, const int64_t* qn, int nq, const int16_t* r,
              const int64_t* rn, int nr, int64_t ld, int k, int32_t* idx, float* score,
              double* cosv, int32_t* n_res, hipStream_t st) {
    const int q_tiles = (nq + kPlainMT - 1) / kPlainMT;
    const int r_tiles = (nr + kPlainNT - 1) / kPlainNT;
    const int S = choose_splits(ctx, q_tiles, r_tiles, 4);
    const int split_tiles = (r_tiles + S - 1) / S;
    const int S_used = (r_tiles + split_tiles - 1) / split_tiles;
    SearchBufs b;
    int rc = prepare(ctx, nq, nr, S_used, KP, b);
    if (rc) return rc;
    launch_inv_norm(qn, nq, rn, nr, b, BoundRows{}, false, st);
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL((plain_kernel<KP>), dim3(q_tiles, S_used), dim3(256), 0, st, q, b.iq, nq, r,
                       b.ir, nr, ld, split_tiles, b.part);
    timing_mark(ctx, 1, true, st);
    HVK_HIP(hipGetLastError());
    return finish<RowsI16, KP>(ctx, RowsI16{q, r, ld}, b, S_used, qn, nq, rn, nr, k, idx, score,
                               cosv, n_res, st);
}

template <int QL, int RL, int KP>
int run_mfma(hvk_ctx* ctx, const int8_t* qlo, const int8_t* qhi, const int64_t* qn, int nq,
             const int8_t* rlo, const int8_t* rhi, const int64_t* rn, int nr, int64_t ld, int k,
             int32_t* idx, float* score, double* cosv, int32_t* n_res, hipStream_t st) {
    using T = MfmaTile<QL, RL>;
    const MfmaPlan pl = plan_mfma(ctx, (nq + T::MT - 1) / T::MT, (nr + T::NT - 1) / T::NT, 0);
    SearchBufs b;
    int rc = prepare(ctx, nq, nr, pl.map.n_splits, KP, b);
    if (rc) return rc;
    launch_inv_norm(qn, nq, rn, nr, b, BoundRows{}, false, st);
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL((mfma_kernel<QL, RL, KP>), dim3(pl.blocks), dim3(kMfmaThreads), 0, st,
                       qlo, qhi, b.iq, nq, rlo, rhi, b.ir, nr, ld, pl.map, b.part, nullptr,
                       nullptr, nullptr, nullptr, b.qthr);
    timing_mark(ctx, 1, true, st);
    HVK_HIP(hipGetLastError());
    return finish<RowsLimb, KP>(ctx, RowsLimb{qlo, QL > 1 ? qhi : nullptr, rlo,
                                              RL > 1 ? rhi : nullptr, ld},
                                b, pl.map.n_splits, qn, nq, rn, nr, k, idx, score, cosv, n_res, st);
}

// Rows and columns of the one-limb tile of shape sh (0 .. kMfmaShapes - 1).
inline void shape_tile(int sh, int& mt, int& nt) {
    switch (sh) {
        case 1: mt = MfmaTile<1, 1, 1>::MT; nt = MfmaTile<1, 1, 1>::NT; return;
        case 2: mt = MfmaTile<1, 1, 2>::MT; nt = MfmaTile<1, 1, 2>::NT; return;
        case 3: mt = MfmaTile<1, 1, 3>::MT; nt = MfmaTile<1, 1, 3>::NT; return;
        case 4: mt = MfmaTile<1, 1, 4>::MT; nt = MfmaTile<1, 1, 4>::NT; return;
        case 5: mt = MfmaTile<1, 1, 5>::MT; nt = MfmaTile<1, 1, 5>::NT; return;
        default: mt = MfmaTile<1, 1>::MT; nt = MfmaTile<1, 1>::NT; return;
    }
}

// Reference slices and tile map of a bounded one-limb search on tile shape sh.
inline MfmaPlan plan_bounded(hvk_ctx* ctx, int nq, int nr, int sh = 0) {
    int mt, nt;
    shape_tile(sh, mt, nt);
    return plan_mfma(ctx, (nq + mt - 1) / mt, (nr + nt - 1) / nt, 0);
}

// The launches of a bounded search on buffers already laid out (b sized by
// search_sizes(nq, nr, pl.map.n_splits, KP)), on tile shape SH.
template <int KP, int SH = 0>
int launch_mfma_bounded(hvk_ctx* ctx, const MfmaPlan& pl, const hvk_bundles& q,
                        const hvk_bundles& r, int64_t ld, int k, const SearchBufs& b, int32_t* idx,
                        float* score, double* cosv, int32_t* n_res, hipStream_t st) {
    const int nq = q.n, nr = r.n;
    launch_inv_norm(q.sqnorm, nq, r.sqnorm, nr, b,
                    BoundRows{q.maxabs, q.resid1, r.maxabs, r.resid1}, true, st);
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL((mfma_kernel<1, 1, KP, true, false, false, SH>), dim3(pl.blocks),
                       dim3(kMfmaThreads), 0, st, q.clamp8, nullptr, b.iq, nq, r.clamp8, nullptr,
                       b.ir, nr, ld, pl.map, b.part, b.bq, b.br, nullptr, nullptr, b.qthr, nullptr,
                       nullptr, nullptr);
    timing_mark(ctx, 1, true, st);
    HVK_HIP(hipGetLastError());
    const RowsLimb rows{q.lo, q.hi, r.lo, r.hi, ld};
    if (int rc = finalize<RowsLimb, KP>(ctx, rows, b, pl.map.n_splits, q.sqnorm, nq, r.sqnorm, nr,
                                        k, idx, score, cosv, st, true))
        return rc;
    return exact_pass2(ctx, q, r, ld, k, b, idx, score, cosv, n_res, st);
}

// Tile shapes other than the default are compiled for 8-entry lists (top-5).
template <int KP>
int run_mfma_bounded(hvk_ctx* ctx, const hvk_bundles& q, const hvk_bundles& r, int64_t ld, int k,
                     int32_t* idx, float* score, double* cosv, int32_t* n_res, hipStream_t st) {
    const int sh = KP == 8 ? ctx->tune_shape : 0;
    const MfmaPlan pl = plan_bounded(ctx, q.n, r.n, sh);
    SearchBufs b;
    if (int rc = prepare(ctx, q.n, r.n, pl.map.n_splits, KP, b)) return rc;
    if constexpr (KP == 8) {
        switch (sh) {
            case 1: return launch_mfma_bounded<KP, 1>(ctx, pl, q, r, ld, k, b, idx, score, cosv, n_res, st);
            case 2: return launch_mfma_bounded<KP, 2>(ctx, pl, q, r, ld, k, b, idx, score, cosv, n_res, st);
            case 3: return launch_mfma_bounded<KP, 3>(ctx, pl, q, r, ld, k, b, idx, score, cosv, n_res, st);
            case 4: return launch_mfma_bounded<KP, 4>(ctx, pl, q, r, ld, k, b, idx, score, cosv, n_res, st);
            case 5: return launch_mfma_bounded<KP, 5>(ctx, pl, q, r, ld, k, b, idx, score, cosv, n_res, st);
            default: break;
        }
    }
    return launch_mfma_bounded<KP>(ctx, pl, q, r, ld, k, b, idx, score, cosv, n_res, st);
}

template <int QL, int RL>
int dispatch_mfma_kp(hvk_ctx* ctx, int kp, const int8_t* qlo, const int8_t* qhi, const int64_t* qn,
                     int nq, const int8_t* rlo, const int8_t* rhi, const int64_t* rn, int nr,
                     int64_t ld, int k, int32_t* idx, float* score, double* cosv, int32_t* n_res,
                     hipStream_t st) {
    switch (kp) {
        case 8: return run_mfma<QL, RL, 8>(ctx, qlo, qhi, qn, nq, rlo, rhi, rn, nr, ld, k, idx, score, cosv, n_res, st);
        case 16: return run_mfma<QL, RL, 16>(ctx, qlo, qhi, qn, nq, rlo, rhi, rn, nr, ld, k, idx, score, cosv, n_res, st);
        default: return run_mfma<QL, RL, 32>(ctx, qlo, qhi, qn, nq, rlo, rhi, rn, nr, ld, k, idx, score, cosv, n_res, st);
    }
}


// Explicit instantiation (EXT empty) or its suppression (EXT extern) of the
// run functions of one candidate-list length; each length is compiled in a
// translation unit of its own.
#define HVK_SEARCH_RUNS(KP, EXT)                                                                  \
    EXT template int run_plain<KP>(hvk_ctx*, const int16_t*, const int64_t*, int, const int16_t*, \
                                   const int64_t*, int, int64_t, int, int32_t*, float*, double*,  \
                                   int32_t*, hipStream_t);                                         \
    EXT template int run_mfma<1, 1, KP>(HVK_RUN_MFMA_ARGS);                                        \
    EXT template int run_mfma<1, 2, KP>(HVK_RUN_MFMA_ARGS);                                        \
    EXT template int run_mfma<2, 1, KP>(HVK_RUN_MFMA_ARGS);                                        \
    EXT template int run_mfma<2, 2, KP>(HVK_RUN_MFMA_ARGS);
#define HVK_SEARCH_BOUNDED(KP, EXT)                                                               \
    EXT template int run_mfma_bounded<KP>(hvk_ctx*, const hvk_bundles&, const hvk_bundles&,       \
                                          int64_t, int, int32_t*, float*, double*, int32_t*,      \
                                          hipStream_t);                                            \
    EXT template int launch_mfma_bounded<KP>(hvk_ctx*, const MfmaPlan&, const hvk_bundles&,       \
                                             const hvk_bundles&, int64_t, int, const SearchBufs&, \
                                             int32_t*, float*, double*, int32_t*, hipStream_t);
#define HVK_RUN_MFMA_ARGS                                                                          \
    hvk_ctx*, const int8_t*, const int8_t*, const int64_t*, int, const int8_t*, const int8_t*,    \
        const int64_t*, int, int64_t, int, int32_t*, float*, double*, int32_t*, hipStream_t

}  // namespace hvk
