// search.hip -- exact cosine top-k of query bundles against reference bundles.
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
#include <cmath>
#include <utility>

#include "hvk_internal.h"

namespace hvk {
namespace {

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
// per-workgroup top lists in LDS (keys only; 0 = empty)
// ------------------------------------------------------------------------
template <int MT, int KP, int CAP>
struct TopLds {
    uint64_t top[MT][KP];  // descending
    uint64_t buf[MT][CAP];
    uint64_t thr[MT];      // top[m][KP-1]
    float thr_f[MT];       // its FP32 score, -inf while the list is not full
    int cnt[MT];
    int ovf[2];

    __device__ void init(int tid, int nthr) {
        for (int i = tid; i < MT * KP; i += nthr) top[i / KP][i % KP] = 0;
        for (int i = tid; i < MT; i += nthr) { thr[i] = 0; thr_f[i] = -INFINITY; cnt[i] = 0; }
        if (tid == 0) { ovf[0] = 0; ovf[1] = 0; }
    }

    // Offers one (query m, reference) key.  Returns true when it beat the
    // threshold but the buffer was full: the caller keeps it for the next
    // pass, which runs after the merge has raised the threshold.  A key is
    // therefore offered at most once until accepted or fallen below.
    __device__ __forceinline__ bool offer(int m, uint64_t key, int pass) {
        if (key <= thr[m]) return false;
        const int slot = atomicAdd(&cnt[m], 1);
        if (slot < CAP) {
            buf[m][slot] = key;
            return false;
        }
        ovf[pass & 1] = 1;
        return true;
    }

    __device__ void merge(int m) {
        const int n = cnt[m] < CAP ? cnt[m] : CAP;
        for (int q = 0; q < n; ++q) {
            const uint64_t c = buf[m][q];
            if (c <= top[m][KP - 1]) continue;
            int p = KP - 1;
            while (p > 0 && top[m][p - 1] < c) {
                top[m][p] = top[m][p - 1];
                --p;
            }
            top[m][p] = c;
        }
        thr[m] = top[m][KP - 1];
        thr_f[m] = thr[m] ? f32_of_ord((uint32_t)(thr[m] >> 32)) : -INFINITY;
        cnt[m] = 0;
    }
};

// End of one offer pass: merge the buffers; true when a key was dropped
// (the caller then runs another pass over its dropped keys).
template <int MT, int KP, int CAP>
__device__ __forceinline__ bool topk_pass_end(TopLds<MT, KP, CAP>& tl, int tid, int pass) {
    __syncthreads();
    const int ov = tl.ovf[pass & 1];
    if (tid == 0) tl.ovf[(pass + 1) & 1] = 0;
    if (tid < MT) tl.merge(tid);
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
// sum_j bq[m].j * br[n].j bounds |dot - dot_clamped| * iq * ir.
__global__ void inv_norm_kernel(const int64_t* __restrict__ qn, int nq, const int64_t* __restrict__ rn,
                                int nr, float* __restrict__ iq, float* __restrict__ ir,
                                int32_t* __restrict__ flag_count, BoundRows bnd,
                                float2* __restrict__ bq, float2* __restrict__ br) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { flag_count[0] = 0; flag_count[32] = 0; }
    if (i < nq) {
        const int64_t v = qn[i];
        const float x = v == 0 ? 0.0f : __fdiv_rn(1.0f, __fsqrt_rn((float)v));
        iq[i] = x;
        if (bq) {
            const int a = bnd.q_maxabs[i] < 127 ? bnd.q_maxabs[i] : 127;
            bq[i] = make_float2(__fmul_rn((float)a, x), __fmul_rn((float)bnd.q_resid1[i], x));
        }
    }
    if (i < nr) {
        const int64_t v = rn[i];
        const float x = v == 0 ? 0.0f : __fdiv_rn(1.0f, __fsqrt_rn((float)v));
        ir[i] = x;
        if (br)
            br[i] = make_float2(__fmul_rn((float)bnd.r_resid1[i], x),
                                __fmul_rn((float)bnd.r_maxabs[i], x));
    }
}

// ------------------------------------------------------------------------
// 2a. plain form: VALU int16 dot (v_dot2_i32_i16), 4 queries x 8 references
//     per thread, K-slab of 64 dims; int32 per slab, int64 across slabs.
// ------------------------------------------------------------------------
constexpr int kPlainMT = 32, kPlainNT = 256, kPlainKS = 64;

template <int KP>
__global__ void __launch_bounds__(256)
plain_kernel(const int16_t* __restrict__ q, const float* __restrict__ iq, int nq,
             const int16_t* __restrict__ r, const float* __restrict__ ir, int nr, int64_t ld,
             int split_tiles, uint64_t* __restrict__ part) {
    constexpr int MT = kPlainMT, NT = kPlainNT, KS = kPlainKS, KW = KS / 2;
    constexpr int CAP = 16;
    __shared__ uint32_t sQ[KW][MT];
    __shared__ uint32_t sR[KW][NT + 1];
    __shared__ TopLds<MT, KP, CAP> tl;
    __shared__ float s_iq[MT];

    const int tid = threadIdx.x;
    const int tx = tid & 31, ty = tid >> 5;
    const int m0 = blockIdx.x * MT;
    const int split = blockIdx.y;
    tl.init(tid, 256);
    if (tid < MT) s_iq[tid] = (m0 + tid < nq) ? iq[m0 + tid] : 0.0f;
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
                    if (tl.offer(m, make_key(s, n), pass)) dropped |= bit;
                    else dropped &= ~bit;
                }
            });
            if (!topk_pass_end(tl, tid, pass)) break;
        }
    }
    for (int i = tid; i < MT * KP; i += 256) {
        const int m = i / KP, j = i % KP;
        if (m0 + m < nq) part[((int64_t)split * nq + m0 + m) * KP + j] = tl.top[m][j];
    }
}

// ------------------------------------------------------------------------
// 2b. MFMA form: v_mfma_i32_32x32x32_i8.  Workgroup of 4 waves (one per
//     SIMD, 512 registers each) on an MT-query x 128-reference tile; wave
//     grid 2 (queries) x 2 (references), wave tile (32 WX) x 64, i.e. WX x 2
//     MFMA blocks per limb pair with the accumulators in AGPRs.
//     MT = 256 (WX = 4) for one query limb, 128 (WX = 2) for two.
//     K-slabs of 64 bytes, double-buffered in LDS (rows XOR-swizzled in
//     16-byte chunks), the next slab prefetched into registers while the
//     current one is multiplied: one barrier per slab.
//     Queries are the A operand (rows), references the B operand (columns).
// ------------------------------------------------------------------------
constexpr int kMfmaNT = 128, kMfmaKS = 64, kMfmaThreads = 256;
template <int QL> constexpr int mfma_wx() { return 2; }
template <int QL> constexpr int mfma_mt() { return 64 * mfma_wx<QL>(); }

// 16-byte global -> LDS copy (global_load_lds_dwordx4): lane l of the wave
// writes LDS bytes [lds + 16 l, lds + 16 l + 16); lds must be wave-uniform.
// The builtin exists only for the device pass.
__device__ __forceinline__ void load_lds16(const int8_t* g, int8_t* lds) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0);
#endif
}

__device__ __forceinline__ int swz(int row, int chunk) {
    return row * 64 + ((chunk ^ ((row >> 1) & 3)) << 4);
}

template <int QL, int RL, int KP>
struct MfmaSmem {
    int8_t q[2][QL][mfma_mt<QL>() * kMfmaKS];
    int8_t r[2][RL][kMfmaNT * kMfmaKS];
    TopLds<mfma_mt<QL>(), KP, 8> tl;
    float iq[mfma_mt<QL>()];
    float2 bq[mfma_mt<QL>()];
};

// BND: one-limb bounded form -- both operands hold clamped counts and each
// score is raised by bq[m].x br[n].x + bq[m].y br[n].y, an upper bound on
// |dot - dot_clamped| * iq * ir.
template <int QL, int RL, int KP, bool BND = false>
__global__ void __launch_bounds__(kMfmaThreads, 2)
mfma_kernel(const int8_t* __restrict__ qlo, const int8_t* __restrict__ qhi,
            const float* __restrict__ iq, int nq, const int8_t* __restrict__ rlo,
            const int8_t* __restrict__ rhi, const float* __restrict__ ir, int nr, int64_t ld,
            int split_tiles, uint64_t* __restrict__ part, const float2* __restrict__ bq,
            const float2* __restrict__ br, const int32_t* __restrict__ q_rows,
            const int32_t* __restrict__ q_count) {
    // q_rows (nullable): query m of this launch is stored row q_rows[m], and
    // the launch covers m < *q_count (a count on the device, at most nq).
    constexpr int WX = mfma_wx<QL>(), MT = mfma_mt<QL>(), NT = kMfmaNT, KS = kMfmaKS;
    constexpr int NTH = kMfmaThreads;
    constexpr int NPAIR = WX * 2 * 16;       // (query, reference) pairs per lane
    __shared__ __attribute__((aligned(16))) MfmaSmem<QL, RL, KP> sm;
    auto& tl = sm.tl;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = blockIdx.x * MT;
    const int split = blockIdx.y;
    if (q_rows) {
        const int c = *q_count;
        nq = c < nq ? c : nq;
        if (m0 >= nq) return;  // the whole workgroup leaves before any barrier
    }
    auto qrow = [&](int m) { return q_rows ? q_rows[m] : m; };
    tl.init(tid, NTH);
    if (tid < MT) {
        const bool in = m0 + tid < nq;
        sm.iq[tid] = in ? iq[qrow(m0 + tid)] : 0.0f;
        sm.bq[tid] = (BND && in) ? bq[qrow(m0 + tid)] : make_float2(0.0f, 0.0f);
    }
    const int n_begin = split * split_tiles * NT;
    int n_end = n_begin + split_tiles * NT;
    if (n_end > nr) n_end = nr;
    const int8_t* qsrc[2] = {qlo, qhi};
    const int8_t* rsrc[2] = {rlo, rhi};
    const int nk = (int)(ld / KS);

    // Direct global->LDS loads: one wave instruction fills 1 KB of LDS (16
    // rows of 64 B), lane l writing bytes [16 l, 16 l + 16).  Lane l of
    // wave-load j therefore fetches row 16 j + l / 4, logical chunk
    // (l & 3) ^ swizzle(row), so that LDS holds the swizzled layout swz().
    constexpr int QJ = MT * KS / 1024 / 4;  // wave-loads per wave per Q limb
    constexpr int RJ = NT * KS / 1024 / 4;  // per R limb
    int q_off[QJ], r_off[RJ];               // byte offset of the lane's chunk in its row
    int q_rowi[QJ], r_rowi[RJ];             // row within the tile
#pragma unroll
    for (int t = 0; t < QJ; ++t) {
        const int j = wave * QJ + t, row = 16 * j + (lane >> 2);
        q_rowi[t] = row;
        q_off[t] = 16 * ((lane & 3) ^ ((row >> 1) & 3));
    }
#pragma unroll
    for (int t = 0; t < RJ; ++t) {
        const int j = wave * RJ + t, row = 16 * j + (lane >> 2);
        r_rowi[t] = row;
        r_off[t] = 16 * ((lane & 3) ^ ((row >> 1) & 3));
    }
    __syncthreads();

    for (int n0 = n_begin; n0 < n_end; n0 += NT) {
        // rows past the end read a valid row; their results are never offered
        const int8_t* gq[QL][QJ];
        const int8_t* gr[RL][RJ];
#pragma unroll
        for (int l = 0; l < QL; ++l)
#pragma unroll
            for (int t = 0; t < QJ; ++t) {
                const int row = qrow(min(m0 + q_rowi[t], nq - 1));
                gq[l][t] = qsrc[l] + (int64_t)row * ld + q_off[t];
            }
#pragma unroll
        for (int l = 0; l < RL; ++l)
#pragma unroll
            for (int t = 0; t < RJ; ++t) {
                const int row = min(n0 + r_rowi[t], nr - 1);
                gr[l][t] = rsrc[l] + (int64_t)row * ld + r_off[t];
            }
        auto issue_slab = [&](int kk, int buf) {
            const int64_t k0 = (int64_t)kk * KS;
#pragma unroll
            for (int l = 0; l < QL; ++l)
#pragma unroll
                for (int t = 0; t < QJ; ++t)
                    load_lds16(gq[l][t] + k0, &sm.q[buf][l][1024 * (wave * QJ + t)]);
#pragma unroll
            for (int l = 0; l < RL; ++l)
#pragma unroll
                for (int t = 0; t < RJ; ++t)
                    load_lds16(gr[l][t] + k0, &sm.r[buf][l][1024 * (wave * RJ + t)]);
        };

        v16i acc[QL][RL][WX][2];
        static_for<QL * RL * WX * 2>([&](auto I) {
            constexpr int i = decltype(I)::value;
            acc[i / (RL * WX * 2)][(i / (WX * 2)) % RL][(i / 2) % WX][i % 2] = v16i{};
        });

        issue_slab(0, 0);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __syncthreads();
        for (int kk = 0; kk < nk; ++kk) {
            const int cur = kk & 1;
            if (kk + 1 < nk) issue_slab(kk + 1, cur ^ 1);
#pragma unroll
            for (int ks = 0; ks < KS / 32; ++ks) {
                const int ch = ks * 2 + (lane >> 5);
                v4i a[QL][WX], b[RL][2];
#pragma unroll
                for (int l = 0; l < QL; ++l)
#pragma unroll
                    for (int x = 0; x < WX; ++x)
                        a[l][x] = *reinterpret_cast<const v4i*>(&sm.q[cur][l][swz(wm * 32 * WX + x * 32 + (lane & 31), ch)]);
#pragma unroll
                for (int l = 0; l < RL; ++l)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
                        b[l][y] = *reinterpret_cast<const v4i*>(&sm.r[cur][l][swz(wn * 64 + y * 32 + (lane & 31), ch)]);
                static_for<QL * RL * WX * 2>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    constexpr int qa = i / (RL * WX * 2), rb = (i / (WX * 2)) % RL;
                    constexpr int x = (i / 2) % WX, y = i % 2;
                    acc[qa][rb][x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                        a[qa][x], b[rb][y], acc[qa][rb][x][y], 0, 0, 0);
                });
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);  // this wave's loads of slab kk+1 landed
            __syncthreads();
        }

        float irv[2];
        float2 brv[2];
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int n = n0 + wn * 64 + y * 32 + (lane & 31);
            irv[y] = n < nr ? ir[n] : 0.0f;
            brv[y] = (BND && n < nr) ? br[n] : make_float2(0.0f, 0.0f);
        }
        // Epilogue.  Pair p = (x * 2 + y) * 16 + e is query row
        // wm*32*WX + x*32 + 8(e/4) + 4(lane/32) + e%4, reference column
        // n0 + wn*64 + y*32 + lane%32 (the 32x32 MFMA output layout).  A float
        // prefilter against the query's current FP32 threshold rejects almost
        // every pair; only survivors build keys and enter the LDS lists.
        uint64_t dropped[(NPAIR + 63) / 64];
#pragma unroll
        for (int i = 0; i < (NPAIR + 63) / 64; ++i) dropped[i] = 0;
        for (int pass = 0;; ++pass) {
            static_for<WX * 2>([&](auto XY) {
                constexpr int x = decltype(XY)::value / 2, y = decltype(XY)::value % 2;
                const int n = n0 + wn * 64 + y * 32 + (lane & 31);
                const int mb = wm * 32 * WX + x * 32 + 4 * (lane >> 5);
                float sv[16];
                bool hit = false;
                static_for<4>([&](auto G) {
                    constexpr int g = decltype(G)::value;
                    const float4 iq4 = *reinterpret_cast<const float4*>(&sm.iq[mb + 8 * g]);
                    const float4 th4 = *reinterpret_cast<const float4*>(&tl.thr_f[mb + 8 * g]);
                    const float iqv[4] = {iq4.x, iq4.y, iq4.z, iq4.w};
                    const float thv[4] = {th4.x, th4.y, th4.z, th4.w};
                    float2 bqv[4];
                    if constexpr (BND) {
                        const float4 b01 = *reinterpret_cast<const float4*>(&sm.bq[mb + 8 * g]);
                        const float4 b23 = *reinterpret_cast<const float4*>(&sm.bq[mb + 8 * g + 2]);
                        bqv[0] = make_float2(b01.x, b01.y); bqv[1] = make_float2(b01.z, b01.w);
                        bqv[2] = make_float2(b23.x, b23.y); bqv[3] = make_float2(b23.z, b23.w);
                    }
                    static_for<4>([&](auto H) {
                        constexpr int h = decltype(H)::value, e = 4 * g + h;
                        constexpr int p = (x * 2 + y) * 16 + e;
                        float v = 0.0f;  // sum_{qa,rb} acc * 256^(qa+rb) in FP32
                        static_for<QL * RL>([&](auto J) {
                            constexpr int qa = decltype(J)::value / RL, rb = decltype(J)::value % RL;
                            constexpr float scale = (float)(1u << (8 * (qa + rb)));
                            v = __fmaf_rn((float)acc[qa][rb][x][y][e], scale, v);
                        });
                        sv[e] = __fmul_rn(__fmul_rn(v, iqv[h]), irv[y]);
                        if constexpr (BND)
                            sv[e] = __fmaf_rn(bqv[h].y, brv[y].y, __fmaf_rn(bqv[h].x, brv[y].x, sv[e]));
                        const bool want = pass == 0 || ((dropped[p / 64] >> (p % 64)) & 1);
                        hit |= want && sv[e] >= thv[h];
                    });
                });
                if (hit && n < nr) {
                    static_for<16>([&](auto E) {
                        constexpr int e = decltype(E)::value, p = (x * 2 + y) * 16 + e;
                        constexpr uint64_t bit = 1ull << (p % 64);
                        const int m = mb + 8 * (e >> 2) + (e & 3);
                        const bool want = pass == 0 || (dropped[p / 64] & bit);
                        if (want && m0 + m < nq) {
                            if (tl.offer(m, make_key(sv[e], n), pass)) dropped[p / 64] |= bit;
                            else dropped[p / 64] &= ~bit;
                        }
                    });
                }
            });
            if (!topk_pass_end(tl, tid, pass)) break;
        }
    }
    for (int i = tid; i < MT * KP; i += NTH) {
        const int m = i / KP, j = i % KP;
        if (m0 + m < nq) part[((int64_t)split * nq + m0 + m) * KP + j] = tl.top[m][j];
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
                const int32_t* __restrict__ q_count, int part_rows) {
    // q_rows (nullable): query m is stored row q_rows[m], m < *q_count; the
    // partial lists hold part_rows queries, later ones are forwarded to the
    // flag list unexamined.
    __shared__ uint64_t s_key[4][KP];
    __shared__ double s_c[4][KP];
    __shared__ int s_i[4][KP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = blockIdx.x * 4 + wv;
    int n_all = nq;
    if (q_rows) n_all = *q_count < nq ? *q_count : nq;
    if (m >= n_all) return;  // whole wave exits together
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
    for (int j = 0; j < nv; ++j) {
        const int idx = idx_of_key(key[j]);
        const int64_t d = rows.dot(lane, row, idx);
        if (lane == 0) { s_i[wv][j] = idx; s_c[wv][j] = cos64(d, qnm, rn[idx]); }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane != 0) return;
    double* cs = s_c[wv];
    int* ix = s_i[wv];
    for (int a = 1; a < nv; ++a) {  // insertion sort: c desc, index asc
        const double c = cs[a];
        const int i = ix[a];
        int p = a;
        while (p > 0 && (cs[p - 1] < c || (cs[p - 1] == c && ix[p - 1] > i))) {
            cs[p] = cs[p - 1];
            ix[p] = ix[p - 1];
            --p;
        }
        cs[p] = c;
        ix[p] = i;
    }
    bool flag = false;
    if (nv == KP && nr > KP) {
        const float s_kp = f32_of_ord((uint32_t)(key[KP - 1] >> 32));
        flag = !(cs[k - 1] > (double)s_kp + (double)kMarginEps);
    }
    for (int j = 0; j < k; ++j) {
        const int64_t o = (int64_t)row * k + j;
        if (j < nv) {
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
int pick_kp(int k) { return k <= 5 ? 8 : (k <= 13 ? 16 : 32); }

struct SearchBufs {
    float* iq;
    float* ir;
    float2* bq;
    float2* br;
    uint64_t* part;
    int32_t* flag_list;   // queued by the first finalize
    int32_t* flag_count;
    uint64_t* part2;      // second (exact) pass
    int32_t* flag_list2;  // queued by the second finalize
    int32_t* flag_count2;
};

constexpr int kPass2Cap = 16384;  // queries the exact second pass takes at once
constexpr int kPass2Splits = 16;
constexpr int kPass2KP = 32;

int prepare(hvk_ctx* ctx, int nq, int nr, int n_splits, int kp, SearchBufs& b) {
    const size_t need_n = sizeof(float) * (3 * ((size_t)nq + nr) + 512);
    int st = ctx->ws_norms.ensure(need_n);
    if (st) return st;
    st = ctx->ws_search.ensure(sizeof(uint64_t) * (size_t)n_splits * nq * kp);
    if (st) return st;
    st = ctx->ws_flags.ensure(sizeof(int32_t) * (2 * (size_t)nq + 128));
    if (st) return st;
    st = ctx->ws_search2.ensure(sizeof(uint64_t) * (size_t)kPass2Splits * kPass2Cap * kPass2KP);
    if (st) return st;
    b.iq = (float*)ctx->ws_norms.ptr;
    b.ir = b.iq + ((nq + 63) / 64) * 64;
    b.bq = reinterpret_cast<float2*>(b.ir + ((nr + 63) / 64) * 64);
    b.br = b.bq + ((nq + 63) / 64) * 64;
    b.part = (uint64_t*)ctx->ws_search.ptr;
    b.flag_count = (int32_t*)ctx->ws_flags.ptr;
    b.flag_count2 = b.flag_count + 32;
    b.flag_list = b.flag_count + 64;
    b.flag_list2 = b.flag_list + nq + 64;
    b.part2 = (uint64_t*)ctx->ws_search2.ptr;
    return HVK_OK;
}

// splits of the reference range so that q_tiles * splits fills the device
int choose_splits(hvk_ctx* ctx, int q_tiles, int r_tiles, int wg_per_cu) {
    const int want = ctx->n_cu * wg_per_cu;
    int s = (want + q_tiles - 1) / q_tiles;
    if (s < 1) s = 1;
    if (s > r_tiles) s = r_tiles;
    if (s > 64) s = 64;
    return s;
}

template <typename Rows, int KP>
int finalize(hvk_ctx* ctx, const Rows& rows, const SearchBufs& b, int S, const int64_t* qn, int nq,
             const int64_t* rn, int nr, int k, int32_t* idx, float* score, double* cosv,
             hipStream_t st) {
    hipLaunchKernelGGL((finalize_kernel<KP, Rows>), dim3((nq + 3) / 4), dim3(256), 0, st, rows,
                       b.part, S, nq, qn, rn, nr, k, idx, score, cosv, b.flag_list, b.flag_count,
                       nullptr, nullptr, nq);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

// Exact FP64 brute force of the queries left in (list, count); reports the
// number of queries that took any exact pass.
template <typename Rows>
int rescan(hvk_ctx* ctx, const Rows& rows, const int32_t* list, const int32_t* count,
           const SearchBufs& b, const int64_t* qn, const int64_t* rn, int nr, int k, int32_t* idx,
           float* score, double* cosv, int32_t* n_res, hipStream_t st) {
    hipLaunchKernelGGL((rescan_kernel<Rows>), dim3(ctx->n_cu), dim3(256), 0, st, rows, qn, rn, nr,
                       k, list, count, idx, score, cosv);
    HVK_HIP(hipGetLastError());
    if (n_res)
        HVK_HIP(hipMemcpyAsync(n_res, b.flag_count, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    return HVK_OK;
}

template <typename Rows, int KP>
int finish(hvk_ctx* ctx, const Rows& rows, const SearchBufs& b, int S, const int64_t* qn, int nq,
           const int64_t* rn, int nr, int k, int32_t* idx, float* score, double* cosv,
           int32_t* n_res, hipStream_t st) {
    if (int rc = finalize<Rows, KP>(ctx, rows, b, S, qn, nq, rn, nr, k, idx, score, cosv, st))
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
    const int nmax = nq > nr ? nq : nr;
    hipLaunchKernelGGL(inv_norm_kernel, dim3((nmax + 255) / 256), dim3(256), 0, st, qn, nq, rn, nr,
                       b.iq, b.ir, b.flag_count, BoundRows{}, nullptr, nullptr);
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
    const int q_tiles = (nq + mfma_mt<QL>() - 1) / mfma_mt<QL>();
    const int r_tiles = (nr + kMfmaNT - 1) / kMfmaNT;
    const int S = choose_splits(ctx, q_tiles, r_tiles, 1);
    const int split_tiles = (r_tiles + S - 1) / S;
    const int S_used = (r_tiles + split_tiles - 1) / split_tiles;
    SearchBufs b;
    int rc = prepare(ctx, nq, nr, S_used, KP, b);
    if (rc) return rc;
    const int nmax = nq > nr ? nq : nr;
    hipLaunchKernelGGL(inv_norm_kernel, dim3((nmax + 255) / 256), dim3(256), 0, st, qn, nq, rn, nr,
                       b.iq, b.ir, b.flag_count, BoundRows{}, nullptr, nullptr);
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL((mfma_kernel<QL, RL, KP>), dim3(q_tiles, S_used), dim3(kMfmaThreads), 0, st,
                       qlo, qhi, b.iq, nq, rlo, rhi, b.ir, nr, ld, split_tiles, b.part, nullptr,
                       nullptr, nullptr, nullptr);
    timing_mark(ctx, 1, true, st);
    HVK_HIP(hipGetLastError());
    return finish<RowsLimb, KP>(ctx, RowsLimb{qlo, QL > 1 ? qhi : nullptr, rlo,
                                              RL > 1 ? rhi : nullptr, ld},
                                b, S_used, qn, nq, rn, nr, k, idx, score, cosv, n_res, st);
}

template <int KP>
int run_mfma_bounded(hvk_ctx* ctx, const hvk_bundles& q, const hvk_bundles& r, int64_t ld, int k,
                     int32_t* idx, float* score, double* cosv, int32_t* n_res, hipStream_t st) {
    const int nq = q.n, nr = r.n;
    const int q_tiles = (nq + mfma_mt<1>() - 1) / mfma_mt<1>();
    const int r_tiles = (nr + kMfmaNT - 1) / kMfmaNT;
    const int S = choose_splits(ctx, q_tiles, r_tiles, 2);
    const int split_tiles = (r_tiles + S - 1) / S;
    const int S_used = (r_tiles + split_tiles - 1) / split_tiles;
    SearchBufs b;
    int rc_ = prepare(ctx, nq, nr, S_used, KP, b);
    if (rc_) return rc_;
    const int nmax = nq > nr ? nq : nr;
    hipLaunchKernelGGL(inv_norm_kernel, dim3((nmax + 255) / 256), dim3(256), 0, st, q.sqnorm, nq,
                       r.sqnorm, nr, b.iq, b.ir, b.flag_count,
                       BoundRows{q.maxabs, q.resid1, r.maxabs, r.resid1}, b.bq, b.br);
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL((mfma_kernel<1, 1, KP, true>), dim3(q_tiles, S_used), dim3(kMfmaThreads), 0,
                       st, q.clamp8, nullptr, b.iq, nq, r.clamp8, nullptr, b.ir, nr, ld,
                       split_tiles, b.part, b.bq, b.br, nullptr, nullptr);
    timing_mark(ctx, 1, true, st);
    HVK_HIP(hipGetLastError());
    const RowsLimb rows{q.lo, q.hi, r.lo, r.hi, ld};
    if (int rc = finalize<RowsLimb, KP>(ctx, rows, b, S_used, q.sqnorm, nq, r.sqnorm, nr, k, idx,
                                        score, cosv, st))
        return rc;
    // exact second pass over the queued queries: two-limb MFMA, KP = 32
    const bool q2 = q.hi != nullptr, r2 = r.hi != nullptr;
    const int cap = nq < kPass2Cap ? nq : kPass2Cap;
    const int tiles2 = (cap + mfma_mt<2>() - 1) / mfma_mt<2>();
    const int split2 = (r_tiles + kPass2Splits - 1) / kPass2Splits;
    const int S2 = (r_tiles + split2 - 1) / split2;
    const dim3 g2(tiles2, S2), blk(kMfmaThreads);
    if (q2 && r2)
        hipLaunchKernelGGL((mfma_kernel<2, 2, kPass2KP>), g2, blk, 0, st, q.lo, q.hi, b.iq, cap,
                           r.lo, r.hi, b.ir, nr, ld, split2, b.part2, nullptr, nullptr,
                           b.flag_list, b.flag_count);
    else if (q2)
        hipLaunchKernelGGL((mfma_kernel<2, 1, kPass2KP>), g2, blk, 0, st, q.lo, q.hi, b.iq, cap,
                           r.lo, nullptr, b.ir, nr, ld, split2, b.part2, nullptr, nullptr,
                           b.flag_list, b.flag_count);
    else if (r2)
        hipLaunchKernelGGL((mfma_kernel<1, 2, kPass2KP>), g2, blk, 0, st, q.lo, nullptr, b.iq, cap,
                           r.lo, r.hi, b.ir, nr, ld, split2, b.part2, nullptr, nullptr,
                           b.flag_list, b.flag_count);
    else
        hipLaunchKernelGGL((mfma_kernel<1, 1, kPass2KP>), g2, blk, 0, st, q.lo, nullptr, b.iq, cap,
                           r.lo, nullptr, b.ir, nr, ld, split2, b.part2, nullptr, nullptr,
                           b.flag_list, b.flag_count);
    HVK_HIP(hipGetLastError());
    hipLaunchKernelGGL((finalize_kernel<kPass2KP, RowsLimb>), dim3((nq + 3) / 4), dim3(256), 0, st,
                       rows, b.part2, S2, nq, q.sqnorm, r.sqnorm, nr, k, idx, score, cosv,
                       b.flag_list2, b.flag_count2, b.flag_list, b.flag_count, cap);
    HVK_HIP(hipGetLastError());
    return rescan(ctx, rows, b.flag_list2, b.flag_count2, b, q.sqnorm, r.sqnorm, nr, k, idx, score,
                  cosv, n_res, st);
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

}  // namespace
}  // namespace hvk

extern "C" int hvk_search_plain(hvk_ctx* ctx, const int16_t* d_q, const int64_t* d_q_sqnorm,
                                int n_q, int q_maxabs, const int16_t* d_r,
                                const int64_t* d_r_sqnorm, int n_r, int r_maxabs, int64_t ld,
                                int k, int32_t* d_idx, float* d_score, double* d_cos64,
                                int32_t* d_n_rescored, hvk_stream stream) {
    if (!ctx || !d_q || !d_q_sqnorm || !d_r || !d_r_sqnorm || !d_idx || !d_score) return HVK_ERR_ARG;
    if (n_q < 0 || n_r < 0 || k < 1 || k > 29 || q_maxabs < 0 || r_maxabs < 0) return HVK_ERR_ARG;
    if (ld == 0) ld = ctx->ld;
    if (ld < ctx->dim || ld % hvk::kPlainKS) return HVK_ERR_ARG;
    if ((int64_t)q_maxabs * r_maxabs * hvk::kPlainKS >= (int64_t(1) << 31)) return HVK_ERR_RANGE;
    if (n_q == 0) return HVK_OK;
    if (n_r == 0) return HVK_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = (hipStream_t)stream;
    switch (hvk::pick_kp(k)) {
        case 8: return hvk::run_plain<8>(ctx, d_q, d_q_sqnorm, n_q, d_r, d_r_sqnorm, n_r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
        case 16: return hvk::run_plain<16>(ctx, d_q, d_q_sqnorm, n_q, d_r, d_r_sqnorm, n_r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
        default: return hvk::run_plain<32>(ctx, d_q, d_q_sqnorm, n_q, d_r, d_r_sqnorm, n_r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
    }
}

extern "C" int hvk_search_mfma(hvk_ctx* ctx, const int8_t* d_q_lo, const int8_t* d_q_hi,
                               const int64_t* d_q_sqnorm, int n_q, const int8_t* d_r_lo,
                               const int8_t* d_r_hi, const int64_t* d_r_sqnorm, int n_r,
                               int64_t ld, int k, int32_t* d_idx, float* d_score,
                               double* d_cos64, int32_t* d_n_rescored, hvk_stream stream) {
    if (!ctx || !d_q_lo || !d_q_sqnorm || !d_r_lo || !d_r_sqnorm || !d_idx || !d_score)
        return HVK_ERR_ARG;
    if (n_q < 0 || n_r < 0 || k < 1 || k > 29) return HVK_ERR_ARG;
    if (ld == 0) ld = ctx->ld;
    if (ld < ctx->dim || ld % hvk::kMfmaKS) return HVK_ERR_ARG;
    if (n_q == 0) return HVK_OK;
    if (n_r == 0) return HVK_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = (hipStream_t)stream;
    const int kp = hvk::pick_kp(k);
    const bool q2 = d_q_hi != nullptr, r2 = d_r_hi != nullptr;
    if (!q2 && !r2) return hvk::dispatch_mfma_kp<1, 1>(ctx, kp, d_q_lo, nullptr, d_q_sqnorm, n_q, d_r_lo, nullptr, d_r_sqnorm, n_r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
    if (!q2 && r2) return hvk::dispatch_mfma_kp<1, 2>(ctx, kp, d_q_lo, nullptr, d_q_sqnorm, n_q, d_r_lo, d_r_hi, d_r_sqnorm, n_r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
    if (q2 && !r2) return hvk::dispatch_mfma_kp<2, 1>(ctx, kp, d_q_lo, d_q_hi, d_q_sqnorm, n_q, d_r_lo, nullptr, d_r_sqnorm, n_r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
    return hvk::dispatch_mfma_kp<2, 2>(ctx, kp, d_q_lo, d_q_hi, d_q_sqnorm, n_q, d_r_lo, d_r_hi, d_r_sqnorm, n_r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
}

extern "C" int hvk_search_mfma_bounded(hvk_ctx* ctx, const hvk_bundles* q, const hvk_bundles* r,
                                       int64_t ld, int k, int32_t* d_idx, float* d_score,
                                       double* d_cos64, int32_t* d_n_rescored, hvk_stream stream) {
    if (!ctx || !q || !r || !d_idx || !d_score) return HVK_ERR_ARG;
    for (const hvk_bundles* x : {q, r})
        if (!x->clamp8 || !x->lo || !x->sqnorm || !x->maxabs || !x->resid1 || x->n < 0)
            return HVK_ERR_ARG;
    if (k < 1 || k > 29) return HVK_ERR_ARG;
    if (ld == 0) ld = ctx->ld;
    if (ld < ctx->dim || ld % hvk::kMfmaKS) return HVK_ERR_ARG;
    if (q->n == 0) return HVK_OK;
    if (r->n == 0) return HVK_ERR_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipStream_t st = (hipStream_t)stream;
    // the bounded keys are upper bounds: a longer candidate list keeps the
    // margin check conclusive for more queries
    switch (hvk::pick_kp(k) * 2 > 32 ? 32 : hvk::pick_kp(k) * 2) {
        case 8: return hvk::run_mfma_bounded<8>(ctx, *q, *r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
        case 16: return hvk::run_mfma_bounded<16>(ctx, *q, *r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
        default: return hvk::run_mfma_bounded<32>(ctx, *q, *r, ld, k, d_idx, d_score, d_cos64, d_n_rescored, st);
    }
}
