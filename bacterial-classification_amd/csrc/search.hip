// search.hip -- exact cosine top-k of query bundles against reference bundles.
//
// Order: (c desc, reference index asc), c = dot / (sqrt(|q|^2) sqrt(|r|^2))
// in FP64 from the exact integer dot (the reference's cosine, src/HV.cpp:
// 219-242, before its 2c-1 map).  Pipeline per call:
//   1. inv_norm_kernel   FP32 inverse norms 1/sqrt(|v|^2) (correctly rounded)
//   2. plain_kernel | mfma_kernel
//        tile of MT queries x NT references, K-slabs staged through LDS; the
//        tile's exact int64 dots are scored in FP32, s = dot * iq * ir, and
//        offered to per-query LDS top lists keyed by (s, -index).  Each entry
//        carries its exact dot.  A workgroup sweeps a slice ("split") of the
//        references and writes its per-query top-KP list.
//   3. finalize_kernel   merges the split lists, re-scores the KP survivors
//        in FP64 from their exact dots, sorts, and checks the margin: every
//        excluded reference has s <= s_KP, so |s - c| <= 8 ulp guarantees
//        exactness when c_k > s_KP + 1e-6.  Queries failing the check are
//        queued for
//   4. rescan_kernel     an exact FP64 brute-force pass over all references.
#include <cmath>
#include <utility>

#include "hvk_internal.h"

namespace hvk {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef short v2s __attribute__((ext_vector_type(2)));

constexpr float kMarginEps = 1e-6f;

struct Cand {
    uint64_t key;  // (ord(s) << 32) | (~index); 0 = empty
    int64_t dot;
};

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

// ------------------------------------------------------------------------
// per-workgroup top lists in LDS
// ------------------------------------------------------------------------
template <int MT, int KP, int CAP>
struct TopLds {
    Cand top[MT][KP];  // descending by key
    Cand buf[MT][CAP];
    uint64_t thr[MT];  // top[m][KP-1].key
    int cnt[MT];
    int ovf[2];

    __device__ void init(int tid, int nthr) {
        for (int i = tid; i < MT * KP; i += nthr) top[i / KP][i % KP] = Cand{0, 0};
        for (int i = tid; i < MT; i += nthr) { thr[i] = 0; cnt[i] = 0; }
        if (tid == 0) { ovf[0] = 0; ovf[1] = 0; }
    }

    // Offers one (query m, reference) pair.  Returns true when the pair beat
    // the threshold but the buffer was full: the caller keeps it for the next
    // pass, which runs after the merge has raised the threshold.  A pair is
    // therefore offered at most once until it is accepted or falls below.
    __device__ __forceinline__ bool offer(int m, uint64_t key, int64_t dot, int pass) {
        if (key <= thr[m]) return false;
        const int slot = atomicAdd(&cnt[m], 1);
        if (slot < CAP) {
            buf[m][slot] = Cand{key, dot};
            return false;
        }
        ovf[pass & 1] = 1;
        return true;
    }

    __device__ void merge(int m) {
        const int n = cnt[m] < CAP ? cnt[m] : CAP;
        for (int q = 0; q < n; ++q) {
            const Cand c = buf[m][q];
            if (c.key <= top[m][KP - 1].key) continue;
            int p = KP - 1;
            while (p > 0 && top[m][p - 1].key < c.key) {
                top[m][p] = top[m][p - 1];
                --p;
            }
            top[m][p] = c;
        }
        thr[m] = top[m][KP - 1].key;
        cnt[m] = 0;
    }
};

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

// End of one offer pass: merge the buffers, report whether any pair was
// dropped (then the caller runs another pass over its dropped pairs).
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
__global__ void inv_norm_kernel(const int64_t* __restrict__ qn, int nq, const int64_t* __restrict__ rn,
                                int nr, float* __restrict__ iq, float* __restrict__ ir,
                                int32_t* __restrict__ flag_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *flag_count = 0;
    if (i < nq) {
        const int64_t v = qn[i];
        iq[i] = v == 0 ? 0.0f : __fdiv_rn(1.0f, __fsqrt_rn((float)v));
    }
    if (i < nr) {
        const int64_t v = rn[i];
        ir[i] = v == 0 ? 0.0f : __fdiv_rn(1.0f, __fsqrt_rn((float)v));
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
             int n_splits, int split_tiles, Cand* __restrict__ part) {
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
                    if (tl.offer(m, make_key(s, n), acc[i][j], pass)) dropped |= bit;
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
// 2b. MFMA form: v_mfma_i32_32x32x32_i8, 128 x 128 tile, 4 waves (2 x 2),
//     wave tile 64 x 64, K-slab of 64 bytes, XOR-swizzled LDS rows.
//     Queries are the A operand (rows), references the B operand (columns).
// ------------------------------------------------------------------------
constexpr int kMfmaMT = 128, kMfmaNT = 128, kMfmaKS = 64;

__device__ __forceinline__ int swz(int row, int chunk) {
    return row * 64 + ((chunk ^ ((row >> 1) & 3)) << 4);
}

template <int QL, int RL, int KP>
__global__ void __launch_bounds__(256, 1)
mfma_kernel(const int8_t* __restrict__ qlo, const int8_t* __restrict__ qhi,
            const float* __restrict__ iq, int nq, const int8_t* __restrict__ rlo,
            const int8_t* __restrict__ rhi, const float* __restrict__ ir, int nr, int64_t ld,
            int n_splits, int split_tiles, Cand* __restrict__ part) {
    constexpr int MT = kMfmaMT, NT = kMfmaNT, KS = kMfmaKS;
    constexpr int CAP = 16;
    __shared__ __attribute__((aligned(16))) int8_t sQ[QL][MT * KS];
    __shared__ __attribute__((aligned(16))) int8_t sR[RL][NT * KS];
    __shared__ TopLds<MT, KP, CAP> tl;
    __shared__ float s_iq[MT];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = blockIdx.x * MT;
    const int split = blockIdx.y;
    tl.init(tid, 256);
    if (tid < MT) s_iq[tid] = (m0 + tid < nq) ? iq[m0 + tid] : 0.0f;
    const int n_begin = split * split_tiles * NT;
    int n_end = n_begin + split_tiles * NT;
    if (n_end > nr) n_end = nr;
    const int8_t* qsrc[2] = {qlo, qhi};
    const int8_t* rsrc[2] = {rlo, rhi};
    __syncthreads();

    for (int n0 = n_begin; n0 < n_end; n0 += NT) {
        v16i acc[QL][RL][2][2];
#pragma unroll
        for (int a = 0; a < QL; ++a)
#pragma unroll
            for (int b = 0; b < RL; ++b)
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) acc[a][b][x][y] = v16i{};

        for (int64_t k0 = 0; k0 < ld; k0 += KS) {
#pragma unroll
            for (int l = 0; l < QL; ++l)
#pragma unroll
                for (int it = 0; it < 2; ++it) {
                    const int c = it * 256 + tid;
                    const int row = c >> 2, ch = c & 3;
                    v4i v = v4i{};
                    if (m0 + row < nq)
                        v = *reinterpret_cast<const v4i*>(qsrc[l] + (int64_t)(m0 + row) * ld + k0 + ch * 16);
                    *reinterpret_cast<v4i*>(&sQ[l][swz(row, ch)]) = v;
                }
#pragma unroll
            for (int l = 0; l < RL; ++l)
#pragma unroll
                for (int it = 0; it < 2; ++it) {
                    const int c = it * 256 + tid;
                    const int row = c >> 2, ch = c & 3;
                    v4i v = v4i{};
                    if (n0 + row < nr)
                        v = *reinterpret_cast<const v4i*>(rsrc[l] + (int64_t)(n0 + row) * ld + k0 + ch * 16);
                    *reinterpret_cast<v4i*>(&sR[l][swz(row, ch)]) = v;
                }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < KS / 32; ++ks) {
                const int ch = ks * 2 + (lane >> 5);
                v4i a[QL][2], b[RL][2];
#pragma unroll
                for (int l = 0; l < QL; ++l)
#pragma unroll
                    for (int x = 0; x < 2; ++x)
                        a[l][x] = *reinterpret_cast<const v4i*>(&sQ[l][swz(wm * 64 + x * 32 + (lane & 31), ch)]);
#pragma unroll
                for (int l = 0; l < RL; ++l)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
                        b[l][y] = *reinterpret_cast<const v4i*>(&sR[l][swz(wn * 64 + y * 32 + (lane & 31), ch)]);
#pragma unroll
                for (int qa = 0; qa < QL; ++qa)
#pragma unroll
                    for (int rb = 0; rb < RL; ++rb)
#pragma unroll
                        for (int x = 0; x < 2; ++x)
#pragma unroll
                            for (int y = 0; y < 2; ++y)
                                acc[qa][rb][x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                                    a[qa][x], b[rb][y], acc[qa][rb][x][y], 0, 0, 0);
            }
            __syncthreads();
        }

        float irv[2];
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int n = n0 + wn * 64 + y * 32 + (lane & 31);
            irv[y] = n < nr ? ir[n] : 0.0f;
        }
        uint64_t dropped = 0;  // bit 32*x+16*y+e: pair waits for another pass
        for (int pass = 0;; ++pass) {
            static_for<64>([&](auto I) {
                constexpr int x = decltype(I)::value / 32, y = (decltype(I)::value / 16) % 2,
                              e = decltype(I)::value % 16;
                constexpr uint64_t bit = 1ull << decltype(I)::value;
                const int n = n0 + wn * 64 + y * 32 + (lane & 31);
                const int m = wm * 64 + x * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                if ((pass == 0 || (dropped & bit)) && n < nr && m0 + m < nq) {
                    int64_t dot = 0;
                    static_for<QL>([&](auto QA) {
                        static_for<RL>([&](auto RB) {
                            constexpr int qa = decltype(QA)::value, rb = decltype(RB)::value;
                            dot += (int64_t)acc[qa][rb][x][y][e] << (8 * (qa + rb));
                        });
                    });
                    const float s = __fmul_rn(__fmul_rn((float)dot, s_iq[m]), irv[y]);
                    if (tl.offer(m, make_key(s, n), dot, pass)) dropped |= bit;
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
// 3. finalize: merge splits, FP64 re-score of the KP survivors, margin check
// ------------------------------------------------------------------------
template <int KP>
__global__ void __launch_bounds__(128)
finalize_kernel(const Cand* __restrict__ part, int n_splits, int nq,
                const int64_t* __restrict__ qn, const int64_t* __restrict__ rn, int nr, int k,
                int32_t* __restrict__ out_idx, float* __restrict__ out_score,
                double* __restrict__ out_cos, int32_t* __restrict__ flag_list,
                int32_t* __restrict__ flag_count) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nq) return;
    Cand best[KP];
#pragma unroll
    for (int j = 0; j < KP; ++j) best[j] = part[(int64_t)m * KP + j];
    for (int s = 1; s < n_splits; ++s) {
        for (int j = 0; j < KP; ++j) {
            const Cand c = part[((int64_t)s * nq + m) * KP + j];
            if (c.key <= best[KP - 1].key) break;  // lists are sorted
            int p = KP - 1;
            while (p > 0 && best[p - 1].key < c.key) { best[p] = best[p - 1]; --p; }
            best[p] = c;
        }
    }
    double cs[KP];
    int ix[KP];
    int nv = 0;
    const int64_t qnm = qn[m];
    for (int j = 0; j < KP; ++j) {
        if (best[j].key == 0) break;
        ix[nv] = idx_of_key(best[j].key);
        cs[nv] = cos64(best[j].dot, qnm, rn[ix[nv]]);
        ++nv;
    }
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
        const float s_kp = f32_of_ord((uint32_t)(best[KP - 1].key >> 32));
        flag = !(cs[k - 1] > (double)s_kp + (double)kMarginEps);
    }
    for (int j = 0; j < k; ++j) {
        const int64_t o = (int64_t)m * k + j;
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
    if (flag) flag_list[atomicAdd(flag_count, 1)] = m;
}

// ------------------------------------------------------------------------
// 4. exact re-scan of queued queries: FP64 brute force over all references
// ------------------------------------------------------------------------
template <bool LIMB>
__device__ __forceinline__ int value_at(const int16_t* c16, const int8_t* lo, const int8_t* hi,
                                        int64_t o) {
    if (LIMB) return (int)lo[o] + (hi ? 256 * (int)hi[o] : 0);
    return (int)c16[o];
}

template <bool LIMB>
__global__ void __launch_bounds__(256)
rescan_kernel(const int16_t* __restrict__ q16, const int8_t* __restrict__ qlo,
              const int8_t* __restrict__ qhi, const int64_t* __restrict__ qn,
              const int16_t* __restrict__ r16, const int8_t* __restrict__ rlo,
              const int8_t* __restrict__ rhi, const int64_t* __restrict__ rn, int nr, int64_t ld,
              int dim, int k, const int32_t* __restrict__ flag_list,
              const int32_t* __restrict__ flag_count, int32_t* __restrict__ out_idx,
              float* __restrict__ out_score, double* __restrict__ out_cos) {
    __shared__ double s_c[256 * kMaxTopK / 8];
    __shared__ int s_i[256 * kMaxTopK / 8];
    const int tid = threadIdx.x;
    const int nflag = *flag_count;
    for (int f = blockIdx.x; f < nflag; f += gridDim.x) {
        const int m = flag_list[f];
        double bc[kMaxTopK];
        int bi[kMaxTopK];
        int cnt = 0;
        for (int n = tid; n < nr; n += 256) {
            int64_t dot = 0;
            for (int d = 0; d < dim; ++d)
                dot += (int64_t)value_at<LIMB>(q16, qlo, qhi, (int64_t)m * ld + d) *
                       (int64_t)value_at<LIMB>(r16, rlo, rhi, (int64_t)n * ld + d);
            const double c = cos64(dot, qn[m], rn[n]);
            if (cnt == k && (c < bc[k - 1] || (c == bc[k - 1] && n > bi[k - 1]))) continue;
            int p = cnt < k ? cnt : k - 1;
            while (p > 0 && (bc[p - 1] < c || (bc[p - 1] == c && bi[p - 1] > n))) {
                bc[p] = bc[p - 1];
                bi[p] = bi[p - 1];
                --p;
            }
            bc[p] = c;
            bi[p] = n;
            if (cnt < k) ++cnt;
        }
        // merge the per-thread lists 8 threads at a time through LDS
        const int per = kMaxTopK / 8 < k ? kMaxTopK / 8 : k;  // unused guard
        (void)per;
        for (int j = cnt; j < k; ++j) { bc[j] = -INFINITY; bi[j] = 0x7FFFFFFF; }
        // sequential merge by thread 0 over all threads' lists, in rounds
        __shared__ double m_c[kMaxTopK];
        __shared__ int m_i[kMaxTopK];
        if (tid == 0)
            for (int j = 0; j < k; ++j) { m_c[j] = -INFINITY; m_i[j] = 0x7FFFFFFF; }
        __syncthreads();
        const int group = (256 * kMaxTopK / 8) / k;  // threads per round
        for (int g0 = 0; g0 < 256; g0 += group) {
            if (tid >= g0 && tid < g0 + group)
                for (int j = 0; j < k; ++j) {
                    s_c[(tid - g0) * k + j] = bc[j];
                    s_i[(tid - g0) * k + j] = bi[j];
                }
            __syncthreads();
            if (tid == 0) {
                const int lim = (256 - g0 < group ? 256 - g0 : group) * k;
                for (int e = 0; e < lim; ++e) {
                    const double c = s_c[e];
                    const int n = s_i[e];
                    if (c < m_c[k - 1] || (c == m_c[k - 1] && n > m_i[k - 1])) continue;
                    int p = k - 1;
                    while (p > 0 && (m_c[p - 1] < c || (m_c[p - 1] == c && m_i[p - 1] > n))) {
                        m_c[p] = m_c[p - 1];
                        m_i[p] = m_i[p - 1];
                        --p;
                    }
                    m_c[p] = c;
                    m_i[p] = n;
                }
            }
            __syncthreads();
        }
        if (tid < k) {
            const int64_t o = (int64_t)m * k + tid;
            const bool ok = m_i[tid] != 0x7FFFFFFF && m_i[tid] < nr;
            out_idx[o] = ok ? m_i[tid] : -1;
            out_score[o] = ok ? (float)m_c[tid] : -INFINITY;
            if (out_cos) out_cos[o] = ok ? m_c[tid] : -INFINITY;
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
    Cand* part;
    int32_t* flag_list;
    int32_t* flag_count;
};

int prepare(hvk_ctx* ctx, int nq, int nr, int n_splits, int kp, SearchBufs& b) {
    size_t need_n = sizeof(float) * ((size_t)nq + nr + 64);
    int st = ctx->ws_norms.ensure(need_n);
    if (st) return st;
    st = ctx->ws_search.ensure(sizeof(Cand) * (size_t)n_splits * nq * kp);
    if (st) return st;
    st = ctx->ws_flags.ensure(sizeof(int32_t) * ((size_t)nq + 64));
    if (st) return st;
    b.iq = (float*)ctx->ws_norms.ptr;
    b.ir = b.iq + ((nq + 63) / 64) * 64;
    b.part = (Cand*)ctx->ws_search.ptr;
    b.flag_count = (int32_t*)ctx->ws_flags.ptr;
    b.flag_list = b.flag_count + 64;
    return HVK_OK;
}

int choose_splits(hvk_ctx* ctx, int q_tiles, int r_tiles, int wg_per_cu) {
    const int want = 2 * ctx->n_cu * wg_per_cu;
    int s = (want + q_tiles - 1) / q_tiles;
    if (s < 1) s = 1;
    if (s > r_tiles) s = r_tiles;
    if (s > 64) s = 64;
    return s;
}

template <int KP>
int run_plain(hvk_ctx* ctx, const int16_t* q, const int64_t* qn, int nq, const int16_t* r,
              const int64_t* rn, int nr, int64_t ld, int k, int32_t* idx, float* score,
              double* cosv, int32_t* n_res, hipStream_t st) {
    const int q_tiles = (nq + kPlainMT - 1) / kPlainMT;
    const int r_tiles = (nr + kPlainNT - 1) / kPlainNT;
    const int S = choose_splits(ctx, q_tiles, r_tiles, 2);
    const int split_tiles = (r_tiles + S - 1) / S;
    const int S_used = (r_tiles + split_tiles - 1) / split_tiles;
    SearchBufs b;
    int rc = prepare(ctx, nq, nr, S_used, KP, b);
    if (rc) return rc;
    const int nmax = nq > nr ? nq : nr;
    hipLaunchKernelGGL(inv_norm_kernel, dim3((nmax + 255) / 256), dim3(256), 0, st, qn, nq, rn, nr,
                       b.iq, b.ir, b.flag_count);
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL((plain_kernel<KP>), dim3(q_tiles, S_used), dim3(256), 0, st, q, b.iq, nq, r,
                       b.ir, nr, ld, S_used, split_tiles, b.part);
    timing_mark(ctx, 1, true, st);
    hipLaunchKernelGGL((finalize_kernel<KP>), dim3((nq + 127) / 128), dim3(128), 0, st, b.part,
                       S_used, nq, qn, rn, nr, k, idx, score, cosv, b.flag_list, b.flag_count);
    hipLaunchKernelGGL((rescan_kernel<false>), dim3(ctx->n_cu), dim3(256), 0, st, q, nullptr,
                       nullptr, qn, r, nullptr, nullptr, rn, nr, ld, ctx->dim, k, b.flag_list,
                       b.flag_count, idx, score, cosv);
    HVK_HIP(hipGetLastError());
    if (n_res)
        HVK_HIP(hipMemcpyAsync(n_res, b.flag_count, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    return HVK_OK;
}

template <int QL, int RL, int KP>
int run_mfma(hvk_ctx* ctx, const int8_t* qlo, const int8_t* qhi, const int64_t* qn, int nq,
             const int8_t* rlo, const int8_t* rhi, const int64_t* rn, int nr, int64_t ld, int k,
             int32_t* idx, float* score, double* cosv, int32_t* n_res, hipStream_t st) {
    const int q_tiles = (nq + kMfmaMT - 1) / kMfmaMT;
    const int r_tiles = (nr + kMfmaNT - 1) / kMfmaNT;
    const int S = choose_splits(ctx, q_tiles, r_tiles, (QL * RL > 2) ? 1 : 2);
    const int split_tiles = (r_tiles + S - 1) / S;
    const int S_used = (r_tiles + split_tiles - 1) / split_tiles;
    SearchBufs b;
    int rc = prepare(ctx, nq, nr, S_used, KP, b);
    if (rc) return rc;
    const int nmax = nq > nr ? nq : nr;
    hipLaunchKernelGGL(inv_norm_kernel, dim3((nmax + 255) / 256), dim3(256), 0, st, qn, nq, rn, nr,
                       b.iq, b.ir, b.flag_count);
    timing_mark(ctx, 1, false, st);
    hipLaunchKernelGGL((mfma_kernel<QL, RL, KP>), dim3(q_tiles, S_used), dim3(256), 0, st, qlo, qhi,
                       b.iq, nq, rlo, rhi, b.ir, nr, ld, S_used, split_tiles, b.part);
    timing_mark(ctx, 1, true, st);
    hipLaunchKernelGGL((finalize_kernel<KP>), dim3((nq + 127) / 128), dim3(128), 0, st, b.part,
                       S_used, nq, qn, rn, nr, k, idx, score, cosv, b.flag_list, b.flag_count);
    hipLaunchKernelGGL((rescan_kernel<true>), dim3(ctx->n_cu), dim3(256), 0, st, nullptr, qlo, qhi,
                       qn, nullptr, rlo, rhi, rn, nr, ld, ctx->dim, k, b.flag_list, b.flag_count,
                       idx, score, cosv);
    HVK_HIP(hipGetLastError());
    if (n_res)
        HVK_HIP(hipMemcpyAsync(n_res, b.flag_count, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    return HVK_OK;
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
    const int kp = hvk::pick_kp(k);
    switch (kp) {
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
