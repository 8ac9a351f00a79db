// select.hip -- exact top-k for any k up to the number of references (the
// candidate-list kernels of the search keep at most kListMaxK results).
//
// The reference's launch_cosine_topk (src/cosine.cu:441-496) writes the full
// score matrix and runs a k-pass selection per query; it accepts any
// k <= num_db.  Here, per chunk of queries:
//   1. the plain or MFMA search kernel in its key-matrix mode writes, for
//      every (query, reference) pair, the FP64 cosine of the exact integer
//      dot as an order-preserving 64-bit key;
//   2. select_kernel, one workgroup per query row: a radix select (8 passes
//      of 8-bit digits) finds the k-th largest key T and how many of the keys
//      equal to T belong to the top k; the selected entries are compacted in
//      reference-index order (keys > T, then the first ties at T), sorted by
//      a stable LSD radix sort on the inverted key (4-bit digits, digits that
//      every key shares skipped), so the order is (c desc, index asc), and
//      written out.  Every value is exact: nothing is re-scored.
#include "search_impl.h"

namespace hvk {

namespace {

constexpr int kSelThreads = 256;
constexpr int kSelWaves = kSelThreads / 64;

// Exclusive prefix of a per-thread flag over the workgroup and its total.
// Every thread of the workgroup calls it (it holds barriers).
__device__ __forceinline__ int block_excl_flag(bool f, int* s_w, int& total) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t b = __ballot(f);
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0));
    if (lane == 0) s_w[wv] = __popcll(b);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kSelWaves; ++w) {
        const int c = s_w[w];
        off += w < wv ? c : 0;
        tot += c;
    }
    __syncthreads();  // s_w is reused by the next call
    total = tot;
    return off + below;
}

__global__ void __launch_bounds__(kSelThreads)
select_kernel(const uint64_t* __restrict__ mat, int nrows, int nr, int k_eff, int k, int row0,
              uint64_t* __restrict__ key_a, uint64_t* __restrict__ key_b,
              int32_t* __restrict__ idx_a, int32_t* __restrict__ idx_b,
              int32_t* __restrict__ out_idx, float* __restrict__ out_score,
              double* __restrict__ out_cos) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t cnt[16][kSelThreads];
    __shared__ int32_t scan[kSelThreads];
    __shared__ int s_w[kSelWaves];
    __shared__ uint64_t s_prefix;
    __shared__ int s_kk;
    __shared__ uint64_t s_and[kSelThreads], s_or[kSelThreads];
    const int tid = threadIdx.x;

    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const uint64_t* row = mat + (int64_t)r * nr;
        // ---- radix select of the k_eff-th largest key ----
        uint64_t prefix = 0, mask = 0;
        int kk = k_eff;  // rank of the wanted key among those matching prefix
        for (int shift = 56; shift >= 0; shift -= 8) {
            hist[tid] = 0;
            __syncthreads();
            for (int e = tid; e < nr; e += kSelThreads) {
                const uint64_t key = row[e];
                if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1u);
            }
            __syncthreads();
            if (tid == 0) {
                int above = 0, d = 255;
                for (; d > 0; --d) {
                    if (above + (int)hist[d] >= kk) break;
                    above += (int)hist[d];
                }
                s_kk = kk - above;
                s_prefix = prefix | ((uint64_t)d << shift);
            }
            __syncthreads();
            kk = s_kk;
            prefix = s_prefix;
            mask |= 0xFFull << shift;
        }
        const uint64_t T = prefix;
        const int need = kk;  // keys equal to T that belong to the top k_eff

        // ---- compaction in index order: keys > T, and the first `need` ties ----
        uint64_t* ka = key_a + (int64_t)r * k_eff;
        int32_t* ia = idx_a + (int64_t)r * k_eff;
        int out_n = 0, ties = 0;
        for (int base = 0; base < nr; base += kSelThreads) {
            const int e = base + tid;
            const uint64_t key = e < nr ? row[e] : 0;
            const bool gt = e < nr && key > T;
            const bool eq = e < nr && key == T;
            int n_eq = 0, n_sel = 0;
            const int tie_rank = ties + block_excl_flag(eq, s_w, n_eq);
            const bool sel = gt || (eq && tie_rank < need);
            const int pos = out_n + block_excl_flag(sel, s_w, n_sel);
            if (sel) {
                ka[pos] = key;
                ia[pos] = e;
            }
            ties += n_eq;
            out_n += n_sel;
        }
        __syncthreads();

        // ---- stable LSD radix sort by ~key (descending key, index ascending) ----
        uint64_t a_and = ~0ull, a_or = 0;
        for (int i = tid; i < k_eff; i += kSelThreads) {
            const uint64_t x = ka[i];
            a_and &= x;
            a_or |= x;
        }
        s_and[tid] = a_and;
        s_or[tid] = a_or;
        __syncthreads();
        for (int o = kSelThreads / 2; o > 0; o >>= 1) {
            if (tid < o) {
                s_and[tid] &= s_and[tid + o];
                s_or[tid] |= s_or[tid + o];
            }
            __syncthreads();
        }
        const uint64_t varies = s_and[0] ^ s_or[0];
        __syncthreads();
        uint64_t* ks = ka;
        int32_t* is = ia;
        uint64_t* kd = key_b + (int64_t)r * k_eff;
        int32_t* id = idx_b + (int64_t)r * k_eff;
        const int C = (k_eff + kSelThreads - 1) / kSelThreads;  // items per thread
        const int c0 = min(tid * C, k_eff), c1 = min(c0 + C, k_eff);
        for (int dg = 0; dg < 16; ++dg) {
            const int sh = 4 * dg;
            if (((varies >> sh) & 15) == 0) continue;  // uniform: every key shares it
#pragma unroll
            for (int d = 0; d < 16; ++d) cnt[d][tid] = 0;
            for (int i = c0; i < c1; ++i) ++cnt[(~ks[i] >> sh) & 15][tid];
            __syncthreads();
            // exclusive scan over (digit, thread): thread t owns flat entries
            // [16 t, 16 t + 16) of cnt viewed as [16][256]
            uint32_t* flat = &cnt[0][0];
            int s = 0;
#pragma unroll
            for (int j = 0; j < 16; ++j) s += (int)flat[16 * tid + j];
            scan[tid] = s;
            __syncthreads();
            for (int o = 1; o < kSelThreads; o <<= 1) {
                const int v = tid >= o ? scan[tid - o] : 0;
                __syncthreads();
                scan[tid] += v;
                __syncthreads();
            }
            int run = scan[tid] - s;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int c = (int)flat[16 * tid + j];
                flat[16 * tid + j] = (uint32_t)run;
                run += c;
            }
            __syncthreads();
            for (int i = c0; i < c1; ++i) {
                const uint64_t x = ks[i];
                const uint32_t p = cnt[(~x >> sh) & 15][tid]++;
                kd[p] = x;
                id[p] = is[i];
            }
            __syncthreads();
            uint64_t* tk = ks; ks = kd; kd = tk;
            int32_t* ti = is; is = id; id = ti;
        }

        // ---- outputs: [k] per query, -1 / -inf past the references ----
        const int64_t o = (int64_t)(row0 + r) * k;
        for (int j = tid; j < k; j += kSelThreads) {
            if (j < k_eff) {
                const double c = f64_of_ord(ks[j]);
                out_idx[o + j] = is[j];
                out_score[o + j] = (float)c;
                if (out_cos) out_cos[o + j] = c;
            } else {
                out_idx[o + j] = -1;
                out_score[o + j] = -INFINITY;
                if (out_cos) out_cos[o + j] = -INFINITY;
            }
        }
        __syncthreads();
    }
}

__global__ void zero_i32(int32_t* p) {
    if (threadIdx.x == 0) *p = 0;
}

// Rows of queries per chunk: the key matrix and the sort buffers of a chunk
// stay within the budget (1 GiB unless tuned; at least one row), in whole
// tiles of rows when a tile fits.
int chunk_rows(const hvk_ctx* ctx, int nq, int nr, int k_eff, int tile) {
    const size_t budget = ctx->tune_select_kib > 0 ? (size_t)ctx->tune_select_kib << 10
                                                   : size_t(1) << 30;
    const size_t per_row = (size_t)nr * 8 + (size_t)k_eff * 24;
    size_t rows = budget / per_row;
    if (rows >= (size_t)tile) rows = rows / tile * tile;
    if (rows < 1) rows = 1;
    return rows < (size_t)nq ? (int)rows : nq;
}

struct SelBufs {
    uint64_t* mat;
    uint64_t *key_a, *key_b;
    int32_t *idx_a, *idx_b;
};

int select_bufs(hvk_ctx* ctx, int qc, int nr, int k_eff, SelBufs& b) {
    const size_t mat = (size_t)qc * nr * 8, keys = (size_t)qc * k_eff * 8,
                 idx = (size_t)qc * k_eff * 4;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    if (int rc = ctx->ws_select.ensure(up(mat) + 2 * up(keys) + 2 * up(idx))) return rc;
    char* w = (char*)ctx->ws_select.ptr;
    b.mat = (uint64_t*)w;
    w += up(mat);
    b.key_a = (uint64_t*)w;
    w += up(keys);
    b.key_b = (uint64_t*)w;
    w += up(keys);
    b.idx_a = (int32_t*)w;
    w += up(idx);
    b.idx_b = (int32_t*)w;
    return HVK_OK;
}

void launch_select(hvk_ctx* ctx, const SelBufs& b, int qc, int nr, int k_eff, int k, int row0,
                   int32_t* idx, float* score, double* cosv, hipStream_t st) {
    const int grid = qc < 4 * ctx->n_cu ? qc : 4 * ctx->n_cu;
    hipLaunchKernelGGL(select_kernel, dim3(grid), dim3(kSelThreads), 0, st, b.mat, qc, nr, k_eff,
                       k, row0, b.key_a, b.key_b, b.idx_a, b.idx_b, idx, score, cosv);
}

template <int QL, int RL>
int select_limb_impl(hvk_ctx* ctx, const int8_t* qlo, const int8_t* qhi, const int64_t* qn, int nq,
                     const int8_t* rlo, const int8_t* rhi, const int64_t* rn, int nr, int64_t ld,
                     int k, int32_t* idx, float* score, double* cosv, int32_t* n_res,
                     hipStream_t st) {
    using T = MfmaTile<QL, RL>;
    const int k_eff = k < nr ? k : nr;
    const int qc_max = chunk_rows(ctx, nq, nr, k_eff, T::MT);
    SelBufs b;
    if (int rc = select_bufs(ctx, qc_max, nr, k_eff, b)) return rc;
    timing_mark(ctx, 1, false, st);
    for (int m0 = 0; m0 < nq; m0 += qc_max) {
        const int qc = nq - m0 < qc_max ? nq - m0 : qc_max;
        const MfmaPlan pl = plan_mfma(ctx, (qc + T::MT - 1) / T::MT, (nr + T::NT - 1) / T::NT, 0);
        hipLaunchKernelGGL((mfma_kernel<QL, RL, 8, false, false, true>), dim3(pl.blocks),
                           dim3(kMfmaThreads), 0, st, qlo + (int64_t)m0 * ld,
                           QL > 1 ? qhi + (int64_t)m0 * ld : nullptr, nullptr, qc, rlo,
                           RL > 1 ? rhi : nullptr, nullptr, nr, ld, pl.map, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, qn + m0, rn, b.mat);
        HVK_HIP(hipGetLastError());
        launch_select(ctx, b, qc, nr, k_eff, k, m0, idx, score, cosv, st);
        HVK_HIP(hipGetLastError());
    }
    timing_mark(ctx, 1, true, st);
    if (n_res) hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, st, n_res);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

}  // namespace

int select_topk_plain(hvk_ctx* ctx, const int16_t* q, const int64_t* qn, int nq, const int16_t* r,
                      const int64_t* rn, int nr, int64_t ld, int k, int32_t* idx, float* score,
                      double* cosv, int32_t* n_res, hipStream_t st) {
    const int k_eff = k < nr ? k : nr;
    const int qc_max = chunk_rows(ctx, nq, nr, k_eff, kPlainMT);
    SelBufs b;
    if (int rc = select_bufs(ctx, qc_max, nr, k_eff, b)) return rc;
    const int r_tiles = (nr + kPlainNT - 1) / kPlainNT;
    timing_mark(ctx, 1, false, st);
    for (int m0 = 0; m0 < nq; m0 += qc_max) {
        const int qc = nq - m0 < qc_max ? nq - m0 : qc_max;
        const int q_tiles = (qc + kPlainMT - 1) / kPlainMT;
        const int S = choose_splits(ctx, q_tiles, r_tiles, 4);
        const int split_tiles = (r_tiles + S - 1) / S;
        const int S_used = (r_tiles + split_tiles - 1) / split_tiles;
        hipLaunchKernelGGL((plain_kernel<8, true>), dim3(q_tiles, S_used), dim3(256), 0, st,
                           q + (int64_t)m0 * ld, nullptr, qc, r, nullptr, nr, ld, split_tiles,
                           nullptr, qn + m0, rn, b.mat);
        HVK_HIP(hipGetLastError());
        launch_select(ctx, b, qc, nr, k_eff, k, m0, idx, score, cosv, st);
        HVK_HIP(hipGetLastError());
    }
    timing_mark(ctx, 1, true, st);
    if (n_res) hipLaunchKernelGGL(zero_i32, dim3(1), dim3(64), 0, st, n_res);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}

int select_topk_limb(hvk_ctx* ctx, const int8_t* qlo, const int8_t* qhi, const int64_t* qn, int nq,
                     const int8_t* rlo, const int8_t* rhi, const int64_t* rn, int nr, int64_t ld,
                     int k, int32_t* idx, float* score, double* cosv, int32_t* n_res,
                     hipStream_t st) {
    if (!qhi && !rhi)
        return select_limb_impl<1, 1>(ctx, qlo, qhi, qn, nq, rlo, rhi, rn, nr, ld, k, idx, score, cosv, n_res, st);
    if (!qhi)
        return select_limb_impl<1, 2>(ctx, qlo, qhi, qn, nq, rlo, rhi, rn, nr, ld, k, idx, score, cosv, n_res, st);
    if (!rhi)
        return select_limb_impl<2, 1>(ctx, qlo, qhi, qn, nq, rlo, rhi, rn, nr, ld, k, idx, score, cosv, n_res, st);
    return select_limb_impl<2, 2>(ctx, qlo, qhi, qn, nq, rlo, rhi, rn, nr, ld, k, idx, score, cosv, n_res, st);
}

}  // namespace hvk
