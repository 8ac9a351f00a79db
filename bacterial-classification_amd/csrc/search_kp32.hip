// search_kp32.hip -- search run functions with 32-entry candidate lists.
#include "search_impl.h"

namespace hvk {

HVK_SEARCH_RUNS(32, )
HVK_SEARCH_BOUNDED(32, )

// second pass of the bounded form over the queued queries (indirect rows,
// device-side count): the same one-limb bounded MFMA (clamped rows, exact
// bound terms bq / br of the first pass), KP = kPass2KP lists.  Its keys are
// upper bounds like the first pass's, so finalize settles a query once its
// k-th exact score clears the list's last key, and queues the rest for the
// FP64 re-scan.
void launch_pass2(hvk_ctx* ctx, const hvk_bundles& q, const hvk_bundles& r, int64_t ld, int cap,
                  const SearchBufs& b, int& n_splits, hipStream_t st) {
    using T = MfmaTile<1, 1>;
    const MfmaPlan pl = plan_mfma(ctx, (cap + T::MT - 1) / T::MT, (r.n + T::NT - 1) / T::NT,
                                  kPass2Splits);
    n_splits = pl.map.n_splits;
    hipLaunchKernelGGL((mfma_kernel<1, 1, kPass2KP, true, true>), dim3(pl.blocks),
                       dim3(kMfmaThreads), 0, st, q.clamp8, nullptr, b.iq, cap, r.clamp8, nullptr,
                       b.ir, r.n, ld, pl.map, b.part2, b.bq, b.br, b.flag_list, b.flag_count,
                       nullptr);
}

int exact_pass2(hvk_ctx* ctx, const hvk_bundles& q, const hvk_bundles& r, int64_t ld, int k,
                const SearchBufs& b, int32_t* idx, float* score, double* cosv, int32_t* n_res,
                hipStream_t st) {
    const int nq = q.n, nr = r.n;
    const RowsLimb rows{q.lo, q.hi, r.lo, r.hi, ld};
    const int cap = nq < kPass2Cap ? nq : kPass2Cap;
    int S2 = 1;
    launch_pass2(ctx, q, r, ld, cap, b, S2, st);
    HVK_HIP(hipGetLastError());
    hipLaunchKernelGGL((finalize_kernel<kPass2KP, RowsLimb>), dim3((nq + 3) / 4), dim3(256), 0, st,
                       rows, b.part2, S2, nq, q.sqnorm, r.sqnorm, nr, k, idx, score, cosv,
                       b.flag_list2, b.flag_count2, b.flag_list, b.flag_count, cap, nullptr);
    HVK_HIP(hipGetLastError());
    return rescan(ctx, rows, b.flag_list2, b.flag_count2, b, q.sqnorm, r.sqnorm, nr, k, idx, score,
                  cosv, n_res, st);
}

}  // namespace hvk
