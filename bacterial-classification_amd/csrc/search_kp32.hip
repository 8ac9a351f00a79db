// search_kp32.hip -- search run functions with 32-entry candidate lists.
#include "search_impl.h"

namespace hvk {

HVK_SEARCH_RUNS(32, )
HVK_SEARCH_BOUNDED(32, )

// exact second pass of the bounded form over the queued queries (indirect
// rows, device-side count): MFMA over the stored limbs, KP = kPass2KP
template <int QL, int RL>
void launch_pass2(hvk_ctx* ctx, const hvk_bundles& q, const hvk_bundles& r, int64_t ld, int cap,
                  const SearchBufs& b, int& n_splits, hipStream_t st) {
    using T = MfmaTile<QL, RL>;
    const MfmaPlan pl = plan_mfma(ctx, (cap + T::MT - 1) / T::MT, (r.n + T::NT - 1) / T::NT,
                                  kPass2Splits);
    n_splits = pl.map.n_splits;
    hipLaunchKernelGGL((mfma_kernel<QL, RL, kPass2KP, false, true>), dim3(pl.blocks),
                       dim3(kMfmaThreads), 0, st,
                       q.lo, QL > 1 ? q.hi : nullptr, b.iq, cap, r.lo, RL > 1 ? r.hi : nullptr,
                       b.ir, r.n, ld, pl.map, b.part2, nullptr, nullptr, b.flag_list,
                       b.flag_count, nullptr);
}

int exact_pass2(hvk_ctx* ctx, const hvk_bundles& q, const hvk_bundles& r, int64_t ld, int k,
                const SearchBufs& b, int32_t* idx, float* score, double* cosv, int32_t* n_res,
                hipStream_t st) {
    const int nq = q.n, nr = r.n;
    const RowsLimb rows{q.lo, q.hi, r.lo, r.hi, ld};
    const bool q2 = q.hi != nullptr, r2 = r.hi != nullptr;
    const int cap = nq < kPass2Cap ? nq : kPass2Cap;
    int S2 = 1;
    if (q2 && r2) launch_pass2<2, 2>(ctx, q, r, ld, cap, b, S2, st);
    else if (q2) launch_pass2<2, 1>(ctx, q, r, ld, cap, b, S2, st);
    else if (r2) launch_pass2<1, 2>(ctx, q, r, ld, cap, b, S2, st);
    else launch_pass2<1, 1>(ctx, q, r, ld, cap, b, S2, st);
    HVK_HIP(hipGetLastError());
    hipLaunchKernelGGL((finalize_kernel<kPass2KP, RowsLimb>), dim3((nq + 3) / 4), dim3(256), 0, st,
                       rows, b.part2, S2, nq, q.sqnorm, r.sqnorm, nr, k, idx, score, cosv,
                       b.flag_list2, b.flag_count2, b.flag_list, b.flag_count, cap, nullptr);
    HVK_HIP(hipGetLastError());
    return rescan(ctx, rows, b.flag_list2, b.flag_count2, b, q.sqnorm, r.sqnorm, nr, k, idx, score,
                  cosv, n_res, st);
}

}  // namespace hvk
