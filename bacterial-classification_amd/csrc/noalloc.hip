// noalloc.hip -- hvk_search_mfma_bounded_noalloc: the bounded MFMA search
// with its workspace in a fixed device array, for callers that cannot
// allocate device memory (code under hipGraph capture, serving loops that
// forbid hipMalloc).  Nothing here allocates: the search buffers (inverse
// norms, candidate lists, queues, the second pass's lists) are laid
// out in `scratch`, a file-scope device array each GPU holds once.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; the kernels clear what must start at zero (queue counters, shared
// thresholds, order histogram) on every call.  Calls that overlap on one
// GPU (two streams, two threads) race on the array: serialise them.
#include "search_impl.h"

namespace hvk {
namespace {

constexpr size_t kScratchBytes = HVK_NOALLOC_SCRATCH_BYTES;  // per GPU, held until the process ends
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// scratch's address on each device (each GPU holds its own copy of the array)
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Resolves scratch's address on `device` (the current device) once.
int scratch_on(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return HVK_ERR_NODEVICE;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            set_last_error("hipGetSymbolAddress(scratch)", e);
            return HVK_ERR_HIP;
        }
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return HVK_OK;
}

inline size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

// Called by hvk_ctx_create with the context's device current.
int noalloc_setup(int device) {
    void* p = nullptr;
    return scratch_on(device, &p);
}

}  // namespace hvk

extern "C" int hvk_search_mfma_bounded_noalloc(hvk_ctx* ctx, const hvk_bundles* q,
                                               const hvk_bundles* r, int64_t ld, int k,
                                               int32_t* d_idx, float* d_score, double* d_cos64,
                                               int32_t* d_n_rescored, hvk_stream stream) {
    using namespace hvk;
    if (!ctx || !q || !r || !d_idx || !d_score) return HVK_ERR_ARG;
    for (const hvk_bundles* x : {q, r})
        if (!x->clamp8 || !x->lo || !x->sqnorm || !x->maxabs || !x->resid1 || x->n < 0)
            return HVK_ERR_ARG;
    if (k < 1 || k > kListMaxK) return HVK_ERR_ARG;
    if (ld == 0) ld = ctx->ld;
    if (ld < ctx->dim || ld % kMfmaKS) return HVK_ERR_ARG;
    if (ld > (int64_t(1) << 24)) return HVK_ERR_RANGE;
    if (q->n == 0) return HVK_OK;
    if (r->n == 0) return HVK_ERR_ARG;
    void* base = nullptr;
    if (int rc = scratch_on(ctx->device, &base)) return rc;
    std::lock_guard<std::mutex> g(ctx->mu);
    const int kp = pick_kp(k);
    const MfmaPlan pl = plan_bounded(ctx, q->n, r->n);
    const SearchSizes z = search_sizes(q->n, r->n, pl.map.n_splits, kp);
    const size_t need = up256(z.norms) + up256(z.part) + up256(z.flags) + up256(z.part2);
    if (need > kScratchBytes) {
        set_last_message("hvk_search_mfma_bounded_noalloc: these sizes need more than the "
                         "HVK_NOALLOC_SCRATCH_BYTES scratch array");
        return HVK_ERR_RANGE;
    }
    char* p = (char*)base;
    SearchBufs b;
    carve(b, p, p + up256(z.norms), p + up256(z.norms) + up256(z.part),
          p + up256(z.norms) + up256(z.part) + up256(z.flags), q->n, r->n);
    hipStream_t st = (hipStream_t)stream;
    switch (kp) {
        case 8: return launch_mfma_bounded<8>(ctx, pl, *q, *r, ld, k, b, d_idx, d_score, d_cos64, d_n_rescored, st);
        case 16: return launch_mfma_bounded<16>(ctx, pl, *q, *r, ld, k, b, d_idx, d_score, d_cos64, d_n_rescored, st);
        default: return launch_mfma_bounded<32>(ctx, pl, *q, *r, ld, k, b, d_idx, d_score, d_cos64, d_n_rescored, st);
    }
}
