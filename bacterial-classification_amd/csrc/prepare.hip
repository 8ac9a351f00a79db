// prepare.hip -- search operands from int16 count rows supplied by a caller
// (HV_16 vectors of the C++ API rather than hvk_encode output): limbs, the
// clamped row, exact sum of squares, max |count| and the clamp residual.
// Replaces the norm pass of the reference's search, compute_inv_norms_kernel
// (src/cosine.cu:35-50), which computed float inverse norms on the fly.
#include "hvk_internal.h"

namespace hvk {
namespace {

// one workgroup (256 threads) per row; 8 counts per thread per step
__global__ void __launch_bounds__(256)
prepare_kernel(const int16_t* __restrict__ counts, int n, int dim, int64_t ld, EncodeOut out) {
    __shared__ int64_t r_sq[4];
    __shared__ int r_mx[4], r_res[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int row = blockIdx.x; row < n; row += gridDim.x) {
        int64_t sq = 0;
        int mx = 0, res = 0;
        const int16_t* src = counts + (int64_t)row * ld;
        for (int64_t c = tid; c < ld / 8; c += 256) {
            const uint4 v = reinterpret_cast<const uint4*>(src)[c];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
            uint32_t lo[2] = {0, 0}, hi[2] = {0, 0}, cl[2] = {0, 0};
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int d = (int)(c * 8 + e);
                int x = (int)(int16_t)(w[e >> 1] >> (16 * (e & 1)));
                if (d >= dim) x = 0;
                sq += (int64_t)x * x;
                const int a = x < 0 ? -x : x;
                mx = max(mx, a);
                res += a > 127 ? a - 127 : 0;
                const int l8 = (int)(int8_t)(x & 0xFF);
                const int h8 = (x - l8) >> 8;
                const int c8 = x > 127 ? 127 : (x < -127 ? -127 : x);
                lo[e >> 2] |= (uint32_t)(uint8_t)l8 << (8 * (e & 3));
                hi[e >> 2] |= (uint32_t)(uint8_t)h8 << (8 * (e & 3));
                cl[e >> 2] |= (uint32_t)(uint8_t)c8 << (8 * (e & 3));
            }
            const int64_t o = (int64_t)row * ld + c * 8;
            if (out.lo) *reinterpret_cast<uint2*>(out.lo + o) = make_uint2(lo[0], lo[1]);
            if (out.hi) *reinterpret_cast<uint2*>(out.hi + o) = make_uint2(hi[0], hi[1]);
            if (out.clamp8) *reinterpret_cast<uint2*>(out.clamp8 + o) = make_uint2(cl[0], cl[1]);
        }
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) {
            sq += (int64_t)(((uint64_t)(uint32_t)__shfl_xor((int)(sq >> 32), s, 64) << 32) |
                            (uint32_t)__shfl_xor((int)(uint32_t)sq, s, 64));
            mx = max(mx, __shfl_xor(mx, s, 64));
            res += __shfl_xor(res, s, 64);
        }
        if (lane == 0) { r_sq[wv] = sq; r_mx[wv] = mx; r_res[wv] = res; }
        __syncthreads();
        if (tid == 0) {
            const int64_t t = r_sq[0] + r_sq[1] + r_sq[2] + r_sq[3];
            if (out.sqnorm) out.sqnorm[row] = t;
            if (out.maxabs) out.maxabs[row] = max(max(r_mx[0], r_mx[1]), max(r_mx[2], r_mx[3]));
            if (out.resid1) out.resid1[row] = r_res[0] + r_res[1] + r_res[2] + r_res[3];
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace hvk

extern "C" int hvk_prepare_counts(hvk_ctx* ctx, const int16_t* d_counts, int n,
                                  const hvk_encode_out* out, hvk_stream stream) {
    if (!ctx || !d_counts || !out || n < 0) return HVK_ERR_ARG;
    if (out->counts) return HVK_ERR_ARG;  // counts are the input
    if (n == 0) return HVK_OK;
    hvk::EncodeOut o{};
    o.lo = out->lo;
    o.hi = out->hi;
    o.clamp8 = out->clamp8;
    o.ld = out->ld ? out->ld : ctx->ld;
    o.sqnorm = out->sqnorm;
    o.maxabs = out->maxabs;
    o.resid1 = out->resid1;
    if (o.ld < ctx->dim || o.ld % 8) return HVK_ERR_ARG;
    const int grid = n < 8 * ctx->n_cu ? n : 8 * ctx->n_cu;
    hipLaunchKernelGGL(hvk::prepare_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                       d_counts, n, ctx->dim, o.ld, o);
    HVK_HIP(hipGetLastError());
    return HVK_OK;
}
