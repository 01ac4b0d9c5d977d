// Plain copy kernel for bandwidth and counter calibration (frm_diag.h).
#include "frm_diag.h"
#include "frm_internal.hpp"

namespace frm {
__global__ __launch_bounds__(256) void k_copy16(const double2* __restrict__ src,
                                                double2* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}
}  // namespace frm

extern "C" int frm_diag_copy(frm_ctx* ctx, void* d_dst, const void* d_src, size_t bytes,
                             int reps, double* avg_ms) {
    FRM_REQUIRE(ctx, ctx && d_dst && d_src && reps > 0 && bytes % 16 == 0,
                "frm_diag_copy: bad arguments");
    const int64_t n = (int64_t)(bytes / 16);
    hipEvent_t a, b;
    FRM_HIP(ctx, hipEventCreate(&a));
    FRM_HIP(ctx, hipEventCreate(&b));
    const dim3 grid(256 * 64), block(256);
    hipLaunchKernelGGL(frm::k_copy16, grid, block, 0, ctx->stream, (const double2*)d_src,
                       (double2*)d_dst, n);  // warm
    FRM_HIP(ctx, hipEventRecord(a, ctx->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(frm::k_copy16, grid, block, 0, ctx->stream, (const double2*)d_src,
                           (double2*)d_dst, n);
    FRM_HIP(ctx, hipEventRecord(b, ctx->stream));
    FRM_HIP(ctx, hipEventSynchronize(b));
    float ms = 0.f;
    FRM_HIP(ctx, hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (avg_ms) *avg_ms = ms / reps;
    return FRM_OK;
}
