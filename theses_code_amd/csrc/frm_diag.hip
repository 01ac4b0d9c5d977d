// Plain copy kernel for bandwidth and counter calibration, and a
// read-dominated stream (frm_diag.h).
#include "frm_diag.h"
#include "frm_internal.hpp"

namespace frm {
// The buffer read as rows of kReadCols 16-byte pieces (8 MiB); lane t sums
// column t over all rows with non-temporal loads, four rows in flight, and
// writes its sum once: the access pattern of the history sums (one pass per
// node column over every slot, all wavefronts walking the same rows
// together, 512-thread workgroups).
constexpr int64_t kReadCols = int64_t(1) << 19;
typedef double diag_dv2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(512) void k_read16(const diag_dv2* __restrict__ src,
                                                diag_dv2* __restrict__ dst, int64_t rows) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const diag_dv2* p = src + t;
    double ax = 0.0, ay = 0.0;
    int64_t r = 0;
    for (; r + 4 <= rows; r += 4) {
        diag_dv2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(p + (r + u) * kReadCols);
#pragma unroll
        for (int u = 0; u < 4; ++u) ax += v[u].x, ay += v[u].y;
    }
    for (; r < rows; ++r) {
        const diag_dv2 v = __builtin_nontemporal_load(p + r * kReadCols);
        ax += v.x, ay += v.y;
    }
    dst[t] = diag_dv2{ax, ay};
}
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

extern "C" int frm_diag_read(frm_ctx* ctx, const void* d_src, size_t bytes, int reps,
                             double* avg_ms) {
    FRM_REQUIRE(ctx, ctx && d_src && reps > 0 && bytes > 0 && bytes % (16 * frm::kReadCols) == 0,
                "frm_diag_read: bad arguments (bytes a multiple of 8 MiB)");
    const int64_t rows = (int64_t)(bytes / (16 * frm::kReadCols));
    const dim3 grid((unsigned)(frm::kReadCols / 512)), block(512);
    void* sums = nullptr;
    FRM_HIP(ctx, hipMalloc(&sums, (size_t)frm::kReadCols * 16));
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipFree(sums);
        return frm::check_hip(ctx, e, "frm_diag_read: hipEventCreate");
    }
    hipLaunchKernelGGL(frm::k_read16, grid, block, 0, ctx->stream, (const frm::diag_dv2*)d_src,
                       (frm::diag_dv2*)sums, rows);  // warm
    (void)hipEventRecord(a, ctx->stream);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(frm::k_read16, grid, block, 0, ctx->stream,
                           (const frm::diag_dv2*)d_src, (frm::diag_dv2*)sums, rows);
    (void)hipEventRecord(b, ctx->stream);
    e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(sums);
    if (e != hipSuccess) return frm::check_hip(ctx, e, "frm_diag_read");
    if (avg_ms) *avg_ms = ms / reps;
    return FRM_OK;
}
