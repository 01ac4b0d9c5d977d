// Plain copy kernel for bandwidth and counter calibration, and a
// read-dominated stream (frm_diag.h).
#include "frm_diag.h"
#include "frm_internal.hpp"

namespace frm {
// The buffer read as rows of kReadCols 4-byte columns (8 MiB); lane t sums
// column t over all rows with non-temporal loads, 16 rows in flight, and
// writes its sum once (8 MiB of sums).  Measured over 4 GiB on MI355X, the
// narrow columns with many rows in flight read fastest: 6.5-6.6 TB/s,
// against 6.3 for 8-byte columns (8 rows in flight) and 5.9-6.1 for
// 16-byte columns (4 to 16 rows), the layout of the history sums' loads.
constexpr int64_t kReadCols = int64_t(1) << 21;
constexpr int kReadRows = 16;  // rows in flight per lane
__global__ __launch_bounds__(256) void k_read4(const float* __restrict__ src,
                                               float* __restrict__ dst, int64_t rows) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float* p = src + t;
    float a = 0.f;
    int64_t r = 0;
    for (; r + kReadRows <= rows; r += kReadRows) {
        float v[kReadRows];
#pragma unroll
        for (int u = 0; u < kReadRows; ++u) v[u] = __builtin_nontemporal_load(p + (r + u) * kReadCols);
#pragma unroll
        for (int u = 0; u < kReadRows; ++u) a += v[u];
    }
    for (; r < rows; ++r) a += __builtin_nontemporal_load(p + r * kReadCols);
    dst[t] = a;
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
    FRM_REQUIRE(ctx, ctx && d_src && reps > 0 && bytes > 0 && bytes % (4 * frm::kReadCols) == 0,
                "frm_diag_read: bad arguments (bytes a multiple of 8 MiB)");
    const int64_t rows = (int64_t)(bytes / (4 * frm::kReadCols));
    const dim3 grid((unsigned)(frm::kReadCols / 256)), block(256);
    void* sums = nullptr;
    FRM_HIP(ctx, hipMalloc(&sums, (size_t)frm::kReadCols * 4));
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipFree(sums);
        return frm::check_hip(ctx, e, "frm_diag_read: hipEventCreate");
    }
    hipLaunchKernelGGL(frm::k_read4, grid, block, 0, ctx->stream, (const float*)d_src,
                       (float*)sums, rows);  // warm
    (void)hipEventRecord(a, ctx->stream);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(frm::k_read4, grid, block, 0, ctx->stream, (const float*)d_src,
                           (float*)sums, rows);
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
