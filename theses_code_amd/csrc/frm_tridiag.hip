// Batched general tridiagonal solver (Thomas / progonka), one wavefront per
// system.  Replaces the reference's al1 / bt1 / back-substitution triple
// (mt_minimal.cpp:829-858, 942-1024; fr_d_t_g_2d.cpp:3267-3304, 3494-3553)
// and its OpenCL Al/Bt/U kernels (fr_d_t_g_2d.cpp:1604-1658, 1896-1922),
// which run one work-item per system.  Each lane owns a contiguous chunk of
// the system; the three Thomas recurrences (alpha: Moebius, beta and
// back-substitution: affine) get their chunk-entry values from wavefront
// scans, then every lane replays the sequential formulas over its chunk.
#include "frm_internal.hpp"
#include "frm_wavescan.hpp"

namespace frm {

struct TriArgs {
    const double *a, *b, *c, *d;
    double* x;
    int nsys, n;
    int64_t ss, se;
};

template <int CM>
__global__ __launch_bounds__(256) void k_tridiag(TriArgs p) {
    const int lane = threadIdx.x & 63;
    const int sys = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (sys >= p.nsys) return;
    const int n = p.n;
    const int e0 = lane * CM;
    double a[CM], b[CM], c[CM], v[CM];
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        const int e = e0 + k;
        if (e < n) {
            const int64_t o = (int64_t)sys * p.ss + (int64_t)e * p.se;
            a[k] = e == 0 ? 0.0 : p.a[o];
            b[k] = p.b[o];
            c[k] = e == n - 1 ? 0.0 : p.c[o];
            v[k] = p.d[o];
        } else {
            a[k] = 0.0;
            b[k] = 1.0;
            c[k] = 0.0;
            v[k] = 0.0;
        }
    }
    // alpha: cp_k = c_k / (b_k - a_k cp_{k-1}), cp_{-1} = 0
    Mob m{1.0, 0.0, 0.0, 1.0};
#pragma unroll
    for (int k = 0; k < CM; ++k)
        if (e0 + k < n) m = mob_mul(Mob{0.0, c[k], -a[k], b[k]}, m);
    const Mob pre = scan_mob_up(m, lane);
    double cp = pre.m01 / pre.m11;  // applied to (0, 1)
    // reuse c[] for cp_k and b[] for 1/(b_k - a_k cp_{k-1})
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        if (e0 + k < n) {
            const double den = 1.0 / (b[k] - a[k] * cp);
            cp = c[k] * den;
            c[k] = cp;
            b[k] = den;
        }
    }
    // beta: dp_k = (d_k - a_k dp_{k-1}) / den_k
    double P = 1.0, Q = 0.0;
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        if (e0 + k < n) {
            const double pk = -a[k] * b[k], qk = v[k] * b[k];
            Q = pk * Q + qk;
            P = pk * P;
        }
    }
    scan_affine_up(P, Q, lane);
    double y = Q;  // dp_{-1} = 0
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        if (e0 + k < n) {
            y = (v[k] - a[k] * y) * b[k];
            v[k] = y;
        }
    }
    // back substitution: x_k = dp_k - cp_k x_{k+1}, x_n = 0
    P = 1.0;
    Q = 0.0;
#pragma unroll
    for (int k = CM - 1; k >= 0; --k) {
        if (e0 + k < n) {
            Q = -c[k] * Q + v[k];
            P = -c[k] * P;
        }
    }
    scan_affine_down(P, Q, lane);
    y = Q;
#pragma unroll
    for (int k = CM - 1; k >= 0; --k) {
        if (e0 + k < n) {
            y = v[k] - c[k] * y;
            p.x[(int64_t)sys * p.ss + (int64_t)(e0 + k) * p.se] = y;
        }
    }
}

}  // namespace frm

using namespace frm;

extern "C" int frm_tridiag(frm_ctx* ctx, const frm_tridiag_desc* d, const double* a,
                           const double* b, const double* c, const double* dd, double* x) {
    FRM_REQUIRE(ctx, ctx && d && a && b && c && dd && x, "frm_tridiag: null argument");
    FRM_REQUIRE(ctx, d->nsys >= 0 && d->n >= 1, "frm_tridiag: nsys %d n %d", d->nsys, d->n);
    if (d->nsys == 0) return FRM_OK;
    TriArgs p{a, b, c, dd, x, d->nsys, d->n, d->stride_sys, d->stride_elem};
    const dim3 grid((unsigned)((d->nsys + 3) / 4)), block(256);
    LaunchTimer t(ctx, FRM_K_SWEEP1D, 0.0);
    if (d->n <= 64)
        hipLaunchKernelGGL(k_tridiag<1>, grid, block, 0, ctx->stream, p);
    else if (d->n <= 128)
        hipLaunchKernelGGL(k_tridiag<2>, grid, block, 0, ctx->stream, p);
    else if (d->n <= 256)
        hipLaunchKernelGGL(k_tridiag<4>, grid, block, 0, ctx->stream, p);
    else if (d->n <= 512)
        hipLaunchKernelGGL(k_tridiag<8>, grid, block, 0, ctx->stream, p);
    else if (d->n <= 1024)
        hipLaunchKernelGGL(k_tridiag<16>, grid, block, 0, ctx->stream, p);
    else if (d->n <= 2048)
        hipLaunchKernelGGL(k_tridiag<32>, grid, block, 0, ctx->stream, p);
    else
        return set_error(ctx, FRM_E_RANGE, "frm_tridiag: n = %d > 2048", d->n);
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}
