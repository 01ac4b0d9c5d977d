// Time sums of the 3D solver of irrigation_fractional.cpp (the Sum family).
//
// The reference computes, per node and per step, the Caputo history term
// that its explicit right part needs (H_solver::time_summing, 4099-4170, with
// the weights of precalc_kb_3, 4081-4098), on the CPU or through its
// OpenCL/CUDA Sum kernel (1254-1320, 1416-1489) driven by fr_ocl_call_sum
// (1828-1905), which uploads the newest time level each step and reads all
// sums back.  Three algorithms (sum_alg):
//   0  S[id] = sum_{t=old_size-2}^{0} kb_3_cache[t] (old[t+1][id] - old[t][id]) / tau[t]
//   1  the same, newest first, stopping before the first t with
//      |kb_3_cache[t]| < sum_param (the older levels are never read again)
//   2  per-node moments Htmp[id][j] += gt^(j+1) coefs[j] (old[1][id] - old[0][id])
//      and S[id] = sum_j v1_j Htmp[id][j] / tau[tstep] with
//      v1_0 = gtj^-alpha, v1_j = v1_{j-1} (alpha + j - 1) / (j gtj)
//      (_kb_row2_fixed_coefs, 1168-1188).
// Here the levels stay on the device: the store keeps the increments
// (level - previous level), so sums 0 and 1 are the shared history kernels
// (FP64, or FP32 with compensated sums), and sum 2 is one streaming pass over
// the moments.  The weights are the caller's (its kb()), as in the reference.
#include <algorithm>
#include <cmath>

#include "frm_history.hpp"

struct frm_timesum {
    frm_ctx* ctx = nullptr;
    int64_t n = 0;
    int32_t sum_alg = 0;
    double sum_param = 0.0;
    int32_t precision = FRM_PREC_F64;
    int32_t levels = 0;        // old_size of the reference
    frm_hist* hist = nullptr;  // increments (sum_alg 0/1)
    double* prev = nullptr;    // newest level
    double* w = nullptr;       // lag-indexed weights kb/tau (device)
    float* w32 = nullptr;
    double* hw = nullptr;      // pinned staging of the weights
    float* hw32 = nullptr;
    hipEvent_t staged{};       // the last upload of hw has completed
    // sum_alg 2
    int32_t nterms = 0;
    int64_t ldm = 0;           // moment stride (term-major, node-contiguous)
    double* inc = nullptr;     // newest increment
    double* mom = nullptr;     // nterms * ldm moments
    double* tab = nullptr;     // device: fold factors m[j], then v1[j]
    double* htab = nullptr;    // pinned staging of tab
    bool folded = true;        // the newest increment is in the moments
};

namespace frm {

// slot[i] = new[i] - prev[i]; prev[i] = new[i]
template <typename T>
__global__ void k_ts_push(const double* __restrict__ lev, double* __restrict__ prev, int64_t n,
                          T* __restrict__ slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = lev[i];
    slot[i] = (T)(v - prev[i]);
    prev[i] = v;
}

__global__ void k_ts_copy(const double* __restrict__ a, int64_t n, double* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_ts_zero(double* __restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// sum_alg 2.  Rounds like the reference's x86-64 build: no contraction, the
// fold as (gt^(j+1) coefs[j]) * diff added to the moment, the terms summed in
// increasing j, one division by tau[tstep].
__global__ __launch_bounds__(256) void k_ts_series(const double* __restrict__ inc,
                                                   double* __restrict__ mom, int64_t ldm,
                                                   const double* __restrict__ tab, int nfold,
                                                   int nterms, double tau_n,
                                                   double* __restrict__ S, int64_t n) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = nfold > 0 ? inc[i] : 0.0;
    const double* v1 = tab + nterms;
    double sum = 0.0;
    int j = 0;
    // eight moments in flight per lane
    for (; j + 8 <= nterms; j += 8) {
        double h[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) h[u] = mom[(int64_t)(j + u) * ldm + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (j + u < nfold) {
                h[u] = h[u] + tab[j + u] * d;
                mom[(int64_t)(j + u) * ldm + i] = h[u];
            }
            sum = sum + v1[j + u] * h[u];
        }
    }
    for (; j < nterms; ++j) {
        double h = mom[(int64_t)j * ldm + i];
        if (j < nfold) {
            h = h + tab[j] * d;
            mom[(int64_t)j * ldm + i] = h;
        }
        sum = sum + v1[j] * h;
    }
    S[i] = sum / tau_n;
}

static void ts_free(frm_timesum* s) {
    if (s->staged) (void)hipEventDestroy(s->staged);
    if (s->hist) frm_hist_destroy(s->hist);
    (void)hipFree(s->prev);
    (void)hipFree(s->w);
    (void)hipFree(s->w32);
    (void)hipFree(s->inc);
    (void)hipFree(s->mom);
    (void)hipFree(s->tab);
    (void)hipHostFree(s->hw);
    (void)hipHostFree(s->hw32);
    (void)hipHostFree(s->htab);
    delete s;
}

static unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace frm

using namespace frm;

extern "C" {

int frm_timesum_create(frm_ctx* ctx, const frm_timesum_desc* d, frm_timesum** out) {
    FRM_REQUIRE(ctx, ctx && d && out, "frm_timesum_create: null argument");
    FRM_REQUIRE(ctx, d->n_nodes > 0, "frm_timesum_create: n_nodes %lld",
                (long long)d->n_nodes);
    FRM_REQUIRE(ctx, d->sum_alg >= 0 && d->sum_alg <= 2, "frm_timesum_create: sum_alg %d",
                d->sum_alg);
    FRM_REQUIRE(ctx, d->precision == FRM_PREC_F64 || d->precision == FRM_PREC_F32C,
                "frm_timesum_create: precision %d (FP64 or compensated FP32)", d->precision);
    FRM_REQUIRE(ctx, d->sum_alg == 2 || d->capacity >= 1, "frm_timesum_create: capacity %d",
                d->capacity);
    FRM_REQUIRE(ctx, d->sum_alg != 2 || d->precision == FRM_PREC_F64,
                "frm_timesum_create: the series moments are FP64");
    FRM_REQUIRE(ctx, d->sum_alg != 2 || (d->sum_param >= 1.0 && d->sum_param <= 4096.0),
                "frm_timesum_create: sum_alg 2 takes 1..4096 terms (sum_param %g)",
                d->sum_param);
    frm_timesum* s = new frm_timesum();
    s->ctx = ctx;
    s->n = d->n_nodes;
    s->sum_alg = d->sum_alg;
    s->sum_param = d->sum_param;
    s->precision = d->precision;
    hipError_t e = hipMalloc(&s->prev, (size_t)s->n * sizeof(double));
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->staged, hipEventDisableTiming);
    if (e != hipSuccess) {
        ts_free(s);
        return check_hip(ctx, e, "frm_timesum_create");
    }
    if (d->sum_alg != 2) {
        // capacity counts time levels; the store holds the increments between them
        const int32_t cap = std::max(1, d->capacity - 1);
        int rc = hist_alloc(ctx, s->n, cap, d->precision, &s->hist);
        if (rc) {
            ts_free(s);
            return rc;
        }
        const bool f32 = d->precision == FRM_PREC_F32C;
        e = hipMalloc(&s->w, (size_t)cap * sizeof(double));
        if (e == hipSuccess) e = hipHostMalloc(&s->hw, (size_t)cap * sizeof(double));
        if (e == hipSuccess && f32) e = hipMalloc(&s->w32, (size_t)cap * sizeof(float));
        if (e == hipSuccess && f32) e = hipHostMalloc(&s->hw32, (size_t)cap * sizeof(float));
    } else {
        // Htmp[id] holds j < sum_param terms (irrigation_fractional.cpp:3396-3402)
        s->nterms = (int32_t)std::ceil(d->sum_param);
        s->ldm = (s->n + 63) / 64 * 64;
        e = hipMalloc(&s->inc, (size_t)s->n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&s->mom, (size_t)s->nterms * s->ldm * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&s->tab, (size_t)2 * s->nterms * sizeof(double));
        if (e == hipSuccess) e = hipHostMalloc(&s->htab, (size_t)2 * s->nterms * sizeof(double));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_ts_zero, dim3(blocks((int64_t)s->nterms * s->ldm, 256)),
                               dim3(256), 0, ctx->stream, s->mom, (int64_t)s->nterms * s->ldm);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipEventRecord(s->staged, ctx->stream);
    if (e != hipSuccess) {
        ts_free(s);
        return check_hip(ctx, e, "frm_timesum_create");
    }
    *out = s;
    return FRM_OK;
}

int frm_timesum_destroy(frm_timesum* s) {
    if (!s) return FRM_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    ts_free(s);
    return FRM_OK;
}

int frm_timesum_clear(frm_timesum* s) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_timesum_clear: null argument");
    frm_ctx* ctx = s->ctx;
    s->levels = 0;
    s->folded = true;
    if (s->hist) s->hist->depth = 0;
    if (s->mom) {
        hipLaunchKernelGGL(k_ts_zero, dim3(blocks((int64_t)s->nterms * s->ldm, 256)), dim3(256),
                           0, ctx->stream, s->mom, (int64_t)s->nterms * s->ldm);
        FRM_HIP(ctx, hipGetLastError());
    }
    return FRM_OK;
}

int frm_timesum_levels(const frm_timesum* s, int32_t* old_size) {
    if (!s || !old_size)
        return set_error(nullptr, FRM_E_INVALID, "frm_timesum_levels: null argument");
    *old_size = s->levels;
    return FRM_OK;
}

int frm_timesum_push(frm_timesum* s, const double* d_level) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_timesum_push: null argument");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, d_level, "frm_timesum_push: null level");
    const dim3 g(blocks(s->n, 256)), b(256);
    if (s->levels == 0) {
        hipLaunchKernelGGL(k_ts_copy, g, b, 0, ctx->stream, d_level, s->n, s->prev);
        FRM_HIP(ctx, hipGetLastError());
        s->levels = 1;
        return FRM_OK;
    }
    if (s->sum_alg == 2) {
        hipLaunchKernelGGL(k_ts_push<double>, g, b, 0, ctx->stream, d_level, s->prev, s->n,
                           s->inc);
        s->folded = false;
    } else {
        frm_hist* h = s->hist;
        if (h->depth >= h->capacity)
            return set_error(ctx, FRM_E_RANGE, "frm_timesum_push: capacity of %d levels reached",
                             h->capacity + 1);
        if (s->precision == FRM_PREC_F32C)
            hipLaunchKernelGGL(k_ts_push<float>, g, b, 0, ctx->stream, d_level, s->prev, s->n,
                               (float*)h->data + (int64_t)h->depth * h->ld);
        else
            hipLaunchKernelGGL(k_ts_push<double>, g, b, 0, ctx->stream, d_level, s->prev, s->n,
                               (double*)h->data + (int64_t)h->depth * h->ld);
        h->depth += 1;
    }
    FRM_HIP(ctx, hipGetLastError());
    s->levels += 1;
    return FRM_OK;
}

int frm_timesum_eval(frm_timesum* s, const frm_timesum_args* a, double* d_S) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_timesum_eval: null argument");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, a && d_S, "frm_timesum_eval: null argument");
    const dim3 g(blocks(s->n, 256)), b(256);
    if (s->levels < 2) {  // time_sum = 0 until two levels exist
        hipLaunchKernelGGL(k_ts_zero, g, b, 0, ctx->stream, d_S, s->n);
        FRM_HIP(ctx, hipGetLastError());
        return FRM_OK;
    }
    // the staging buffers are rewritten below: the previous upload must be done
    FRM_HIP(ctx, hipEventSynchronize(s->staged));
    if (s->sum_alg == 2) {
        FRM_REQUIRE(ctx, a->ncoefs == 0 || a->coefs, "frm_timesum_eval: null coefs");
        FRM_REQUIRE(ctx, a->tau_n != 0.0, "frm_timesum_eval: tau_n is zero");
        const int nt = s->nterms;
        // fold factors gt^(j+1) coefs[j] for the coefficients that exist
        // (the row2_precalc size guard of 4158), then v1_j
        const int nfold = s->folded ? 0 : std::min(nt, std::max(0, (int)a->ncoefs));
        double mult = a->gt;
        for (int j = 0; j < nfold; ++j) {
            s->htab[j] = mult * a->coefs[j];
            mult *= a->gt;
        }
        double v1 = std::pow(a->gtj, -a->alpha);
        for (int j = 0; j < nt; ++j) {
            s->htab[nt + j] = v1;
            const double i = j + 1.0;
            v1 *= -(-a->alpha - i + 1.0) / (i * a->gtj);
        }
        FRM_HIP(ctx, hipMemcpyAsync(s->tab, s->htab, (size_t)2 * nt * sizeof(double),
                                    hipMemcpyHostToDevice, ctx->stream));
        FRM_HIP(ctx, hipEventRecord(s->staged, ctx->stream));
        LaunchTimer t(ctx, FRM_K_MEMTERM,
                      (double)s->n * sizeof(double) * (2 + nt + (nfold > 0 ? nfold : 0)));
        hipLaunchKernelGGL(k_ts_series, g, b, 0, ctx->stream, s->inc, s->mom, s->ldm, s->tab,
                           nfold, nt, a->tau_n, d_S, s->n);
        FRM_HIP(ctx, hipGetLastError());
        s->folded = true;
        return FRM_OK;
    }
    frm_hist* h = s->hist;
    const int depth = h->depth;  // old_size - 1
    FRM_REQUIRE(ctx, a->kb_3_cache && a->tau, "frm_timesum_eval: null kb_3_cache or tau");
    // restricted summing: the newest terms down to the first small weight
    int L = depth;
    if (s->sum_alg == 1) {
        L = 0;
        for (int t = depth - 1; t >= 0 && !(std::fabs(a->kb_3_cache[t]) < s->sum_param); --t)
            ++L;
    }
    // lag-indexed weights: lag 1 is the newest increment, t = depth - 1
    for (int k = 0; k < L; ++k) {
        const int t = depth - 1 - k;
        s->hw[k] = a->kb_3_cache[t] / a->tau[t];
    }
    const bool f32 = s->precision == FRM_PREC_F32C;
    if (f32) {
        for (int k = 0; k < L; ++k) s->hw32[k] = (float)s->hw[k];
        if (L > 0)
            FRM_HIP(ctx, hipMemcpyAsync(s->w32, s->hw32, (size_t)L * sizeof(float),
                                        hipMemcpyHostToDevice, ctx->stream));
    } else if (L > 0) {
        FRM_HIP(ctx, hipMemcpyAsync(s->w, s->hw, (size_t)L * sizeof(double),
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    FRM_HIP(ctx, hipEventRecord(s->staged, ctx->stream));
    int nch = 0;
    int rc = memterm_partials(ctx, h, 1, s->w, depth, &nch, depth - L, f32 ? s->w32 : nullptr);
    if (rc) return rc;
    return memterm_reduce1(ctx, h, nch, d_S);
}

}  // extern "C"
