// Standalone sum-of-exponentials history: per node, a set of exponentially
// decaying accumulators that replace a weighted sum over the whole history
// when the weights are (close to) a sum of exponentials in the lag.
//
// The reference's series forms keep exactly such per-node accumulators: the
// Lc2 exponential sums of MF_intF_Fv (MF_int_series.cpp:720-836), advanced
// with the newest increment only and combined with step-dependent factors,
// and the dt_alg 3 moments of TG_solver (fr_d_t_g_2d.cpp:3385-3428).  Here the
// recurrence is generic: per node i and mode l < nterms,
//   H_l[i] <- q[l] H_l[i] + incr[i]
//   out[t n + i] (+)= scale sum_l c[t nterms + l] H_l[i],   t < ntargets.
// q and c travel as kernel arguments (at most 64 modes, two targets), so a
// caller may change them every step with no copy.
//
// Layout: the modes are node-major planes H[l * ld + i], ld = n rounded up to
// 64 nodes, so each mode is one coalesced stream; one lane per node.  The
// kernel is HBM-bound: per node-step it reads and writes every mode once
// (2 * nterms * 8 B, or 4 B per value with FP32 mode states) and reads the
// increment and writes the targets.
#include "frm_internal.hpp"
#include "frm_stream.hpp"

namespace frm {

constexpr int kSoeStepMaxTerms = 64;
constexpr int kSoeStepThreads = 256;

struct SoeCoefs {
    double q[kSoeStepMaxTerms];
    double c[2 * kSoeStepMaxTerms];
};

}  // namespace frm

using namespace frm;

struct frm_soe {
    frm_ctx* ctx = nullptr;
    int64_t n = 0, ld = 0;
    int32_t nterms = 0, ntargets = 1, f32 = 0;
    void* H = nullptr;
};

// One lane per node; the mode loop is uniform across the wavefront, so q[l]
// and c[l] are scalar loads from the kernel arguments.  With FP32 states the
// recurrence runs in FP64 and the stored (rounded) value is the one summed,
// so the output is the sum of the states the next step starts from.
template <typename HT, int NT, bool BYPASS>
__global__ __launch_bounds__(kSoeStepThreads) void k_soe_step(
    HT* __restrict__ H, int64_t ld, int64_t n, int nterms, const double* __restrict__ incr,
    SoeCoefs cf, double scale, int accumulate, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * kSoeStepThreads + threadIdx.x;
    if (i >= n) return;
    const bool fold = incr != nullptr;
    const double x = fold ? __builtin_nontemporal_load(incr + i) : 0.0;
    double s0 = 0.0, s1 = 0.0;
    for (int l = 0; l < nterms; ++l) {
        // the planes are streamed once per step (non-temporal when they
        // exceed the cache)
        double h = (double)stream_load<BYPASS>(H + (int64_t)l * ld + i);
        if (fold) {
            h = cf.q[l] * h + x;
            const HT hs = (HT)h;
            stream_store<BYPASS>(hs, H + (int64_t)l * ld + i);
            h = (double)hs;
        }
        s0 += cf.c[l] * h;
        if (NT == 2) s1 += cf.c[nterms + l] * h;
    }
    if (accumulate) {
        out[i] += scale * s0;
        if (NT == 2) out[n + i] += scale * s1;
    } else {
        out[i] = scale * s0;
        if (NT == 2) out[n + i] = scale * s1;
    }
}

// Zeroes the mode planes (a kernel rather than a memset, so a captured
// stream records a kernel node).
template <typename HT>
__global__ __launch_bounds__(kSoeStepThreads) void k_soe_zero(HT* __restrict__ H, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * kSoeStepThreads + threadIdx.x;
    if (i < count) H[i] = (HT)0;
}

static size_t soe_esz(const frm_soe* s) { return s->f32 ? sizeof(float) : sizeof(double); }

int frm_soe_create(frm_ctx* ctx, const frm_soe_desc* d, frm_soe** out) {
    FRM_REQUIRE(ctx, ctx && d && out, "frm_soe_create: null argument");
    FRM_REQUIRE(ctx, d->n_nodes > 0, "frm_soe_create: n_nodes %lld", (long long)d->n_nodes);
    FRM_REQUIRE(ctx, d->nterms >= 1 && d->nterms <= kSoeStepMaxTerms,
                "frm_soe_create: nterms %d (1..%d)", d->nterms, kSoeStepMaxTerms);
    FRM_REQUIRE(ctx, d->ntargets == 1 || d->ntargets == 2, "frm_soe_create: ntargets %d",
                d->ntargets);
    FRM_REQUIRE(ctx, d->f32 == 0 || d->f32 == 1, "frm_soe_create: f32 %d", d->f32);
    frm_soe* s = new frm_soe();
    s->ctx = ctx;
    s->n = d->n_nodes;
    s->ld = (d->n_nodes + 63) / 64 * 64;
    s->nterms = d->nterms;
    s->ntargets = d->ntargets;
    s->f32 = d->f32;
    hipError_t e = hipMalloc(&s->H, (size_t)s->nterms * s->ld * soe_esz(s));
    if (e != hipSuccess) {
        delete s;
        return check_hip(ctx, e, "frm_soe_create");
    }
    if (int rc = frm_soe_clear(s)) {
        (void)hipFree(s->H);
        delete s;
        return rc;
    }
    *out = s;
    return FRM_OK;
}

int frm_soe_destroy(frm_soe* s) {
    if (!s) return FRM_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(s->H);
    delete s;
    return FRM_OK;
}

int frm_soe_clear(frm_soe* s) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_soe_clear: null");
    frm_ctx* ctx = s->ctx;
    const int64_t count = (int64_t)s->nterms * s->ld;
    const dim3 grid((unsigned)((count + kSoeStepThreads - 1) / kSoeStepThreads));
    if (s->f32)
        hipLaunchKernelGGL(k_soe_zero<float>, grid, dim3(kSoeStepThreads), 0, ctx->stream,
                           (float*)s->H, count);
    else
        hipLaunchKernelGGL(k_soe_zero<double>, grid, dim3(kSoeStepThreads), 0, ctx->stream,
                           (double*)s->H, count);
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}

int frm_soe_step(frm_ctx* ctx, frm_soe* s, const double* d_incr, const double* h_q,
                 const double* h_coef, double scale, int32_t accumulate, double* d_out) {
    FRM_REQUIRE(ctx, ctx && s && h_coef && d_out, "frm_soe_step: null argument");
    FRM_REQUIRE(ctx, s->ctx == ctx, "frm_soe_step: state belongs to another context");
    FRM_REQUIRE(ctx, !d_incr || h_q, "frm_soe_step: a fold needs the decays q");
    SoeCoefs cf{};
    for (int l = 0; l < s->nterms; ++l) {
        cf.q[l] = h_q ? h_q[l] : 0.0;
        for (int t = 0; t < s->ntargets; ++t) cf.c[t * s->nterms + l] = h_coef[t * s->nterms + l];
    }
    const size_t esz = soe_esz(s);
    // algorithmic bytes: every mode read (and written on a fold), the
    // increment, the targets (read too when accumulating)
    const double bytes = (double)s->n * ((d_incr ? 2.0 : 1.0) * s->nterms * esz +
                                         (d_incr ? 8.0 : 0.0) +
                                         s->ntargets * 8.0 * (accumulate ? 2.0 : 1.0));
    const dim3 grid((unsigned)((s->n + kSoeStepThreads - 1) / kSoeStepThreads));
    const int acc = accumulate != 0;
    const bool nt = (double)s->nterms * s->ld * esz >= kStreamBypassBytes;
    auto launch = [&](auto kern, auto* H) {
        hipLaunchKernelGGL(kern, grid, dim3(kSoeStepThreads), 0, ctx->stream, H, s->ld, s->n,
                           s->nterms, d_incr, cf, scale, acc, d_out);
    };
    LaunchTimer t(ctx, FRM_K_MEMTERM, bytes);
    if (s->f32) {
        float* H = (float*)s->H;
        if (s->ntargets == 2) nt ? launch(k_soe_step<float, 2, true>, H) : launch(k_soe_step<float, 2, false>, H);
        else nt ? launch(k_soe_step<float, 1, true>, H) : launch(k_soe_step<float, 1, false>, H);
    } else {
        double* H = (double*)s->H;
        if (s->ntargets == 2) nt ? launch(k_soe_step<double, 2, true>, H) : launch(k_soe_step<double, 2, false>, H);
        else nt ? launch(k_soe_step<double, 1, true>, H) : launch(k_soe_step<double, 1, false>, H);
    }
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}
