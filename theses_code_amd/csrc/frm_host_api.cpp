// C ABI of the host-side weight generators (include/frm_host.h).
#include <cmath>
#include <vector>

#include "frm.h"
#include "frm_host.h"
#include "frm_gpow.hpp"
#include "frm_soe.hpp"
#include "frm_weights.hpp"

using namespace frm;

extern "C" {

int frm_host_gamma(double x, double* out) {
    if (!out) return FRM_E_INVALID;
    *out = gamma_cody(x);
    return FRM_OK;
}

int frm_host_mittag_leffler(double a, double b, double z, double eps, double* out) {
    if (!out || !(a > 0) || !(eps > 0)) return FRM_E_INVALID;
    *out = ml_eval(a, b, z, eps);
    return FRM_OK;
}

int frm_host_ab_integral(double a, int32_t maxl, int32_t maxkiter, double eps, double tj,
                         double ts1, double ts, double* out) {
    if (!out || !(a > 0 && a < 1) || maxl < 3 || maxkiter < 1 || !(eps > 0)) return FRM_E_INVALID;
    AbIntegrator ab(a, maxl, maxkiter, eps);
    *out = ab.integral(tj, ts1, ts);
    return FRM_OK;
}

int frm_host_ab_weights(double a, double tau, int32_t maxl, int32_t maxkiter, double eps,
                        int32_t exact, int32_t D, double* out) {
    if (!out || D < 1 || !(a > 0 && a < 1) || !(tau > 0)) return FRM_E_INVALID;
    if (exact) {
        ab_exact_weights(a, tau, D, out);
        return FRM_OK;
    }
    if (maxl < 3 || maxkiter < 1 || !(eps > 0)) return FRM_E_INVALID;
    AbIntegrator cur(a, maxl, maxkiter, eps), hist(a, maxl, maxkiter, 1e-10);
    out[0] = cur.integral(tau, tau, 0.0);
    for (int d = 2; d <= D; ++d) out[d - 1] = hist.integral(d * tau, tau, 0.0);
    return FRM_OK;
}

int frm_host_caputo_lod_weights(double alpha, double tau, int32_t D, double* out) {
    if (!out || D < 0 || !(alpha > 0 && alpha < 1) || !(tau > 0)) return FRM_E_INVALID;
    out[0] = caputo_kb(alpha, tau, 0.5, 0.5);
    out[1] = caputo_kb(alpha, tau, 0.5, 0.5);
    out[2] = caputo_kb(alpha, tau, 1.0, 0.5);
    for (int k = 1; k <= D; ++k) {
        out[3 + k - 1] = caputo_kb(alpha, tau, k + 0.5, 1.0);
        out[3 + D + k - 1] = caputo_kb(alpha, tau, k + 1.0, 1.0);
    }
    return FRM_OK;
}

int frm_host_lod_weights_pow(double alpha, double tau, double p, double int_err, int32_t n,
                             double* out) {
    if (!out || n < 0 || !(alpha > 0 && alpha < 1) || !(tau > 0) || !(p > 0 && p <= 4) ||
        !(int_err > 0))
        return FRM_E_INVALID;
    const double Da = 1.0 / gamma_cody(1.0 - alpha);
    bool capped = false;
    out[0] = Da * kb_pow(alpha, tau, p, int_err, n, n, 0.5, 0.0, 0.5, &capped);
    out[1] = Da * kb_pow(alpha, tau, p, int_err, n, n, 1.0, 0.5, 0.5, &capped);
    out[2] = Da * kb_pow(alpha, tau, p, int_err, n, n, 1.0, 0.0, 0.5, &capped);
    for (int k = 0; k < n; ++k) {
        out[3 + k] = Da * kb_pow(alpha, tau, p, int_err, n, n - 1 - k, 0.5, 0.0, 1.0, &capped);
        out[3 + n + k] = Da * kb_pow(alpha, tau, p, int_err, n, n - 1 - k, 1.0, 0.0, 1.0, &capped);
    }
    return capped ? FRM_E_STATE : FRM_OK;
}

int frm_host_soe_fit(const double* w, int32_t ntargets, int32_t K, int32_t k0, int32_t L,
                     double tau, double* q, double* c, double* max_rel) {
    if (!w || !q || !c || !max_rel || ntargets < 1 || K <= k0 || k0 < 1 || L < 1 || !(tau > 0))
        return FRM_E_INVALID;
    std::vector<const double*> tw;
    for (int t = 0; t < ntargets; ++t) tw.push_back(w + (size_t)t * K);
    const SoeFit f = soe_fit(tw, K, k0, L, tau);
    if (f.q.size() != (size_t)L || f.c.size() != (size_t)ntargets * L) return FRM_E_INVALID;
    for (int l = 0; l < L; ++l) q[l] = f.q[l];
    for (size_t i = 0; i < f.c.size(); ++i) c[i] = f.c[i];
    *max_rel = f.max_rel;
    // no finite fit over the rate ranges tried
    return std::isfinite(f.max_rel) ? FRM_OK : FRM_E_INVALID;
}

}  // extern "C"
