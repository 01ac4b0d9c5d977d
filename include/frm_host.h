/* Host-side FP64 generators of the weights and series coefficients that the
 * GPU path uses (no GPU involved; callable without a device).  Each is the
 * product's own restatement of the reference routine named beside it. */
#ifndef FRM_HOST_H
#define FRM_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gamma(x): W. J. Cody's rational approximation with argument reduction and
 * Stirling's series above 12 (Gamma/LogGamma of fr_d_t_g_2d.cpp:75-244,
 * MF_int_series.cpp:30-85). */
int frm_host_gamma(double x, double* out);

/* E_{a,b}(z) by the method of mittag_leffler (MF_int_series.cpp:247-302). */
int frm_host_mittag_leffler(double a, double b, double z, double eps, double* out);

/* MF_int_choice(tj, ts1, ts, a, a, maxl, maxkiter, eps)
 * (MF_int_series.cpp:523-530): integral over [ts, ts1] of
 * E_a(-a/(1-a) (tj - s)^a) ds by the series or the integral representation. */
int frm_host_ab_integral(double a, int32_t maxl, int32_t maxkiter, double eps, double tj,
                         double ts1, double ts, double* out);

/* I(d) = integral over [(d-1) tau, d tau] of E_a(-a/(1-a) u^a) du for
 * d = 1..D: exact = 0 by MF_int_choice(d tau, tau, 0) (the cached I of
 * MF_intF_simple, MF_int_series.cpp:614-718, history eps 1e-10 for d >= 2),
 * exact = 1 to FP64 accuracy (series for d = 1, spectral representation
 * otherwise). */
int frm_host_ab_weights(double a, double tau, int32_t maxl, int32_t maxkiter, double eps,
                        int32_t exact, int32_t D, double* out);

/* Caputo LOD weights for g(t) = t (TG_solver::kb, fr_d_t_g_2d.cpp:2937-2963
 * closed form): out[0..2] = kb(n,1/2,0,1/2), kb(n,1,1/2,1/2), kb(n,1,0,1/2)
 * (row current, column current, half-step correction), then the row history
 * weights kb(s,1/2,0,1) and the column ones kb(s,1,0,1) for lags 1..D
 * (3 + 2D values). */
int frm_host_caputo_lod_weights(double alpha, double tau, int32_t D, double* out);

/* Caputo LOD weights of step n for g(t) = t^p (TG_solver_<3, ...>::kb,
 * int_alg 1: the _kb_row series of fr_d_t_g_2d.cpp:2937-2963 to int_err):
 * the same 3 + 2n values as frm_host_caputo_lod_weights for step n, lag k
 * multiplying the increment of step n - k (the weights change every step).
 * FRM_E_STATE when a series reaches its safety bound. */
int frm_host_lod_weights_pow(double alpha, double tau, double p, double int_err, int32_t n,
                             double* out);

/* Sum-of-exponentials fit of lag-indexed weights w[t*K + k-1] (targets
 * t < ntargets, lags k = k0..K) by L decays q_l = exp(-r_l tau) with
 * log-spaced rates and FP64 least squares in relative weighting:
 * w_t(k) ~ sum_l c[t*L + l] q_l^(k - k0).  Stores the largest relative
 * error over the fitted lags in *max_rel.  With fewer lags than modes
 * (K - k0 + 1 < L) only K - k0 + 1 modes are fitted and the others get
 * q = c = 0.  FRM_E_INVALID for bad arguments or when no fit is finite. */
int frm_host_soe_fit(const double* w, int32_t ntargets, int32_t K, int32_t k0, int32_t L,
                     double tau, double* q, double* c, double* max_rel);

#ifdef __cplusplus
}
#endif
#endif /* FRM_HOST_H */
