/* CPU oracle for the fractional-solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference's CPU
 * algorithms (sevaboh/theses_code) so that tests, the smoke check and the
 * bench's CPU baseline can check the HIP path against them.  The product
 * (theses_code_amd/) never links, loads or calls it.
 *
 * Each function names the reference file:line it follows.  Where the
 * reference's numerics depend on the exact operation order (truncated series,
 * cancellation-prone differences) the restatement keeps that order, so that
 * the oracle matches the reference to round-off; the oracle is pinned against
 * golden vectors written by the reference itself (tests/golden/).
 */
#ifndef FRM_ORACLE_H
#define FRM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- L0 numerics ---------------------------------------------------- */
/* Cody's rational Gamma as used by every reference file
 * (mt_minimal.cpp:20-189; MF_int_series.cpp:25-193 extends it to x <= 0). */
double orc_gamma(double x);
double orc_log_gamma(double x);

/* ---- W3/H3/S2: 2D LOD solver of fr_d_t_g_2d.cpp ---------------------- */
/* kb(s, ja, sa, l) at step n with int_alg = 1 (_kb_row,
 * fr_d_t_g_2d.cpp:2937-2963, driven from kb() 3059-3239).
 * func: 0 -> g(t)=t, 3 -> g(t)=t^fpow. Da = 1/Gamma(1-alpha). */
double orc_tg_kb(int func, double fpow, double alpha, double tau, double int_err,
                 int n, int s, double ja, double sa, double l);

typedef struct orc_tg orc_tg;
/* Coefficient arrays are (sN+2)*(M+2), idx(i,j) = i + j*(sN+2): the lower (A_)
 * and upper (B) off-diagonals of the row (x) and column (y) sweeps. */
orc_tg* orc_tg_create(int sN, int M, int func, double fpow, double alpha, double tau,
                      double sigma, double C1, double S, double int_err,
                      const double* rowA, const double* rowB,
                      const double* colA, const double* colB, const double* C0);
void orc_tg_destroy(orc_tg* t);
/* One calc_step() (fr_d_t_g_2d.cpp:2552-2718), dt_alg = 0. */
void orc_tg_step(orc_tg* t);
void orc_tg_get(const orc_tg* t, double* C);
/* Append 2*depth seeded synthetic snapshots and advance tstep by depth
 * (used to time steps at a given history depth). */
void orc_tg_prefill(orc_tg* t, int depth, uint64_t seed);
int orc_tg_tstep(const orc_tg* t);
/* Reference coefficient form (fr_d_t_g_2d.cpp:3242-3265, fixed_d = 1):
 * rows A_=(V2/dL)(dv/dL+1/2), B=(V2/dL)(dv/dL-1/2); columns V2*dv/dLx^2. */
void orc_tg_coefs_ref(int sN, int M, const double* V2, double dv, double dL, double dLx,
                      double* rowA, double* rowB, double* colA, double* colB);
/* Switch the LOD restatement to the Atangana-Baleanu kernel: both
 * half-steps use the full-step weights I(d) = MF_int_choice(d tau, tau, 0,
 * a, a, maxl, maxkiter, eps)/(1-a) of MF_intF_simple (MF_int_series.cpp:
 * 614-718): current I(1), history I(k+1)/2, no half-step correction. */
enum { ORC_KERNEL_AB = 10 };
void orc_tg_set_ab(orc_tg* t, int maxl, int maxkiter, double eps);
/* dt_alg 3: the binomial-series history with `terms` moments per node and
 * direction (fr_d_t_g_2d.cpp:3385-3428; int_alg 6), before the first step */
void orc_tg_set_dt3(orc_tg* t, int terms);
/* dt_alg 2: the newest-first history sum stops at the first weight below eps */
void orc_tg_set_dt2(orc_tg* t, double eps);
/* dt_alg 4: the ia/ibk series approximations of the history per node and
 * direction (fr_d_t_g_2d.cpp:246-741, 3438-3469), before the first step */
void orc_tg_set_dt4(orc_tg* t);
/* dt_alg 4 test hook: one node's approximated sums V[n-1] at x = n for
 * n = 1..nsteps (direction 0: rows, 1: columns), increments F[0..nsteps) */
void orc_dt4_series(double alpha, int dir, const double* F, int nsteps, double* V);
/* Grid spacings the reference constructor derives (fr_d_t_g_2d.cpp:2337-2388). */
void orc_tg_ref_spacing(int sN, int M, double* dL, double* dLx);

/* ---- W1/W2/H1/H2/S1: 1D H_solver of mt_minimal.cpp / mt_psi_caputo.cpp - */
typedef struct orc_mt orc_mt;
enum { ORC_MT_CONST = 0, ORC_MT_CV2 = 1, ORC_MT_KEXP = 2 };
enum { ORC_TOP_AUTO = 0, ORC_TOP_DIRICHLET = 1, ORC_TOP_FLUX = 2 };
/* psi = 0: mt_minimal L1 weights; psi = 1: mt_psi_caputo with g(t)=t^beta.
 * form: CONST (KK=D, 1/C=1), CV2 (KK=k, 1/C piecewise as cv_form 2),
 * KEXP (KK=k0*exp(b*u), 1/C = 1).  top: AUTO follows the reference rule
 * (first-order flux unless 1/C(U0) == 1e20); Ev is the top flux term.
 * bottom_cond: 1 Dirichlet Hbottom, 0 flux (mt_minimal.cpp:987-1020). */
orc_mt* orc_mt_create(int N, double L, int psi, double beta, double alpha, double tau,
                      int form, double D, double k, double k0, double b,
                      int top, double top_val, double Ev, int bottom_cond, double Hbottom,
                      int picard, const double* U0);
void orc_mt_destroy(orc_mt* m);
void orc_mt_step(orc_mt* m);          /* calc_step(), mt_minimal.cpp:1032 / mt_psi_caputo.cpp:1435 */
void orc_mt_get(const orc_mt* m, double* U);  /* N+2 values */
double orc_mt_da(const orc_mt* m);
/* History sum alone, as Om1() forms it (mt_minimal.cpp:906-939 /
 * mt_psi_caputo.cpp:1315-1342): out[i-1] = Da*time_sum for i = 1..N-1, from
 * os snapshots of length N+2.  For psi, weights are kb_time of step tstep. */
void orc_mt_history(int N, int psi, double beta, double alpha, double tau, int os, int tstep,
                    const double* snaps, double* out);
/* psi weight calc_int(alpha, a, a+1, A, tau, 1) (mt_psi_caputo.cpp:720-740)
 * for a = 0..A-1, g(t) = t^beta. */
void orc_psi_weights(double alpha, double beta, double tau, int A, double* out);
/* the adaptive-quadrature fallback of calc_int (_kb, mt_psi_caputo.cpp:476-494) */
double orc_psi_kb_quad(double alpha, double beta, double t0, double t1, double gtj, int sing);

/* ---- W4/H5/S3: Atangana-Baleanu path of MF_int_series.cpp -------------- */
double orc_mittag_leffler(double a, double b, double z, double eps);   /* :247-302 */
/* OpenMP threads of the oracle's parallel loops (for the CPU baselines) */
void orc_set_threads(int n);
int orc_get_threads(void);
/* I(d) = MF_int_choice(d*tau, tau, 0, a, a, maxl, maxkiter, eps), d = 1..D */
void orc_ab_weights(double a, double tau, int maxl, int maxkiter, double eps, int D, double* out);
/* solve_and_check(a, T, n, m, simple=1, ...) (MF_int_series.cpp:875-958):
 * returns the mean squared error; if C_out != NULL writes the final C[0..m]. */
double orc_ab_solve(double a, double T, int n, int m, int maxl, int maxkiter, double eps,
                    int F_func, double* C_out);
/* per-node history of MF_intF_simple (mn = 1) for n = 2..nmax: out[n-2] */
void orc_ab_history(double a, double tau, int maxl, int maxkiter, double eps, int nmax,
                    const double* fvs, double* out);
/* one node's series history of solve_and_check with simple = 0 (MF_intF_choice
 * with on = i-2, mn = 1) for i = 2..nmax: out[i-2], choice[i-2] (1 series, 2
 * integral form; may be NULL).  Returns the count of non-finite recounts. */
long orc_ab_series_history(double a, double tau, int maxl, int maxkiter, double eps, int nmax,
                           const double* fvs, double* out, double* choice);
/* solve_and_check(a, T, n, m, simple=0, ...): as orc_ab_solve with the series history */
double orc_ab_solve_series(double a, double T, int n, int m, int maxl, int maxkiter, double eps,
                           int F_func, double* C_out);
/* MF_int_choice(tj, ts1, ts, a, a, maxl, maxkiter, eps) */
double orc_ab_int(double tj, double ts1, double ts, double a, int maxl, int maxkiter, double eps);

/* ---- Sum family of irrigation_fractional.cpp ------------------------- */
/* H_solver::time_summing for sum_alg 0 and 1 (irrigation_fractional.cpp:
 * 4121-4148) at every node: old holds nlev levels of n values (level-major),
 * kb the nlev-1 weights kb_3_cache[t], tau the steps tau[t]. */
void orc_timesum01(int sum_alg, double sum_param, int64_t n, int nlev, const double* old,
                   const double* kb, const double* tau, double* S);
/* time_summing for sum_alg 2 (4150-4164): folds the newest increment diff
 * into the moments H (nterms x n, term-major) with the first
 * min(nterms, ncoefs) coefficients, then evaluates _kb_row2_fixed_coefs
 * (1168-1188) over all nterms moments and divides by tau_n. */
void orc_timesum2(int64_t n, int nterms, double* H, const double* diff, const double* coefs,
                  int ncoefs, double gtj, double gt, double alpha, double tau_n, double* S);

#ifdef __cplusplus
}
#endif
#endif
