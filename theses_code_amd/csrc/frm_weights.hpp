// Host-side FP64 weight generators.
#pragma once
#include <vector>

namespace frm {

// Integrals of the Atangana-Baleanu kernel E_a(-a/(1-a) u^a) by the methods
// of MF_int_series.cpp (series / integral representation), with the
// per-(k, l) caches of the series kept across calls.
class AbIntegrator {
   public:
    AbIntegrator(double a, int maxl, int maxkiter, double eps);
    // integral over [ts, ts1] of E_a(-a/(1-a) (tj - s)^a) ds
    double integral(double tj, double ts1, double ts);

    // coefficient of (ts1^(l+1) - ts^(l+1)) in the series form
    // (MF_int_ksum, MF_int_series.cpp:332-435)
    double ksum(int l, double tj);

   private:
    double a_;
    int maxl_, maxkiter_;
    double eps_;
    std::vector<double> logi_, lgk_, prodlog_, prodsgn_;
};

// Upper limit of the integral representation over [ts, ts1] at tj.
double fv_upper(double a, double tj, double ts1, double ts, double eps);

// The maxl quadrature nodes (weight wk, abscissa vk) below the upper limit C.
class FvNodes {
   public:
    FvNodes(double a, int maxl, double C);
    void node(int l, double& wk, double& vk) const;

   private:
    int maxl_;
    double wk1_, wk2_, wk3_, mC1_, mC2_;
};

// Rows per step of ab_series_tables: Kc, its log flag, dp, sk, log sk.
constexpr int kAbSeriesRows = 5;

// Node-independent coefficients of the series history of solve_and_check
// (simple = 0) for steps i = 2..cap: tab[(i-2)*kAbSeriesRows*maxl ...],
// choice[i-2] = 1 (series) or 2 (integral form).
void ab_series_tables(double a, double tau, int maxl, int maxkiter, double eps, int cap,
                      double* tab, int* choice);

// Per-step coefficients of the binomial-series history of TG_solver with
// dt_alg = 3 (int_alg 6), g(t) = t, `terms` moments per node and direction.
// For step n = 1..cap-1, row (n-1): upd[k] = (n tau)^(k+1) v3_k (the moment
// update of the newest full-step increment), er[k] / ec[k] = the evaluation
// factors of _kb_row2_fixed_coefs at the row / column targets (n+1/2) tau and
// (n+1) tau, and cnt[4(n-1)..] = {moments updated in the row and column
// direction (how many coefficients row2_precalc holds), terms evaluated in
// each (up to a non-finite factor)}.
struct BinomialTables {
    int terms = 0;
    std::vector<double> upd, er, ec;  // [cap-1][terms]
    std::vector<int> cnt;             // [cap-1][4]
};
BinomialTables binomial_tables(double alpha, double tau, double int_err, int terms, int cap);

// E_{a,b}(z) by the method of mittag_leffler (MF_int_series.cpp:247-302):
// the power series for z > 0; for z <= 0 the integral representation when
// |z| > pi a (0 < a < 1, b < 1 + a), else the series summed in pairs, the
// pair sums ordered by decreasing magnitude and added level by level.
double ml_eval(double a, double b, double z, double eps);

// kb(s, ja, sa, l) of TG_solver for g(t) = t as a function of
// d = n + ja - s - sa: Da tau^(1-a)/(1-a) (d^(1-a) - (d-l)^(1-a)), Da =
// 1/Gamma(1-a), with the difference of powers in a cancellation-free form.
double caputo_kb(double alpha, double tau, double d, double l);

// Gamma(x), W. J. Cody's rational approximation (Stirling above 12).
double gamma_cody(double x);

// Accurate FP64 integrals I(d) = int_{(d-1)tau}^{d tau} E_a(-a/(1-a) u^a) du,
// d = 1..D (series for d = 1, spectral representation otherwise).
void ab_exact_weights(double a, double tau, int D, double* out);

}  // namespace frm
