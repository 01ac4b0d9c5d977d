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

   private:
    double ksum(int l, double tj);
    double a_;
    int maxl_, maxkiter_;
    double eps_;
    std::vector<double> logi_, lgk_, prodlog_, prodsgn_;
};

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
