// Host FP64 generators of Atangana-Baleanu kernel integrals.
//
// I(tj, ts1, ts) = integral over [ts, ts1] of E_a(-a/(1-a) (tj - s)^a) ds is
// evaluated with the two methods of MF_int_series.cpp: the double series of
// MF_int / MF_int_ksum (332-455), used when (a/(1-a)) tj^a < 1, and the
// maxl-node quadrature of the ML function's integral representation of
// MF_intFv (457-521) otherwise (MF_int_choice, 523-530).  The series stops
// after ten consecutive terms below eps/l!, so its value depends on the
// operation order; the arithmetic below follows the reference's order so
// that the GPU path sees the reference's weights.
#include <algorithm>
#include <cmath>
#include <vector>

#include "frm_weights.hpp"

namespace frm {

namespace {
constexpr double kPi = 3.14159265358979323846;

double cody_gamma(double x);

double cody_log_gamma(double x) {
    if (x <= 0.0) return cody_log_gamma(x + 1.0) - std::log(x);
    if (x < 12.0) return std::log(std::fabs(cody_gamma(x)));
    static const double c[8] = {1.0 / 12.0,   -1.0 / 360.0,        1.0 / 1260.0, -1.0 / 1680.0,
                                1.0 / 1188.0, -691.0 / 360360.0, 1.0 / 156.0,  -3617.0 / 122400.0};
    const double z = 1.0 / (x * x);
    double acc = c[7];
    for (int k = 6; k >= 0; --k) {
        acc *= z;
        acc += c[k];
    }
    return (x - 0.5) * std::log(x) - x + 0.91893853320467274178032973640562 + acc / x;
}

// W. J. Cody's rational approximation on (1,2) with argument reduction (the
// published coefficient table every reference file carries).
double cody_gamma(double x) {
    static const double p[8] = {-1.71618513886549492533811E+0, 2.47656508055759199108314E+1,
                                -3.79804256470945635097577E+2, 6.29331155312818442661052E+2,
                                8.66966202790413211295064E+2,  -3.14512729688483675254357E+4,
                                -3.61444134186911729807069E+4, 6.64561438202405440627855E+4};
    static const double q[8] = {-3.08402300119738975254353E+1, 3.15350626979604161529144E+2,
                                -1.01515636749021914166146E+3, -3.10777167157231109440444E+3,
                                2.25381184209801510330112E+4,  4.75584627752788110767815E+3,
                                -1.34659959864969306392456E+5, -1.15132259675553483497211E+5};
    if (x <= 0.0) return (1.0 / x) * cody_gamma(x + 1.0);
    if (x < 0.001) return 1.0 / (x * (1.0 + 0.577215664901532860606512090 * x));
    if (x < 12.0) {
        double y = x;
        int shift = 0;
        const bool below = y < 1.0;
        if (below)
            y += 1.0;
        else {
            shift = (int)std::floor(y) - 1;
            y -= shift;
        }
        double num = 0.0, den = 1.0;
        const double z = y - 1;
        for (int k = 0; k < 8; ++k) {
            num = (num + p[k]) * z;
            den = den * z + q[k];
        }
        double r = num / den + 1.0;
        if (below)
            r /= (y - 1.0);
        else
            for (int k = 0; k < shift; ++k) r *= y++;
        return r;
    }
    if (x > 171.624) return 1e308 * 10.0;
    return std::exp(cody_log_gamma(x));
}

double sgn(double x) { return x > 0.0 ? 1.0 : -1.0; }

// Pair sums ordered by decreasing magnitude, then added neighbour by
// neighbour, level by level; an odd leftover joins the last new sum (a
// single value is therefore doubled, as in sum_sort_pyramid,
// MF_int_series.cpp:200-217).
double pyramid_sum(std::vector<double> v) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end(), [](double x, double y) { return std::fabs(x) > std::fabs(y); });
    size_t lo = 0, hi = 0;
    do {
        lo = hi;
        hi = v.size();
        const size_t odd = (hi - lo) % 2;
        for (size_t i = lo; i < hi - odd; i += 2) v.push_back(v[i] + v[i + 1]);
        if (odd) v.back() += v[hi - 1];
    } while (v.size() - hi > 1);
    return v.back();
}

// Integrand of the integral representation of E_{a,b}(z), z < 0.
double ml_kernel(double a, double b, double z, double r) {
    const double k1 = (1.0 / (kPi * a)) * std::pow(r, (1 - b) / a) * std::exp(-std::pow(r, 1.0 / a));
    const double k2 = r * std::sin(kPi * (1 - b)) - z * std::sin(kPi * (1 - b + a));
    const double k3 = r * r - 2 * r * z * std::cos(kPi * a) + z * z;
    return k1 * k2 / k3;
}

// Endpoint average plus sd equally spaced interior points, with sd doubled
// until two estimates differ by less than eps (or sd reaches 1000).
double ml_quadrature(double a, double b, double z, double r0, double r1, double eps) {
    int sd = 1;
    double prev = 0.0;
    for (;;) {
        const double ends = 0.5 * (r1 - r0) * (ml_kernel(a, b, z, r0) + ml_kernel(a, b, z, r1));
        double v = 0.0;
        for (int i = 1; i <= sd; i++) v += ml_kernel(a, b, z, r0 + (i / ((double)(sd + 1))) * (r1 - r0));
        v *= (r1 - r0);
        if (std::isfinite(ends)) v += ends;
        v /= (double)(sd + 1);
        if (sd >= 1000) return v;
        if (sd == 1) prev = ends;
        if (std::fabs(v - prev) < eps) return v;
        sd *= 2;
        prev = v;
    }
}

double spectral(double a, double v) {  // Fv of the integral representation
    const double p = -a / (1.0 - a);
    return (1.0 / (kPi * a)) * (p * std::sin(kPi * a)) / (v * v - 2 * v * p * std::cos(kPi * a) + p * p);
}
}  // namespace

// Upper limit C of the integral representation over [ts, ts1] at tj: grown
// by 1.5 from 1 until the integrand's part beyond C falls to eps
// (MF_intFv, MF_int_series.cpp:477-489; MF_intF_Fv uses ts = tj - tau,
// ts1 = tj, :739-750).
double fv_upper(double a, double tj, double ts1, double ts, double eps) {
    double C = 1.0;
    for (;;) {
        const double ks = ts * std::pow(C, 1.0 / a), ks1 = ts1 * std::pow(C, 1.0 / a);
        const double k2 = std::log(std::fabs(spectral(a, C))) - (1.0 / a) * std::log(C) -
                          std::pow(C, 1.0 / a) * tj;
        if (std::fabs(std::exp(ks + k2) - std::exp(ks1 + k2)) > eps)
            C *= 1.5;
        else
            break;
    }
    return C;
}

// Quadrature nodes below C: maxl/3 equal steps up to mC2 = Fm - Fd, maxl/3 up
// to mC1 = Fm + Fd (Fm the maximum of the spectral factor, Fd five times
// its distance to the inflection points), then linearly growing steps
// (MF_int_series.cpp:490-510, 744-753).
FvNodes::FvNodes(double a, int maxl, double C) : maxl_(maxl) {
    const double Fm = -a * std::cos(kPi * a) / (1.0 - a);
    const double Fd = 5 * (std::sqrt(3.0) / 3) * std::sqrt((a * a / ((1.0 - a) * (1.0 - a))) - Fm * Fm);
    mC1_ = Fm + Fd;
    mC2_ = Fm - Fd;
    if (mC2_ < 0.0) mC2_ = 0.0;
    if (mC1_ < 0.0) mC1_ = 0.0;
    if (mC2_ == 0) mC2_ = mC1_ / 2.0;
    if (mC1_ == 0) {
        mC1_ = C / 3.0;
        mC2_ = 2.0 * C / 3.0;
    }
    wk1_ = mC2_ / (maxl / 3);
    wk2_ = (mC1_ - mC2_) / (maxl / 3);
    wk3_ = 2.0 * (C - mC2_) / (maxl / 3);
}

void FvNodes::node(int l, double& wk, double& vk) const {
    const int third = maxl_ / 3;
    if (l < third) {
        wk = wk1_;
        vk = l * wk + 0.5 * wk;
    } else if (l < 2 * third) {
        wk = wk2_;
        vk = mC2_ + (l - third) * wk + 0.5 * wk;
    } else {
        wk = wk2_ + (wk3_ - wk2_) * (double)(l - 2 * third) / third;
        vk = mC1_ + (l - 2 * third) * wk + 0.5 * wk;
    }
}

AbIntegrator::AbIntegrator(double a, int maxl, int maxkiter, double eps)
    : a_(a), maxl_(maxl), maxkiter_(maxkiter), eps_(eps) {
    logi_.resize(maxl);
    logi_[0] = 0.0;
    for (int i = 1; i < maxl; ++i) logi_[i] = std::log((double)i);
    lgk_.assign(maxkiter, -1.0);
    prodlog_.assign((size_t)maxkiter * maxl, 0.0);
    prodsgn_.assign((size_t)maxkiter * maxl, 2.0);
}

// coefficient of (ts1^(l+1) - ts^(l+1)) in the series (MF_int_ksum)
double AbIntegrator::ksum(int l, double tj) {
    const double a = a_;
    double ret = 0.0;
    int small_run = 0;
    const double lfact = cody_log_gamma(l + 1.0);
    const double la = std::log(a / (1.0 - a));
    const double ltj = std::log(tj);
    const double leps = std::log(eps_);
    double s = std::pow(-1.0, l);
    for (double k = 0; k < maxkiter_; k++) {
        const int ki = (int)k;
        const double k1 = k * la;
        double k2;
        if (lgk_[ki] < 0) {
            k2 = cody_log_gamma(a * k + 1.0);
            lgk_[ki] = k2;
        } else
            k2 = lgk_[ki];
        const double k3 = (a * k - l) * ltj;
        double term = k1 + k3 - k2;
        const size_t q = (size_t)ki * maxl_ + l;
        double k4 = 0, s4 = 1;
        if (prodsgn_[q] == 2) {
            for (int i = 1; i <= l; i++) {
                k4 += std::log(std::fabs(a * k - i + 1)) - logi_[i];
                s4 *= sgn(a * k - i + 1);
            }
            prodlog_[q] = k4;
            prodsgn_[q] = s4;
        } else {
            k4 = prodlog_[q];
            s4 = prodsgn_[q];
        }
        term += k4;
        term = std::exp(term);
        term *= s * s4 / (l + 1.0);
        if (std::log(std::fabs(term)) < leps - lfact)
            small_run++;
        else
            small_run = 0;
        if (small_run == 10) break;
        if (!std::isfinite(term)) break;
        ret += term;
        s *= -1.0;
    }
    return ret;
}

double AbIntegrator::integral(double tj, double ts1, double ts) {
    const double a = a_;
    const double maxz = (a / (1 - a)) * std::pow(tj, a);
    if (maxz < 1.0) {
        double ret = 0.0;
        for (int l = 0; l < maxl_; l++) {
            const double lk = std::pow(ts1, (l + 1.0)) - std::pow(ts, (l + 1.0));
            const double lt = lk * ksum(l, tj);
            if (std::isfinite(lt)) ret += lt;
        }
        return ret;
    }
    // integral representation, quadrature over maxl nodes in three segments
    const double C = fv_upper(a, tj, ts1, ts, 1e-10);
    const FvNodes nd(a, maxl_, C);
    const int maxl = maxl_;
    double ret = 0.0;
    for (int l = 0; l < maxl; l++) {
        double wk, vk;
        nd.node(l, wk, vk);
        const double ks = ts * std::pow(vk, 1.0 / a), ks1 = ts1 * std::pow(vk, 1.0 / a);
        const double k2 = std::log(std::fabs(spectral(a, vk))) - (1.0 / a) * std::log(vk) -
                          std::pow(vk, 1.0 / a) * tj;
        double lt = wk * (std::exp(ks + k2) - std::exp(ks1 + k2));
        if (spectral(a, vk) < 0) lt = -lt;
        if (!std::isfinite(lt)) break;
        ret += lt;
    }
    return ret;
}

// Per-step coefficients of the series history of solve_and_check with
// simple = 0: MF_intF_choice(i, tau, a, a, maxl, maxkiter, eps, ..., on =
// i - 2, mn = 1, append_all = 1) (MF_int_series.cpp:839-853), which every
// node of a problem calls with its own accumulators.  Everything that does
// not depend on the node is formed here, with the reference's operations:
//   series phase ((a/(1-a)) tj^a < 1): Kc = MF_int_ksum(l, tj) (MF_intF,
//     :576-590) and dp = ((t+1) tau)^(l+1) - (t tau)^(l+1) (:598);
//   integral phase: Kc = wk v^(-1/a) e^(-v^(1/a) tj) Fv(v), or its log with
//     the flag sign(Fv) when that underflows to 0 (MF_intF_Fv, :755-777);
//   both phases, MF_intF_Fv's accumulator (:779-804): sk = e^(v^(1/a) t tau)
//     - e^(v^(1/a) (t+1) tau) and its log form v^(1/a) t tau +
//     log(e^(v^(1/a) tau) - 1), to which the node adds log|F|,
// with t = i - 2, tj = i tau and the node set v (wk) of step i.
void ab_series_tables(double a, double tau, int maxl, int maxkiter, double eps, int cap,
                      double* tab, int* choice) {
    AbIntegrator ks(a, maxl, maxkiter, eps);
    for (int i = 2; i <= cap; ++i) {
        double* T = tab + (size_t)(i - 2) * kAbSeriesRows * maxl;
        const double tj = i * tau;
        const int t = i - 2;
        const bool series = (a / (1 - a)) * std::pow(tj, a) < 1.0;
        choice[i - 2] = series ? 1 : 2;
        const double C = fv_upper(a, tj, tj, tj - tau, eps);
        const FvNodes nd(a, maxl, C);
        for (int l = 0; l < maxl; ++l) {
            double wk, vk;
            nd.node(l, wk, vk);
            double kc = 0.0, lg = 0.0, dp = 0.0;
            if (series) {
                kc = ks.ksum(l, tj);
                dp = std::pow((t + 1) * tau, (l + 1.0)) - std::pow(t * tau, l + 1.0);
            } else {
                kc = wk * std::pow(vk, -(1.0 / a)) * std::exp(-std::pow(vk, 1.0 / a) * tj) * spectral(a, vk);
                if (kc == 0.0) {
                    lg = sgn(spectral(a, vk));
                    kc = std::log(wk) - (1.0 / a) * std::log(vk) - std::pow(vk, 1.0 / a) * tj +
                         std::log(std::fabs(spectral(a, vk)));
                }
            }
            T[0 * maxl + l] = kc;
            T[1 * maxl + l] = lg;
            T[2 * maxl + l] = dp;
            T[3 * maxl + l] = std::exp(std::pow(vk, 1.0 / a) * t * tau) - std::exp(std::pow(vk, 1.0 / a) * ((t + 1) * tau));
            T[4 * maxl + l] = std::pow(vk, 1.0 / a) * t * tau + std::log(std::exp(std::pow(vk, 1.0 / a) * tau) - 1);
        }
    }
}

// Coefficient k of _kb_row2 (fr_d_t_g_2d.cpp:2970-3038) for g(t) = t on
// [t0, t1]: the integral of x^k over [t0, t1] over t1^(k+1); k = 0 as
// (t1 - t0)/t1, else (1 - (t0/t1)^(k+1))/(k+1) with v6 = (t0/t1)^(k+1) the
// caller's running power.  The reference then runs its series in powers of
// (x - t1) of the inverse's derivatives, which for g = t adds a zero term
// (the second derivative of t is 0) and stops.
static double binomial_v3(int k, double t0, double t1, double v6) {
    if (k == 0) return (t1 - t0) / t1;
    return 1.0 * ((1.0 / (k + 1.0)) * (1 - v6));
}

BinomialTables binomial_tables(double alpha, double tau, double int_err, int terms, int cap) {
    BinomialTables bt;
    bt.terms = terms;
    const int rows = std::max(0, cap - 1);
    bt.upd.assign((size_t)rows * terms, 0.0);
    bt.er.assign((size_t)rows * terms, 0.0);
    bt.ec.assign((size_t)rows * terms, 0.0);
    bt.cnt.assign((size_t)rows * 4, 0);
    std::vector<double> v3(terms);
    for (int n = 1; n < cap; ++n) {
        const int i = n - 1;  // the newest full step, interval [tau i, tau(i+1)]
        const double t0 = tau * (i + 0.0), t1 = tau * (i + 0.0 + 1.0);
        double v6 = t0 / t1;
        for (int k = 0; k < terms; ++k) {
            if (k != 0) v6 *= t0 / t1;
            v3[k] = binomial_v3(k, t0, t1, v6);
        }
        // terms one _kb_row2 call at target gtj takes: until `terms` of them
        // or until |v1 v2 v3| <= int_err (its `continue` goes to that test);
        // row2_precalc keeps the largest count of the calls so far
        auto taken = [&](double gtj) {
            double v1 = 1, v2 = std::pow(gtj, -alpha) * t1;
            int k = 0;
            do {
                if (k != 0) v1 *= -(-alpha - k + 1) / k;
                if (k != 0) v2 *= t1 / gtj;
                const double v = v3[k];
                k++;
                if (k > 10000) break;  // integr_max_niter
                if (k < terms) {
                    if (std::fabs(v1 * v2 * v) > int_err) continue;
                    break;
                }
                break;
            } while (true);
            return k;
        };
        // position 1 (ja 1/2): called once, at step i + 1; position 5 (ja 1):
        // at step i and again at step i + 1 (kb()'s block, 3112-3140)
        const int c1 = taken(tau * (n + 0.5));
        const int c5 = std::max(taken(tau * (i + 1.0)), taken(tau * (n + 1.0)));
        double* u = bt.upd.data() + (size_t)i * terms;
        const double g1 = n * tau;
        double mult = g1;
        for (int k = 0; k < std::max(c1, c5); ++k) {
            u[k] = mult * v3[k];
            mult *= g1;
        }
        int* cn = bt.cnt.data() + (size_t)i * 4;
        cn[0] = c1;
        cn[1] = c5;
        // evaluation factors of _kb_row2_fixed_coefs (3041-3057)
        for (int dir = 0; dir < 2; ++dir) {
            const double gtj = dir == 0 ? (n + 0.5) * tau : (n + 1.0) * tau;
            double* e = (dir == 0 ? bt.er.data() : bt.ec.data()) + (size_t)i * terms;
            double f = std::pow(gtj, -alpha), fi = 0;
            int used = 0;
            do {
                e[used] = f;
                fi += 1.0;
                used++;
                f *= -(-alpha - fi + 1.0) / (fi * gtj);
                if (!std::isfinite(f)) break;
            } while (used != terms);
            cn[2 + dir] = used;
        }
    }
    return bt;
}

}  // namespace frm

namespace frm {

// I(d) = integral over [(d-1) tau, d tau] of E_a(-lam u^a) du, lam = a/(1-a),
// to FP64 accuracy: d = 1 by the series t E_{a,2}(-lam t^a) =
// sum_k (-lam)^k t^(a k + 1) / Gamma(a k + 2); d >= 2 from the spectral
// representation E_a(-lam u^a) = int_0^inf mu(r) e^(-r u) dr with
// mu(r) = lam sin(pi a) r^(a-1) / (pi (r^(2a) + 2 lam r^a cos(pi a) + lam^2)),
// integrated over x = log r by the trapezoid rule (the integrand is analytic
// in a strip around the real x axis and decays exponentially at both ends,
// so the rule converges geometrically in the step).
void ab_exact_weights(double a, double tau, int D, double* out) {
    const double lam = a / (1.0 - a);
    {
        const double t = tau, z = -lam * std::pow(t, a);
        double s = 0.0;
        for (int k = 0; k < 400; ++k) {
            const double tk = std::pow(z, k) / std::exp(std::lgamma(a * k + 2));
            s += tk;
            if (k > 5 && std::fabs(tk) < 1e-18 * std::fabs(s)) break;
        }
        out[0] = t * s;
    }
    if (D < 2) return;
    const double h = 0.01;
    const double x0 = std::log(1e-16) / a + std::log(lam) / a - 5.0;
    const double x1 = std::log(60.0 / tau);
    const int nx = (int)((x1 - x0) / h) + 1;
    std::vector<double> base(nx), fac(nx), q(nx);
    const double sa = std::sin(kPi * a), ca = std::cos(kPi * a);
    for (int j = 0; j < nx; ++j) {
        const double r = std::exp(x0 + j * h);
        const double ra = std::pow(r, a);
        const double mu = lam * sa / kPi * std::pow(r, a - 1) / (ra * ra + 2 * lam * ra * ca + lam * lam);
        base[j] = mu * (-std::expm1(-r * tau)) * h;  // (1 - e^{-r tau})/r * r dx
        q[j] = std::exp(-r * tau);
        fac[j] = q[j];  // e^{-r (d-1) tau} for d = 2
    }
    for (int d = 2; d <= D; ++d) {
        double s = 0.0;
        for (int j = 0; j < nx; ++j) {
            s += base[j] * fac[j];
            fac[j] *= q[j];
        }
        out[d - 1] = s;
    }
}

}  // namespace frm

namespace frm {

double gamma_cody(double x) { return cody_gamma(x); }

// The difference of powers is evaluated as -d^(1-a) expm1((1-a) log1p(-l/d)),
// which keeps full relative precision at large lags where the direct
// difference cancels (the closed form of _kb_row, fr_d_t_g_2d.cpp:2937-2963).
double caputo_kb(double alpha, double tau, double d, double l) {
    const double c = tau == 0 ? 0 : std::pow(tau, 1.0 - alpha) / (1.0 - alpha) / std::tgamma(1.0 - alpha);
    if (d <= 0) return 0.0;
    if (l >= d) return c * std::pow(d, 1.0 - alpha);
    return -c * std::pow(d, 1.0 - alpha) * std::expm1((1.0 - alpha) * std::log1p(-l / d));
}

double ml_eval(double a, double b, double z, double eps) {
    int k0 = (int)(std::log(eps * (1 - z)) / std::log(z));
    const int kmin = 1 + (int)((1 - b) / a);
    if (z >= 1.0 || z < 0.0) k0 = (int)(-2.0 * std::log(eps));
    if (k0 < kmin) k0 = kmin;
    if (z > 0) {
        double res = 0.0, zz = 1.0;
        for (int i = 0; i < k0; i++) {
            const double v = zz / cody_gamma(b + a * i);
            if (!std::isfinite(v)) break;
            res += v;
            zz *= z;
        }
        return res;
    }
    if (std::fabs(z) > kPi * a && b < 1 + a && a > 0 && a < 1) {
        const double r1 = std::max({1.0, 2.0 * std::fabs(z), std::pow(-std::log(kPi * eps / 6.0), a)});
        return ml_quadrature(a, b, z, 0.0, r1, eps);
    }
    std::vector<double> pairs;
    double zz = 1.0;
    for (int i = 0; i < k0; i += 2) {
        const double v1 = zz / cody_gamma(b + a * i);
        zz *= z;
        const double v2 = zz / cody_gamma(b + a * (i + 1));
        zz *= z;
        if (!std::isfinite(v1 + v2)) break;
        pairs.push_back(v1 + v2);
    }
    return pyramid_sum(pairs);
}

}  // namespace frm
