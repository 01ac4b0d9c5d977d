// Sum-of-exponentials fit of a lag-indexed weight sequence:
//   w_t(k) ~ sum_l c_{t,l} q_l^(k - k0),  k = k0..K,  q_l = exp(-r_l tau),
// with the rates r_l log-spaced over [r_lo, r_hi] and the coefficients from
// a relative-weighted linear least-squares problem solved by Householder QR
// in FP64.  A small search over the rate range keeps the best fit.  The
// reference approximates long histories by series too (dt_alg 3/4 in
// fr_d_t_g_2d.cpp:3385-3469, the Lc accumulators of MF_int_series.cpp:
// 569-836); this is the exponential-sum form the GPU recurrences need.
#include <algorithm>
#include <cmath>
#include <vector>

#include "frm_soe.hpp"

namespace frm {

namespace {
// Least squares min |A x - b| for A m x n (column-major, m >= n), several
// right-hand sides; A and B are overwritten.  Returns x (n x nrhs).
std::vector<double> householder_lsq(int m, int n, std::vector<double>& A, int nrhs,
                                    std::vector<double>& B) {
    if (m < n) return std::vector<double>((size_t)n * nrhs, 0.0);  // callers fit at most m modes
    for (int j = 0; j < n; ++j) {
        double* col = &A[(size_t)j * m];
        double norm = 0.0;
        for (int i = j; i < m; ++i) norm += col[i] * col[i];
        norm = std::sqrt(norm);
        if (norm == 0.0) continue;
        const double alpha = col[j] > 0 ? -norm : norm;
        std::vector<double> v(m - j);
        v[0] = col[j] - alpha;
        for (int i = j + 1; i < m; ++i) v[i - j] = col[i];
        double vv = 0.0;
        for (double e : v) vv += e * e;
        if (vv == 0.0) continue;
        auto apply = [&](double* x) {
            double dot = 0.0;
            for (int i = j; i < m; ++i) dot += v[i - j] * x[i];
            const double f = 2.0 * dot / vv;
            for (int i = j; i < m; ++i) x[i] -= f * v[i - j];
        };
        for (int k = j; k < n; ++k) apply(&A[(size_t)k * m]);
        for (int r = 0; r < nrhs; ++r) apply(&B[(size_t)r * m]);
    }
    std::vector<double> x((size_t)n * nrhs, 0.0);
    for (int r = 0; r < nrhs; ++r) {
        for (int j = n - 1; j >= 0; --j) {
            double s = B[(size_t)r * m + j];
            for (int k = j + 1; k < n; ++k) s -= A[(size_t)k * m + j] * x[(size_t)r * n + k];
            const double d = A[(size_t)j * m + j];
            x[(size_t)r * n + j] = d != 0.0 ? s / d : 0.0;
        }
    }
    return x;
}
}  // namespace

SoeFit soe_fit(const std::vector<const double*>& w, int K, int k0, int L, double tau) {
    SoeFit best;
    best.max_rel = HUGE_VAL;
    const int m = K - k0 + 1;
    const int nt = (int)w.size();
    best.q.assign(L > 0 ? L : 0, 0.0);
    best.c.assign((size_t)nt * (L > 0 ? L : 0), 0.0);
    if (m <= 0 || L <= 0) {
        best.max_rel = 0.0;
        return best;
    }
    // A history with fewer fitted lags than modes is fitted by m modes (the
    // system is then square); the remaining modes keep q = 0 and c = 0, so
    // their recurrences add nothing.
    const int Lf = std::min(L, m);
    const double lo_f[] = {1e-3, 1e-2, 1e-1}, hi_f[] = {1.0, 3.0, 10.0};
    for (double flo : lo_f) {
        for (double fhi : hi_f) {
            const double rlo = flo / (K * tau), rhi = fhi / (k0 * tau);
            std::vector<double> q(L, 0.0);
            for (int l = 0; l < Lf; ++l) {
                const double x = Lf == 1 ? std::log(rlo) : std::log(rlo) + (std::log(rhi) - std::log(rlo)) * l / (Lf - 1);
                q[l] = std::exp(-std::exp(x) * tau);
            }
            // columns q_l^(k-k0), each scaled to unit max; rows scaled by 1/|w|
            // (done per target by dividing the row of A and b; A depends on the
            // target through the row scale, so solve target by target)
            SoeFit cur;
            cur.q = q;
            cur.c.assign((size_t)nt * L, 0.0);
            cur.max_rel = 0.0;
            for (int t = 0; t < nt; ++t) {
                std::vector<double> A((size_t)m * Lf), B(m);
                std::vector<double> colmax(Lf, 0.0);
                for (int l = 0; l < Lf; ++l) {
                    double p = 1.0;
                    for (int i = 0; i < m; ++i) {
                        A[(size_t)l * m + i] = p;
                        p *= q[l];
                    }
                }
                for (int i = 0; i < m; ++i) {
                    const double y = w[t][k0 - 1 + i];
                    const double sc = y != 0.0 ? 1.0 / std::fabs(y) : 1.0;
                    B[i] = y * sc;
                    for (int l = 0; l < Lf; ++l) A[(size_t)l * m + i] *= sc;
                }
                for (int l = 0; l < Lf; ++l) {
                    for (int i = 0; i < m; ++i) colmax[l] = std::max(colmax[l], std::fabs(A[(size_t)l * m + i]));
                    if (colmax[l] > 0)
                        for (int i = 0; i < m; ++i) A[(size_t)l * m + i] /= colmax[l];
                }
                std::vector<double> x = householder_lsq(m, Lf, A, 1, B);
                for (int l = 0; l < Lf; ++l) cur.c[(size_t)t * L + l] = colmax[l] > 0 ? x[l] / colmax[l] : 0.0;
                // residual; a non-finite one counts as an infinite error
                std::vector<double> p(Lf, 1.0);
                for (int i = 0; i < m; ++i) {
                    double s = 0.0;
                    for (int l = 0; l < Lf; ++l) {
                        s += cur.c[(size_t)t * L + l] * p[l];
                        p[l] *= q[l];
                    }
                    const double y = w[t][k0 - 1 + i];
                    double e = std::fabs(s - y) / std::max(std::fabs(y), 1e-300);
                    if (!std::isfinite(e)) e = HUGE_VAL;
                    cur.max_rel = std::max(cur.max_rel, e);
                }
            }
            if (cur.max_rel < best.max_rel) best = cur;
        }
    }
    return best;
}

}  // namespace frm
