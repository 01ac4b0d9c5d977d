// LOD weights for g(t) = t^p (TG_solver_<func = 3, ...>, psi-Caputo in time):
// the kb series of TG_solver::_kb_row (fr_d_t_g_2d.cpp:2937-2963) with the
// derivatives of the inverse of g (inv_g_der, func 3: 2857-2864), shared by
// the host (current-step weights) and the weight kernel (history weights).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace frm {

// Safety bound on the series' terms: the terms fall as a power of the index
// for every p > 0 and int_err > 0, so a finite input stops long before.
constexpr int kGpowMaxTerms = 1 << 22;

// The integral of (g(tj) - g(x))^-alpha over [t0, t1] for g = x^p, by the
// Taylor series of the inverse of g around gtj = g(tj), summed until a term
// is at most int_err in magnitude -- the reference's operations in its order.
// *capped is set when the safety bound ends the series.
__host__ __device__ inline double kb_row_pow(double t0, double t1, double gtj, double alpha,
                                             double p, double int_err, bool* capped) {
    const double ip = 1.0 / p;
    double sum = 0.0, v, v2, v3, v4;
    int i = 0;
    {
        const double f0 = pow(gtj, 1.0 / p);  // inv_g_der(gtj, 0, 1)
        v = ((1.0 / p) * pow(gtj, (1.0 / p) - 1)) * (f0 / pow(gtj, 1.0 / p));
    }
    v *= pow(gtj, 1 - alpha);
    const double r1 = 1.0 - pow(t1, p) / gtj, r0 = 1.0 - pow(t0, p) / gtj;
    v3 = pow(r1, 1 - alpha);
    v4 = pow(r0, 1 - alpha);
    v2 = -(v3 - v4) / (1 - alpha);
    sum += v2 * v;
    while (fabs(v * v2) > int_err) {
        if (i >= kGpowMaxTerms) {
            *capped = true;
            break;
        }
        i++;
        v = v * (ip - (i + 1) + 1) / gtj;  // inv_g_der(gtj, i + 1, v)
        v *= -gtj / i;
        v3 *= r1;
        v4 *= r0;
        v2 = -(v3 - v4) / (i - alpha + 1);
        sum += v2 * v;
    }
    return sum;
}

// kb(s, ja, sa, l) at step n (fr_d_t_g_2d.cpp:3112-3131, int_alg 1) without
// the factor Da = 1/Gamma(1 - alpha): zero when the interval lies past the
// target.
__host__ __device__ inline double kb_pow(double alpha, double tau, double p, double int_err,
                                         int n, int s, double ja, double sa, double l,
                                         bool* capped) {
    if (tau * (s + sa + l) <= tau * (n + ja))
        return kb_row_pow(tau * (s + sa), tau * (s + sa + l), pow(tau * (n + ja), p), alpha, p,
                          int_err, capped);
    return 0.0;
}

}  // namespace frm
