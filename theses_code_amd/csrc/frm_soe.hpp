// Sum-of-exponentials fits of lag-indexed memory weights (host, FP64).
#pragma once
#include <vector>

namespace frm {

struct SoeFit {
    std::vector<double> q;  // L decay factors exp(-r_l tau)
    std::vector<double> c;  // [targets][L] coefficients of q_l^(k - k0)
    double max_rel = 0.0;   // max relative error over lags k0..K
};

// w[t] points at weights for lags 1..K (w[t][k-1]); fits lags k0..K.
SoeFit soe_fit(const std::vector<const double*>& w, int K, int k0, int L, double tau);

}  // namespace frm
