// Fits sums of exponentials to the LOD Caputo and Atangana-Baleanu history
// weights for the capacity/term pairs given on the command line and prints
// "K L max_rel c_abs_max" per pair, for tests/test_soe_fit.py (built with
// AddressSanitizer there, so any read outside the fit's arrays aborts).
//   soe_fit_check <alpha> <tau> <kernel: caputo|ab> K:L [K:L ...]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "frm_soe.hpp"
#include "frm_weights.hpp"

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    const double alpha = std::atof(argv[1]), tau = std::atof(argv[2]);
    const bool ab = std::strcmp(argv[3], "ab") == 0;
    std::vector<double> abx;
    for (int a = 4; a < argc; ++a) {
        int K = 0, L = 0;
        if (std::sscanf(argv[a], "%d:%d", &K, &L) != 2 || K < 1 || L < 1) return 2;
        // lag-indexed weights of the two LOD targets, as frm_lod2d_create
        // builds them: rows kb(d = k + 1/2, l = 1), columns kb(d = k + 1, l = 1)
        std::vector<double> w(2 * (size_t)K);
        if (ab) {
            abx.assign(K + 2, 0.0);
            frm::ab_exact_weights(alpha, tau, K + 2, abx.data());
        }
        for (int k = 1; k <= K; ++k) {
            if (ab) {
                w[k - 1] = w[K + k - 1] = 0.5 * abx[k] / (1.0 - alpha);
            } else {
                w[k - 1] = frm::caputo_kb(alpha, tau, k + 0.5, 1.0);
                w[K + k - 1] = frm::caputo_kb(alpha, tau, k + 1.0, 1.0);
            }
        }
        const frm::SoeFit f = frm::soe_fit({w.data(), w.data() + K}, K, 2, L, tau);
        double cmax = 0.0;
        for (double c : f.c) cmax = std::fmax(cmax, std::fabs(c));
        if (f.q.size() != (size_t)L || f.c.size() != 2 * (size_t)L) return 3;
        std::printf("%d %d %.6e %.6e\n", K, L, f.max_rel, cmax);
    }
    return 0;
}
