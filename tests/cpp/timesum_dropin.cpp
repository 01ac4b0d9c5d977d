// Drives the C++ drop-in frm::TimeSum (include/frm_solvers.hpp) the way the
// 3D irrigation solver calls its time sums: a level pushed per step
// (oldH.push_back), then the sums of every node (time_summing /
// fr_ocl_call_sum, irrigation_fractional.cpp:4099-4170, 1828-1905).
// Input (float64, argv[1]): n, nlev, sum_alg, sum_param, ncoefs, gtj, gt,
// alpha, tau_n; levels[nlev][n]; kb_3_cache[nlev-1]; tau[nlev-1];
// coefs[ncoefs].  sum_alg 2 evaluates after every push from the second on
// with the same coefficients.  Output (argv[2]): the sums after the last
// level.
#include <cstdio>
#include <vector>

#include "frm_solvers.hpp"

int main(int argc, char** argv) {
    if (argc != 3) return 1;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    double h[9];
    if (std::fread(h, sizeof(double), 9, f) != 9) return 3;
    const long n = (long)h[0];
    const int nlev = (int)h[1], sum_alg = (int)h[2], ncoefs = (int)h[4];
    std::vector<double> lev((size_t)nlev * n), kb(nlev - 1), tau(nlev - 1), coefs(ncoefs);
    if (std::fread(lev.data(), sizeof(double), lev.size(), f) != lev.size()) return 4;
    if (std::fread(kb.data(), sizeof(double), kb.size(), f) != kb.size()) return 4;
    if (std::fread(tau.data(), sizeof(double), tau.size(), f) != tau.size()) return 4;
    if (std::fread(coefs.data(), sizeof(double), coefs.size(), f) != coefs.size()) return 4;
    std::fclose(f);
    std::vector<double> Sums(n, 0.0);
    try {
        frm::Context ctx(0);
        frm::TimeSum ts(ctx, n, sum_alg, h[3], nlev);
        for (int k = 0; k < nlev; ++k) {
            ts.push(lev.data() + (size_t)k * n);
            if (sum_alg == 2 && k >= 1) ts.sums_series(coefs, h[5], h[6], h[7], h[8], Sums.data());
        }
        if (sum_alg != 2) ts.sums(kb.data(), tau.data(), Sums.data());
        std::printf("old_size %d nodes %ld\n", ts.old_size(), n);
    } catch (const frm::error& e) {
        std::fprintf(stderr, "frm error %d: %s\n", e.status, e.what());
        return 5;
    }
    FILE* o = std::fopen(argv[2], "wb");
    std::fwrite(Sums.data(), sizeof(double), Sums.size(), o);
    std::fclose(o);
    return 0;
}
