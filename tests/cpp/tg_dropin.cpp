// Drives the C++ drop-in class frm::TG_solver (include/frm_solvers.hpp):
// reads sN, M, steps, alpha, tau, sigma, C1, S and the five arrays
// (rowA, rowB, colA, colB, C0) as float64 from argv[1], runs calc_step()
// `steps` times, fetches b_C and writes it to argv[2].
#include <cstdio>
#include <vector>

#include "frm_solvers.hpp"

int main(int argc, char** argv) {
    if (argc != 3) return 1;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    double h[8];
    if (std::fread(h, sizeof(double), 8, f) != 8) return 3;
    const int sN = (int)h[0], M = (int)h[1], steps = (int)h[2];
    const size_t nn = (size_t)(sN + 2) * (M + 2);
    std::vector<double> a(5 * nn);
    if (std::fread(a.data(), sizeof(double), a.size(), f) != a.size()) return 4;
    std::fclose(f);
    try {
        frm::Context ctx(0);
        frm_lod2d_desc d{sN, M, steps, FRM_PREC_F64, h[3], h[4], h[5], h[6], h[7],
                         a.data(), a.data() + nn, a.data() + 2 * nn, a.data() + 3 * nn};
        frm::TG_solver s(ctx, d, a.data() + 4 * nn);
        for (int t = 0; t < steps; ++t) s.calc_step();
        s.fetch();
        FILE* o = std::fopen(argv[2], "wb");
        std::fwrite(s.b_C.data(), sizeof(double), s.b_C.size(), o);
        std::fclose(o);
        std::printf("tstep %d\n", s.tstep);
    } catch (const frm::error& e) {
        std::fprintf(stderr, "frm error %d: %s\n", e.status, e.what());
        return 5;
    }
    return 0;
}
