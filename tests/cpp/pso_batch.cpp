// Drives the C++ drop-in batch frm::H_batch (include/frm_solvers.hpp) the way
// a PSO iteration of the reference's fitting driver scores its particles
// (fit_and_solve / solve_and_test, mt_minimal.cpp:1277-1330), one problem per
// particle.  Input (float64, argv[1]): nprob, N, steps, tau, nobs; per
// particle alpha, D; the shared U0[N+2]; T[nobs], Z[nobs], F[nobs].
// Output (argv[2]): the per-particle fitting errors.
#include <cstdio>
#include <vector>

#include "frm_solvers.hpp"

int main(int argc, char** argv) {
    if (argc != 3) return 1;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    double h[5];
    if (std::fread(h, sizeof(double), 5, f) != 5) return 3;
    const int nprob = (int)h[0], N = (int)h[1], steps = (int)h[2], nobs = (int)h[4];
    const double tau = h[3];
    std::vector<double> par(2 * (size_t)nprob), U0(N + 2), obs(3 * (size_t)nobs);
    if (std::fread(par.data(), sizeof(double), par.size(), f) != par.size()) return 4;
    if (std::fread(U0.data(), sizeof(double), U0.size(), f) != U0.size()) return 4;
    if (std::fread(obs.data(), sizeof(double), obs.size(), f) != obs.size()) return 4;
    std::fclose(f);
    try {
        frm::Context ctx(0);
        std::vector<frm_mt1d_problem> probs(nprob);
        for (int p = 0; p < nprob; ++p) {
            frm_mt1d_problem q{};
            q.N = N;
            q.L = 1.0;
            q.alpha = par[2 * p];
            q.form = FRM_FORM_CONST;
            q.D = par[2 * p + 1];
            q.top = FRM_TOP_DIRICHLET;
            q.top_val = U0[0];
            q.bottom_cond = 1;
            q.Hbottom = U0[N];
            probs[p] = q;
        }
        frm_mt1d_desc d{nprob, FRM_MT_L1, steps, FRM_PREC_F64, 1, tau, probs.data()};
        frm::H_batch b(ctx, d, std::vector<std::vector<double>>(nprob, U0));
        b.set_observations({obs.begin(), obs.begin() + nobs}, {obs.begin() + nobs, obs.begin() + 2 * nobs},
                           {obs.begin() + 2 * nobs, obs.end()});
        for (int t = 0; t < steps; ++t) b.calc_step();
        const std::vector<double> err = b.fit_error();
        FILE* o = std::fopen(argv[2], "wb");
        std::fwrite(err.data(), sizeof(double), err.size(), o);
        std::fclose(o);
        std::printf("tstep %d particles %d\n", b.tstep, nprob);
    } catch (const frm::error& e) {
        std::fprintf(stderr, "frm error %d: %s\n", e.status, e.what());
        return 5;
    }
    return 0;
}
