// Golden-vector harness for the Atangana-Baleanu path of MF_int_series.cpp.
// Modes (first input value):
//   0: weights I(d) = MF_int_choice(d*tau, tau, 0, a, a, maxl, maxkiter, eps)
//      (MF_int_series.cpp:523-530) for d = 1..D.
//      Input: mode, a, tau, maxl, maxkiter, eps, D.  Output: I[1..D].
//   1: full-history sums as the solver calls them (MF_int_series.cpp:924-927):
//      for n = 2..nmax, MF_intF_simple(n, tau, a, a, 0, maxl, maxkiter, eps,
//      NULL, 1, Fvs) in increasing n, so that I_cache is filled exactly as in
//      solve_and_check.  Input: mode, a, tau, maxl, maxkiter, eps, nmax,
//      Fvs[nmax].  Output: the nmax-1 sums.
//   2: solve_and_check(a, T, n, m, simple, maxl, maxkiter, eps)
//      (MF_int_series.cpp:875-958).  Input: mode, a, T, n, m, simple, maxl,
//      maxkiter, eps, F_func.  Output: the returned mean squared error.
//   3: mittag_leffler(a, b, z, eps) (MF_int_series.cpp:247-302) for a list.
//      Input: mode, count, then (a, b, z, eps) per entry.  Output: values.
//   4: series history of one node as solve_and_check calls it with simple = 0
//      (MF_int_series.cpp:917-922): for i = 2..nmax, MF_intF_choice(i, tau, a,
//      a, maxl, maxkiter, eps, NULL, &Lc1, &Lc2, &logLc2, i-2, 1, 1, 1,
//      Fvs[i-2]) in increasing i on one node's accumulators.  Input: mode, a,
//      tau, maxl, maxkiter, eps, nmax, Fvs[nmax-1].  Output: the nmax-1 values
//      and, per i, last_choice (1 series, 2 integral form).
#include "harness_io.h"
#define main ref_main
#include "MF_int_series.cpp"
#undef main

int main(int argc, char** argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 1; }
    HarnessIn in(argv[1]);
    HarnessOut out(argv[2]);
    F_notimes = 1;
    no_prints = 1;
    F_nosorting = 1;
    F_noquadrature = 1;
    F_nosimple = 0;
    int mode = in.i();
    if (mode == 0) {
        double a = in.d(), tau = in.d();
        int maxl = in.i(), maxkiter = in.i();
        double eps = in.d();
        int D = in.i();
        for (int d = 1; d <= D; ++d) out.d(MF_int_choice(d * tau, tau, 0, a, a, maxl, maxkiter, eps));
    } else if (mode == 1) {
        double a = in.d(), tau = in.d();
        int maxl = in.i(), maxkiter = in.i();
        double eps = in.d();
        int nmax = in.i();
        std::vector<double> fvs(nmax);
        in.arr(fvs.data(), nmax);
        for (int n = 2; n <= nmax; ++n)
            out.d(MF_intF_simple(n, tau, a, a, 0, maxl, maxkiter, eps, NULL, 1, fvs.data()));
    } else if (mode == 2) {
        double a = in.d(), T = in.d();
        int n = in.i(), m = in.i(), simple = in.i(), maxl = in.i(), maxkiter = in.i();
        double eps = in.d();
        F_func = in.i();
        out.d(solve_and_check(a, T, n, m, simple, maxl, maxkiter, eps));
    } else if (mode == 4) {
        double a = in.d(), tau = in.d();
        int maxl = in.i(), maxkiter = in.i();
        double eps = in.d();
        int nmax = in.i();
        std::vector<double> fvs(nmax - 1), choice;
        in.arr(fvs.data(), nmax - 1);
        std::vector<double> Lc1(maxl, 0.0), Lc2(maxl, 0.0), logLc2(maxl, 0.0);
        for (int i = 2; i <= nmax; ++i) {
            out.d(MF_intF_choice(i, tau, a, a, maxl, maxkiter, eps, NULL, &Lc1, &Lc2, &logLc2, i - 2,
                                 1, 1, 1, fvs[i - 2]));
            choice.push_back(last_choice);
        }
        for (double c : choice) out.d(c);
    } else if (mode == 3) {
        int cnt = in.i();
        for (int q = 0; q < cnt; ++q) {
            double a = in.d(), b = in.d(), z = in.d(), eps = in.d();
            out.d(mittag_leffler(a, b, z, eps));
        }
    }
    return 0;
}
