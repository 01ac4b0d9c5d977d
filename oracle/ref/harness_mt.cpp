// Golden-vector harness for the 1D moisture solvers H_solver of mt_minimal.cpp
// (default) and mt_psi_caputo.cpp (built with -DREF_PSI).
//
// The reference's Richards-equation physics is reduced to forms whose
// coefficients the oracle restates exactly:
//   * gamma = 1 (integer-order space derivative): calc_v() returns zeros, and
//     Dg = dL^-2 / Gamma(1);
//   * k_form = 0: KK(i) = k (constant filtration coefficient);
//   * cv_form = 2: inv_dw_dh(i) = 1/Ch(P), P = -U[i], piecewise linear Ch
//     (mt_minimal.cpp:499-503), 1e20 when P < 0;
//   * Tr = 0, no irrigation, no precipitation, flags|4 (day_night_coef = 1):
//     Uc() = Gamma(1) * dL * Ev / k.
// Modes (first input value):
//   0: trajectory. Input: mode, BS, NB, alpha, tau, L, k, Ev, Hbottom,
//      bottom_cond, nsteps, every, [beta (psi only)], U0[N+2].
//      Output: U[0..N+1] every `every` steps; for psi also kb_time of the
//      final step.
//   1: history sum isolation (mt_minimal.cpp:906-939). Input: mode, BS, NB,
//      alpha, tau, os, tstep, [beta], snapshots[os][N+2]. U == 0 so that
//      Om1() leaves Om[i] = Da * time_sum. Output: Da, Om[1..N-1].
//   2: psi weights only (mt_psi_caputo.cpp:720-740). Input: mode, alpha,
//      beta, tau, nA, A[nA]. Output for each A: calc_int(alpha,a,a+1,A,tau,1)
//      for a = 0..A-1.
//   3: calc_int's quadrature fallback alone (mt_psi_caputo.cpp:476-494).
//      Input: mode, n, then n x (alpha, beta, t0, t1, gtj, sing).  Output:
//      _kb(1, t0, t1, gtj, sing, alpha) for each.
#include "harness_io.h"
#define main ref_main
#ifdef REF_PSI
#include "mt_psi_caputo.cpp"
#else
#include "mt_minimal.cpp"
#endif
#undef main
#include <cstring>

static H_solver* make_solver(int bs, int nb, double alpha, double tau, double L,
                             double k, double Ev, double Hb, int bc) {
    BS = bs;
    NB = nb;
    sN = N = BS * NB;
    sNB = NB;
    F = new double[N + 2];
    BVK = new double[BS];
    flags = 4;
    H_solver* s = new H_solver(/*rpf*/ 1, /*cvf*/ 2, /*no_irr*/ 1, bc, tau, alpha,
                               /*gamma*/ 1.0, NULL, /*kpf*/ 0, L, 0.5, Hb);
    s->Tr = 0.0;
    s->Ev = Ev;
    s->k = s->saved_k = k;
    return s;
}

int main(int argc, char** argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 1; }
#if defined(_OPENMP) && defined(REF_PSI)
    omp_init_lock(&lock);
#endif
    HarnessIn in(argv[1]);
    HarnessOut out(argv[2]);
    int mode = in.i();
    if (mode == 0) {
        int bs = in.i(), nb = in.i();
        double alpha = in.d(), tau = in.d(), L = in.d(), k = in.d(), Ev = in.d(), Hb = in.d();
        int bc = in.i(), nsteps = in.i(), every = in.i();
#ifdef REF_PSI
        func_in_kernel = 3;
        func_power = 1.0;
        func_power_t = in.d();
#endif
        H_solver* s = make_solver(bs, nb, alpha, tau, L, k, Ev, Hb, bc);
        in.arr(s->b_U, N + 2);
        for (int t = 1; t <= nsteps; ++t) {
            s->calc_step();
            if (t % every == 0 || t == nsteps) out.arr(s->b_U, N + 2);
        }
#ifdef REF_PSI
        out.arr(s->kb_time, s->tstep);  // kb_time of the last step: tstep entries
#endif
    } else if (mode == 1) {
        int bs = in.i(), nb = in.i();
        double alpha = in.d(), tau = in.d();
        int os = in.i(), tstep = in.i();
#ifdef REF_PSI
        func_in_kernel = 3;
        func_power = 1.0;
        func_power_t = in.d();
#endif
        H_solver* s = make_solver(bs, nb, alpha, tau, 1.0, 1e-3, 0.0, 1e300, 0);
        for (int q = 0; q < os; ++q) {
            double* hs = new double[N + 2];
            in.arr(hs, N + 2);
            s->oldH.push_back(hs);
        }
        for (int i = 0; i < N + 2; ++i) s->b_U[i] = 0.0;
        s->tstep = tstep;
#ifdef REF_PSI
        // kb_time as calc_step() builds it (mt_psi_caputo.cpp:1473-1490)
        double sfp = func_power;
        func_power = func_power_t;
        s->kb_time = new double[tstep + 1];
        for (int aa = 0; aa <= tstep; ++aa)
            s->kb_time[aa] = s->calc_int(alpha, aa, aa + 1, tstep + 1, tau, 1);
        func_power = sfp;
#endif
        s->Om1();
        out.d(s->Da);
        out.arr(s->Om + 1, N - 1);
    } else if (mode == 2) {
#ifdef REF_PSI
        double alpha = in.d(), beta = in.d(), tau = in.d();
        int nA = in.i();
        func_in_kernel = 3;
        func_power = beta;
        H_solver* s = make_solver(1, 4, alpha, tau, 1.0, 1e-3, 0.0, 1e300, 0);
        for (int q = 0; q < nA; ++q) {
            int A = in.i();
            for (int aa = 0; aa < A; ++aa) out.d(s->calc_int(alpha, aa, aa + 1, A, tau, 1));
        }
#else
        std::fprintf(stderr, "mode 2 needs REF_PSI\n");
        return 1;
#endif
    } else if (mode == 3) {
#ifdef REF_PSI
        int n = in.i();
        func_in_kernel = 3;
        H_solver* s = make_solver(1, 4, 0.5, 1e-3, 1.0, 1e-3, 0.0, 1e300, 0);
        for (int q = 0; q < n; ++q) {
            double alpha = in.d(), beta = in.d(), t0 = in.d(), t1 = in.d(), gtj = in.d();
            int sing = in.i();
            func_power = beta;
            out.d(s->_kb(1, t0, t1, gtj, sing, alpha));
        }
#else
        std::fprintf(stderr, "mode 3 needs REF_PSI\n");
        return 1;
#endif
    }
    return 0;
}
