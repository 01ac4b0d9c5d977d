// Golden-vector harness for the 2D LOD solver of fr_d_t_g_2d.cpp.
//
// Drives the reference's own TG_solver_<func,fixed_d=1,kform=0> (CPU path,
// dt_alg = 0 full-history summation, or 3) on injected state:
//   * V2 (the per-node squared-speed field) is overwritten after construction,
//   * the initial field b_C is overwritten and pushed as oldC[0], exactly as
//     TG_solver::initial() does (fr_d_t_g_2d.cpp:2294-2297) with noinit=1,
//   * calc_step() (fr_d_t_g_2d.cpp:2552) is then called nsteps times.
// Output: b_C snapshots every `every` steps, then the eight kb() weight
// variants for s = 0..n at the final step n (fr_d_t_g_2d.cpp:3059-3239).
//
// Input layout (float64): func, sN, M, alpha, tau, sigma, d_value, C1, S,
//   nsteps, every, int_alg, func_power, dt_alg, dt_eps, V2[(sN+2)(M+2)],
//   C0[(sN+2)(M+2)].  dt_alg 3 is the binomial-series history (Htmp1/Htmp2,
//   fr_d_t_g_2d.cpp:3385-3428) with dt_eps terms; update() then takes
//   int_alg 6 unless 2 is given (2421-2424).
#include "harness_io.h"
#define main ref_main
#include "fr_d_t_g_2d.cpp"
#undef main
#include <cstring>

template <int func>
static void run(HarnessIn& in, HarnessOut& out, int sN_, int M_) {
    double alpha = in.d(), tau = in.d(), sigma = in.d(), dv = in.d();
    double c1 = in.d(), S = in.d();
    int nsteps = in.i(), every = in.i();
    int ialg = in.i();
    double fpow = in.d();
    int dalg = in.i();
    double deps = in.d();
    sN = sN_;
    _M = M_;
    int_alg = ialg;
    dt_alg = dalg;
    dt_eps = deps;
    int_err = 1e-6;
    _sigma = sigma;
    TG_solver_<func, 1, 0>* s = new TG_solver_<func, 1, 0>(
        tau, alpha, 0.0, /*k_der*/ 0, dv, -1, 5, fpow, -1, -1, NULL, 0, -1,
        /*noinit*/ 1, c1, S);
    if (func == 3) s->func_power = fpow;
    size_t nn = (size_t)(sN + 2) * (_M + 2);
    in.arr(s->V2, nn);
    in.arr(s->b_C, nn);
    double* hs = new double[nn];
    std::memcpy(hs, s->b_C, nn * sizeof(double));
    s->oldC.push_back(hs);
    s->fixedV = NULL;  // initial() sets these (fr_d_t_g_2d.cpp:2285-2289)
    s->nfixedV = 0;
    s->N = sN;
    for (int k = 1; k <= nsteps; ++k) {
        s->calc_step();
        if (k % every == 0 || k == nsteps) out.arr(s->b_C, nn);
    }
    // weights at the final step n = tstep for every s <= n
    int n = s->tstep;
    for (int v = 0; v < 8; ++v) {
        double ja = (v & 1) ? 1.0 : 0.5;
        double sa = (v & 2) ? 0.5 : 0.0;
        double l = (v & 4) ? 1.0 : 0.5;
        for (int q = 0; q <= n; ++q) out.d(s->kb(q, ja, sa, l));
    }
    delete s;
}

int main(int argc, char** argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 1; }
    HarnessIn in(argv[1]);
    HarnessOut out(argv[2]);
    debug_level = 0;
    int func = in.i();
    int sN_ = in.i(), M_ = in.i();
    if (func == 0) run<0>(in, out, sN_, M_);
    else if (func == 3) run<3>(in, out, sN_, M_);
    else { std::fprintf(stderr, "unsupported func %d\n", func); return 1; }
    return 0;
}
