// Golden-vector harness for the fitting input reader of mt_minimal.cpp
// (main, 1955-2043): runs the reference's own main in fitting mode on a
// fit_input file and captures what it parsed.
//
// The reference's translation unit is built as a shared library
// (libref_mt_main.so, main renamed ref_main); this executable defines
// fit_and_solve (mt_minimal.cpp:1401) with the reference's signature, so the
// call at the end of the reader resolves here (the library's own definition
// is interposed, nothing of it runs) and the parsed fields are written out.
//
// Usage: ref_fitio out.bin <arguments of the reference's main>
// Output (float64): 1 if fit_and_solve was reached (the file parsed), else 0;
// then to_fit nvalues ninit ncheck nperc nirr nEV_T nEV_F ev_mu pso_n pso_o
// pso_fi_p pso_fi_g pso_eps pso_max_iter, values_T/Z/F, init_Z/F,
// perc_T/A, irr_T/A, check_T/Z/F, EV_T, LAI, EV_F[0..nEV_F) (nEV_T each),
// nparam_values, param_values[0..nparam_values) -- when the file holds
// exactly 19 values the reference reports 20 and the 20th is unset (stack
// contents).
#include "harness_io.h"

int ref_main(int argc, char** argv);

static const char* g_out = nullptr;
static int g_called = 0;

void fit_and_solve(double t, double save_tau, double out_tau, int irr, int bc, double tau_m,
                   double* init_Z, double* init_F, int ninit, double* perc_T, double* perc_A,
                   int nperc, double* irr_T, double* irr_A, int nirr, double* values_T,
                   double* values_Z, double* values_F, int nvalues, double* check_T,
                   double* check_Z, double* check_F, int ncheck, double* EV_T, double** EV_F,
                   double* LAI, double ev_mu, int nEV_T, int nEV_F, int to_fit, int pso_n,
                   double pso_o, double pso_fi_p, double pso_fi_g, double pso_eps,
                   int pso_max_iter, int fit, double* param_values, int nparam_values,
                   char* et_file, char* pv_file) {
    HarnessOut out(g_out);
    g_called = 1;
    out.d(1.0);
    for (double x : {(double)to_fit, (double)nvalues, (double)ninit, (double)ncheck, (double)nperc,
                     (double)nirr, (double)nEV_T, (double)nEV_F, ev_mu, (double)pso_n, pso_o,
                     pso_fi_p, pso_fi_g, pso_eps, (double)pso_max_iter})
        out.d(x);
    out.arr(values_T, nvalues);
    out.arr(values_Z, nvalues);
    out.arr(values_F, nvalues);
    out.arr(init_Z, ninit);
    out.arr(init_F, ninit);
    out.arr(perc_T, nperc);
    out.arr(perc_A, nperc);
    out.arr(irr_T, nirr);
    out.arr(irr_A, nirr);
    out.arr(check_T, ncheck);
    out.arr(check_Z, ncheck);
    out.arr(check_F, ncheck);
    out.arr(EV_T, nEV_T);
    out.arr(LAI, nEV_T);
    for (int j = 0; j < nEV_F; ++j) out.arr(EV_F[j], nEV_T);
    out.d(nparam_values);
    out.arr(param_values, nparam_values);
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: %s out.bin <main arguments>\n", argv[0]); return 1; }
    g_out = argv[1];
    argv[1] = argv[0];
    ref_main(argc - 1, argv + 1);
    if (!g_called) {
        HarnessOut out(g_out);
        out.d(0.0);
    }
    return 0;
}
