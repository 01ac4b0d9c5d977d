// Reads a fitting input file with frm::read_fit_input (include/frm_fitio.hpp)
// and prints what it parsed, one field per line, for tests/test_fitio.py.
//   fitio_check <fit_input> <Hmult> [et_file|-] [pv_file|-]
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "frm_fitio.hpp"

static void row(const char* name, const std::vector<double>& v) {
    std::printf("%s", name);
    for (double x : v) std::printf(" %.17g", x);
    std::printf("\n");
}

int main(int argc, char** argv) {
    if (argc < 3) return 1;
    const char* et = argc > 3 && std::strcmp(argv[3], "-") ? argv[3] : nullptr;
    const char* pv = argc > 4 && std::strcmp(argv[4], "-") ? argv[4] : nullptr;
    frm::FitInput in;
    const bool ok = frm::read_fit_input(argv[1], std::atof(argv[2]), in, et, pv);
    std::printf("ok %d\n", ok ? 1 : 0);
    if (!ok) return 0;
    std::printf("counts %d %d %d %d %d %d %d %d\n", in.to_fit, in.nvalues, in.ninit, in.ncheck,
                in.nperc, in.nirr, in.nEV_T, in.nEV_F);
    std::printf("ev_mu %.17g\n", in.ev_mu);
    std::printf("pso %d %.17g %.17g %.17g %.17g %d\n", in.pso_n, in.pso_o, in.pso_fi_p,
                in.pso_fi_g, in.pso_eps, in.pso_max_iter);
    row("values_T", in.values_T);
    row("values_Z", in.values_Z);
    row("values_F", in.values_F);
    row("init_Z", in.init_Z);
    row("init_F", in.init_F);
    row("perc_T", in.perc_T);
    row("perc_A", in.perc_A);
    row("irr_T", in.irr_T);
    row("irr_A", in.irr_A);
    row("check_T", in.check_T);
    row("check_Z", in.check_Z);
    row("check_F", in.check_F);
    row("EV_T", in.EV_T);
    row("LAI", in.LAI);
    for (size_t j = 0; j < in.EV_F.size(); ++j) row("EV_F", in.EV_F[j]);
    row("params", in.param_values);
    // one result time level, as the checking run writes it
    std::vector<double> U = {-1.5, -0.25, 0.0}, w = {0.3, 0.35, 0.4};
    frm::write_res_rows(stdout, 1000.0, 0.05, U.data(), w.data(), 2);
    // one log row over the check values: model values are the check values
    // shifted by a fixed pattern (one positive, clamped to 0)
    frm::CheckError e;
    std::vector<double> v(in.ncheck);
    for (int i = 0; i < in.ncheck; ++i)
        v[i] = e.add(in.check_F[i], in.check_F[i] * 0.9 + (i % 3 == 0 ? 0.5 : -0.25));
    if (in.ncheck > 0)
        frm::write_log_row(stdout, in.check_T[0], in.check_Z.data(), in.check_F.data(), v.data(),
                           in.ncheck, 0.31, 12.5, e);
    return 0;
}
