// Applies the ia/ibk plan of frm_iaibk.cpp to one node's increments on the
// CPU, with the GPU kernel's operation semantics, and prints the unscaled
// sum V of every step n = 1..N (for tests/test_iaibk_plan.py, which
// compares them with the oracle's restatement of calc_va).
//   iaibk_check <alpha> <dir: 0 rows, 1 columns>  < F (N doubles, text)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "frm_iaibk.hpp"

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    const double alpha = std::atof(argv[1]);
    const int dir = std::atoi(argv[2]);
    std::vector<double> F;
    double v;
    while (std::scanf("%lf", &v) == 1) F.push_back(v);
    const int N = (int)F.size();
    const frm::IaPlan P = frm::iaibk_plan(1.0 - alpha, dir ? 2.0 : 1.5, dir ? 1.0 : 0.5, N);
    std::vector<double> c((size_t)P.nslots * frm::kIaMaxAk, 0.0);
    for (int n = 1; n <= N; ++n) {
        const double f = F[n - 1];
        double V = 0.0;
        for (int o = P.step_off[n - 1]; o < P.step_off[n]; ++o) {
            const frm::IaOp& op = P.ops[o];
            double* cs = &c[(size_t)op.slot * frm::kIaMaxAk];
            const double* w = &P.vals[op.off];
            switch (op.kind) {
                case frm::kIaEvalIa:
                case frm::kIaEvalIbk: {
                    double r = 0.0;
                    for (int k = 0; k < op.nt; ++k) r += cs[k] * w[k];
                    V += r;
                    break;
                }
                case frm::kIaExact: V += w[0] * f; break;
                case frm::kIaCreate:
                    for (int k = 0; k < frm::kIaMaxAk; ++k) cs[k] = k < op.nt ? P.ak[k] * (f * w[k]) : 0.0;
                    break;
                case frm::kIaAdd:
                    for (int k = 0; k < op.nt; ++k) cs[k] += P.ak[k] * (f * w[k]);
                    break;
                case frm::kIaCopy:
                    for (int k = 0; k < frm::kIaMaxAk; ++k) cs[k] = c[(size_t)op.src * frm::kIaMaxAk + k];
                    break;
                default: return 3;
            }
        }
        std::printf("%.17g\n", V);
    }
    std::fprintf(stderr, "slots %d ops %zu vals %zu\n", P.nslots, P.ops.size(), P.vals.size());
    return 0;
}
