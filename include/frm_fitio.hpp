// frm_fitio.hpp -- the reference's fitting input file and result rows
// (header-only, host code; no GPU calls).
//
// read_fit_input reads the fit_input.txt format of the 1D fitting drivers
// (mt_minimal.cpp:1955-2043, the same block in mt_psi_caputo.cpp's main):
//   to_fit nvalues ninit ncheck nperc nirr nEV_T nEV_F ev_mu
//   pso_n pso_o pso_fi_p pso_fi_g pso_eps pso_max_iter
//   nvalues x (T Z F)        values to fit, F scaled by Hmult
//   ninit   x (Z F)          initial H, F scaled by Hmult
//   nperc   x (T A)          precipitation
//   nirr    x (T A)          irrigation
//   ncheck  x (T Z F)        values to check, F scaled by Hmult
//   nEV_T   x (T LAI EV_1 .. EV_nEV_F)   evapotranspiration (or from et_file)
//   up to 20 fixed parameter values      (or from pv_file)
// Fields are whitespace-separated and read with the conversions the
// reference uses (%d, %lg), so a file the reference accepts parses to the
// same values here; where the reference's reader returns 0 (a short or
// malformed file) this returns false.  Differences: a file that cannot be
// opened is an error here (the reference dereferences the null FILE*), an
// et_file that cannot be opened falls back to the main file without the
// reference's fclose(NULL), and exactly 19 parameter values give 19 here
// (the reference's counter then reports 20, the last one unset).
//
// write_res_rows writes one time level of a result file (res_*.txt) as the
// checking run does (mt_minimal.cpp:1631-1635, 1697-1702): N+1 rows
// "t z U wetness " with z = i dL, %g formatting.  The soil model behind
// `wetness` is the caller's.
//
// CheckError accumulates the checking run's error measure (1655-1684): a
// model value v at a check point is clamped to v <= 0, err += (F - v)^2,
// and the relative term is 1 where F == 0 and v != 0, min(|(F - v)/F|, 1)
// where F != 0 (nothing where both are 0).  write_log_row writes one row of
// the checking log (log_*.txt, 1678-1704): "T", the "Z F v" triplets of the
// row, then "rlw .. twc .. sum_err .. avg_err .. avg_rel_err ..", %g
// formatting; rlw and twc (root-layer wetness and total water content) come
// from the caller's soil model, as do the model values (the depth and time
// interpolation of 1644-1668 is the caller's, as in frm_mt1d_fit_error).
#pragma once
#include <cmath>
#include <cstdio>
#include <vector>

namespace frm {

struct FitInput {
    int to_fit = 0, nvalues = 0, ninit = 0, ncheck = 0, nperc = 0, nirr = 0, nEV_T = 0,
        nEV_F = 0;
    double ev_mu = 0.0;
    int pso_n = 0, pso_max_iter = 0;
    double pso_o = 0.0, pso_fi_p = 0.0, pso_fi_g = 0.0, pso_eps = 0.0;
    std::vector<double> values_T, values_Z, values_F;
    std::vector<double> init_Z, init_F;
    std::vector<double> perc_T, perc_A, irr_T, irr_A;
    std::vector<double> check_T, check_Z, check_F;
    std::vector<double> EV_T, LAI;
    std::vector<std::vector<double>> EV_F;  // [nEV_F][nEV_T]
    std::vector<double> param_values;       // at most 20
};

namespace fitio_detail {

inline bool read_rows(FILE* f, int n, int cols, double scale_last,
                      std::vector<double>* out[3]) {
    for (int c = 0; c < cols; ++c) out[c]->assign((size_t)(n > 0 ? n : 0), 0.0);
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < cols; ++c)
            if (std::fscanf(f, " %lg", &(*out[c])[i]) != 1) return false;
        (*out[cols - 1])[i] *= scale_last;
    }
    return true;
}

}  // namespace fitio_detail

inline bool read_fit_input(const char* path, double Hmult, FitInput& in,
                           const char* et_file = nullptr, const char* pv_file = nullptr) {
    using fitio_detail::read_rows;
    FILE* f = std::fopen(path, "rt");
    if (!f) return false;
    bool ok = std::fscanf(f, "%d %d %d %d %d %d %d %d %lg", &in.to_fit, &in.nvalues, &in.ninit,
                          &in.ncheck, &in.nperc, &in.nirr, &in.nEV_T, &in.nEV_F,
                          &in.ev_mu) == 9 &&
              std::fscanf(f, " %d %lg %lg %lg %lg %d", &in.pso_n, &in.pso_o, &in.pso_fi_p,
                          &in.pso_fi_g, &in.pso_eps, &in.pso_max_iter) == 6;
    if (ok) {
        std::vector<double>* v[3] = {&in.values_T, &in.values_Z, &in.values_F};
        std::vector<double>* ini[3] = {&in.init_Z, &in.init_F, nullptr};
        std::vector<double>* pr[3] = {&in.perc_T, &in.perc_A, nullptr};
        std::vector<double>* ir[3] = {&in.irr_T, &in.irr_A, nullptr};
        std::vector<double>* ck[3] = {&in.check_T, &in.check_Z, &in.check_F};
        ok = read_rows(f, in.nvalues, 3, Hmult, v) && read_rows(f, in.ninit, 2, Hmult, ini) &&
             read_rows(f, in.nperc, 2, 1.0, pr) && read_rows(f, in.nirr, 2, 1.0, ir) &&
             read_rows(f, in.ncheck, 3, Hmult, ck);
    }
    if (ok) {
        // evapotranspiration rows: T, LAI, then one value per EV column
        FILE* ef = et_file ? std::fopen(et_file, "rt") : nullptr;
        FILE* src = ef ? ef : f;
        const size_t nt = (size_t)(in.nEV_T > 0 ? in.nEV_T : 0);
        in.EV_T.assign(nt, 0.0);
        in.LAI.assign(nt, 0.0);
        in.EV_F.assign((size_t)(in.nEV_F > 0 ? in.nEV_F : 0), std::vector<double>(nt, 0.0));
        for (size_t i = 0; ok && i < nt; ++i) {
            ok = std::fscanf(src, " %lg", &in.EV_T[i]) == 1 &&
                 std::fscanf(src, " %lg", &in.LAI[i]) == 1;
            for (size_t j = 0; ok && j < in.EV_F.size(); ++j)
                ok = std::fscanf(src, " %lg", &in.EV_F[j][i]) == 1;
        }
        if (ef) std::fclose(ef);
    }
    if (ok) {
        // fixed parameter values: as many numbers as follow, at most 20
        FILE* pf = pv_file ? std::fopen(pv_file, "rt") : nullptr;
        FILE* src = pf ? pf : f;
        in.param_values.clear();
        double x;
        while (in.param_values.size() < 20 && std::fscanf(src, " %lg", &x) == 1)
            in.param_values.push_back(x);
        if (pf) std::fclose(pf);
    }
    std::fclose(f);
    return ok;
}

inline void write_res_rows(FILE* f, double t, double dL, const double* U, const double* wetness,
                           int N) {
    for (int i = 0; i < N + 1; i++)
        std::fprintf(f, "%g %g %g %g \n", t, (double)i * dL, U[i], wetness[i]);
    std::fflush(f);
}

struct CheckError {
    double err = 0.0;   // sum of (F - v)^2
    double srel = 0.0;  // sum of the relative terms
    int n = 0;          // check values counted
    // adds one check value F against the model value v; returns v as counted
    // (clamped to v <= 0)
    double add(double F, double v) {
        if (v > 0.0) v = 0.0;
        err += (F - v) * (F - v);
        if ((F == 0) && (v != 0.0)) srel += 1.0;
        if (F != 0) srel += std::fabs((F - v) / F) < 1 ? std::fabs((F - v) / F) : 1.0;
        n++;
        return v;
    }
};

// One log row: T, then n triplets (Z[i], F[i], v[i]), then the soil-model
// figures and the accumulated error measure.
inline void write_log_row(FILE* f, double T, const double* Z, const double* F, const double* v,
                          int n, double rlw, double twc, const CheckError& e) {
    std::fprintf(f, "%g ", T);
    for (int i = 0; i < n; i++) std::fprintf(f, "%g %g %g ", Z[i], F[i], v[i]);
    std::fprintf(f, "rlw %g twc %g sum_err %g avg_err %g avg_rel_err %g\n", rlw, twc, e.err,
                 (e.n ? e.err / e.n : 0), (e.n ? e.srel / e.n : 0));
}

}  // namespace frm
