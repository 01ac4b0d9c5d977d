// frm_solvers.hpp -- C++ solver classes over the C ABI (header-only).
//
// They keep the calc_step() interface and the public state names of the
// reference's classes, so a caller written against TG_solver
// (fr_d_t_g_2d.cpp:2037-2771) or H_solver (mt_minimal.cpp:276-1048,
// mt_psi_caputo.cpp:369-1507) can switch to the GPU path:
//   frm::TG_solver s(desc, C0);  s.calc_step();  s.fetch();  s.b_C[idx]
// Errors from the ABI are thrown as frm::error (the reference has no status
// codes; its callers test results with finite()).
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "frm.h"

namespace frm {

struct error : std::runtime_error {
    int status;
    error(int st, const std::string& what) : std::runtime_error(what), status(st) {}
};

inline void check(int st, const frm_ctx* ctx) {
    if (st != FRM_OK) throw error(st, std::string(frm_strerror(st)) + ": " + frm_last_error(ctx));
}

// One device + one stream.
class Context {
   public:
    explicit Context(int device = 0) { check(frm_ctx_create(device, &ctx_), nullptr); }
    ~Context() { frm_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    frm_ctx* get() const { return ctx_; }
    void sync() { check(frm_sync(ctx_), ctx_); }

   private:
    frm_ctx* ctx_ = nullptr;
};

// TG_solver (2D LOD, g(t) = t, full history) on the GPU.  b_C mirrors the
// reference's field (sN+2)*(M+2), idx(i,j) = i + j*(sN+2); it is refreshed
// from the device by fetch() (the reference's fr_ocl_get, 4386-4392).
class TG_solver {
   public:
    TG_solver(Context& ctx, const frm_lod2d_desc& d, const double* C0)
        : ctx_(ctx), sN(d.sN), M(d.M), b_C((size_t)(d.sN + 2) * (d.M + 2)) {
        check(frm_lod2d_create(ctx.get(), &d, &s_), ctx.get());
        check(frm_lod2d_set_state(s_, C0), ctx.get());
        b_C.assign(C0, C0 + b_C.size());
    }
    ~TG_solver() { frm_lod2d_destroy(s_); }
    TG_solver(const TG_solver&) = delete;
    TG_solver& operator=(const TG_solver&) = delete;

    void calc_step() { check(frm_lod2d_step(s_, 1), ctx_.get()); ++tstep; }
    void calc_steps(int n) { check(frm_lod2d_step(s_, n), ctx_.get()); tstep += n; }
    void fetch() { check(frm_lod2d_get_state(s_, b_C.data()), ctx_.get()); }

    const int sN, M;
    int tstep = 0;
    std::vector<double> b_C;

   private:
    Context& ctx_;
    frm_lod2d* s_ = nullptr;
};

// A batch of H_solver problems; b_U[p] mirrors problem p's U[0..N+1].
class H_batch {
   public:
    H_batch(Context& ctx, const frm_mt1d_desc& d, const std::vector<std::vector<double>>& U0)
        : ctx_(ctx), b_U(U0) {
        check(frm_mt1d_create(ctx.get(), &d, &s_), ctx.get());
        std::vector<double> packed;
        for (const auto& u : U0) packed.insert(packed.end(), u.begin(), u.end());
        check(frm_mt1d_set_state(s_, packed.data()), ctx.get());
    }
    ~H_batch() { frm_mt1d_destroy(s_); }
    H_batch(const H_batch&) = delete;
    H_batch& operator=(const H_batch&) = delete;

    void calc_step() { check(frm_mt1d_step(s_, 1), ctx_.get()); ++tstep; }
    void fetch() {
        size_t n = 0;
        for (const auto& u : b_U) n += u.size();
        std::vector<double> packed(n);
        check(frm_mt1d_get_state(s_, packed.data()), ctx_.get());
        size_t o = 0;
        for (auto& u : b_U) {
            std::copy(packed.begin() + o, packed.begin() + o + u.size(), u.begin());
            o += u.size();
        }
    }

    // kind FRM_MT_AB: solve_and_check's error measure per problem
    std::vector<double> ab_error() {
        std::vector<double> e(b_U.size());
        check(frm_mt1d_ab_error(s_, e.data()), ctx_.get());
        return e;
    }

    // The PSO / steepest-descent drivers (fit_and_solve, mt_minimal.cpp:1401)
    // score every particle with solve_and_test (1277-1330); with one problem
    // per particle, one batch scores a whole iteration:
    //   H_batch b(ctx, desc_of_particles, U0s); b.set_observations(T, Z, F);
    //   for (...) b.calc_step();  auto err = b.fit_error();
    void set_observations(const std::vector<double>& T, const std::vector<double>& Z,
                          const std::vector<double>& F) {
        check(frm_mt1d_set_observations(s_, (int32_t)T.size(), T.data(), Z.data(), F.data()),
              ctx_.get());
    }
    std::vector<double> fit_error() {
        std::vector<double> e(b_U.size());
        check(frm_mt1d_fit_error(s_, e.data()), ctx_.get());
        return e;
    }

    int tstep = 1;  // the reference's H_solver starts at tstep = 1
    std::vector<std::vector<double>> b_U;

   private:
    Context& ctx_;
    frm_mt1d* s_ = nullptr;
};

// The time sums of the 3D irrigation solver (irrigation_fractional.cpp):
// what H_solver::time_summing returns per node (4099-4170) and
// fr_ocl_call_sum computes on the GPU (1828-1905).  The levels stay on the
// device: push() takes each new level as oldH.push_back does (a host array
// of n_nodes values here; frm_timesum_push takes device pointers), sums()
// fills Sums[] for every node from the weights the caller's precalc_kb_3
// produced:
//   TimeSum ts(ctx, nodes, sum_alg, sum_param, max_steps + 1);
//   ts.push(level0);  ...  ts.push(level);  ts.sums(kb_3_cache, tau, Sums);
class TimeSum {
   public:
    TimeSum(Context& ctx, int64_t n_nodes, int sum_alg, double sum_param, int capacity,
            int precision = FRM_PREC_F64)
        : ctx_(ctx), n_(n_nodes) {
        frm_timesum_desc d{n_nodes, capacity, sum_alg, sum_param, precision};
        check(frm_timesum_create(ctx.get(), &d, &s_), ctx.get());
        check(frm_malloc(ctx.get(), (size_t)n_nodes * sizeof(double), &lev_), ctx.get());
        check(frm_malloc(ctx.get(), (size_t)n_nodes * sizeof(double), &out_), ctx.get());
    }
    ~TimeSum() {
        frm_timesum_destroy(s_);
        frm_free(ctx_.get(), lev_);
        frm_free(ctx_.get(), out_);
    }
    TimeSum(const TimeSum&) = delete;
    TimeSum& operator=(const TimeSum&) = delete;

    void push(const double* level) {
        check(frm_upload(ctx_.get(), lev_, level, (size_t)n_ * sizeof(double)), ctx_.get());
        check(frm_timesum_push(s_, (const double*)lev_), ctx_.get());
    }
    int old_size() const {
        int32_t v = 0;
        check(frm_timesum_levels(s_, &v), ctx_.get());
        return v;
    }
    // sum_alg 0 and 1: kb_3_cache[t] and tau[t] for t < old_size - 1
    void sums(const double* kb_3_cache, const double* tau, double* Sums) {
        frm_timesum_args a{};
        a.kb_3_cache = kb_3_cache;
        a.tau = tau;
        eval(a, Sums);
    }
    // sum_alg 2: row2_precalc[time_i1][0], g(Time[tstep+1]), g(Time[tstep]),
    // alpha and tau[tstep]
    void sums_series(const std::vector<double>& coefs, double gtj, double gt, double alpha,
                     double tau_n, double* Sums) {
        frm_timesum_args a{};
        a.coefs = coefs.data();
        a.ncoefs = (int32_t)coefs.size();
        a.gtj = gtj;
        a.gt = gt;
        a.alpha = alpha;
        a.tau_n = tau_n;
        eval(a, Sums);
    }

   private:
    void eval(const frm_timesum_args& a, double* Sums) {
        check(frm_timesum_eval(s_, &a, (double*)out_), ctx_.get());
        check(frm_download(ctx_.get(), Sums, out_, (size_t)n_ * sizeof(double)), ctx_.get());
    }
    Context& ctx_;
    int64_t n_;
    frm_timesum* s_ = nullptr;
    void* lev_ = nullptr;
    void* out_ = nullptr;
};

}  // namespace frm
