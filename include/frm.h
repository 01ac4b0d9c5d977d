/* frm.h -- C ABI of the MI355X-native fractional-memory hot path.
 *
 * The reference (sevaboh/theses_code) keeps this path inside C++ solver
 * classes: TG_solver::calc_step (fr_d_t_g_2d.cpp:2552), H_solver::calc_step
 * (mt_minimal.cpp:1032, mt_psi_caputo.cpp:1435) and the step loop of
 * solve_and_check (MF_int_series.cpp:909-954), with an OpenCL/CUDA runtime
 * wrapper (opencl_cuda_base.cpp) underneath.  This header is the boundary
 * those classes call instead: plain C, plain pointers and sizes, int status
 * (0 = FRM_OK, negative = error), no exceptions across it.
 *
 * Threading: one frm_ctx owns one HIP stream; calls on one context are not
 * thread-safe, calls on different contexts are.  Device buffers passed in are
 * owned by the caller; objects created here (history stores, solvers) own
 * their device memory until destroyed.
 */
#ifndef FRM_H
#define FRM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    FRM_OK = 0,
    FRM_E_INVALID = -1, /* bad argument or shape */
    FRM_E_HIP = -2,     /* HIP runtime error (message in frm_last_error) */
    FRM_E_NOMEM = -3,   /* device allocation failed */
    FRM_E_RANGE = -4,   /* history capacity or size limit exceeded */
    FRM_E_STATE = -5    /* call not valid in the object's current state */
};

/* memory-term precision forms */
enum {
    FRM_PREC_F64 = 0,      /* plain FP64 history and FP64 accumulation */
    FRM_PREC_F32C = 1,     /* FP32 history, compensated (Kahan) FP32 sums */
    FRM_PREC_MFMA_BF16 = 2, /* blocked Toeplitz on MFMA, split hi+lo BF16, FP32 acc */
    FRM_PREC_SOE = 3,       /* sum-of-exponentials recurrences (series history) */
    FRM_PREC_BINOMIAL = 4,  /* the reference's single-series expansion of the history
                               (TG_solver dt_alg = 3, 2D LOD Caputo only) */
    FRM_PREC_IAIBK = 5      /* the reference's per-node "long" (ia) and "short" (ibk)
                               series approximations of the history (TG_solver
                               dt_alg = 4, fr_d_t_g_2d.cpp:246-741; 2D LOD Caputo,
                               g(t) = t) */
};

/* kernel-class ids for frm_timing_get */
enum {
    FRM_K_MEMTERM = 0, /* history sum (the dominant kernel) */
    FRM_K_ROWS = 1,    /* row half-step sweep */
    FRM_K_COLS = 2,    /* column half-step sweep (+ history push) */
    FRM_K_SWEEP1D = 3, /* batched 1D sweep */
    FRM_K_OTHER = 4,
    FRM_K_BLOCKFAR = 5, /* blocked Toeplitz far-part product on MFMA (algorithmic
                           bytes: 4 B per settled slot and node read, 4 B per F row
                           and node written) */
    FRM_K_COUNT = 6
};

typedef struct frm_ctx frm_ctx;
typedef struct frm_hist frm_hist;
typedef struct frm_lod2d frm_lod2d;
typedef struct frm_mt1d frm_mt1d;

/* ------------------------------------------------- runtime (G1 row) ----
 * Replaces OpenCL_program / OpenCL_commandqueue / OpenCL_buffer of
 * opencl_cuda_base.cpp:23-455: device selection, one in-order stream,
 * buffers and transfers.  Unlike the reference (a Finish() after every
 * launch, fr_d_t_g_2d.cpp:4046-4048) work stays asynchronous until
 * frm_sync or a host read-back. */
int frm_ctx_create(int device, frm_ctx** out);
int frm_ctx_destroy(frm_ctx* ctx);
int frm_sync(frm_ctx* ctx);
const char* frm_strerror(int status);
const char* frm_last_error(const frm_ctx* ctx);
int frm_malloc(frm_ctx* ctx, size_t bytes, void** dptr);
int frm_free(frm_ctx* ctx, void* dptr);
int frm_upload(frm_ctx* ctx, void* dptr, const void* hptr, size_t bytes);
int frm_download(frm_ctx* ctx, void* hptr, const void* dptr, size_t bytes);
/* Per-kernel-class timing with HIP events recorded on the context's stream
 * around every launch of the classes in `mask` (bit FRM_K_*); 0 disables. */
int frm_timing_enable(frm_ctx* ctx, uint32_t mask);
int frm_timing_reset(frm_ctx* ctx);
/* Synchronises, then returns launches, total milliseconds and the algorithmic
 * bytes those launches were charged with. */
int frm_timing_get(frm_ctx* ctx, int kclass, int64_t* launches, double* total_ms,
                   double* bytes);

/* ------------------------------------------ history store (H1-H3, H5) ----
 * Device-resident list of per-node increments Delta_s (s = 0..depth-1) of
 * n_nodes values.  Replaces the per-step heap snapshots oldH.push_back
 * (mt_minimal.cpp:1025-1030), oldC.push_back (fr_d_t_g_2d.cpp:2617-2622,
 * 2682-2691) and the save1/save2 kernels (fr_d_t_g_2d.cpp:1947-1962):
 * the sums only ever read differences of consecutive snapshots, so the
 * store keeps the differences, one read per term. */
int frm_hist_create(frm_ctx* ctx, int64_t n_nodes, int32_t capacity, int32_t precision,
                    frm_hist** out);
int frm_hist_destroy(frm_hist* h);
/* Appends Delta = d_new - d_old (device pointers, n_nodes values each). */
int frm_hist_push(frm_ctx* ctx, frm_hist* h, const double* d_new, const double* d_old);
int frm_hist_depth(const frm_hist* h, int32_t* depth);
int frm_hist_clear(frm_hist* h);

/* Memory term over the store (H1/H2/H3/H5 history loops):
 *   out[t*n_nodes + i] (+)= scale * sum_{s=0}^{depth-1} w[t*ldw + (depth-1-s)] * Delta_s[i]
 * for targets t < ntargets (<= 2), i.e. w is indexed by lag-1 with w[0] the
 * newest increment's weight.  d_w and d_out are device pointers (FP64). */
typedef struct frm_memterm_desc {
    int32_t ntargets;   /* 1 or 2 (the two LOD half-step targets share one pass) */
    int32_t ldw;        /* leading dimension of d_w (>= depth) */
    const double* d_w;  /* lag-indexed FP64 weights */
    double scale;
    int32_t accumulate; /* 0: out = ..., 1: out += ... */
} frm_memterm_desc;
int frm_memterm(frm_ctx* ctx, frm_hist* h, const frm_memterm_desc* desc, double* d_out);
/* frm_memterm for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It has no counterpart in the reference.  The same kernels and results as
 * frm_memterm, bit for bit, but the partial sums of its time chunks go to a
 * static per-GPU array of 16 MiB (2^21 FP64 values) instead of the store's
 * workspace, and the call allocates nothing and records no timing events.
 * FRM_E_RANGE when time chunks x ntargets x padded nodes exceed 2^21 (the
 * one-chunk case fits up to 2^20 nodes with two targets, 2^21 with one).
 * Every call on one GPU shares the array: calls that overlap on one GPU
 * (other streams or contexts) race. */
int frm_memterm_static(frm_ctx* ctx, frm_hist* h, const frm_memterm_desc* desc, double* d_out);

/* ------------------------------------------------ batched tridiagonal ----
 * a[k] x[k-1] + b[k] x[k] + c[k] x[k+1] = d[k], k = 0..n-1 (a[0], c[n-1]
 * ignored), nsys independent systems; element k of system s at
 * s*stride_sys + k*stride_elem.  Replaces the al1/bt1/back-substitution
 * triple (mt_minimal.cpp:829-858, 942-1024) and the OpenCL Al/Bt/U kernels
 * (fr_d_t_g_2d.cpp:1604-1658, 1896-1922).  Thomas (progonka) recurrences,
 * one wavefront per system, chunk seeds from wavefront scans. */
typedef struct frm_tridiag_desc {
    int32_t nsys, n;
    int64_t stride_sys, stride_elem;
} frm_tridiag_desc;
int frm_tridiag(frm_ctx* ctx, const frm_tridiag_desc* desc, const double* a, const double* b,
                const double* c, const double* d, double* x);

/* memory kernels of the 2D LOD solver */
enum {
    FRM_KERNEL_CAPUTO = 0, /* (t-s)^-alpha / Gamma(1-alpha): fr_d_t_g_2d.cpp kb(), g(t) = t */
    FRM_KERNEL_AB = 1      /* E_alpha(-alpha/(1-alpha) (t-s)^alpha) / (1-alpha): MF_int_series.cpp */
};

/* --------------------------------------- 2D LOD solver (TG_solver) ------
 * fr_d_t_g_2d.cpp TG_solver_<func=0 (g(t)=t), fixed_d, kform=0> with
 * dt_alg = 0 (full history).  Grid nodes i = 0..sN+1, j = 0..M+1 at
 * idx(i,j) = i + j*(sN+2) (fr_d_t_g_2d.cpp:2003).  Row systems j = 1..M-1
 * (Dirichlet C1 at i = 0, zero gradient at i = sN), column systems
 * i = 1..sN-1 (zero gradient at both ends).  The sweep coefficients are the
 * lower (A_) and upper (B) off-diagonals per node and direction
 * (fr_d_t_g_2d.cpp:3242-3265); the diagonal is (sigma/tau)*kb + A_ + B.
 * Caputo weights kb(s, ja, sa, l) are the kernel's integrals over the LOD
 * sub-intervals (fr_d_t_g_2d.cpp:3112-3131).  With FRM_KERNEL_AB both
 * half-steps use the full-step weights of MF_intF_simple (MF_int_series.cpp:
 * 614-718), I(d) = MF_int_choice(d tau, tau, 0, alpha, alpha, ab_maxl,
 * ab_maxkiter, ab_eps) / (1 - alpha): current weight I(1), history weights
 * I(k+1)/2 for the increment k steps back, no half-step correction; the two
 * half-steps add up to the full-step AB discretisation (a half-interval
 * current weight, as for Caputo, would make the constant mode grow because
 * the AB kernel is bounded). */
typedef struct frm_lod2d_desc {
    int32_t sN, M;
    int32_t capacity;  /* max number of steps (history depth) */
    int32_t precision; /* FRM_PREC_* for the memory term (FRM_PREC_F32C: FP32
                          increments and weights, both targets' sums
                          Kahan-compensated in FP32, 4 B per node and slot) */
    double alpha, tau, sigma, C1, S;
    const double* rowA; /* host arrays of (sN+2)*(M+2) values */
    const double* rowB;
    const double* colA;
    const double* colB;
    int32_t kernel;      /* FRM_KERNEL_*; 0 = Caputo */
    int32_t ab_maxl;     /* AB series terms / quadrature nodes (0: 50) */
    int32_t ab_maxkiter; /* AB inner series limit (0: 1000) */
    double ab_eps;       /* AB series tolerance (0: 1e-10) */
    int32_t ab_exact;    /* 1: AB weights to FP64 accuracy instead of the MF_int_choice methods */
    int32_t soe_terms;   /* FRM_PREC_SOE: exponentials (0: 32, max 64); a
                            history of fewer lags than terms is fitted by
                            capacity - 1 of them.  Creation fails with
                            FRM_E_INVALID when the fit misses the weights by
                            more than 1e-6 relative (e.g. the truncated AB
                            weights without ab_exact) */
    int32_t soe_f32;     /* FRM_PREC_SOE: 1 stores the mode states in FP32 */
    int32_t mfma_parts;  /* FRM_PREC_MFMA_BF16: BF16 parts per operand, 2 (three
                            products, ~2^-16) or 3 (six products, exact FP32
                            increments, ~2^-24); 0: 2 for Caputo, 3 for AB */
    int32_t sweep_w;     /* wavefronts per sweep line (1, 2, 4); 0: chosen at
                            creation by timing each (the autotuner) */
    int32_t series_terms; /* FRM_PREC_BINOMIAL: moments per node and direction
                             (the reference's dt_eps; 0: 25, max 1000) */
    double series_tol;    /* FRM_PREC_BINOMIAL: int_err, where _kb_row2 stops
                             taking coefficients (0: 1e-5, the reference's default) */
    double trunc_eps;     /* FRM_PREC_F64, Caputo: the reference's restricted
                             summation (dt_alg = 2, fr_d_t_g_2d.cpp:3356-3381): each
                             direction's history sum, newest first, stops at the
                             first weight below trunc_eps; 0: the full history */
    int32_t g_func;       /* time function g of the kernel (g(t) - g(s))^-alpha:
                             0 = t (default); 3 = t^g_pow (TG_solver_<3, ...>,
                             fr_d_t_g_2d.cpp:2823-2834).  With 3 the weights are
                             not Toeplitz and the current-step weights change every
                             step: every step rebuilds the history weights on the
                             GPU (_kb_row series, 2937-2963, int_alg 1) and the
                             sweeps' Thomas coefficients; FRM_PREC_F64 and the
                             Caputo kernel only */
    double g_pow;         /* g_func 3: the power (0 < g_pow <= 4) */
    double int_err;       /* g_func 3: the series' term tolerance (the reference's
                             int_err; 0: 1e-6, its default, 5322) */
} frm_lod2d_desc;
int frm_lod2d_create(frm_ctx* ctx, const frm_lod2d_desc* desc, frm_lod2d** out);
int frm_lod2d_destroy(frm_lod2d* s);
/* Initial field (host, (sN+2)*(M+2)); resets the step counter and history,
 * as TG_solver::initial() pushes C_0 (fr_d_t_g_2d.cpp:2294-2297). */
int frm_lod2d_set_state(frm_lod2d* s, const double* h_C);
/* nsteps calls of calc_step(), queued on the context's stream. */
int frm_lod2d_step(frm_lod2d* s, int32_t nsteps);
int frm_lod2d_get_state(frm_lod2d* s, double* h_C);
int frm_lod2d_tstep(const frm_lod2d* s, int32_t* tstep);
/* Device pointer of the state (reference layout), for callers that keep
 * the field on the GPU. */
int frm_lod2d_state_ptr(frm_lod2d* s, double** d_C);
/* FRM_PREC_SOE: largest relative error of the fitted exponential sum over
 * the history weights it replaces (lags 2..capacity); 0 for full history. */
int frm_lod2d_fit_error(const frm_lod2d* s, double* max_rel);
/* Host copies of the weights in use: out[0] = row current weight, out[1] =
 * column current weight, out[2] = half-step correction, then the row and
 * column history weights for lags 1..n (2 + 1 + 2n values).  With g_func 3
 * the weights change every step: those of step n, computed on the host as
 * the step computes them (lag k multiplies the increment of step n - k). */
int frm_lod2d_weights(const frm_lod2d* s, int32_t n, double* out);
/* Launch shapes of the sweeps in use: out[0..3] = elements per lane and
 * wavefronts per line of the row sweep, then of the column sweep. */
int frm_lod2d_sweep_shape(const frm_lod2d* s, int32_t* out);
/* Node slices in which the blocked Toeplitz form (FRM_PREC_MFMA_BF16)
 * forms a 64-step block's far product during the block before it, one
 * every 64 / slices steps; each slice is about one round of the product's
 * workgroups on the device (0 for the other forms). */
int frm_lod2d_far_slices(const frm_lod2d* s, int32_t* slices);

/* ------------------------- batched 1D H_solver (mt_minimal / mt_psi_caputo)
 * A batch of independent 1D problems stepped together, each one
 * H_solver::calc_step (mt_minimal.cpp:1032-1038, mt_psi_caputo.cpp:1435-1497)
 * with gamma = 1: nodes i = 0..N, ghost N+1, dL = L/N,
 *   A_(i) U_{i-1} - R(i) U_i + B(i) U_{i+1} = Om_i,  i = 1..N-1,
 *   A_ = Dg KK(i-1)/C(i), B = Dg KK(i+1)/C(i), R = (KK(i+1)+KK(i-1)) Dg/C(i) + d0,
 *   Om = -d0 U^n + (KK(i)-KK(i-1))/dL + Da * history   (mt_minimal.cpp:731-743, 860-940)
 * with d0 = Da (kind L1: Da = tau^-alpha/Gamma(2-alpha)) or
 * Da (1-alpha) kb_time[tstep] (kind PSI: Da = tau^-1/Gamma(2-alpha)).
 * Kind PSI forms kb_time every step on the GPU by calc_int's series
 * (mt_psi_caputo.cpp:720-740), and by its _kb quadrature fallback where a
 * series is not finite; frm_mt1d_get_state returns FRM_E_STATE if a weight
 * is still not finite.  The weights of step k + 1 are formed on the
 * context's second stream while step k runs (in order while per-launch
 * timing is on or in graph mode; the results are the same).
 * The history sum keeps the reference's index convention: the first two
 * snapshot increments never enter (mt_minimal.cpp:917-921).  Coefficient
 * forms: CONST (KK = D, C = 1), CV2 (KK = k, 1/C piecewise in U as cv_form 2,
 * mt_minimal.cpp:499-503), KEXP (KK = k0 exp(b U), C = 1).  `picard` > 1
 * re-solves the step with coefficients from the latest iterate (1 = the
 * reference's lagged coefficients).  Problem p's nodes are packed at offset
 * sum_{q<p} (N_q + 2) of the state array. */
enum { FRM_MT_L1 = 0, FRM_MT_PSI = 1, FRM_MT_AB = 2 };
enum { FRM_FORM_CONST = 0, FRM_FORM_CV2 = 1, FRM_FORM_KEXP = 2 };
enum { FRM_TOP_AUTO = 0, FRM_TOP_DIRICHLET = 1, FRM_TOP_FLUX = 2 };

typedef struct frm_mt1d_problem {
    int32_t N;           /* intervals (nodes 0..N), 2 <= N <= 1025 */
    int32_t form;        /* FRM_FORM_* */
    int32_t top;         /* FRM_TOP_*: AUTO is the reference's rule (al1/bt1) */
    int32_t bottom_cond; /* 1: U_N = Hbottom; 0: flux rule of U1 (mt_minimal.cpp:987-1020) */
    double L, alpha, beta; /* beta: psi(t) = t^beta for kind PSI */
    double D, k, k0, b;
    double top_val;      /* Dirichlet value for FRM_TOP_DIRICHLET */
    double Ev;           /* top flux term of Uc() (evaporation minus inflow) */
    double Hbottom;
    /* kind FRM_MT_AB only (see below) */
    int32_t src;         /* manufactured solution F_func: 0, 1 or 2 */
    int32_t ab_maxl;     /* MF_int_choice series terms / quadrature nodes (0: 1000) */
    int32_t ab_maxkiter; /* inner series limit (0: 1000) */
    int32_t ab_exact;    /* 1: FP64-accurate kernel integrals */
    double ab_eps;       /* series tolerance of the current-step weight (0: 1e-10) */
    int32_t ab_series;   /* 1: the series history of simple = 0 (see below) */
} frm_mt1d_problem;

typedef struct frm_mt1d_desc {
    int32_t nprob;
    int32_t kind;      /* FRM_MT_L1, FRM_MT_PSI or FRM_MT_AB */
    int32_t capacity;  /* max steps */
    int32_t precision; /* FRM_PREC_F64 or FRM_PREC_F32C */
    int32_t picard;    /* Picard iterations per step, >= 1 */
    double tau;
    const frm_mt1d_problem* probs; /* host array of nprob */
    int32_t graph;     /* 1: steps run as HIP graphs of up to 32 steps each:
                          one graph per distinct sequence of kernels (a handful
                          per run), built and instantiated the first time, then
                          replayed with these steps' arguments and grids set on
                          its nodes (hipGraphExecKernelNodeSetParams); kind psi
                          forms the next step's weights on a branch beside the
                          step.  Not while per-launch timing events are on */
    int32_t hist_block; /* history block length: the shortest time chunk, in
                           slots, into which the memory term splits the history
                           when the nodes alone give too few wavefronts (the
                           sweep adds one partial sum per chunk).  0: chosen at
                           creation by timing the memory term and sweep at depth
                           capacity/2 for 16..512 (frm_mt1d_hist_block reports
                           it); > 0: that length */
} frm_mt1d_desc;

int frm_mt1d_create(frm_ctx* ctx, const frm_mt1d_desc* desc, frm_mt1d** out);
int frm_mt1d_destroy(frm_mt1d* s);
/* Packed initial states (host, sum of N_p + 2 values); resets tstep to 1. */
int frm_mt1d_set_state(frm_mt1d* s, const double* h_U);
int frm_mt1d_step(frm_mt1d* s, int32_t nsteps);
int frm_mt1d_get_state(frm_mt1d* s, double* h_U);
int frm_mt1d_tstep(const frm_mt1d* s, int32_t* tstep);
/* the history block length in use (frm_mt1d_desc.hist_block) */
int frm_mt1d_hist_block(const frm_mt1d* s, int32_t* len);

/* Kind FRM_MT_AB: the Atangana-Baleanu test problem of solve_and_check
 * (MF_int_series.cpp:875-958, simple = 1) for each problem of the batch:
 * D_a C = C_xx - C_x + f on [0,1], h = 1/N, C = 0 at t = 0, exact Dirichlet
 * ends C = CC(x, t), full history of (C^{k+1}-C^k)/tau weighted by
 * I(d) = MF_int_choice(d tau, tau, 0) (MF_intF_simple, 614-718), current
 * weight I(1)/(1-alpha) without 1/tau, and the textbook Thomas sweep with
 * A = -1/h^2 - 1/(2h), B = I(1)/(1-alpha) + 2/h^2, Cc = -1/h^2 + 1/(2h).
 * f and CC are the reference's manufactured solutions F_func = src
 * (860-873; their Mittag-Leffler factors are evaluated in FP64 on the host
 * at creation).  Fields used: N, alpha, src, ab_*; form/top/... ignored;
 * picard must be 1; precision FRM_PREC_F64.  frm_mt1d_ab_error returns per
 * problem the reference's error measure sum_steps sum_j (C_j - CC)^2 /
 * ((N+1) steps) over the steps taken since the last set_state.
 *
 * ab_series = 1 takes the history as solve_and_check does with simple = 0
 * (MF_int_series.cpp:917-922): per node, MF_intF_choice's accumulators
 * (MF_intF's power series Lc1 while (a/(1-a)) t^a < 1, MF_intF_Fv's
 * exponential series Lc2 with its log form on overflow, 569-853), fed with
 * (C^{i-1}-C^{i-2})/tau; ab_maxl is the series length (and ab_eps the
 * tolerance of its coefficients).  The node-independent coefficients of
 * every step are formed on the host at creation (ab_maxl * 5 doubles per
 * step and distinct (alpha, ab_maxl, ab_maxkiter, ab_eps)).  ab_exact must
 * be 0.  Where a term of MF_intF's sum is not finite the sum restarts from
 * the sorted pyramid sum of all finite terms (600-609), on the GPU as in the
 * reference, for ab_maxl up to 4096; above that frm_mt1d_get_state returns
 * FRM_E_STATE when it happens. */
int frm_mt1d_ab_error(frm_mt1d* s, double* out);

/* Fitting error of every problem against one set of observations, as the
 * reference's fitting drivers score a particle (solve_and_test,
 * mt_minimal.cpp:1277-1330; the PSO loop calls it once per particle, here
 * the particles are the problems of one batch): for each observation (T, Z,
 * F) with tt <= T < tt + tau in the step that starts at tt (tt accumulated
 * as tt += tau from the last set_state), U at depth Z is interpolated
 * linearly between nodes j, j+1 with dL j < Z <= dL (j+1) before the step
 * (j < N-1) and after it (j < N), blended by k1 = (T - tt)/tau, clamped to
 * <= 0, and (F - v)^2 is added; a problem whose U[0] is not finite after a
 * step scores 1e300 from then on.  Observations: host arrays of nobs
 * values, copied; replacing them or set_state clears the errors. */
int frm_mt1d_set_observations(frm_mt1d* s, int32_t nobs, const double* T, const double* Z,
                              const double* F);
/* Per-problem error (nprob values) over the steps taken since the errors
 * were cleared. */
int frm_mt1d_fit_error(frm_mt1d* s, double* out);

/* ------------------------------------- time sums of the 3D solver (Sum) ----
 * Drop-in for fr_ocl_call_sum and its Sum kernel (irrigation_fractional.cpp:
 * driver 1828-1905, OpenCL kernel 1254-1320, CUDA kernel 1416-1489) and for
 * the CPU loop they replace, H_solver::time_summing (4099-4170), which reads
 * the weights precalc_kb_3 leaves in kb_3_cache (4081-4098).  The time
 * levels (oldH, n_nodes = (sN+2)(_M+2)(K+2) values each) stay on the device;
 * per node id, with old_size levels pushed:
 *   sum_alg 0: S[id] = sum_{t=old_size-2}^{0} kb_3_cache[t] (old[t+1][id] - old[t][id]) / tau[t]
 *   sum_alg 1: the same, newest first, stopping before the first t with
 *              |kb_3_cache[t]| < sum_param (as the CPU loop; the reference's
 *              GPU kernel also adds the first such term)
 *   sum_alg 2: Htmp[id][j] += gt^(j+1) coefs[j] (old[last][id] - old[last-1][id])
 *              for j < min(ceil(sum_param), ncoefs), then
 *              S[id] = _kb_row2_fixed_coefs(gtj, gtj^-alpha, Htmp[id], ceil(sum_param), alpha) / tau_n
 * S is 0 while fewer than two levels exist.  sum_alg 2 folds each increment
 * into the moments once: the first eval after a push folds it, later evals
 * before the next push only re-evaluate; as in the reference, which folds
 * only the newest difference per call, an increment pushed over without an
 * eval never enters the moments.  The reference's double / float /
 * half kernel variants map to FRM_PREC_F64 and FRM_PREC_F32C (FP32 levels
 * differences, compensated sums; sum_alg 0 and 1). */
typedef struct frm_timesum frm_timesum;
typedef struct frm_timesum_desc {
    int64_t n_nodes;
    int32_t capacity;  /* time levels held (sum_alg 0 and 1); push beyond it fails */
    int32_t sum_alg;   /* 0, 1 or 2 */
    double sum_param;  /* sum_alg 1: weight threshold; 2: number of series terms */
    int32_t precision; /* FRM_PREC_F64 or FRM_PREC_F32C (sum_alg 0 and 1) */
} frm_timesum_desc;
typedef struct frm_timesum_args {
    const double* kb_3_cache; /* host, old_size-1 weights by level t (sum_alg 0, 1) */
    const double* tau;        /* host, tau[t] for t < old_size-1 (sum_alg 0, 1) */
    const double* coefs;      /* host, row2_precalc[time_i1][0] (sum_alg 2) */
    int32_t ncoefs;           /* its length (sum_alg 2) */
    double gtj;               /* g(Time[tstep+1]) (sum_alg 2) */
    double gt;                /* g(Time[tstep]) (sum_alg 2) */
    double alpha;             /* time derivative order (sum_alg 2) */
    double tau_n;             /* tau[tstep] (sum_alg 2) */
} frm_timesum_args;
int frm_timesum_create(frm_ctx* ctx, const frm_timesum_desc* desc, frm_timesum** out);
int frm_timesum_destroy(frm_timesum* s);
/* Clears the levels and the moments (a new run on the same grid). */
int frm_timesum_clear(frm_timesum* s);
/* oldH.push_back(level): d_level is a device array of n_nodes values. */
int frm_timesum_push(frm_timesum* s, const double* d_level);
int frm_timesum_levels(const frm_timesum* s, int32_t* old_size);
/* S (device, n_nodes values) for the levels pushed so far. */
int frm_timesum_eval(frm_timesum* s, const frm_timesum_args* args, double* d_S);

/* ------------------------------------------- sum-of-exponentials history ----
 * Per-node exponential accumulators (the series forms H4/H5: the Lc2 sums of
 * MF_intF_Fv, MF_int_series.cpp:720-836, advanced with the newest increment
 * only; the dt_alg 3 moments, fr_d_t_g_2d.cpp:3385-3428).  Per node i and
 * mode l < nterms:
 *   H_l[i] = q[l] * H_l[i] + d_incr[i]                 (skipped if d_incr is NULL)
 *   d_out[t*n_nodes + i] (+)= scale * sum_l h_coef[t*nterms + l] * H_l[i]
 * for targets t < ntargets.  h_q (nterms) and h_coef (ntargets*nterms) are
 * host arrays read at the call; d_incr and d_out are device arrays.  The
 * states start at zero.  frm_host_soe_fit gives q and c for lag-indexed
 * weights. */
typedef struct frm_soe frm_soe;
typedef struct frm_soe_desc {
    int64_t n_nodes;
    int32_t nterms;   /* 1..64 */
    int32_t ntargets; /* 1 or 2 */
    int32_t f32;      /* 1: FP32 mode states (recurrence and sums in FP64) */
} frm_soe_desc;
int frm_soe_create(frm_ctx* ctx, const frm_soe_desc* desc, frm_soe** out);
int frm_soe_destroy(frm_soe* s);
/* Zeroes the states (a new run). */
int frm_soe_clear(frm_soe* s);
int frm_soe_step(frm_ctx* ctx, frm_soe* s, const double* d_incr, const double* h_q,
                 const double* h_coef, double scale, int32_t accumulate, double* d_out);

#ifdef __cplusplus
}
#endif
#endif /* FRM_H */
