// 2D locally-one-dimensional fractional solver (the hot path of
// TG_solver::calc_step, fr_d_t_g_2d.cpp:2552-2718, dt_alg = 0, g(t) = t).
//
// Per step n the reference, for every row j, forms Om (3306-3492) with the
// full history sum over s = n-1..0, builds Al/Bt (3267-3304, 3494-3553) and
// back-substitutes (2548-2550); then the same for every column with the
// extra half-step term kb(n,1,0,1/2)*(C_n - C_{n+1/2}) (3430-3436).  Its
// OpenCL version launches Om, Al, Bt and U separately, with Al/Bt/U on one
// work-item per system and a Finish() after each launch.
//
// Here one step is three launches on one stream:
//   1. memterm: both half-step history sums in one pass over the increments
//      (rows use kb(s,1/2,0,1), columns kb(s,1,0,1); both depend only on the
//      lag n-s for g(t) = t), split along time into FP64 partials -- or, for
//      the chosen forms, the sum-of-exponentials recurrences or the blocked
//      Toeplitz parts (frm_mfma.hip);
//   2. rows: one to four wavefronts per row -- RHS assembly, Thomas forward
//      and back substitution (chunked over lanes, scan seeds), boundary
//      nodes, and a copy of C_n for step 3;
//   3. cols: the same per column, plus the push of Delta_n = C_{n+1} - C_n
//      into the history store.
// With g(t) = t and a time-independent operator the tridiagonal matrices do
// not change between steps, so their Thomas alpha coefficients are computed
// once on the host in FP64, with the reference's formula and order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "frm_gpow.hpp"
#include "frm_history.hpp"
#include "frm_iaibk.hpp"
#include "frm_stream.hpp"
#include "frm_mfma.hpp"
#include "frm_wavescan.hpp"
#include "frm_soe.hpp"
#include "frm_weights.hpp"

namespace frm {
// Launch shape of a sweep: chunk length CM and wavefronts per line W.
struct SweepShape {
    int cm = 0, w = 1;
};
}  // namespace frm

struct frm_lod2d {
    frm_ctx* ctx = nullptr;
    int sN = 0, M = 0, ld = 0;
    int NX = 0, NY = 0;  // interior nodes per row / per column
    int capacity = 0;
    int tstep = 0;
    double alpha = 0, tau = 0, sigma = 0, C1 = 0, S = 0;
    double sg = 0;                  // sigma / tau
    double kr0 = 0, kc0 = 0, kch = 0;  // current-step weights
    frm_hist* hist = nullptr;
    double* C = nullptr;       // state, reference layout
    double* Cprev = nullptr;   // C_n at interior nodes, compact node order
    double* w = nullptr;       // [2][capacity] lag-indexed history weights
    float* w32 = nullptr;      // FRM_PREC_F32C: the same in FP32
    double* rcoef = nullptr;   // rows:    a, f, al interleaved per line (load_coefs)
    double* ccoef = nullptr;   // columns: the same
    frm::SweepShape rsh, csh;  // launch shapes of the row / column sweeps
    std::vector<double> rplain, cplain;  // [3][lines][n] coefficients, for re-layout
    int precision = FRM_PREC_F64;
    std::vector<double> hw;    // host copy: kr0, kc0, kch, then [2][capacity] weights
    // sum-of-exponentials form
    int soe_L = 0, soe_f32 = 0;
    double soe_err = 0.0;
    void* soe_H = nullptr;     // [L][ld] mode states (FP32 or FP64)
    double* soe_q = nullptr;   // [L]
    double* soe_c = nullptr;   // [2][L]
    double soe_w1r = 0.0, soe_w1c = 0.0;  // exact lag-1 weights
    // FRM_PREC_BINOMIAL: moments [2][terms][ld] and per-step coefficient tables
    int bin_terms = 0;
    double bin_Da = 0.0;
    double* bin_H = nullptr;
    double* bin_upd = nullptr;  // [capacity-1][terms]
    double* bin_er = nullptr;
    double* bin_ec = nullptr;
    int* bin_cnt = nullptr;     // [capacity-1][4]
    // restricted summation: lags beyond trunc_L are not read (weights past a
    // direction's own cut-off are zero)
    int trunc_L = 0;
    // blocked MFMA form: FP32 increments in hist, BF16-exact weight parts,
    // far part of the current 64-step block
    void* atab = nullptr;      // weight-operand fragments (atab_build)
    int atab_chunks = 0;
    int mfma_nprod = 6;        // 3 (two parts) or 6 (three parts)
    int far_slices = 8;        // node slices of a block's far product (1 to 64)
    float* F = nullptr;        // [2][128][ld]: far parts of this block and the next
    float* G = nullptr;        // [16][ld] mid part of the current 8-step sub-block
    // g(t) = t^p (g_func 3): the weights and the sweeps' Thomas coefficients
    // are rebuilt every step
    int g_func = 0;
    double g_pow = 1.0, int_err = 1e-6, Da = 0.0;
    double* lineAB = nullptr;  // off-diagonals per line: rows A, B [M-1][NX], cols A, B [sN-1][NY]
    int* flag = nullptr;       // set when a weight series reached its safety bound
    // FRM_PREC_IAIBK: per-node coefficient slots [2][slots][MAXAK][ld] and
    // both directions' operation plans
    double* ia_H = nullptr;
    frm::IaOp* ia_ops[2] = {nullptr, nullptr};
    double* ia_vals[2] = {nullptr, nullptr};
    std::vector<int> ia_off[2];
    int ia_slots = 0;
    double ia_ak[frm::kIaMaxAk] = {};
    double ia_scale = 0.0;     // Da tau^(1-alpha) / (1-alpha)
};

namespace frm {

// ------------------------------------------------------------- kernels
// A line of n unknowns is split into NL = 64 W chunks of CM elements, one
// per lane of the W wavefronts that share the line (W = 1, 2 or 4: the
// number of lanes that run the parallel scan levels before each lane's
// sequential Thomas replay of CM elements).
//
// Coefficient layout: for line L and array c (0: a, 1: f, 2: al), element
// e = g*CM + k of the line (chunk g = 64 wl + lane, position k) is stored at
//   ((L*3 + c)*CM + k)*NL + g,
// so that the k-th load of a coefficient is NL consecutive doubles.
// The state, the partial sums and Cprev are staged through LDS: read in
// element order (consecutive lanes, consecutive nodes), written to LDS at
// e + e/CM (chunk-major with one pad double per chunk, so that the chunk
// reads g*(CM+1) + k spread over the banks), read back per chunk.
template <int CM, int W>
__device__ inline void load_coefs(const double* coef, int line, int g, double (&a)[CM],
                                  double (&f)[CM], double (&al)[CM]) {
    constexpr int NL = kWave * W;
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        a[k] = coef[((int64_t)(line * 3 + 0) * CM + k) * NL + g];
        f[k] = coef[((int64_t)(line * 3 + 1) * CM + k) * NL + g];
        al[k] = coef[((int64_t)(line * 3 + 2) * CM + k) * NL + g];
    }
}

struct RowArgs {
    double* C;
    double* Cprev;
    const double* coef;  // interleaved, see load_coefs
    const double* part;  // partial sums [nch][2][ldh], target 0
    int64_t ldh;
    int nch;
    int ld, lines, n, NX;
    double negsk, S_half, sg, b0, al1;
};

// Rows: 4 wavefronts per workgroup, 4/W lines.
template <int CM, int W>
__global__ __launch_bounds__(256) void k_lod_rows(RowArgs p) {
    constexpr int S = CM + 1;
    constexpr int LPB = 4 / W;
    constexpr int NL = kWave * W;
    __shared__ double lds[LPB][NL * S];
    __shared__ LineShare<W> share[LPB];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = wv / W, wl = wv % W;
    const int line = blockIdx.x * LPB + li;
    const bool live = line < p.lines;  // uniform per line; dead lines still reach the barriers
    const int n = p.n;
    const int g = wl * kWave + lane;
    const int e0 = g * CM;
    double* buf = lds[li];
    const int64_t qbase = (int64_t)line * p.NX;
    double* Crow = p.C + (int64_t)(line + 1) * p.ld;
    // The coefficients do not depend on the staging: their loads go out
    // first and overlap the state loads (up to 16 elements per lane; at 32
    // the registers would spill).
    constexpr bool kEarly = CM <= 16;
    double a[CM], f[CM], al[CM], v[CM];
#pragma unroll
    for (int k = 0; k < CM; ++k) a[k] = f[k] = al[k] = v[k] = 0.0;
    if (kEarly && live) load_coefs<CM, W>(p.coef, line, g, a, f, al);
    if (live) {
        // om in element order: C_n (copied to Cprev) and the history sum.
        // Every load is issued before the first store (clamped, unguarded
        // indices; the pointers do not alias), so the row costs one memory
        // latency instead of one per element group.  The first partial-sum
        // chunk is loaded with the state (0 + p0 = p0 exactly, so the sum
        // order is unchanged); further chunks follow.
        const double* __restrict__ cr = Crow + 1;
        const double* __restrict__ pt = p.part + qbase;
        double* __restrict__ cpv = p.Cprev + qbase;
        const bool any = p.nch > 0;
        double u[CM], ts[CM];
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = min(r * NL + g, n - 1);
            u[r] = cr[e];
            ts[r] = any ? pt[e] : 0.0;
        }
        for (int c = 1; c < p.nch; ++c) {
            const double* __restrict__ pc = pt + (int64_t)c * 2 * p.ldh;
            double x[CM];
#pragma unroll
            for (int r = 0; r < CM; ++r) x[r] = pc[min(r * NL + g, n - 1)];
#pragma unroll
            for (int r = 0; r < CM; ++r) ts[r] += x[r];
        }
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = r * NL + g;
            if (e < n) {
                cpv[e] = u[r];
                double om = p.negsk * u[r];
                om += p.S_half;
                om += p.sg * ts[r];
                buf[e + e / CM] = om;
            }
        }
    }
    __syncthreads();
    if (live) {
        if (!kEarly) load_coefs<CM, W>(p.coef, line, g, a, f, al);
#pragma unroll
        for (int k = 0; k < CM; ++k) v[k] = e0 + k < n ? buf[g * S + k] : 0.0;
    }
    const double un = line_solve_mw<CM, W>(a, f, al, v, e0, n, p.b0, lane, wl, &share[li], live);
    if (live) {
#pragma unroll
        for (int k = 0; k < CM; ++k)
            if (e0 + k < n) buf[g * S + k] = v[k];
    }
    __syncthreads();
    if (live) {
        for (int r = 0; r < CM; ++r) {
            const int e = r * NL + g;
            if (e < n) Crow[e + 1] = buf[e + e / CM];
        }
        if (g == 0) {
            Crow[0] = p.al1 * v[0] + p.b0;
            Crow[n + 1] = un;
            Crow[n + 2] = un;
        }
    }
}

struct ColArgs {
    double* C;
    const double* Cprev;
    double* slot;        // history slot for Delta_n (compact node order)
    float* slotf;        // FP32 history slot instead of slot when non-null
    const double* coef;  // interleaved, see load_coefs
    const double* part;  // target 1 partials at part + ldh
    int64_t ldh;
    int nch;
    int ld, lines, n, NX;
    int groups;          // column groups of col_width lines
    double negsk, S_half, sg, kch, b0, al1;
};

// Column lines per workgroup: a row segment of col_width columns is one
// contiguous piece of the state, Cprev and partial sums; workgroups stay at
// 512 threads and within the LDS the staging needs.
template <int CM, int W>
__host__ __device__ constexpr int col_width() {
    return (CM * W >= 32 ? 4 : 8) / W > 0 ? (CM * W >= 32 ? 4 : 8) / W : 1;
}

template <int CM, int W>
__global__ __launch_bounds__((col_width<CM, W>() * kWave * W)) void k_lod_cols(ColArgs p) {
    constexpr int kColW = col_width<CM, W>();
    constexpr int S = CM + 1;
    constexpr int NL = kWave * W;
    __shared__ double lds[kColW][NL * S];
    __shared__ LineShare<W> share[kColW];
    // Workgroups are dealt to the 8 XCDs round-robin (b -> XCD b % 8); the
    // group order below gives each XCD a contiguous run of column groups so
    // that neighbouring groups, which share cache lines, share an L2.
    const int per = gridDim.x / 8;
    const int grp = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (grp >= p.groups) return;  // whole workgroup, before any barrier
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = wv / W, wl = wv % W;
    const int line = grp * kColW + li;
    const bool live = line < p.lines;
    const int n = p.n;
    const int g = wl * kWave + lane;
    const int e0 = g * CM;
    const int64_t ld = p.ld;
    // staging map: column tc of the group, rows tr + NL r
    const int tc = threadIdx.x % kColW, tr = threadIdx.x / kColW;
    const int tline = grp * kColW + tc;
    const bool tlive = tline < p.lines;
    // coefficient loads first (they overlap the staging loads); the staging
    // thread keeps its C_n values of Cprev for the increment push
    constexpr bool kEarly = CM <= 8;
    double a[CM], f[CM], al[CM], v[CM], cp[CM];
#pragma unroll
    for (int k = 0; k < CM; ++k) a[k] = f[k] = al[k] = v[k] = cp[k] = 0.0;
    if (kEarly && live) load_coefs<CM, W>(p.coef, line, g, a, f, al);
    if (tlive) {
        // all loads first (clamped indices), then the LDS writes
        const int i = tline + 1;
        const double* __restrict__ cc = p.C + i;
        const double* __restrict__ cpv = p.Cprev + (i - 1);
        const bool any = p.nch > 0;
        const double* __restrict__ p0 = p.part + p.ldh + (i - 1);
        double u[CM], ts[CM];
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = min(r * NL + tr, n - 1);
            u[r] = cc[(int64_t)(e + 1) * ld];
            cp[r] = cpv[(int64_t)e * p.NX];  // node (i, j = e+1)
            ts[r] = any ? p0[(int64_t)e * p.NX] : 0.0;
        }
        for (int c = 1; c < p.nch; ++c) {
            const double* __restrict__ pc = p.part + ((int64_t)c * 2 + 1) * p.ldh + (i - 1);
            double x[CM];
#pragma unroll
            for (int r = 0; r < CM; ++r) x[r] = pc[(int64_t)min(r * NL + tr, n - 1) * p.NX];
#pragma unroll
            for (int r = 0; r < CM; ++r) ts[r] += x[r];
        }
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = r * NL + tr;
            if (e < n) {
                const double t = ts[r] + p.kch * (cp[r] - u[r]);
                double om = p.negsk * u[r];
                om += p.S_half;
                om += p.sg * t;
                lds[tc][e + e / CM] = om;
            }
        }
    }
    __syncthreads();
    if (live) {
        if (!kEarly) load_coefs<CM, W>(p.coef, line, g, a, f, al);
#pragma unroll
        for (int k = 0; k < CM; ++k) v[k] = e0 + k < n ? lds[li][g * S + k] : 0.0;
    }
    const double un = line_solve_mw<CM, W>(a, f, al, v, e0, n, p.b0, lane, wl, &share[li], live);
    if (live) {
#pragma unroll
        for (int k = 0; k < CM; ++k)
            if (e0 + k < n) lds[li][g * S + k] = v[k];
        double* Ccol = p.C + (line + 1);
        if (g == 0) {
            Ccol[0] = p.al1 * v[0] + p.b0;
            Ccol[(int64_t)(n + 1) * ld] = un;
            Ccol[(int64_t)(n + 2) * ld] = un;
        }
    }
    __syncthreads();
    if (tlive) {
        const int i = tline + 1;
        double* __restrict__ cc = p.C + i;
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = r * NL + tr;
            if (e < n) {
                const int64_t q = (int64_t)e * p.NX + (i - 1);
                const double u = lds[tc][e + e / CM];
                cc[(int64_t)(e + 1) * ld] = u;
                const double d = u - cp[r];
                if (p.slotf)
                    p.slotf[q] = (float)d;
                else
                    p.slot[q] = d;
            }
        }
    }
}

// ------------------------------------------------------------- g(t) = t^p
// History weights of step n, lag-indexed (w[t*cap + k] multiplies the
// increment of full step s = n-1-k): rows kb(s, 1/2, 0, 1), columns
// kb(s, 1, 0, 1), each by the _kb_row series (frm_gpow.hpp), one thread per
// weight; the reference rebuilds them for every s <= n each step
// (fr_d_t_g_2d.cpp:3059-3239).
__global__ __launch_bounds__(256) void k_gpow_weights(int n, double alpha, double tau, double p,
                                                      double int_err, double Da, int cap,
                                                      double* __restrict__ w,
                                                      int* __restrict__ flag) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * n) return;
    const int t = idx / n, k = idx % n;
    bool capped = false;
    const double v = kb_pow(alpha, tau, p, int_err, n, n - 1 - k, t ? 1.0 : 0.5, 0.0, 1.0, &capped);
    w[(int64_t)t * cap + k] = Da * v;
    if (capped) atomicOr(flag, 1);
}

// Thomas alpha coefficients (a, f, al) of every line for the diagonal
// sgk + A + B, in the reference's order (al1, fr_d_t_g_2d.cpp:3267-3304) --
// line_coefs on the GPU, one thread per line, written in the interleaved
// layout of load_coefs for chunk length cm and nl lanes per line.
__global__ __launch_bounds__(64) void k_line_coefs(const double* __restrict__ A,
                                                   const double* __restrict__ B, int lines, int n,
                                                   double sgk, double al_first, int cm, int nl,
                                                   double* __restrict__ out) {
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= lines) return;
    const double* a = A + (int64_t)L * n;
    const double* b = B + (int64_t)L * n;
    double Al = al_first;
    for (int e = 0; e < n; ++e) {
        const double ae = a[e], be = b[e];
        const double R = sgk + ae + be;
        const double den = R - ae * Al;
        const double Aln = be / den;
        const double f = (Aln == 0.0 && be == 0.0) ? 1.0 / den : Aln / be;
        const int g = e / cm, k = e % cm;
        out[(((int64_t)L * 3 + 0) * cm + k) * nl + g] = ae;
        out[(((int64_t)L * 3 + 1) * cm + k) * nl + g] = f;
        out[(((int64_t)L * 3 + 2) * cm + k) * nl + g] = Aln;
        Al = Aln;
    }
}

// ------------------------------------------------------------- ia/ibk memory term
// The reference's dt_alg 4 (calc_va, fr_d_t_g_2d.cpp:490-741) for every node:
// each direction's plan step (frm_iaibk.cpp) evaluates the node's valid
// approximations at x = n, adds the newest block exactly, then folds the
// newest increment Delta_{n-1} into the coefficient slots.  One lane per
// node; the operations are wavefront-uniform.
struct IaArgs {
    double* H;                 // [2][slots][MAXAK][ld]
    const double* ring;        // two-slot increment ring
    int64_t ld;
    int n, slots;
    const IaOp* ops[2];
    int o0[2], o1[2];          // this step's operation range per direction
    const double* vals[2];
    double ak[kIaMaxAk];
    double scale;
    double* part;              // [2][ld]
};

__global__ __launch_bounds__(256) void k_iaibk_memterm(IaArgs p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.ld) return;
    const int64_t ld = p.ld;
    const double f = p.ring[(int64_t)((p.n - 1) % 2) * ld + i];  // the two-slot ring
#pragma unroll 1
    for (int t = 0; t < 2; ++t) {
        double* H = p.H + (int64_t)t * p.slots * kIaMaxAk * ld + i;
        const IaOp* ops = p.ops[t];
        const double* vals = p.vals[t];
        double V = 0.0;
        for (int o = p.o0[t]; o < p.o1[t]; ++o) {
            const IaOp op = ops[o];
            double* c = H + (int64_t)op.slot * kIaMaxAk * ld;
            const double* w = vals + op.off;
            if (op.kind == kIaEvalIa || op.kind == kIaEvalIbk) {
                double r = 0.0;
                for (int k = 0; k < op.nt; ++k) r += c[(int64_t)k * ld] * w[k];
                V += r;
            } else if (op.kind == kIaExact) {
                V += w[0] * f;
            } else if (op.kind == kIaCreate) {
                for (int k = 0; k < kIaMaxAk; ++k) c[(int64_t)k * ld] = k < op.nt ? p.ak[k] * (f * w[k]) : 0.0;
            } else if (op.kind == kIaAdd) {
                for (int k = 0; k < op.nt; ++k) c[(int64_t)k * ld] += p.ak[k] * (f * w[k]);
            } else if (op.kind == kIaCopy) {
                const double* src = H + (int64_t)op.src * kIaMaxAk * ld;
                for (int k = 0; k < kIaMaxAk; ++k) c[(int64_t)k * ld] = src[(int64_t)k * ld];
            }
        }
        p.part[(int64_t)t * ld + i] = V * p.scale;
    }
}

// ------------------------------------------------------------- SOE memory term
// History weights for lags k >= 2 replaced by sum_l c_{t,l} q_l^(k-2): per
// node, the mode states H_l = sum_{s <= n-2} q_l^(n-2-s) Delta_s follow
// H_l <- q_l H_l + Delta_{n-2} (the increment that just reached lag 2), and
// lag 1 is summed exactly from a two-slot ring of increments.  The decay
// multiply is done in FP64 so that slow modes keep their exact factor when
// the states are stored in FP32.
constexpr int kSoeLag0 = 2;
// largest relative error of a fitted exponential sum over the weights it
// replaces that frm_lod2d_create accepts
constexpr double kSoeFitTol = 1e-6;

template <typename HT, bool NT>
__global__ __launch_bounds__(256) void k_soe_memterm(HT* __restrict__ H, const double* __restrict__ ring,
                                                     int64_t ld, int n, int L,
                                                     const double* __restrict__ q,
                                                     const double* __restrict__ c, double w1r,
                                                     double w1c, double* __restrict__ part) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ld) return;
    const bool fold = n >= kSoeLag0;
    const double x = fold ? ring[(int64_t)((n - kSoeLag0) % kSoeLag0) * ld + i] : 0.0;
    double s0 = 0.0, s1 = 0.0;
    for (int l = 0; l < L; ++l) {
        // the mode planes are streamed once per step (non-temporal when
        // they exceed the cache)
        double h = (double)stream_load<NT>(H + (int64_t)l * ld + i);
        if (fold) {
            h = q[l] * h + x;
            stream_store<NT>((HT)h, H + (int64_t)l * ld + i);
            h = (double)(HT)h;
        }
        s0 += c[l] * h;
        s1 += c[L + l] * h;
    }
    if (n >= 1) {
        const double y = ring[(int64_t)((n - 1) % kSoeLag0) * ld + i];
        s0 += w1r * y;
        s1 += w1c * y;
    }
    part[i] = s0;
    part[ld + i] = s1;
}

// ------------------------------------------------------------- binomial-series memory term
// TG_solver's dt_alg = 3 (fr_d_t_g_2d.cpp:3385-3428): per node and sweep
// direction, moments H_k = sum over full steps s of g(t_{s+1})^(k+1) v3_k(s)
// Delta_s (Htmp1 for rows, Htmp2 for columns), each step taking the newest
// increment Delta_{n-1} = C_n - C_{n-1} from the two-slot ring into the
// first cnt moments, and the memory term Da sum_k f_k H_k with the binomial
// factors f_k of the direction's target (_kb_row2_fixed_coefs).  Products
// and sums are rounded separately, as the reference's x86-64 build does.
__global__ __launch_bounds__(256) void k_binomial_memterm(
    double* __restrict__ H, const double* __restrict__ ring, int64_t ld, int n, int terms,
    const double* __restrict__ upd, const double* __restrict__ er, const double* __restrict__ ec,
    const int* __restrict__ cnt, double Da, double* __restrict__ part) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ld) return;
    if (n == 0) {  // no full step yet: the row term is 0, the column one only its half step
        part[i] = 0.0;
        part[ld + i] = 0.0;
        return;
    }
    const int64_t r = (int64_t)(n - 1) * terms;
    const int* c = cnt + 4 * (int64_t)(n - 1);
    const double x = ring[(int64_t)((n - 1) % kSoeLag0) * ld + i];
    constexpr int B = 8;  // moments loaded together
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
        double* __restrict__ Hd = H + (int64_t)dir * terms * ld + i;
        const double* __restrict__ f = (dir == 0 ? er : ec) + r;
        const int cu = c[dir], ce = c[2 + dir];
        double acc = 0.0;
        for (int k0 = 0; k0 < terms; k0 += B) {
            double h[B];
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (k0 + u < terms) h[u] = Hd[(int64_t)(k0 + u) * ld];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int k = k0 + u;
                if (k >= terms) break;
                if (k < cu) {
                    h[u] = h[u] + upd[r + k] * x;
                    Hd[(int64_t)k * ld] = h[u];
                }
                if (k < ce) acc += f[k] * h[u];
            }
        }
        part[(int64_t)dir * ld + i] = Da * acc;
    }
}

// ------------------------------------------------------------- host helpers
// Thomas alpha coefficients of one line in the reference's order
// (al1: fr_d_t_g_2d.cpp:3267-3304, R: 3257-3265); f = Al[i+1]/B(i), or the
// zero-pivot form 1/(R - A Al[i]) (3519-3522).
static void line_coefs(int n, const double* A, const double* B, int64_t stride, double sgk,
                       double al_first, double* oa, double* of, double* oal) {
    double Al = al_first;
    for (int e = 0; e < n; ++e) {
        const double a = A[e * stride], b = B[e * stride];
        const double R = sgk + a + b;
        const double Aln = b / (R - a * Al);
        oa[e] = a;
        if (Aln == 0.0 && b == 0.0)
            of[e] = 1.0 / (R - a * Al);
        else
            of[e] = Aln / b;
        oal[e] = Aln;
        Al = Aln;
    }
}

// nearest BF16 value (round to nearest even) of x, as a float
static float bf16_round(double x) {
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

// [3][lines][n] (a, f, al per line) -> the interleaved layout of load_coefs
static std::vector<double> interleave_coefs(const std::vector<double>& plain, int lines, int n,
                                            int cm, int nl) {
    std::vector<double> out((size_t)lines * 3 * cm * nl, 0.0);
    for (int L = 0; L < lines; ++L)
        for (int c = 0; c < 3; ++c)
            for (int e = 0; e < n; ++e) {
                const int g = e / cm, k = e % cm;
                out[(((size_t)L * 3 + c) * cm + k) * nl + g] = plain[((size_t)c * lines + L) * n + e];
            }
    return out;
}

// smallest chunk length that fits n unknowns on 64 W lanes (0: none)
static int pick_cm(int n, int W) {
    for (int cm : {4, 8, 16, 32})
        if (n <= kWave * W * cm) return cm;
    return 0;
}


#define FRM_SWEEP_CASES(X) X(4, 1) X(8, 1) X(16, 1) X(32, 1) X(4, 2) X(8, 2) X(16, 2) X(32, 2) \
    X(4, 4) X(8, 4) X(16, 4) X(32, 4)

static hipError_t launch_rows(SweepShape sh, hipStream_t st, const RowArgs& a) {
    const int lpb = 4 / sh.w;
    const dim3 grid((unsigned)((a.lines + lpb - 1) / lpb));
#define X(CM, W)                                                                     \
    if (sh.cm == CM && sh.w == W) {                                                  \
        hipLaunchKernelGGL((k_lod_rows<CM, W>), grid, dim3(256), 0, st, a);          \
        return hipGetLastError();                                                    \
    }
    FRM_SWEEP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

static int col_groups(SweepShape sh, int lines) {
    int cw = 0;
#define X(CM, W) \
    if (sh.cm == CM && sh.w == W) cw = col_width<CM, W>();
    FRM_SWEEP_CASES(X)
#undef X
    return cw ? (lines + cw - 1) / cw : 0;
}

static hipError_t launch_cols(SweepShape sh, hipStream_t st, const ColArgs& a) {
    const dim3 grid((unsigned)((a.groups + 7) / 8 * 8));
#define X(CM, W)                                                                                 \
    if (sh.cm == CM && sh.w == W) {                                                              \
        hipLaunchKernelGGL((k_lod_cols<CM, W>), grid, dim3(col_width<CM, W>() * kWave * W), 0, \
                           st, a);                                                               \
        return hipGetLastError();                                                                \
    }
    FRM_SWEEP_CASES(X)
#undef X
    return hipErrorInvalidValue;
}

// Kernel arguments of one step's sweeps (n = step index, for the push slot).
static RowArgs row_args(frm_lod2d* s, int nch) {
    RowArgs a;
    a.C = s->C;
    a.Cprev = s->Cprev;
    a.coef = s->rcoef;
    a.part = s->hist->part;
    a.ldh = s->hist->ld;
    a.nch = nch;
    a.ld = s->ld;
    a.lines = s->M - 1;
    a.n = s->NX;
    a.NX = s->NX;
    a.negsk = -s->sg * s->kr0;
    a.S_half = s->S * 0.5;
    a.sg = s->sg;
    a.b0 = s->C1;
    a.al1 = 0.0;
    return a;
}

static ColArgs col_args(frm_lod2d* s, int nch, int n) {
    const bool soe = s->precision == FRM_PREC_SOE;
    const bool f32 = s->precision == FRM_PREC_MFMA_BF16 || s->precision == FRM_PREC_F32C;
    ColArgs a;
    a.C = s->C;
    a.Cprev = s->Cprev;
    const bool ring = soe || s->precision == FRM_PREC_BINOMIAL || s->precision == FRM_PREC_IAIBK;
    a.slot = (double*)s->hist->data +
             (int64_t)(ring ? (n % kSoeLag0) : s->hist->depth) * s->hist->ld;
    a.slotf = f32 ? (float*)s->hist->data + (int64_t)s->hist->depth * s->hist->ld : nullptr;
    a.coef = s->ccoef;
    a.part = s->hist->part;
    a.ldh = s->hist->ld;
    a.nch = nch;
    a.ld = s->ld;
    a.lines = s->sN - 1;
    a.n = s->NY;
    a.NX = s->NX;
    a.negsk = -s->sg * s->kc0;
    a.S_half = s->S * 0.5;
    a.sg = s->sg;
    a.kch = s->kch;
    a.b0 = 0.0;
    a.al1 = 1.0;
    a.groups = col_groups(s->csh, a.lines);
    return a;
}

static hipError_t upload_coefs(frm_lod2d* s, bool rows, SweepShape sh) {
    const int lines = rows ? s->M - 1 : s->sN - 1, n = rows ? s->NX : s->NY;
    const std::vector<double> lay =
        interleave_coefs(rows ? s->rplain : s->cplain, lines, n, sh.cm, kWave * sh.w);
    return hipMemcpy(rows ? s->rcoef : s->ccoef, lay.data(), lay.size() * sizeof(double),
                     hipMemcpyHostToDevice);
}

// Autotuner of the sweep launch shape: for every wavefronts-per-line count
// W in {1, 2, 4} that fits the line (or only `force`), lay the coefficients
// out for it and time the row and the column sweep on the solver's own
// buffers (zero state, one chunk of memory-term partials as every step
// after the first has: a step's full memory traffic and arithmetic), then
// keep the fastest shape of each.  Runs at creation, before the initial
// state is set; the state and the history slot it writes are overwritten
// by frm_lod2d_set_state and the first step.
static int tune_sweeps(frm_lod2d* s, int force) {
    frm_ctx* ctx = s->ctx;
    // the one chunk of partials the timed sweeps read: zeros
    FRM_HIP(ctx, hipMemsetAsync(s->hist->part, 0, sizeof(double) * 2 * s->hist->ld, ctx->stream));
    hipEvent_t e0, e1;
    FRM_HIP(ctx, hipEventCreate(&e0));
    FRM_HIP(ctx, hipEventCreate(&e1));
    constexpr int kWarm = 3, kReps = 20;
    int rc = FRM_OK;
    for (int pass = 0; pass < 2 && rc == FRM_OK; ++pass) {
        const bool rows = pass == 0;
        const int n = rows ? s->NX : s->NY;
        float best = 1e30f;
        SweepShape pick;
        for (int W : {1, 2, 4}) {
            if ((force && W != force) || !pick_cm(n, W)) continue;
            const SweepShape sh{pick_cm(n, W), W};
            hipError_t e = upload_coefs(s, rows, sh);
            if (rows) s->rsh = sh;
            else s->csh = sh;
            float ms = 0.f;
            for (int it = 0; it < kWarm + kReps && e == hipSuccess; ++it) {
                if (it == kWarm) e = hipEventRecord(e0, ctx->stream);
                if (e != hipSuccess) break;
                e = rows ? launch_rows(sh, ctx->stream, row_args(s, 1))
                         : launch_cols(sh, ctx->stream, col_args(s, 1, 0));
            }
            if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                rc = check_hip(ctx, e, "sweep autotuning");
                break;
            }
            if (ms < best) {
                best = ms;
                pick = sh;
            }
        }
        if (rc == FRM_OK) {
            if (rows) s->rsh = pick;
            else s->csh = pick;
            const hipError_t e = upload_coefs(s, rows, pick);
            if (e != hipSuccess) rc = check_hip(ctx, e, "sweep coefficients");
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// g(t) = t^p: the current-step weights of step n on the host -- rows
// kb(n, 1/2, 0, 1/2), columns kb(n, 1, 1/2, 1/2), the column half-step term
// kb(n, 1, 0, 1/2) -- as TG_solver::kb computes them (int_alg 1)
static void gpow_current(const frm_lod2d* s, int n, double* kr, double* kc, double* kh,
                         bool* capped) {
    const double a = s->alpha, t = s->tau, p = s->g_pow, ie = s->int_err;
    *kr = s->Da * kb_pow(a, t, p, ie, n, n, 0.5, 0.0, 0.5, capped);
    *kc = s->Da * kb_pow(a, t, p, ie, n, n, 1.0, 0.5, 0.5, capped);
    *kh = s->Da * kb_pow(a, t, p, ie, n, n, 1.0, 0.0, 0.5, capped);
}

// g(t) = t^p: the history weights of step n (GPU) and both directions'
// Thomas coefficients for its current weights
static int gpow_prepare(frm_lod2d* s, int n) {
    frm_ctx* ctx = s->ctx;
    bool capped = false;
    gpow_current(s, n, &s->kr0, &s->kc0, &s->kch, &capped);
    if (capped)
        return set_error(ctx, FRM_E_STATE, "frm_lod2d_step: kb series did not reach int_err at step %d", n);
    if (n > 0) {
        LaunchTimer t(ctx, FRM_K_OTHER, 0.0);
        hipLaunchKernelGGL(k_gpow_weights, dim3((unsigned)((2 * n + 255) / 256)), dim3(256), 0,
                           ctx->stream, n, s->alpha, s->tau, s->g_pow, s->int_err, s->Da,
                           s->capacity, s->w, s->flag);
        FRM_HIP(ctx, hipGetLastError());
    }
    const int64_t nr = (int64_t)(s->M - 1) * s->NX, nc = (int64_t)(s->sN - 1) * s->NY;
    LaunchTimer t(ctx, FRM_K_OTHER, 0.0);
    hipLaunchKernelGGL(k_line_coefs, dim3((unsigned)((s->M - 1 + 63) / 64)), dim3(64), 0,
                       ctx->stream, s->lineAB, s->lineAB + nr, s->M - 1, s->NX, s->sg * s->kr0,
                       0.0, s->rsh.cm, kWave * s->rsh.w, s->rcoef);
    FRM_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_line_coefs, dim3((unsigned)((s->sN - 1 + 63) / 64)), dim3(64), 0,
                       ctx->stream, s->lineAB + 2 * nr, s->lineAB + 2 * nr + nc, s->sN - 1, s->NY,
                       s->sg * s->kc0, 1.0, s->csh.cm, kWave * s->csh.w, s->ccoef);
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}

static void lod_free(frm_lod2d* s) {
    if (!s) return;
    if (s->hist) frm_hist_destroy(s->hist);
    (void)hipFree(s->C);
    (void)hipFree(s->Cprev);
    (void)hipFree(s->w);
    (void)hipFree(s->w32);
    (void)hipFree(s->rcoef);
    (void)hipFree(s->ccoef);
    (void)hipFree(s->soe_H);
    (void)hipFree(s->soe_q);
    (void)hipFree(s->soe_c);
    (void)hipFree(s->bin_H);
    (void)hipFree(s->bin_upd);
    (void)hipFree(s->bin_er);
    (void)hipFree(s->bin_ec);
    (void)hipFree(s->bin_cnt);
    (void)hipFree(s->atab);
    (void)hipFree(s->F);
    (void)hipFree(s->G);
    (void)hipFree(s->lineAB);
    (void)hipFree(s->flag);
    (void)hipFree(s->ia_H);
    for (int t = 0; t < 2; ++t) {
        (void)hipFree(s->ia_ops[t]);
        (void)hipFree(s->ia_vals[t]);
    }
    delete s;
}

}  // namespace frm

using namespace frm;

extern "C" {

int frm_lod2d_create(frm_ctx* ctx, const frm_lod2d_desc* d, frm_lod2d** out) {
    FRM_REQUIRE(ctx, ctx && d && out, "frm_lod2d_create: null argument");
    *out = nullptr;
    FRM_REQUIRE(ctx, d->sN >= 2 && d->M >= 2, "frm_lod2d_create: grid %dx%d too small", d->sN,
                d->M);
    FRM_REQUIRE(ctx, d->capacity >= 1, "frm_lod2d_create: capacity %d", d->capacity);
    FRM_REQUIRE(ctx, d->alpha > 0 && d->alpha < 1 && d->tau > 0,
                "frm_lod2d_create: need 0 < alpha < 1 and tau > 0");
    FRM_REQUIRE(ctx, d->rowA && d->rowB && d->colA && d->colB,
                "frm_lod2d_create: coefficient arrays missing");
    FRM_REQUIRE(ctx,
                d->precision == FRM_PREC_F64 || d->precision == FRM_PREC_F32C ||
                    d->precision == FRM_PREC_SOE || d->precision == FRM_PREC_MFMA_BF16 ||
                    d->precision == FRM_PREC_BINOMIAL || d->precision == FRM_PREC_IAIBK,
                "frm_lod2d_create: precision %d not built", d->precision);
    FRM_REQUIRE(ctx,
                d->trunc_eps == 0 ||
                    (d->trunc_eps > 0 && d->precision == FRM_PREC_F64 && d->kernel == FRM_KERNEL_CAPUTO),
                "frm_lod2d_create: trunc_eps takes FRM_PREC_F64 and the Caputo kernel");
    FRM_REQUIRE(ctx,
                d->precision != FRM_PREC_BINOMIAL ||
                    (d->kernel == FRM_KERNEL_CAPUTO && d->series_terms >= 0 &&
                     d->series_terms <= 1000 && d->series_tol >= 0),
                "frm_lod2d_create: FRM_PREC_BINOMIAL takes the Caputo kernel and 0..1000 terms");
    FRM_REQUIRE(ctx, d->soe_terms >= 0 && d->soe_terms <= 64, "frm_lod2d_create: soe_terms %d",
                d->soe_terms);
    FRM_REQUIRE(ctx, d->mfma_parts == 0 || d->mfma_parts == 2 || d->mfma_parts == 3,
                "frm_lod2d_create: mfma_parts %d (2 or 3)", d->mfma_parts);
    FRM_REQUIRE(ctx, d->kernel == FRM_KERNEL_CAPUTO || d->kernel == FRM_KERNEL_AB,
                "frm_lod2d_create: kernel %d", d->kernel);
    FRM_REQUIRE(ctx, d->g_func == 0 || d->g_func == 3, "frm_lod2d_create: g_func %d (0 or 3)",
                d->g_func);
    FRM_REQUIRE(ctx,
                d->precision != FRM_PREC_IAIBK ||
                    (d->kernel == FRM_KERNEL_CAPUTO && d->g_func == 0 && d->trunc_eps == 0),
                "frm_lod2d_create: FRM_PREC_IAIBK takes the Caputo kernel with g(t) = t");
    FRM_REQUIRE(ctx,
                d->g_func == 0 ||
                    (d->precision == FRM_PREC_F64 && d->kernel == FRM_KERNEL_CAPUTO &&
                     d->trunc_eps == 0 && d->g_pow > 0 && d->g_pow <= 4 && d->int_err >= 0),
                "frm_lod2d_create: g_func 3 takes FRM_PREC_F64, the Caputo kernel, the full "
                "history, 0 < g_pow <= 4 and int_err >= 0");
    FRM_REQUIRE(ctx, pick_cm(d->sN - 1, 4) && pick_cm(d->M - 1, 4),
                "frm_lod2d_create: lines longer than %d unknowns", 4 * 64 * 32);
    FRM_REQUIRE(ctx, d->sweep_w == 0 || d->sweep_w == 1 || d->sweep_w == 2 || d->sweep_w == 4,
                "frm_lod2d_create: sweep_w %d (0, 1, 2 or 4)", d->sweep_w);
    FRM_REQUIRE(ctx, d->sweep_w == 0 || (pick_cm(d->sN - 1, d->sweep_w) && pick_cm(d->M - 1, d->sweep_w)),
                "frm_lod2d_create: lines too long for sweep_w %d", d->sweep_w);
    FRM_HIP(ctx, hipSetDevice(ctx->device));

    frm_lod2d* s = new frm_lod2d();
    s->ctx = ctx;
    s->sN = d->sN;
    s->M = d->M;
    s->ld = d->sN + 2;
    s->NX = d->sN - 1;
    s->NY = d->M - 1;
    s->capacity = d->capacity;
    s->alpha = d->alpha;
    s->tau = d->tau;
    s->sigma = d->sigma;
    s->C1 = d->C1;
    s->S = d->S;
    s->sg = d->sigma / d->tau;
    s->g_func = d->g_func;
    s->g_pow = d->g_func == 3 ? d->g_pow : 1.0;
    s->int_err = d->int_err > 0 ? d->int_err : 1e-6;
    s->Da = 1.0 / frm::gamma_cody(1.0 - d->alpha);
    // Caputo: the kernel's integrals over the LOD sub-intervals, as
    // TG_solver::kb (rows target t_{n+1/2}, columns t_{n+1}).
    // Atangana-Baleanu: the kernel is bounded, so a half-interval current
    // weight is smaller than the first history weight and the constant mode
    // of the zero-gradient columns would grow; both half-steps therefore take
    // the full-step weights of MF_intF_simple, I(d) = MF_int_choice(d tau,
    // tau, 0)/(1-alpha), current I(1), with half of the memory term each, so
    // that the two half-steps add up to the full-step AB discretisation.
    frm::AbIntegrator ab(d->alpha, d->ab_maxl > 0 ? d->ab_maxl : 50,
                         d->ab_maxkiter > 0 ? d->ab_maxkiter : 1000,
                         d->ab_eps > 0 ? d->ab_eps : 1e-10);
    const bool is_ab = d->kernel == FRM_KERNEL_AB;
    std::vector<double> abx;
    if (is_ab && d->ab_exact) {
        abx.resize(d->capacity + 2);
        frm::ab_exact_weights(d->alpha, d->tau, d->capacity + 2, abx.data());
    }
    auto abI = [&](int dd) {
        if (!abx.empty()) return abx[dd - 1] / (1.0 - d->alpha);
        return ab.integral(dd * d->tau, d->tau, 0.0) / (1.0 - d->alpha);
    };
    // current-step weights: kb(n,1/2,0,1/2), kb(n,1,1/2,1/2), kb(n,1,0,1/2)
    s->kr0 = is_ab ? abI(1) : caputo_kb(d->alpha, d->tau, 0.5, 0.5);
    s->kc0 = is_ab ? abI(1) : caputo_kb(d->alpha, d->tau, 0.5, 0.5);
    s->kch = is_ab ? 0.0 : caputo_kb(d->alpha, d->tau, 1.0, 0.5);
    if (s->g_func == 3) {
        bool capped = false;
        gpow_current(s, 0, &s->kr0, &s->kc0, &s->kch, &capped);
        if (capped) {
            lod_free(s);
            return set_error(ctx, FRM_E_INVALID, "frm_lod2d_create: kb series did not reach int_err");
        }
    }
    const int64_t nodes = (int64_t)s->NX * s->NY;
    const size_t nn = (size_t)s->ld * (d->M + 2);
    s->precision = d->precision;
    // the SOE and binomial forms keep only a two-slot ring of increments,
    // the blocked MFMA and compensated FP32 forms FP32 increments
    const bool mf = d->precision == FRM_PREC_MFMA_BF16;
    const bool ring = d->precision == FRM_PREC_SOE || d->precision == FRM_PREC_BINOMIAL ||
                      d->precision == FRM_PREC_IAIBK;
    int rc = hist_alloc(ctx, nodes, ring ? kSoeLag0 : d->capacity,
                        mf || d->precision == FRM_PREC_F32C ? d->precision : FRM_PREC_F64,
                        &s->hist);
    if (rc) {
        lod_free(s);
        return rc;
    }
    hipError_t e = hipMalloc(&s->C, nn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->Cprev, nodes * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->w, (size_t)2 * d->capacity * sizeof(double));
    // coefficient buffers sized for the longest padded line of any shape
    auto padded = [](int n) {
        int mx = 0;
        for (int W : {1, 2, 4})
            if (pick_cm(n, W)) mx = std::max(mx, pick_cm(n, W) * kWave * W);
        return mx;
    };
    if (e == hipSuccess)
        e = hipMalloc(&s->rcoef, (size_t)3 * (d->M - 1) * padded(s->NX) * sizeof(double));
    if (e == hipSuccess)
        e = hipMalloc(&s->ccoef, (size_t)3 * (d->sN - 1) * padded(s->NY) * sizeof(double));
    if (e != hipSuccess) {
        lod_free(s);
        return check_hip(ctx, e, "hipMalloc(lod2d)");
    }
    // lag-indexed history weights: rows kb(s,1/2,0,1) at lag k = n-s is
    // d = k+1/2, l = 1; columns kb(s,1,0,1) is d = k+1, l = 1
    std::vector<double> w(2 * (size_t)d->capacity);
    for (int k = 1; k <= d->capacity && s->g_func == 0; ++k) {
        if (is_ab) {
            w[k - 1] = w[d->capacity + k - 1] = 0.5 * abI(k + 1);
        } else {
            w[k - 1] = caputo_kb(d->alpha, d->tau, k + 0.5, 1.0);
            w[d->capacity + k - 1] = caputo_kb(d->alpha, d->tau, k + 1.0, 1.0);
        }
    }
    // restricted summation (dt_alg = 2): the sums run newest first and stop
    // at the first |kb| < trunc_eps; the Caputo weights fall with the lag,
    // so each direction keeps its leading lags and the kernel reads the
    // newest trunc_L = max of the two
    s->trunc_L = d->capacity;
    if (d->trunc_eps > 0) {
        int keep[2] = {0, 0};
        for (int t = 0; t < 2; ++t) {
            double* wt = w.data() + (size_t)t * d->capacity;
            int k = 0;
            while (k < d->capacity && !(std::fabs(wt[k]) < d->trunc_eps)) ++k;
            keep[t] = k;
            for (; k < d->capacity; ++k) wt[k] = 0.0;
        }
        s->trunc_L = std::max(keep[0], keep[1]);
    }
    // line coefficients
    std::vector<double> rc3((size_t)3 * (d->M - 1) * s->NX), cc3((size_t)3 * (d->sN - 1) * s->NY);
    const int64_t lrows = d->M - 1, lcols = d->sN - 1;
    for (int j = 1; j < d->M; ++j) {
        const int64_t base = (int64_t)j * s->ld + 1;
        double* oa = rc3.data() + (int64_t)(j - 1) * s->NX;
        line_coefs(s->NX, d->rowA + base, d->rowB + base, 1, s->sg * s->kr0, 0.0, oa,
                   oa + lrows * s->NX, oa + 2 * lrows * s->NX);
    }
    for (int i = 1; i < d->sN; ++i) {
        const int64_t base = (int64_t)1 * s->ld + i;
        double* oa = cc3.data() + (int64_t)(i - 1) * s->NY;
        line_coefs(s->NY, d->colA + base, d->colB + base, s->ld, s->sg * s->kc0, 1.0, oa,
                   oa + lcols * s->NY, oa + 2 * lcols * s->NY);
    }
    if (s->g_func == 3) {
        // the lines' off-diagonals for the per-step coefficient kernel
        const int64_t nr = lrows * s->NX, nc = lcols * s->NY;
        std::vector<double> ab((size_t)2 * (nr + nc));
        for (int j = 1; j < d->M; ++j)
            for (int i = 1; i < d->sN; ++i) {
                const int64_t q = (int64_t)j * s->ld + i;
                ab[(j - 1) * s->NX + (i - 1)] = d->rowA[q];
                ab[nr + (j - 1) * s->NX + (i - 1)] = d->rowB[q];
                ab[2 * nr + (i - 1) * s->NY + (j - 1)] = d->colA[q];
                ab[2 * nr + nc + (i - 1) * s->NY + (j - 1)] = d->colB[q];
            }
        e = hipMalloc(&s->lineAB, ab.size() * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&s->flag, sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpy(s->lineAB, ab.data(), ab.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(s->flag, 0, sizeof(int));
        if (e != hipSuccess) {
            lod_free(s);
            return check_hip(ctx, e, "hipMalloc(lod2d g = t^p)");
        }
    }
    s->rplain = std::move(rc3);
    s->cplain = std::move(cc3);
    s->hw.assign({s->kr0, s->kc0, s->kch});
    s->hw.insert(s->hw.end(), w.begin(), w.end());
    if (d->precision == FRM_PREC_SOE) {
        s->soe_L = d->soe_terms > 0 ? d->soe_terms : 32;
        s->soe_f32 = d->soe_f32;
        s->soe_w1r = w[0];
        s->soe_w1c = w[d->capacity];
        frm::SoeFit fit = frm::soe_fit({w.data(), w.data() + d->capacity}, d->capacity, kSoeLag0,
                                       s->soe_L, d->tau);
        s->soe_err = fit.max_rel;
        // a fit that misses the weights by more than the chosen forms' error
        // bar (or is not finite) is refused rather than run
        if (!(fit.max_rel <= kSoeFitTol)) {
            lod_free(s);
            return set_error(ctx, FRM_E_INVALID,
                             "frm_lod2d_create: the %d-term exponential sum fits the history "
                             "weights to %.3g relative (limit %.0e)%s",
                             d->soe_terms > 0 ? d->soe_terms : 32, fit.max_rel, kSoeFitTol,
                             is_ab && !d->ab_exact ? "; truncated AB weights: set ab_exact" : "");
        }
        const size_t hsz = (s->soe_f32 ? sizeof(float) : sizeof(double)) * (size_t)s->soe_L * s->hist->ld;
        e = hipMalloc(&s->soe_H, hsz);
        if (e == hipSuccess) e = hipMalloc(&s->soe_q, sizeof(double) * s->soe_L);
        if (e == hipSuccess) e = hipMalloc(&s->soe_c, sizeof(double) * 2 * s->soe_L);
        if (e == hipSuccess) e = hipMemcpy(s->soe_q, fit.q.data(), sizeof(double) * s->soe_L, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(s->soe_c, fit.c.data(), sizeof(double) * 2 * s->soe_L, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(s->soe_H, 0, hsz);
        if (e != hipSuccess) {
            lod_free(s);
            return check_hip(ctx, e, "hipMalloc(lod2d SOE)");
        }
    }
    if (d->precision == FRM_PREC_BINOMIAL) {
        s->bin_terms = d->series_terms > 0 ? d->series_terms : 25;
        s->bin_Da = 1.0 / gamma_cody(1.0 - d->alpha);
        const frm::BinomialTables bt = frm::binomial_tables(
            d->alpha, d->tau, d->series_tol > 0 ? d->series_tol : 1e-5, s->bin_terms, d->capacity);
        const size_t hsz = sizeof(double) * 2 * (size_t)s->bin_terms * s->hist->ld;
        const size_t tsz = std::max<size_t>(1, bt.upd.size());
        e = hipMalloc(&s->bin_H, hsz);
        if (e == hipSuccess) e = hipMalloc(&s->bin_upd, sizeof(double) * tsz);
        if (e == hipSuccess) e = hipMalloc(&s->bin_er, sizeof(double) * tsz);
        if (e == hipSuccess) e = hipMalloc(&s->bin_ec, sizeof(double) * tsz);
        if (e == hipSuccess) e = hipMalloc(&s->bin_cnt, sizeof(int) * std::max<size_t>(1, bt.cnt.size()));
        if (e == hipSuccess && !bt.upd.empty()) {
            e = hipMemcpy(s->bin_upd, bt.upd.data(), sizeof(double) * bt.upd.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(s->bin_er, bt.er.data(), sizeof(double) * bt.er.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(s->bin_ec, bt.ec.data(), sizeof(double) * bt.ec.size(), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(s->bin_cnt, bt.cnt.data(), sizeof(int) * bt.cnt.size(), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess) e = hipMemset(s->bin_H, 0, hsz);
        if (e != hipSuccess) {
            lod_free(s);
            return check_hip(ctx, e, "hipMalloc(lod2d binomial)");
        }
    }
    if (d->precision == FRM_PREC_IAIBK) {
        // both directions' plans for the whole capacity, (a, b) = (3/2, 1/2)
        // for rows and (2, 1) for columns (Om1, fr_d_t_g_2d.cpp:3443-3458)
        const double A = 1.0 - d->alpha;
        frm::IaPlan pl[2] = {frm::iaibk_plan(A, 1.5, 0.5, d->capacity),
                             frm::iaibk_plan(A, 2.0, 1.0, d->capacity)};
        s->ia_slots = std::max(1, std::max(pl[0].nslots, pl[1].nslots));
        for (int k = 0; k < frm::kIaMaxAk; ++k) s->ia_ak[k] = pl[0].ak[k];
        s->ia_scale = (1.0 / gamma_cody(1.0 - d->alpha)) * (std::pow(d->tau, 1.0 - d->alpha) / (1.0 - d->alpha));
        e = hipMalloc(&s->ia_H, sizeof(double) * 2 * (size_t)s->ia_slots * frm::kIaMaxAk * s->hist->ld);
        for (int t = 0; t < 2 && e == hipSuccess; ++t) {
            s->ia_off[t] = pl[t].step_off;
            e = hipMalloc(&s->ia_ops[t], sizeof(frm::IaOp) * std::max<size_t>(1, pl[t].ops.size()));
            if (e == hipSuccess) e = hipMalloc(&s->ia_vals[t], sizeof(double) * std::max<size_t>(1, pl[t].vals.size()));
            if (e == hipSuccess && !pl[t].ops.empty())
                e = hipMemcpy(s->ia_ops[t], pl[t].ops.data(), sizeof(frm::IaOp) * pl[t].ops.size(),
                              hipMemcpyHostToDevice);
            if (e == hipSuccess && !pl[t].vals.empty())
                e = hipMemcpy(s->ia_vals[t], pl[t].vals.data(), sizeof(double) * pl[t].vals.size(),
                              hipMemcpyHostToDevice);
        }
        if (e != hipSuccess) {
            lod_free(s);
            return check_hip(ctx, e, "hipMalloc(lod2d ia/ibk)");
        }
    }
    if (mf) {
        // BF16 parts of the FP64 weights: w = p0 + p1 + p2 + O(2^-24 w)
        // default parts: two for the Caputo weights (whose products are
        // dominated by the few large recent lags), three for the bounded AB
        // kernel, whose slowly decaying weights make the far part a large
        // share of the memory term
        const int nparts = d->mfma_parts ? d->mfma_parts : (is_ab ? 3 : 2);
        s->mfma_nprod = nparts == 2 ? 3 : 6;
        std::vector<float> parts(3 * w.size());
        for (size_t k = 0; k < w.size(); ++k) {
            double r = w[k];
            for (int p = 0; p < 3; ++p) {
                parts[p * w.size() + k] = bf16_round(r);
                r -= (double)parts[p * w.size() + k];
            }
        }
        // d = 32 .. N0 + 32 for the largest block start N0 below the capacity
        s->atab_chunks = std::max(1, (d->capacity - 1) / kBlockSteps * kBlockSteps / kFarChunk) + 1;
        // slices of about one round of far-product workgroups (two per CU:
        // a launch of whole rounds keeps every CU busy to its end), at most
        // one per step
        {
            const int groups = block_far_groups(s->hist->ld);
            const int round = block_far_round(ctx->device);
            if (round <= 0) {
                lod_free(s);
                return set_error(ctx, FRM_E_HIP, "frm_lod2d_create: CU count of device %d",
                                 ctx->device);
            }
            s->far_slices = std::max(1, std::min(kBlockSteps, (groups + round - 1) / round));
        }
        const std::vector<uint16_t> tab = atab_build(parts.data(), d->capacity, s->atab_chunks);
        e = hipMalloc(&s->atab, tab.size() * sizeof(uint16_t));
        // two far buffers: the current block's and the next one's, formed in slices
        if (e == hipSuccess) e = hipMalloc(&s->F, sizeof(float) * 2 * 2 * kBlockSteps * s->hist->ld);
        if (e == hipSuccess) e = hipMalloc(&s->G, sizeof(float) * 2 * kSubSteps * s->hist->ld);
        if (e == hipSuccess)
            e = hipMemcpy(s->atab, tab.data(), tab.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            lod_free(s);
            return check_hip(ctx, e, "hipMalloc(lod2d MFMA)");
        }
    }
    e = hipMemcpy(s->w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && d->precision == FRM_PREC_F32C) {
        std::vector<float> w32(w.begin(), w.end());
        e = hipMalloc(&s->w32, w32.size() * sizeof(float));
        if (e == hipSuccess)
            e = hipMemcpy(s->w32, w32.data(), w32.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemset(s->C, 0, nn * sizeof(double));
    if (e != hipSuccess) {
        lod_free(s);
        return check_hip(ctx, e, "hipMemcpy(lod2d setup)");
    }
    rc = tune_sweeps(s, d->sweep_w);
    if (rc) {
        lod_free(s);
        return rc;
    }
    // leave no trace of the tuning runs: zero state, Cprev and history slots
    {
        const size_t hb = (s->precision == FRM_PREC_MFMA_BF16 || s->precision == FRM_PREC_F32C
                               ? sizeof(float)
                               : sizeof(double)) *
                          (size_t)std::min(s->hist->capacity, kSoeLag0) * s->hist->ld;
        e = hipMemset(s->C, 0, nn * sizeof(double));
        if (e == hipSuccess) e = hipMemset(s->Cprev, 0, nodes * sizeof(double));
        if (e == hipSuccess) e = hipMemset(s->hist->data, 0, hb);
        if (e != hipSuccess) {
            lod_free(s);
            return check_hip(ctx, e, "hipMemset(lod2d after tuning)");
        }
    }
    *out = s;
    return FRM_OK;
}

int frm_lod2d_destroy(frm_lod2d* s) {
    if (!s) return FRM_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    lod_free(s);
    return FRM_OK;
}

int frm_lod2d_set_state(frm_lod2d* s, const double* h_C) {
    if (!s || !h_C) return set_error(s ? s->ctx : nullptr, FRM_E_INVALID, "frm_lod2d_set_state: null");
    frm_ctx* ctx = s->ctx;
    FRM_HIP(ctx, hipMemcpyAsync(s->C, h_C, (size_t)s->ld * (s->M + 2) * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->tstep = 0;
    s->hist->depth = 0;
    if (s->soe_H) {
        const size_t hsz = (s->soe_f32 ? sizeof(float) : sizeof(double)) * (size_t)s->soe_L * s->hist->ld;
        FRM_HIP(ctx, hipMemsetAsync(s->soe_H, 0, hsz, ctx->stream));
        FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (s->bin_H) {
        const size_t hsz = sizeof(double) * 2 * (size_t)s->bin_terms * s->hist->ld;
        FRM_HIP(ctx, hipMemsetAsync(s->bin_H, 0, hsz, ctx->stream));
        FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return FRM_OK;
}

int frm_lod2d_step(frm_lod2d* s, int32_t nsteps) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_step: null solver");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, nsteps >= 0, "frm_lod2d_step: nsteps %d", nsteps);
    const bool soe = s->precision == FRM_PREC_SOE;
    const bool mf = s->precision == FRM_PREC_MFMA_BF16;
    const bool bin = s->precision == FRM_PREC_BINOMIAL;
    const bool iab = s->precision == FRM_PREC_IAIBK;
    if (s->tstep + nsteps > s->capacity)
        return set_error(ctx, FRM_E_RANGE, "frm_lod2d_step: %d steps exceed capacity %d",
                         s->tstep + nsteps, s->capacity);
    for (int it = 0; it < nsteps; ++it) {
        const int n = s->tstep;
        int nch = 0;
        if (s->g_func == 3) {
            const int rc = gpow_prepare(s, n);
            if (rc) return rc;
        }
        if (soe) {
            nch = 1;
            const int64_t ld = s->hist->ld;
            LaunchTimer t(ctx, FRM_K_MEMTERM,
                          (double)ld * (2.0 * s->soe_L * (s->soe_f32 ? 4 : 8) + 2 * 8 + 2 * 8));
            const dim3 grid((unsigned)((ld + 255) / 256));
            const bool nt = (double)ld * s->soe_L * (s->soe_f32 ? 4 : 8) >= kStreamBypassBytes;
            auto launch = [&](auto kern, auto* H) {
                hipLaunchKernelGGL(kern, grid, dim3(256), 0, ctx->stream, H,
                                   (const double*)s->hist->data, ld, n, s->soe_L, s->soe_q,
                                   s->soe_c, s->soe_w1r, s->soe_w1c, s->hist->part);
            };
            if (s->soe_f32) {
                if (nt) launch(k_soe_memterm<float, true>, (float*)s->soe_H);
                else launch(k_soe_memterm<float, false>, (float*)s->soe_H);
            } else {
                if (nt) launch(k_soe_memterm<double, true>, (double*)s->soe_H);
                else launch(k_soe_memterm<double, false>, (double*)s->soe_H);
            }
            FRM_HIP(ctx, hipGetLastError());
        } else if (iab) {
            // nothing to sum before the first increment
            nch = n > 0 ? 1 : 0;
            if (n > 0) {
                IaArgs a;
                a.H = s->ia_H;
                a.ring = (const double*)s->hist->data;
                a.ld = s->hist->ld;
                a.n = n;
                a.slots = s->ia_slots;
                for (int t = 0; t < 2; ++t) {
                    a.ops[t] = s->ia_ops[t];
                    a.vals[t] = s->ia_vals[t];
                    a.o0[t] = s->ia_off[t][n - 1];
                    a.o1[t] = s->ia_off[t][n];
                }
                for (int k = 0; k < kIaMaxAk; ++k) a.ak[k] = s->ia_ak[k];
                a.scale = s->ia_scale;
                a.part = s->hist->part;
                LaunchTimer t(ctx, FRM_K_MEMTERM, 0.0);
                hipLaunchKernelGGL(k_iaibk_memterm, dim3((unsigned)(a.ld / 256)), dim3(256), 0,
                                   ctx->stream, a);
                FRM_HIP(ctx, hipGetLastError());
            }
        } else if (bin) {
            nch = 1;
            const int64_t ld = s->hist->ld;
            LaunchTimer t(ctx, FRM_K_MEMTERM, (double)ld * (2.0 * s->bin_terms * 16 + 8 + 2 * 8));
            hipLaunchKernelGGL(k_binomial_memterm, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0,
                               ctx->stream, s->bin_H, (const double*)s->hist->data, ld, n,
                               s->bin_terms, s->bin_upd, s->bin_er, s->bin_ec, s->bin_cnt, s->bin_Da,
                               s->hist->part);
            FRM_HIP(ctx, hipGetLastError());
        } else if (mf) {
            // block b = n / 64 uses the far part of the slots before P0 =
            // N0 - 64 (F of parity b), formed during block b - 1; this
            // block's steps N0 + q 64/S form, slice q of S, the next block's
            // far part (slots before N0, all settled) into F of parity b + 1
            const int64_t ld = s->hist->ld;
            const int N0 = n / kBlockSteps * kBlockSteps;
            const int P0 = std::max(0, N0 - kBlockSteps);
            const float* h32 = (const float*)s->hist->data;
            const size_t fsz = (size_t)2 * kBlockSteps * ld;
            float* Fcur = s->F + (size_t)((N0 / kBlockSteps) & 1) * fsz;
            float* Fnext = s->F + (size_t)(((N0 / kBlockSteps) + 1) & 1) * fsz;
            const int every = kBlockSteps / s->far_slices;
            const int q = (n - N0) / every;
            if (N0 > 0 && N0 + kBlockSteps < s->capacity && (n - N0) % every == 0) {
                const int groups = block_far_groups(ld);
                const int per = (groups + s->far_slices - 1) / s->far_slices;
                const int wg0 = q * per, nwg = std::min(per, groups - wg0);
                if (nwg > 0) {
                    const double nodes = (double)s->hist->n_nodes * nwg / groups;
                    LaunchTimer t(ctx, FRM_K_BLOCKFAR, nodes * (4.0 * N0 + 4.0 * 2 * kBlockSteps));
                    if (block_far(ctx->stream, h32, ld, N0 + kBlockSteps, N0, s->atab, s->atab_chunks,
                                  s->mfma_nprod, Fnext, wg0, nwg))
                        return check_hip(ctx, hipGetLastError(), "block_far");
                }
            }
            // the slots from P0 up to the sub-block start, every 8 steps, summed in FP64
            const int N1 = n / kSubSteps * kSubSteps;
            if (N1 > P0 && n == N1) {
                LaunchTimer t(ctx, FRM_K_MEMTERM, (double)ld * ((N1 - P0) * 4.0 + 2 * kSubSteps * 4));
                if (block_mid(ctx->stream, h32, ld, P0, N1, s->w, s->capacity, s->G))
                    return check_hip(ctx, hipGetLastError(), "block_mid");
            }
            nch = 1;
            const bool useF = P0 > 0, useG = N1 > P0;
            LaunchTimer t(ctx, FRM_K_MEMTERM,
                          (double)ld * ((n - N1) * 4.0 + (useF ? 2 * 4 : 0) + (useG ? 2 * 4 : 0) +
                                        2 * 8));
            if (block_near(ctx->stream, h32, ld, N0, N1, n, s->w, s->capacity, Fcur, useF, s->G, useG,
                           s->hist->part))
                return check_hip(ctx, hipGetLastError(), "block_near");
        } else {
            int rc = memterm_partials(ctx, s->hist, 2, s->w, s->capacity, &nch,
                                      s->hist->depth - s->trunc_L, s->w32);
            if (rc) return rc;
        }
        {
            LaunchTimer t(ctx, FRM_K_ROWS, 0.0);
            FRM_HIP(ctx, launch_rows(s->rsh, ctx->stream, row_args(s, nch)));
        }
        {
            LaunchTimer t(ctx, FRM_K_COLS, 0.0);
            FRM_HIP(ctx, launch_cols(s->csh, ctx->stream, col_args(s, nch, n)));
        }
        if (!soe && !bin && !iab) s->hist->depth += 1;
        s->tstep = n + 1;
    }
    return FRM_OK;
}

int frm_lod2d_get_state(frm_lod2d* s, double* h_C) {
    if (!s || !h_C) return set_error(s ? s->ctx : nullptr, FRM_E_INVALID, "frm_lod2d_get_state: null");
    frm_ctx* ctx = s->ctx;
    FRM_HIP(ctx, hipMemcpyAsync(h_C, s->C, (size_t)s->ld * (s->M + 2) * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    int flag = 0;
    if (s->flag)
        FRM_HIP(ctx, hipMemcpyAsync(&flag, s->flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (flag)
        return set_error(ctx, FRM_E_STATE,
                         "frm_lod2d_get_state: a history-weight series did not reach int_err");
    return FRM_OK;
}

int frm_lod2d_tstep(const frm_lod2d* s, int32_t* tstep) {
    if (!s || !tstep) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_tstep: null");
    *tstep = s->tstep;
    return FRM_OK;
}

int frm_lod2d_fit_error(const frm_lod2d* s, double* max_rel) {
    if (!s || !max_rel) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_fit_error: null");
    *max_rel = s->soe_err;
    return FRM_OK;
}

int frm_lod2d_weights(const frm_lod2d* s, int32_t n, double* out) {
    if (!s || !out || n < 0 || n > s->capacity)
        return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_weights: bad arguments");
    if (s->g_func == 3) {
        // the weights of step n, computed on the host as the step computes them
        bool capped = false;
        gpow_current(s, n, &out[0], &out[1], &out[2], &capped);
        for (int k = 0; k < n; ++k) {
            out[3 + k] = s->Da * kb_pow(s->alpha, s->tau, s->g_pow, s->int_err, n, n - 1 - k, 0.5,
                                        0.0, 1.0, &capped);
            out[3 + n + k] = s->Da * kb_pow(s->alpha, s->tau, s->g_pow, s->int_err, n, n - 1 - k,
                                            1.0, 0.0, 1.0, &capped);
        }
        return capped ? set_error(nullptr, FRM_E_STATE, "frm_lod2d_weights: series bound") : FRM_OK;
    }
    out[0] = s->hw[0];
    out[1] = s->hw[1];
    out[2] = s->hw[2];
    for (int k = 0; k < n; ++k) {
        out[3 + k] = s->hw[3 + k];
        out[3 + n + k] = s->hw[3 + s->capacity + k];
    }
    return FRM_OK;
}

int frm_lod2d_far_slices(const frm_lod2d* s, int32_t* slices) {
    if (!s || !slices) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_far_slices: null");
    *slices = s->precision == FRM_PREC_MFMA_BF16 ? s->far_slices : 0;
    return FRM_OK;
}

int frm_lod2d_sweep_shape(const frm_lod2d* s, int32_t* out) {
    if (!s || !out) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_sweep_shape: null");
    out[0] = s->rsh.cm;
    out[1] = s->rsh.w;
    out[2] = s->csh.cm;
    out[3] = s->csh.w;
    return FRM_OK;
}

int frm_lod2d_state_ptr(frm_lod2d* s, double** d_C) {
    if (!s || !d_C) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_state_ptr: null");
    *d_C = s->C;
    return FRM_OK;
}

}  // extern "C"
