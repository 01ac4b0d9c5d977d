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
//      lag n-s for g(t) = t), split along time into FP64 partials;
//   2. rows: one wavefront per row -- RHS assembly, Thomas forward and back
//      substitution (chunked over lanes, wavefront-scan seeds), boundary
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

#include "frm_history.hpp"
#include "frm_mfma.hpp"
#include "frm_wavescan.hpp"
#include "frm_soe.hpp"
#include "frm_weights.hpp"

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
    double* rcoef = nullptr;   // rows:    a, f, al interleaved per line (coef_at)
    double* ccoef = nullptr;   // columns: the same
    int cm_rows = 0, cm_cols = 0;
    int precision = FRM_PREC_F64;
    std::vector<double> hw;    // host copy: kr0, kc0, kch, then [2][capacity] weights
    // sum-of-exponentials form
    int soe_L = 0, soe_f32 = 0;
    double soe_err = 0.0;
    void* soe_H = nullptr;     // [L][ld] mode states (FP32 or FP64)
    double* soe_q = nullptr;   // [L]
    double* soe_c = nullptr;   // [2][L]
    double soe_w1r = 0.0, soe_w1c = 0.0;  // exact lag-1 weights
    // blocked MFMA form: FP32 increments in hist, BF16-exact weight parts,
    // far part of the current 64-step block
    void* atab = nullptr;      // weight-operand fragments (atab_build)
    int atab_chunks = 0;
    int mfma_nprod = 6;        // 3 (two parts) or 6 (three parts)
    float* F = nullptr;        // [128][ld]
    double* G = nullptr;       // [16][ld] mid part of the current 8-step sub-block
};

namespace frm {

// ------------------------------------------------------------- weights
// kb(s, ja, sa, l) for g(t) = t depends on d = n + ja - s - sa only:
//   Da tau^(1-a)/(1-a) * (d^(1-a) - (d-l)^(1-a)),   Da = 1/Gamma(1-a)
// (the closed form of _kb_row, fr_d_t_g_2d.cpp:2937-2963).  The difference
// of powers is evaluated as -d^(1-a) expm1((1-a) log1p(-l/d)), which keeps
// full relative precision at large lags where the direct difference cancels.
static double caputo_kb(double alpha, double tau, double d, double l) {
    const double c = tau == 0 ? 0 : std::pow(tau, 1.0 - alpha) / (1.0 - alpha) / std::tgamma(1.0 - alpha);
    if (d <= 0) return 0.0;
    if (l >= d) return c * std::pow(d, 1.0 - alpha);
    return -c * std::pow(d, 1.0 - alpha) * std::expm1((1.0 - alpha) * std::log1p(-l / d));
}

// ------------------------------------------------------------- kernels
// Coefficient layout: for line L and array c (0: a, 1: f, 2: al), element
// e = lane*CM + k of the line (lane's chunk, position k) is stored at
//   ((L*3 + c)*CM + k)*64 + lane,
// so that a wavefront's k-th load of a coefficient is 64 consecutive doubles.
// The state, the partial sums and Cprev are staged through LDS: read in
// element order (consecutive lanes, consecutive nodes), written to LDS at
// e + e/CM (chunk-major with one pad double per chunk, so that the chunk
// reads lane*(CM+1) + k spread over the banks), read back per chunk.
template <int CM>
__device__ inline int64_t coef_at(int line, int c, int k, int lane) {
    return ((int64_t)(line * 3 + c) * CM + k) * kWave + lane;
}

template <int CM>
__device__ inline void load_coefs(const double* coef, int line, int lane, int e0, int n,
                                  double (&a)[CM], double (&f)[CM], double (&al)[CM]) {
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        a[k] = coef[coef_at<CM>(line, 0, k, lane)];
        f[k] = coef[coef_at<CM>(line, 1, k, lane)];
        al[k] = coef[coef_at<CM>(line, 2, k, lane)];
    }
}

struct RowArgs {
    double* C;
    double* Cprev;
    const double* coef;  // interleaved, see coef_at
    const double* part;  // partial sums [nch][2][ldh], target 0
    int64_t ldh;
    int nch;
    int ld, lines, n, NX;
    double negsk, S_half, sg, b0, al1;
};

constexpr int kRowWaves = 4;  // row lines per workgroup

template <int CM>
__global__ __launch_bounds__(256) void k_lod_rows(RowArgs p) {
    constexpr int S = CM + 1;
    __shared__ double lds[kRowWaves][kWave * S];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int line = blockIdx.x * kRowWaves + wv;
    const bool live = line < p.lines;  // wave-uniform; dead waves still reach the barriers
    const int n = p.n;
    const int e0 = lane * CM;
    double* buf = lds[wv];
    double* Crow = p.C + (int64_t)(line + 1) * p.ld;
    const int64_t qbase = (int64_t)line * p.NX;
    if (live) {
        // om in element order: C_n (copied to Cprev) and the history sum.
        // Every load is issued before the first store (clamped, unguarded
        // indices; the pointers do not alias), so the row costs one memory
        // latency instead of one per element group.
        const double* __restrict__ cr = Crow + 1;
        const double* __restrict__ pt = p.part + qbase;
        double* __restrict__ cpv = p.Cprev + qbase;
        // the first partial-sum chunk is loaded with the state (0 + p0 = p0
        // exactly, so the sum order is unchanged); further chunks follow
        const bool any = p.nch > 0;
        double u[CM], ts[CM];
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = min(r * kWave + lane, n - 1);
            u[r] = cr[e];
            ts[r] = any ? pt[e] : 0.0;
        }
        for (int c = 1; c < p.nch; ++c) {
            const double* __restrict__ pc = pt + (int64_t)c * 2 * p.ldh;
            double x[CM];
#pragma unroll
            for (int r = 0; r < CM; ++r) x[r] = pc[min(r * kWave + lane, n - 1)];
#pragma unroll
            for (int r = 0; r < CM; ++r) ts[r] += x[r];
        }
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = r * kWave + lane;
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
    double a[CM], f[CM], al[CM], v[CM];
    double un = 0.0;
    if (live) {
        load_coefs<CM>(p.coef, line, lane, e0, n, a, f, al);
#pragma unroll
        for (int k = 0; k < CM; ++k) v[k] = e0 + k < n ? buf[lane * S + k] : 0.0;
        un = line_solve<CM>(a, f, al, v, e0, n, p.b0, lane);
#pragma unroll
        for (int k = 0; k < CM; ++k)
            if (e0 + k < n) buf[lane * S + k] = v[k];
    }
    __syncthreads();
    if (live) {
        for (int r = 0; r < CM; ++r) {
            const int e = r * kWave + lane;
            if (e < n) Crow[e + 1] = buf[e + e / CM];
        }
        if (lane == 0) Crow[0] = p.al1 * v[0] + p.b0;
        if (lane == 63) {
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
    const double* coef;  // interleaved, see coef_at
    const double* part;  // target 1 partials at part + ldh
    int64_t ldh;
    int nch;
    int ld, lines, n, NX;
    int groups;          // column groups of kColW lines
    double negsk, S_half, sg, kch, b0, al1;
};

// Column lines per workgroup: a row segment of W columns is one contiguous
// W*8-byte piece of the state, Cprev and partial sums.  W = 8 up to 1024
// unknowns per line; 4 above, where a lane's registers hold 32 elements.
template <int CM>
constexpr int col_width() {
    return CM >= 32 ? 4 : 8;
}

template <int CM>
__global__ __launch_bounds__(col_width<CM>() * 64) void k_lod_cols(ColArgs p) {
    constexpr int kColW = col_width<CM>();
    constexpr int S = CM + 1;
    __shared__ double lds[kColW][kWave * S];
    // Workgroups are dealt to the 8 XCDs round-robin (b -> XCD b % 8); the
    // group order below gives each XCD a contiguous run of column groups so
    // that neighbouring groups, which share cache lines, share an L2.
    const int per = gridDim.x / 8;
    const int g = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (g >= p.groups) return;  // whole workgroup, before any barrier
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int line = g * kColW + wv;
    const bool live = line < p.lines;
    const int n = p.n;
    const int e0 = lane * CM;
    const int64_t ld = p.ld;
    // staging map: column tc of the group, rows tr + 64 r
    const int tc = threadIdx.x % kColW, tr = threadIdx.x / kColW;
    const int tline = g * kColW + tc;
    const bool tlive = tline < p.lines;
    constexpr int kRowsPerPass = kColW * 64 / kColW;  // 64
    if (tlive) {
        // all loads first (clamped indices), then the LDS writes
        const int i = tline + 1;
        const double* __restrict__ cc = p.C + i;
        const double* __restrict__ cpv = p.Cprev + (i - 1);
        const bool any = p.nch > 0;
        const double* __restrict__ p0 = p.part + p.ldh + (i - 1);
        double u[CM], cp[CM], ts[CM];
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = min(r * kRowsPerPass + tr, n - 1);
            u[r] = cc[(int64_t)(e + 1) * ld];
            cp[r] = cpv[(int64_t)e * p.NX];  // node (i, j = e+1)
            ts[r] = any ? p0[(int64_t)e * p.NX] : 0.0;
        }
        for (int c = 1; c < p.nch; ++c) {
            const double* __restrict__ pc = p.part + ((int64_t)c * 2 + 1) * p.ldh + (i - 1);
            double x[CM];
#pragma unroll
            for (int r = 0; r < CM; ++r) x[r] = pc[(int64_t)min(r * kRowsPerPass + tr, n - 1) * p.NX];
#pragma unroll
            for (int r = 0; r < CM; ++r) ts[r] += x[r];
        }
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = r * kRowsPerPass + tr;
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
    double a[CM], f[CM], al[CM], v[CM];
    double un = 0.0;
    if (live) {
        double* buf = lds[wv];
        load_coefs<CM>(p.coef, line, lane, e0, n, a, f, al);
#pragma unroll
        for (int k = 0; k < CM; ++k) v[k] = e0 + k < n ? buf[lane * S + k] : 0.0;
        un = line_solve<CM>(a, f, al, v, e0, n, p.b0, lane);
#pragma unroll
        for (int k = 0; k < CM; ++k)
            if (e0 + k < n) buf[lane * S + k] = v[k];
        double* Ccol = p.C + (line + 1);
        if (lane == 0) Ccol[0] = p.al1 * v[0] + p.b0;
        if (lane == 63) {
            Ccol[(int64_t)(n + 1) * ld] = un;
            Ccol[(int64_t)(n + 2) * ld] = un;
        }
    }
    __syncthreads();
    if (tlive) {
        const int i = tline + 1;
        const double* __restrict__ cpv = p.Cprev + (i - 1);
        double cp[CM];
#pragma unroll
        for (int r = 0; r < CM; ++r) cp[r] = cpv[(int64_t)min(r * kRowsPerPass + tr, n - 1) * p.NX];
        double* __restrict__ cc = p.C + i;
#pragma unroll
        for (int r = 0; r < CM; ++r) {
            const int e = r * kRowsPerPass + tr;
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

// ------------------------------------------------------------- SOE memory term
// History weights for lags k >= 2 replaced by sum_l c_{t,l} q_l^(k-2): per
// node, the mode states H_l = sum_{s <= n-2} q_l^(n-2-s) Delta_s follow
// H_l <- q_l H_l + Delta_{n-2} (the increment that just reached lag 2), and
// lag 1 is summed exactly from a two-slot ring of increments.  The decay
// multiply is done in FP64 so that slow modes keep their exact factor when
// the states are stored in FP32.
constexpr int kSoeLag0 = 2;

template <typename HT>
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
        double h = (double)H[(int64_t)l * ld + i];
        if (fold) {
            h = q[l] * h + x;
            H[(int64_t)l * ld + i] = (HT)h;
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

// [3][lines][n] (a, f, al per line) -> the interleaved layout of coef_at
static std::vector<double> interleave_coefs(const std::vector<double>& plain, int lines, int n,
                                            int cm) {
    std::vector<double> out((size_t)lines * 3 * cm * kWave, 0.0);
    for (int L = 0; L < lines; ++L)
        for (int c = 0; c < 3; ++c)
            for (int e = 0; e < n; ++e) {
                const int lane = e / cm, k = e % cm;
                out[(((size_t)L * 3 + c) * cm + k) * kWave + lane] =
                    plain[((size_t)c * lines + L) * n + e];
            }
    return out;
}

static int pick_cm(int n) {
    if (n <= 64 * 4) return 4;
    if (n <= 64 * 8) return 8;
    if (n <= 64 * 16) return 16;
    if (n <= 64 * 32) return 32;
    return 0;
}

template <typename Args>
static hipError_t launch_rows(int cm, dim3 grid, hipStream_t st, const Args& a) {
    switch (cm) {
        case 4: hipLaunchKernelGGL(k_lod_rows<4>, grid, dim3(256), 0, st, a); break;
        case 8: hipLaunchKernelGGL(k_lod_rows<8>, grid, dim3(256), 0, st, a); break;
        case 16: hipLaunchKernelGGL(k_lod_rows<16>, grid, dim3(256), 0, st, a); break;
        case 32: hipLaunchKernelGGL(k_lod_rows<32>, grid, dim3(256), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename Args>
static hipError_t launch_cols(int cm, dim3 grid, hipStream_t st, const Args& a) {
    switch (cm) {
        case 4: hipLaunchKernelGGL(k_lod_cols<4>, grid, dim3(col_width<4>() * 64), 0, st, a); break;
        case 8: hipLaunchKernelGGL(k_lod_cols<8>, grid, dim3(col_width<8>() * 64), 0, st, a); break;
        case 16: hipLaunchKernelGGL(k_lod_cols<16>, grid, dim3(col_width<16>() * 64), 0, st, a); break;
        case 32: hipLaunchKernelGGL(k_lod_cols<32>, grid, dim3(col_width<32>() * 64), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static void lod_free(frm_lod2d* s) {
    if (!s) return;
    if (s->hist) frm_hist_destroy(s->hist);
    hipFree(s->C);
    hipFree(s->Cprev);
    hipFree(s->w);
    hipFree(s->rcoef);
    hipFree(s->ccoef);
    hipFree(s->soe_H);
    hipFree(s->soe_q);
    hipFree(s->soe_c);
    hipFree(s->atab);
    hipFree(s->F);
    hipFree(s->G);
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
                d->precision == FRM_PREC_F64 || d->precision == FRM_PREC_SOE ||
                    d->precision == FRM_PREC_MFMA_BF16,
                "frm_lod2d_create: precision %d not built", d->precision);
    FRM_REQUIRE(ctx, d->soe_terms >= 0 && d->soe_terms <= 64, "frm_lod2d_create: soe_terms %d",
                d->soe_terms);
    FRM_REQUIRE(ctx, d->mfma_parts == 0 || d->mfma_parts == 2 || d->mfma_parts == 3,
                "frm_lod2d_create: mfma_parts %d (2 or 3)", d->mfma_parts);
    FRM_REQUIRE(ctx, d->kernel == FRM_KERNEL_CAPUTO || d->kernel == FRM_KERNEL_AB,
                "frm_lod2d_create: kernel %d", d->kernel);
    const int cmr = pick_cm(d->sN - 1), cmc = pick_cm(d->M - 1);
    FRM_REQUIRE(ctx, cmr && cmc, "frm_lod2d_create: lines longer than %d unknowns", 64 * 32);
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
    s->cm_rows = cmr;
    s->cm_cols = cmc;
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
    const int64_t nodes = (int64_t)s->NX * s->NY;
    const size_t nn = (size_t)s->ld * (d->M + 2);
    s->precision = d->precision;
    // the SOE form keeps only a two-slot ring of increments, the blocked
    // MFMA form FP32 increments
    const bool mf = d->precision == FRM_PREC_MFMA_BF16;
    int rc = hist_alloc(ctx, nodes, d->precision == FRM_PREC_SOE ? kSoeLag0 : d->capacity,
                        mf ? FRM_PREC_MFMA_BF16 : FRM_PREC_F64, &s->hist);
    if (rc) {
        lod_free(s);
        return rc;
    }
    hipError_t e = hipMalloc(&s->C, nn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->Cprev, nodes * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->w, (size_t)2 * d->capacity * sizeof(double));
    if (e == hipSuccess)
        e = hipMalloc(&s->rcoef, (size_t)3 * (d->M - 1) * cmr * kWave * sizeof(double));
    if (e == hipSuccess)
        e = hipMalloc(&s->ccoef, (size_t)3 * (d->sN - 1) * cmc * kWave * sizeof(double));
    if (e != hipSuccess) {
        lod_free(s);
        return check_hip(ctx, e, "hipMalloc(lod2d)");
    }
    // lag-indexed history weights: rows kb(s,1/2,0,1) at lag k = n-s is
    // d = k+1/2, l = 1; columns kb(s,1,0,1) is d = k+1, l = 1
    std::vector<double> w(2 * (size_t)d->capacity);
    for (int k = 1; k <= d->capacity; ++k) {
        if (is_ab) {
            w[k - 1] = w[d->capacity + k - 1] = 0.5 * abI(k + 1);
        } else {
            w[k - 1] = caputo_kb(d->alpha, d->tau, k + 0.5, 1.0);
            w[d->capacity + k - 1] = caputo_kb(d->alpha, d->tau, k + 1.0, 1.0);
        }
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
    rc3 = interleave_coefs(rc3, (int)lrows, s->NX, cmr);
    cc3 = interleave_coefs(cc3, (int)lcols, s->NY, cmc);
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
        s->atab_chunks = std::max(1, (d->capacity - 1) / kBlockSteps * kBlockSteps / kFarChunk);
        const std::vector<uint16_t> tab = atab_build(parts.data(), d->capacity, s->atab_chunks);
        e = hipMalloc(&s->atab, tab.size() * sizeof(uint16_t));
        if (e == hipSuccess) e = hipMalloc(&s->F, sizeof(float) * 2 * kBlockSteps * s->hist->ld);
        if (e == hipSuccess) e = hipMalloc(&s->G, sizeof(double) * 2 * kSubSteps * s->hist->ld);
        if (e == hipSuccess)
            e = hipMemcpy(s->atab, tab.data(), tab.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            lod_free(s);
            return check_hip(ctx, e, "hipMalloc(lod2d MFMA)");
        }
    }
    e = hipMemcpy(s->w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(s->rcoef, rc3.data(), rc3.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(s->ccoef, cc3.data(), cc3.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        lod_free(s);
        return check_hip(ctx, e, "hipMemcpy(lod2d setup)");
    }
    *out = s;
    return FRM_OK;
}

int frm_lod2d_destroy(frm_lod2d* s) {
    if (!s) return FRM_OK;
    hipStreamSynchronize(s->ctx->stream);
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
    return FRM_OK;
}

int frm_lod2d_step(frm_lod2d* s, int32_t nsteps) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_step: null solver");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, nsteps >= 0, "frm_lod2d_step: nsteps %d", nsteps);
    const bool soe = s->precision == FRM_PREC_SOE;
    const bool mf = s->precision == FRM_PREC_MFMA_BF16;
    if (s->tstep + nsteps > s->capacity)
        return set_error(ctx, FRM_E_RANGE, "frm_lod2d_step: %d steps exceed capacity %d",
                         s->tstep + nsteps, s->capacity);
    const double S_half = s->S * 0.5;
    const int wpb = 256 / kWave;
    for (int it = 0; it < nsteps; ++it) {
        const int n = s->tstep;
        int nch = 0;
        if (soe) {
            nch = 1;
            const int64_t ld = s->hist->ld;
            LaunchTimer t(ctx, FRM_K_MEMTERM,
                          (double)ld * (2.0 * s->soe_L * (s->soe_f32 ? 4 : 8) + 2 * 8 + 2 * 8));
            const dim3 grid((unsigned)((ld + 255) / 256));
            if (s->soe_f32)
                hipLaunchKernelGGL(k_soe_memterm<float>, grid, dim3(256), 0, ctx->stream,
                                   (float*)s->soe_H, (const double*)s->hist->data, ld, n, s->soe_L,
                                   s->soe_q, s->soe_c, s->soe_w1r, s->soe_w1c, s->hist->part);
            else
                hipLaunchKernelGGL(k_soe_memterm<double>, grid, dim3(256), 0, ctx->stream,
                                   (double*)s->soe_H, (const double*)s->hist->data, ld, n, s->soe_L,
                                   s->soe_q, s->soe_c, s->soe_w1r, s->soe_w1c, s->hist->part);
            FRM_HIP(ctx, hipGetLastError());
        } else if (mf) {
            const int64_t ld = s->hist->ld;
            const int N0 = n / kBlockSteps * kBlockSteps;
            const float* h32 = (const float*)s->hist->data;
            if (N0 > 0 && n == N0) {
                LaunchTimer t(ctx, FRM_K_BLOCKFAR, 2.0 * 2 * kBlockSteps * N0 * (double)s->hist->n_nodes);
                if (block_far(ctx->stream, h32, ld, N0, s->atab, s->atab_chunks, s->mfma_nprod, s->F))
                    return check_hip(ctx, hipGetLastError(), "block_far");
            }
            // the settled part of the block, every 8 steps, in FP64
            const int N1 = n / kSubSteps * kSubSteps;
            if (N1 > N0 && n == N1) {
                LaunchTimer t(ctx, FRM_K_MEMTERM, (double)ld * ((N1 - N0) * 4.0 + 2 * kSubSteps * 8));
                if (block_mid(ctx->stream, h32, ld, N0, N1, s->w, s->capacity, s->G))
                    return check_hip(ctx, hipGetLastError(), "block_mid");
            }
            nch = 1;
            LaunchTimer t(ctx, FRM_K_MEMTERM,
                          (double)ld * ((n - N1) * 4.0 + (N0 > 0 ? 2 * 4 : 0) + (N1 > N0 ? 2 * 8 : 0) +
                                        2 * 8));
            if (block_near(ctx->stream, h32, ld, N0, N1, n, s->w, s->capacity, s->F, N0 > 0, s->G,
                           N1 > N0, s->hist->part))
                return check_hip(ctx, hipGetLastError(), "block_near");
        } else {
            int rc = memterm_partials(ctx, s->hist, 2, s->w, s->capacity, &nch);
            if (rc) return rc;
        }
        {
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
            a.S_half = S_half;
            a.sg = s->sg;
            a.b0 = s->C1;
            a.al1 = 0.0;
            LaunchTimer t(ctx, FRM_K_ROWS, 0.0);
            FRM_HIP(ctx, launch_rows(s->cm_rows, dim3((a.lines + wpb - 1) / wpb), ctx->stream, a));
        }
        {
            ColArgs a;
            a.C = s->C;
            a.Cprev = s->Cprev;
            a.slot = (double*)s->hist->data +
                     (int64_t)(soe ? (n % kSoeLag0) : s->hist->depth) * s->hist->ld;
            a.slotf = mf ? (float*)s->hist->data + (int64_t)s->hist->depth * s->hist->ld : nullptr;
            a.coef = s->ccoef;
            a.part = s->hist->part;
            a.ldh = s->hist->ld;
            a.nch = nch;
            a.ld = s->ld;
            a.lines = s->sN - 1;
            a.n = s->NY;
            a.NX = s->NX;
            a.negsk = -s->sg * s->kc0;
            a.S_half = S_half;
            a.sg = s->sg;
            a.kch = s->kch;
            a.b0 = 0.0;
            a.al1 = 1.0;
            const int cw = s->cm_cols >= 32 ? col_width<32>() : col_width<16>();
            a.groups = (a.lines + cw - 1) / cw;
            LaunchTimer t(ctx, FRM_K_COLS, 0.0);
            FRM_HIP(ctx, launch_cols(s->cm_cols, dim3((a.groups + 7) / 8 * 8), ctx->stream, a));
        }
        if (!soe) s->hist->depth += 1;
        s->tstep = n + 1;
    }
    return FRM_OK;
}

int frm_lod2d_get_state(frm_lod2d* s, double* h_C) {
    if (!s || !h_C) return set_error(s ? s->ctx : nullptr, FRM_E_INVALID, "frm_lod2d_get_state: null");
    frm_ctx* ctx = s->ctx;
    FRM_HIP(ctx, hipMemcpyAsync(h_C, s->C, (size_t)s->ld * (s->M + 2) * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
    out[0] = s->hw[0];
    out[1] = s->hw[1];
    out[2] = s->hw[2];
    for (int k = 0; k < n; ++k) {
        out[3 + k] = s->hw[3 + k];
        out[3 + n + k] = s->hw[3 + s->capacity + k];
    }
    return FRM_OK;
}

int frm_lod2d_state_ptr(frm_lod2d* s, double** d_C) {
    if (!s || !d_C) return set_error(nullptr, FRM_E_INVALID, "frm_lod2d_state_ptr: null");
    *d_C = s->C;
    return FRM_OK;
}

}  // extern "C"
