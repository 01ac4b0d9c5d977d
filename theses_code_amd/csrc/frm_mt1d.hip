// Batched 1D time-fractional moisture solver: the hot path of
// H_solver::calc_step of mt_minimal.cpp (Caputo L1 weights) and
// mt_psi_caputo.cpp (psi-Caputo weights, psi(t) = t^beta), gamma = 1.
//
// The reference steps one problem at a time on the CPU: coefficient calls
// per node (al1 829-858, bt1 942-980), the history loop of Om1 (906-939 /
// 1315-1342) over heap snapshots, the progonka back-substitution (1021-1022)
// and, for psi, a full rebuild of kb_time with series integrals every step
// (mt_psi_caputo.cpp:1481-1489).  Here a whole batch of problems of uneven
// size advances together, three launches per step:
//   1. (psi only) weights: one lane per (problem, interval), FP64 series;
//   2. memory term: per node, over the stored increments; the problems are
//      padded to 128-node groups so that each wavefront belongs to one
//      problem and reads its weights through the scalar cache;
//   3. sweep: one wavefront per problem -- coefficients from the current
//      iterate (held in LDS), Thomas alpha/beta/back-substitution with
//      wavefront-scan chunk seeds, boundary rules, Picard iterations, and
//      the push of the new increment.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

#include "frm_diag.h"
#include "frm_history.hpp"
#include "frm_wavescan.hpp"
#include "frm_weights.hpp"

namespace frm {

constexpr int kGroup = 128;       // nodes per memory-term wavefront
constexpr int kMtCM = 16;         // largest unknowns per lane in the sweep (N - 1 <= 1024)
constexpr int kMtMaxN = 64 * kMtCM + 1;

// unknowns per lane for a batch whose largest problem has maxN intervals
static int mt_pick_cm(int maxN) {
    if (maxN - 1 <= 64 * 4) return 4;
    if (maxN - 1 <= 64 * 8) return 8;
    return 16;
}

struct MtProbDev {
    int N, form, top, bottom_cond, src;
    int64_t uoff, hoff;
    double dL, Dg, Da, alpha, beta, D, k, k0, b, top_val, Ev, Hbottom;
    // kind AB with the series history: series length, offsets of the
    // problem's per-step coefficient rows and phase choices
    int ser, ser_maxl;
    int64_t ser_off, ser_coff;
};

// The memory term's time-chunk partials added per node, in chunk order (the
// consumers of a chunked launch), for the CM nodes q0 + e0 + k (k < CM,
// e0 + k < n) of a lane: the loads of 32 / CM chunks of every node are
// issued together (clamped indices, so every load is in bounds).
template <int CM>
__device__ __forceinline__ void chunk_sums(const double* __restrict__ part, int64_t ldh, int nch,
                                           int64_t q0, int e0, int n, double (&hs)[CM]) {
    constexpr int B = CM >= 32 ? 1 : 32 / CM;
#pragma unroll
    for (int k = 0; k < CM; ++k) hs[k] = 0.0;
    for (int c0 = 0; c0 < nch; c0 += B) {
        double x[CM][B];
#pragma unroll
        for (int k = 0; k < CM; ++k)
#pragma unroll
            for (int u = 0; u < B; ++u)
                x[k][u] = part[(int64_t)min(c0 + u, nch - 1) * ldh + q0 + min(e0 + k, n - 1)];
#pragma unroll
        for (int k = 0; k < CM; ++k)
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (c0 + u < nch && e0 + k < n) hs[k] += x[k][u];
    }
}

// ---------------------------------------------------------------- physics
// KK(i) and inv_dw_dh(i) of H_solver at state value x (mt_minimal.cpp:481-562
// for cv_form 2 and k_form 0; CONST and KEXP are this build's forms).
__device__ inline double mt_KK(const MtProbDev& P, double x) {
    if (P.form == FRM_FORM_CV2) return 1.0 * P.k;
    if (P.form == FRM_FORM_KEXP) return P.k0 * exp(P.b * x);
    return P.D;
}
__device__ inline double mt_inv(const MtProbDev& P, double x) {
    if (P.form != FRM_FORM_CV2) return 1.0;
    const double Pr = -x;
    double Ch = 0.0;
    if (Pr >= 10) Ch = 0.05;
    if ((Pr < 10) && (Pr >= 0.0)) Ch = 0.5 - Pr * (0.5 - 0.05) / 10.0;
    return Ch != 0.0 ? 1.0 / Ch : 1e20;
}

// ---------------------------------------------------------------- memory term
// part[c*ldh + j] = sum over slots of chunk c of w[prob(j)*ldw + depth-1-s] * Delta_s[j]
// One wavefront = one 128-node group = one problem (gprob), so the weight
// address is wavefront-uniform.  gvalid holds each group's count of real
// nodes: the lanes past it (the padding of a problem's last group) issue no
// loads, so the stream carries only the problems' own nodes (with 128-node
// padding it carried 12% more on the 4096 uneven problems of the psi batch).
typedef double mt_dv2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(kMemtermThreads) void k_mt_memterm_f64(
    const double* __restrict__ hist, int64_t ldh, int depth, const double* __restrict__ w,
    int ldw, const int* __restrict__ gprob, const int* __restrict__ gvalid, int chunk,
    double* __restrict__ part) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = t * 2;
    if (j >= ldh) return;
    const int p = gprob[t >> 6];
    const bool live = p >= 0 && (int)(j & (kGroup - 1)) < gvalid[t >> 6];
    const int c = blockIdx.y;
    const int s0 = c * chunk;
    const int s1 = min(depth, s0 + chunk);
    double a0 = 0.0, a1 = 0.0;
    if (live) {
        const int64_t ldv = ldh / 2;
        const mt_dv2* q = reinterpret_cast<const mt_dv2*>(hist + (int64_t)s0 * ldh + j);
        const double* wp = w + (int64_t)p * ldw + (depth - 1 - s0);
        int s = s0;
        for (; s + 4 <= s1; s += 4) {
            mt_dv2 x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = __builtin_nontemporal_load(q + (int64_t)u * ldv);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a0 = fma(wp[-u], x[u].x, a0);
                a1 = fma(wp[-u], x[u].y, a1);
            }
            q += 4 * ldv;
            wp -= 4;
        }
        for (; s < s1; ++s) {
            const mt_dv2 x = __builtin_nontemporal_load(q);
            a0 = fma(wp[0], x.x, a0);
            a1 = fma(wp[0], x.y, a1);
            q += ldv;
            wp -= 1;
        }
    }
    part[(int64_t)c * ldh + j] = a0;
    part[(int64_t)c * ldh + j + 1] = a1;
}

// FP32 increments, FP32 weights, Kahan-compensated FP32 sums; the chunk's
// compensated sum is returned in FP64.  NPL nodes per lane: 2 (8-byte
// loads, a wavefront is one 128-node group) or 4 (16-byte loads, each half
// wavefront one group, each lane taking its half's problem weights).  Large
// batches take 4 (4096 uneven psi problems: 5.8 -> 6.4 TB/s); small ones
// keep 2, whose twice as many wavefronts fill the GPU with fewer time
// chunks.

__device__ inline void kahan_add(float& s, float& c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

template <int NPL>
__global__ __launch_bounds__(kMemtermThreads) void k_mt_memterm_f32c(
    const float* __restrict__ hist, int64_t ldh, int depth, const float* __restrict__ w, int ldw,
    const int* __restrict__ gprob, const int* __restrict__ gvalid, int chunk,
    double* __restrict__ part) {
    static_assert(NPL == 2 || NPL == 4, "two or four nodes per lane");
    typedef float vt __attribute__((ext_vector_type(NPL)));
    constexpr int SL = NPL == 2 ? 8 : 4;  // slots in flight per lane: 64 bytes
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = t * NPL;
    const bool in = j < ldh;
    const int64_t grp = in ? j / kGroup : 0;
    const int p = in ? gprob[grp] : -1;
    const bool live = in && p >= 0 && (int)(j & (kGroup - 1)) < gvalid[grp];
    // the problems of the wavefront's groups, wavefront-uniform values (lanes
    // past ldh carry p = -1; a dead group's weight pointer is problem 0's,
    // in bounds and never used)
    const int pa = __builtin_amdgcn_readfirstlane(p);
    const int pb = NPL == 4 ? __builtin_amdgcn_readlane(p, 32) : pa;
    const bool upper = NPL == 4 && (threadIdx.x & 63) >= 32;
    const int c = blockIdx.y;
    const int s0 = c * chunk;
    const int s1 = min(depth, s0 + chunk);
    float sm[NPL], cm[NPL];
#pragma unroll
    for (int k = 0; k < NPL; ++k) sm[k] = cm[k] = 0.f;
    if (live) {
        const int64_t ldv = ldh / NPL;
        const vt* q = reinterpret_cast<const vt*>(hist + (int64_t)s0 * ldh + j);
        const float* wa = w + (int64_t)max(pa, 0) * ldw + (depth - 1 - s0);
        const float* wb = w + (int64_t)max(pb, 0) * ldw + (depth - 1 - s0);
        int s = s0;
        for (; s + SL <= s1; s += SL) {
            vt x[SL];
#pragma unroll
            for (int u = 0; u < SL; ++u) x[u] = __builtin_nontemporal_load(q + (int64_t)u * ldv);
#pragma unroll
            for (int u = 0; u < SL; ++u) {
                const float wu = upper ? wb[-u] : wa[-u];
#pragma unroll
                for (int k = 0; k < NPL; ++k) kahan_add(sm[k], cm[k], wu * x[u][k]);
            }
            q += SL * ldv;
            wa -= SL;
            wb -= SL;
        }
        for (; s < s1; ++s) {
            const vt x = __builtin_nontemporal_load(q);
            const float wu = upper ? wb[0] : wa[0];
#pragma unroll
            for (int k = 0; k < NPL; ++k) kahan_add(sm[k], cm[k], wu * x[k]);
            q += ldv;
            wa -= 1;
            wb -= 1;
        }
    }
    if (in) {
#pragma unroll
        for (int k = 0; k < NPL; ++k) part[(int64_t)c * ldh + j + k] = (double)sm[k] - (double)cm[k];
    }
}

// ---------------------------------------------------------------- psi weights
// calc_int(alpha, a, a+1, A, tau, 1) of mt_psi_caputo.cpp:720-740 for
// g(t) = t^beta: Taylor series around g(tj) (_kb_row, 497-566) or, for
// intervals in the first half (t1 <= A/2), Newton binomial in g(tj) times the
// Taylor series of the inverse's derivative around g(t1) (_kb_row2,
// 567-719); both stop at term size global_eps = 1e-5 as the reference does.
// The reference caches the inner integrals of _kb_row2 per interval; they
// depend only on the interval, so recomputing them gives the same values.
constexpr double kPsiEps = 1e-5;
constexpr int kPsiMaxIt = 100000;

__device__ inline double psi_inv_der(double beta, double t, int n, double d) {
    if (n == 0) return pow(t, 1.0 / beta);
    if (n == 1) return ((1.0 / beta) * pow(t, (1.0 / beta) - 1)) * (d / pow(t, 1.0 / beta));
    return d * ((1.0 / beta) - n + 1) / t;
}

// v_first = f'(gtj) gtj^(1-alpha), the series' first factor (the same for
// every interval of a step)
__device__ double psi_kb_row(double beta, double t0, double t1, double gtj, double alpha,
                             double v_first) {
    const double gt1 = pow(t1, beta), gt0 = pow(t0, beta);
    double sum = 0.0, i = 0.0;
    int niter = 0;
    double v = v_first;
    double v3 = pow(fabs(1.0 - gt1 / gtj), 1 - alpha);
    double v4 = pow(fabs(1.0 - gt0 / gtj), 1 - alpha);
    double v2 = -(v3 - v4) / (1 - alpha);
    sum += v2 * v;
    // term i: the reference forms f^(i+1)(gtj) = f^(i)(gtj) (1/beta - i)/gtj
    // and multiplies by -gtj/i; gtj cancels, leaving one division per term
    // (the terms agree with the reference's to rounding)
    const double q1 = fabs(1.0 - gt1 / gtj), q0 = fabs(1.0 - gt0 / gtj), ib = 1.0 / beta;
    while (fabs(v * v2) > kPsiEps) {
        i += 1.0;
        v *= (i - ib) / i;
        v3 *= q1;
        v4 *= q0;
        v2 = -(v3 - v4) / (i - alpha + 1.0);
        sum += v2 * v;
        if (++niter > kPsiMaxIt) break;
        if (!isfinite(sum)) break;
    }
    return sum;
}

// gtj_na = gtj^-alpha (the same for every interval of a step)
__device__ double psi_kb_row2(double beta, double t0, double t1, double gtj, double alpha,
                              double gtj_na) {
    const double gt0 = pow(t0, beta), gt1 = pow(t1, beta);
    double sum = 0.0, v1 = 1.0, v2, v3, v4, v5, v6, v7, i = 0.0, m;
    int niter = 0;
    v2 = gtj_na * gt1;
    v6 = gt0 / gt1;
    // the first derivative of the inverse at g(t1) starts every term's inner
    // series; it depends on the interval only (three pow calls, made once)
    const double v4_first = psi_inv_der(beta, gt1, 1, psi_inv_der(beta, gt1, 0, 0));
    const double ib = 1.0 / beta, rg = (gt0 - gt1) / gt1;
    do {
        if (i != 0.0) {
            v1 *= -(-alpha - i + 1.0) / i;
            v2 *= gt1 / gtj;
            v6 *= gt0 / gt1;
        }
        if (i == 0.0) {
            v3 = (t1 - t0) / gt1;
        } else {
            v3 = 0.0;
            m = 0.0;
            v4 = v4_first;
            v5 = (1.0 / (i + 1.0)) * (1.0 - v6);
            v3 += v4 * v5;
            v7 = 1;
            // inner term m: the reference forms f^(m+1)(gt1) / m / ((i+m+1) /
            // (gt1 m)) and v7 (gt0-gt1) ((m-1)/m) ((i+m)/(gt1 (m-1))); gt1, m
            // and m-1 cancel, leaving one division each
            while (fabs(v4 * v5) > (fabs(v3) * kPsiEps)) {
                m += 1.0;
                v4 *= (ib - m) / (i + m + 1.0);
                if (m == 1.0)
                    v7 = (1.0 / (gt1)) * (v6 * (gt0 - gt1));
                else
                    v7 *= rg * (i + m) / m;
                v5 = -(v5 + v7);
                v3 += v4 * v5;
                if (++niter > kPsiMaxIt) break;
                if (!isfinite(v3)) break;
            }
        }
        sum += v1 * v2 * v3;
        i += 1.0;
        if (++niter > kPsiMaxIt) break;
        if (!isfinite(sum)) break;
    } while (fabs(v1 * v2 * v3) > kPsiEps);
    return sum;
}

// The factors of calc_int that depend only on the target g(A tau) of the
// step, formed once per problem and step: gtj, f'(gtj) gtj^(1-alpha) and
// gtj^-alpha (the same pow calls the per-interval series would make).
struct PsiStep {
    double gtj, v_first, gtj_na;
};

__device__ inline PsiStep psi_step(double alpha, double beta, int A, double tau) {
    PsiStep q;
    q.gtj = pow(A * tau, beta);
    double v = psi_inv_der(beta, q.gtj, 1, psi_inv_der(beta, q.gtj, 0, 1));
    v *= pow(q.gtj, 1 - alpha);
    q.v_first = v;
    q.gtj_na = pow(q.gtj, -alpha);
    return q;
}

__device__ double psi_calc_int(double alpha, double beta, int a, int A, double tau, const PsiStep& q) {
    if ((a + 1 <= 0.5 * A) && !(beta == 1))
        return psi_kb_row2(beta, a * tau, (a + 1) * tau, q.gtj, alpha, q.gtj_na);
    return psi_kb_row(beta, a * tau, (a + 1) * tau, q.gtj, alpha, q.v_first);
}

// The reference's fallback where a series is not finite: _kb(1, t0, t1,
// gtj, sing, alpha) (mt_psi_caputo.cpp:476-494), the integral of
// (gtj - g(t))^-alpha over [t0, t1] by midpoint rules on 1 and 2 panels,
// halving the interval while they differ by more than global_eps; below
// global_eps in width the two rules' mean, or on the interval that ends at
// the target (sing) the adaptive 3/8-rule march of ___kb (447-473).  The
// recursion runs on an explicit stack of frames, left half first, each
// interval's value the sum of its halves' as in the reference.  Returns NaN
// if the subdivision would pass kPsiQuadDepth levels or the march
// kPsiQuadMaxIt steps (the reference has no such bounds; widths below
// global_eps end the subdivision after log2(width / 1e-5) levels).
constexpr int kPsiQuadDepth = 48;
constexpr int kPsiQuadMaxIt = 1 << 20;

__device__ inline double psi_kb_mid(double beta, int n, double t0, double t1, double gtj,
                                    double alpha) {
#pragma clang fp contract(off)
    double ret = 0.0;
    const double step = (t1 - t0) / n;
    for (int i = 0; i < n; i++) ret += step * pow(gtj - pow(t0 + (i + 0.5) * step, beta), -alpha);
    return ret;
}

__device__ double psi_kb_sing(double beta, int n, double t0, double t1, double gtj, double alpha) {
#pragma clang fp contract(off)
    double f = pow(gtj - pow(t0, beta), -alpha), f1, f2, f3, f4;
    double sum = 0.0, v, t = t0, step = (t1 - t0) / n;
    int it = 0;
    do {
        f1 = pow(gtj - pow(t + 0.5 * step, beta), -alpha);
        f3 = pow(gtj - pow(t + step / 3.0, beta), -alpha);
        f4 = pow(gtj - pow(t + 2.0 * step / 3.0, beta), -alpha);
        f2 = pow(gtj - pow(t + step, beta), -alpha);
        v = (f + 3 * f3 + 3 * f4 + f2) * step / 8.0;
        if (!isfinite(v)) break;
        sum += v;
        t += step;
        if ((f1 / f) < (f2 / f1)) step /= (f2 / f1) / (f1 / f);
        // a finite step reaches 0 within 1100 halvings; an infinite one
        // (t1 - t0 overflowing) is cut off here
        for (int hv = 0; (t + step) > t1; ++hv) {
            if (hv > 1100) return NAN;
            step /= 2.0;
        }
        f = f2;
        if (++it > kPsiQuadMaxIt) return NAN;
    } while ((t1 - t) > kPsiEps * kPsiEps * kPsiEps);
    return sum;
}

__device__ double psi_kb_quad(double beta, double t0, double t1, double gtj, bool sing,
                              double alpha) {
#pragma clang fp contract(off)
    struct Frame {
        double t0, t1, left;
        int phase;  // 0: not yet evaluated, 1: left half pending, 2: right half pending
    };
    Frame fr[kPsiQuadDepth];
    int sp = 0;
    fr[0] = {t0, t1, 0.0, 0};
    double ret = 0.0;
    for (;;) {
        Frame& F = fr[sp];
        if (F.phase == 0) {
            const double v1 = psi_kb_mid(beta, 1, F.t0, F.t1, gtj, alpha);
            const double v2 = psi_kb_mid(beta, 2, F.t0, F.t1, gtj, alpha);
            if ((F.t1 - F.t0) < kPsiEps) {
                ret = sing ? psi_kb_sing(beta, 16, F.t0, F.t1, gtj, alpha) : 0.5 * (v1 + v2);
            } else if (fabs(v2 - v1) > kPsiEps) {
                if (sp + 1 >= kPsiQuadDepth) return NAN;
                F.phase = 1;
                fr[sp + 1] = {F.t0, 0.5 * (F.t0 + F.t1), 0.0, 0};
                ++sp;
                continue;
            } else {
                ret = v2;
            }
        } else if (F.phase == 1) {  // the left half's value arrived
            F.left = ret;
            F.phase = 2;
            fr[sp + 1] = {0.5 * (F.t0 + F.t1), F.t1, 0.0, 0};
            ++sp;
            continue;
        } else {  // the right half's value arrived
            ret = F.left + ret;
        }
        if (sp == 0) return ret;
        --sp;
    }
}

// Stores the value of interval x of problem p at step k as a history weight
// (x < depth) or the current-step factor d0.
__device__ inline void psi_store(const MtProbDev& P, int p, int x, int depth, double v,
                                 double* __restrict__ w, int ldw, float* __restrict__ w32,
                                 double* __restrict__ d0) {
    if (x < depth) {
        const double wl = v * (1 - P.alpha);
        const int l = depth - (x + 1);
        w[(int64_t)p * ldw + l] = wl;
        if (w32) w32[(int64_t)p * ldw + l] = (float)wl;
    } else {
        d0[p] = P.Da * v * (1 - P.alpha);
    }
}

// Step k (= tstep): w[p*ldw + l] = (1-alpha) kb_time[depth-l] for l < depth
// (depth = k-3 history terms) and d0[p] = Da (1-alpha) kb_time[k].
// Thread x of row p handles interval a = x+1 (x < depth) or a = k (x = depth).
__global__ void k_psi_weights(const MtProbDev* __restrict__ pr, int nprob, int k, int depth,
                              double tau, double* __restrict__ w, int ldw, float* __restrict__ w32,
                              double* __restrict__ d0, int* __restrict__ bad,
                              int2* __restrict__ queue, int qcap, int* __restrict__ qcount) {
    const int p = blockIdx.y;
    if (p >= nprob) return;  // the whole workgroup: one problem per grid row
    const MtProbDev P = pr[p];
    __shared__ PsiStep q;
    if (threadIdx.x == 0) q = psi_step(P.alpha, P.beta, k + 1, tau);
    __syncthreads();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x > depth) return;
    const int a = (x < depth) ? x + 1 : k;
    const double v = psi_calc_int(P.alpha, P.beta, a, k + 1, tau, q);
    if (!isfinite(v)) {
        // calc_int falls back to _kb here: k_psi_quad computes and stores it
        const int slot = atomicAdd(&qcount[k & 1], 1);
        if (slot < qcap)
            queue[slot] = make_int2(p, x);
        else
            atomicOr(bad, 4);
        return;
    }
    psi_store(P, p, x, depth, v, w, ldw, w32, d0);
}

// One thread per queued (problem, interval) whose series was not finite:
// calc_int's _kb fallback, stored as k_psi_weights stores a weight.  A value
// that is still not finite sets bad bit 1.  Workgroup 0 clears the queue
// length that step k + 1 appends to (step k - 1's k_psi_quad, its last
// reader, has finished).
__global__ __launch_bounds__(64) void k_psi_quad(const MtProbDev* __restrict__ pr, int k,
                                                 int depth, double tau, double* __restrict__ w,
                                                 int ldw, float* __restrict__ w32,
                                                 double* __restrict__ d0, int* __restrict__ bad,
                                                 const int2* __restrict__ queue, int qcap,
                                                 int* __restrict__ qcount) {
    const int nq = min(qcount[k & 1], qcap);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) qcount[(k + 1) & 1] = 0;
    for (int q = t; q < nq; q += gridDim.x * blockDim.x) {
        const int2 e = queue[q];
        const MtProbDev P = pr[e.x];
        const int a = (e.y < depth) ? e.y + 1 : k;
        const double gtj = pow((k + 1) * tau, P.beta);
        const double v = psi_kb_quad(P.beta, a * tau, (a + 1) * tau, gtj, a == k, P.alpha);
        if (!isfinite(v)) atomicOr(bad, 1);
        psi_store(P, e.x, e.y, depth, v, w, ldw, w32, d0);
    }
}

__global__ void k_diag_psi_kb_quad(const double* __restrict__ c, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* r = c + 6 * (int64_t)i;
    out[i] = psi_kb_quad(r[1], r[2], r[3], r[4], r[5] != 0.0, r[0]);
}

// ---------------------------------------------------------------- sweep
struct MtSweepArgs {
    const MtProbDev* pr;
    int nprob;
    double* U;
    const double* part;
    int nch;
    int64_t ldh;
    const double* d0;
    void* slot;  // history slot for the new increment, or null
    int slot_f32;
    int picard;
};

// CM = 16 (problems of up to 1024 nodes) is held to 256 registers, two
// wavefronts per SIMD: one problem per wavefront leaves few wavefronts to
// hide the scans' latency, and at 290 registers (one per SIMD) the 4096
// problems of the psi batch ran in four rounds (0.83 s over 3000 steps
// against 0.61 s with a small spill).
template <int CM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CM == 16 ? 2 : 1, 8))) void k_mt_sweep(MtSweepArgs a) {
    __shared__ double xs_all[4][64 * CM + 3];  // nodes 0..N+1, N - 1 <= 64*CM
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + wv;
    if (p >= a.nprob) return;
    double* xs = xs_all[wv];
    const MtProbDev P = a.pr[p];
    const int N = P.N, n = N - 1, e0 = lane * CM;
    double* Ug = a.U + P.uoff;
    {
        // N + 2 <= 64 CM + 3 nodes: CM + 1 loads per lane, all in flight
        double ux[CM + 1];
#pragma unroll
        for (int r = 0; r <= CM; ++r) ux[r] = Ug[min(lane + 64 * r, N + 1)];
#pragma unroll
        for (int r = 0; r <= CM; ++r)
            if (lane + 64 * r < N + 2) xs[lane + 64 * r] = ux[r];
    }
    double u0[CM], hs[CM];
    chunk_sums<CM>(a.part, a.ldh, a.nch, P.hoff, e0, n, hs);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        const int e = e0 + k;
        u0[k] = 0.0;
        if (e < n) u0[k] = xs[e + 1];
    }
    const double d0 = a.d0[p];
    for (int it = 0; it < a.picard; ++it) {
        double A[CM], B[CM], R[CM], om[CM];
#pragma unroll
        for (int k = 0; k < CM; ++k) {
            const int e = e0 + k;
            if (e < n) {
                const int i = e + 1;
                const double inv = mt_inv(P, xs[i]);
                const double Km = mt_KK(P, xs[i - 1]), K0 = mt_KK(P, xs[i]), Kp = mt_KK(P, xs[i + 1]);
                A[k] = P.Dg * Km * inv;
                B[k] = P.Dg * Kp * inv;
                R[k] = inv * (Kp + Km) * P.Dg + d0;
                double o = -d0 * u0[k];
                o += (K0 - Km) / P.dL;
                o += P.Da * hs[k];
                om[k] = o;
            } else {
                A[k] = 0.0;
                B[k] = 1.0;
                R[k] = 1.0;
                om[k] = 0.0;
            }
        }
        // top boundary: al1 / bt1 (mt_minimal.cpp:831-834, 944-947)
        const double x0 = xs[0];
        int first_order;
        if (P.top == FRM_TOP_DIRICHLET) first_order = 0;
        else if (P.top == FRM_TOP_FLUX) first_order = 1;
        else first_order = (mt_inv(P, x0) != 1e20);
        const double al_first = first_order ? 1.0 : 0.0;
        const double Uc = 1.0 * P.dL * (P.Ev - 0.0) / mt_KK(P, x0);
        double b0;
        if (P.top == FRM_TOP_DIRICHLET) b0 = P.top_val;
        else if (first_order) b0 = -Uc;
        else b0 = -Uc * mt_KK(P, xs[1]) / P.dL;
        // alpha: Al[i+1] = B/(R - A Al[i]), a Moebius map of Al
        Mob m{1.0, 0.0, 0.0, 1.0};
#pragma unroll
        for (int k = 0; k < CM; ++k)
            if (e0 + k < n) m = mob_mul(Mob{0.0, B[k], -A[k], R[k]}, m);
        const Mob pre = scan_mob_up(m, lane);
        double al = (pre.m00 * al_first + pre.m01) / (pre.m10 * al_first + pre.m11);
#pragma unroll
        for (int k = 0; k < CM; ++k) {
            if (e0 + k < n) {
                al = B[k] / (R[k] - A[k] * al);
                R[k] = al;  // R now holds Al[i+1]
            }
        }
        // beta: Bt[i+1] = (Al[i+1]/B)(A Bt[i] - Om)
        double Pp = 1.0, Qq = 0.0;
#pragma unroll
        for (int k = 0; k < CM; ++k) {
            if (e0 + k < n) {
                const double pk = R[k] * A[k] / B[k], qk = -R[k] * om[k] / B[k];
                Qq = pk * Qq + qk;
                Pp = pk * Pp;
            }
        }
        scan_affine_up(Pp, Qq, lane);
        double bt = Pp * b0 + Qq;
#pragma unroll
        for (int k = 0; k < CM; ++k) {
            if (e0 + k < n) {
                bt = (R[k] / B[k]) * (A[k] * bt - om[k]);
                om[k] = bt;  // om now holds Bt[i+1]
            }
        }
        // bottom boundary (mt_minimal.cpp:987-1020), gamma = 1 so Qb = -0
        const int last = (n - 1) / CM;
        double UN_l = 0.0;
#pragma unroll
        for (int k = 0; k < CM; ++k) {
            if (e0 + k == n - 1) {
                const double BtN = om[k], AlN = R[k], Qb = -0.0;
                if (P.bottom_cond == 1)
                    UN_l = P.Hbottom;
                else if (fabs(BtN + Qb) > 1e-15)
                    UN_l = (fabs(1.0 - AlN) > 1e-15) ? (BtN + Qb) / (1.0 - AlN) : xs[N - 1] + Qb;
                else
                    UN_l = (fabs(1.0 - AlN) > 1e-15) ? 0.0 : xs[N - 1] + Qb;
            }
        }
        const double UN = __shfl(UN_l, last);
        // back substitution U_i = Al[i+1] U_{i+1} + Bt[i+1]
        Pp = 1.0;
        Qq = 0.0;
#pragma unroll
        for (int k = CM - 1; k >= 0; --k) {
            if (e0 + k < n) {
                Qq = R[k] * Qq + om[k];
                Pp = R[k] * Pp;
            }
        }
        scan_affine_down(Pp, Qq, lane);
        double y = Pp * UN + Qq;
        double v[CM];
#pragma unroll
        for (int k = CM - 1; k >= 0; --k) {
            v[k] = 0.0;
            if (e0 + k < n) {
                y = R[k] * y + om[k];
                v[k] = y;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int k = 0; k < CM; ++k)
            if (e0 + k < n) xs[e0 + k + 1] = v[k];
        if (lane == 0) {
            xs[0] = al_first * v[0] + b0;
            xs[N] = UN;
            xs[N + 1] = UN;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    for (int q = lane; q < N + 2; q += 64) Ug[q] = xs[q];
    if (a.slot) {
#pragma unroll
        for (int k = 0; k < CM; ++k) {
            const int e = e0 + k;
            if (e < n) {
                const double d = xs[e + 1] - u0[k];
                if (a.slot_f32)
                    static_cast<float*>(a.slot)[P.hoff + e] = (float)d;
                else
                    static_cast<double*>(a.slot)[P.hoff + e] = d;
            }
        }
    }
}

// ---------------------------------------------------------------- AB kind
// The manufactured solutions of MF_int_series.cpp:860-873 (F_func = src);
// E is the Mittag-Leffler factor of the source at time t.
constexpr double kAbPi = 3.14159265358979323846;

__host__ __device__ inline double ab_CC(int F, double x, double t) {
    if (F == 0) return t * t * t * t;
    if (F == 1) return x * (x - 1.0) * t * t;
    if (F == 2) return t * t * t * t * t * sin(kAbPi * x);
    return 0.0;
}

__device__ inline double ab_f(int F, double x, double t, double a, double E) {
    if (F == 0) return (24.0 / (1.0 - a)) * t * t * t * t * E;
    if (F == 1) return (2.0 * x - 1.0) * t * t - 2 * t * t + (2.0 / (1.0 - a)) * x * (x - 1.0) * t * t * E;
    if (F == 2)
        return (120.0 / (1.0 - a)) * t * t * t * t * t * sin(kAbPi * x) * E +
               kAbPi * t * t * t * t * t * (kAbPi * sin(kAbPi * x) + cos(kAbPi * x));
    return 0.0;
}

struct AbSweepArgs {
    const MtProbDev* pr;
    int nprob;
    double* U;
    const double* part;  // history sums sum_k I(lag+1) (C^{k+1}-C^k)/tau, [nch][ldh];
                         // series problems: MF_intF_choice's value in row 0
    int nch;
    int64_t ldh;
    const double* d0;    // I(1)/(1-alpha) per problem
    const double* cf;    // per history node: -1/(B - A Al_{j-1})
    const double* cal;   // per history node: -Al_j
    const double* ml;    // [nprob][ldml] source factor at step k
    int ldml;
    int k;               // step being taken (1-based), t = k tau
    double tau;
    double* slot;        // history slot for (C^k - C^{k-1})/tau
    double* err;         // per problem: running sum of (C_j - CC)^2
};

// One wavefront per problem: right-hand side Rp_j = I(1)/(1-a) C_j + f -
// history/(1-a) (MF_int_series.cpp:922-931), the Thomas sweep with the
// exact Dirichlet ends (935-945) through the scan-seeded line solve (the
// constant alpha recurrence is precomputed), the increment push and the
// squared error against CC (947-951).
template <int CM>
__global__ __launch_bounds__(256) void k_ab1d_sweep(AbSweepArgs a) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + wv;
    if (p >= a.nprob) return;  // whole wavefront; the kernel has no block barrier
    const MtProbDev P = a.pr[p];
    const int m = P.N, n = m - 1, e0 = lane * CM;
    const double h = P.dL, al = P.alpha;
    const double t = a.k * a.tau;
    const double E = a.ml[(int64_t)p * a.ldml + (a.k - 1)];
    const double Ii1a = a.d0[p];
    const double A = -(1.0 / (h * h)) - (1.0 / (2.0 * h));
    double* C = a.U + P.uoff;
    double ca[CM], cf[CM], cal[CM], v[CM], c0[CM], hs[CM];
    chunk_sums<CM>(a.part, a.ldh, (a.k >= 2 && !P.ser) ? a.nch : 0, P.hoff, e0, n, hs);
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        const int e = e0 + k;
        ca[k] = cf[k] = cal[k] = v[k] = c0[k] = 0.0;
        if (e < n) {
            const int j = e + 1;
            c0[k] = C[j];
            double rp = Ii1a * c0[k] + ab_f(P.src, j * h, t, al, E);
            if (a.k >= 2) {
                if (P.ser)  // MF_intF_choice's value, already over (1 - a)
                    rp -= a.part[P.hoff + e];
                else
                    rp -= hs[k] / (1.0 - al);
            }
            v[k] = rp;
            ca[k] = A;
            cf[k] = a.cf[P.hoff + e];
            cal[k] = a.cal[P.hoff + e];
        }
    }
    const double b0 = ab_CC(P.src, 0.0, t), un = ab_CC(P.src, m * h, t);
    line_solve<CM>(ca, cf, cal, v, e0, n, b0, lane, &un);
    double es = 0.0;
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        const int e = e0 + k;
        if (e < n) {
            const int j = e + 1;
            C[j] = v[k];
            if (a.slot) a.slot[P.hoff + e] = (v[k] - c0[k]) / a.tau;
            const double d = fabs(v[k] - ab_CC(P.src, h * j, t));
            es += d * d;
        }
    }
    if (lane == 0) {
        C[0] = b0;  // bt[0] - al[0] C[1] with al[0] = 0
        C[m] = un;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) es += __shfl_xor(es, o);
    if (lane == 0) a.err[p] += es;  // the two ends equal CC exactly
}

// ---------------------------------------------------------------- AB series history
struct AbSeriesArgs {
    const MtProbDev* pr;
    const int* gprob;       // 128-node group -> problem
    int64_t ldh;
    const double* tab;      // per-step coefficient rows (ab_series_tables)
    const int* choice;      // per step: 1 series, 2 integral form
    const double* F;        // (C^{i-1} - C^{i-2}) / tau per history node
    double* lc1;            // [maxL][ldh] MF_intF's accumulators
    double* lc2;            // [maxL][ldh] MF_intF_Fv's accumulators
    signed char* lflag;     // [maxL][ldh] 0, or the sign of a log-form Lc2
    double* part;           // out: the series value per history node
    int i;                  // step, >= 2
    int* bad;               // bit 2: a non-finite term with ab_maxl above kRecountMaxL
};

// The longest series the sorted-pyramid recount holds in LDS.
constexpr int kRecountMaxL = 4096;

// The non-finite branch of MF_intF (MF_int_series.cpp:594-609 with
// F_nosorting = 1, where the `else` binds to the finiteness test) for node
// e.x, whose last non-finite term is e.y: each non-finite term replaces the
// running sum by sum_sort_pyramid (200-218) of all products Lc1[l] Kc[l]
// (0 where either is not finite), and the finite terms after the last such
// term are added to it in increasing l.  The whole workgroup works on the
// node: the products are sorted by decreasing magnitude in LDS (bitonic),
// then added pairwise level by level, an odd level's last element folded
// into the level's last sum.  Elements of equal magnitude and opposite sign
// may be ordered differently from std::sort; their sums are the same unless
// such a pair straddles a pair boundary.  Every thread of the workgroup
// calls it with the same node.
__device__ void ab_block_recount(const AbSeriesArgs& a, int2 e, double* v, double* w) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    const int64_t h = e.x;
    const MtProbDev& P = a.pr[a.gprob[h / kGroup]];
    const int maxl = P.ser_maxl;
    const double* __restrict__ T = a.tab + P.ser_off + (int64_t)(a.i - 2) * kAbSeriesRows * maxl;
    int n2 = 1;
    while (n2 < maxl) n2 <<= 1;
    for (int l = tid; l < n2; l += blockDim.x) {
        double x = 0.0;
        if (l < maxl) {
            const double c1 = a.lc1[(int64_t)l * a.ldh + h], kc = T[l];
            if (isfinite(c1) && isfinite(kc)) x = c1 * kc;
        }
        v[l] = x;
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int l = tid; l < n2; l += blockDim.x) {
                const int m = l ^ j;
                if (m > l) {
                    const double x = v[l], y = v[m];
                    const bool desc = (l & k) == 0;
                    if (desc ? fabs(x) < fabs(y) : fabs(x) > fabs(y)) {
                        v[l] = y;
                        v[m] = x;
                    }
                }
            }
            __syncthreads();
        }
    // the levels: v -> w -> v ...; the zeros padding v beyond maxl are never
    // read
    double* src = v;
    double* dst = w;
    int m = maxl;
    if (m == 1 && tid == 0) v[0] += v[0];  // a single element is doubled
    while (m > 1) {
        const int m2 = m >> 1;
        for (int j = tid; j < m2; j += blockDim.x) dst[j] = src[2 * j] + src[2 * j + 1];
        __syncthreads();
        if ((m & 1) && tid == 0) dst[m2 - 1] += src[m - 1];
        __syncthreads();
        double* t = src;
        src = dst;
        dst = t;
        m = m2;
    }
    if (tid == 0) {
        double ret = src[0];
        for (int l = e.y + 1; l < maxl; ++l) {
            const double c1 = a.lc1[(int64_t)l * a.ldh + h], kc = T[l];
            if (isfinite(c1) && isfinite(kc)) ret += c1 * kc;
        }
        a.part[h] = ret / (1.0 - P.alpha);
    }
    __syncthreads();  // v and w are reused by the next node
}

// One thread per node: MF_intF_choice(i, ..., on = i - 2, mn = 1, append_all
// = 1, Fvalue = F) (MF_int_series.cpp:839-853) on the node's accumulators.
// Both phases append to MF_intF_Fv's Lc2 (779-804: add sk F, or keep log|Lc2|
// with its sign once the sum overflows, after which the newest term alone
// is kept in log form); the series phase adds dp F to MF_intF's Lc1 and
// returns sum Lc1 Kc (598-603), the integral phase returns sum of the
// products of Lc2 and Kc with either in log form (806-822), terms in
// increasing l, over (1 - a).  Products and sums are rounded separately, as
// the reference's x86-64 build does.  Returns in *last_bad the last term
// whose Lc1 or Kc was not finite (-1: none) and in *maxl_out the series
// length.
__device__ inline void ab_series_node(const AbSeriesArgs& a, int64_t h, int* last_bad_out,
                                      int* maxl_out) {
#pragma clang fp contract(off)
    if (h >= a.ldh) return;
    const int p = a.gprob[h / kGroup];
    if (p < 0) return;
    const MtProbDev& P = a.pr[p];
    if (!P.ser || h - P.hoff >= P.N - 1) return;
    const int maxl = P.ser_maxl;
    const double* __restrict__ T = a.tab + P.ser_off + (int64_t)(a.i - 2) * kAbSeriesRows * maxl;
    const bool series = a.choice[P.ser_coff + a.i - 2] == 1;
    const double F = a.F[h];
    // the accumulators are only this node's: no access aliases another, so
    // the loads of several terms can be in flight together
    double* __restrict__ lc1 = a.lc1 + h;
    double* __restrict__ lc2 = a.lc2 + h;
    signed char* __restrict__ lflag = a.lflag + h;
    const int64_t ld = a.ldh;
    double ret = 0.0;
    int last_bad = -1;  // the last term whose Lc1 or Kc was not finite
    // one term: the accumulators of term l as loaded (c2, fl, c1)
    auto term = [&](int l, double c2, int fl, double c1) {
        const int64_t q = (int64_t)l * ld;
        if (fl == 0) {
            const double nv = c2 + T[3 * maxl + l] * F;
            if (!isfinite(nv)) {
                fl = c2 > 0.0 ? 1 : -1;
                c2 = log(fabs(c2));
            } else {
                c2 = nv;
            }
        } else {
            c2 = T[4 * maxl + l] + log(fabs(F));
            fl = F > 0.0 ? -1 : 1;
        }
        lc2[q] = c2;
        lflag[q] = (signed char)fl;
        const double kc = T[l];
        if (series) {
            c1 = c1 + T[2 * maxl + l] * F;
            lc1[q] = c1;
            if (isfinite(c1) && isfinite(kc))
                ret += c1 * kc;
            else
                last_bad = l;
        } else {
            const double lk = T[maxl + l];
            double v = 0.0;
            if (fl == 0 && lk == 0) v = c2 * kc;
            if (fl != 0 && lk == 0) v = fl * (kc > 0.0 ? 1.0 : -1.0) * exp(c2 + log(fabs(kc)));
            if (fl == 0 && lk != 0) v = lk * (c2 > 0.0 ? 1.0 : -1.0) * exp(kc + log(fabs(c2)));
            if (fl != 0 && lk != 0) v = lk * fl * exp(kc + c2);
            if (!isfinite(v)) v = 0.0;
            ret += v;
        }
    };
    // the accumulators of 8 terms are loaded before any is updated (the
    // terms are still added in increasing l)
    constexpr int B = 8;
    int l = 0;
    for (; l + B <= maxl; l += B) {
        double c2[B], c1[B];
        int fl[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int64_t q = (int64_t)(l + u) * ld;
            c2[u] = lc2[q];
            fl[u] = lflag[q];
            c1[u] = series ? lc1[q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < B; ++u) term(l + u, c2[u], fl[u], c1[u]);
    }
    for (; l < maxl; ++l) {
        const int64_t q = (int64_t)l * ld;
        term(l, lc2[q], lflag[q], series ? lc1[q] : 0.0);
    }
    a.part[h] = ret / (1.0 - P.alpha);
    *last_bad_out = last_bad;
    *maxl_out = maxl;
}

// One thread per node (ab_series_node); then the workgroup recounts, one
// node after another, the sums of its nodes that met a non-finite term
// (ab_block_recount), overwriting their part.  Three workgroups per CU: the
// recount's code would otherwise raise the kernel to 175 VGPRs (two waves
// per SIMD instead of three); held to 168 it spills 36 bytes per lane
// outside the term loop.
__global__ __launch_bounds__(256, 3) void k_ab_series(AbSeriesArgs a) {
    __shared__ double v[kRecountMaxL];
    __shared__ double w[kRecountMaxL / 2];
    __shared__ int2 todo[256];
    __shared__ int ntodo;
    if (threadIdx.x == 0) ntodo = 0;
    __syncthreads();
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int last_bad = -1, maxl = 0;
    ab_series_node(a, h, &last_bad, &maxl);
    if (last_bad >= 0) {
        if (maxl <= kRecountMaxL)
            todo[atomicAdd(&ntodo, 1)] = make_int2((int)h, last_bad);
        else
            atomicOr(a.bad, 2);
    }
    __syncthreads();
    const int n = ntodo;
    for (int q = 0; q < n; ++q) ab_block_recount(a, todo[q], v, w);
}

// ---------------------------------------------------------------- fitting error
// solve_and_test (mt_minimal.cpp:1277-1330) for every problem of the batch:
// one thread per problem walks the observations of the current step in
// order (so the sum is the reference's, term by term).
struct FitArgs {
    const MtProbDev* pr;
    int nprob;
    const double* U;
    const double* T;   // observation times, depths, values
    const double* Z;
    const double* F;
    const int* idx;    // the observations of this step, in order
    int cnt;
    double tt, tau;
    double* vs;        // [nprob][nobs] values before the step
    int nobs;
    double* err;       // [nprob]
    int* dead;         // [nprob]: U[0] was not finite after a step
};

// U at depth z between nodes j, j+1 with dL j < z <= dL (j+1), j < jmax;
// false when no such j
__device__ inline bool fit_interp(const double* U, double dL, int jmax, double z, double& v) {
    const double r = z / dL;
    if (!(r > 0.0) || r > jmax + 2.0) return false;  // no interval (also NaN)
    int j = (int)floor(r) - 1;
    if (j < 0) j = 0;
    for (int q = 0; q < 3 && j < jmax; ++q, ++j)
        if (((dL * j) < z) && ((dL * (j + 1)) >= z)) {
            const double k2 = (z - (dL * j)) / dL;
            v = (1 - k2) * U[j] + k2 * U[j + 1];
            return true;
        }
    return false;
}

__global__ __launch_bounds__(256) void k_fit_before(FitArgs a) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nprob) return;
    const MtProbDev& P = a.pr[p];
    const double* U = a.U + P.uoff;
    for (int q = 0; q < a.cnt; ++q) {
        const int o = a.idx[q];
        double v;
        if (fit_interp(U, P.dL, P.N - 1, a.Z[o], v)) a.vs[(int64_t)p * a.nobs + o] = v;
    }
}

__global__ __launch_bounds__(256) void k_fit_after(FitArgs a) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.nprob || a.dead[p]) return;
    const MtProbDev& P = a.pr[p];
    const double* U = a.U + P.uoff;
    double e = a.err[p];
    for (int q = 0; q < a.cnt; ++q) {
        const int o = a.idx[q];
        const double k1 = (a.T[o] - a.tt) / a.tau;
        double v;
        if (fit_interp(U, P.dL, P.N, a.Z[o], v)) {
            v = (1 - k1) * a.vs[(int64_t)p * a.nobs + o] + k1 * v;
            if (v > 0.0) v = 0.0;
            e += (a.F[o] - v) * (a.F[o] - v);
        }
    }
    if (!isfinite(U[0])) {
        e = 1e300;
        a.dead[p] = 1;
    }
    a.err[p] = e;
}

}  // namespace frm

using namespace frm;

namespace frm {
// A kernel launch of one step: function, shape and a copy of its argument
// values (hipKernelNodeParams takes pointers to them).
struct KRec {
    const void* func = nullptr;
    dim3 grid, block;
    int step = 0;         // step of the recorded run it belongs to
    bool branch = false;  // weights formed ahead (kind psi): a branch of the graph
    alignas(16) unsigned char buf[448];
    size_t used = 0;
    size_t off[16];
    int n = 0;
    template <class T>
    void arg(const T& v) {
        static_assert(sizeof(T) <= 448, "kernel argument too large");
        used = (used + alignof(T) - 1) / alignof(T) * alignof(T);
        std::memcpy(buf + used, &v, sizeof(T));
        off[n++] = used;
        used += sizeof(T);
    }
    void ptrs(void** p) {
        for (int i = 0; i < n; ++i) p[i] = buf + off[i];
    }
};
// The launches of up to kGraphSteps consecutive steps as an instantiated
// graph (each step's kernel nodes in a chain; for kind psi the next step's
// weights on a branch beside it), reused for every run of steps that issues
// the same kernels in the same order: only the nodes' arguments and grids
// change between launches.
constexpr int kGraphSteps = 32;
struct StepGraph {
    std::vector<const void*> funcs;
    std::vector<int> tags;  // step * 2 + branch of every node
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphNode_t> nodes;
};
}  // namespace frm

struct frm_mt1d {
    frm_ctx* ctx = nullptr;
    int nprob = 0, kind = 0, capacity = 0, precision = 0, picard = 1;
    double tau = 0;
    int tstep = 1;
    int depth = 0;
    int slots = 0;
    int64_t nU = 0, ldh = 0, ngroups = 0;
    int maxN = 0;
    std::vector<MtProbDev> hprobs;
    MtProbDev* probs = nullptr;
    int* gprob = nullptr;
    int* gvalid = nullptr;    // real nodes per group (0 for padding groups)
    int min_chunk = kMinChunk;      // history block length: shortest time chunk of the memory term
    int graph = 0;                  // steps as HIP graphs, built once per launch sequence
    std::vector<frm::StepGraph> graphs;  // one instantiated graph per launch sequence
    std::vector<frm::KRec>* rec = nullptr;  // while set, launches are recorded, not issued
    int rec_step = 0;                        // step of the recorded run being recorded
    hipStream_t lstream = nullptr;          // while set, launches go to this stream
    // kind psi: the weights of step k + 1 are formed on the context's second
    // stream while step k's memory term and sweep run on its stream (the
    // weights do not depend on the state); w, w32 and d0 then hold two
    // buffers, step k using buffer k & 1.  wready[b]: buffer b's weights are
    // formed; wfree[b]: the last step that read buffer b is done.
    hipStream_t wstream = nullptr;
    hipEvent_t wready[2] = {nullptr, nullptr}, wfree[2] = {nullptr, nullptr};
    int wahead = 0;                         // step whose weights were launched ahead (0: none)
    size_t wsz = 0;                         // elements of one w buffer (0: one buffer)
    int64_t hnodes = 0, hnodes_mt = 0;  // real history nodes (all; full-history problems)
    double* U = nullptr;
    void* hist = nullptr;     // slots * ldh (FP64 or FP32)
    double* part = nullptr;   // kMaxChunks * ldh
    double* w = nullptr;      // nprob * ldw FP64 weights (lag-indexed)
    float* w32 = nullptr;     // FP32 copy for FRM_PREC_F32C
    double* d0 = nullptr;     // per-problem current-step coefficient
    int* bad = nullptr;       // set when a psi weight is not finite
    int ldw = 0;
    // kind AB
    double* ab_cf = nullptr;  // [ldh] -1/(B - A Al_{j-1})
    double* ab_cal = nullptr; // [ldh] -Al_j
    double* ab_ml = nullptr;  // [nprob][capacity] source factors
    double* ab_err = nullptr; // [nprob]
    // kind AB, series history (ab_series)
    int ser_maxL = 0, ser_count = 0;
    int* gprob_mt = nullptr;  // group -> problem for the full-history sums (-1: series)
    double* ser_tab = nullptr;
    int* ser_choice = nullptr;
    double* ser_lc = nullptr;        // [2][maxL][ldh]: Lc1, Lc2
    signed char* ser_flag = nullptr; // [maxL][ldh]
    // kind psi: (problem, interval) pairs whose weight series was not finite
    int2* psi_queue = nullptr;
    int* psi_qcount = nullptr;       // [2] queue lengths (by step parity)
    int psi_qcap = 0;
    std::vector<double> ser_bytes;   // per step i >= 2: algorithmic bytes of k_ab_series
    // fitting error (solve_and_test)
    int nobs = 0;
    std::vector<double> obs_T;       // host copy: which observations fall in which step
    double tt = 0.0;                 // start of the next step, accumulated as the reference does
    double* fit_obs = nullptr;       // [3][nobs] T, Z, F
    int* fit_idx = nullptr;          // [nobs] scratch: the observations of one step
    double* fit_vs = nullptr;        // [nprob][nobs]
    double* fit_err = nullptr;       // [nprob]
    int* fit_dead = nullptr;         // [nprob]
    std::vector<int> fit_off;        // [capacity+1]: observations of step k at fit_idx[off[k-1] ..]
};

// Issues a kernel launch on the context's stream, or records it for the
// step graph while s->rec is set; arguments are converted to the kernel's
// parameter types.
template <typename... KArgs, typename... Args>
static hipError_t mt_launch(frm_mt1d* s, void (*kern)(KArgs...), dim3 grid, dim3 block,
                            Args&&... args) {
    static_assert(sizeof...(KArgs) == sizeof...(Args), "argument count");
    frm::KRec r;
    r.func = (const void*)kern;
    r.grid = grid;
    r.block = block;
    (r.arg(static_cast<std::decay_t<KArgs>>(args)), ...);
    if (s->rec) {
        r.step = s->rec_step;
        r.branch = s->lstream != nullptr;
        s->rec->push_back(r);
        return hipSuccess;
    }
    void* p[16];
    r.ptrs(p);
    return hipLaunchKernel(r.func, grid, block, p, 0, s->lstream ? s->lstream : s->ctx->stream);
}

// The weight buffers of step k (kind psi: two, by step parity).
static double* mt_w(const frm_mt1d* s, int k) {
    return s->w + (s->wsz ? (size_t)(k & 1) * s->wsz : 0);
}
static float* mt_w32(const frm_mt1d* s, int k) {
    return s->w32 ? s->w32 + (s->wsz ? (size_t)(k & 1) * s->wsz : 0) : nullptr;
}
static double* mt_d0(const frm_mt1d* s, int k) {
    return s->d0 + (s->wsz ? (size_t)(k & 1) * s->nprob : 0);
}

// Kind psi: the weights of step k at `depth` history slots, then the
// quadrature fallback of the intervals whose series was not finite.
static hipError_t mt_launch_psi_weights(frm_mt1d* s, int k, int depth) {
    const int threads = 256;
    const dim3 grid((unsigned)((depth + 1 + threads - 1) / threads), (unsigned)s->nprob);
    hipError_t e = mt_launch(s, k_psi_weights, grid, dim3(threads), s->probs, s->nprob, k, depth,
                             s->tau, mt_w(s, k), s->ldw, mt_w32(s, k), mt_d0(s, k), s->bad,
                             s->psi_queue, s->psi_qcap, s->psi_qcount);
    // a fixed grid: threads beyond the queue's length exit at once
    if (e == hipSuccess)
        e = mt_launch(s, k_psi_quad, dim3(64), dim3(64), s->probs, k, depth, s->tau, mt_w(s, k),
                      s->ldw, mt_w32(s, k), mt_d0(s, k), s->bad, s->psi_queue, s->psi_qcap,
                      s->psi_qcount);
    return e;
}

// The same on the weight stream, once the last reader of step k's buffers
// is done; wready of the buffer marks the end.  While a graph is recorded
// the launches become a branch of it, ordered by the graph's edges instead.
static hipError_t mt_launch_psi_weights_ahead(frm_mt1d* s, int k, int depth) {
    hipError_t e = s->rec ? hipSuccess : hipStreamWaitEvent(s->wstream, s->wfree[k & 1], 0);
    s->lstream = s->wstream;
    if (e == hipSuccess) e = mt_launch_psi_weights(s, k, depth);
    s->lstream = nullptr;
    if (e == hipSuccess && !s->rec) e = hipEventRecord(s->wready[k & 1], s->wstream);
    return e;
}

// Time chunks of the memory term at `depth` history slots: one, unless the
// nodes alone give fewer than kMemtermWaves wavefronts; chunks are at least
// `min_chunk` slots (the history block length) and at most kMaxChunks.
static int mt_chunks(const frm_mt1d* s, int depth, int min_chunk) {
    const int64_t waves = s->ldh / (2 * kWave);
    int nch = 1;
    if (waves < kMemtermWaves) nch = (int)((kMemtermWaves + waves - 1) / waves);
    return std::max(1, std::min<int>({nch, (depth + min_chunk - 1) / min_chunk, kMaxChunks}));
}

// The memory term of step k: `depth` history slots in `nch` time chunks.
static hipError_t mt_launch_memterm(frm_mt1d* s, int k, int depth, int nch) {
    const int chunk = (depth + nch - 1) / nch;
    const dim3 grid((unsigned)(s->ldh / 2 / kMemtermThreads), (unsigned)nch);
    if (s->precision == FRM_PREC_F32C) {
        // four nodes per lane where the batch fills the GPU with half the wavefronts
        if (nch == 1 && s->ldh / (4 * kWave) >= kMemtermWaves) {
            const dim3 g4((unsigned)((s->ldh / 4 + kMemtermThreads - 1) / kMemtermThreads), 1);
            return mt_launch(s, k_mt_memterm_f32c<4>, g4, dim3(kMemtermThreads),
                             (const float*)s->hist, s->ldh, depth, mt_w32(s, k), s->ldw, s->gprob,
                             s->gvalid, chunk, s->part);
        }
        return mt_launch(s, k_mt_memterm_f32c<2>, grid, dim3(kMemtermThreads), (const float*)s->hist,
                         s->ldh, depth, mt_w32(s, k), s->ldw, s->gprob, s->gvalid, chunk, s->part);
    }
    return mt_launch(s, k_mt_memterm_f64, grid, dim3(kMemtermThreads), (const double*)s->hist,
                     s->ldh, depth, mt_w(s, k), s->ldw, s->gprob_mt ? s->gprob_mt : s->gprob,
                     s->gvalid, chunk, s->part);
}

// The sweep of step k over the memory term's `nch` chunks; the step's
// increment goes to history slot `depth` when `push`.
static hipError_t mt_launch_sweep(frm_mt1d* s, int k, int depth, int nch, bool push) {
    const dim3 g((unsigned)((s->nprob + 3) / 4));
    const int cm = mt_pick_cm(s->maxN);
    const size_t esz = s->precision == FRM_PREC_F32C ? sizeof(float) : sizeof(double);
    if (s->kind == FRM_MT_AB) {
        AbSweepArgs b;
        b.pr = s->probs;
        b.nprob = s->nprob;
        b.U = s->U;
        b.part = s->part;
        b.nch = nch;
        b.ldh = s->ldh;
        b.d0 = mt_d0(s, k);
        b.cf = s->ab_cf;
        b.cal = s->ab_cal;
        b.ml = s->ab_ml;
        b.ldml = s->capacity;
        b.k = k;
        b.tau = s->tau;
        b.slot = (double*)s->hist + (size_t)depth * s->ldh;
        b.err = s->ab_err;
        switch (cm) {
            case 4: return mt_launch(s, k_ab1d_sweep<4>, g, dim3(256), b);
            case 8: return mt_launch(s, k_ab1d_sweep<8>, g, dim3(256), b);
            default: return mt_launch(s, k_ab1d_sweep<16>, g, dim3(256), b);
        }
    }
    MtSweepArgs a;
    a.pr = s->probs;
    a.nprob = s->nprob;
    a.U = s->U;
    a.part = s->part;
    a.nch = nch;
    a.ldh = s->ldh;
    a.d0 = mt_d0(s, k);
    a.slot = push ? (void*)((char*)s->hist + esz * (size_t)depth * s->ldh) : nullptr;
    a.slot_f32 = s->precision == FRM_PREC_F32C;
    a.picard = s->picard;
    switch (cm) {
        case 4: return mt_launch(s, k_mt_sweep<4>, g, dim3(256), a);
        case 8: return mt_launch(s, k_mt_sweep<8>, g, dim3(256), a);
        default: return mt_launch(s, k_mt_sweep<16>, g, dim3(256), a);
    }
}

// Autotuner of the history block length, the shortest time chunk of the
// memory term.  With few nodes the history is split into time chunks so that
// enough wavefronts stream it, and the sweep then adds the chunks' partial
// sums: shorter chunks give the memory term more wavefronts and the sweep
// more partials to add.  Each candidate length's memory term and sweep are
// timed at the run's mean depth (capacity / 2 history slots) on the solver's
// own buffers, before any state is set (the state is zeroed first; the
// sweeps of kinds L1 and psi write no history slot, the AB sweep writes slot
// D, which a step overwrites before it is read), and the fastest is kept.
// A batch whose nodes alone fill the GPU takes one chunk whatever the length
// and is not timed.
static int mt_tune_block(frm_ctx* ctx, frm_mt1d* s, int force) {
    if (force > 0) {
        s->min_chunk = force;
        return FRM_OK;
    }
    s->min_chunk = kMinChunk;
    const int D = std::min(s->slots - 1, s->capacity / 2);
    if (s->ser_count == s->nprob || s->ldh / (2 * kWave) >= kMemtermWaves || D < 32) return FRM_OK;
    static const int kCand[] = {16, 32, 64, 128, 256, 512};
    constexpr int kReps = 20;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipMemsetAsync(s->U, 0, sizeof(double) * s->nU, ctx->stream);
    float best = 0.f;
    int last_nch = -1;
    for (int L : kCand) {
        if (e != hipSuccess) break;
        const int nch = mt_chunks(s, D, L);
        if (nch == last_nch) continue;  // the same launches as the shorter length
        last_nch = nch;
        for (int r = -2; r < kReps && e == hipSuccess; ++r) {
            if (r == 0) e = hipEventRecord(a, ctx->stream);
            if (e == hipSuccess) e = mt_launch_memterm(s, 0, D, nch);
            if (e == hipSuccess) e = mt_launch_sweep(s, 0, D, nch, false);
        }
        if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess && (best == 0.f || ms < best)) {
            best = ms;
            s->min_chunk = L;
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return e == hipSuccess ? FRM_OK : check_hip(ctx, e, "frm_mt1d_create: history block tuning");
}

static void mt_free(frm_mt1d* s) {
    if (!s) return;
    (void)hipFree(s->probs);
    (void)hipFree(s->gprob);
    (void)hipFree(s->gvalid);
    for (auto& g : s->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    (void)hipFree(s->U);
    (void)hipFree(s->hist);
    (void)hipFree(s->part);
    if (s->wstream) (void)hipStreamSynchronize(s->wstream);  // the context's, shared
    for (int b = 0; b < 2; ++b) {
        if (s->wready[b]) (void)hipEventDestroy(s->wready[b]);
        if (s->wfree[b]) (void)hipEventDestroy(s->wfree[b]);
    }
    (void)hipFree(s->w);
    (void)hipFree(s->w32);
    (void)hipFree(s->d0);
    (void)hipFree(s->bad);
    (void)hipFree(s->ab_cf);
    (void)hipFree(s->ab_cal);
    (void)hipFree(s->ab_ml);
    (void)hipFree(s->ab_err);
    (void)hipFree(s->gprob_mt);
    (void)hipFree(s->fit_obs);
    (void)hipFree(s->fit_idx);
    (void)hipFree(s->fit_vs);
    (void)hipFree(s->fit_err);
    (void)hipFree(s->fit_dead);
    (void)hipFree(s->ser_tab);
    (void)hipFree(s->ser_choice);
    (void)hipFree(s->ser_lc);
    (void)hipFree(s->ser_flag);
    (void)hipFree(s->psi_queue);
    (void)hipFree(s->psi_qcount);
    delete s;
}

// Kind AB: kernel integrals I(d) (history weights I(lag + 1), current
// I(1)/(1-a)), the constant Thomas alpha recurrence of solve_and_check
// (MF_int_series.cpp:906-940) and the Mittag-Leffler factors of the source
// for every step, all in FP64 on the host.
static int ab_setup(frm_ctx* ctx, frm_mt1d* s, const frm_mt1d_desc* d) {
    const int cap = s->capacity;
    std::vector<double> w((size_t)s->nprob * s->ldw, 0.0), d0(s->nprob);
    std::vector<double> cf(s->ldh, 0.0), cal(s->ldh, 0.0), ml((size_t)s->nprob * cap);
    // problems with the same order and weight method share I(1..cap+1), and
    // with the same order and source the Mittag-Leffler factors
    std::map<std::tuple<double, int, int, double, int>, std::vector<double>> icache;
    std::map<std::pair<double, int>, std::vector<double>> mcache;
    for (int p = 0; p < s->nprob; ++p) {
        const frm_mt1d_problem& q = d->probs[p];
        const MtProbDev& P = s->hprobs[p];
        const double a = q.alpha, tau = d->tau;
        const int maxl = q.ab_maxl > 0 ? q.ab_maxl : 1000;
        const int maxk = q.ab_maxkiter > 0 ? q.ab_maxkiter : 1000;
        const double eps = q.ab_eps > 0 ? q.ab_eps : 1e-10;
        const auto key = std::make_tuple(a, maxl, maxk, eps, q.ab_exact ? 1 : q.ab_series ? 2 : 0);
        auto it = icache.find(key);
        if (it == icache.end()) {
            std::vector<double> I(cap + 1, 0.0);  // I[d-1] = I(d)
            if (q.ab_exact) {
                ab_exact_weights(a, tau, cap + 1, I.data());
            } else {
                // the current weight takes the problem's eps, the history
                // 1e-10 (MF_int_series.cpp:917 and the cached I of
                // MF_intF_simple); the series history needs I(1) only
                AbIntegrator cur(a, maxl, maxk, eps);
                I[0] = cur.integral(tau, tau, 0.0);
                if (!q.ab_series) {
                    AbIntegrator hist(a, maxl, maxk, 1e-10);
                    for (int dd = 2; dd <= cap + 1; ++dd) I[dd - 1] = hist.integral(dd * tau, tau, 0.0);
                }
            }
            it = icache.emplace(key, std::move(I)).first;
        }
        const std::vector<double>& I = it->second;
        const double Ii = I[0];
        for (int l = 1; l <= s->ldw; ++l) w[(size_t)p * s->ldw + l - 1] = I[l];
        d0[p] = Ii / (1.0 - a);
        const int m = P.N;
        const double h = 1.0 / m;
        const double A = -(1.0 / (h * h)) - (1.0 / (2.0 * h));
        const double B = (Ii / (1.0 - a)) + (2.0 / (h * h));
        const double Cc = -(1.0 / (h * h)) + (1.0 / (2.0 * h));
        double al = 0.0;
        for (int j = 1; j < m; ++j) {
            const double den = B - A * al;
            al = Cc / den;
            cf[P.hoff + j - 1] = -1.0 / den;
            cal[P.hoff + j - 1] = -al;
        }
        const double beta = q.src == 0 ? 5.0 : (q.src == 1 ? 3.0 : 6.0);
        auto mt = mcache.find({a, q.src});
        if (mt == mcache.end()) {
            std::vector<double> v(cap);
            for (int k = 1; k <= cap; ++k)
                v[k - 1] = ml_eval(a, beta, (-a / (1.0 - a)) * std::pow(k * tau, a), 1e-20);
            mt = mcache.emplace(std::make_pair(a, q.src), std::move(v)).first;
        }
        std::copy(mt->second.begin(), mt->second.end(), ml.begin() + (size_t)p * cap);
    }
    // series history: the node-independent coefficients of every step, per
    // distinct (alpha, maxl, maxkiter, eps)
    std::vector<double> stab;
    std::vector<int> sch;
    if (s->ser_count > 0 && cap >= 2) {
        std::map<std::tuple<double, int, int, double>, std::pair<int64_t, int64_t>> scache;
        for (int p = 0; p < s->nprob; ++p) {
            MtProbDev& P = s->hprobs[p];
            if (!P.ser) continue;
            const frm_mt1d_problem& q = d->probs[p];
            const int maxk = q.ab_maxkiter > 0 ? q.ab_maxkiter : 1000;
            const double eps = q.ab_eps > 0 ? q.ab_eps : 1e-10;
            const auto key = std::make_tuple(q.alpha, P.ser_maxl, maxk, eps);
            auto it = scache.find(key);
            if (it == scache.end()) {
                const int64_t off = (int64_t)stab.size(), coff = (int64_t)sch.size();
                const int64_t rows = (int64_t)(cap - 1) * kAbSeriesRows * P.ser_maxl;
                if (off + rows > (int64_t(1) << 28))
                    return set_error(ctx, FRM_E_INVALID,
                                     "frm_mt1d_create: series coefficient tables of %.1f GB "
                                     "(capacity x ab_maxl x %d doubles per distinct order)",
                                     8.0 * (off + rows) / 1e9, kAbSeriesRows);
                stab.resize(off + rows);
                sch.resize(coff + cap - 1);
                ab_series_tables(q.alpha, d->tau, P.ser_maxl, maxk, eps, cap, stab.data() + off,
                                 sch.data() + coff);
                it = scache.emplace(key, std::make_pair(off, coff)).first;
            }
            P.ser_off = it->second.first;
            P.ser_coff = it->second.second;
        }
        // per node and series term: Lc2 and its flag read and written, and in
        // the power-series phase Lc1 too; per node the increment and the value
        s->ser_bytes.assign(cap - 1, 0.0);
        for (int p = 0; p < s->nprob; ++p) {
            const MtProbDev& P = s->hprobs[p];
            if (!P.ser) continue;
            for (int i = 2; i <= cap; ++i) {
                const bool series = sch[P.ser_coff + i - 2] == 1;
                s->ser_bytes[i - 2] += (double)(P.N - 1) * (P.ser_maxl * (series ? 34.0 : 18.0) + 16.0);
            }
        }
    }
    hipError_t e = hipSuccess;
    if (s->ser_count > 0) {
        // the full-history sums skip the series problems' groups
        std::vector<int> gm(s->ngroups);
        e = hipMemcpy(gm.data(), s->gprob, sizeof(int) * gm.size(), hipMemcpyDeviceToHost);
        for (int& g : gm)
            if (g >= 0 && s->hprobs[g].ser) g = -1;
        const size_t nl = (size_t)s->ser_maxL * s->ldh;
        if (e == hipSuccess) e = hipMalloc(&s->gprob_mt, sizeof(int) * gm.size());
        if (e == hipSuccess) e = hipMemcpy(s->gprob_mt, gm.data(), sizeof(int) * gm.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess && !stab.empty()) e = hipMalloc(&s->ser_tab, sizeof(double) * stab.size());
        if (e == hipSuccess && !sch.empty()) e = hipMalloc(&s->ser_choice, sizeof(int) * sch.size());
        if (e == hipSuccess && !stab.empty())
            e = hipMemcpy(s->ser_tab, stab.data(), sizeof(double) * stab.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess && !sch.empty())
            e = hipMemcpy(s->ser_choice, sch.data(), sizeof(int) * sch.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&s->ser_lc, sizeof(double) * 2 * nl);
        if (e == hipSuccess) e = hipMalloc(&s->ser_flag, nl);
        if (e == hipSuccess) e = hipMemset(s->ser_lc, 0, sizeof(double) * 2 * nl);
        if (e == hipSuccess) e = hipMemset(s->ser_flag, 0, nl);
        if (e == hipSuccess)
            e = hipMemcpy(s->probs, s->hprobs.data(), sizeof(MtProbDev) * s->nprob, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMalloc(&s->ab_cf, sizeof(double) * s->ldh);
    if (e == hipSuccess) e = hipMalloc(&s->ab_cal, sizeof(double) * s->ldh);
    if (e == hipSuccess) e = hipMalloc(&s->ab_ml, sizeof(double) * ml.size());
    if (e == hipSuccess) e = hipMalloc(&s->ab_err, sizeof(double) * s->nprob);
    if (e == hipSuccess) e = hipMemcpy(s->w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d0, d0.data(), sizeof(double) * d0.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->ab_cf, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->ab_cal, cal.data(), sizeof(double) * cal.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->ab_ml, ml.data(), sizeof(double) * ml.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->ab_err, 0, sizeof(double) * s->nprob);
    return e == hipSuccess ? FRM_OK : check_hip(ctx, e, "mt1d AB setup");
}

extern "C" {

int frm_mt1d_create(frm_ctx* ctx, const frm_mt1d_desc* d, frm_mt1d** out) {
    FRM_REQUIRE(ctx, ctx && d && out && d->probs, "frm_mt1d_create: null argument");
    *out = nullptr;
    FRM_REQUIRE(ctx, d->nprob >= 1 && d->capacity >= 1 && d->picard >= 1 && d->tau > 0,
                "frm_mt1d_create: bad batch descriptor");
    FRM_REQUIRE(ctx, d->kind == FRM_MT_L1 || d->kind == FRM_MT_PSI || d->kind == FRM_MT_AB,
                "frm_mt1d_create: kind %d", d->kind);
    FRM_REQUIRE(ctx, d->kind != FRM_MT_AB || (d->picard == 1 && d->precision == FRM_PREC_F64),
                "frm_mt1d_create: kind AB takes picard 1 and FRM_PREC_F64");
    FRM_REQUIRE(ctx, d->precision == FRM_PREC_F64 || d->precision == FRM_PREC_F32C,
                "frm_mt1d_create: precision %d", d->precision);
    FRM_REQUIRE(ctx, d->hist_block >= 0, "frm_mt1d_create: hist_block %d", d->hist_block);
    FRM_HIP(ctx, hipSetDevice(ctx->device));
    frm_mt1d* s = new frm_mt1d();
    s->ctx = ctx;
    s->nprob = d->nprob;
    s->kind = d->kind;
    s->capacity = d->capacity;
    s->precision = d->precision;
    s->picard = d->picard;
    s->tau = d->tau;
    s->graph = d->graph != 0;
    // L1/psi: increments enter from step 3 on; AB: every increment
    const bool is_ab = d->kind == FRM_MT_AB;
    s->slots = is_ab ? d->capacity : std::max(1, d->capacity - 2);
    s->ldw = s->slots;
    std::vector<int> gp, gv;
    int64_t uoff = 0, hoff = 0;
    for (int p = 0; p < d->nprob; ++p) {
        const frm_mt1d_problem& q = d->probs[p];
        const bool ser = is_ab && q.ab_series != 0;
        const int ser_maxl = q.ab_maxl > 0 ? q.ab_maxl : 1000;
        if (q.N < 2 || q.N > kMtMaxN || !(q.alpha > 0 && q.alpha < 1) || !(q.L > 0) ||
            q.form < 0 || q.form > 2 || q.top < 0 || q.top > 2 ||
            (is_ab && (q.src < 0 || q.src > 2)) || (ser && (q.ab_exact || ser_maxl < 3))) {
            mt_free(s);
            return set_error(ctx, FRM_E_INVALID, "frm_mt1d_create: problem %d invalid (N=%d)", p,
                             q.N);
        }
        MtProbDev P{};
        P.N = q.N;
        P.form = q.form;
        P.top = q.top;
        P.bottom_cond = q.bottom_cond;
        P.src = q.src;
        P.uoff = uoff;
        P.hoff = hoff;
        P.dL = is_ab ? 1.0 / q.N : q.L / q.N;
        P.Dg = std::pow(P.dL, -2) / 1.0;  // Gamma(2 - gamma) = Gamma(1) = 1
        P.alpha = q.alpha;
        P.beta = q.beta;
        P.Da = (d->kind == FRM_MT_PSI ? std::pow(d->tau, -1) : std::pow(d->tau, -q.alpha)) /
               std::tgamma(2.0 - q.alpha);
        P.D = q.D;
        P.k = q.k;
        P.k0 = q.k0;
        P.b = q.b;
        P.top_val = q.top_val;
        P.Ev = q.Ev;
        P.Hbottom = q.Hbottom;
        P.ser = ser ? 1 : 0;
        P.ser_maxl = ser ? ser_maxl : 0;
        if (ser) {
            s->ser_count += 1;
            s->ser_maxL = std::max(s->ser_maxL, ser_maxl);
        }
        s->hprobs.push_back(P);
        const int ng = (q.N - 1 + kGroup - 1) / kGroup;
        for (int g = 0; g < ng; ++g) {
            gp.push_back(p);
            gv.push_back(std::min(kGroup, q.N - 1 - g * kGroup));
        }
        s->hnodes += q.N - 1;
        if (!ser) s->hnodes_mt += q.N - 1;
        hoff += (int64_t)ng * kGroup;
        uoff += q.N + 2;
        s->maxN = std::max(s->maxN, q.N);
    }
    s->nU = uoff;
    // whole memory-term workgroups: 64-node groups of kMemtermThreads/64 waves
    const int64_t wg_nodes = 2 * kMemtermThreads;
    s->ldh = (hoff + wg_nodes - 1) / wg_nodes * wg_nodes;
    s->ngroups = s->ldh / kGroup;
    while ((int64_t)gp.size() < s->ngroups) {
        gp.push_back(-1);
        gv.push_back(0);
    }
    const size_t esz = d->precision == FRM_PREC_F32C ? sizeof(float) : sizeof(double);
    hipError_t e = hipMalloc(&s->probs, sizeof(MtProbDev) * d->nprob);
    if (e == hipSuccess) e = hipMalloc(&s->gprob, sizeof(int) * gp.size());
    if (e == hipSuccess) e = hipMalloc(&s->gvalid, sizeof(int) * gv.size());
    if (e == hipSuccess) e = hipMalloc(&s->U, sizeof(double) * s->nU);
    if (e == hipSuccess) e = hipMalloc(&s->hist, esz * (size_t)s->slots * s->ldh);
    if (e == hipSuccess) e = hipMalloc(&s->part, sizeof(double) * (size_t)kMaxChunks * s->ldh);
    // kind psi: two buffers of weights, the next step's formed ahead
    const size_t nwb = d->kind == FRM_MT_PSI ? 2 : 1;
    if (d->kind == FRM_MT_PSI) s->wsz = (size_t)d->nprob * s->ldw;
    if (e == hipSuccess) e = hipMalloc(&s->w, sizeof(double) * nwb * (size_t)d->nprob * s->ldw);
    if (e == hipSuccess && d->precision == FRM_PREC_F32C)
        e = hipMalloc(&s->w32, sizeof(float) * nwb * (size_t)d->nprob * s->ldw);
    if (e == hipSuccess) e = hipMalloc(&s->d0, sizeof(double) * nwb * d->nprob);
    if (e == hipSuccess && d->kind == FRM_MT_PSI) {
        if (!ctx->aux_stream) e = hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking);
        s->wstream = ctx->aux_stream;
        for (int b = 0; b < 2 && e == hipSuccess; ++b) {
            e = hipEventCreateWithFlags(&s->wready[b], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&s->wfree[b], hipEventDisableTiming);
        }
    }
    if (e == hipSuccess) e = hipMalloc(&s->bad, sizeof(int));
    if (e == hipSuccess && d->kind == FRM_MT_PSI) {
        // one entry per weight of a step, up to 2^20 (more set bad bit 4)
        s->psi_qcap = (int)std::min<int64_t>((int64_t)d->nprob * (s->capacity + 1), 1 << 20);
        e = hipMalloc(&s->psi_queue, sizeof(int2) * s->psi_qcap);
        if (e == hipSuccess) e = hipMalloc(&s->psi_qcount, 2 * sizeof(int));
        if (e == hipSuccess) e = hipMemset(s->psi_qcount, 0, 2 * sizeof(int));
    }
    if (e != hipSuccess) {
        mt_free(s);
        return check_hip(ctx, e, "hipMalloc(mt1d)");
    }
    // L1 weights: kappa(l+2) = (l+2)^(1-a) - (l+1)^(1-a) as the reference forms
    // them (mt_minimal.cpp:926-931); psi weights are rebuilt every step.
    std::vector<double> w((size_t)d->nprob * s->ldw, 0.0), d0(d->nprob);
    std::vector<float> w32(d->precision == FRM_PREC_F32C ? w.size() : 0);
    for (int p = 0; p < d->nprob; ++p) {
        const double a = d->probs[p].alpha;
        d0[p] = s->hprobs[p].Da;
        if (d->kind == FRM_MT_L1)
            for (int l = 0; l < s->ldw; ++l) {
                const double kv = std::pow((double)(l + 2), 1.0 - a) - std::pow((double)(l + 1), 1.0 - a);
                w[(size_t)p * s->ldw + l] = kv;
                if (!w32.empty()) w32[(size_t)p * s->ldw + l] = (float)kv;
            }
    }
    e = hipMemcpy(s->probs, s->hprobs.data(), sizeof(MtProbDev) * d->nprob, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->gprob, gp.data(), sizeof(int) * gp.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->gvalid, gv.data(), sizeof(int) * gv.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !w32.empty())
        e = hipMemcpy(s->w32, w32.data(), sizeof(float) * w32.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->d0, d0.data(), sizeof(double) * d0.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->bad, 0, sizeof(int));
    if (e == hipSuccess) e = hipMemset(s->hist, 0, esz * (size_t)s->slots * s->ldh);
    if (e != hipSuccess) {
        mt_free(s);
        return check_hip(ctx, e, "mt1d setup copies");
    }
    if (is_ab) {
        const int rc = ab_setup(ctx, s, d);
        if (rc) {
            mt_free(s);
            return rc;
        }
    }
    if (const int rc = mt_tune_block(ctx, s, d->hist_block)) {
        mt_free(s);
        return rc;
    }
    *out = s;
    return FRM_OK;
}

int frm_mt1d_destroy(frm_mt1d* s) {
    if (!s) return FRM_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    mt_free(s);
    return FRM_OK;
}

int frm_mt1d_set_state(frm_mt1d* s, const double* h_U) {
    if (!s || !h_U) return set_error(s ? s->ctx : nullptr, FRM_E_INVALID, "frm_mt1d_set_state: null");
    frm_ctx* ctx = s->ctx;
    FRM_HIP(ctx, hipMemcpyAsync(s->U, h_U, sizeof(double) * s->nU, hipMemcpyHostToDevice, ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->tstep = 1;
    s->depth = 0;
    s->wahead = 0;  // weights formed ahead were for the old run (done: the stream waited on them)
    if (s->psi_qcount) {
        FRM_HIP(ctx, hipMemsetAsync(s->psi_qcount, 0, 2 * sizeof(int), ctx->stream));
        FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (s->ab_err) {
        FRM_HIP(ctx, hipMemsetAsync(s->ab_err, 0, sizeof(double) * s->nprob, ctx->stream));
        FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    s->tt = 0.0;
    if (s->fit_err) {
        FRM_HIP(ctx, hipMemsetAsync(s->fit_err, 0, sizeof(double) * s->nprob, ctx->stream));
        FRM_HIP(ctx, hipMemsetAsync(s->fit_dead, 0, sizeof(int) * s->nprob, ctx->stream));
        FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (s->ser_lc) {
        const size_t nl = (size_t)s->ser_maxL * s->ldh;
        FRM_HIP(ctx, hipMemsetAsync(s->ser_lc, 0, sizeof(double) * 2 * nl, ctx->stream));
        FRM_HIP(ctx, hipMemsetAsync(s->ser_flag, 0, nl, ctx->stream));
        FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return FRM_OK;
}

int frm_mt1d_step(frm_mt1d* s, int32_t nsteps) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_mt1d_step: null solver");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, nsteps >= 0, "frm_mt1d_step: nsteps %d", nsteps);
    if (s->tstep - 1 + nsteps > s->capacity)
        return set_error(ctx, FRM_E_RANGE, "frm_mt1d_step: %d steps exceed capacity %d",
                         s->tstep - 1 + nsteps, s->capacity);
    const size_t esz = s->precision == FRM_PREC_F32C ? sizeof(float) : sizeof(double);
    // one step: the launches of step s->tstep, queued on the context's stream
    auto step_once = [&]() -> int {
        const int k = s->tstep;
        const int depth = s->depth;  // = max(0, k - 3)
        FitArgs fa{};
        const int fcnt = s->nobs ? s->fit_off[k] - s->fit_off[k - 1] : 0;
        if (fcnt > 0) {
            fa.pr = s->probs;
            fa.nprob = s->nprob;
            fa.U = s->U;
            fa.T = s->fit_obs;
            fa.Z = s->fit_obs + s->nobs;
            fa.F = s->fit_obs + 2 * (size_t)s->nobs;
            fa.idx = s->fit_idx + s->fit_off[k - 1];
            fa.cnt = fcnt;
            fa.tt = s->tt;
            fa.tau = s->tau;
            fa.vs = s->fit_vs;
            fa.nobs = s->nobs;
            fa.err = s->fit_err;
            fa.dead = s->fit_dead;
            LaunchTimer t(ctx, FRM_K_OTHER, 0.0);
            FRM_HIP(ctx, mt_launch(s, k_fit_before, dim3((unsigned)((s->nprob + 255) / 256)),
                                   dim3(256), fa));
        }
        // the step's fitting error, once U holds the new state
        auto fit_after = [&]() -> int {
            if (fcnt > 0) {
                LaunchTimer t(ctx, FRM_K_OTHER, 0.0);
                FRM_HIP(ctx, mt_launch(s, k_fit_after, dim3((unsigned)((s->nprob + 255) / 256)),
                                       dim3(256), fa));
            }
            s->tt += s->tau;
            return FRM_OK;
        };
        // kind psi: the weights overlap the previous step unless launches are
        // timed one by one (recorded for a graph: a branch of it, which the
        // next step's first launch depends on)
        const bool ahead = s->kind == FRM_MT_PSI && s->wstream && !ctx->timing_mask;
        if (s->kind == FRM_MT_PSI) {
            if (s->wahead == k) {
                if (!s->rec) FRM_HIP(ctx, hipStreamWaitEvent(ctx->stream, s->wready[k & 1], 0));
            } else if (ahead && !s->rec) {
                FRM_HIP(ctx, mt_launch_psi_weights_ahead(s, k, depth));
                FRM_HIP(ctx, hipStreamWaitEvent(ctx->stream, s->wready[k & 1], 0));
            } else {
                LaunchTimer t(ctx, FRM_K_OTHER, 0.0);
                FRM_HIP(ctx, mt_launch_psi_weights(s, k, depth));
            }
            s->wahead = 0;
            // the next step's weights (depth + 1 slots from step 3 on), formed
            // while this step's memory term and sweep run
            if (ahead && k + 1 <= s->capacity) {
                FRM_HIP(ctx, mt_launch_psi_weights_ahead(s, k + 1, k >= 3 ? depth + 1 : depth));
                s->wahead = k + 1;
            }
        }
        int nch = 0;
        if (depth > 0 && s->ser_count < s->nprob) {
            nch = mt_chunks(s, depth, s->min_chunk);
            // algorithmic bytes: the problems' own nodes (padding lanes load nothing)
            LaunchTimer t(ctx, FRM_K_MEMTERM,
                          (double)depth * (s->gprob_mt ? s->hnodes_mt : s->hnodes) * esz);
            FRM_HIP(ctx, mt_launch_memterm(s, k, depth, nch));
        }
        if (s->kind == FRM_MT_AB && s->ser_count > 0 && k >= 2) {
            AbSeriesArgs c;
            c.pr = s->probs;
            c.gprob = s->gprob;
            c.ldh = s->ldh;
            c.tab = s->ser_tab;
            c.choice = s->ser_choice;
            c.F = (const double*)s->hist + (size_t)(depth - 1) * s->ldh;  // step k - 1's increment
            c.lc1 = s->ser_lc;
            c.lc2 = s->ser_lc + (size_t)s->ser_maxL * s->ldh;
            c.lflag = s->ser_flag;
            c.part = s->part;
            c.i = k;
            c.bad = s->bad;
            LaunchTimer t(ctx, FRM_K_MEMTERM, s->ser_bytes.empty() ? 0.0 : s->ser_bytes[k - 2]);
            FRM_HIP(ctx, mt_launch(s, k_ab_series, dim3((unsigned)((s->ldh + 255) / 256)), dim3(256),
                                   c));
        }
        if (s->kind == FRM_MT_AB) {
            {
                LaunchTimer t(ctx, FRM_K_SWEEP1D, 0.0);
                FRM_HIP(ctx, mt_launch_sweep(s, k, depth, nch, true));
            }
            s->depth = depth + 1;
            s->tstep = k + 1;
            if (int rc = fit_after()) return rc;
            return FRM_OK;
        }
        {
            LaunchTimer t(ctx, FRM_K_SWEEP1D, 0.0);
            FRM_HIP(ctx, mt_launch_sweep(s, k, depth, nch, k >= 3));
        }
        // the buffers of step k are free for step k + 2's weights
        if (s->wstream && !s->rec) FRM_HIP(ctx, hipEventRecord(s->wfree[k & 1], ctx->stream));
        if (k >= 3) s->depth = depth + 1;
        s->tstep = k + 1;
        if (int rc = fit_after()) return rc;
        return FRM_OK;
    };
    // Graph mode: the launches of up to kGraphSteps steps are recorded,
    // matched with the graph built for the same kernel sequence (a chain of
    // kernel nodes, instantiated once), the nodes get these steps' arguments
    // and grids (hipGraphExecKernelNodeSetParams) and the graph is launched
    // once.  Runs of steps differ only in arguments and grids (depth,
    // time-chunk count, slots), so a run uses a handful of graphs.  Not while
    // per-launch timing is on.
    if (!s->graph || ctx->timing_mask != 0) {
        for (int it = 0; it < nsteps; ++it)
            if (int rc = step_once()) return rc;
        // a synchronisation of the context's stream also covers the weights
        // formed ahead
        if (s->wahead) FRM_HIP(ctx, hipStreamWaitEvent(ctx->stream, s->wready[s->wahead & 1], 0));
        return FRM_OK;
    }
    std::vector<KRec> recs;
    std::vector<void*> ptrs;
    for (int it = 0; it < nsteps;) {
        recs.clear();
        const int tstep0 = s->tstep, depth0 = s->depth;
        const double tt0 = s->tt;
        // record steps until the run is full; a step that fails is taken
        // back (the solver stays where it was before it) and ends the run
        int done = 0, rc = FRM_OK;
        const int wahead0 = s->wahead;
        s->rec = &recs;
        for (; done < kGraphSteps && it + done < nsteps; ++done) {
            const size_t mark = recs.size();
            const int ts = s->tstep, dp = s->depth, wa = s->wahead;
            const double tt = s->tt;
            s->rec_step = done;
            rc = step_once();
            if (rc) {
                recs.resize(mark);
                s->tstep = ts;
                s->depth = dp;
                s->tt = tt;
                s->wahead = wa;
                break;
            }
        }
        s->rec = nullptr;
        // none of the run's steps is launched: the solver goes back to its start
        auto unrun = [&](int code) {
            s->tstep = tstep0;
            s->depth = depth0;
            s->tt = tt0;
            s->wahead = wahead0;
            return code;
        };
        if (!recs.empty()) {
            StepGraph* g = nullptr;
            for (auto& c : s->graphs) {
                if (c.funcs.size() != recs.size()) continue;
                bool same = true;
                for (size_t i = 0; i < recs.size() && same; ++i)
                    same = c.funcs[i] == recs[i].func && c.tags[i] == recs[i].step * 2 + recs[i].branch;
                if (same) {
                    g = &c;
                    break;
                }
            }
            ptrs.resize(16 * recs.size());
            auto params = [&](size_t i) {
                void** pp = ptrs.data() + 16 * i;
                recs[i].ptrs(pp);
                hipKernelNodeParams q{};
                q.blockDim = recs[i].block;
                q.extra = nullptr;
                q.func = const_cast<void*>(recs[i].func);
                q.gridDim = recs[i].grid;
                q.kernelParams = pp;
                q.sharedMemBytes = 0;
                return q;
            };
            if (!g) {
                // Edges: each step's own launches in order.  The weights
                // formed ahead in step j (for step j + 1, into the buffers
                // step j - 1 read) start after step j - 1's last launch (or
                // after step j's weights when step j formed them itself) and
                // after the weights formed ahead in step j - 1 (the weight
                // kernels share the fallback queue, so they stay in order);
                // step j + 1's first launch waits for them.  Without weights
                // formed ahead the graph is a chain.
                StepGraph ng;
                FRM_HIP(ctx, hipGraphCreate(&ng.graph, 0));
                hipError_t e = hipSuccess;
                hipGraphNode_t main_last = nullptr, br_last = nullptr;
                std::vector<hipGraphNode_t> step_last(done + 1, nullptr), step_br(done + 1, nullptr);
                int cur = -1;
                for (size_t i = 0; i < recs.size() && e == hipSuccess; ++i) {
                    const KRec& r = recs[i];
                    hipGraphNode_t dep[2];
                    size_t nd = 0;
                    if (r.branch) {
                        if (step_br[r.step]) {
                            dep[nd++] = step_br[r.step];
                        } else {
                            if (cur == r.step && main_last) dep[nd++] = main_last;
                            else if (r.step > 0 && step_last[r.step - 1]) dep[nd++] = step_last[r.step - 1];
                            if (br_last) dep[nd++] = br_last;
                        }
                    } else {
                        if (main_last) dep[nd++] = main_last;
                        if (r.step != cur && r.step > 0 && step_br[r.step - 1]) dep[nd++] = step_br[r.step - 1];
                    }
                    hipKernelNodeParams q = params(i);
                    hipGraphNode_t node = nullptr;
                    e = hipGraphAddKernelNode(&node, ng.graph, nd ? dep : nullptr, nd, &q);
                    if (r.branch) {
                        br_last = node;
                        step_br[r.step] = node;
                    } else {
                        main_last = node;
                        step_last[r.step] = node;
                        cur = r.step;
                    }
                    ng.nodes.push_back(node);
                    ng.funcs.push_back(r.func);
                    ng.tags.push_back(r.step * 2 + r.branch);
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&ng.exec, ng.graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    if (ng.exec) (void)hipGraphExecDestroy(ng.exec);
                    (void)hipGraphDestroy(ng.graph);
                    return unrun(check_hip(ctx, e, "frm_mt1d_step: step graph"));
                }
                s->graphs.push_back(ng);
                g = &s->graphs.back();
            } else {
                for (size_t i = 0; i < recs.size(); ++i) {
                    hipKernelNodeParams q = params(i);
                    const hipError_t e = hipGraphExecKernelNodeSetParams(g->exec, g->nodes[i], &q);
                    if (e != hipSuccess) return unrun(check_hip(ctx, e, "frm_mt1d_step: node update"));
                }
            }
            const hipError_t e = hipGraphLaunch(g->exec, ctx->stream);
            if (e != hipSuccess) return unrun(check_hip(ctx, e, "frm_mt1d_step: graph launch"));
        }
        if (rc) return rc;
        it += done;
    }
    // the weight buffers' events, for steps launched eagerly later: every
    // graph so far is ahead of them on the context's stream
    if (s->wstream)
        for (int b = 0; b < 2; ++b) {
            FRM_HIP(ctx, hipEventRecord(s->wfree[b], ctx->stream));
            FRM_HIP(ctx, hipEventRecord(s->wready[b], ctx->stream));
        }
    return FRM_OK;
}

int frm_mt1d_get_state(frm_mt1d* s, double* h_U) {
    if (!s || !h_U) return set_error(s ? s->ctx : nullptr, FRM_E_INVALID, "frm_mt1d_get_state: null");
    frm_ctx* ctx = s->ctx;
    int bad = 0;
    FRM_HIP(ctx, hipMemcpyAsync(h_U, s->U, sizeof(double) * s->nU, hipMemcpyDeviceToHost, ctx->stream));
    FRM_HIP(ctx, hipMemcpyAsync(&bad, s->bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (bad & 1)
        return set_error(ctx, FRM_E_STATE,
                         "frm_mt1d: a psi weight was not finite after the adaptive quadrature "
                         "fallback of calc_int (mt_psi_caputo.cpp:736-737)");
    if (bad & 4)
        return set_error(ctx, FRM_E_STATE,
                         "frm_mt1d: more psi weight series than the fallback queue holds (2^20) "
                         "were not finite in one step");
    if (bad & 2)
        return set_error(ctx, FRM_E_STATE,
                         "frm_mt1d: a term of the AB power series was not finite with ab_maxl "
                         "above 4096, the longest series whose sorted-pyramid recount "
                         "(MF_int_series.cpp:600-609) this path holds in LDS");
    return FRM_OK;
}

int frm_mt1d_ab_error(frm_mt1d* s, double* out) {
    if (!s || !out) return set_error(s ? s->ctx : nullptr, FRM_E_INVALID, "frm_mt1d_ab_error: null");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, s->kind == FRM_MT_AB, "frm_mt1d_ab_error: batch kind %d is not AB", s->kind);
    std::vector<double> e(s->nprob);
    FRM_HIP(ctx, hipMemcpyAsync(e.data(), s->ab_err, sizeof(double) * s->nprob, hipMemcpyDeviceToHost,
                                ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int steps = s->tstep - 1;
    for (int p = 0; p < s->nprob; ++p)
        out[p] = steps > 0 ? e[p] / ((s->hprobs[p].N + 1.0) * steps) : 0.0;
    return FRM_OK;
}

int frm_mt1d_set_observations(frm_mt1d* s, int32_t nobs, const double* T, const double* Z,
                              const double* F) {
    if (!s) return set_error(nullptr, FRM_E_INVALID, "frm_mt1d_set_observations: null solver");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, nobs >= 0 && (nobs == 0 || (T && Z && F)),
                "frm_mt1d_set_observations: bad arguments");
    (void)hipFree(s->fit_obs);
    (void)hipFree(s->fit_idx);
    (void)hipFree(s->fit_vs);
    s->fit_obs = s->fit_vs = nullptr;
    s->fit_idx = nullptr;
    s->nobs = 0;
    s->fit_off.clear();
    // which step each observation falls in: the reference's loop variable
    // tt starts at 0 and grows by tau per step; step k covers [tt, tt + tau)
    std::vector<std::vector<int>> per(s->capacity + 1);
    double tt = 0.0;
    for (int k = 1; k <= s->capacity; ++k) {
        for (int o = 0; o < nobs; ++o)
            if ((T[o] >= tt) && (T[o] < (tt + s->tau))) per[k].push_back(o);
        tt += s->tau;
    }
    std::vector<int> idx;
    s->fit_off.assign(s->capacity + 1, 0);
    for (int k = 1; k <= s->capacity; ++k) {
        idx.insert(idx.end(), per[k].begin(), per[k].end());
        s->fit_off[k] = (int)idx.size();
    }
    hipError_t e = hipSuccess;
    if (!s->fit_err) {
        e = hipMalloc(&s->fit_err, sizeof(double) * s->nprob);
        if (e == hipSuccess) e = hipMalloc(&s->fit_dead, sizeof(int) * s->nprob);
    }
    if (e == hipSuccess) e = hipMemset(s->fit_err, 0, sizeof(double) * s->nprob);
    if (e == hipSuccess) e = hipMemset(s->fit_dead, 0, sizeof(int) * s->nprob);
    if (nobs > 0) {
        std::vector<double> obs((size_t)3 * nobs);
        std::copy(T, T + nobs, obs.begin());
        std::copy(Z, Z + nobs, obs.begin() + nobs);
        std::copy(F, F + nobs, obs.begin() + 2 * (size_t)nobs);
        if (e == hipSuccess) e = hipMalloc(&s->fit_obs, sizeof(double) * obs.size());
        if (e == hipSuccess) e = hipMalloc(&s->fit_idx, sizeof(int) * std::max<size_t>(1, idx.size()));
        if (e == hipSuccess) e = hipMalloc(&s->fit_vs, sizeof(double) * (size_t)nobs * s->nprob);
        if (e == hipSuccess)
            e = hipMemcpy(s->fit_obs, obs.data(), sizeof(double) * obs.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess && !idx.empty())
            e = hipMemcpy(s->fit_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(s->fit_vs, 0, sizeof(double) * (size_t)nobs * s->nprob);
        s->nobs = nobs;
    }
    return e == hipSuccess ? FRM_OK : check_hip(ctx, e, "frm_mt1d_set_observations");
}

int frm_mt1d_fit_error(frm_mt1d* s, double* out) {
    if (!s || !out) return set_error(s ? s->ctx : nullptr, FRM_E_INVALID, "frm_mt1d_fit_error: null");
    frm_ctx* ctx = s->ctx;
    FRM_REQUIRE(ctx, s->fit_err != nullptr, "frm_mt1d_fit_error: no observations were set");
    std::vector<double> e(s->nprob);
    FRM_HIP(ctx, hipMemcpyAsync(e.data(), s->fit_err, sizeof(double) * s->nprob, hipMemcpyDeviceToHost,
                                ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < s->nprob; ++p) out[p] = std::isfinite(e[p]) ? e[p] : 1e300;
    return FRM_OK;
}

int frm_mt1d_hist_block(const frm_mt1d* s, int32_t* len) {
    if (!s || !len) return set_error(nullptr, FRM_E_INVALID, "frm_mt1d_hist_block: null");
    *len = s->min_chunk;
    return FRM_OK;
}

int frm_mt1d_tstep(const frm_mt1d* s, int32_t* tstep) {
    if (!s || !tstep) return set_error(nullptr, FRM_E_INVALID, "frm_mt1d_tstep: null");
    *tstep = s->tstep;
    return FRM_OK;
}

}  // extern "C"

extern "C" int frm_diag_psi_kb_quad(frm_ctx* ctx, int32_t n, const double* h_cases,
                                    double* h_out) {
    FRM_REQUIRE(ctx, ctx && h_cases && h_out && n > 0 && n <= (1 << 20),
                "frm_diag_psi_kb_quad: bad arguments (1 <= n <= 2^20)");
    for (int32_t i = 0; i < n; ++i) {
        const double* r = h_cases + 6 * (int64_t)i;
        FRM_REQUIRE(ctx, std::isfinite(r[2]) && std::isfinite(r[3]) && std::isfinite(r[4]) &&
                             r[1] > 0.0 && r[0] >= 0.0 && r[0] < 1.0 && r[2] <= r[3],
                    "frm_diag_psi_kb_quad: case %d: need 0 <= alpha < 1, beta > 0, finite "
                    "t0 <= t1 and gtj", i);
    }
    double *dc = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&dc, sizeof(double) * 6 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&dout, sizeof(double) * (size_t)n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dc, h_cases, sizeof(double) * 6 * (size_t)n, hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(frm::k_diag_psi_kb_quad, dim3((unsigned)((n + 63) / 64)), dim3(64), 0,
                           ctx->stream, dc, n, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_out, dout, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dc);
    (void)hipFree(dout);
    return e == hipSuccess ? FRM_OK : check_hip(ctx, e, "frm_diag_psi_kb_quad");
}
