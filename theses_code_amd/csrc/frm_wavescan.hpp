// Wavefront-level building blocks for tridiagonal (Thomas) sweeps.
//
// A line of n unknowns is split into 64 contiguous chunks of CM elements,
// one chunk per lane.  The Thomas forward and back-substitution recurrences
// are first-order affine maps, so each lane composes the maps of its chunk,
// a 6-step Kogge-Stone scan over the wavefront gives every lane the value
// entering its chunk, and the lane then replays the exact sequential Thomas
// formulas over its chunk (method: Stone 1973 recurrence scan; the in-chunk
// arithmetic is the reference's progonka, mt_minimal.cpp:942-1024).
#pragma once
#include <hip/hip_runtime.h>

namespace frm {

// Exclusive composition of affine maps x -> P x + Q over lanes 0..lane-1
// (forward: earlier lanes first).
__device__ inline void scan_affine_up(double& P, double& Q, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Po = __shfl_up(P, d);
        const double Qo = __shfl_up(Q, d);
        if (lane >= d) {
            Q = P * Qo + Q;
            P = P * Po;
        }
    }
    const double Pe = __shfl_up(P, 1);
    const double Qe = __shfl_up(Q, 1);
    P = lane == 0 ? 1.0 : Pe;
    Q = lane == 0 ? 0.0 : Qe;
}

// Exclusive composition over lanes lane+1..63 (backward: later lanes first).
__device__ inline void scan_affine_down(double& P, double& Q, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Po = __shfl_down(P, d);
        const double Qo = __shfl_down(Q, d);
        if (lane + d < 64) {
            Q = P * Qo + Q;
            P = P * Po;
        }
    }
    const double Pe = __shfl_down(P, 1);
    const double Qe = __shfl_down(Q, 1);
    P = lane == 63 ? 1.0 : Pe;
    Q = lane == 63 ? 0.0 : Qe;
}

// Both scans with the wavefront's total map as well (lane 63's inclusive
// forward composition, lane 0's inclusive backward composition), for lines
// that span several wavefronts.
__device__ inline void scan_affine_up_total(double& P, double& Q, int lane, double& Pt, double& Qt) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Po = __shfl_up(P, d);
        const double Qo = __shfl_up(Q, d);
        if (lane >= d) {
            Q = P * Qo + Q;
            P = P * Po;
        }
    }
    Pt = __shfl(P, 63);
    Qt = __shfl(Q, 63);
    const double Pe = __shfl_up(P, 1);
    const double Qe = __shfl_up(Q, 1);
    P = lane == 0 ? 1.0 : Pe;
    Q = lane == 0 ? 0.0 : Qe;
}

__device__ inline void scan_affine_down_total(double& P, double& Q, int lane, double& Pt,
                                              double& Qt) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double Po = __shfl_down(P, d);
        const double Qo = __shfl_down(Q, d);
        if (lane + d < 64) {
            Q = P * Qo + Q;
            P = P * Po;
        }
    }
    Pt = __shfl(P, 0);
    Qt = __shfl(Q, 0);
    const double Pe = __shfl_down(P, 1);
    const double Qe = __shfl_down(Q, 1);
    P = lane == 63 ? 1.0 : Pe;
    Q = lane == 63 ? 0.0 : Qe;
}

// Scratch in LDS for one line that spans W wavefronts.
template <int W>
struct LineShare {
    double fP[W], fQ[W];  // forward wave totals
    double bP[W], bQ[W];  // backward wave totals
    double un;
};

// line_solve for a line spread over W wavefronts of one workgroup: lane
// `lane` of wavefront `wl` (0..W-1) owns elements e0 = (64 wl + lane) CM ..
// + CM - 1.  The wavefront totals meet in `sh`; every wavefront of the
// workgroup must call this (it contains workgroup barriers), `active`
// false for wavefronts without a line.
template <int CM, int W>
__device__ inline double line_solve_mw(const double (&a)[CM], const double (&f)[CM],
                                       const double (&al)[CM], double (&v)[CM], int e0, int n,
                                       double b0, int lane, int wl, LineShare<W>* sh, bool active,
                                       const double* un_fixed = nullptr) {
    // ---- forward
    double P = 1.0, Q = 0.0, Pt = 1.0, Qt = 0.0;
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        if (e0 + k < n) {
            const double p = f[k] * a[k];
            const double q = -f[k] * v[k];
            Q = p * Q + q;
            P = p * P;
        }
    }
    if (active) {
        scan_affine_up_total(P, Q, lane, Pt, Qt);
        if (lane == 0) {
            sh->fP[wl] = Pt;
            sh->fQ[wl] = Qt;
        }
    }
    __syncthreads();
    double x = 0.0;
    if (active) {
        double y = b0;  // value entering this wavefront's part of the line
        for (int w = 0; w < wl; ++w) y = sh->fP[w] * y + sh->fQ[w];
        x = P * y + Q;
#pragma unroll
        for (int k = 0; k < CM; ++k) {
            if (e0 + k < n) {
                x = f[k] * (a[k] * x - v[k]);
                v[k] = x;
            }
        }
#pragma unroll
        for (int k = 0; k < CM; ++k)
            if (e0 + k == n - 1) sh->un = un_fixed ? *un_fixed : v[k] / (1.0 - al[k]);
    }
    __syncthreads();
    const double un = active ? sh->un : 0.0;
    // ---- backward
    P = 1.0;
    Q = 0.0;
#pragma unroll
    for (int k = CM - 1; k >= 0; --k) {
        if (e0 + k < n) {
            Q = al[k] * Q + v[k];
            P = al[k] * P;
        }
    }
    if (active) {
        scan_affine_down_total(P, Q, lane, Pt, Qt);
        if (lane == 0) {
            sh->bP[wl] = Pt;
            sh->bQ[wl] = Qt;
        }
    }
    __syncthreads();
    if (active) {
        double y = un;  // value just above this wavefront's part of the line
        for (int w = W - 1; w > wl; --w) y = sh->bP[w] * y + sh->bQ[w];
        y = P * y + Q;
#pragma unroll
        for (int k = CM - 1; k >= 0; --k) {
            if (e0 + k < n) {
                y = al[k] * y + v[k];
                v[k] = y;
            }
        }
    }
    return un;
}

// Solve one line whose Thomas alpha coefficients are known.
//   forward : bt_e = f_e * (a_e * bt_{e-1} - om_e),  bt_{-1} = b0
//   backward: u_e  = al_e * u_{e+1} + bt_e,           u_n = bt_{n-1} / (1 - al_{n-1})
// (fr_d_t_g_2d.cpp:3494-3553 and 2548-2550 with f_e = Al[i+1]/B(i)).
// On entry v[] holds om for the lane's chunk; on exit it holds u.
// Returns u_n (the value at i = N) to every lane; with un_fixed the caller
// gives u_n (a Dirichlet end) instead.
template <int CM>
__device__ inline double line_solve(const double (&a)[CM], const double (&f)[CM],
                                    const double (&al)[CM], double (&v)[CM], int e0, int n,
                                    double b0, int lane, const double* un_fixed = nullptr) {
    // ---- forward: compose chunk maps x -> p x + q, p = f a, q = -f om
    double P = 1.0, Q = 0.0;
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        if (e0 + k < n) {
            const double p = f[k] * a[k];
            const double q = -f[k] * v[k];
            Q = p * Q + q;
            P = p * P;
        }
    }
    scan_affine_up(P, Q, lane);
    double x = P * b0 + Q;  // bt entering the chunk
#pragma unroll
    for (int k = 0; k < CM; ++k) {
        if (e0 + k < n) {
            x = f[k] * (a[k] * x - v[k]);
            v[k] = x;  // v now holds bt
        }
    }
    // ---- u_n from the lane holding element n-1
    const int last = (n - 1) / CM;
    double un_local = 0.0;
#pragma unroll
    for (int k = 0; k < CM; ++k)
        if (e0 + k == n - 1) un_local = v[k] / (1.0 - al[k]);
    const double un = un_fixed ? *un_fixed : __shfl(un_local, last);
    // ---- backward: compose y -> al y + bt from the top of the chunk down
    P = 1.0;
    Q = 0.0;
#pragma unroll
    for (int k = CM - 1; k >= 0; --k) {
        if (e0 + k < n) {
            Q = al[k] * Q + v[k];
            P = al[k] * P;
        }
    }
    scan_affine_down(P, Q, lane);
    double y = P * un + Q;  // u just above the chunk
#pragma unroll
    for (int k = CM - 1; k >= 0; --k) {
        if (e0 + k < n) {
            y = al[k] * y + v[k];
            v[k] = y;  // v now holds u
        }
    }
    return un;
}

}  // namespace frm

namespace frm {

// 2x2 matrix [[m00, m01], [m10, m11]] acting on homogeneous (y, z); used for
// the Thomas alpha recurrence c'_k = c_k / (b_k - a_k c'_{k-1}), a Moebius
// map.  Products are rescaled by an exact power of two to stay in range.
struct Mob {
    double m00, m01, m10, m11;
};

__device__ inline Mob mob_mul(const Mob& x, const Mob& y) {  // x * y (apply y first)
    Mob r;
    r.m00 = x.m00 * y.m00 + x.m01 * y.m10;
    r.m01 = x.m00 * y.m01 + x.m01 * y.m11;
    r.m10 = x.m10 * y.m00 + x.m11 * y.m10;
    r.m11 = x.m10 * y.m01 + x.m11 * y.m11;
    const double mx = fmax(fmax(fabs(r.m00), fabs(r.m01)), fmax(fabs(r.m10), fabs(r.m11)));
    if (mx > 0.0 && isfinite(mx)) {
        int ex;
        frexp(mx, &ex);
        r.m00 = ldexp(r.m00, -ex);
        r.m01 = ldexp(r.m01, -ex);
        r.m10 = ldexp(r.m10, -ex);
        r.m11 = ldexp(r.m11, -ex);
    }
    return r;
}

__device__ inline Mob shfl_up_mob(const Mob& m, int d) {
    return Mob{__shfl_up(m.m00, d), __shfl_up(m.m01, d), __shfl_up(m.m10, d), __shfl_up(m.m11, d)};
}

// Exclusive forward scan of Moebius maps (earlier lanes applied first).
__device__ inline Mob scan_mob_up(Mob m, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Mob o = shfl_up_mob(m, d);
        if (lane >= d) m = mob_mul(m, o);
    }
    Mob e = shfl_up_mob(m, 1);
    if (lane == 0) e = Mob{1.0, 0.0, 0.0, 1.0};
    return e;
}

}  // namespace frm
