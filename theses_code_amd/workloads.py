"""Seeded synthetic inputs for the benchmark configurations (splitmix64-seeded
xoshiro256**, rng.py, seed 0x5EED0000 + configuration index).

2D Caputo convection-diffusion on [0,1]^2 (the fr_d_t_g_2d LOD form): a grid
of sN = M = n+1 intervals gives n x n interior unknowns per half-step (n row
systems of n unknowns, then n column systems), nodes i = 0..sN+1 at x = i*h,
h = 1/sN.  The velocity is the curl of a Gaussian random stream function
(FFT-filtered white noise, Gaussian spectrum, correlation length 0.1) scaled
to max|v|; the initial field is a sum of 8 Gaussian blobs.
"""
from __future__ import annotations

import numpy as np

from .rng import Xoshiro
from .solvers import coefs_advection

SEED_BASE = 0x5EED0000


def gaussian_random_field(n: int, corr: float, h: float, rng) -> np.ndarray:
    """Periodic Gaussian random field on an n x n grid of spacing h."""
    noise = rng.normal(size=(n, n))
    k = np.fft.fftfreq(n, d=h) * 2 * np.pi
    kx, ky = np.meshgrid(k, k, indexing="xy")
    spec = np.exp(-0.25 * (kx ** 2 + ky ** 2) * corr ** 2)
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * spec))
    return f / np.std(f)


def lod2d_case(n_interior: int, seed: int, vmax: float = 0.1, D: float = 1e-3,
               blobs: int = 8, blob_sigma: float = 0.05):
    """Returns (sN, M, coefs, C0) for an n_interior^2 LOD problem."""
    rng = Xoshiro(seed)
    sN = M = n_interior + 1
    h = 1.0 / sN
    ng = sN + 2  # nodes 0..sN+1 in both directions, layout [j][i]
    psi = gaussian_random_field(ng, 0.1, h, rng)
    # v = curl psi = (d psi/dy, -d psi/dx), central differences (periodic)
    vx = (np.roll(psi, -1, axis=0) - np.roll(psi, 1, axis=0)) / (2 * h)
    vy = -(np.roll(psi, -1, axis=1) - np.roll(psi, 1, axis=1)) / (2 * h)
    s = vmax / np.max(np.hypot(vx, vy))
    vx, vy = vx * s, vy * s
    x = np.arange(ng) * h
    X, Y = np.meshgrid(x, x, indexing="xy")  # X[j, i] = x_i, Y[j, i] = y_j
    C0 = np.zeros((ng, ng))
    amps = rng.uniform(0, 1, blobs)
    ctr = rng.uniform(0, 1, (blobs, 2))
    for a, (cx, cy) in zip(amps, ctr):
        C0 += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * blob_sigma ** 2))
    coefs = coefs_advection(vx, vy, D, h)
    return sN, M, coefs, C0


# the 2D Caputo convection-diffusion workload quoted by the headline metric
C1_WORKLOAD = dict(name="lod2d_caputo_512x512_fp64", n_interior=512, alpha=0.7, tau=1e-3,
                   steps=5000, sigma=1.0, D=1e-3, vmax=0.1, seed=SEED_BASE + 1)


# 1D Caputo diffusion in the mt_minimal form: N = 200 intervals on [0,1],
# Dirichlet 0 at both ends, u0 = sin(pi x) + 0.01 N(0,1) zeroed at the ends
C0_WORKLOAD = dict(name="mt1d_caputo_200_fp64", N=200, alpha=0.8, tau=1e-3, steps=2000, D=1e-3,
                   seed=SEED_BASE + 0)


def mt1d_c0_case(seed: int = C0_WORKLOAD["seed"], N: int = 200):
    rng = Xoshiro(seed)
    x = np.arange(N + 2) / N
    u0 = np.sin(np.pi * x) + 0.01 * rng.normal(size=N + 2)
    u0[0] = 0.0
    u0[N] = 0.0
    u0[N + 1] = 0.0
    prob = dict(N=N, L=1.0, alpha=C0_WORKLOAD["alpha"], form=0, D=C0_WORKLOAD["D"], top=1,
                top_val=0.0, bottom_cond=1, Hbottom=0.0)
    return prob, u0


# 4096 independent psi-Caputo moisture problems (mt_psi_caputo form) with
# nonlinear diffusivity k(u) = k0 exp(b u); per-problem parameters drawn
# from seeded sub-streams.  Domain [0,1], wet Dirichlet top u = 0.4, flux
# (zero-gradient) bottom.
C3_WORKLOAD = dict(name="mt1d_psi_batch4096_kexp", nprob=4096, Nmin=64, Nmax=1024, tau=1e-3,
                   steps=3000, picard=3, seed=SEED_BASE + 3)


def mt1d_c3_case(nprob: int = 4096, seed: int = C3_WORKLOAD["seed"], Nmin: int = 64,
                 Nmax: int = 1024):
    probs, U0 = [], []
    for r in Xoshiro(seed).spawn(nprob):
        n = int(r.integers(Nmin, Nmax + 1))
        probs.append(dict(N=n, L=1.0, alpha=float(r.uniform(0.5, 0.99)),
                          beta=float(r.uniform(0.5, 1.5)), form=2, k0=float(r.uniform(1e-4, 1e-2)),
                          b=float(r.uniform(0.0, 3.0)), top=1, top_val=0.4, bottom_cond=0))
        u = r.uniform(0.1, 0.4, n + 2)
        u[0] = 0.4
        u[n + 1] = u[n]
        U0.append(u)
    return probs, U0


# 2D Caputo model at 1024 x 1024, alpha 0.6, 4000 steps (precision study)
C2_WORKLOAD = dict(name="lod2d_caputo_1024x1024", n_interior=1024, alpha=0.6, tau=1e-3,
                   steps=4000, sigma=1.0, D=1e-3, vmax=0.1, seed=SEED_BASE + 2, soe_terms=32)

# 16 Atangana-Baleanu diffusion parameter sets (LOD form), sides 1024..2048,
# 2000 steps, Gaussian random initial fields (correlation length 0.1)
C4_WORKLOAD = dict(name="lod2d_ab_16sets", nsets=16, side_min=1024, side_max=2048, tau=1e-3,
                   steps=2000, sigma=1.0, D=1e-3, seed=SEED_BASE + 4, soe_terms=64)


def c4_sets(nsets: int = 16, seed: int = C4_WORKLOAD["seed"], side_min: int = 1024,
            side_max: int = 2048):
    """Per-set (alpha, n_interior, seed) drawn from seeded sub-streams."""
    out = []
    for r in Xoshiro(seed).spawn(nsets):
        out.append(dict(alpha=float(r.uniform(0.3, 0.9)),
                        n_interior=int(r.integers(side_min, side_max + 1)),
                        seed=int(r.integers(0, 2 ** 31))))
    return out


def ab_diffusion_case(n_interior: int, seed: int, D: float = 1e-3):
    """Pure diffusion (no velocity) with a Gaussian random initial field."""
    rng = Xoshiro(seed)
    sN = M = n_interior + 1
    h = 1.0 / sN
    ng = sN + 2
    C0 = gaussian_random_field(ng, 0.1, h, rng)
    C0 = 0.5 + 0.1 * C0
    zero = np.zeros((ng, ng))
    return sN, M, coefs_advection(zero, zero, D, h), C0


# Batch of 1D Atangana-Baleanu test problems (solve_and_check form): 1024
# problems, N uniform in [64, 1024] intervals, orders from 16 values in
# [0.3, 0.9], manufactured solutions F_func uniform in {0, 1, 2}, tau = 1/2000,
# 2000 steps (T = 1); MF_int_choice weights with 50 series terms.
AB1D_WORKLOAD = dict(name="ab1d_batch1024", nprob=1024, Nmin=64, Nmax=1024, tau=1.0 / 2000,
                     steps=2000, maxl=50, maxkiter=1000, eps=1e-10, seed=SEED_BASE + 5)


# The series history of solve_and_check (simple = 0) at the shape of the
# reference's convection_diffusion_test_time (MF_int_series.cpp:1128-1146:
# 6000 steps on [0, 1], 100 intervals, 50 series terms, orders 0.9 .. 0.1),
# batched: every problem one of those nine orders and one of the three
# manufactured solutions.
AB1D_SERIES_WORKLOAD = dict(name="ab1d_series_batch1024", nprob=1024, N=100, tau=1.0 / 6000,
                            steps=6000, maxl=50, maxkiter=1000, eps=1e-10, seed=SEED_BASE + 6)


def ab1d_series_case(nprob: int = AB1D_SERIES_WORKLOAD["nprob"],
                     seed: int = AB1D_SERIES_WORKLOAD["seed"]):
    rng = Xoshiro(seed)
    alphas = np.round(np.arange(0.9, 0.05, -0.1), 1)
    cfg = AB1D_SERIES_WORKLOAD
    return [dict(N=cfg["N"], alpha=float(alphas[q % alphas.size]), src=int(rng.integers(0, 3)),
                 ab_maxl=cfg["maxl"], ab_maxkiter=cfg["maxkiter"], ab_eps=cfg["eps"], ab_series=1)
            for q in range(nprob)]


def ab1d_case(nprob: int = AB1D_WORKLOAD["nprob"], seed: int = AB1D_WORKLOAD["seed"]):
    rng = Xoshiro(seed)
    alphas = np.linspace(0.3, 0.9, 16)
    probs = []
    for _ in range(nprob):
        probs.append(dict(N=int(rng.integers(AB1D_WORKLOAD["Nmin"], AB1D_WORKLOAD["Nmax"] + 1)),
                          alpha=float(alphas[rng.integers(0, 16)]), src=int(rng.integers(0, 3)),
                          ab_maxl=AB1D_WORKLOAD["maxl"], ab_maxkiter=AB1D_WORKLOAD["maxkiter"],
                          ab_eps=AB1D_WORKLOAD["eps"]))
    return probs


# Time sums of the 3D irrigation solver (irrigation_fractional.cpp, the Sum
# family): a (sN+2)(M+2)(K+2) = 128 x 128 x 64 node grid, its default step of
# 50 s (tau_m = 1000/20, 9569), g(t) = t, the full-history sum (sum_alg 0)
# over 2000 steps with the exact Caputo interval weights as kb_3_cache, and
# the series form (sum_alg 2) with 16 moments per node.
TIMESUM3D_WORKLOAD = dict(name="timesum3d_128x128x64", sN=126, M=126, K=62, tau=50.0,
                          alpha=0.8, steps=2000, series_terms=16, pool=16,
                          seed=SEED_BASE + 7)


def timesum3d_levels(n: int, count: int, seed: int = TIMESUM3D_WORKLOAD["seed"]):
    """`count` distinct synthetic time levels of n nodes (smooth in time)."""
    rng = Xoshiro(seed)
    base = rng.uniform(-1, 1, n)
    rate = rng.uniform(-1, 1, n)
    return [base + rate * np.sqrt(k + 1.0) for k in range(count)]


def caputo_interval_weights(T: np.ndarray, T_end: float, alpha: float) -> np.ndarray:
    """kb(alpha, t, t+1, n+1) for g(t) = t without the 1/Gamma(1-alpha)
    factor: the integral of (T_end - x)^-alpha over [T[t], T[t+1]]."""
    return ((T_end - T[:-1]) ** (1 - alpha) - (T_end - T[1:]) ** (1 - alpha)) / (1 - alpha)

