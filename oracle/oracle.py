"""ctypes wrapper of the CPU oracle (oracle/frm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product package never
imports it.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_D = ct.c_double
_I = ct.c_int
_P = ct.POINTER(ct.c_double)
_V = ct.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        L.orc_gamma.restype = _D
        L.orc_gamma.argtypes = [_D]
        L.orc_log_gamma.restype = _D
        L.orc_log_gamma.argtypes = [_D]
        L.orc_tg_kb.restype = _D
        L.orc_tg_kb.argtypes = [_I, _D, _D, _D, _D, _I, _I, _D, _D, _D]
        L.orc_tg_create.restype = _V
        L.orc_tg_create.argtypes = [_I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _P, _P, _P, _P, _P]
        L.orc_tg_destroy.argtypes = [_V]
        L.orc_tg_step.argtypes = [_V]
        L.orc_tg_get.argtypes = [_V, _P]
        L.orc_tg_prefill.argtypes = [_V, _I, ct.c_uint64]
        L.orc_tg_set_ab.argtypes = [_V, _I, _I, _D]
        L.orc_tg_set_dt3.argtypes = [_V, _I]
        L.orc_tg_set_dt2.argtypes = [_V, _D]
        L.orc_tg_set_dt4.argtypes = [_V]
        L.orc_dt4_series.argtypes = [_D, _I, _P, _I, _P]
        L.orc_tg_coefs_ref.argtypes = [_I, _I, _P, _D, _D, _D, _P, _P, _P, _P]
        L.orc_tg_ref_spacing.argtypes = [_I, _I, _P, _P]
        L.orc_mt_create.restype = _V
        L.orc_mt_create.argtypes = [_I, _D, _I, _D, _D, _D, _I, _D, _D, _D, _D, _I, _D, _D,
                                    _I, _D, _I, _P]
        L.orc_mt_destroy.argtypes = [_V]
        L.orc_mt_step.argtypes = [_V]
        L.orc_mt_get.argtypes = [_V, _P]
        L.orc_mt_da.restype = _D
        L.orc_mt_da.argtypes = [_V]
        L.orc_mt_history.argtypes = [_I, _I, _D, _D, _D, _I, _I, _P, _P]
        L.orc_psi_weights.argtypes = [_D, _D, _D, _I, _P]
        L.orc_psi_kb_quad.restype = _D
        L.orc_psi_kb_quad.argtypes = [_D, _D, _D, _D, _D, _I]
        if hasattr(L, "orc_mittag_leffler"):
            L.orc_mittag_leffler.restype = _D
            L.orc_mittag_leffler.argtypes = [_D, _D, _D, _D]
            L.orc_ab_weights.argtypes = [_D, _D, _I, _I, _D, _I, _P]
            L.orc_ab_solve.restype = _D
            L.orc_ab_solve.argtypes = [_D, _D, _I, _I, _I, _I, _D, _I, _P]
            L.orc_ab_history.argtypes = [_D, _D, _I, _I, _D, _I, _P, _P]
            L.orc_ab_int.restype = _D
            L.orc_ab_int.argtypes = [_D, _D, _D, _D, _I, _I, _D]
            L.orc_ab_series_history.restype = ct.c_long
            L.orc_ab_series_history.argtypes = [_D, _D, _I, _I, _D, _I, _P, _P, _P]
            L.orc_ab_solve_series.restype = _D
            L.orc_ab_solve_series.argtypes = [_D, _D, _I, _I, _I, _I, _D, _I, _P]
        L.orc_timesum01.argtypes = [_I, _D, ct.c_int64, _I, _P, _P, _P, _P]
        L.orc_timesum2.argtypes = [ct.c_int64, _I, _P, _P, _P, _I, _D, _D, _D, _D, _P]
        if hasattr(L, "orc_set_threads"):
            L.orc_set_threads.argtypes = [_I]
            L.orc_get_threads.restype = _I
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def gamma(x: float) -> float:
    return lib().orc_gamma(x)


def tg_kb(func, fpow, alpha, tau, n, s, ja, sa, l, int_err=1e-6) -> float:
    return lib().orc_tg_kb(func, fpow, alpha, tau, int_err, n, s, ja, sa, l)


def dt4_series(alpha, direction, F):
    """calc_va's approximated sums of one node for steps 1..len(F) (the
    dt_alg 4 history, unscaled)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    V = np.empty(F.size)
    lib().orc_dt4_series(alpha, direction, _p(F), F.size, _p(V))
    return V


def tg_kb_table(func, fpow, alpha, tau, n, int_err=1e-6) -> np.ndarray:
    """The eight kb(s, ja, sa, l) variants for s = 0..n, variant bit0: ja=1,
    bit1: sa=1/2, bit2: l=1 (fr_d_t_g_2d.cpp:3077-3080)."""
    out = np.empty((8, n + 1))
    for v in range(8):
        ja = 1.0 if v & 1 else 0.5
        sa = 0.5 if v & 2 else 0.0
        l = 1.0 if v & 4 else 0.5
        for s in range(n + 1):
            out[v, s] = tg_kb(func, fpow, alpha, tau, n, s, ja, sa, l, int_err)
    return out


def tg_ref_spacing(sN: int, M: int):
    a, b = ct.c_double(), ct.c_double()
    lib().orc_tg_ref_spacing(sN, M, ct.byref(a), ct.byref(b))
    return a.value, b.value


def tg_coefs_ref(sN, M, V2, dv, dL, dLx):
    nn = (sN + 2) * (M + 2)
    V2 = np.ascontiguousarray(V2, dtype=np.float64).reshape(nn)
    outs = [np.empty(nn) for _ in range(4)]
    lib().orc_tg_coefs_ref(sN, M, _p(V2), dv, dL, dLx, *[_p(o) for o in outs])
    return outs


class TGSolver:
    """Oracle of TG_solver::calc_step (fr_d_t_g_2d.cpp:2552), dt_alg = 0, dt_alg
    = 2 with the cut-off `dt2`, dt_alg = 3 with `dt3` series terms, or dt_alg
    = 4 (the ia/ibk approximations) with `dt4`."""

    def __init__(self, sN, M, alpha, tau, sigma, C1, S, rowA, rowB, colA, colB, C0,
                 func=0, fpow=1.0, int_err=1e-6, ab=None, dt3=0, dt2=0.0, dt4=False):
        self.sN, self.M = sN, M
        nn = (sN + 2) * (M + 2)
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(nn)
                for a in (rowA, rowB, colA, colB, C0)]
        self._h = lib().orc_tg_create(sN, M, func, fpow, alpha, tau, sigma, C1, S, int_err,
                                      *[_p(a) for a in arrs])
        if ab is not None:  # (maxl, maxkiter, eps): Atangana-Baleanu kernel
            lib().orc_tg_set_ab(self._h, *ab)
        if dt3:
            lib().orc_tg_set_dt3(self._h, int(dt3))
        if dt2:
            lib().orc_tg_set_dt2(self._h, float(dt2))
        if dt4:
            lib().orc_tg_set_dt4(self._h)

    def step(self, n=1):
        for _ in range(n):
            lib().orc_tg_step(self._h)

    def prefill(self, depth: int, seed: int = 1):
        lib().orc_tg_prefill(self._h, depth, seed)

    def state(self) -> np.ndarray:
        out = np.empty((self.M + 2) * (self.sN + 2))
        lib().orc_tg_get(self._h, _p(out))
        return out.reshape(self.M + 2, self.sN + 2)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_tg_destroy(self._h)
            self._h = None


FORM_CONST, FORM_CV2, FORM_KEXP = 0, 1, 2
TOP_AUTO, TOP_DIRICHLET, TOP_FLUX = 0, 1, 2


class MTSolver:
    """Oracle of H_solver::calc_step (mt_minimal.cpp:1032 / mt_psi_caputo.cpp:1435)."""

    def __init__(self, N, L, alpha, tau, U0, psi=0, beta=1.0, form=FORM_CONST, D=0.0, k=0.0,
                 k0=0.0, b=0.0, top=TOP_AUTO, top_val=0.0, Ev=0.0, bottom_cond=1,
                 Hbottom=0.0, picard=1):
        self.N = N
        U0 = np.ascontiguousarray(U0, dtype=np.float64)
        assert U0.shape == (N + 2,)
        self._h = lib().orc_mt_create(N, L, psi, beta, alpha, tau, form, D, k, k0, b, top,
                                      top_val, Ev, bottom_cond, Hbottom, picard, _p(U0))

    def step(self, n=1):
        for _ in range(n):
            lib().orc_mt_step(self._h)

    def state(self) -> np.ndarray:
        out = np.empty(self.N + 2)
        lib().orc_mt_get(self._h, _p(out))
        return out

    @property
    def Da(self) -> float:
        return lib().orc_mt_da(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mt_destroy(self._h)
            self._h = None


def mt_fit_error(solver, N, L, tau, nsteps, T, Z, F) -> float:
    """The error loop of solve_and_test (mt_minimal.cpp:1288-1330) around an
    oracle solver, step by step: observations with tt <= T < tt + tau are
    interpolated in depth before (nodes j < N-1) and after (j < N) the step,
    blended in time, clamped to <= 0 and scored by squares; 1e300 once U[0]
    is not finite.  Plain Python floats: IEEE double, no fused operations.
    (Where the reference's `vs` is never written -- a depth in the last
    interval -- it reads uninitialised memory; here it reads 0.)"""
    dL = L / N
    vs = [0.0] * len(T)
    err, tt = 0.0, 0.0
    for _ in range(nsteps):
        U = solver.state()
        for i in range(len(T)):
            if (T[i] >= tt) and (T[i] < (tt + tau)):
                for j in range(N - 1):
                    if ((dL * j) < Z[i]) and ((dL * (j + 1)) >= Z[i]):
                        k2 = (Z[i] - (dL * j)) / dL
                        vs[i] = (1 - k2) * U[j] + k2 * U[j + 1]
                        break
        solver.step(1)
        U = solver.state()
        for i in range(len(T)):
            if (T[i] >= tt) and (T[i] < (tt + tau)):
                k1 = (T[i] - tt) / tau
                for j in range(N):
                    if ((dL * j) < Z[i]) and ((dL * (j + 1)) >= Z[i]):
                        k2 = (Z[i] - (dL * j)) / dL
                        v = (1 - k2) * U[j] + k2 * U[j + 1]
                        v = (1 - k1) * vs[i] + k1 * v
                        if v > 0.0:
                            v = 0.0
                        err += (F[i] - v) * (F[i] - v)
        if not np.isfinite(U[0]):
            err = 1e300
            break
        tt += tau
    return err if np.isfinite(err) else 1e300


def mt_history(N, alpha, tau, snaps, tstep, psi=0, beta=1.0) -> np.ndarray:
    snaps = np.ascontiguousarray(snaps, dtype=np.float64)
    os_ = snaps.shape[0]
    out = np.empty(N - 1)
    lib().orc_mt_history(N, psi, beta, alpha, tau, os_, tstep, _p(snaps.reshape(-1)), _p(out))
    return out


def psi_weights(alpha, beta, tau, A) -> np.ndarray:
    out = np.empty(A)
    lib().orc_psi_weights(alpha, beta, tau, A, _p(out))
    return out


def psi_kb_quad(alpha, beta, t0, t1, gtj, sing) -> float:
    """calc_int's adaptive-quadrature fallback _kb(1, t0, t1, gtj, sing, alpha)."""
    return lib().orc_psi_kb_quad(alpha, beta, t0, t1, gtj, int(sing))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()


def mittag_leffler(a, b, z, eps) -> float:
    return lib().orc_mittag_leffler(a, b, z, eps)


def ab_weights(a, tau, maxl, maxkiter, eps, D) -> np.ndarray:
    out = np.empty(D)
    lib().orc_ab_weights(a, tau, maxl, maxkiter, eps, D, _p(out))
    return out


def ab_solve(a, T, n, m, maxl, maxkiter, eps, F_func=0):
    C = np.empty(m + 1)
    err = lib().orc_ab_solve(a, T, n, m, maxl, maxkiter, eps, F_func, _p(C))
    return err, C


def ab_history(a, tau, maxl, maxkiter, eps, fvs) -> np.ndarray:
    fvs = np.ascontiguousarray(fvs, dtype=np.float64)
    out = np.empty(fvs.size - 1)
    lib().orc_ab_history(a, tau, maxl, maxkiter, eps, fvs.size, _p(fvs), _p(out))
    return out


def ab_series_history(a, tau, maxl, maxkiter, eps, fvs):
    """Series history of one node for i = 2..len(fvs)+1 (solve_and_check,
    simple = 0): (values, choice per step, non-finite recount count)."""
    fvs = np.ascontiguousarray(fvs, dtype=np.float64)
    out, ch = np.empty(fvs.size), np.empty(fvs.size)
    r = lib().orc_ab_series_history(a, tau, maxl, maxkiter, eps, fvs.size + 1, _p(fvs), _p(out),
                                    _p(ch))
    return out, ch, int(r)


def ab_solve_series(a, T, n, m, maxl, maxkiter, eps, F_func=0):
    C = np.empty(m + 1)
    err = lib().orc_ab_solve_series(a, T, n, m, maxl, maxkiter, eps, F_func, _p(C))
    return err, C


def ab_int(tj, ts1, ts, a, maxl=50, maxkiter=1000, eps=1e-10) -> float:
    return lib().orc_ab_int(tj, ts1, ts, a, maxl, maxkiter, eps)


def timesum01(sum_alg, sum_param, old, kb, tau):
    """H_solver::time_summing, sum_alg 0 or 1, at every node of the levels
    old[nlev, n] (irrigation_fractional.cpp:4121-4148)."""
    old = np.ascontiguousarray(old, dtype=np.float64)
    nlev, n = old.shape
    kb = np.ascontiguousarray(kb, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    assert kb.size >= nlev - 1 and tau.size >= nlev - 1
    S = np.empty(n)
    lib().orc_timesum01(sum_alg, sum_param, n, nlev, _p(old), _p(kb), _p(tau), _p(S))
    return S


class TimeSum2:
    """time_summing with sum_alg 2 (irrigation_fractional.cpp:4150-4164): the
    per-node moments Htmp, folded with the newest increment at each call."""

    def __init__(self, n, sum_param):
        self.n = n
        self.nterms = int(np.ceil(sum_param))
        self.H = np.zeros((self.nterms, n))

    def __call__(self, diff, coefs, gtj, gt, alpha, tau_n):
        diff = np.ascontiguousarray(diff, dtype=np.float64)
        coefs = np.ascontiguousarray(coefs, dtype=np.float64)
        if coefs.size == 0:
            coefs = np.zeros(1)
            nc = 0
        else:
            nc = coefs.size
        S = np.empty(self.n)
        lib().orc_timesum2(self.n, self.nterms, _p(self.H), _p(diff), _p(coefs), nc, gtj, gt,
                           alpha, tau_n, _p(S))
        return S

