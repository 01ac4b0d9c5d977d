"""Host-side mirrors of the reference's solver classes over the HIP library.

TGSolver mirrors TG_solver (fr_d_t_g_2d.cpp:2037-2771): it keeps the field in
the reference layout (b_C[(sN+2)*(M+2)], idx(i,j) = i + j*(sN+2)), counts
tstep, and calc_step() advances one LOD step -- on the GPU, through
frm_lod2d_step.  Coefficients are given as the per-node off-diagonals of the
row and column sweeps, built by one of the coefficient forms below.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _native as N


def tg_ref_spacing(sN: int, M: int):
    """Grid spacings TG_solver's constructor derives (fr_d_t_g_2d.cpp:2337-2388):
    dL = 2 Lx/(2 sN + 1) with Lx = 5, dLx = Q/M with Q = k (L/2 - H),
    k = 2 sqrt(5)/(2.5 pi), L = 22, H = 1/sqrt(5)."""
    k = 2.0 * np.sqrt(5.0) / (np.pi * 2.5)
    H = 1 / np.sqrt(5.0)
    Q = k * (0.5 * 22.0 - H)
    return 2 * 5.0 / (2.0 * sN + 1.0), Q / M


def coefs_ref(V2: np.ndarray, dv: float, dL: float, dLx: float):
    """Sweep coefficients of TG_solver_<func,fixed_d=1,kform=0>
    (fr_d_t_g_2d.cpp:3242-3265): rows A_=(V2/dL)(dv/dL+1/2), B=(V2/dL)(dv/dL-1/2);
    columns A_=B=V2 dv/dLx^2 (V2 = squared speed per node)."""
    V2 = np.asarray(V2, dtype=np.float64)
    rowA = (V2 / dL) * ((0.5 * (dv + dv) / dL) + 0.5)
    rowB = (V2 / dL) * ((0.5 * (dv + dv) / dL) - 0.5)
    colA = (V2 / (dLx * dLx)) * 0.5 * (dv + dv)
    return rowA, rowB, colA, colA.copy()


def coefs_advection(vx: np.ndarray, vy: np.ndarray, D: float, h: float):
    """Central-difference advection-diffusion sweep coefficients on a uniform
    grid: row A_ = D/h^2 + vx/(2h), B = D/h^2 - vx/(2h); columns with vy.
    The LOD split treats x-terms in the row half-step and y-terms in the
    column half-step, exactly as the reference splits its operator."""
    vx = np.asarray(vx, dtype=np.float64)
    vy = np.asarray(vy, dtype=np.float64)
    d = D / (h * h)
    return d + vx / (2 * h), d - vx / (2 * h), d + vy / (2 * h), d - vy / (2 * h)


class TGSolver:
    """TG_solver::calc_step on the MI355X (dt_alg = 0; g(t) = t, or t^g_pow
    with g_func = 3)."""

    def __init__(self, ctx: N.Context, sN: int, M: int, alpha: float, tau: float, sigma: float,
                 coefs, C0: np.ndarray, capacity: int, C1: float = 1.0, S: float = 0.0,
                 precision: int = N.PREC_F64, kernel: int = N.KERNEL_CAPUTO, ab=(50, 1000, 1e-10),
                 ab_exact: bool = False, soe_terms: int = 32, soe_f32: bool = False,
                 mfma_parts: int = 0, sweep_w: int = 0, series_terms: int = 0,
                 series_tol: float = 0.0, trunc_eps: float = 0.0, g_func: int = 0,
                 g_pow: float = 1.0, int_err: float = 0.0):
        self.ctx, self.sN, self.M = ctx, sN, M
        nn = (sN + 2) * (M + 2)
        self._coefs = [np.ascontiguousarray(c, dtype=np.float64).reshape(nn) for c in coefs]
        d = N.Lod2dDesc(sN, M, capacity, precision, alpha, tau, sigma, C1, S,
                        *[N._dp(c) for c in self._coefs], kernel, ab[0], ab[1], ab[2],
                        int(ab_exact), soe_terms, int(soe_f32), mfma_parts, sweep_w,
                        series_terms, series_tol, trunc_eps, g_func, g_pow, int_err)
        self.capacity = capacity
        h = N._V()
        N._check(N.lib().frm_lod2d_create(ctx.h, ct.byref(d), ct.byref(h)), ctx.h)
        self.h = h
        ctx.adopt(self)
        self.set_state(C0)

    def set_state(self, C: np.ndarray):
        C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1)
        assert C.size == (self.sN + 2) * (self.M + 2)
        N._check(N.lib().frm_lod2d_set_state(self.h, N._dp(C)), self.ctx.h)

    def calc_step(self, nsteps: int = 1):
        N._check(N.lib().frm_lod2d_step(self.h, nsteps), self.ctx.h)

    @property
    def tstep(self) -> int:
        t = ct.c_int32()
        N._check(N.lib().frm_lod2d_tstep(self.h, ct.byref(t)), self.ctx.h)
        return t.value

    @property
    def b_C(self) -> np.ndarray:
        out = np.empty((self.M + 2) * (self.sN + 2))
        N._check(N.lib().frm_lod2d_get_state(self.h, N._dp(out)), self.ctx.h)
        return out.reshape(self.M + 2, self.sN + 2)

    @property
    def fit_error(self) -> float:
        """Largest relative error of the SOE fit over the weights it replaces."""
        v = ct.c_double()
        N._check(N.lib().frm_lod2d_fit_error(self.h, ct.byref(v)), self.ctx.h)
        return v.value

    @property
    def far_slices(self) -> int:
        """Node slices of the MFMA form's far product per 64-step block (0: other forms)."""
        v = ct.c_int32()
        N._check(N.lib().frm_lod2d_far_slices(self.h, ct.byref(v)), self.ctx.h)
        return v.value

    @property
    def sweep_shape(self):
        """(elements per lane, wavefronts per line) of the row and the column sweep."""
        out = (ct.c_int32 * 4)()
        N._check(N.lib().frm_lod2d_sweep_shape(self.h, out), self.ctx.h)
        return (out[0], out[1]), (out[2], out[3])

    def weights(self, n: int):
        """(row current, column current, half-step correction, row weights,
        column weights) for lags 1..n."""
        out = np.empty(3 + 2 * n)
        N._check(N.lib().frm_lod2d_weights(self.h, n, N._dp(out)), self.ctx.h)
        return out[0], out[1], out[2], out[3:3 + n], out[3 + n:]

    def close(self):
        if getattr(self, "h", None):
            N.lib().frm_lod2d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class MT1DBatch:
    """A batch of H_solver problems (mt_minimal.cpp / mt_psi_caputo.cpp) stepped
    together on the GPU.  `problems` is a list of dicts with keys of
    frm_mt1d_problem (N, L, alpha, form, ...); the state is the list of the
    problems' U arrays (N+2 values each, nodes 0..N and the ghost N+1)."""

    FIELDS = ("N", "form", "top", "bottom_cond", "L", "alpha", "beta", "D", "k", "k0", "b",
              "top_val", "Ev", "Hbottom", "src", "ab_maxl", "ab_maxkiter", "ab_exact", "ab_eps",
              "ab_series")

    def __init__(self, ctx: N.Context, problems, tau: float, capacity: int, U0,
                 kind: int = N.MT_L1, precision: int = N.PREC_F64, picard: int = 1,
                 graph: bool = False, hist_block: int = 0):
        self.ctx = ctx
        defaults = dict(form=N.FORM_CONST, top=N.TOP_AUTO, bottom_cond=1, L=1.0, beta=1.0, D=0.0,
                        k=0.0, k0=0.0, b=0.0, top_val=0.0, Ev=0.0, Hbottom=0.0, src=0,
                        ab_maxl=0, ab_maxkiter=0, ab_exact=0, ab_eps=0.0, ab_series=0)
        arr = (N.MT1DProblem * len(problems))()
        for i, pr in enumerate(problems):
            q = {**defaults, **pr}
            for f in self.FIELDS:
                setattr(arr[i], f, q[f])
        self.sizes = [int(pr["N"]) + 2 for pr in problems]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(int)
        self._arr = arr
        d = N.MT1DDesc(len(problems), kind, capacity, precision, picard, tau, arr, int(graph),
                       int(hist_block))
        h = N._V()
        N._check(N.lib().frm_mt1d_create(ctx.h, ct.byref(d), ct.byref(h)), ctx.h)
        self.h = h
        ctx.adopt(self)
        self.set_state(U0)

    def _pack(self, Us):
        flat = np.concatenate([np.asarray(u, dtype=np.float64).reshape(-1) for u in Us])
        assert flat.size == self.offsets[-1]
        return np.ascontiguousarray(flat)

    def set_state(self, Us):
        N._check(N.lib().frm_mt1d_set_state(self.h, N._dp(self._pack(Us))), self.ctx.h)

    def calc_step(self, nsteps: int = 1):
        N._check(N.lib().frm_mt1d_step(self.h, nsteps), self.ctx.h)

    @property
    def tstep(self) -> int:
        t = ct.c_int32()
        N._check(N.lib().frm_mt1d_tstep(self.h, ct.byref(t)), self.ctx.h)
        return t.value

    @property
    def hist_block(self) -> int:
        """The history block length in use (autotuned unless given)."""
        t = ct.c_int32()
        N._check(N.lib().frm_mt1d_hist_block(self.h, ct.byref(t)), self.ctx.h)
        return t.value

    def state(self):
        out = np.empty(self.offsets[-1])
        N._check(N.lib().frm_mt1d_get_state(self.h, N._dp(out)), self.ctx.h)
        return [out[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.sizes))]

    def ab_error(self) -> np.ndarray:
        """Kind MT_AB: per problem, the reference's error measure of
        solve_and_check over the steps taken so far."""
        out = np.empty(len(self.sizes))
        N._check(N.lib().frm_mt1d_ab_error(self.h, N._dp(out)), self.ctx.h)
        return out

    def set_observations(self, T, Z, F):
        """Observations (time, depth, value) scored after every step as the
        reference's fitting drivers score a particle (solve_and_test); each
        problem of the batch is one particle.  Clears the errors."""
        self._obs = [np.ascontiguousarray(a, dtype=np.float64) for a in (T, Z, F)]
        n = self._obs[0].size
        assert all(a.size == n for a in self._obs)
        N._check(N.lib().frm_mt1d_set_observations(self.h, n, *[N._dp(a) for a in self._obs]),
                 self.ctx.h)

    def fit_error(self) -> np.ndarray:
        """Per problem: sum of squared misfits so far (1e300 once U[0] was
        not finite), as solve_and_test returns it."""
        out = np.empty(len(self.sizes))
        N._check(N.lib().frm_mt1d_fit_error(self.h, N._dp(out)), self.ctx.h)
        return out

    def close(self):
        if getattr(self, "h", None):
            N.lib().frm_mt1d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class TimeSum:
    """The Sum family of irrigation_fractional.cpp on the MI355X: the time
    sums H_solver::time_summing computes (4099-4170) and fr_ocl_call_sum
    offloads (1828-1905), over time levels kept on the device.

    push(level) is oldH.push_back(level) (a host array of n_nodes values, or a
    DeviceArray); sums(...) returns S[id] for every node, given what the
    reference's caller holds at that step: kb_3_cache and tau (sum_alg 0, 1),
    or row2_precalc's coefficients with g(Time[tstep+1]), g(Time[tstep]),
    alpha and tau[tstep] (sum_alg 2)."""

    def __init__(self, ctx: N.Context, n_nodes: int, sum_alg: int = 0, sum_param: float = 0.0,
                 capacity: int = 1, precision: int = N.PREC_F64):
        self.ctx, self.n = ctx, int(n_nodes)
        d = N.TimesumDesc(self.n, capacity, sum_alg, sum_param, precision)
        h = N._V()
        N._check(N.lib().frm_timesum_create(ctx.h, ct.byref(d), ct.byref(h)), ctx.h)
        self.h = h
        ctx.adopt(self)
        self._lev = N.DeviceArray(ctx, self.n)
        self._S = N.DeviceArray(ctx, self.n)

    def push(self, level):
        if isinstance(level, N.DeviceArray):
            assert level.n == self.n
            ptr = level.ptr
        else:
            level = np.ascontiguousarray(level, dtype=np.float64).reshape(-1)
            assert level.size == self.n
            self.ctx.upload(self._lev.ptr, level)
            ptr = self._lev.ptr
        N._check(N.lib().frm_timesum_push(self.h, ptr), self.ctx.h)

    @property
    def old_size(self) -> int:
        v = ct.c_int32()
        N._check(N.lib().frm_timesum_levels(self.h, ct.byref(v)), self.ctx.h)
        return v.value

    def eval_device(self, kb_3_cache=None, tau=None, coefs=None, gtj=0.0, gt=0.0, alpha=0.0,
                    tau_n=1.0) -> N.DeviceArray:
        """The sums into a device buffer (no copy back)."""
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return N._dp(a)

        nc = 0 if coefs is None else int(np.asarray(coefs).size)
        a = N.TimesumArgs(ptr(kb_3_cache), ptr(tau), ptr(coefs), nc, gtj, gt, alpha, tau_n)
        N._check(N.lib().frm_timesum_eval(self.h, ct.byref(a), self._S.ptr), self.ctx.h)
        return self._S

    def sums(self, **kw) -> np.ndarray:
        return self.eval_device(**kw).to_host()

    def clear(self):
        N._check(N.lib().frm_timesum_clear(self.h), self.ctx.h)

    def close(self):
        if getattr(self, "h", None):
            N.lib().frm_timesum_destroy(self.h)
            self.h = None
            self._lev.free()
            self._S.free()

    def __del__(self):
        self.close()



class SoeHistory:
    """Per-node sum-of-exponentials accumulators (frm_soe_*): the form the
    reference's series histories take (the Lc2 sums of MF_intF_Fv,
    MF_int_series.cpp:720-836; the dt_alg 3 moments, fr_d_t_g_2d.cpp:3385-3428).

    step(incr, q, c) folds the increment into every mode, H_l = q_l H_l + incr,
    and returns scale * sum_l c[t, l] H_l for each target t; step(None, c=c)
    evaluates without folding."""

    def __init__(self, ctx: N.Context, n_nodes: int, nterms: int, ntargets: int = 1,
                 f32: bool = False):
        self.ctx, self.n, self.nterms, self.ntargets = ctx, int(n_nodes), int(nterms), int(ntargets)
        d = N.SoeDesc(self.n, self.nterms, self.ntargets, int(f32))
        h = N._V()
        N._check(N.lib().frm_soe_create(ctx.h, ct.byref(d), ct.byref(h)), ctx.h)
        self.h = h
        ctx.adopt(self)
        self._inc = N.DeviceArray(ctx, self.n)
        self._out = N.DeviceArray(ctx, self.n * self.ntargets)

    def step_device(self, incr, q=None, c=None, scale: float = 1.0,
                    accumulate: bool = False) -> N.DeviceArray:
        """As step(), leaving the targets in a device buffer (n_nodes values
        per target).  `incr` is a host array, a DeviceArray or None."""
        if incr is None:
            iptr = None
        elif isinstance(incr, N.DeviceArray):
            assert incr.n == self.n
            iptr = incr.ptr
        else:
            a = np.ascontiguousarray(incr, dtype=np.float64).reshape(-1)
            assert a.size == self.n
            self.ctx.upload(self._inc.ptr, a)
            iptr = self._inc.ptr
        qa = None if q is None else np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        ca = np.ascontiguousarray(c, dtype=np.float64).reshape(-1)
        assert ca.size == self.nterms * self.ntargets and (qa is None or qa.size == self.nterms)
        N._check(N.lib().frm_soe_step(self.ctx.h, self.h, iptr, None if qa is None else N._dp(qa),
                                      N._dp(ca), scale, int(accumulate), self._out.ptr),
                 self.ctx.h)
        return self._out

    def step(self, incr, q=None, c=None, scale: float = 1.0) -> np.ndarray:
        return self.step_device(incr, q, c, scale).to_host().reshape(self.ntargets, self.n)

    def clear(self):
        N._check(N.lib().frm_soe_clear(self.h), self.ctx.h)

    def close(self):
        if getattr(self, "h", None):
            N.lib().frm_soe_destroy(self.h)
            self.h = None
            self._inc.free()
            self._out.free()

    def __del__(self):
        self.close()
