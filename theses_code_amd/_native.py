"""ctypes binding of libfrm.so (include/frm.h).

The HIP library is the only implementation: if libfrm.so is missing or a call
fails, this module raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfrm.so")

FRM_OK, FRM_E_INVALID, FRM_E_HIP, FRM_E_NOMEM, FRM_E_RANGE, FRM_E_STATE = 0, -1, -2, -3, -4, -5
PREC_F64, PREC_F32C, PREC_MFMA_BF16, PREC_SOE, PREC_BINOMIAL, PREC_IAIBK = 0, 1, 2, 3, 4, 5
KERNEL_CAPUTO, KERNEL_AB = 0, 1
K_MEMTERM, K_ROWS, K_COLS, K_SWEEP1D, K_OTHER, K_BLOCKFAR = 0, 1, 2, 3, 4, 5

_D, _I, _I64, _U32 = ct.c_double, ct.c_int, ct.c_int64, ct.c_uint32
_P, _V, _S = ct.POINTER(ct.c_double), ct.c_void_p, ct.c_size_t


class FrmError(RuntimeError):
    pass


class MemtermDesc(ct.Structure):
    _fields_ = [("ntargets", ct.c_int32), ("ldw", ct.c_int32), ("d_w", _V), ("scale", _D),
                ("accumulate", ct.c_int32)]


class TridiagDesc(ct.Structure):
    _fields_ = [("nsys", ct.c_int32), ("n", ct.c_int32), ("stride_sys", ct.c_int64),
                ("stride_elem", ct.c_int64)]


class Lod2dDesc(ct.Structure):
    _fields_ = [("sN", ct.c_int32), ("M", ct.c_int32), ("capacity", ct.c_int32),
                ("precision", ct.c_int32), ("alpha", _D), ("tau", _D), ("sigma", _D),
                ("C1", _D), ("S", _D), ("rowA", _P), ("rowB", _P), ("colA", _P), ("colB", _P),
                ("kernel", ct.c_int32), ("ab_maxl", ct.c_int32), ("ab_maxkiter", ct.c_int32),
                ("ab_eps", _D), ("ab_exact", ct.c_int32), ("soe_terms", ct.c_int32),
                ("soe_f32", ct.c_int32), ("mfma_parts", ct.c_int32), ("sweep_w", ct.c_int32),
                ("series_terms", ct.c_int32), ("series_tol", _D), ("trunc_eps", _D),
                ("g_func", ct.c_int32), ("g_pow", _D), ("int_err", _D)]


class MT1DProblem(ct.Structure):
    _fields_ = [("N", ct.c_int32), ("form", ct.c_int32), ("top", ct.c_int32),
                ("bottom_cond", ct.c_int32), ("L", _D), ("alpha", _D), ("beta", _D), ("D", _D),
                ("k", _D), ("k0", _D), ("b", _D), ("top_val", _D), ("Ev", _D), ("Hbottom", _D),
                ("src", ct.c_int32), ("ab_maxl", ct.c_int32), ("ab_maxkiter", ct.c_int32),
                ("ab_exact", ct.c_int32), ("ab_eps", _D), ("ab_series", ct.c_int32)]


class MT1DDesc(ct.Structure):
    _fields_ = [("nprob", ct.c_int32), ("kind", ct.c_int32), ("capacity", ct.c_int32),
                ("precision", ct.c_int32), ("picard", ct.c_int32), ("tau", _D),
                ("probs", ct.POINTER(MT1DProblem)), ("graph", ct.c_int32),
                ("hist_block", ct.c_int32)]


class TimesumDesc(ct.Structure):
    _fields_ = [("n_nodes", ct.c_int64), ("capacity", ct.c_int32), ("sum_alg", ct.c_int32),
                ("sum_param", _D), ("precision", ct.c_int32)]


class TimesumArgs(ct.Structure):
    _fields_ = [("kb_3_cache", _P), ("tau", _P), ("coefs", _P), ("ncoefs", ct.c_int32),
                ("gtj", _D), ("gt", _D), ("alpha", _D), ("tau_n", _D)]


class SoeDesc(ct.Structure):
    _fields_ = [("n_nodes", ct.c_int64), ("nterms", ct.c_int32), ("ntargets", ct.c_int32),
                ("f32", ct.c_int32)]


MT_L1, MT_PSI, MT_AB = 0, 1, 2
FORM_CONST, FORM_CV2, FORM_KEXP = 0, 1, 2
TOP_AUTO, TOP_DIRICHLET, TOP_FLUX = 0, 1, 2

# name -> (restype, argtypes); every symbol include/frm.h declares
SIGNATURES = {
    # host-side generators (include/frm_host.h)
    "frm_host_gamma": (_I, [_D, _P]),
    "frm_host_mittag_leffler": (_I, [_D, _D, _D, _D, _P]),
    "frm_host_ab_integral": (_I, [_D, ct.c_int32, ct.c_int32, _D, _D, _D, _D, _P]),
    "frm_host_ab_weights": (_I, [_D, _D, ct.c_int32, ct.c_int32, _D, ct.c_int32, ct.c_int32, _P]),
    "frm_host_caputo_lod_weights": (_I, [_D, _D, ct.c_int32, _P]),
    "frm_host_soe_fit": (_I, [_P, ct.c_int32, ct.c_int32, ct.c_int32, ct.c_int32, _D, _P, _P, _P]),
    "frm_host_lod_weights_pow": (_I, [_D, _D, _D, _D, ct.c_int32, _P]),
    "frm_ctx_create": (_I, [_I, ct.POINTER(_V)]),
    "frm_ctx_destroy": (_I, [_V]),
    "frm_sync": (_I, [_V]),
    "frm_strerror": (ct.c_char_p, [_I]),
    "frm_last_error": (ct.c_char_p, [_V]),
    "frm_malloc": (_I, [_V, _S, ct.POINTER(_V)]),
    "frm_free": (_I, [_V, _V]),
    "frm_upload": (_I, [_V, _V, _V, _S]),
    "frm_download": (_I, [_V, _V, _V, _S]),
    "frm_timing_enable": (_I, [_V, _U32]),
    "frm_timing_reset": (_I, [_V]),
    "frm_timing_get": (_I, [_V, _I, ct.POINTER(_I64), _P, _P]),
    "frm_hist_create": (_I, [_V, _I64, ct.c_int32, ct.c_int32, ct.POINTER(_V)]),
    "frm_hist_destroy": (_I, [_V]),
    "frm_hist_push": (_I, [_V, _V, _V, _V]),
    "frm_hist_depth": (_I, [_V, ct.POINTER(ct.c_int32)]),
    "frm_hist_clear": (_I, [_V]),
    "frm_memterm": (_I, [_V, _V, ct.POINTER(MemtermDesc), _V]),
    "frm_memterm_static": (_I, [_V, _V, ct.POINTER(MemtermDesc), _V]),
    "frm_tridiag": (_I, [_V, ct.POINTER(TridiagDesc), _V, _V, _V, _V, _V]),
    "frm_lod2d_create": (_I, [_V, ct.POINTER(Lod2dDesc), ct.POINTER(_V)]),
    "frm_lod2d_destroy": (_I, [_V]),
    "frm_lod2d_set_state": (_I, [_V, _P]),
    "frm_lod2d_step": (_I, [_V, ct.c_int32]),
    "frm_lod2d_get_state": (_I, [_V, _P]),
    "frm_lod2d_tstep": (_I, [_V, ct.POINTER(ct.c_int32)]),
    "frm_lod2d_state_ptr": (_I, [_V, ct.POINTER(_V)]),
    "frm_diag_copy": (_I, [_V, _V, _V, _S, _I, _P]),
    "frm_diag_read": (_I, [_V, _V, _S, _I, _P]),
    "frm_diag_psi_kb_quad": (_I, [_V, ct.c_int32, _P, _P]),
    "frm_lod2d_fit_error": (_I, [_V, _P]),
    "frm_lod2d_sweep_shape": (_I, [_V, ct.POINTER(ct.c_int32)]),
    "frm_lod2d_far_slices": (_I, [_V, ct.POINTER(ct.c_int32)]),
    "frm_lod2d_weights": (_I, [_V, ct.c_int32, _P]),
    "frm_mt1d_create": (_I, [_V, ct.POINTER(MT1DDesc), ct.POINTER(_V)]),
    "frm_mt1d_destroy": (_I, [_V]),
    "frm_mt1d_set_state": (_I, [_V, _P]),
    "frm_mt1d_step": (_I, [_V, ct.c_int32]),
    "frm_mt1d_get_state": (_I, [_V, _P]),
    "frm_mt1d_tstep": (_I, [_V, ct.POINTER(ct.c_int32)]),
    "frm_mt1d_hist_block": (_I, [_V, ct.POINTER(ct.c_int32)]),
    "frm_mt1d_ab_error": (_I, [_V, _P]),
    "frm_mt1d_set_observations": (_I, [_V, _I, _P, _P, _P]),
    "frm_mt1d_fit_error": (_I, [_V, _P]),
    "frm_timesum_create": (_I, [_V, ct.POINTER(TimesumDesc), ct.POINTER(_V)]),
    "frm_timesum_destroy": (_I, [_V]),
    "frm_timesum_clear": (_I, [_V]),
    "frm_timesum_push": (_I, [_V, _V]),
    "frm_timesum_levels": (_I, [_V, ct.POINTER(ct.c_int32)]),
    "frm_timesum_eval": (_I, [_V, ct.POINTER(TimesumArgs), _V]),
    "frm_soe_create": (_I, [_V, ct.POINTER(SoeDesc), ct.POINTER(_V)]),
    "frm_soe_destroy": (_I, [_V]),
    "frm_soe_clear": (_I, [_V]),
    "frm_soe_step": (_I, [_V, _V, _V, _P, _P, _D, ct.c_int32, _V]),
}

_lib = None


def lib():
    """Load libfrm.so; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FrmError(f"{LIB_PATH} is missing: build it with `python -m theses_code_amd.build` "
                           "(the HIP library is required; there is no CPU fallback)")
        L = ct.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, ctx=None):
    if rc != FRM_OK:
        msg = lib().frm_last_error(ctx)
        raise FrmError(f"{lib().frm_strerror(rc).decode()} ({rc}): {msg.decode() if msg else ''}")


def _dp(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_P)


class Context:
    """One device and one HIP stream (frm_ctx)."""

    def __init__(self, device: int = 0):
        h = _V()
        _check(lib().frm_ctx_create(device, ct.byref(h)))
        self.h = h
        self._owned = weakref.WeakSet()

    def adopt(self, obj):
        """Registers an object with a close() that must run before the context
        is destroyed (solvers hold the context's stream)."""
        self._owned.add(obj)

    def sync(self):
        _check(lib().frm_sync(self.h), self.h)

    def malloc(self, nbytes: int) -> int:
        p = _V()
        _check(lib().frm_malloc(self.h, nbytes, ct.byref(p)), self.h)
        return p.value

    def free(self, ptr: int):
        _check(lib().frm_free(self.h, ptr), self.h)

    def upload(self, dptr: int, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        _check(lib().frm_upload(self.h, dptr, arr.ctypes.data, arr.nbytes), self.h)

    def download(self, dptr: int, arr: np.ndarray):
        _check(lib().frm_download(self.h, arr.ctypes.data, dptr, arr.nbytes), self.h)
        return arr

    def timing_enable(self, mask: int):
        _check(lib().frm_timing_enable(self.h, mask), self.h)

    def timing_reset(self):
        _check(lib().frm_timing_reset(self.h), self.h)

    def timing(self, kclass: int):
        n, ms, by = _I64(), _D(), _D()
        _check(lib().frm_timing_get(self.h, kclass, ct.byref(n), ct.byref(ms), ct.byref(by)), self.h)
        return n.value, ms.value, by.value

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 20) -> float:
        """Measured device copy bandwidth in GB/s (read + write bytes)."""
        src, dst = self.malloc(nbytes), self.malloc(nbytes)
        ms = _D()
        try:
            _check(lib().frm_diag_copy(self.h, dst, src, nbytes, reps, ct.byref(ms)), self.h)
        finally:
            self.free(src)
            self.free(dst)
        return 2.0 * nbytes / (ms.value * 1e-3) / 1e9

    def read_bandwidth(self, nbytes: int = 4 << 30, reps: int = 10) -> float:
        """Measured device read bandwidth in GB/s (a non-temporal stream of
        4-byte columns, 16 rows in flight per lane, one small write per lane);
        nbytes a multiple of 8 MiB."""
        src = self.malloc(nbytes)
        ms = _D()
        try:
            _check(lib().frm_diag_read(self.h, src, nbytes, reps, ct.byref(ms)), self.h)
        finally:
            self.free(src)
        return nbytes / (ms.value * 1e-3) / 1e9

    def psi_kb_quad(self, cases) -> np.ndarray:
        """The device quadrature fallback of the psi weights on rows (alpha,
        beta, t0, t1, gtj, sing)."""
        c = np.ascontiguousarray(cases, dtype=np.float64).reshape(-1, 6)
        out = np.empty(len(c))
        _check(lib().frm_diag_psi_kb_quad(self.h, len(c), _dp(c), _dp(out)), self.h)
        return out

    def close(self):
        if getattr(self, "h", None):
            for obj in list(getattr(self, "_owned", ())):
                obj.close()
            lib().frm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class DeviceArray:
    """A float64 device buffer owned by a Context (for tests of the low-level ABI)."""

    def __init__(self, ctx: Context, n: int):
        self.ctx, self.n = ctx, int(n)
        self.ptr = ctx.malloc(max(1, self.n) * 8)

    @classmethod
    def from_host(cls, ctx: Context, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(ctx, a.size)
        ctx.upload(d.ptr, a)
        return d

    def to_host(self) -> np.ndarray:
        return self.ctx.download(self.ptr, np.empty(self.n))

    def free(self):
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = 0


# ---------------------------------------------------------------- host generators
def host_gamma(x: float) -> float:
    out = _D()
    _check(lib().frm_host_gamma(x, ct.byref(out)))
    return out.value


def host_mittag_leffler(a: float, b: float, z: float, eps: float) -> float:
    out = _D()
    _check(lib().frm_host_mittag_leffler(a, b, z, eps, ct.byref(out)))
    return out.value


def host_ab_integral(a, maxl, maxkiter, eps, tj, ts1, ts) -> float:
    out = _D()
    _check(lib().frm_host_ab_integral(a, maxl, maxkiter, eps, tj, ts1, ts, ct.byref(out)))
    return out.value


def host_ab_weights(a, tau, maxl, maxkiter, eps, exact, D) -> np.ndarray:
    out = np.empty(D)
    _check(lib().frm_host_ab_weights(a, tau, maxl, maxkiter, eps, int(exact), D, _dp(out)))
    return out


def host_caputo_lod_weights(alpha, tau, D):
    out = np.empty(3 + 2 * D)
    _check(lib().frm_host_caputo_lod_weights(alpha, tau, D, _dp(out)))
    return out[0], out[1], out[2], out[3:3 + D], out[3 + D:]


def host_lod_weights_pow(alpha, tau, p, int_err, n):
    """Step-n LOD weights for g(t) = t^p: (row current, column current,
    half-step correction, row weights, column weights), lags 1..n."""
    out = np.empty(3 + 2 * n)
    _check(lib().frm_host_lod_weights_pow(alpha, tau, p, int_err, n, _dp(out)))
    return out[0], out[1], out[2], out[3:3 + n], out[3 + n:]


def host_soe_fit(w: np.ndarray, k0: int, L: int, tau: float):
    """w: [ntargets][K] lag-indexed weights; returns (q[L], c[ntargets][L], max_rel)."""
    w = np.ascontiguousarray(np.atleast_2d(w), dtype=np.float64)
    nt, K = w.shape
    q, c, mr = np.empty(L), np.empty(nt * L), _D()
    _check(lib().frm_host_soe_fit(_dp(w), nt, K, k0, L, tau, _dp(q), _dp(c), ct.byref(mr)))
    return q, c.reshape(nt, L), mr.value
