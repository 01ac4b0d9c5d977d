"""MI355X-native hot path of the time-fractional solvers of sevaboh/theses_code:
memory-term history sums and tridiagonal sweeps, as hand-written HIP kernels
for gfx950 behind the C ABI in include/frm.h."""
from ._native import (DeviceArray, Context, FrmError, lib, PREC_F64, PREC_F32C, PREC_MFMA_BF16,
                      PREC_SOE, PREC_BINOMIAL, K_MEMTERM, K_ROWS, K_COLS)
from .solvers import MT1DBatch, SoeHistory, TGSolver, TimeSum, coefs_ref, coefs_advection, tg_ref_spacing

__all__ = ["DeviceArray", "Context", "FrmError", "lib", "TGSolver", "MT1DBatch", "SoeHistory", "TimeSum", "coefs_ref",
           "coefs_advection", "tg_ref_spacing", "PREC_F64", "PREC_F32C", "PREC_MFMA_BF16",
           "PREC_SOE", "PREC_BINOMIAL", "K_MEMTERM", "K_ROWS", "K_COLS"]
