// Blocked Toeplitz memory term on MFMA (launchers in frm_mfma.hip).
#pragma once
#include <cstdint>

#include "frm_internal.hpp"

namespace frm {

constexpr int kBlockSteps = 64;  // steps that share one far-part product
constexpr int kFarNodes = 64;    // nodes per far-part workgroup (ld is a multiple)

// F[t*64 + j][i] = sum_{s < N0} W_t(N0 + j - s) Delta_s[i], t = 0, 1, j < 64,
// with W_t(lag) = (whi + wlo)[t*ldw + lag - 1] (BF16-exact hi/lo parts) and
// zero for lags beyond ldw; hist holds FP32 increments with slot stride ld.
int block_far(hipStream_t st, const float* hist, int64_t ld, int N0, const float* whi,
              const float* wlo, int ldw, float* F);
// part[t*ld + i] = F[t*64 + n - N0][i] (if use_F) + sum_{N0 <= s < n} w_t(n - s) Delta_s[i]
int block_near(hipStream_t st, const float* hist, int64_t ld, int N0, int n, const double* w,
               int ldw, const float* F, int use_F, double* part);

}  // namespace frm
