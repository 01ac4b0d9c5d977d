// Blocked Toeplitz memory term on MFMA (launchers in frm_mfma.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "frm_internal.hpp"

namespace frm {

constexpr int kBlockSteps = 64;  // steps that share one far-part product
constexpr int kSubSteps = 8;     // steps that share one mid-part sum

constexpr int kFarChunk = 32;    // history slots per product k-chunk

// F[t*64 + j][i] = sum_{s < K} W_t(N0 + j - s) Delta_s[i], t = 0, 1, j < 64,
// N0 and K <= N0 multiples of 64, for the nodes of workgroups wg0 ..
// wg0 + nwg - 1 (128 nodes each, block_far_groups(ld) in all), from the
// weight-fragment table of atab_build (which must cover N0/32 + 1 chunks);
// hist holds FP32 increments with slot stride ld.  nprod = 3 uses two BF16
// parts per operand, 6 uses three.
int block_far(hipStream_t st, const float* hist, int64_t ld, int N0, int K, const void* atab,
              int atab_chunks, int nprod, float* F, int wg0, int nwg);
int block_far_groups(int64_t ld);
// Workgroups of block_far resident on the whole of `device` at once (one
// round); -1 if the device cannot be queried.
int block_far_round(int device);
// Host: weight-fragment table for chunks d = 32 .. 32*chunks from the BF16
// parts of the weights, wparts[(p*2 + t)*ldw + lag - 1], p < 3.
std::vector<uint16_t> atab_build(const float* wparts, int ldw, int chunks);
// G[t*8 + j][i] = sum_{N0 <= s < N1} w_t(N1 + j - s) Delta_s[i], j < 8, summed
// in FP64, stored in FP32; N1 - N0 a positive multiple of 8 below 128.
int block_mid(hipStream_t st, const float* hist, int64_t ld, int N0, int N1, const double* w,
              int ldw, float* G);
// part[t*ld + i] = F[t*64 + n - N0][i] (if use_F) + G[t*8 + n - N1][i] (if use_G)
//                  + sum_{N1 <= s < n} w_t(n - s) Delta_s[i]
int block_near(hipStream_t st, const float* hist, int64_t ld, int N0, int N1, int n,
               const double* w, int ldw, const float* F, int use_F, const float* G, int use_G,
               double* part);

}  // namespace frm
