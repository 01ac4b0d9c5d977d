// Blocked Toeplitz memory term on MFMA (split hi+lo BF16, FP32 accumulation).
//
// For Toeplitz weights the history sums of a block of future steps share
// their settled past: for the steps n = N0 + j, j < 64, of a block starting
// at N0,
//   S_t(n) = sum_{s < N0} w_t(N0 + j - s) Delta_s  +  sum_{N0 <= s < n} w_t(n - s) Delta_s,
// and the first ("far") part for all 64 j and both LOD targets t is one
// dense product  F[t*64 + j][i] = sum_{s < N0} W[t*64 + j][s] Delta_s[i]  with
// M = 128, K = N0, N = nodes -- read once per 64 steps instead of every step.
// The reference forms every step's sum from scratch (fr_d_t_g_2d.cpp:
// 3340-3437).  W and Delta are split into BF16 hi and lo parts and the three
// products hi*hi + hi*lo + lo*hi accumulate in FP32 on
// v_mfma_f32_16x16x32_bf16; the near part (at most 63 slots) is summed per
// step in FP64.
#include "frm_mfma.hpp"

namespace frm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kKc = 32;   // slots per MFMA k-chunk
constexpr int kNt = 64;   // nodes per workgroup
constexpr int kBpad = 40; // LDS row length of a node's 32 slots (bf16), padded

// Workgroup: 4 wavefronts, 128 rows (2 targets x 64 steps) x 64 nodes.
// Wavefront w owns row tiles 2w, 2w+1 (16 rows each) and all 4 column
// tiles; lane l holds A[l%16][8(l/16)+q], B[8(l/16)+q][l%16] and
// D[4(l/16)+r][l%16] of each 16x16x32 product.
__global__ __launch_bounds__(256) void k_block_far(const float* __restrict__ hist, int64_t ld,
                                                   int N0, const float* __restrict__ whi,
                                                   const float* __restrict__ wlo, int ldw,
                                                   float* __restrict__ F) {
    __shared__ __attribute__((aligned(16))) __bf16 sBh[kNt][kBpad];
    __shared__ __attribute__((aligned(16))) __bf16 sBl[kNt][kBpad];
    __shared__ float sW[2][2][96];  // [hi/lo][target][y], y = 63 - j + k
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int kg = lane >> 4, lc = lane & 15;
    const int64_t i0 = (int64_t)blockIdx.x * kNt;
    f32x4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int c = tid & 63, kq = tid >> 6;  // staging: node c, slots kq*8 .. kq*8+7
    for (int s0 = 0; s0 < N0; s0 += kKc) {
        // ---- stage 32 slots x 64 nodes of history as BF16 hi/lo, node-major in LDS
        bf16x8 hv, lv;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int s = s0 + kq * 8 + q;
            const float x = s < N0 ? hist[(int64_t)s * ld + i0 + c] : 0.f;
            const __bf16 h = (__bf16)x;
            hv[q] = h;
            lv[q] = (__bf16)(x - (float)h);
        }
        *reinterpret_cast<bf16x8*>(&sBh[c][kq * 8]) = hv;
        *reinterpret_cast<bf16x8*>(&sBl[c][kq * 8]) = lv;
        // ---- lag table of the chunk: lag = N0 - s0 + 63 - y (>= 1 inside the history)
        for (int e = tid; e < 2 * 95; e += 256) {
            const int t = e / 95, y = e - t * 95;
            const int lag = N0 - s0 + 63 - y;
            const bool ok = lag >= 1 && lag <= ldw;
            sW[0][t][y] = ok ? whi[(int64_t)t * ldw + lag - 1] : 0.f;
            sW[1][t][y] = ok ? wlo[(int64_t)t * ldw + lag - 1] : 0.f;
        }
        __syncthreads();
        bf16x8 bh[4], bl[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            bh[ct] = *reinterpret_cast<const bf16x8*>(&sBh[ct * 16 + lc][kg * 8]);
            bl[ct] = *reinterpret_cast<const bf16x8*>(&sBl[ct * 16 + lc][kg * 8]);
        }
#pragma unroll
        for (int rr = 0; rr < 2; ++rr) {
            const int rt = wave * 2 + rr;
            const int t = rt >> 2, j = (rt & 3) * 16 + lc;
            bf16x8 ah, al;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int y = 63 - j + kg * 8 + q;
                ah[q] = (__bf16)sW[0][t][y];
                al[q] = (__bf16)sW[1][t][y];
            }
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                acc[rr][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[ct], acc[rr][ct], 0, 0, 0);
                acc[rr][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[ct], acc[rr][ct], 0, 0, 0);
                acc[rr][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[ct], acc[rr][ct], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int rt = wave * 2 + rr;
        const int t = rt >> 2;
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = (rt & 3) * 16 + kg * 4 + r;
                F[(int64_t)(t * 64 + j) * ld + i0 + ct * 16 + lc] = acc[rr][ct][r];
            }
    }
}

// part[t*ld + i] = F[t*64 + n - N0][i] + sum_{N0 <= s < n} w_t(n - s) Delta_s[i],
// four nodes per thread (16-byte loads of the FP32 slots), FP64 sums.
__global__ __launch_bounds__(256) void k_block_near(const float* __restrict__ hist, int64_t ld,
                                                    int N0, int n, const double* __restrict__ w,
                                                    int ldw, const float* __restrict__ F,
                                                    int use_F, double* __restrict__ part) {
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= ld) return;
    const int j = n - N0;
    double s0[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0};
    if (use_F) {
        const float4 f0 = *reinterpret_cast<const float4*>(F + (int64_t)j * ld + i);
        const float4 f1 = *reinterpret_cast<const float4*>(F + (int64_t)(kBlockSteps + j) * ld + i);
        s0[0] = f0.x, s0[1] = f0.y, s0[2] = f0.z, s0[3] = f0.w;
        s1[0] = f1.x, s1[1] = f1.y, s1[2] = f1.z, s1[3] = f1.w;
    }
    for (int s = N0; s < n; ++s) {
        const float4 x = *reinterpret_cast<const float4*>(hist + (int64_t)s * ld + i);
        const double a = w[n - s - 1], b = w[ldw + n - s - 1];
        s0[0] += a * (double)x.x, s1[0] += b * (double)x.x;
        s0[1] += a * (double)x.y, s1[1] += b * (double)x.y;
        s0[2] += a * (double)x.z, s1[2] += b * (double)x.z;
        s0[3] += a * (double)x.w, s1[3] += b * (double)x.w;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        part[i + q] = s0[q];
        part[ld + i + q] = s1[q];
    }
}

int block_far(hipStream_t st, const float* hist, int64_t ld, int N0, const float* whi,
              const float* wlo, int ldw, float* F) {
    if (ld % kNt || N0 <= 0) return -1;
    const dim3 grid((unsigned)(ld / kNt)), block(256);
    hipLaunchKernelGGL(k_block_far, grid, block, 0, st, hist, ld, N0, whi, wlo, ldw, F);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int block_near(hipStream_t st, const float* hist, int64_t ld, int N0, int n, const double* w,
               int ldw, const float* F, int use_F, double* part) {
    if (ld % 4 || n < N0 || n - N0 >= kBlockSteps) return -1;
    const int64_t threads = ld / 4;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_block_near, grid, block, 0, st, hist, ld, N0, n, w, ldw, F, use_F, part);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace frm
