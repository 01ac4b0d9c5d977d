// Blocked Toeplitz memory term on MFMA (split hi+lo BF16, FP32 accumulation).
//
// For Toeplitz weights the history sums of a block of future steps share
// their settled past: for the steps n = N0 + j, j < 64, of a block starting
// at N0, and P0 = N0 - 64,
//   S_t(n) = sum_{s < P0} w_t(N0 + j - s) Delta_s  +  sum_{P0 <= s < n} w_t(n - s) Delta_s,
// and the first ("far") part for all 64 j and both LOD targets t is one
// dense product  F[t*64 + j][i] = sum_{s < P0} W[t*64 + j][s] Delta_s[i]  with
// M = 128, K = P0, N = nodes -- read once per 64 steps instead of every step.
// Its slots are settled one block before it is needed, so it is formed
// during the previous block in node slices of about one round of
// workgroups each (two per CU), one every 64 / slices steps: the cost is
// spread evenly over the steps instead of falling on the block's first.
// The reference forms every step's sum from scratch (fr_d_t_g_2d.cpp:
// 3340-3437).  W and Delta are split into two or three BF16 parts whose
// cross products accumulate in FP32 on v_mfma_f32_16x16x32_bf16.  The rest
// is summed in FP64: every 8 steps the settled slots from P0 on
// (k_block_mid, up to 120 slots), every step the F row, the mid row and at
// most 7 recent slots (k_block_near).
#include <cstring>

#include "frm_mfma.hpp"

namespace frm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kKc = 32;   // slots per MFMA k-chunk
constexpr int kNt = 128;  // nodes per workgroup (8 column tiles)
constexpr int kBpad = 40; // LDS row length of a node's 32 slots (bf16), padded
static_assert(kBlockSteps == 2 * kKc, "k_block_far takes the chunks of a block in pairs");
constexpr int kFarWgPerCu = 2;  // k_block_far workgroups resident per CU

// Workgroup: 8 wavefronts for a tile of 128 rows (2 targets x 64 steps) x
// 128 nodes.  Wavefronts 0-3 run the products: wavefront w owns row tiles
// r0, r0 + 2 (16 rows each, r0 = 0, 1, 4, 5) and all 8 column tiles; lane l
// holds A[l%16][8(l/16)+q], B[8(l/16)+q][l%16] and D[4(l/16)+r][l%16] of
// each 16x16x32 product.  Wavefronts 4-7 only stream the history: they keep
// two 32-slot chunks in flight in registers, split each into BF16 parts and
// store it into one of two LDS buffers, which the product wavefronts read;
// one barrier per chunk hands a buffer over.  With the stream and the
// products in different wavefronts neither waits for the other except at
// the barrier (measured on 2^20 nodes at depth 2048: 5.1-5.2 TB/s with 512
// workgroups per launch, one round at two per CU, against 4.4-4.5 for the
// earlier kernel, whose four wavefronts each loaded, split and multiplied;
// F is bit-identical).  The history stream sets the pace: the stream
// wavefronts' loads alone, without the split, the LDS stores or the
// products, take the whole kernel's time, and neither a third chunk in
// flight nor 16-byte loads change it; the same loads spread over all eight
// wavefronts stream about 15% faster.
//
// Operands are split into BF16 parts x = x0 + x1 (+ x2): NP = 3 forms
// x0 y0 + x0 y1 + x1 y0 (relative error ~2^-16); NP = 6 adds x0 y2 + x2 y0 +
// x1 y1 and takes each FP32 increment exactly (three 8-bit parts), leaving
// ~2^-24 -- the FP32 accumulation's own level.
//
// The history chunk (32 slots x 128 nodes, FP32) is loaded with
// non-temporal loads (the stream is read once; the weight table stays in
// L2): stream thread (g = t % 64, q = t / 64) reads slots 8q..8q+7 of nodes
// g and g + 64 (four bytes per lane, 256 contiguous bytes per wavefront
// load), splits them into BF16 parts and stores each node's 8 slots as one
// 16-byte LDS write into node-major rows [node][slot].  With 80-byte rows,
// the 64 nodes of one store fall on all 8 groups of 4 banks (a layout with
// 8-byte stores 4 rows apart spent 75% of the LDS cycles in bank conflicts).
// Loads are unguarded: K is a multiple of two chunks (block_far checks it),
// and a guarded load compiles to a select between global and private
// addresses and a flat load, which serialises the prefetch.
//
// The weight operand of a chunk depends only on d = N0 - s0 (Toeplitz): its
// fragments come ready-made from a table built once per solver (atab_build).
// Row tile rt + 2 of target t at lag base d equals row tile rt at d + 32
// (rows j and j + 32, same k), i.e. the fragment rt held one chunk earlier;
// so a product wavefront loads one fragment per chunk, a chunk ahead, and
// keeps the previous one.
template <int NP>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_block_far(
    const float* __restrict__ hist, int64_t ld, int N0, int K, const bf16x8* __restrict__ atab,
    float* __restrict__ F, int wg0) {
    constexpr int NS = NP == 6 ? 3 : 2;    // BF16 parts per operand
    constexpr int CT = kNt / 16;           // column tiles per workgroup
    static_assert(kNt == 128 && kKc == 32, "the stream map takes 2 x 64 nodes, 4 x 8 slots");
    __shared__ __attribute__((aligned(16))) __bf16 sB[2][NS][kNt][kBpad];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t i0 = (int64_t)(blockIdx.x + wg0) * kNt;
    const int nch = K / kKc;  // even
    if (wave >= 4) {
        // stream wavefronts: chunk c + 1 is split into buffer (c + 1) & 1
        // while the products run on chunk c; chunk c + 3 is then loaded
        const int t = tid - 256, ng = t & 63, sq = t >> 6;
        float x[8][2], y[8][2];
        auto load = [&](float (&d)[8][2], int c) {
            const float* q = hist + (int64_t)(c * kKc + 8 * sq) * ld + i0 + ng;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                d[u][0] = __builtin_nontemporal_load(q + (int64_t)u * ld);
                d[u][1] = __builtin_nontemporal_load(q + (int64_t)u * ld + 64);
            }
        };
        auto split = [&](float (&d)[8][2], int b) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float r[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) r[u] = d[u][h];
#pragma unroll
                for (int p = 0; p < NS; ++p) {
                    bf16x8 v;
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        v[u] = (__bf16)r[u];
                        r[u] -= (float)v[u];
                    }
                    *reinterpret_cast<bf16x8*>(&sB[b][p][ng + 64 * h][8 * sq]) = v;
                }
            }
        };
        load(x, 0);
        load(y, 1);
        split(x, 0);
        if (2 < nch) load(x, 2);
        __syncthreads();
        // the barriers pair with the product wavefronts' (same count)
        for (int c = 0; c < nch; c += 2) {
            split(y, 1);  // chunk c + 1
            if (c + 3 < nch) load(y, c + 3);
            __syncthreads();
            if (c + 2 < nch) {
                split(x, 0);  // chunk c + 2
                if (c + 4 < nch) load(x, c + 4);
            }
            __syncthreads();
        }
        return;
    }
    const int kg = lane >> 4, lc = lane & 15;
    const int r0 = (wave & 1) + 4 * (wave >> 1);  // row tiles r0, r0 + 2
    f32x4 acc[2][CT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < CT; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    // an: fragment of row tile r0 for the chunk at s0 (d = N0 - s0); acar:
    // the one of the chunk before, which is row tile r0 + 2's for this chunk
    bf16x8 an[NS], acar[NS];
    auto afrag = [&](bf16x8 (&dst)[NS], int d) {
        const bf16x8* at = atab + ((int64_t)(d / kKc - 1) * 8 + r0) * 3 * 64 + lane;
#pragma unroll
        for (int p = 0; p < NS; ++p) dst[p] = at[p * 64];
    };
    afrag(acar, N0 + kKc);  // row tile r0 + 2 of the first chunk
    afrag(an, N0);
    auto chunk = [&](int c, int b) {
        bf16x8 ap[2][NS];
#pragma unroll
        for (int p = 0; p < NS; ++p) {
            ap[0][p] = an[p];
            ap[1][p] = acar[p];
            acar[p] = an[p];
        }
        if (c + 1 < nch) afrag(an, N0 - (c + 1) * kKc);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            bf16x8 bp[NS];
#pragma unroll
            for (int p = 0; p < NS; ++p)
                bp[p] = *reinterpret_cast<const bf16x8*>(&sB[b][p][ct * 16 + lc][kg * 8]);
#pragma unroll
            for (int rr = 0; rr < 2; ++rr) {
                f32x4 cc = acc[rr][ct];
                if constexpr (NP == 6) {  // small terms first
                    cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rr][1], bp[1], cc, 0, 0, 0);
                    cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rr][0], bp[2], cc, 0, 0, 0);
                    cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rr][2], bp[0], cc, 0, 0, 0);
                }
                cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rr][0], bp[1], cc, 0, 0, 0);
                cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rr][1], bp[0], cc, 0, 0, 0);
                cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[rr][0], bp[0], cc, 0, 0, 0);
                acc[rr][ct] = cc;
            }
        }
    };
    __syncthreads();
    for (int c = 0; c < nch; c += 2) {
        chunk(c, 0);
        __syncthreads();
        chunk(c + 1, 1);
        __syncthreads();
    }
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int rt = r0 + 2 * rr;
        const int t = rt >> 2;
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = (rt & 3) * 16 + kg * 4 + r;
                F[(int64_t)(t * 64 + j) * ld + i0 + ct * 16 + lc] = acc[rr][ct][r];
            }
    }
}

// Mid part of a 64-step block, refreshed every kSubSteps steps: for the
// sub-block starting at N1 and the first slot P0 the far part leaves out
// (P0 < N1 <= P0 + 120),
//   G[t*8 + j][i] = sum_{P0 <= s < N1} w_t(N1 + j - s) Delta_s[i],  j < 8,
// summed in FP64 and stored in FP32 like the far part F (two nodes per
// thread, 8-byte loads of the FP32 slots; the weights are
// wavefront-uniform).  FP32 storage rounds G at 2^-24, below the far part's
// split-BF16 error; the FP64 store cost 64 more bytes per node and sub-block
// and 8 more per node and step on the reading side.  N1 - P0 is a multiple
// of 8: the slots go in batches of 8 whose loads are all issued before the
// first product, and a batch's 8 x 8 products touch only the 15 lags
// N1 - s0 - 7 .. N1 - s0 + 7.  Terms are added in increasing s, as a plain
// loop over s would add them.
__global__ __launch_bounds__(256) void k_block_mid(const float* __restrict__ hist, int64_t ld,
                                                   int N0, int N1, const double* __restrict__ w,
                                                   int ldw, float* __restrict__ G) {
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2;
    if (i >= ld) return;
    double g[2][kSubSteps][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int j = 0; j < kSubSteps; ++j) g[t][j][0] = g[t][j][1] = 0.0;
    for (int s0 = N0; s0 < N1; s0 += kSubSteps) {
        float2 x[kSubSteps];
#pragma unroll
        for (int u = 0; u < kSubSteps; ++u)
            x[u] = *reinterpret_cast<const float2*>(hist + (int64_t)(s0 + u) * ld + i);
        // wl[t][k] = w_t(L0 + k), L0 = N1 - s0 - 7 >= 1; lags past the
        // capacity belong to no step and read as 0
        const int L0 = N1 - s0 - (kSubSteps - 1);
        double wl[2][2 * kSubSteps - 1];
#pragma unroll
        for (int k = 0; k < 2 * kSubSteps - 1; ++k) {
            const bool ok = L0 + k <= ldw;
            const int l = ok ? L0 + k : ldw;  // in bounds either way: unconditional loads
            const double v0 = w[l - 1], v1 = w[ldw + l - 1];
            wl[0][k] = ok ? v0 : 0.0;
            wl[1][k] = ok ? v1 : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kSubSteps; ++u) {
            const double x0 = (double)x[u].x, x1 = (double)x[u].y;
#pragma unroll
            for (int j = 0; j < kSubSteps; ++j) {
                // lag of step N1 + j from slot s0 + u
                const int k = kSubSteps - 1 - u + j;
                g[0][j][0] += wl[0][k] * x0;
                g[0][j][1] += wl[0][k] * x1;
                g[1][j][0] += wl[1][k] * x0;
                g[1][j][1] += wl[1][k] * x1;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int j = 0; j < kSubSteps; ++j)
            *reinterpret_cast<float2*>(G + (int64_t)(t * kSubSteps + j) * ld + i) =
                make_float2((float)g[t][j][0], (float)g[t][j][1]);
}

// part[t*ld + i] = F[t*64 + n - N0][i] + G[t*8 + n - N1][i]
//                  + sum_{N1 <= s < n} w_t(n - s) Delta_s[i],
// four nodes per thread (16-byte loads of the FP32 slots), FP64 sums.  The
// slot count CNT = n - N1 (< 8) is a template parameter, so the CNT slot
// loads, the F row and the G row are all in flight together.
template <int CNT>
__device__ __forceinline__ void near_part(const float* __restrict__ hist, int64_t ld, int64_t i,
                                          int N0, int N1, int n, const double* __restrict__ w,
                                          int ldw, const float* __restrict__ F, int use_F,
                                          const float* __restrict__ G, int use_G,
                                          double* __restrict__ part) {
    float4 x[CNT > 0 ? CNT : 1];
#pragma unroll
    for (int u = 0; u < CNT; ++u)
        x[u] = *reinterpret_cast<const float4*>(hist + (int64_t)(N1 + u) * ld + i);
    double s0[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0};
    if (use_F) {
        const int j = n - N0;
        const float4 f0 = *reinterpret_cast<const float4*>(F + (int64_t)j * ld + i);
        const float4 f1 = *reinterpret_cast<const float4*>(F + (int64_t)(kBlockSteps + j) * ld + i);
        s0[0] = f0.x, s0[1] = f0.y, s0[2] = f0.z, s0[3] = f0.w;
        s1[0] = f1.x, s1[1] = f1.y, s1[2] = f1.z, s1[3] = f1.w;
    }
    if (use_G) {
        const int j = n - N1;
        const float4 g0 = *reinterpret_cast<const float4*>(G + (int64_t)j * ld + i);
        const float4 g1 = *reinterpret_cast<const float4*>(G + (int64_t)(kSubSteps + j) * ld + i);
        s0[0] += g0.x, s0[1] += g0.y, s0[2] += g0.z, s0[3] += g0.w;
        s1[0] += g1.x, s1[1] += g1.y, s1[2] += g1.z, s1[3] += g1.w;
    }
#pragma unroll
    for (int u = 0; u < CNT; ++u) {
        const int lag = CNT - u;  // slot N1 + u, step n = N1 + CNT
        const double a = w[lag - 1], b = w[ldw + lag - 1];
        s0[0] += a * (double)x[u].x, s1[0] += b * (double)x[u].x;
        s0[1] += a * (double)x[u].y, s1[1] += b * (double)x[u].y;
        s0[2] += a * (double)x[u].z, s1[2] += b * (double)x[u].z;
        s0[3] += a * (double)x[u].w, s1[3] += b * (double)x[u].w;
    }
    *reinterpret_cast<double2*>(part + i) = make_double2(s0[0], s0[1]);
    *reinterpret_cast<double2*>(part + i + 2) = make_double2(s0[2], s0[3]);
    *reinterpret_cast<double2*>(part + ld + i) = make_double2(s1[0], s1[1]);
    *reinterpret_cast<double2*>(part + ld + i + 2) = make_double2(s1[2], s1[3]);
}

__global__ __launch_bounds__(256) void k_block_near(const float* __restrict__ hist, int64_t ld,
                                                    int N0, int N1, int n,
                                                    const double* __restrict__ w, int ldw,
                                                    const float* __restrict__ F, int use_F,
                                                    const float* __restrict__ G, int use_G,
                                                    double* __restrict__ part) {
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= ld) return;
#define FRM_NEAR_CASE(c) \
    case c: near_part<c>(hist, ld, i, N0, N1, n, w, ldw, F, use_F, G, use_G, part); break;
    switch (n - N1) {
        FRM_NEAR_CASE(0) FRM_NEAR_CASE(1) FRM_NEAR_CASE(2) FRM_NEAR_CASE(3)
        FRM_NEAR_CASE(4) FRM_NEAR_CASE(5) FRM_NEAR_CASE(6) FRM_NEAR_CASE(7)
        default: break;
    }
#undef FRM_NEAR_CASE
}

int block_far(hipStream_t st, const float* hist, int64_t ld, int N0, int K, const void* atab,
              int atab_chunks, int nprod, float* F, int wg0, int nwg) {
    // the table must reach d = N0 + 32 (the carried fragment of the first
    // chunk); the last chunk's lag base is N0 - K + 32 >= 32
    if (ld % kNt || N0 <= 0 || N0 % kBlockSteps || K <= 0 || K % kBlockSteps || K > N0 ||
        N0 / kKc + 1 > atab_chunks || (nprod != 3 && nprod != 6) || wg0 < 0 || nwg <= 0 ||
        (int64_t)wg0 + nwg > ld / kNt)
        return -1;
    const dim3 grid((unsigned)nwg), block(512);
    const bf16x8* a = (const bf16x8*)atab;
    if (nprod == 6)
        hipLaunchKernelGGL(k_block_far<6>, grid, block, 0, st, hist, ld, N0, K, a, F, wg0);
    else
        hipLaunchKernelGGL(k_block_far<3>, grid, block, 0, st, hist, ld, N0, K, a, F, wg0);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int block_far_groups(int64_t ld) { return (int)(ld / kNt); }

int block_far_round(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        return -1;
    return kFarWgPerCu * cus;
}

// Host: the weight-operand fragments for d = 32, 64, ..., 32*chunks.  Entry
// [(d/32 - 1)*8 + rt][p][lane] holds, for row tile rt (target t = rt/4, steps
// j = 16(rt%4) + lane%16) and k = 8(lane/16) + q, the BF16 part p of
// W_t(d + j - k) (zero outside lags 1..ldw); W parts given as floats
// wparts[(p*2 + t)*ldw + lag - 1].
std::vector<uint16_t> atab_build(const float* wparts, int ldw, int chunks) {
    std::vector<uint16_t> out((size_t)chunks * 8 * 3 * 64 * 8);
    for (int c = 0; c < chunks; ++c) {
        const int d = (c + 1) * kKc;
        for (int rt = 0; rt < 8; ++rt)
            for (int p = 0; p < 3; ++p)
                for (int lane = 0; lane < 64; ++lane)
                    for (int q = 0; q < 8; ++q) {
                        const int t = rt >> 2, j = (rt & 3) * 16 + (lane & 15);
                        const int k = (lane >> 4) * 8 + q;
                        const int lag = d + j - k;
                        float v = 0.f;
                        if (lag >= 1 && lag <= ldw) v = wparts[((size_t)p * 2 + t) * ldw + lag - 1];
                        uint32_t u;
                        std::memcpy(&u, &v, 4);  // v is BF16-exact: the top half is its bits
                        out[((((size_t)c * 8 + rt) * 3 + p) * 64 + lane) * 8 + q] = (uint16_t)(u >> 16);
                    }
    }
    return out;
}

int block_mid(hipStream_t st, const float* hist, int64_t ld, int N0, int N1, const double* w,
              int ldw, float* G) {
    if (ld % 2 || N0 < 0 || N1 <= N0 || N1 - N0 >= 2 * kBlockSteps || (N1 - N0) % kSubSteps) return -1;
    const int64_t threads = ld / 2;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_block_mid, grid, block, 0, st, hist, ld, N0, N1, w, ldw, G);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int block_near(hipStream_t st, const float* hist, int64_t ld, int N0, int N1, int n,
               const double* w, int ldw, const float* F, int use_F, const float* G, int use_G,
               double* part) {
    if (ld % 4 || n < N1 || N1 < N0 || n - N1 >= kSubSteps || n - N0 >= kBlockSteps) return -1;
    const int64_t threads = ld / 4;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    hipLaunchKernelGGL(k_block_near, grid, block, 0, st, hist, ld, N0, N1, n, w, ldw, F, use_F, G,
                       use_G, part);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace frm
