// History store and memory-term kernels.
//
// The reference keeps one heap snapshot per (half-)step and, per node, loops
// over all of them: mt_minimal.cpp:906-939 (H1), mt_psi_caputo.cpp:1315-1342
// (H2), fr_d_t_g_2d.cpp:3340-3437 (H3, CPU) and the OpenCL Om kernel at
// fr_d_t_g_2d.cpp:1704-1840 (one work-item per node, 32-wide LDS weight
// tiles).  Here the store holds the increments Delta_s = C_{s+1} - C_s
// snapshot-major (slot s holds all nodes, so a wavefront reading one slot
// reads 64 consecutive nodes), and each lane sums its nodes over all slots
// for one or two weight vectors (the two LOD half-step targets share the
// pass).  When there are too few nodes to fill the GPU the slots are split
// into time chunks with one FP64 partial each, which the consumer adds in
// chunk order (deterministic).  The weights are uniform across a wavefront,
// so they come through the scalar cache.
#include <algorithm>
#include <mutex>

#include "frm_history.hpp"

namespace frm {

// Partial-sum workspace of frm_memterm_static: a file-scope device array of
// kStaticScratchDoubles FP64 values (16 MiB), one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself (on the host that
// name is a stand-in, not a device address): static_scratch_setup looks up
// each device's address once, with that device current, and the entry
// point hands the kernels that address.
constexpr int64_t kStaticScratchDoubles = int64_t(1) << 21;
static __device__ __align__(256) double scratch[kStaticScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];

int static_scratch_setup(frm_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return set_error(ctx, FRM_E_INVALID, "static scratch: device %d beyond %d", ctx->device,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        double* p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return check_hip(ctx, e, "hipGetSymbolAddress(scratch)");
        if (!p) return set_error(ctx, FRM_E_HIP, "hipGetSymbolAddress(scratch): null address");
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->static_scratch = g_scratch_ptr[ctx->device];
    return FRM_OK;
}

// ---------------------------------------------------------------- kernels
// part[(c*NT + t)*ld + i] = sum_{s in chunk c} w[t*ldw + depth-1-s] * Delta_s[i]
//
// Each lane owns two adjacent nodes and streams their increments slot after
// slot with 16-byte non-temporal loads (the history is read once per step
// and never reused before it is evicted, so it should not displace the
// state, coefficients and partials from the caches), four slots in flight.
// Measured on MI355X: one pass per node column over all slots, with every
// wavefront walking the same slot region at the same time, keeps DRAM rows
// open and sustains ~7.0 TB/s; splitting the slots into chunks that
// different workgroups read concurrently scatters the accesses and drops to
// 5.1-5.5 TB/s.  The split is therefore used only when the node count alone
// cannot fill the GPU.
typedef double frm_dv2 __attribute__((ext_vector_type(2)));

template <int NT>
__global__ __launch_bounds__(kMemtermThreads) void k_memterm_f64(
    const double* __restrict__ hist, int64_t ld, int depth, int first, const double* __restrict__ w,
    int ldw, int chunk, double* __restrict__ part) {
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2;
    if (i >= ld) return;
    const int c = blockIdx.y;
    const int s0 = first + c * chunk;
    const int s1 = min(depth, s0 + chunk);
    double a0[NT], a1[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) a0[t] = a1[t] = 0.0;
    const int64_t ldv = ld / 2;
    const frm_dv2* p = reinterpret_cast<const frm_dv2*>(hist + (int64_t)s0 * ld + i);
    // lag-1 index of slot s is depth-1-s; walk the weights backwards
    const double* wp = w + (depth - 1 - s0);
    int s = s0;
    for (; s + 4 <= s1; s += 4) {
        frm_dv2 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = __builtin_nontemporal_load(p + (int64_t)u * ldv);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double wt = wp[t * ldw - u];
                a0[t] = fma(wt, x[u].x, a0[t]);
                a1[t] = fma(wt, x[u].y, a1[t]);
            }
        }
        p += 4 * ldv;
        wp -= 4;
    }
    for (; s < s1; ++s) {
        const frm_dv2 x = __builtin_nontemporal_load(p);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const double wt = wp[t * ldw];
            a0[t] = fma(wt, x.x, a0[t]);
            a1[t] = fma(wt, x.y, a1[t]);
        }
        p += ldv;
        wp -= 1;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        double* o = part + ((int64_t)c * NT + t) * ld + i;
        o[0] = a0[t];
        o[1] = a1[t];
    }
}

// FP32 increments and weights, one target: each lane's terms go through a
// Kahan-compensated FP32 sum, returned in FP64 per chunk, eight slots in
// flight.  A lane owns NPL = 2 or 4 adjacent nodes (8- or 16-byte loads).
// Measured on MI355X at depth 1000: 2 nodes per lane is best up to ~0.8M
// nodes (7.2 TB/s at 262144, where 4 per lane leaves too few wavefronts),
// 4 per lane from ~1M nodes (6.0 vs 5.6 TB/s at 1048576).
constexpr int64_t kF32WideLd = 1 << 20;

__device__ inline void kahan_add(float& s, float& c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// NT targets (weights w[t*ldw + lag - 1]); part[(c*NT + t)*ld + i].
template <int NPL, int NT>
__global__ __launch_bounds__(kMemtermThreads) void k_memterm_f32c(
    const float* __restrict__ hist, int64_t ld, int depth, int first, const float* __restrict__ w,
    int ldw, int chunk, double* __restrict__ part) {
    typedef float vt __attribute__((ext_vector_type(NPL)));
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * NPL;
    if (i >= ld) return;
    const int c = blockIdx.y;
    const int s0 = first + c * chunk;
    const int s1 = min(depth, s0 + chunk);
    float sm[NT][NPL], cm[NT][NPL];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int q = 0; q < NPL; ++q) sm[t][q] = cm[t][q] = 0.f;
    const int64_t ldv = ld / NPL;
    const vt* p = reinterpret_cast<const vt*>(hist + (int64_t)s0 * ld + i);
    const float* wp = w + (depth - 1 - s0);
    int s = s0;
    for (; s + 8 <= s1; s += 8) {
        vt x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = __builtin_nontemporal_load(p + (int64_t)u * ldv);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const float wt = wp[t * ldw - u];
#pragma unroll
                for (int q = 0; q < NPL; ++q) kahan_add(sm[t][q], cm[t][q], wt * x[u][q]);
            }
        }
        p += 8 * ldv;
        wp -= 8;
    }
    for (; s < s1; ++s) {
        const vt x = __builtin_nontemporal_load(p);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const float wt = wp[t * ldw];
#pragma unroll
            for (int q = 0; q < NPL; ++q) kahan_add(sm[t][q], cm[t][q], wt * x[q]);
        }
        p += ldv;
        wp -= 1;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        double* o = part + ((int64_t)c * NT + t) * ld + i;
#pragma unroll
        for (int q = 0; q < NPL; ++q) o[q] = (double)sm[t][q] - (double)cm[t][q];
    }
}

// out[t*n + i] = scale * sum_c part[(c*NT+t)*ld + i] (+ out)
__global__ void k_memterm_reduce(const double* __restrict__ part, int64_t ld, int nchunks,
                                 int nt, int64_t n, double scale, int accumulate,
                                 double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (i >= n) return;
    double acc = 0.0;
    for (int c = 0; c < nchunks; ++c) acc += part[((int64_t)c * nt + t) * ld + i];
    double v = scale * acc;
    if (accumulate) v += out[(int64_t)t * n + i];
    out[(int64_t)t * n + i] = v;
}

__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                      double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}

// ---------------------------------------------------------------- host side
int hist_alloc(frm_ctx* ctx, int64_t n_nodes, int32_t capacity, int32_t precision,
               frm_hist** out) {
    FRM_REQUIRE(ctx, ctx && out && n_nodes > 0 && capacity > 0, "frm_hist_create: bad arguments");
    FRM_REQUIRE(ctx,
                precision == FRM_PREC_F64 || precision == FRM_PREC_MFMA_BF16 ||
                    precision == FRM_PREC_F32C,
                "frm_hist_create: precision %d not supported", precision);
    // FP64 increments, or FP32 increments for the blocked MFMA and the
    // compensated FP32 forms
    const size_t esz = precision == FRM_PREC_F64 ? sizeof(double) : sizeof(float);
    frm_hist* h = new frm_hist();
    h->ctx = ctx;
    h->n_nodes = n_nodes;
    h->ld = (n_nodes + kLdAlign - 1) / kLdAlign * kLdAlign;
    h->capacity = capacity;
    h->precision = precision;
    h->max_chunks = kMaxChunks;
    size_t bytes = (size_t)capacity * h->ld * esz;
    hipError_t e = hipMalloc(&h->data, bytes);
    if (e != hipSuccess) {
        delete h;
        return check_hip(ctx, e, "hipMalloc(history store)");
    }
    e = hipMalloc(&h->part, (size_t)h->max_chunks * 2 * h->ld * sizeof(double));
    if (e != hipSuccess) {
        (void)hipFree(h->data);
        delete h;
        return check_hip(ctx, e, "hipMalloc(partials)");
    }
    // zero once so padded lanes read zeros, not stale memory
    e = hipMemsetAsync(h->data, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(h->data);
        (void)hipFree(h->part);
        delete h;
        return check_hip(ctx, e, "hipMemset(history store)");
    }
    *out = h;
    return FRM_OK;
}

// Number of time chunks for a launch: one, unless the nodes alone give fewer
// than kMemtermWaves wavefronts; chunks never get shorter than kMinChunk.
static int pick_chunks(const frm_hist* h, int depth) {
    if (depth <= 0) return 0;
    const int64_t waves = h->ld / (2 * kWave);
    int nch = 1;
    if (waves < kMemtermWaves) nch = (int)((kMemtermWaves + waves - 1) / waves);
    nch = std::min<int>(nch, (depth + kMinChunk - 1) / kMinChunk);
    return std::max(1, std::min(nch, h->max_chunks));
}

int memterm_chunks(const frm_hist* h, int first) {
    const int depth = h->depth;
    first = std::max(0, std::min(first, depth));
    return pick_chunks(h, depth - first);
}

int memterm_partials(frm_ctx* ctx, frm_hist* h, int ntargets, const double* d_w, int ldw,
                     int* nchunks_out, int first, const float* d_w32, double* part) {
    const int depth = h->depth;
    first = std::max(0, std::min(first, depth));
    const int span = depth - first;  // slots first .. depth-1, lags depth-first .. 1
    const int nchunks = pick_chunks(h, span);
    *nchunks_out = nchunks;
    if (nchunks == 0) return FRM_OK;
    // a workspace given by the caller: no timing events (recording them may
    // create events)
    frm_ctx* tctx = part ? nullptr : ctx;
    if (!part) part = h->part;
    const int chunk = (span + nchunks - 1) / nchunks;
    const int threads = kMemtermThreads;
    dim3 grid((unsigned)(h->ld / 2 / threads), (unsigned)nchunks);
    if (h->precision == FRM_PREC_F32C) {
        FRM_REQUIRE(ctx, (ntargets == 1 || ntargets == 2) && d_w32,
                    "memterm_partials: FP32 store takes one or two targets and FP32 weights");
        LaunchTimer timer(tctx, FRM_K_MEMTERM,
                          (double)span * h->n_nodes * sizeof(float) +
                              (double)ntargets * h->n_nodes * sizeof(double));
        const float* hf = (const float*)h->data;
        if (h->ld >= kF32WideLd) {
            // ld is a multiple of 4 * 256: the last workgroup may be half idle
            dim3 g4((unsigned)((h->ld / 4 + threads - 1) / threads), (unsigned)nchunks);
            if (ntargets == 1)
                hipLaunchKernelGGL((k_memterm_f32c<4, 1>), g4, dim3(threads), 0, ctx->stream, hf,
                                   h->ld, depth, first, d_w32, ldw, chunk, part);
            else
                hipLaunchKernelGGL((k_memterm_f32c<4, 2>), g4, dim3(threads), 0, ctx->stream, hf,
                                   h->ld, depth, first, d_w32, ldw, chunk, part);
        } else {
            if (ntargets == 1)
                hipLaunchKernelGGL((k_memterm_f32c<2, 1>), grid, dim3(threads), 0, ctx->stream, hf,
                                   h->ld, depth, first, d_w32, ldw, chunk, part);
            else
                hipLaunchKernelGGL((k_memterm_f32c<2, 2>), grid, dim3(threads), 0, ctx->stream, hf,
                                   h->ld, depth, first, d_w32, ldw, chunk, part);
        }
        FRM_HIP(ctx, hipGetLastError());
        return FRM_OK;
    }
    FRM_REQUIRE(ctx, h->precision == FRM_PREC_F64, "memterm_partials: precision %d",
                h->precision);
    LaunchTimer timer(tctx, FRM_K_MEMTERM,
                      (double)span * h->n_nodes * sizeof(double) +
                          (double)ntargets * h->n_nodes * sizeof(double));
    if (ntargets == 1)
        hipLaunchKernelGGL(k_memterm_f64<1>, grid, dim3(threads), 0, ctx->stream,
                           (const double*)h->data, h->ld, depth, first, d_w, ldw, chunk, part);
    else
        hipLaunchKernelGGL(k_memterm_f64<2>, grid, dim3(threads), 0, ctx->stream,
                           (const double*)h->data, h->ld, depth, first, d_w, ldw, chunk, part);
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}

int memterm_reduce1(frm_ctx* ctx, const frm_hist* h, int nchunks, double* d_out) {
    const int threads = 256;
    hipLaunchKernelGGL(k_memterm_reduce, dim3((unsigned)((h->n_nodes + threads - 1) / threads)),
                       dim3(threads), 0, ctx->stream, (const double*)h->part, h->ld, nchunks, 1,
                       h->n_nodes, 1.0, 0, d_out);
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}

}  // namespace frm

using namespace frm;

extern "C" {

int frm_hist_create(frm_ctx* ctx, int64_t n_nodes, int32_t capacity, int32_t precision,
                    frm_hist** out) {
    FRM_REQUIRE(ctx, precision == FRM_PREC_F64, "frm_hist_create: precision %d not supported",
                precision);
    return hist_alloc(ctx, n_nodes, capacity, precision, out);
}

int frm_hist_destroy(frm_hist* h) {
    if (!h) return FRM_OK;
    (void)hipStreamSynchronize(h->ctx->stream);
    (void)hipFree(h->data);
    (void)hipFree(h->part);
    delete h;
    return FRM_OK;
}

int frm_hist_depth(const frm_hist* h, int32_t* depth) {
    if (!h || !depth) return set_error(nullptr, FRM_E_INVALID, "frm_hist_depth: null argument");
    *depth = h->depth;
    return FRM_OK;
}

int frm_hist_clear(frm_hist* h) {
    if (!h) return set_error(nullptr, FRM_E_INVALID, "frm_hist_clear: null argument");
    h->depth = 0;
    return FRM_OK;
}

int frm_hist_push(frm_ctx* ctx, frm_hist* h, const double* d_new, const double* d_old) {
    FRM_REQUIRE(ctx, ctx && h && d_new && d_old, "frm_hist_push: null argument");
    if (h->depth >= h->capacity)
        return set_error(ctx, FRM_E_RANGE, "frm_hist_push: capacity %d reached", h->capacity);
    double* slot = (double*)h->data + (int64_t)h->depth * h->ld;
    const int threads = 256;
    hipLaunchKernelGGL(k_sub, dim3((unsigned)((h->n_nodes + threads - 1) / threads)),
                       dim3(threads), 0, ctx->stream, d_new, d_old, h->n_nodes, slot);
    FRM_HIP(ctx, hipGetLastError());
    h->depth += 1;
    return FRM_OK;
}

int frm_memterm(frm_ctx* ctx, frm_hist* h, const frm_memterm_desc* desc, double* d_out) {
    FRM_REQUIRE(ctx, ctx && h && desc && d_out, "frm_memterm: null argument");
    FRM_REQUIRE(ctx, desc->ntargets == 1 || desc->ntargets == 2, "frm_memterm: ntargets %d",
                desc->ntargets);
    FRM_REQUIRE(ctx, h->depth == 0 || (desc->d_w && desc->ldw >= h->depth),
                "frm_memterm: weights shorter than the history depth");
    int nchunks = 0;
    int rc = memterm_partials(ctx, h, desc->ntargets, desc->d_w, desc->ldw, &nchunks);
    if (rc) return rc;
    const int threads = 256;
    dim3 grid((unsigned)((h->n_nodes + threads - 1) / threads), (unsigned)desc->ntargets);
    hipLaunchKernelGGL(k_memterm_reduce, grid, dim3(threads), 0, ctx->stream,
                       (const double*)h->part, h->ld, nchunks, desc->ntargets, h->n_nodes,
                       desc->scale, desc->accumulate, d_out);
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}

// frm_memterm with the partial sums in this device's static array instead of
// the store's workspace; allocates nothing and records no timing events.
// Every call on one GPU shares the array (each call writes every partial it
// reads before reading it, so nothing needs clearing): calls that overlap
// on one GPU, from different streams or contexts, race.
int frm_memterm_static(frm_ctx* ctx, frm_hist* h, const frm_memterm_desc* desc, double* d_out) {
    FRM_REQUIRE(ctx, ctx && h && desc && d_out, "frm_memterm_static: null argument");
    FRM_REQUIRE(ctx, desc->ntargets == 1 || desc->ntargets == 2, "frm_memterm_static: ntargets %d",
                desc->ntargets);
    FRM_REQUIRE(ctx, h->depth == 0 || (desc->d_w && desc->ldw >= h->depth),
                "frm_memterm_static: weights shorter than the history depth");
    FRM_REQUIRE(ctx, h->precision == FRM_PREC_F64, "frm_memterm_static: precision %d",
                h->precision);
    if (!ctx->static_scratch)
        return set_error(ctx, FRM_E_STATE, "frm_memterm_static: no static workspace on device %d",
                         ctx->device);
    const int nch = memterm_chunks(h);
    const int64_t need = (int64_t)nch * desc->ntargets * h->ld;
    if (need > kStaticScratchDoubles)
        return set_error(ctx, FRM_E_RANGE,
                         "frm_memterm_static: %d time chunks x %d targets x %lld padded nodes need "
                         "%lld partial sums; the static workspace holds %lld (16 MiB): use "
                         "frm_memterm",
                         nch, desc->ntargets, (long long)h->ld, (long long)need,
                         (long long)kStaticScratchDoubles);
    int nchunks = 0;
    int rc = memterm_partials(ctx, h, desc->ntargets, desc->d_w, desc->ldw, &nchunks, 0, nullptr,
                              ctx->static_scratch);
    if (rc) return rc;
    const int threads = 256;
    dim3 grid((unsigned)((h->n_nodes + threads - 1) / threads), (unsigned)desc->ntargets);
    hipLaunchKernelGGL(k_memterm_reduce, grid, dim3(threads), 0, ctx->stream,
                       (const double*)ctx->static_scratch, h->ld, nchunks, desc->ntargets,
                       h->n_nodes, desc->scale, desc->accumulate, d_out);
    FRM_HIP(ctx, hipGetLastError());
    return FRM_OK;
}

}  // extern "C"
