// Runtime layer of libfrm: contexts, device buffers, transfers and per-kernel
// event timing.  Replaces the OpenCL/CUDA wrapper of opencl_cuda_base.cpp
// (OpenCL_program, OpenCL_commandqueue, OpenCL_buffer; NVRTC twin at
// 456-822) with one HIP stream per context and no per-launch host sync.
#include <cstdlib>
#include <cstring>

#include "frm_internal.hpp"

namespace {
thread_local std::string g_err;  // errors raised before a context exists
}

namespace frm {

int set_error(frm_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->last_error = buf;
    else
        g_err = buf;
    return code;
}

int check_hip(frm_ctx* ctx, hipError_t e, const char* what) {
    return set_error(ctx, e == hipErrorOutOfMemory ? FRM_E_NOMEM : FRM_E_HIP, "%s: %s", what,
                     hipGetErrorString(e));
}

static hipEvent_t take_event(frm_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e{};
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

LaunchTimer::LaunchTimer(frm_ctx* c, int k, double bytes)
    : ctx(c), kclass(k), on(false), bytes_(bytes) {
    if (ctx && (ctx->timing_mask & (1u << k))) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (a && b && hipEventRecord(a, ctx->stream) == hipSuccess) on = true;
    }
}

LaunchTimer::~LaunchTimer() {
    if (!on) return;
    if (hipEventRecord(b, ctx->stream) == hipSuccess) ctx->pending.push_back({a, b, kclass, bytes_});
}

static void harvest(frm_ctx* ctx) {
    for (auto& s : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s.start, s.stop) == hipSuccess) {
            ctx->launches[s.kclass] += 1;
            ctx->total_ms[s.kclass] += ms;
            ctx->total_bytes[s.kclass] += s.bytes;
        }
        ctx->event_pool.push_back(s.start);
        ctx->event_pool.push_back(s.stop);
    }
    ctx->pending.clear();
}

}  // namespace frm

extern "C" {

const char* frm_strerror(int status) {
    switch (status) {
        case FRM_OK: return "ok";
        case FRM_E_INVALID: return "invalid argument";
        case FRM_E_HIP: return "HIP runtime error";
        case FRM_E_NOMEM: return "device out of memory";
        case FRM_E_RANGE: return "capacity or size limit exceeded";
        case FRM_E_STATE: return "invalid state";
        default: return "unknown status";
    }
}

const char* frm_last_error(const frm_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : g_err.c_str();
}

int frm_ctx_create(int device, frm_ctx** out) {
    if (!out) return frm::set_error(nullptr, FRM_E_INVALID, "frm_ctx_create: null out");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return frm::check_hip(nullptr, e, "hipGetDeviceCount");
    if (device < 0 || device >= ndev)
        return frm::set_error(nullptr, FRM_E_INVALID, "device %d not present (%d devices)", device,
                              ndev);
    e = hipSetDevice(device);
    if (e != hipSuccess) return frm::check_hip(nullptr, e, "hipSetDevice");
    frm_ctx* ctx = new frm_ctx();
    ctx->device = device;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        int rc = frm::check_hip(nullptr, e, "hipStreamCreate");
        delete ctx;
        return rc;
    }
    // this device's address of frm_memterm_static's array (device current)
    int rc = frm::static_scratch_setup(ctx);
    if (rc) {
        (void)frm::set_error(nullptr, rc, "frm_ctx_create: %s", ctx->last_error.c_str());
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return FRM_OK;
}

int frm_ctx_destroy(frm_ctx* ctx) {
    if (!ctx) return FRM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->aux_stream) (void)hipStreamSynchronize(ctx->aux_stream);
    frm::harvest(ctx);
    for (auto ev : ctx->event_pool) (void)hipEventDestroy(ev);
    if (ctx->aux_stream) (void)hipStreamDestroy(ctx->aux_stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FRM_OK;
}

int frm_sync(frm_ctx* ctx) {
    if (!ctx) return frm::set_error(nullptr, FRM_E_INVALID, "frm_sync: null ctx");
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    frm::harvest(ctx);
    return FRM_OK;
}

int frm_malloc(frm_ctx* ctx, size_t bytes, void** dptr) {
    FRM_REQUIRE(ctx, ctx && dptr, "frm_malloc: null argument");
    FRM_HIP(ctx, hipSetDevice(ctx->device));
    FRM_HIP(ctx, hipMalloc(dptr, bytes ? bytes : 1));
    return FRM_OK;
}

int frm_free(frm_ctx* ctx, void* dptr) {
    FRM_REQUIRE(ctx, ctx, "frm_free: null ctx");
    if (!dptr) return FRM_OK;
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FRM_HIP(ctx, hipFree(dptr));
    return FRM_OK;
}

int frm_upload(frm_ctx* ctx, void* dptr, const void* hptr, size_t bytes) {
    FRM_REQUIRE(ctx, ctx && (bytes == 0 || (dptr && hptr)), "frm_upload: null argument");
    FRM_HIP(ctx, hipMemcpyAsync(dptr, hptr, bytes, hipMemcpyHostToDevice, ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FRM_OK;
}

int frm_download(frm_ctx* ctx, void* hptr, const void* dptr, size_t bytes) {
    FRM_REQUIRE(ctx, ctx && (bytes == 0 || (dptr && hptr)), "frm_download: null argument");
    FRM_HIP(ctx, hipMemcpyAsync(hptr, dptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FRM_OK;
}

int frm_timing_enable(frm_ctx* ctx, uint32_t mask) {
    FRM_REQUIRE(ctx, ctx, "frm_timing_enable: null ctx");
    ctx->timing_mask = mask;
    return FRM_OK;
}

int frm_timing_reset(frm_ctx* ctx) {
    FRM_REQUIRE(ctx, ctx, "frm_timing_reset: null ctx");
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    frm::harvest(ctx);
    for (int k = 0; k < FRM_K_COUNT; ++k) {
        ctx->launches[k] = 0;
        ctx->total_ms[k] = 0;
        ctx->total_bytes[k] = 0;
    }
    return FRM_OK;
}

int frm_timing_get(frm_ctx* ctx, int kclass, int64_t* launches, double* total_ms,
                   double* bytes) {
    FRM_REQUIRE(ctx, ctx && kclass >= 0 && kclass < FRM_K_COUNT, "frm_timing_get: bad class");
    FRM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    frm::harvest(ctx);
    if (launches) *launches = ctx->launches[kclass];
    if (total_ms) *total_ms = ctx->total_ms[kclass];
    if (bytes) *bytes = ctx->total_bytes[kclass];
    return FRM_OK;
}

}  // extern "C"
