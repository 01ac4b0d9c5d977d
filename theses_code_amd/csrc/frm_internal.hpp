// Internal definitions shared by the HIP translation units of libfrm.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "frm.h"

namespace frm {

constexpr int kWave = 64;  // CDNA wavefront width

struct TimingSlot {
    hipEvent_t start, stop;
    int kclass;
    double bytes;
};

}  // namespace frm

struct frm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    uint32_t timing_mask = 0;
    std::vector<frm::TimingSlot> pending;  // recorded, not yet harvested
    std::vector<hipEvent_t> event_pool;
    int64_t launches[FRM_K_COUNT] = {};
    double total_ms[FRM_K_COUNT] = {};
    double total_bytes[FRM_K_COUNT] = {};
    double* static_scratch = nullptr;  // this device's copy of the static partial-sum array
    // second stream (default priority) for work that overlaps the context's
    // stream, ordered against it by events; created on first use
    hipStream_t aux_stream = nullptr;
};

namespace frm {

int set_error(frm_ctx* ctx, int code, const char* fmt, ...);
// Looks up this device's address of the static partial-sum array of
// frm_memterm_static (once per device; the device must be current) and
// stores it in ctx->static_scratch.
int static_scratch_setup(frm_ctx* ctx);
int check_hip(frm_ctx* ctx, hipError_t e, const char* what);

// Brackets one kernel launch of class `kclass` with timing events when
// enabled for that class.
struct LaunchTimer {
    frm_ctx* ctx;
    int kclass;
    bool on;
    hipEvent_t a{}, b{};
    LaunchTimer(frm_ctx* c, int k, double bytes);
    ~LaunchTimer();
    double bytes_;
};

}  // namespace frm

#define FRM_HIP(ctx, call)                                           \
    do {                                                             \
        hipError_t _e = (call);                                      \
        if (_e != hipSuccess) return frm::check_hip((ctx), _e, #call); \
    } while (0)

#define FRM_REQUIRE(ctx, cond, ...)                                   \
    do {                                                              \
        if (!(cond)) return frm::set_error((ctx), FRM_E_INVALID, __VA_ARGS__); \
    } while (0)
