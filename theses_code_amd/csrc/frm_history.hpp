// History store shared between the memory-term and solver translation units.
#pragma once
#include "frm_internal.hpp"

struct frm_hist {
    frm_ctx* ctx = nullptr;
    int64_t n_nodes = 0;
    int64_t ld = 0;  // slot stride in elements (multiple of kLdAlign)
    int32_t capacity = 0;
    int32_t depth = 0;
    int32_t precision = FRM_PREC_F64;
    int32_t max_chunks = 0;  // partial-sum workspace size (time chunks)
    void* data = nullptr;    // capacity * ld elements
    double* part = nullptr;  // max_chunks * 2 * ld partial sums
};

namespace frm {

constexpr int kMemtermThreads = 512;
constexpr int64_t kLdAlign = 2 * kMemtermThreads;  // whole workgroups of 2 nodes per lane
constexpr int64_t kMemtermWaves = 2048;  // wavefronts that keep HBM busy on MI355X
constexpr int32_t kMinChunk = 64;        // shortest time chunk worth a workgroup
constexpr int32_t kMaxChunks = 64;

int hist_alloc(frm_ctx* ctx, int64_t n_nodes, int32_t capacity, int32_t precision,
               frm_hist** out);
// Launches the split-K partial sums for the current depth into h->part,
// over slots first .. depth-1 (the newest depth - first lags); returns the
// number of chunks written (0 when there is nothing to sum).  An FP32 store
// (FRM_PREC_F32C) takes one target and its FP32 weights in d_w32.  With
// `part` given, the partials go there instead (nchunks * ntargets * ld
// doubles, see memterm_chunks) and no timing events are recorded.
int memterm_partials(frm_ctx* ctx, frm_hist* h, int ntargets, const double* d_w, int ldw,
                     int* nchunks, int first = 0, const float* d_w32 = nullptr,
                     double* part = nullptr);
// Time chunks memterm_partials splits slots first .. depth-1 into.
int memterm_chunks(const frm_hist* h, int first = 0);
// out[i] = sum_c part[c*ld + i] for one target (chunk order, deterministic)
int memterm_reduce1(frm_ctx* ctx, const frm_hist* h, int nchunks, double* d_out);

}  // namespace frm
