/* frm_diag.h -- diagnostics of libfrm with no counterpart in the reference.
 *
 * frm_diag_copy runs a plain device-to-device copy kernel (one 16-byte load
 * and store per lane) `reps` times on the context's stream and reports the
 * average kernel time from HIP events: the measured achievable HBM
 * bandwidth the memory-term roofline is compared against, and the known-size
 * kernel that scales rocprof FETCH_SIZE / WRITE_SIZE counters.
 *
 * frm_diag_read streams `bytes` (a multiple of 8 MiB) of device memory as
 * rows of 8 MiB, each lane summing one 4-byte column over all rows with
 * non-temporal loads, 16 rows in flight, and writing its sum once (the
 * fastest read pattern measured on MI355X): the measured read bandwidth,
 * which the read-dominated kernels are compared against.  It allocates its
 * own buffer of per-lane sums (8 MiB).
 *
 * frm_diag_psi_kb_quad runs the device restatement of the psi-Caputo
 * weights' quadrature fallback, _kb(1, t0, t1, gtj, sing, alpha)
 * (mt_psi_caputo.cpp:476-494, which calc_int calls where a weight series is
 * not finite, 736-737), on n host-given cases of six doubles each (alpha,
 * beta, t0, t1, gtj, sing; g = t^beta), one GPU thread per case, and
 * returns the n values to the host: the check of that code path against the
 * reference's own values.  Requires 0 <= alpha < 1, beta > 0 and finite
 * t0 <= t1, gtj. */
#ifndef FRM_DIAG_H
#define FRM_DIAG_H
#include <stddef.h>

#include "frm.h"

#ifdef __cplusplus
extern "C" {
#endif

int frm_diag_copy(frm_ctx* ctx, void* d_dst, const void* d_src, size_t bytes, int reps,
                  double* avg_ms);
int frm_diag_read(frm_ctx* ctx, const void* d_src, size_t bytes, int reps, double* avg_ms);
int frm_diag_psi_kb_quad(frm_ctx* ctx, int32_t n, const double* h_cases, double* h_out);

#ifdef __cplusplus
}
#endif
#endif
