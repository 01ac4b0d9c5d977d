// Loads and stores for arrays that a kernel streams once per step.
//
// A plane set that does not fit in the 256 MB last-level cache next to the
// other streams of a step is evicted before the next step reads it again, so
// caching it only displaces other data: such streams are read and written
// non-temporally.  A smaller set stays cached between steps and keeps
// ordinary accesses.  Measured: the SOE mode planes at 67 MB ran 1.4x slower
// non-temporal, the 142 MB moments of the series time sums 1.16x slower;
// 268 MB of SOE planes and 10-20 GB at 39M nodes ran 3-12% faster.
#pragma once
#include <hip/hip_runtime.h>

namespace frm {

constexpr double kStreamBypassBytes = 200.0 * 1024 * 1024;

template <bool NT, typename T>
__device__ __forceinline__ T stream_load(const T* p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool NT, typename T>
__device__ __forceinline__ void stream_store(T v, T* p) {
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

}  // namespace frm
