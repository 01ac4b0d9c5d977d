// Loads and stores for arrays that a kernel streams once per step.
//
// A plane set larger than about half of the 256 MB last-level cache is
// evicted before the next step reads it again, so caching it only displaces
// other data: such streams are read and written non-temporally.  A smaller
// set stays cached between steps and keeps ordinary accesses (measured on
// the SOE mode planes: 67 MB at 512^2 ran 1.4x slower non-temporal, 268 MB
// at 1024^2 and 10-20 GB at 39M nodes ran 3-12% faster).
#pragma once
#include <hip/hip_runtime.h>

namespace frm {

constexpr double kStreamBypassBytes = 128.0 * 1024 * 1024;

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
