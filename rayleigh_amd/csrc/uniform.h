// Wave-uniform copies of values loaded from memory.  Work items and problem
// descriptors are read from device memory with wave-uniform addresses, but values
// loaded through a loaded pointer are not always recognised as uniform; buffer
// resources and scalar offsets built from them would then be wrapped in
// read-first-lane loops.  Passing them through v_readfirstlane makes them scalar.
#pragma once

#include <hip/hip_runtime.h>

namespace raysht {

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ long uniform(long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (long)(((unsigned long)hi << 32) | lo);
}

template <typename T>
__device__ __forceinline__ T *uniform(T *p)
{
    return reinterpret_cast<T *>(uniform(reinterpret_cast<long>(p)));
}

}  // namespace raysht
