// Device helpers shared by the assembly kernels (gfx950).
#pragma once
#include "fx_common.hpp"

namespace fxk
{

// Gathers through a 32-bit byte offset from a wave-uniform base: the base
// stays in SGPRs and each lane supplies one VGPR offset (the saddr form of
// global_load), so a gather costs one shift instead of 64-bit address
// arithmetic.  Every array read this way is under 4 GiB (the entry points
// check the sizes).
template <class T>
__device__ __forceinline__ T gat32(const T* __restrict__ base, int i)
{
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base)
                                     + static_cast<unsigned>(i) * unsigned(sizeof(T)));
}

// LDS floating-point atomic add whose result is not used (ds_add_f64): the
// LDS unit performs old + v, rounded as a VALU add, and the issuing wave does
// not wait for it.  One wave's LDS operations execute in issue order, so a
// wave's adds to one entry happen in program order.
__device__ __forceinline__ void lds_add(double* p, double v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Copy n consecutive values of a stream read once (non-temporal loads) into
// LDS, NT threads cooperating: each thread has four loads in flight before
// their LDS stores, instead of one load, wait and store per step.
template <int NT, class D, class S>
__device__ __forceinline__ void stage_nt(D* __restrict__ dst, const S* __restrict__ src, int n)
{
  for (int i = threadIdx.x; i < n; i += 4 * NT)
  {
    S v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = i + u * NT < n ? __builtin_nontemporal_load(src + i + u * NT) : S(0);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i + u * NT < n) dst[i + u * NT] = static_cast<D>(v[u]);
  }
}

} // namespace fxk
