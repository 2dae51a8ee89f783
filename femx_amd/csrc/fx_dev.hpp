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

} // namespace fxk
