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

// Order the LDS accesses of one wave before and after this point (a wave's
// LDS instructions execute in issue order; this keeps the compiler from
// moving them across).
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane i takes lane i+1's v and lane 63 takes fill (DPP wave_shl:1 on both
// halves; every lane of the wave must be active).
__device__ __forceinline__ double wave_shl1(double v, double fill)
{
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x130, 0xf,
                                             0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x130, 0xf,
                                             0xf, false);
  return __hiloint2double(hi, lo);
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


// ---------------------------------------------------------------------------
// Single-pass VJPs over element patches (fx_patches_create): one workgroup of
// kPatchLanes lanes per patch, one element per lane; each local DOF's
// element contributions are summed in LDS in element order, and DOFs on
// seams between patches get a partial per patch, summed by the seam pass.
constexpr int kPatchLanes = 256;

// Local DOFs a lane handles per round of the staging and summing loops: their
// loads are all issued before the first is used.
constexpr int kPatchUnroll = 8;

struct PatchArgs
{
  const int*            patch_off;
  const int*            patch_elem;
  const int*            elem_vert; // 3 per element in patch order
  const unsigned short* elem_ldof; // per element in patch order: local DOF | run rank << 12
  const int*            ldof_off;
  const int*            ldof_dof;
  const int*            ldof_part;
  const unsigned short* ldof_end;
  double*               jt;
  double*               parts;
};

struct SeamArgs
{
  int           num_seam;
  const int*    seam_off;
  const int*    seam_dof;
  const double* parts;
  const double* lam;
  const int*    bc_mask;
  double*       jt;
};

// Seam pass (one lane per seam DOF): the DOF's per-patch partials in patch
// order (lam's pass-through is in the first), the first four loaded at once;
// a DOF in no element gets lam's pass-through alone.
__device__ __forceinline__ void patch_seam_block(int blk, const SeamArgs& a)
{
  const int s = blk * 256 + threadIdx.x;
  if (s >= a.num_seam) return;
  const int t0 = a.seam_off[s], t1 = a.seam_off[s + 1], d = a.seam_dof[s];
  const int n  = t1 - t0;
  double    p[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) p[u] = n > 0 ? __builtin_nontemporal_load(a.parts + t0 + min(u, n - 1)) : 0.0;
  double v = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (u < n) v = v + p[u];
  for (int t = t0 + 4; t < t1; ++t) v = v + a.parts[t];
  if (n == 0 && a.bc_mask && a.bc_mask[d]) v = a.lam[d];
  __builtin_nontemporal_store(v, a.jt + d);
}

// One local DOF's run summed in slot order, then lam's pass-through on a
// constrained DOF -- added by the patch that finishes the DOF, or into the
// partial of its first patch -- and the result stored (jt or the partial).
__device__ __forceinline__ void patch_finish(const double* cs, int r0, int r1, int d, int part,
                                             bool constrained, const double* lam, double* jt,
                                             double* parts)
{
  double v = 0.0;
  for (int s = r0; s < r1; ++s) v = v + cs[s];
  if (constrained && (part < 0 || (part & FX_TH_PART_FIRST))) v = v + lam[d];
  if (part < 0)
    __builtin_nontemporal_store(v, jt + d);
  else
    __builtin_nontemporal_store(v, parts + (part & ~FX_TH_PART_FIRST));
}

} // namespace fxk
