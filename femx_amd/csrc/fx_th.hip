// Taylor-Hood P2/P1 Navier-Stokes: single-instance kernels and their C-ABI
// entry points.  The device code and its design notes are in fx_th_dev.hpp.
#include "fx_th_dev.hpp"

#include <atomic>
#include <mutex>
#include <string>

namespace
{
using namespace fxk;

// kThWaves waves per workgroup, one row block each, consecutive row blocks
// per XCD; each wave has a row tile of `tile` doubles after the tables.
// At most 256 VGPRs (2 waves per SIMD): a register count past 256 would
// halve the occupancy, which costs far more than a few spills.
__global__ __launch_bounds__(kThLanes* kThWaves) __attribute__((amdgpu_waves_per_eu(2))) void k_th_assemble(ThAsmArgs a, int num_blocks,
                                                                    int tile)
{
  extern __shared__ double smem[];
  ThTensor& T = *reinterpret_cast<ThTensor*>(smem);
  th_stage_tensor(T, kThLanes * kThWaves);
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kThLanes); // uniform per wave
  const int b = fx::xcd_block(gridDim.x) * kThWaves + w;
  if (b < num_blocks) th_assemble_block(b, a, T, smem + kThTensorLen + w * tile);
}

// The same over the blocks' round tables: each wave walks a few blocks of its
// XCD's share (xcd_walk), its pipeline running on across block boundaries.
// Blocks per wave at 512^2 (A/B against fx_th_assemble's 290 us): all the
// device holds at once (~20) 292 us, 1: 284, 2: 264, 4: 263, 8: 278; at
// 256^2 (about 5 blocks per resident wave in all) 4 gained nothing (88 vs
// 87 us): the last walks' tail is as long as the rest.  With the next
// block's data loaded one block ahead, longer walks pay: at 512^2 a cap of a
// fifth of each resident wave's share (4 blocks) 247 us, a third (6) 233 us.
constexpr int kThRtBlocksPerWave = 8;

__global__ __launch_bounds__(kThLanes* kThWaves) __attribute__((amdgpu_waves_per_eu(2))) void k_th_assemble_rt(ThAsmArgs a, int num_blocks,
                                                                       int tile)
{
  extern __shared__ double smem[];
  ThTensor& T = *reinterpret_cast<ThTensor*>(smem);
  th_stage_tensor(T, kThLanes * kThWaves);
  __syncthreads();
  const int         w  = __builtin_amdgcn_readfirstlane(threadIdx.x / kThLanes);
  const fx::XcdWalk wk = fx::xcd_walk(num_blocks, kThWaves, w);
  th_tile_clear(smem + kThTensorLen + w * tile, tile);
  th_assemble_walk_rt(a, T, smem + kThTensorLen + w * tile, wk.first, wk.end, wk.stride);
}

static_assert(sizeof(ThTensor) + std::size_t(FX_TH_TILE_NNZ_MAX) * 8 * kThWaves <= 160 * 1024,
              "the largest Taylor-Hood row tile must fit the workgroup's LDS");

__global__ __launch_bounds__(kThLanes) void k_th_vjp_elem(ThVjpArgs a)
{
  th_vjp_elem_block(fx::xcd_block(gridDim.x), a, kThTables);
}

__global__ __launch_bounds__(256) void k_vjp_gather(GatherArgs a)
{
  __shared__ double sval[kGatherStage];
  vjp_gather_block(fx::xcd_block(gridDim.x), a, sval);
}

static_assert(kPatchLanes == FX_TH_PATCH_MAX, "one lane per element of a patch");

__global__ __launch_bounds__(kPatchLanes) __attribute__((amdgpu_waves_per_eu(4))) void k_th_vjp_patch(ThVjpArgs a, PatchArgs q,
                                                                                           int max_patch_ldofs)
{
  extern __shared__ __align__(16) unsigned char lds[];
  th_vjp_patch_block(fx::xcd_block(gridDim.x), a, q, kThTables, lds, max_patch_ldofs);
}

__global__ __launch_bounds__(256) void k_th_vjp_seam(SeamArgs a)
{
  patch_seam_block(fx::xcd_block(gridDim.x), a);
}

// The two VJP phases with the element-column workspace `work` (15 doubles
// per element); shared by fx_th_vjp (caller workspace) and
// fx_th_vjp_noalloc (the library's fixed device workspace).
void th_vjp_launch(fx_index ndof, fx_index num_elems, const double* xv, const double* yv,
                   const fx_index* vconn3, const fx_index* dofs15, const double* state, double nu,
                   const double* nu_e, const double* lam, const fx_index* bc_mask,
                   const fx_index* inc_off, const fx_index* elem_inc, double* jt, double* g,
                   double* work, void* stream, const char* who)
{
  const std::string w(who);
  fx::require(ndof >= 0 && num_elems >= 0, FX_ERR_ARG, (w + ": negative size").c_str());
  if (ndof == 0) return;
  // state, lam and bc_mask are gathered through 32-bit byte offsets
  fx::require(fx::fits32(ndof, 8) && fx::fits32(num_elems, 8)
                  && fx::fits32(15 * static_cast<std::int64_t>(num_elems), 8),
              FX_ERR_OVERFLOW, (w + ": state, element or workspace arrays beyond 4 GiB").c_str());
  fx::require(xv && yv && state && lam && inc_off && jt
                  && (num_elems == 0 || (vconn3 && dofs15 && elem_inc && work)),
              FX_ERR_ARG, (w + ": null array").c_str());
  const std::size_t nd = std::size_t(ndof) * 8, ne = std::size_t(num_elems) * 8;
  fx::require(!fx::overlaps(jt, nd, lam, nd) && !fx::overlaps(jt, nd, state, nd)
                  && !(g && fx::overlaps(g, ne, jt, nd)) && !fx::overlaps(work, 15 * ne, jt, nd)
                  && !fx::overlaps(work, 15 * ne, lam, nd)
                  && !(g && fx::overlaps(work, 15 * ne, g, ne)),
              FX_ERR_ALIAS, (w + ": outputs must not alias inputs or each other").c_str());
  if (num_elems > 0)
  {
    const ThVjpArgs a{num_elems, xv,      yv,       vconn3, dofs15, state, nu,
                      nu_e,      lam,     bc_mask,  elem_inc, work, g};
    hipLaunchKernelGGL(k_th_vjp_elem, dim3(fx::blocks_for(num_elems, kThLanes)), dim3(kThLanes), 0,
                       fx::as_stream(stream), a);
    fx::check_launch(who);
  }
  const GatherArgs ga{ndof, inc_off, work, lam, bc_mask, jt};
  hipLaunchKernelGGL(k_vjp_gather, dim3(fx::blocks_for(ndof, 256)), dim3(256), 0,
                     fx::as_stream(stream), ga);
  fx::check_launch(who);
}

// Fixed device workspace of fx_th_vjp_noalloc: 15 doubles per element for
// up to FX_TH_VJP_NOALLOC_MAX_ELEMS elements (64 MiB), one copy per GPU
// held until the process ends.  In host code the array's name is not a
// device address: kernels get the per-device address that
// hipGetSymbolAddress returns, resolved once per device on first use.
constexpr std::size_t kNoAllocWorkDoubles = 15 * static_cast<std::size_t>(FX_TH_VJP_NOALLOC_MAX_ELEMS);
static __device__ __align__(256) double scratch[kNoAllocWorkDoubles];

constexpr int                     kMaxDevices = 64;
std::mutex                        g_scratch_mutex;
std::atomic<double*>              g_scratch_ptr[kMaxDevices];

double* scratch_on_current_device()
{
  int dev = 0;
  fx::check_hip(hipGetDevice(&dev), "fx_th_vjp_noalloc: hipGetDevice");
  fx::require(dev >= 0 && dev < kMaxDevices, FX_ERR_UNSUPPORTED,
              "fx_th_vjp_noalloc: device ordinal beyond 63");
  double* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (p) return p;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
  if (!p)
  {
    void* addr = nullptr;
    fx::check_hip(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)),
                  "fx_th_vjp_noalloc: hipGetSymbolAddress(scratch)");
    fx::require(addr != nullptr, FX_ERR_HIP, "fx_th_vjp_noalloc: workspace has no device address");
    p = static_cast<double*>(addr);
    g_scratch_ptr[dev].store(p, std::memory_order_release);
  }
  return p;
}

} // namespace

extern "C" {

int fx_th_assemble(fx_index ndof, fx_index num_elems, fx_index num_blocks, fx_index max_block_nnz,
                   const double* xv, const double* yv, const fx_index* vconn3,
                   const fx_index* dofs15, const double* state, double nu, const double* nu_e,
                   const fx_index* inc_off, const fx_index* inc_elem, const fx_index* inc_meta,
                   const fx_index* blk_row, const fx_index* blk_flags, const fx_index* blk_lanes,
                   const fx_index* row_ptr, double* vals, double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(ndof >= 0 && num_elems >= 0 && num_blocks >= 0 && max_block_nnz >= 0, FX_ERR_ARG,
                "fx_th_assemble: negative size");
    if (ndof == 0 || num_blocks == 0) return;
    fx::require(fx::fits32(ndof, 8) && fx::fits32(num_elems, 8)
                    && fx::fits32(15 * static_cast<std::int64_t>(num_elems), 16),
                FX_ERR_OVERFLOW, "fx_th_assemble: state, element or incidence arrays beyond 4 GiB");
    fx::require(xv && yv && state && inc_off && blk_row && blk_flags && row_ptr && res
                    && (num_elems == 0 || (vconn3 && dofs15 && inc_elem && inc_meta)),
                FX_ERR_ARG, "fx_th_assemble: null array");
    fx::require(!fx::overlaps(res, std::size_t(ndof) * 8, state, std::size_t(ndof) * 8)
                    && !(vals && fx::overlaps(vals, 8, res, std::size_t(ndof) * 8)),
                FX_ERR_ALIAS, "Assembly state, residual, and matrix values must not alias");
    const int    tile = vals ? max_block_nnz : 0;
    const size_t lds  = sizeof(ThTensor) + static_cast<size_t>(tile) * 8 * kThWaves;
    fx::require(lds <= 160 * 1024, FX_ERR_UNSUPPORTED, "fx_th_assemble: row tile exceeds LDS");
    const ThAsmArgs a{xv,      yv,       vconn3,   dofs15,   state,
                      nu,      nu_e,     inc_off,  inc_elem, reinterpret_cast<const uint4*>(inc_meta),
                      blk_row, blk_flags, blk_lanes, row_ptr, vals, res};
    hipLaunchKernelGGL(k_th_assemble, dim3(fx::blocks_for(num_blocks, kThWaves)),
                       dim3(kThLanes * kThWaves), lds, fx::as_stream(stream), a, num_blocks, tile);
    fx::check_launch("fx_th_assemble");
  });
}

int fx_th_assemble_rounds(fx_index ndof, fx_index num_elems, fx_index num_blocks,
                          fx_index max_block_nnz, const double* xv, const double* yv,
                          const fx_index* vconn3, const fx_index* dofs15, const double* state,
                          double nu, const double* nu_e, const fx_index* blk_row,
                          const fx_index* blk_flags, const fx_index* blk_lanes,
                          const fx_index* rt_off, const fx_index* rt_elem,
                          const fx_index* rt_meta, const fx_index* row_ptr, double* vals,
                          double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(ndof >= 0 && num_elems >= 0 && num_blocks >= 0 && max_block_nnz >= 0, FX_ERR_ARG,
                "fx_th_assemble_rounds: negative size");
    if (ndof == 0 || num_blocks == 0) return;
    fx::require(fx::fits32(ndof, 8) && fx::fits32(num_elems, 8)
                    && fx::fits32(15 * static_cast<std::int64_t>(num_elems), 16),
                FX_ERR_OVERFLOW,
                "fx_th_assemble_rounds: state, element or incidence arrays beyond 4 GiB");
    fx::require(xv && yv && state && blk_row && blk_flags && blk_lanes && rt_off && row_ptr && res
                    && (num_elems == 0 || (vconn3 && dofs15 && rt_elem && rt_meta)),
                FX_ERR_ARG, "fx_th_assemble_rounds: null array");
    fx::require(!fx::overlaps(res, std::size_t(ndof) * 8, state, std::size_t(ndof) * 8)
                    && !(vals && fx::overlaps(vals, 8, res, std::size_t(ndof) * 8)),
                FX_ERR_ALIAS, "Assembly state, residual, and matrix values must not alias");
    const int    tile = vals ? max_block_nnz : 0;
    const size_t lds  = sizeof(ThTensor) + static_cast<size_t>(tile) * 8 * kThWaves;
    fx::require(lds <= 160 * 1024, FX_ERR_UNSUPPORTED,
                "fx_th_assemble_rounds: row tile exceeds LDS");
    const ThAsmArgs a{xv,       yv,        vconn3,    dofs15,  state,   nu,
                      nu_e,     nullptr,   nullptr,   nullptr, blk_row, blk_flags,
                      blk_lanes, row_ptr,  vals,      res,     rt_off,  rt_elem,
                      reinterpret_cast<const uint4*>(rt_meta), static_cast<int>(num_blocks), 1};
    // each wave walks up to kThRtBlocksPerWave blocks: long enough for the
    // pipeline to run on across block boundaries, short enough that the
    // hardware still balances the walks: at most a third of what each
    // resident wave takes in all (the resident waves: the CUs times the
    // workgroups that fit, at most two by the kernel's 2 waves/SIMD, by LDS),
    // so that the last walks' tail stays short.  A multiple of 8 workgroups,
    // for xcd_walk.
    int dev = 0, cus = 0;
    fx::check_hip(hipGetDevice(&dev), "hipGetDevice");
    fx::check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                  "hipDeviceGetAttribute");
    const std::int64_t resident =
        static_cast<std::int64_t>(std::max(cus, 1)) * kThWaves
        * std::max<std::int64_t>(1, std::min<std::int64_t>(2, (160 * 1024) / static_cast<std::int64_t>(lds)));
    const std::int64_t bpw = std::max<std::int64_t>(
        1, std::min<std::int64_t>(kThRtBlocksPerWave, num_blocks / (3 * std::max<std::int64_t>(resident, 1))));
    const std::int64_t need = (static_cast<std::int64_t>(num_blocks) + kThWaves * bpw - 1) / (kThWaves * bpw);
    const int grid = static_cast<int>((need + 7) / 8 * 8);
    hipLaunchKernelGGL(k_th_assemble_rt, dim3(grid), dim3(kThLanes * kThWaves), lds,
                       fx::as_stream(stream), a, num_blocks, tile);
    fx::check_launch("fx_th_assemble_rounds");
  });
}

int fx_th_vjp(fx_index ndof, fx_index num_elems, const double* xv, const double* yv,
              const fx_index* vconn3, const fx_index* dofs15, const double* state, double nu,
              const double* nu_e, const double* lam, const fx_index* bc_mask,
              const fx_index* inc_off, const fx_index* elem_inc, double* jt, double* g,
              double* work, void* stream)
{
  return fx::guard([&] {
    th_vjp_launch(ndof, num_elems, xv, yv, vconn3, dofs15, state, nu, nu_e, lam, bc_mask, inc_off,
                  elem_inc, jt, g, work, stream, "fx_th_vjp");
  });
}

int fx_th_vjp_patch(fx_index ndof, fx_index num_elems, const double* xv, const double* yv,
                    const double* state, double nu,
                    const double* nu_e, const double* lam, const fx_index* bc_mask,
                    const fx_patch_plan* plan, double* jt, double* g, void* stream)
{
  return fx::guard([&] {
    fx::require(ndof >= 0 && num_elems >= 0 && plan, FX_ERR_ARG,
                "fx_th_vjp_patch: negative size or null plan");
    const fx_patch_plan& q = *plan;
    fx::require(q.num_patches >= 0 && q.num_seam >= 0 && q.num_parts >= 0
                    && q.max_patch_ldofs >= 0 && q.max_patch_ldofs <= 15 * kPatchLanes
                    && (q.num_patches > 0) == (num_elems > 0),
                FX_ERR_ARG, "fx_th_vjp_patch: plan sizes do not fit the mesh");
    if (ndof == 0) return;
    fx::require(fx::fits32(ndof, 8) && fx::fits32(num_elems, 8) && fx::fits32(q.num_parts, 8),
                FX_ERR_OVERFLOW, "fx_th_vjp_patch: state, element or partial arrays beyond 4 GiB");
    fx::require(xv && yv && state && lam && jt && q.patch_off && q.ldof_off && q.seam_off
                    && (num_elems == 0
                        || (q.patch_elem && q.elem_vert && q.elem_ldof
                            && q.ldof_dof && q.ldof_part && q.ldof_end))
                    && (q.num_seam == 0 || q.seam_dof) && (q.num_parts == 0 || q.parts),
                FX_ERR_ARG, "fx_th_vjp_patch: null array");
    const std::size_t nd = std::size_t(ndof) * 8, ne = std::size_t(num_elems) * 8,
                      np = std::size_t(q.num_parts) * 8;
    fx::require(!fx::overlaps(jt, nd, lam, nd) && !fx::overlaps(jt, nd, state, nd)
                    && !(g && fx::overlaps(g, ne, jt, nd)) && !fx::overlaps(q.parts, np, jt, nd)
                    && !fx::overlaps(q.parts, np, lam, nd)
                    && !(g && fx::overlaps(q.parts, np, g, ne)),
                FX_ERR_ALIAS, "fx_th_vjp_patch: outputs must not alias inputs or each other");
    if (q.num_patches > 0)
    {
      const ThVjpArgs   a{num_elems, xv, yv, nullptr, nullptr, state, nu, nu_e,
                          lam,       bc_mask, nullptr, nullptr, g};
      const PatchArgs pa{q.patch_off, q.patch_elem, q.elem_vert, q.elem_ldof, q.ldof_off,
                           q.ldof_dof,  q.ldof_part,  q.ldof_end,  jt,          q.parts};
      const int         mx = static_cast<int>(q.max_patch_ldofs);
      hipLaunchKernelGGL(k_th_vjp_patch, dim3(static_cast<unsigned>(q.num_patches)),
                         dim3(kPatchLanes), th_patch_lds(mx), fx::as_stream(stream), a, pa, mx);
      fx::check_launch("fx_th_vjp_patch");
    }
    if (q.num_seam > 0)
    {
      const SeamArgs sa{q.num_seam, q.seam_off, q.seam_dof, q.parts, lam, bc_mask, jt};
      hipLaunchKernelGGL(k_th_vjp_seam, dim3(fx::blocks_for(q.num_seam, 256)), dim3(256), 0,
                         fx::as_stream(stream), sa);
      fx::check_launch("fx_th_vjp_patch");
    }
  });
}

int fx_th_vjp_noalloc(fx_index ndof, fx_index num_elems, const double* xv, const double* yv,
                      const fx_index* vconn3, const fx_index* dofs15, const double* state,
                      double nu, const double* nu_e, const double* lam, const fx_index* bc_mask,
                      const fx_index* inc_off, const fx_index* elem_inc, double* jt, double* g,
                      void* stream)
{
  return fx::guard([&] {
    fx::require(num_elems >= 0 && num_elems <= FX_TH_VJP_NOALLOC_MAX_ELEMS, FX_ERR_ARG,
                "fx_th_vjp_noalloc: num_elems exceeds FX_TH_VJP_NOALLOC_MAX_ELEMS (the fixed "
                "64 MiB device workspace holds 15 doubles per element); use fx_th_vjp with a "
                "caller workspace");
    if (ndof == 0) return;
    th_vjp_launch(ndof, num_elems, xv, yv, vconn3, dofs15, state, nu, nu_e, lam, bc_mask, inc_off,
                  elem_inc, jt, g, scratch_on_current_device(), stream, "fx_th_vjp_noalloc");
  });
}

} // extern "C"
