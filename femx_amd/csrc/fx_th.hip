// Taylor-Hood P2/P1 Navier-Stokes: single-instance kernels and their C-ABI
// entry points.  The device code and its design notes are in fx_th_dev.hpp.
#include "fx_th_dev.hpp"

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

__global__ __launch_bounds__(kThLanes) void k_th_vjp_elem(ThVjpArgs a)
{
  th_vjp_elem_block(fx::xcd_block(gridDim.x), a, kThTables);
}

__global__ __launch_bounds__(256) void k_vjp_gather(GatherArgs a)
{
  __shared__ double sval[kGatherStage];
  vjp_gather_block(fx::xcd_block(gridDim.x), a, sval);
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

int fx_th_vjp(fx_index ndof, fx_index num_elems, const double* xv, const double* yv,
              const fx_index* vconn3, const fx_index* dofs15, const double* state, double nu,
              const double* nu_e, const double* lam, const fx_index* bc_mask,
              const fx_index* inc_off, const fx_index* elem_inc, double* jt, double* g,
              double* work, void* stream)
{
  return fx::guard([&] {
    fx::require(ndof >= 0 && num_elems >= 0, FX_ERR_ARG, "fx_th_vjp: negative size");
    if (ndof == 0) return;
    // state, lam and bc_mask are gathered through 32-bit byte offsets
    fx::require(fx::fits32(ndof, 8) && fx::fits32(num_elems, 8)
                    && fx::fits32(15 * static_cast<std::int64_t>(num_elems), 8),
                FX_ERR_OVERFLOW, "fx_th_vjp: state, element or workspace arrays beyond 4 GiB");
    fx::require(xv && yv && state && lam && inc_off && jt
                    && (num_elems == 0 || (vconn3 && dofs15 && elem_inc && work)),
                FX_ERR_ARG, "fx_th_vjp: null array");
    const std::size_t nd = std::size_t(ndof) * 8, ne = std::size_t(num_elems) * 8;
    fx::require(!fx::overlaps(jt, nd, lam, nd) && !fx::overlaps(jt, nd, state, nd)
                    && !(g && fx::overlaps(g, ne, jt, nd)) && !fx::overlaps(work, 15 * ne, jt, nd)
                    && !fx::overlaps(work, 15 * ne, lam, nd) && !(g && fx::overlaps(work, 15 * ne, g, ne)),
                FX_ERR_ALIAS, "fx_th_vjp: outputs must not alias inputs or each other");
    if (num_elems > 0)
    {
      const ThVjpArgs a{num_elems, xv,      yv,       vconn3, dofs15, state, nu,
                        nu_e,      lam,     bc_mask,  elem_inc, work, g};
      hipLaunchKernelGGL(k_th_vjp_elem, dim3(fx::blocks_for(num_elems, kThLanes)), dim3(kThLanes), 0,
                         fx::as_stream(stream), a);
      fx::check_launch("fx_th_vjp");
    }
    const GatherArgs ga{ndof, inc_off, work, lam, bc_mask, jt};
    hipLaunchKernelGGL(k_vjp_gather, dim3(fx::blocks_for(ndof, 256)), dim3(256), 0,
                       fx::as_stream(stream), ga);
    fx::check_launch("fx_th_vjp");
  });
}

} // extern "C"
