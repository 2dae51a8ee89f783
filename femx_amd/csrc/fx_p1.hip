// P1 Poisson assembly and log-kappa VJP: single-instance kernels and their
// C-ABI entry points.  The device code and its design notes are in
// fx_p1_dev.hpp.
#include "fx_p1_dev.hpp"

#include <algorithm>

namespace
{
using namespace fxk;

template <bool kLong>
__global__ __launch_bounds__(kP1Block) void k_p1_assemble(P1AsmArgs a)
{
  extern __shared__ double smem[];
  p1_assemble_block<kLong>(fx::xcd_block(gridDim.x), a, smem);
}

// capped at 96 VGPRs: 5 waves/SIMD instead of 4, no spills (74.6 -> 69.1 us at 1024^2)
__global__ __launch_bounds__(kP1Block) __attribute__((amdgpu_waves_per_eu(5))) void k_p1_vjp(P1VjpArgs a)
{
  __shared__ long long sinc[kIncStage];
  p1_vjp_block(fx::xcd_block(gridDim.x), a, sinc);
}

__global__ __launch_bounds__(kPatchLanes) void k_p1_vjp_patch(P1VjpArgs a, PatchArgs q,
                                                              int max_patch_ldofs)
{
  extern __shared__ __align__(16) unsigned char lds[];
  p1_vjp_patch_block(fx::xcd_block(gridDim.x), a, q, lds, max_patch_ldofs);
}

__global__ __launch_bounds__(256) void k_p1_vjp_seam(SeamArgs a)
{
  patch_seam_block(fx::xcd_block(gridDim.x), a);
}

__global__ __launch_bounds__(kPatchLanes) void k_p1_asm_patch(P1AsmArgs a, PatchArgs q, P1AsmPatch s,
                                                              int max_patch_ldofs, int max_tile,
                                                              int max_runs)
{
  extern __shared__ __align__(16) unsigned char lds[];
  p1_asm_patch_block(fx::xcd_block(gridDim.x), a, q, s, lds, max_patch_ldofs, max_tile, max_runs);
}

__global__ __launch_bounds__(kP1Block) void k_p1_assemble_list(P1AsmArgs a, const int* rows, int nrows)
{
  __shared__ double smem[kP1Block * kP1ListStride];
  p1_assemble_list_block(fx::xcd_block(gridDim.x), a, rows, nrows, smem);
}

__global__ __launch_bounds__(64 * kP1StWaves) void k_p1_asm_stencil(P1AsmArgs a,
                                                                      const int* blk_code,
                                                                      const int* words)
{
  __shared__ double bb[kP1StWaves][4 * kP1StBand];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  p1_assemble_stencil_block(fx::xcd_block(gridDim.x) * kP1StWaves + wv, a, blk_code, words, bb[wv]);
}

} // namespace

extern "C" {

int fx_p1_assemble_stencil(fx_index num_nodes, fx_index num_elems, fx_index nnz,
                           fx_index max_row_len, const double* x, const double* y,
                           const fx_index* conn, const double* kappa, int kappa_is_log,
                           const double* f, const double* u, const fx_index* inc_off,
                           const fx_index* inc, const fx_index* row_ptr, const fx_index* blk_code,
                           const fx_index* words, double* vals, double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_elems >= 0 && nnz >= 0, FX_ERR_ARG,
                "fx_p1_assemble_stencil: negative size");
    if (num_nodes == 0) return;
    fx::require(x && y && inc_off && row_ptr && res && blk_code && words
                    && (num_elems == 0 || (conn && inc)),
                FX_ERR_ARG, "fx_p1_assemble_stencil: null array");
    fx::require(fx::fits32(num_nodes, 8) && fx::fits32(num_elems, 8), FX_ERR_OVERFLOW,
                "fx_p1_assemble_stencil: node or element arrays beyond 4 GiB");
    fx::require(max_row_len >= 0 && max_row_len <= 255, FX_ERR_UNSUPPORTED,
                "fx_p1_assemble_stencil: CSR rows longer than 255 entries");
    const std::size_t N = num_nodes;
    fx::require(!fx::overlaps(res, N * 8, vals, std::size_t(nnz) * 8)
                    && !(u && (fx::overlaps(u, N * 8, res, N * 8)
                               || fx::overlaps(u, N * 8, vals, std::size_t(nnz) * 8))),
                FX_ERR_ALIAS, "Assembly state, residual, and matrix values must not alias");
    const P1AsmArgs a{num_nodes, x,  y,       conn, kappa, kappa_is_log, f, u, inc_off,
                      reinterpret_cast<const int2*>(inc), row_ptr, vals, res, 0};
    hipLaunchKernelGGL(k_p1_asm_stencil, dim3(fx::blocks_for(num_nodes, 64 * kP1StWaves)),
                       dim3(64 * kP1StWaves), 0, fx::as_stream(stream), a, blk_code, words);
    fx::check_launch("fx_p1_assemble_stencil");
  });
}

int fx_p1_assemble(fx_index num_nodes, fx_index num_elems, fx_index nnz, fx_index max_row_len,
                   const double* x, const double* y, const fx_index* conn, const double* kappa,
                   int kappa_is_log, const double* f, const double* u, const fx_index* inc_off, const fx_index* inc,
                   const fx_index* row_ptr, double* vals, double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_elems >= 0 && nnz >= 0, FX_ERR_ARG,
                "fx_p1_assemble: negative size");
    if (num_nodes == 0) return;
    fx::require(x && y && inc_off && row_ptr && res && (num_elems == 0 || (conn && inc)),
                FX_ERR_ARG, "fx_p1_assemble: null array");
    fx::require(fx::fits32(num_nodes, 8) && fx::fits32(num_elems, 8), FX_ERR_OVERFLOW,
                "fx_p1_assemble: node or element arrays beyond 4 GiB");
    fx::require(max_row_len >= 0 && max_row_len <= 255, FX_ERR_UNSUPPORTED,
                "fx_p1_assemble: CSR rows longer than 255 entries");
    const std::size_t N = num_nodes;
    fx::require(!fx::overlaps(res, N * 8, vals, std::size_t(nnz) * 8)
                    && !(u && (fx::overlaps(u, N * 8, res, N * 8)
                               || fx::overlaps(u, N * 8, vals, std::size_t(nnz) * 8))),
                FX_ERR_ALIAS, "Assembly state, residual, and matrix values must not alias");
    const unsigned nb      = fx::blocks_for(num_nodes, kP1Block);
    // blocks whose CSR range exceeds the tile add into vals in global memory
    const int      acc_cap = kP1Block * std::min(std::max(max_row_len, 1), kP1MaxRowLen);
    const size_t   lds     = static_cast<size_t>(acc_cap) * 8 + kIncStage * 8;
    const P1AsmArgs a{num_nodes, x,  y,       conn, kappa, kappa_is_log, f, u, inc_off,
                      reinterpret_cast<const int2*>(inc), row_ptr, vals, res, acc_cap};
    if (max_row_len > kP1MaxRowLen)
      hipLaunchKernelGGL(k_p1_assemble<true>, dim3(nb), dim3(kP1Block), lds, fx::as_stream(stream), a);
    else
      hipLaunchKernelGGL(k_p1_assemble<false>, dim3(nb), dim3(kP1Block), lds, fx::as_stream(stream), a);
    fx::check_launch("fx_p1_assemble");
  });
}

int fx_p1_vjp(fx_index num_nodes, fx_index num_elems, const double* x, const double* y,
              const fx_index* conn, const double* m, const double* u, const double* lam,
              const fx_index* bc_mask, const fx_index* inc_off, const fx_index* inc,
              double* jt_lam, double* g, void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_elems >= 0, FX_ERR_ARG, "fx_p1_vjp: negative size");
    if (num_nodes == 0) return;
    fx::require(x && y && u && lam && inc_off && jt_lam
                    && (num_elems == 0 || (conn && inc && m && g)),
                FX_ERR_ARG, "fx_p1_vjp: null array");
    fx::require(fx::fits32(num_nodes, 8) && fx::fits32(num_elems, 8), FX_ERR_OVERFLOW,
                "fx_p1_vjp: node or element arrays beyond 4 GiB");
    const std::size_t N = num_nodes, E = num_elems;
    fx::require(!fx::overlaps(jt_lam, N * 8, g, E * 8) && !fx::overlaps(jt_lam, N * 8, lam, N * 8)
                    && !fx::overlaps(jt_lam, N * 8, u, N * 8)
                    && !fx::overlaps(g, E * 8, m, E * 8),
                FX_ERR_ALIAS, "fx_p1_vjp: outputs must not alias inputs or each other");
    const unsigned nb = fx::blocks_for(num_nodes, kP1Block);
    const P1VjpArgs a{num_nodes, x,  y,   conn, m, u, lam, bc_mask, inc_off,
                      reinterpret_cast<const int2*>(inc), jt_lam, g};
    hipLaunchKernelGGL(k_p1_vjp, dim3(nb), dim3(kP1Block), 0, fx::as_stream(stream), a);
    fx::check_launch("fx_p1_vjp");
  });
}

int fx_p1_vjp_patch(fx_index num_nodes, fx_index num_elems, const double* x, const double* y,
                    const double* m, const double* u, const double* lam, const fx_index* bc_mask,
                    const fx_patch_plan* plan, double* jt_lam, double* g, void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_elems >= 0 && plan, FX_ERR_ARG,
                "fx_p1_vjp_patch: negative size or null plan");
    const fx_patch_plan& q = *plan;
    fx::require(q.num_patches >= 0 && q.num_seam >= 0 && q.num_parts >= 0
                    && q.max_patch_ldofs >= 0 && q.max_patch_ldofs <= 3 * kPatchLanes
                    && (q.num_patches > 0) == (num_elems > 0),
                FX_ERR_ARG, "fx_p1_vjp_patch: plan sizes do not fit the mesh");
    if (num_nodes == 0) return;
    fx::require(fx::fits32(num_nodes, 8) && fx::fits32(num_elems, 8) && fx::fits32(q.num_parts, 8),
                FX_ERR_OVERFLOW, "fx_p1_vjp_patch: node, element or partial arrays beyond 4 GiB");
    fx::require(x && y && u && lam && jt_lam && q.patch_off && q.ldof_off && q.seam_off
                    && (num_elems == 0
                        || (m && g && q.patch_elem && q.elem_ldof && q.ldof_dof && q.ldof_part
                            && q.ldof_end))
                    && (q.num_seam == 0 || q.seam_dof) && (q.num_parts == 0 || q.parts),
                FX_ERR_ARG, "fx_p1_vjp_patch: null array");
    const std::size_t N = num_nodes, E = num_elems, np = std::size_t(q.num_parts) * 8;
    fx::require(!fx::overlaps(jt_lam, N * 8, g, E * 8) && !fx::overlaps(jt_lam, N * 8, lam, N * 8)
                    && !fx::overlaps(jt_lam, N * 8, u, N * 8) && !fx::overlaps(g, E * 8, m, E * 8)
                    && !fx::overlaps(q.parts, np, jt_lam, N * 8)
                    && !fx::overlaps(q.parts, np, lam, N * 8) && !fx::overlaps(q.parts, np, g, E * 8),
                FX_ERR_ALIAS, "fx_p1_vjp_patch: outputs must not alias inputs or each other");
    if (q.num_patches > 0)
    {
      const P1VjpArgs a{num_nodes, x, y, nullptr, m, u, lam, bc_mask, nullptr, nullptr, jt_lam, g};
      const PatchArgs pa{q.patch_off, q.patch_elem, q.elem_vert, q.elem_ldof, q.ldof_off,
                         q.ldof_dof,  q.ldof_part,  q.ldof_end,  jt_lam,      q.parts};
      const int       mx = static_cast<int>(q.max_patch_ldofs);
      hipLaunchKernelGGL(k_p1_vjp_patch, dim3(static_cast<unsigned>(q.num_patches)),
                         dim3(kPatchLanes), p1_patch_lds(mx), fx::as_stream(stream), a, pa, mx);
      fx::check_launch("fx_p1_vjp_patch");
    }
    if (q.num_seam > 0)
    {
      const SeamArgs sa{q.num_seam, q.seam_off, q.seam_dof, q.parts, lam, bc_mask, jt_lam};
      hipLaunchKernelGGL(k_p1_vjp_seam, dim3(fx::blocks_for(q.num_seam, 256)), dim3(256), 0,
                         fx::as_stream(stream), sa);
      fx::check_launch("fx_p1_vjp_patch");
    }
  });
}

int fx_p1_assemble_patch(fx_index num_nodes, fx_index num_elems, fx_index nnz, const double* x,
                         const double* y, const fx_index* conn, const double* kappa,
                         int kappa_is_log, const double* f, const double* u,
                         const fx_index* inc_off, const fx_index* inc, const fx_index* row_ptr,
                         const fx_p1_asm_plan* plan, double* vals, double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_elems >= 0 && nnz >= 0 && plan, FX_ERR_ARG,
                "fx_p1_assemble_patch: negative size or null plan");
    const fx_p1_asm_plan& s = *plan;
    fx::require(s.num_patches >= 0 && s.max_patch_ldofs >= 0
                    && s.max_patch_ldofs <= 3 * kPatchLanes && (num_elems > 0 || s.num_patches == 0)
                    && s.num_left >= 0 && s.num_left <= num_nodes && s.max_tile >= 0
                    && s.max_tile < 0xffff && s.max_runs >= 0 && s.max_runs <= s.max_tile,
                FX_ERR_ARG, "fx_p1_assemble_patch: plan sizes do not fit the mesh");
    if (num_nodes == 0) return;
    fx::require(fx::fits32(num_nodes, 8) && fx::fits32(num_elems, 8), FX_ERR_OVERFLOW,
                "fx_p1_assemble_patch: node or element arrays beyond 4 GiB");
    fx::require(x && y && inc_off && row_ptr && res
                    && (s.num_patches == 0
                        || (s.patch_off && s.patch_elem && s.elem_ldof && s.elem_pos && s.ldof_off
                            && s.ldof_dof && s.ldof_tile && s.patch_tile && s.run_off && s.runs))
                    && (num_elems == 0 || (conn && inc)) && (s.num_left == 0 || s.rows_left),
                FX_ERR_ARG, "fx_p1_assemble_patch: null array");
    const std::size_t N = num_nodes;
    fx::require(!fx::overlaps(res, N * 8, vals, std::size_t(nnz) * 8)
                    && !(u && (fx::overlaps(u, N * 8, res, N * 8)
                               || fx::overlaps(u, N * 8, vals, std::size_t(nnz) * 8))),
                FX_ERR_ALIAS, "Assembly state, residual, and matrix values must not alias");
    const P1AsmArgs a{num_nodes, x,  y,       conn, kappa, kappa_is_log, f, u, inc_off,
                      reinterpret_cast<const int2*>(inc), row_ptr, vals, res, 0};
    if (s.num_patches > 0)
    {
      const PatchArgs  pa{s.patch_off, s.patch_elem, nullptr, s.elem_ldof, s.ldof_off,
                         s.ldof_dof,  nullptr,      nullptr, nullptr,     nullptr};
      const P1AsmPatch ps{s.elem_pos, s.ldof_tile, s.patch_tile, s.run_off,
                          reinterpret_cast<const int2*>(s.runs)};
      const int        mx = static_cast<int>(s.max_patch_ldofs), mt = static_cast<int>(s.max_tile);
      const int        mr = static_cast<int>(s.max_runs);
      const std::size_t lds = p1_asm_patch_lds(mx, mt, mr);
      fx::require(lds <= 160 * 1024, FX_ERR_UNSUPPORTED,
                  "fx_p1_assemble_patch: a patch's tile and nodes exceed the LDS");
      hipLaunchKernelGGL(k_p1_asm_patch, dim3(static_cast<unsigned>(s.num_patches)),
                         dim3(kPatchLanes), lds, fx::as_stream(stream), a, pa, ps, mx, mt, mr);
      fx::check_launch("fx_p1_assemble_patch");
    }
    if (s.num_left > 0)
    {
      hipLaunchKernelGGL(k_p1_assemble_list, dim3(fx::blocks_for(s.num_left, kP1Block)),
                         dim3(kP1Block), 0, fx::as_stream(stream), a, s.rows_left,
                         static_cast<int>(s.num_left));
      fx::check_launch("fx_p1_assemble_patch");
    }
  });
}

} // extern "C"
