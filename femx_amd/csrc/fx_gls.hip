// GLS P1/P1 Navier-Stokes time-step assembly: kernel and C-ABI entry point.
// The device code and its design notes are in fx_gls_dev.hpp.
#include <cmath>

#include "fx_gls_dev.hpp"

namespace
{
using namespace fxk;

__global__ __launch_bounds__(64) void k_gls_assemble(GlsAsmArgs a)
{
  extern __shared__ double tile[];
  gls_assemble_block(fx::xcd_block(gridDim.x), a, tile);
}

template <int LAG>
__global__ __launch_bounds__(64) void k_gls_hist_vjp_elem(GlsVjpArgs a)
{
  gls_hist_vjp_elem<LAG>(fx::xcd_block(gridDim.x) * 64 + threadIdx.x, a);
}

// lag 1 (advection partials only) held to 168 VGPRs: 3 waves/SIMD
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_gls_hist_vjp_elem1(
    GlsVjpArgs a)
{
  gls_hist_vjp_elem<1>(fx::xcd_block(gridDim.x) * 64 + threadIdx.x, a);
}

__global__ __launch_bounds__(kGlsSumNodes) void k_gls_hist_vjp_sum(GlsVjpArgs a)
{
  __shared__ double sw[kGlsSumStage];
  gls_hist_vjp_sum_block(fx::xcd_block(gridDim.x), a, sw);
}

__global__ __launch_bounds__(kPatchLanes) void k_gls_hist_vjp_patch(GlsVjpArgs a, PatchArgs q,
                                                                    GlsPatchOut o,
                                                                    int max_patch_ldofs)
{
  extern __shared__ __align__(16) unsigned char lds[];
  gls_vjp_patch_block(fx::xcd_block(gridDim.x), a, q, o, lds, max_patch_ldofs);
}

__global__ __launch_bounds__(256) void k_gls_hist_vjp_seam(GlsVjpArgs a, SeamArgs s, GlsPatchOut o)
{
  gls_vjp_seam_block(fx::xcd_block(gridDim.x), a, s, o);
}

} // namespace

extern "C" {

int fx_gls_assemble(fx_index num_nodes, fx_index num_states, fx_index num_elems, fx_index nnz,
                    fx_index max_tile, const double* x, const double* y, const fx_index* conn3,
                    const fx_index* dofs9, const fx_index* node_rows, double rho, double mu, double dt, fx_index step,
                    fx_index num_hist, const double* hist, const double* nxt,
                    const fx_index* inc_off, const fx_index* inc_elem, const fx_index* inc_meta,
                    const fx_index* blk_runs, const fx_index* node_tile, double* vals,
                    double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_states >= 0 && num_elems >= 0 && nnz >= 0 && max_tile >= 0,
                FX_ERR_ARG, "fx_gls_assemble: negative size");
    // NavierStokesModel's parameter checks (femx/model/ns/NavierStokesModel.cpp:35-46)
    fx::require(dt > 0.0 && std::isfinite(dt), FX_ERR_ARG,
                "NavierStokesModel requires a positive finite time step");
    fx::require(std::isfinite(rho) && rho > 0.0, FX_ERR_ARG,
                "NavierStokesModel requires positive finite density");
    fx::require(std::isfinite(mu) && mu > 0.0, FX_ERR_ARG,
                "NavierStokesModel requires positive finite viscosity");
    // checkTimeAssemblyInputs (femx/assembly/Assembly.hpp:86-104)
    fx::require(num_hist > 0, FX_ERR_ARG,
                "Time assembly state dimensions do not match AssemblyMap");
    if (num_states == 0) return;
    fx::require(num_nodes > 0 && x && y && node_rows && hist && nxt && inc_off && blk_runs
                    && node_tile && res && (num_elems == 0 || (conn3 && dofs9 && inc_elem && inc_meta)),
                FX_ERR_ARG, "fx_gls_assemble: null array");
    const std::size_t S = num_states, H = S * num_hist * 8, V = std::size_t(nnz) * 8;
    fx::require(!fx::overlaps(res, S * 8, hist, H) && !fx::overlaps(res, S * 8, nxt, S * 8)
                    && !fx::overlaps(res, S * 8, vals, V) && !fx::overlaps(vals, V, hist, H)
                    && !fx::overlaps(vals, V, nxt, S * 8),
                FX_ERR_ALIAS, "Time assembly outputs must not alias inputs or each other");
    fx::require(!vals || max_tile <= kGlsMaxTile, FX_ERR_UNSUPPORTED,
                "fx_gls_assemble: a block's rows exceed the 64 KiB LDS tile");
    const unsigned  nb  = fx::blocks_for(num_nodes, kGlsBlock);
    const size_t    lds = vals ? static_cast<size_t>(max_tile > 0 ? max_tile : 1) * 8 : 0;
    const GlsAsmArgs a{num_nodes,
                       num_states,
                       x,
                       y,
                       conn3,
                       dofs9,
                       node_rows,
                       step,
                       num_hist,
                       hist,
                       nxt,
                       gls_const(rho, mu, dt),
                       inc_off,
                       inc_elem,
                       reinterpret_cast<const uint4*>(inc_meta),
                       reinterpret_cast<const int4*>(blk_runs),
                       node_tile,
                       vals,
                       res,
                       max_tile};
    hipLaunchKernelGGL(k_gls_assemble, dim3(nb), dim3(64), lds, fx::as_stream(stream), a);
    fx::check_launch("fx_gls_assemble");
  });
}

int fx_gls_hist_vjp(fx_index num_nodes, fx_index num_states, fx_index num_elems, const double* x,
                    const double* y, const fx_index* conn3, const fx_index* dofs9,
                    const fx_index* node_rows, double rho, double mu, double dt, fx_index step,
                    fx_index num_hist, fx_index lag, const double* hist, const double* nxt,
                    const double* adj, const fx_index* inc_off, const fx_index* inc_slot,
                    double* out, double* work, void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_states >= 0 && num_elems >= 0, FX_ERR_ARG,
                "fx_gls_hist_vjp: negative size");
    fx::require(dt > 0.0 && std::isfinite(dt), FX_ERR_ARG,
                "NavierStokesModel requires a positive finite time step");
    fx::require(std::isfinite(rho) && rho > 0.0, FX_ERR_ARG,
                "NavierStokesModel requires positive finite density");
    fx::require(std::isfinite(mu) && mu > 0.0, FX_ERR_ARG,
                "NavierStokesModel requires positive finite viscosity");
    // checkTimeAssemblyInputs' variable-block rule (femx/assembly/Assembly.hpp:100-103)
    fx::require(num_hist > 0 && lag >= 0 && lag < num_hist, FX_ERR_ARG,
                "Time assembly variable block is invalid");
    if (num_states == 0) return;
    fx::require(num_nodes > 0 && x && y && node_rows && hist && nxt && adj && inc_off && out
                    && (num_elems == 0 || (conn3 && dofs9 && inc_slot && work)),
                FX_ERR_ARG, "fx_gls_hist_vjp: null array");
    const std::size_t S = std::size_t(num_states) * 8, H = S * num_hist;
    const std::size_t W = std::size_t(num_elems) * 6 * 8;
    fx::require(!fx::overlaps(out, S, hist, H) && !fx::overlaps(out, S, nxt, S)
                    && !fx::overlaps(out, S, adj, S) && !fx::overlaps(work, W, out, S)
                    && !fx::overlaps(work, W, hist, H) && !fx::overlaps(work, W, nxt, S)
                    && !fx::overlaps(work, W, adj, S),
                FX_ERR_ALIAS, "Navier history VJP output must not alias its inputs");
    const GlsVjpArgs a{num_nodes, num_states, x,   y,   conn3, dofs9,
                       node_rows, step,       num_hist, lag, hist,  nxt,
                       adj,       gls_const(rho, mu, dt), inc_off, inc_slot, out, work,
                       num_elems};
    // two phases: per-element columns (partials once per element), then
    // per-node sums in ascending element order
    if (num_elems > 0)
    {
      const dim3 grid(fx::blocks_for(num_elems, 64));
      if (lag == 0)
        hipLaunchKernelGGL(k_gls_hist_vjp_elem<0>, grid, dim3(64), 0, fx::as_stream(stream), a);
      else if (lag == 1)
        hipLaunchKernelGGL(k_gls_hist_vjp_elem1, grid, dim3(64), 0, fx::as_stream(stream), a);
      else
        hipLaunchKernelGGL(k_gls_hist_vjp_elem<2>, grid, dim3(64), 0, fx::as_stream(stream), a);
    }
    hipLaunchKernelGGL(k_gls_hist_vjp_sum, dim3(fx::blocks_for(num_nodes, kGlsSumNodes)),
                       dim3(kGlsSumNodes), 0, fx::as_stream(stream), a);
    fx::check_launch("fx_gls_hist_vjp");
  });
}

int fx_gls_hist_vjp_patch(fx_index num_nodes, fx_index num_states, fx_index num_elems,
                          const double* x, const double* y, const fx_index* node_rows, double rho,
                          double mu, double dt, fx_index step, fx_index num_hist,
                          const double* hist, const double* nxt, const double* adj,
                          const fx_patch_plan* plan, double* out_lag0, double* out_lag1,
                          void* stream)
{
  return fx::guard([&] {
    fx::require(num_nodes >= 0 && num_states >= 0 && num_elems >= 0 && plan, FX_ERR_ARG,
                "fx_gls_hist_vjp_patch: negative size or null plan");
    fx::require(dt > 0.0 && std::isfinite(dt), FX_ERR_ARG,
                "NavierStokesModel requires a positive finite time step");
    fx::require(std::isfinite(rho) && rho > 0.0, FX_ERR_ARG,
                "NavierStokesModel requires positive finite density");
    fx::require(std::isfinite(mu) && mu > 0.0, FX_ERR_ARG,
                "NavierStokesModel requires positive finite viscosity");
    // checkTimeAssemblyInputs' variable-block rule (femx/assembly/Assembly.hpp:100-103)
    fx::require(num_hist > 0 && (out_lag0 || out_lag1) && (!out_lag1 || num_hist > 1), FX_ERR_ARG,
                "Time assembly variable block is invalid");
    const fx_patch_plan& q = *plan;
    fx::require(q.num_patches >= 0 && q.num_seam >= 0 && q.num_parts >= 0
                    && q.max_patch_ldofs >= 0 && q.max_patch_ldofs <= 3 * kPatchLanes
                    && (q.num_patches > 0) == (num_elems > 0),
                FX_ERR_ARG, "fx_gls_hist_vjp_patch: plan sizes do not fit the mesh");
    if (num_states == 0) return;
    fx::require(fx::fits32(num_states, 8 * num_hist) && fx::fits32(num_nodes, 12)
                    && fx::fits32(q.num_parts, 32),
                FX_ERR_OVERFLOW, "fx_gls_hist_vjp_patch: arrays beyond 4 GiB");
    fx::require(num_nodes > 0 && x && y && node_rows && hist && nxt && adj && q.patch_off
                    && q.ldof_off && q.seam_off
                    && (num_elems == 0
                        || (q.patch_elem && q.elem_ldof && q.ldof_dof && q.ldof_part && q.ldof_end))
                    && (q.num_seam == 0 || q.seam_dof) && (q.num_parts == 0 || q.parts),
                FX_ERR_ARG, "fx_gls_hist_vjp_patch: null array");
    const std::size_t S = std::size_t(num_states) * 8, H = S * num_hist;
    const std::size_t W = std::size_t(q.num_parts) * 32;
    for (double* out : {out_lag0, out_lag1})
      fx::require(!out || (!fx::overlaps(out, S, hist, H) && !fx::overlaps(out, S, nxt, S)
                           && !fx::overlaps(out, S, adj, S) && !fx::overlaps(q.parts, W, out, S)),
                  FX_ERR_ALIAS, "Navier history VJP output must not alias its inputs");
    fx::require(!(out_lag0 && out_lag1 && fx::overlaps(out_lag0, S, out_lag1, S))
                    && !fx::overlaps(q.parts, W, hist, H) && !fx::overlaps(q.parts, W, nxt, S)
                    && !fx::overlaps(q.parts, W, adj, S),
                FX_ERR_ALIAS, "Navier history VJP output must not alias its inputs");
    const GlsVjpArgs  a{num_nodes, num_states, x,       y,        nullptr, nullptr,
                       node_rows, step,       num_hist, 0,        hist,    nxt,
                       adj,       gls_const(rho, mu, dt), nullptr, nullptr, nullptr, nullptr,
                       num_elems};
    const GlsPatchOut o{out_lag0, out_lag1};
    if (q.num_patches > 0)
    {
      const PatchArgs pa{q.patch_off, q.patch_elem, q.elem_vert, q.elem_ldof, q.ldof_off,
                         q.ldof_dof,  q.ldof_part,  q.ldof_end,  nullptr,     q.parts};
      const int       mx = static_cast<int>(q.max_patch_ldofs);
      hipLaunchKernelGGL(k_gls_hist_vjp_patch, dim3(static_cast<unsigned>(q.num_patches)),
                         dim3(kPatchLanes), gls_patch_lds(mx), fx::as_stream(stream), a, pa, o, mx);
      fx::check_launch("fx_gls_hist_vjp_patch");
    }
    if (q.num_seam > 0)
    {
      const SeamArgs sa{q.num_seam, q.seam_off, q.seam_dof, q.parts, nullptr, nullptr, nullptr};
      hipLaunchKernelGGL(k_gls_hist_vjp_seam, dim3(fx::blocks_for(q.num_seam, 256)), dim3(256), 0,
                         fx::as_stream(stream), a, sa, o);
      fx::check_launch("fx_gls_hist_vjp_patch");
    }
  });
}

} // extern "C"
