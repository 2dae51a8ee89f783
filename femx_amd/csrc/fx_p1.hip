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

} // namespace

extern "C" {

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

} // extern "C"
