// Dirichlet row replacement and forward elimination for gfx950.
//
// Replaces femx's CUDA boundary kernels (femx/assembly/BoundaryMap.cu:16-94)
// and prepareForwardSolve (:192-211).  femx subtracts the column corrections
// from the right-hand side with atomicAdd (BoundaryMap.cu:63); here each
// unconstrained row that meets a constrained column is owned by one lane,
// which applies its corrections in ascending boundary index -- the order of
// femx's host loop (femx/assembly/BoundaryMap.cpp:227-240) -- so the result
// is deterministic and equal to the host's.  The same launch replaces the
// constrained rows, which no elimination lane touches.
#include "fx_common.hpp"

namespace
{
constexpr int kBcBlock = 256;

__global__ __launch_bounds__(kBcBlock) void k_replace_rows(int nb, const int* __restrict__ bc_rows,
                                                           const int* __restrict__ diag,
                                                           const int* __restrict__ row_ptr,
                                                           double* __restrict__ vals, double dv)
{
  const int ib = blockIdx.x * kBcBlock + threadIdx.x;
  if (ib >= nb) return;
  const int row = bc_rows[ib];
  for (int k = row_ptr[row]; k < row_ptr[row + 1]; ++k) vals[k] = 0.0;
  vals[diag[ib]] = dv;
}

__global__ __launch_bounds__(kBcBlock) void k_replace_res(int nb, const int* __restrict__ bc_rows,
                                                          const double* __restrict__ state,
                                                          const double* __restrict__ bcv,
                                                          double* __restrict__ res)
{
  const int ib = blockIdx.x * kBcBlock + threadIdx.x;
  if (ib >= nb) return;
  const int row = bc_rows[ib];
  res[row]      = state[row] - bcv[ib];
}

__global__ __launch_bounds__(kBcBlock) void k_zero(int nb, const int* __restrict__ bc_rows,
                                                   double* __restrict__ v)
{
  const int ib = blockIdx.x * kBcBlock + threadIdx.x;
  if (ib < nb) v[bc_rows[ib]] = 0.0;
}

// Lanes [0, ne) eliminate constrained columns from unconstrained rows; lanes
// [ne, ne+nb) turn constrained rows into identity rows with rhs = g.
__global__ __launch_bounds__(kBcBlock) void k_prepare_forward(
    int nb, const int* __restrict__ bc_rows, const int* __restrict__ diag,
    const int* __restrict__ row_ptr, int ne, const int* __restrict__ erows,
    const int* __restrict__ eoff, const int* __restrict__ ek, const int* __restrict__ eib,
    double* __restrict__ vals, double* __restrict__ rhs, const double* __restrict__ bcv)
{
  const int t = blockIdx.x * kBcBlock + threadIdx.x;
  if (t < ne)
  {
    const int row = erows[t];
    double    r   = rhs[row];
    for (int i = eoff[t]; i < eoff[t + 1]; ++i)
    {
      const int k = ek[i];
      r           = r - vals[k] * bcv[eib[i]];
      vals[k]     = 0.0;
    }
    rhs[row] = r;
  }
  else if (t < ne + nb)
  {
    const int ib  = t - ne;
    const int row = bc_rows[ib];
    for (int k = row_ptr[row]; k < row_ptr[row + 1]; ++k) vals[k] = 0.0;
    vals[diag[ib]] = 1.0;
    rhs[row]       = bcv[ib];
  }
}
} // namespace

extern "C" {

int fx_bc_replace_rows(fx_index nb, const fx_index* bc_rows, const fx_index* diag,
                       const fx_index* row_ptr, double* vals, double dv, void* stream)
{
  return fx::guard([&] {
    fx::require(nb >= 0, FX_ERR_ARG, "fx_bc_replace_rows: negative size");
    if (nb == 0) return;
    fx::require(bc_rows && diag && row_ptr && vals, FX_ERR_ARG, "fx_bc_replace_rows: null array");
    hipLaunchKernelGGL(k_replace_rows, dim3(fx::blocks_for(nb, kBcBlock)), dim3(kBcBlock), 0,
                       fx::as_stream(stream), nb, bc_rows, diag, row_ptr, vals, dv);
    fx::check_launch("fx_bc_replace_rows");
  });
}

int fx_bc_replace_res(fx_index nb, const fx_index* bc_rows, const double* state,
                      const double* bcv, double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(nb >= 0, FX_ERR_ARG, "fx_bc_replace_res: negative size");
    if (nb == 0) return;
    fx::require(bc_rows && state && bcv && res, FX_ERR_ARG, "fx_bc_replace_res: null array");
    fx::require(state != res && bcv != res, FX_ERR_ALIAS,
                "BoundaryMap residual output must not alias its inputs");
    hipLaunchKernelGGL(k_replace_res, dim3(fx::blocks_for(nb, kBcBlock)), dim3(kBcBlock), 0,
                       fx::as_stream(stream), nb, bc_rows, state, bcv, res);
    fx::check_launch("fx_bc_replace_res");
  });
}

int fx_bc_zero(fx_index nb, const fx_index* bc_rows, double* v, void* stream)
{
  return fx::guard([&] {
    fx::require(nb >= 0, FX_ERR_ARG, "fx_bc_zero: negative size");
    if (nb == 0) return;
    fx::require(bc_rows && v, FX_ERR_ARG, "fx_bc_zero: null array");
    hipLaunchKernelGGL(k_zero, dim3(fx::blocks_for(nb, kBcBlock)), dim3(kBcBlock), 0,
                       fx::as_stream(stream), nb, bc_rows, v);
    fx::check_launch("fx_bc_zero");
  });
}

int fx_bc_prepare_forward(fx_index nb, const fx_index* bc_rows, const fx_index* diag,
                          const fx_index* row_ptr, fx_index ne, const fx_index* erows,
                          const fx_index* eoff, const fx_index* ek, const fx_index* eib,
                          double* vals, double* rhs, const double* bcv, void* stream)
{
  return fx::guard([&] {
    fx::require(nb >= 0 && ne >= 0, FX_ERR_ARG, "fx_bc_prepare_forward: negative size");
    if (nb + ne == 0) return;
    fx::require(bc_rows && diag && row_ptr && vals && rhs && bcv && eoff
                    && (ne == 0 || (erows && ek && eib)),
                FX_ERR_ARG, "fx_bc_prepare_forward: null array");
    fx::require(rhs != bcv, FX_ERR_ALIAS, "BoundaryMap RHS and prescribed values must not alias");
    fx::require(vals != rhs && vals != bcv, FX_ERR_ALIAS,
                "BoundaryMap vectors must not alias matrix values");
    hipLaunchKernelGGL(k_prepare_forward, dim3(fx::blocks_for(std::int64_t(nb) + ne, kBcBlock)),
                       dim3(kBcBlock), 0, fx::as_stream(stream), nb, bc_rows, diag, row_ptr, ne,
                       erows, eoff, ek, eib, vals, rhs, bcv);
    fx::check_launch("fx_bc_prepare_forward");
  });
}

} // extern "C"
