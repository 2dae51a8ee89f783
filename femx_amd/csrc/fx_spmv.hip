// CSR SpMV, transpose apply and vector ops for gfx950.
//
// Replaces cusparseSpMV (femx/linalg/handler/MatrixHandler.cu:335-398),
// cublasDdot and axpbyKernel (femx/linalg/handler/VectorHandler.cu:124-136,
// 297-317).  SpMV keeps one row per lane and sums the row in CSR order, the
// order of femx's host matvec (MatrixHandler.cpp:157-175), so results equal
// the host's bit for bit; FE rows are short (7 entries for structured P1), so
// a block's rows are staged through LDS with coalesced loads.  The transpose apply gathers each column through the transpose
// permutation in ascending row order -- the order femx's host matvecT adds
// into y (MatrixHandler.cpp:177-198) -- instead of scattering with atomics.
#include "fx_dev.hpp"

namespace
{
constexpr int kRowBlock = 256;
constexpr int kDotBlock = 256;
constexpr int kDotGrid  = 1024;

constexpr int kStage = 2048; // CSR entries staged per block (24 KiB of LDS)

// CSR-stream SpMV: a block's 256 rows own one contiguous range of vals and
// col_ind, loaded into LDS with coalesced non-temporal loads (they are read
// once); each lane then sums its row from LDS in CSR order.  Blocks whose
// range exceeds the stage read their rows directly.
__global__ __launch_bounds__(kRowBlock) void k_spmv(int rows, const int* __restrict__ row_ptr,
                                                    const int* __restrict__ col,
                                                    const double* __restrict__ vals,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ y, double alpha,
                                                    double beta)
{
  __shared__ double sv[kStage];
  __shared__ int    sc[kStage];
  const int  r0 = fx::xcd_block(gridDim.x) * kRowBlock, r1 = min(r0 + kRowBlock, rows);
  const int  k0 = row_ptr[r0], kn = row_ptr[r1] - k0;
  const bool staged = kn <= kStage;
  if (staged)
  {
    fxk::stage_nt<kRowBlock>(sv, vals + k0, kn);
    fxk::stage_nt<kRowBlock>(sc, col + k0, kn);
    __syncthreads();
  }
  const int r = r0 + threadIdx.x;
  if (r >= rows) return;
  double v = 0.0;
  if (staged)
  {
    const int a = row_ptr[r] - k0, b = row_ptr[r + 1] - k0;
    for (int k = a; k < b; ++k) v = v + sv[k] * x[sc[k]];
  }
  else
  {
    const int k1 = row_ptr[r + 1];
    for (int k = row_ptr[r]; k < k1; ++k) v = v + vals[k] * x[col[k]];
  }
  y[r] = beta == 0.0 ? alpha * v : alpha * v + beta * y[r];
}

// k_spmv with 16-bit column offsets (col = row + off): the staged index
// stream is half as wide; the products and their order are unchanged.
__global__ __launch_bounds__(kRowBlock) void k_spmv16(int rows, const int* __restrict__ row_ptr,
                                                      const short* __restrict__ off,
                                                      const double* __restrict__ vals,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y, double alpha,
                                                      double beta)
{
  __shared__ double sv[kStage];
  __shared__ short  so[kStage];
  const int  r0 = fx::xcd_block(gridDim.x) * kRowBlock, r1 = min(r0 + kRowBlock, rows);
  const int  k0 = row_ptr[r0], kn = row_ptr[r1] - k0;
  const bool staged = kn <= kStage;
  if (staged)
  {
    fxk::stage_nt<kRowBlock>(sv, vals + k0, kn);
    fxk::stage_nt<kRowBlock>(so, off + k0, kn);
    __syncthreads();
  }
  const int r = r0 + threadIdx.x;
  if (r >= rows) return;
  double v = 0.0;
  if (staged)
  {
    const int a = row_ptr[r] - k0, b = row_ptr[r + 1] - k0;
    for (int k = a; k < b; ++k) v = v + sv[k] * x[r + so[k]];
  }
  else
  {
    const int k1 = row_ptr[r + 1];
    for (int k = row_ptr[r]; k < k1; ++k) v = v + vals[k] * x[r + off[k]];
  }
  y[r] = beta == 0.0 ? alpha * v : alpha * v + beta * y[r];
}

// Several products with one matrix, y_j = alpha A x_j + beta y_j for
// j < nv (nv <= kMaxVec): the block's CSR range is staged once and each lane
// sums its row for every vector, each sum in CSR order exactly as k_spmv /
// k_spmv16 form it, so every y_j equals a single-vector product bit for bit.
// The matrix streams from HBM once instead of nv times.  off (16-bit column
// offsets) or col (int32 columns) gives the columns; the other is null.
constexpr int kMaxVec = 16;

template <int NV>
__global__ __launch_bounds__(kRowBlock) void k_spmv_multi(
    int rows, const int* __restrict__ row_ptr, const int* __restrict__ col,
    const short* __restrict__ off, const double* __restrict__ vals,
    const double* __restrict__ X, long ldx, double* __restrict__ Y, long ldy, double alpha,
    double beta)
{
  __shared__ double sv[kStage];
  __shared__ int    sc[kStage];
  const int  r0 = fx::xcd_block(gridDim.x) * kRowBlock, r1 = min(r0 + kRowBlock, rows);
  const int  k0 = row_ptr[r0], kn = row_ptr[r1] - k0;
  const bool staged = kn <= kStage;
  const int  rr     = r0 + threadIdx.x;
  if (staged)
  {
    // stage the column index of each entry (rows of this block only)
    fxk::stage_nt<kRowBlock>(sv, vals + k0, kn);
    if (col)
      fxk::stage_nt<kRowBlock>(sc, col + k0, kn);
    else
      fxk::stage_nt<kRowBlock>(sc, off + k0, kn);
    __syncthreads();
  }
  if (rr >= rows) return;
  const int r = rr;
  double    acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0.0;
  const int a = row_ptr[r], b = row_ptr[r + 1];
  for (int k = a; k < b; ++k)
  {
    double v;
    int    c;
    if (staged)
    {
      v = sv[k - k0];
      c = sc[k - k0];
    }
    else
    {
      v = vals[k];
      c = col ? col[k] : static_cast<int>(off[k]);
    }
    if (!col) c += r;
    double xv[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) xv[j] = X[j * ldx + c];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = acc[j] + v * xv[j];
  }
#pragma unroll
  for (int j = 0; j < NV; ++j)
  {
    double* y = Y + j * ldy + r;
    __builtin_nontemporal_store(beta == 0.0 ? alpha * acc[j] : alpha * acc[j] + beta * *y, y);
  }
}

// The same products with stencil blocks (fx_csr_stencil_plan).  In a block
// of 64 rows that share one column-offset stencil, entry k of every row
// reads the vectors at row + off_k: a run of consecutive offsets (a band) is
// one coalesced 64-value load per vector, at the band's first offset, and
// each next offset in the band is the same registers moved down one lane
// (DPP wave_shl:1: lane i takes lane i+1) with the one new value, at row 63
// + offset, filled into lane 63.  Gathering instead, each entry is a
// separate request per vector for lines the wave has just read (the L1 does
// not keep them across the resident waves' entries).  One wave per 64-row
// block; the block's matrix range (64 S contiguous values) is staged in the
// wave's LDS.  Other blocks gather as k_spmv_multi does, unstaged.  Every
// sum is formed in CSR order as in k_spmv_multi, so each product equals the
// single-vector one bit for bit.  X's offsets fit 32 bits (checked by the
// entry point).
constexpr int kStWaves = 4;
constexpr int kStMaxLen = FX_STENCIL_MAX_LEN; // longest stencil read as one (stage: 64 x 8 values per wave)

template <int NV>
__global__ __launch_bounds__(64 * kStWaves) void k_spmv_multi_st(
    int rows, const int* __restrict__ row_ptr, const int* __restrict__ col,
    const double* __restrict__ vals, const int* __restrict__ blk_code,
    const int* __restrict__ st_off, const double* __restrict__ X, long ldx,
    double* __restrict__ Y, long ldy, double alpha, double beta)
{
  __shared__ double sv_all[kStWaves][64 * kStMaxLen];
  const int wv   = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b    = fx::xcd_block(gridDim.x) * kStWaves + wv;
  const int r0   = 64 * b;
  if (r0 >= rows) return;
  const int r    = r0 + lane;
  const int code = blk_code[b];
  double    acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0.0;
  if (code >= 0 && (code & 31) <= kStMaxLen)
  {
    const int S = code & 31, so = code >> 5;
    const int base = row_ptr[r0];
    double*   sv   = sv_all[wv];
    const int n    = 64 * S;
    for (int i = lane; i < n; i += 4 * 64)
    {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = i + 64 * u < n ? __builtin_nontemporal_load(vals + base + i + 64 * u) : 0.0;
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i + 64 * u < n) sv[i + 64 * u] = v[u];
    }
    fxk::wave_sync();
    const double* m = sv + lane * S;
    for (int k = 0; k < S;)
    {
      const int o = st_off[so + k];
      int       w = 1;
      while (k + w < S && st_off[so + k + w] == o + w) ++w;
      double cur[NV];
#pragma unroll
      for (int j = 0; j < NV; ++j) cur[j] = fxk::gat32(X + j * ldx, r + o);
#pragma unroll
      for (int j = 0; j < NV; ++j) acc[j] = acc[j] + m[k] * cur[j];
      for (int d = 1; d < w; ++d)
      {
        const int    t = r0 + 63 + o + d; // the value new to the band: lane 63's
        const double mk = m[k + d];
#pragma unroll
        for (int j = 0; j < NV; ++j)
        {
          cur[j] = fxk::wave_shl1(cur[j], X[j * ldx + t]);
          acc[j] = acc[j] + mk * cur[j];
        }
      }
      k += w;
    }
  }
  else if (r < rows)
  {
    const int a = row_ptr[r], e = row_ptr[r + 1];
    for (int k = a; k < e; ++k)
    {
      const double v = vals[k];
      const int    c = col[k];
      double       xv[NV];
#pragma unroll
      for (int j = 0; j < NV; ++j) xv[j] = fxk::gat32(X + j * ldx, c);
#pragma unroll
      for (int j = 0; j < NV; ++j) acc[j] = acc[j] + v * xv[j];
    }
  }
  if (r < rows)
  {
#pragma unroll
    for (int j = 0; j < NV; ++j)
    {
      double* y = Y + j * ldy + r;
      __builtin_nontemporal_store(beta == 0.0 ? alpha * acc[j] : alpha * acc[j] + beta * *y, y);
    }
  }
}

template <int NV = kMaxVec>
void launch_multi_st_n(int n, int rows, const int* row_ptr, const int* col, const double* vals,
                       const int* blk_code, const int* st_off, const double* X, long ldx, double* Y,
                       long ldy, double alpha, double beta, hipStream_t st)
{
  if constexpr (NV >= 1)
  {
    if (n == NV)
    {
      const unsigned nb = fx::blocks_for(rows, 64 * kStWaves);
      hipLaunchKernelGGL(k_spmv_multi_st<NV>, dim3(nb), dim3(64 * kStWaves), 0, st, rows, row_ptr,
                         col, vals, blk_code, st_off, X, ldx, Y, ldy, alpha, beta);
    }
    else
      launch_multi_st_n<NV - 1>(n, rows, row_ptr, col, vals, blk_code, st_off, X, ldx, Y, ldy,
                                alpha, beta, st);
  }
}

template <int NV>
void launch_multi(int rows, const int* row_ptr, const int* col, const short* off, const double* vals,
                  const double* X, long ldx, double* Y, long ldy, double alpha, double beta,
                  hipStream_t st)
{
  hipLaunchKernelGGL(k_spmv_multi<NV>, dim3(fx::blocks_for(rows, kRowBlock)), dim3(kRowBlock), 0,
                     st, rows, row_ptr, col, off, vals, X, ldx, Y, ldy, alpha, beta);
}

// Dispatch on the vector count (compile-time, so each entry's loads issue
// together).
template <int NV = kMaxVec>
void launch_multi_n(int n, int rows, const int* row_ptr, const int* col, const short* off,
                    const double* vals, const double* X, long ldx, double* Y, long ldy,
                    double alpha, double beta, hipStream_t st)
{
  if constexpr (NV >= 1)
  {
    if (n == NV)
      launch_multi<NV>(rows, row_ptr, col, off, vals, X, ldx, Y, ldy, alpha, beta, st);
    else
      launch_multi_n<NV - 1>(n, rows, row_ptr, col, off, vals, X, ldx, Y, ldy, alpha, beta, st);
  }
}

__global__ __launch_bounds__(kRowBlock) void k_spmv_t(int cols, const int* __restrict__ t_ptr,
                                                      const int* __restrict__ t_row,
                                                      const int* __restrict__ t_perm,
                                                      const double* __restrict__ vals,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y, double alpha,
                                                      double beta)
{
  const int c = blockIdx.x * kRowBlock + threadIdx.x;
  if (c >= cols) return;
  double    v  = beta == 0.0 ? 0.0 : y[c] * beta;
  const int k1 = t_ptr[c + 1];
  for (int k = t_ptr[c]; k < k1; ++k) v = v + vals[t_perm[k]] * (alpha * x[t_row[k]]);
  y[c] = v;
}

__global__ __launch_bounds__(kRowBlock) void k_axpby(int n, double a, const double* __restrict__ x,
                                                     double b, double* __restrict__ y)
{
  const int i = blockIdx.x * kRowBlock + threadIdx.x;
  if (i < n) y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

__global__ __launch_bounds__(kRowBlock) void k_gather(int n, const int* __restrict__ idx,
                                                      const double* __restrict__ src,
                                                      double* __restrict__ dst)
{
  const int i = blockIdx.x * kRowBlock + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ __launch_bounds__(kRowBlock) void k_div(int n, const double* x, const double* d,
                                                   double* z)
{
  const int i = blockIdx.x * kRowBlock + threadIdx.x;
  if (i < n) z[i] = x[i] / d[i];
}

__global__ __launch_bounds__(kRowBlock) void k_fill(int n, double v, double* y)
{
  const int i = blockIdx.x * kRowBlock + threadIdx.x;
  if (i < n) y[i] = v;
}

__global__ __launch_bounds__(kRowBlock) void k_scatter(int n, const int* __restrict__ idx,
                                                       const double* __restrict__ src,
                                                       double* __restrict__ dst)
{
  const int i = blockIdx.x * kRowBlock + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}

// Fixed-shape two-stage tree reduction: the grid, the stride pattern and the
// tree are independent of timing, so the sum is reproducible bit for bit.
__device__ __forceinline__ double block_sum(double v)
{
  __shared__ double s[kDotBlock];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1)
  {
    if (threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  return s[0];
}

__global__ __launch_bounds__(kDotBlock) void k_dot_partial(int n, const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           double* __restrict__ work)
{
  double v = 0.0;
  for (int i = blockIdx.x * kDotBlock + threadIdx.x; i < n; i += kDotGrid * kDotBlock)
    v = v + x[i] * y[i];
  v = block_sum(v);
  if (threadIdx.x == 0) work[blockIdx.x] = v;
}

__global__ __launch_bounds__(kDotBlock) void k_dot_final(const double* __restrict__ work,
                                                         double* __restrict__ out)
{
  double v = 0.0;
  for (int i = threadIdx.x; i < kDotGrid; i += kDotBlock) v = v + work[i];
  v = block_sum(v);
  if (threadIdx.x == 0) out[0] = v;
}
} // namespace

extern "C" {

int fx_csr_matvec(fx_index rows, fx_index cols, const fx_index* row_ptr, const fx_index* col_ind,
                  const double* vals, const double* x, double* y, double alpha, double beta,
                  void* stream)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && cols >= 0, FX_ERR_ARG, "fx_csr_matvec: negative size");
    if (rows == 0) return;
    fx::require(row_ptr && y && (cols == 0 || x), FX_ERR_ARG, "fx_csr_matvec: null array");
    fx::require(!fx::overlaps(x, std::size_t(cols) * 8, y, std::size_t(rows) * 8),
                FX_ERR_ALIAS, "CSR matvec does not support in-place views");
    fx::require(vals != y, FX_ERR_ALIAS, "CSR matvec output must not alias matrix values");
    hipLaunchKernelGGL(k_spmv, dim3(fx::blocks_for(rows, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), rows, row_ptr, col_ind, vals, x, y, alpha, beta);
    fx::check_launch("fx_csr_matvec");
  });
}

int fx_csr_matvec16(fx_index rows, fx_index cols, const fx_index* row_ptr, const int16_t* col_off,
                    const double* vals, const double* x, double* y, double alpha, double beta,
                    void* stream)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && cols >= 0, FX_ERR_ARG, "fx_csr_matvec16: negative size");
    if (rows == 0) return;
    fx::require(row_ptr && y && (cols == 0 || x), FX_ERR_ARG, "fx_csr_matvec16: null array");
    fx::require(!fx::overlaps(x, std::size_t(cols) * 8, y, std::size_t(rows) * 8),
                FX_ERR_ALIAS, "CSR matvec does not support in-place views");
    fx::require(vals != y, FX_ERR_ALIAS, "CSR matvec output must not alias matrix values");
    hipLaunchKernelGGL(k_spmv16, dim3(fx::blocks_for(rows, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), rows, row_ptr, reinterpret_cast<const short*>(col_off),
                       vals, x, y, alpha, beta);
    fx::check_launch("fx_csr_matvec16");
  });
}

int fx_csr_col_off16(fx_index rows, const fx_index* row_ptr, const fx_index* col_ind,
                     int16_t* col_off)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && row_ptr, FX_ERR_ARG, "fx_csr_col_off16: invalid arguments");
    fx::require(row_ptr[rows] == 0 || (col_ind && col_off), FX_ERR_ARG,
                "fx_csr_col_off16: null array");
    for (fx_index r = 0; r < rows; ++r)
      for (fx_index k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      {
        const long d = static_cast<long>(col_ind[k]) - r;
        fx::require(d >= -32768 && d <= 32767, FX_ERR_UNSUPPORTED,
                    "fx_csr_col_off16: a column offset does not fit 16 bits");
        col_off[k] = static_cast<int16_t>(d);
      }
  });
}

int fx_csr_matvec_t(fx_index rows, fx_index cols, const fx_index* t_ptr, const fx_index* t_row,
                    const fx_index* t_perm, const double* vals, const double* x, double* y,
                    double alpha, double beta, void* stream)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && cols >= 0, FX_ERR_ARG, "fx_csr_matvec_t: negative size");
    if (cols == 0) return;
    fx::require(t_ptr && y && (rows == 0 || x), FX_ERR_ARG, "fx_csr_matvec_t: null array");
    fx::require(!fx::overlaps(x, std::size_t(rows) * 8, y, std::size_t(cols) * 8),
                FX_ERR_ALIAS, "CSR matvecT does not support in-place views");
    fx::require(vals != y, FX_ERR_ALIAS, "CSR matvecT output must not alias matrix values");
    hipLaunchKernelGGL(k_spmv_t, dim3(fx::blocks_for(cols, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), cols, t_ptr, t_row, t_perm, vals, x, y, alpha, beta);
    fx::check_launch("fx_csr_matvec_t");
  });
}

int fx_vec_axpby(fx_index n, double a, const double* x, double b, double* y, void* stream)
{
  return fx::guard([&] {
    fx::require(n >= 0, FX_ERR_ARG, "fx_vec_axpby: negative size");
    if (n == 0) return;
    fx::require(x && y, FX_ERR_ARG, "fx_vec_axpby: null array");
    hipLaunchKernelGGL(k_axpby, dim3(fx::blocks_for(n, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), n, a, x, b, y);
    fx::check_launch("fx_vec_axpby");
  });
}

int fx_csr_matvec_multi(fx_index rows, fx_index cols, const fx_index* row_ptr,
                        const fx_index* col_ind, const int16_t* col_off, const double* vals,
                        fx_index nv, const double* X, fx_index ldx, double* Y, fx_index ldy,
                        double alpha, double beta, void* stream)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && cols >= 0 && nv >= 0, FX_ERR_ARG, "fx_csr_matvec_multi: negative size");
    if (rows == 0 || nv == 0) return;
    fx::require(row_ptr && vals && X && Y && ((col_ind != nullptr) != (col_off != nullptr)),
                FX_ERR_ARG, "fx_csr_matvec_multi: null array (give col_ind or col_off, not both)");
    fx::require(ldx >= cols && ldy >= rows, FX_ERR_ARG,
                "fx_csr_matvec_multi: leading dimension smaller than the vector length");
    const std::size_t xs = std::size_t(ldx) * (nv - 1) + cols, ys = std::size_t(ldy) * (nv - 1) + rows;
    fx::require(!fx::overlaps(X, xs * 8, Y, ys * 8), FX_ERR_ALIAS,
                "CSR matvec input and output must not alias");
    for (fx_index j0 = 0; j0 < nv; j0 += kMaxVec)
    {
      const int n = static_cast<int>(std::min<fx_index>(kMaxVec, nv - j0));
      launch_multi_n(n, rows, row_ptr, col_ind, reinterpret_cast<const short*>(col_off), vals,
                     X + j0 * ldx, static_cast<long>(ldx), Y + j0 * ldy, static_cast<long>(ldy),
                     alpha, beta, fx::as_stream(stream));
      fx::check_launch("fx_csr_matvec_multi");
    }
  });
}

int fx_csr_matvec_multi_st(fx_index rows, fx_index cols, const fx_index* row_ptr,
                           const fx_index* col_ind, const double* vals, const fx_index* blk_code,
                           const fx_index* st_off, fx_index nv, const double* X, fx_index ldx,
                           double* Y, fx_index ldy, double alpha, double beta, void* stream)
{
  return fx::guard([&] {
    fx::require(rows >= 0 && cols >= 0 && nv >= 0, FX_ERR_ARG,
                "fx_csr_matvec_multi_st: negative size");
    if (rows == 0 || nv == 0) return;
    fx::require(row_ptr && col_ind && vals && blk_code && st_off && X && Y, FX_ERR_ARG,
                "fx_csr_matvec_multi_st: null array");
    fx::require(ldx >= cols && ldy >= rows, FX_ERR_ARG,
                "fx_csr_matvec_multi_st: leading dimension smaller than the vector length");
    fx::require(fx::fits32(cols, 8), FX_ERR_OVERFLOW,
                "fx_csr_matvec_multi_st: a vector exceeds 4 GiB");
    const std::size_t xs = std::size_t(ldx) * (nv - 1) + cols, ys = std::size_t(ldy) * (nv - 1) + rows;
    fx::require(!fx::overlaps(X, xs * 8, Y, ys * 8), FX_ERR_ALIAS,
                "CSR matvec input and output must not alias");
    for (fx_index j0 = 0; j0 < nv; j0 += kMaxVec)
    {
      const int n = static_cast<int>(std::min<fx_index>(kMaxVec, nv - j0));
      launch_multi_st_n(n, rows, row_ptr, col_ind, vals, blk_code, st_off, X + j0 * ldx,
                        static_cast<long>(ldx), Y + j0 * ldy, static_cast<long>(ldy), alpha, beta,
                        fx::as_stream(stream));
      fx::check_launch("fx_csr_matvec_multi_st");
    }
  });
}

int fx_vec_gather(fx_index n, const fx_index* idx, const double* src, fx_index src_len,
                  double* dst, void* stream)
{
  return fx::guard([&] {
    fx::require(n >= 0 && src_len >= 0, FX_ERR_ARG, "fx_vec_gather: negative size");
    if (n == 0) return;
    fx::require(idx && src && dst, FX_ERR_ARG, "fx_vec_gather: null array");
    fx::require(!fx::overlaps(src, std::size_t(src_len) * 8, dst, std::size_t(n) * 8),
                FX_ERR_ALIAS, "Host gather does not support aliased vectors");
    hipLaunchKernelGGL(k_gather, dim3(fx::blocks_for(n, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), n, idx, src, dst);
    fx::check_launch("fx_vec_gather");
  });
}

int fx_vec_scatter(fx_index n, const fx_index* idx, const double* src, double* dst,
                   fx_index dst_len, void* stream)
{
  return fx::guard([&] {
    fx::require(n >= 0 && dst_len >= 0, FX_ERR_ARG, "fx_vec_scatter: negative size");
    if (n == 0) return;
    fx::require(idx && src && dst, FX_ERR_ARG, "fx_vec_scatter: null array");
    fx::require(!fx::overlaps(src, std::size_t(n) * 8, dst, std::size_t(dst_len) * 8),
                FX_ERR_ALIAS, "Host scatter does not support aliased vectors");
    hipLaunchKernelGGL(k_scatter, dim3(fx::blocks_for(n, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), n, idx, src, dst);
    fx::check_launch("fx_vec_scatter");
  });
}

int fx_vec_div(fx_index n, const double* x, const double* d, double* z, void* stream)
{
  return fx::guard([&] {
    fx::require(n >= 0, FX_ERR_ARG, "fx_vec_div: negative size");
    if (n == 0) return;
    fx::require(x && d && z, FX_ERR_ARG, "fx_vec_div: null array");
    hipLaunchKernelGGL(k_div, dim3(fx::blocks_for(n, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), n, x, d, z);
    fx::check_launch("fx_vec_div");
  });
}

int fx_vec_fill(fx_index n, double v, double* y, void* stream)
{
  return fx::guard([&] {
    fx::require(n >= 0, FX_ERR_ARG, "fx_vec_fill: negative size");
    if (n == 0) return;
    fx::require(y, FX_ERR_ARG, "fx_vec_fill: null array");
    hipLaunchKernelGGL(k_fill, dim3(fx::blocks_for(n, kRowBlock)), dim3(kRowBlock), 0,
                       fx::as_stream(stream), n, v, y);
    fx::check_launch("fx_vec_fill");
  });
}

fx_index fx_dot_work_len(void) { return kDotGrid; }

int fx_vec_dot(fx_index n, const double* x, const double* y, double* work, double* out,
               void* stream)
{
  return fx::guard([&] {
    fx::require(n >= 0 && work && out && (n == 0 || (x && y)), FX_ERR_ARG,
                "fx_vec_dot: invalid arguments");
    hipLaunchKernelGGL(k_dot_partial, dim3(kDotGrid), dim3(kDotBlock), 0, fx::as_stream(stream), n,
                       x, y, work);
    fx::check_launch("fx_vec_dot");
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kDotBlock), 0, fx::as_stream(stream), work, out);
    fx::check_launch("fx_vec_dot");
  });
}

} // extern "C"
