// Taylor-Hood P2/P1 steady Navier-Stokes: Newton Jacobian + residual
// assembly and the adjoint VJP, for gfx950.
//
// Physics (per element, nu = 1/Re, 7-point degree-5 rule, affine geometry):
//   R_(a,c) = int nu grad u_c . grad phi_a + (u . grad u_c) phi_a - p d_c phi_a
//   R_b     = -int psi_b div u
//   J_(a,c),(b,d) = int nu delta_cd grad phi_a . grad phi_b
//                   + phi_a (phi_b d_d u_c + delta_cd u . grad phi_b)
//   J_(a,c),b' = -int psi_b' d_c phi_a,   J_b',(b,d) = -int psi_b' d_d phi_b,
//   J_b',b'' = 0 (present in femx's dense element pattern).
// Local DOF order: 2a+c for the 6 P2 velocity nodes, then 12+b for the 3
// pressure vertices (FESpace::localDof, femx/fem/FESpace.cpp:96-101).
//
// Assembly design: one lane per work item -- the x and y rows of one velocity
// node (they share every column position and every quadrature-point value)
// or one single row -- gathering over the row's incident elements in
// ascending element order.  A 64-lane block owns a contiguous range of rows;
// their CSR values are accumulated in an LDS tile and written back with
// coalesced non-temporal stores.  Each CSR value is a sum over its elements
// in ascending element order with no atomics, so repeated launches give
// bitwise equal results.  Element geometry is recomputed from the three
// vertex coordinates; the 7-point shape tables are built once per block in
// LDS.  Arithmetic uses fused multiply-adds.
//
// VJP design: one lane per column work item, contracting the element matrix
// columns with the boundary-masked adjoint through its quadrature-point
// interpolant (lambda_h, grad lambda_h, pi_h), which avoids forming J.
#include "fx_common.hpp"

namespace
{

constexpr int kThLanes = 64;

struct ThTables
{
  double N[7][6];     // P2 shape values
  double dNdr[7][12]; // reference gradients, (shape, r/s) row-major
  double psi[7][3];   // P1 pressure shapes
  double w[7];        // weights on the reference triangle (sum 1/2)
};

// 7-point rule: centroid and two 3-point orbits with barycentric
// a = (6 -+ sqrt 15)/21, weights 9/80 and (155 -+ sqrt 15)/2400 (Radon;
// Stroud T2:5-1).  P2 shapes l_i(2 l_i - 1) and 4 l_i l_j (edges 01, 12, 20).
__device__ void th_build_tables(ThTables& T)
{
  const int q = threadIdx.x;
  if (q >= 7) return;
  const double s15 = sqrt(15.0);
  const double a1 = (6.0 - s15) / 21.0, b1 = (9.0 + 2.0 * s15) / 21.0;
  const double a2 = (6.0 + s15) / 21.0, b2 = (9.0 - 2.0 * s15) / 21.0;
  const double w1 = (155.0 - s15) / 2400.0, w2 = (155.0 + s15) / 2400.0;
  double       r, s, w;
  switch (q)
  {
  case 0: r = 1.0 / 3.0, s = 1.0 / 3.0, w = 9.0 / 80.0; break;
  case 1: r = a1, s = a1, w = w1; break;
  case 2: r = b1, s = a1, w = w1; break;
  case 3: r = a1, s = b1, w = w1; break;
  case 4: r = a2, s = a2, w = w2; break;
  case 5: r = b2, s = a2, w = w2; break;
  default: r = a2, s = b2, w = w2; break;
  }
  const double l0 = 1.0 - r - s, l1 = r, l2 = s;
  T.w[q]       = w;
  T.N[q][0]    = l0 * (2.0 * l0 - 1.0);
  T.N[q][1]    = l1 * (2.0 * l1 - 1.0);
  T.N[q][2]    = l2 * (2.0 * l2 - 1.0);
  T.N[q][3]    = 4.0 * l0 * l1;
  T.N[q][4]    = 4.0 * l1 * l2;
  T.N[q][5]    = 4.0 * l2 * l0;
  double* d    = T.dNdr[q];
  d[0]         = -(4.0 * l0 - 1.0), d[1] = -(4.0 * l0 - 1.0);
  d[2]         = 4.0 * l1 - 1.0, d[3] = 0.0;
  d[4]         = 0.0, d[5] = 4.0 * l2 - 1.0;
  d[6]         = 4.0 * (l0 - l1), d[7] = -4.0 * l1;
  d[8]         = 4.0 * l2, d[9] = 4.0 * l1;
  d[10]        = -4.0 * l2, d[11] = 4.0 * (l0 - l2);
  T.psi[q][0]  = l0;
  T.psi[q][1]  = l1;
  T.psi[q][2]  = l2;
}

// Affine P1 geometry: inverse Jacobian (row-major) and |det J| as
// ElementValues computes them (femx/fem/ElementValues.cpp:138-215).
struct ThGeom
{
  double i00, i01, i10, i11, adet;
};

__device__ __forceinline__ ThGeom th_geom(const double* __restrict__ xv,
                                          const double* __restrict__ yv, int v0, int v1, int v2)
{
  const double x0 = xv[v0], y0 = yv[v0];
  const double j00 = xv[v1] - x0, j01 = xv[v2] - x0, j10 = yv[v1] - y0, j11 = yv[v2] - y0;
  const double det = j00 * j11 - j01 * j10;
  const double id  = 1.0 / det;
  return {j11 * id, -j01 * id, -j10 * id, j00 * id, fabs(det)};
}

struct ThElem
{
  double s[15];
  ThGeom g;
  double nu;
};

__device__ __forceinline__ void th_gather(ThElem& E, int e, const double* __restrict__ xv,
                                          const double* __restrict__ yv,
                                          const int* __restrict__ vconn,
                                          const int* __restrict__ dofs,
                                          const double* __restrict__ state, double nu,
                                          const double* __restrict__ nu_e)
{
  const int* d = dofs + 15 * static_cast<long>(e);
#pragma unroll
  for (int k = 0; k < 15; ++k) E.s[k] = state[d[k]];
  const int* v = vconn + 3 * static_cast<long>(e);
  E.g          = th_geom(xv, yv, v[0], v[1], v[2]);
  E.nu         = nu_e ? nu_e[e] : nu;
}

// Physical P2 gradients at one quadrature point.
__device__ __forceinline__ void th_grads(const ThTables& T, int q, const ThGeom& g, double* dNx,
                                         double* dNy)
{
#pragma unroll
  for (int b = 0; b < 6; ++b)
  {
    const double dr = T.dNdr[q][2 * b], ds = T.dNdr[q][2 * b + 1];
    dNx[b]          = fma(dr, g.i00, ds * g.i10);
    dNy[b]          = fma(dr, g.i01, ds * g.i11);
  }
}

// Velocity rows (a,0) and (a,1) of the element Jacobian and residual.
__device__ void th_velocity_rows(const ThTables& T, const ThElem& E, int a, double* Kx,
                                 double* Ky, double& Rx, double& Ry)
{
#pragma unroll
  for (int j = 0; j < 15; ++j) Kx[j] = 0.0, Ky[j] = 0.0;
  Rx = 0.0, Ry = 0.0;
  const double nu = E.nu;
  for (int q = 0; q < 7; ++q)
  {
    double dNx[6], dNy[6];
    th_grads(T, q, E.g, dNx, dNy);
    const double wq = E.g.adet * T.w[q];
    double       u0 = 0.0, u1 = 0.0, d00 = 0.0, d01 = 0.0, d10 = 0.0, d11 = 0.0, p = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b)
    {
      const double nb = T.N[q][b], sx = E.s[2 * b], sy = E.s[2 * b + 1];
      u0 = fma(nb, sx, u0), u1 = fma(nb, sy, u1);
      d00 = fma(dNx[b], sx, d00), d01 = fma(dNy[b], sx, d01);
      d10 = fma(dNx[b], sy, d10), d11 = fma(dNy[b], sy, d11);
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) p = fma(T.psi[q][b], E.s[12 + b], p);
    const double na = T.N[q][a], gax = dNx[a], gay = dNy[a];
    const double nwq = na * wq;
    Rx = fma(fma(nu, fma(gax, d00, gay * d01), na * fma(u0, d00, u1 * d01)) - p * gax, wq, Rx);
    Ry = fma(fma(nu, fma(gax, d10, gay * d11), na * fma(u0, d10, u1 * d11)) - p * gay, wq, Ry);
    const double nugx = nu * gax * wq, nugy = nu * gay * wq;
#pragma unroll
    for (int b = 0; b < 6; ++b)
    {
      const double la  = fma(nugx, dNx[b], fma(nugy, dNy[b], nwq * fma(u0, dNx[b], u1 * dNy[b])));
      const double nn  = nwq * T.N[q][b];
      Kx[2 * b]        = fma(nn, d00, Kx[2 * b] + la);
      Kx[2 * b + 1]    = fma(nn, d01, Kx[2 * b + 1]);
      Ky[2 * b]        = fma(nn, d10, Ky[2 * b]);
      Ky[2 * b + 1]    = fma(nn, d11, Ky[2 * b + 1] + la);
    }
#pragma unroll
    for (int b = 0; b < 3; ++b)
    {
      const double pw = T.psi[q][b] * wq;
      Kx[12 + b]      = fma(-pw, gax, Kx[12 + b]);
      Ky[12 + b]      = fma(-pw, gay, Ky[12 + b]);
    }
  }
}

// Pressure row b' of the element Jacobian and residual.
__device__ void th_pressure_row(const ThTables& T, const ThElem& E, int bp, double* Kp, double& Rp)
{
#pragma unroll
  for (int j = 0; j < 15; ++j) Kp[j] = 0.0;
  Rp = 0.0;
  for (int q = 0; q < 7; ++q)
  {
    double dNx[6], dNy[6];
    th_grads(T, q, E.g, dNx, dNy);
    const double wq = E.g.adet * T.w[q];
    double       div = 0.0;
#pragma unroll
    for (int b = 0; b < 6; ++b) div = fma(dNx[b], E.s[2 * b], fma(dNy[b], E.s[2 * b + 1], div));
    const double pw = T.psi[q][bp] * wq;
    Rp              = fma(-pw, div, Rp);
#pragma unroll
    for (int b = 0; b < 6; ++b)
    {
      Kp[2 * b]     = fma(-pw, dNx[b], Kp[2 * b]);
      Kp[2 * b + 1] = fma(-pw, dNy[b], Kp[2 * b + 1]);
    }
  }
}

__device__ __forceinline__ void th_add_row(double* acc, uint4 meta, const double* K)
{
  const unsigned w[4] = {meta.x, meta.y, meta.z, meta.w};
#pragma unroll
  for (int j = 0; j < 15; ++j)
  {
    const int pos = (w[j >> 2] >> (8 * (j & 3))) & 0xff;
    acc[pos]      = acc[pos] + K[j];
  }
}

__global__ __launch_bounds__(kThLanes) void k_th_assemble(
    const double* __restrict__ xv, const double* __restrict__ yv, const int* __restrict__ vconn,
    const int* __restrict__ dofs, const double* __restrict__ state, double nu,
    const double* __restrict__ nu_e, const int* __restrict__ inc_off,
    const int* __restrict__ inc_elem, const uint4* __restrict__ inc_meta,
    const int* __restrict__ blk_row, const int* __restrict__ blk_flags,
    const int* __restrict__ row_ptr, double* __restrict__ vals, double* __restrict__ res)
{
  extern __shared__ double smem[];
  ThTables& T      = *reinterpret_cast<ThTables*>(smem);
  double*   acc    = smem + sizeof(ThTables) / sizeof(double);
  const int b      = blockIdx.x;
  const int rs     = blk_row[b], re = blk_row[b + 1];
  const bool pairs = blk_flags[b] != 0;
  const int base   = row_ptr[rs];
  const int bn     = vals ? row_ptr[re] - base : 0;
  th_build_tables(T);
  for (int i = threadIdx.x; i < bn; i += kThLanes) acc[i] = 0.0;
  __syncthreads();

  const int r = rs + (pairs ? 2 * static_cast<int>(threadIdx.x) : static_cast<int>(threadIdx.x));
  if (r < re)
  {
    double*   ax = acc + (row_ptr[r] - base);
    double*   ay = pairs ? acc + (row_ptr[r + 1] - base) : nullptr;
    double    sx = 0.0, sy = 0.0;
    const int t1 = inc_off[r + 1];
    for (int t = inc_off[r]; t < t1; ++t)
    {
      const int   e    = inc_elem[t];
      const uint4 meta = inc_meta[t];
      const int   li   = meta.w >> 24;
      ThElem      E;
      th_gather(E, e, xv, yv, vconn, dofs, state, nu, nu_e);
      double K0[15], K1[15], R0, R1;
      if (li < 12)
      {
        th_velocity_rows(T, E, li >> 1, K0, K1, R0, R1);
        if (!pairs && (li & 1))
        {
#pragma unroll
          for (int j = 0; j < 15; ++j) K0[j] = K1[j];
          R0 = R1;
        }
      }
      else
        th_pressure_row(T, E, li - 12, K0, R0);
      sx = sx + R0;
      if (pairs) sy = sy + R1;
      if (vals)
      {
        th_add_row(ax, meta, K0);
        if (pairs) th_add_row(ay, meta, K1);
      }
    }
    __builtin_nontemporal_store(sx, res + r);
    if (pairs) __builtin_nontemporal_store(sy, res + r + 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < bn; i += kThLanes) __builtin_nontemporal_store(acc[i], vals + base + i);
}

// Adjoint VJP: jt = J^T lam~ (+ lam on constrained DOFs), g_e = lam~_e^T dR_e/dnu_e.
__global__ __launch_bounds__(kThLanes) void k_th_vjp(
    const double* __restrict__ xv, const double* __restrict__ yv, const int* __restrict__ vconn,
    const int* __restrict__ dofs, const double* __restrict__ state, double nu,
    const double* __restrict__ nu_e, const double* __restrict__ lam,
    const int* __restrict__ bc_mask, const int* __restrict__ inc_off,
    const int* __restrict__ inc_elem, const uint4* __restrict__ inc_meta,
    const int* __restrict__ blk_row, const int* __restrict__ blk_flags, double* __restrict__ jt,
    double* __restrict__ gout)
{
  __shared__ ThTables T;
  const int b      = blockIdx.x;
  const int rs     = blk_row[b], re = blk_row[b + 1];
  const bool pairs = blk_flags[b] != 0;
  th_build_tables(T);
  __syncthreads();
  const int c = rs + (pairs ? 2 * static_cast<int>(threadIdx.x) : static_cast<int>(threadIdx.x));
  if (c >= re) return;
  double    gx = 0.0, gy = 0.0;
  const int t1 = inc_off[c + 1];
  for (int t = inc_off[c]; t < t1; ++t)
  {
    const int e  = inc_elem[t];
    const int lj = inc_meta[t].w >> 24;
    ThElem    E;
    th_gather(E, e, xv, yv, vconn, dofs, state, nu, nu_e);
    double    l[15];
    const int* d = dofs + 15 * static_cast<long>(e);
#pragma unroll
    for (int k = 0; k < 15; ++k)
    {
      const int dk = d[k];
      l[k]         = (bc_mask && bc_mask[dk]) ? 0.0 : lam[dk];
    }
    const bool vel = lj < 12;
    const int  bb  = vel ? lj >> 1 : lj - 12;
    double     cx = 0.0, cy = 0.0, ge = 0.0;
    for (int q = 0; q < 7; ++q)
    {
      double dNx[6], dNy[6];
      th_grads(T, q, E.g, dNx, dNy);
      const double wq = E.g.adet * T.w[q];
      double u0 = 0.0, u1 = 0.0, d00 = 0.0, d01 = 0.0, d10 = 0.0, d11 = 0.0;
      double L0 = 0.0, L1 = 0.0, g00 = 0.0, g01 = 0.0, g10 = 0.0, g11 = 0.0, P = 0.0;
#pragma unroll
      for (int a = 0; a < 6; ++a)
      {
        const double na = T.N[q][a], sx = E.s[2 * a], sy = E.s[2 * a + 1];
        const double lx = l[2 * a], ly = l[2 * a + 1];
        u0 = fma(na, sx, u0), u1 = fma(na, sy, u1);
        d00 = fma(dNx[a], sx, d00), d01 = fma(dNy[a], sx, d01);
        d10 = fma(dNx[a], sy, d10), d11 = fma(dNy[a], sy, d11);
        L0 = fma(na, lx, L0), L1 = fma(na, ly, L1);
        g00 = fma(dNx[a], lx, g00), g01 = fma(dNy[a], lx, g01);
        g10 = fma(dNx[a], ly, g10), g11 = fma(dNy[a], ly, g11);
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) P = fma(T.psi[q][a], l[12 + a], P);
      if (vel)
      {
        const double nb = T.N[q][bb], bx = dNx[bb], by = dNy[bb];
        const double ub = fma(u0, bx, u1 * by);
        const double c0 = fma(nb, fma(d00, L0, d10 * L1), fma(E.nu, fma(bx, g00, by * g01), fma(L0, ub, -P * bx)));
        const double c1 = fma(nb, fma(d01, L0, d11 * L1), fma(E.nu, fma(bx, g10, by * g11), fma(L1, ub, -P * by)));
        cx = fma(c0, wq, cx);
        cy = fma(c1, wq, cy);
        if (lj == 0)
          ge = fma(fma(g00, d00, g01 * d01) + fma(g10, d10, g11 * d11), wq, ge);
      }
      else
        cx = fma(-T.psi[q][bb] * (g00 + g11), wq, cx);
    }
    gx = gx + ((!pairs && vel && (lj & 1)) ? cy : cx);
    if (pairs) gy = gy + cy;
    if (lj == 0 && gout) gout[e] = ge;
  }
  if (bc_mask && bc_mask[c]) gx = gx + lam[c];
  jt[c] = gx;
  if (pairs)
  {
    if (bc_mask && bc_mask[c + 1]) gy = gy + lam[c + 1];
    jt[c + 1] = gy;
  }
}

} // namespace

extern "C" {

int fx_th_assemble(fx_index ndof, fx_index num_elems, fx_index num_blocks, fx_index max_block_nnz,
                   const double* xv, const double* yv, const fx_index* vconn3,
                   const fx_index* dofs15, const double* state, double nu, const double* nu_e,
                   const fx_index* inc_off, const fx_index* inc_elem, const fx_index* inc_meta,
                   const fx_index* blk_row, const fx_index* blk_flags, const fx_index* row_ptr,
                   double* vals, double* res, void* stream)
{
  return fx::guard([&] {
    fx::require(ndof >= 0 && num_elems >= 0 && num_blocks >= 0 && max_block_nnz >= 0, FX_ERR_ARG,
                "fx_th_assemble: negative size");
    if (ndof == 0 || num_blocks == 0) return;
    fx::require(xv && yv && state && inc_off && blk_row && blk_flags && row_ptr && res
                    && (num_elems == 0 || (vconn3 && dofs15 && inc_elem && inc_meta)),
                FX_ERR_ARG, "fx_th_assemble: null array");
    fx::require(!fx::overlaps(res, std::size_t(ndof) * 8, state, std::size_t(ndof) * 8)
                    && !(vals && fx::overlaps(vals, 8, res, std::size_t(ndof) * 8)),
                FX_ERR_ALIAS, "Assembly state, residual, and matrix values must not alias");
    const size_t lds = sizeof(ThTables) + static_cast<size_t>(vals ? max_block_nnz : 0) * 8;
    fx::require(lds <= 160 * 1024, FX_ERR_UNSUPPORTED, "fx_th_assemble: row tile exceeds LDS");
    hipLaunchKernelGGL(k_th_assemble, dim3(num_blocks), dim3(kThLanes), lds, fx::as_stream(stream),
                       xv, yv, vconn3, dofs15, state, nu, nu_e, inc_off, inc_elem,
                       reinterpret_cast<const uint4*>(inc_meta), blk_row, blk_flags, row_ptr, vals,
                       res);
    fx::check_launch("fx_th_assemble");
  });
}

int fx_th_vjp(fx_index ndof, fx_index num_elems, fx_index num_blocks, const double* xv,
              const double* yv, const fx_index* vconn3, const fx_index* dofs15,
              const double* state, double nu, const double* nu_e, const double* lam,
              const fx_index* bc_mask, const fx_index* inc_off, const fx_index* inc_elem,
              const fx_index* inc_meta, const fx_index* blk_row, const fx_index* blk_flags,
              double* jt, double* g, void* stream)
{
  return fx::guard([&] {
    fx::require(ndof >= 0 && num_elems >= 0 && num_blocks >= 0, FX_ERR_ARG,
                "fx_th_vjp: negative size");
    if (ndof == 0 || num_blocks == 0) return;
    fx::require(xv && yv && state && lam && inc_off && blk_row && blk_flags && jt
                    && (num_elems == 0 || (vconn3 && dofs15 && inc_elem && inc_meta)),
                FX_ERR_ARG, "fx_th_vjp: null array");
    fx::require(!fx::overlaps(jt, std::size_t(ndof) * 8, lam, std::size_t(ndof) * 8)
                    && !fx::overlaps(jt, std::size_t(ndof) * 8, state, std::size_t(ndof) * 8)
                    && !(g && fx::overlaps(g, std::size_t(num_elems) * 8, jt, std::size_t(ndof) * 8)),
                FX_ERR_ALIAS, "fx_th_vjp: outputs must not alias inputs or each other");
    hipLaunchKernelGGL(k_th_vjp, dim3(num_blocks), dim3(kThLanes), 0, fx::as_stream(stream), xv, yv,
                       vconn3, dofs15, state, nu, nu_e, lam, bc_mask, inc_off, inc_elem,
                       reinterpret_cast<const uint4*>(inc_meta), blk_row, blk_flags, jt, g);
    fx::check_launch("fx_th_vjp");
  });
}

} // extern "C"
