// Taylor-Hood P2/P1 steady Navier-Stokes: Newton Jacobian + residual
// assembly and the adjoint VJP, for gfx950.  Device code shared by the
// single-instance entry points (fx_th.hip) and the batched ones (fx_batch.hip).
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
// Assembly design: two lanes per work item -- the x and y rows of one
// velocity node (they share every column position and every quadrature-point
// value) or one single row -- gathering over the row's incident elements in
// ascending element order.  A 64-lane block owns a contiguous range of rows;
// their CSR values are accumulated in an LDS tile and written back with
// coalesced non-temporal stores.  Each CSR value is a sum over its elements
// in ascending element order with no atomics, so repeated launches give
// bitwise equal results.  Element geometry is recomputed from the three
// vertex coordinates; the 7-point shape tables are built once per block in
// LDS.  Arithmetic uses fused multiply-adds.
//
// VJP design: two phases.  One lane per element contracts every column of
// its element matrix with the boundary-masked adjoint through the adjoint's
// quadrature-point interpolant (lambda_h, grad lambda_h, pi_h) -- the
// quadrature-point state is formed once per element and J is never formed --
// into a 15-per-element workspace; then one lane per DOF sums its column's
// element contributions in ascending element order (a segmented reduction
// over the row incidence) instead of scattering with atomics.
#pragma once
#include "fx_common.hpp"

namespace fxk
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
__device__ __forceinline__ void th_velocity_rows(const ThTables& T, const ThElem& E, int a, double* Kx,
                                 double* Ky, double& Rx, double& Ry)
{
#pragma unroll
  for (int j = 0; j < 15; ++j) Kx[j] = 0.0, Ky[j] = 0.0;
  Rx = 0.0, Ry = 0.0;
  const double nu = E.nu;
#pragma unroll 1
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
    const double na  = T.N[q][a];
    const double gax = fma(T.dNdr[q][2 * a], E.g.i00, T.dNdr[q][2 * a + 1] * E.g.i10);
    const double gay = fma(T.dNdr[q][2 * a], E.g.i01, T.dNdr[q][2 * a + 1] * E.g.i11);
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
__device__ __forceinline__ void th_pressure_row(const ThTables& T, const ThElem& E, int bp, double* Kp, double& Rp)
{
#pragma unroll
  for (int j = 0; j < 15; ++j) Kp[j] = 0.0;
  Rp = 0.0;
#pragma unroll 1
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

__device__ __forceinline__ double shfl1(double v) { return __shfl_xor(v, 1); }

struct ThAsmArgs
{
  const double* xv;
  const double* yv;
  const int*    vconn;
  const int*    dofs;
  const double* state;
  double        nu;
  const double* nu_e;
  const int*    inc_off;
  const int*    inc_elem;
  const uint4*  inc_meta;
  const int*    blk_row;
  const int*    blk_flags;
  const int*    row_ptr;
  double*       vals;
  double*       res;
};

// Two lanes per work item: lane role 0 takes the item's incidences t0, t0+2,
// ..., role 1 takes t0+1, t0+3, ...; after each round role 0 applies its own
// contribution and then its partner's (fetched by a cross-lane swap), so every
// CSR value still receives its element contributions in ascending element
// order, from one thread, in program order.
__device__ __forceinline__ void th_assemble_block(int b, const ThAsmArgs& a, const ThTables& T,
                                                  double* acc)
{
  const int rs     = a.blk_row[b], re = a.blk_row[b + 1];
  const bool pairs = a.blk_flags[b] != 0;
  const int base   = a.row_ptr[rs];
  const int bn     = a.vals ? a.row_ptr[re] - base : 0;
  for (int i = threadIdx.x; i < bn; i += kThLanes) acc[i] = 0.0;
  __syncthreads();

  const int role   = threadIdx.x & 1;
  const int nitems = pairs ? (re - rs) >> 1 : re - rs;
  for (int item = threadIdx.x >> 1; item < nitems; item += kThLanes / 2)
  {
    const int r  = rs + (pairs ? 2 * item : item);
    double*   ax = acc + (a.row_ptr[r] - base);
    double*   ay = pairs ? acc + (a.row_ptr[r + 1] - base) : nullptr;
    double    sx = 0.0, sy = 0.0;
    const int t0 = a.inc_off[r], t1 = a.inc_off[r + 1];
    for (int t = t0; t < t1; t += 2)
    {
      const int  tm   = t + role;
      const bool have = tm < t1;
      double     K0[15], K1[15], R0 = 0.0, R1 = 0.0;
      uint4      meta = make_uint4(0, 0, 0, 0);
      if (have)
      {
        const int e = a.inc_elem[tm];
        meta        = a.inc_meta[tm];
        const int li = meta.w >> 24;
        ThElem    E;
        th_gather(E, e, a.xv, a.yv, a.vconn, a.dofs, a.state, a.nu, a.nu_e);
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
      }
      else
      {
#pragma unroll
        for (int j = 0; j < 15; ++j) K0[j] = 0.0, K1[j] = 0.0;
      }
      // Apply in element order: role 0 (incidence t) first, then role 1
      // (incidence t+1).  The wave barrier keeps the compiler from merging or
      // reordering the two predicated LDS updates; a wave executes them in
      // program order.
      if (role == 0 && have)
      {
        sx = sx + R0;
        if (pairs) sy = sy + R1;
        if (a.vals)
        {
          th_add_row(ax, meta, K0);
          if (pairs) th_add_row(ay, meta, K1);
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (role == 1 && have && a.vals)
      {
        th_add_row(ax, meta, K0);
        if (pairs) th_add_row(ay, meta, K1);
      }
      __builtin_amdgcn_wave_barrier();
      const double pR0 = shfl1(R0), pR1 = shfl1(R1);
      if (role == 0 && t + 1 < t1)
      {
        sx = sx + pR0;
        if (pairs) sy = sy + pR1;
      }
    }
    if (role == 0)
    {
      __builtin_nontemporal_store(sx, a.res + r);
      if (pairs) __builtin_nontemporal_store(sy, a.res + r + 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < bn; i += kThLanes) __builtin_nontemporal_store(acc[i], a.vals + base + i);
  __syncthreads();
}

struct ThVjpArgs
{
  int           num_elems;
  const double* xv;
  const double* yv;
  const int*    vconn;
  const int*    dofs;
  const double* state;
  double        nu;
  const double* nu_e;
  const double* lam;
  const int*    bc_mask;
  double*       cbuf;
  double*       gout;
};

// Adjoint VJP, phase 1 (one lane per element): the element's contribution to
// every column of J^T lam~, c_e[j] = sum_i K^e_ij lam~_e[i], contracted at
// the quadrature points through the adjoint interpolant lambda_h (no K^e is
// formed), and the viscosity sensitivity g_e = lam~_e^T dR_e/dnu_e.
//   column (b,d): N_b (du_0/dx_d L_0 + du_1/dx_d L_1) + nu grad phi_b . grad L_d
//                 + L_d (u . grad phi_b) - P d_d phi_b
//   column p_b' : -psi_b' div L
__device__ __forceinline__ void th_vjp_elem_block(int blk, const ThVjpArgs& a, const ThTables& T)
{
  const int e = blk * kThLanes + threadIdx.x;
  if (e >= a.num_elems) return;
  ThElem E;
  th_gather(E, e, a.xv, a.yv, a.vconn, a.dofs, a.state, a.nu, a.nu_e);
  double     l[15], c[15];
  const int* d = a.dofs + 15 * static_cast<long>(e);
#pragma unroll
  for (int k = 0; k < 15; ++k)
  {
    const int dk = d[k];
    l[k]         = (a.bc_mask && a.bc_mask[dk]) ? 0.0 : a.lam[dk];
    c[k]         = 0.0;
  }
  double ge = 0.0;
#pragma unroll 1
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
#pragma unroll
    for (int b = 0; b < 6; ++b)
    {
      const double nb = T.N[q][b], bx = dNx[b], by = dNy[b];
      const double ub = fma(u0, bx, u1 * by);
      const double c0 = fma(nb, fma(d00, L0, d10 * L1), fma(E.nu, fma(bx, g00, by * g01), fma(L0, ub, -P * bx)));
      const double c1 = fma(nb, fma(d01, L0, d11 * L1), fma(E.nu, fma(bx, g10, by * g11), fma(L1, ub, -P * by)));
      c[2 * b]        = fma(c0, wq, c[2 * b]);
      c[2 * b + 1]    = fma(c1, wq, c[2 * b + 1]);
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) c[12 + b] = fma(-T.psi[q][b] * (g00 + g11), wq, c[12 + b]);
    ge = fma(fma(g00, d00, g01 * d01) + fma(g10, d10, g11 * d11), wq, ge);
  }
  double* out = a.cbuf + 15 * static_cast<long>(e);
#pragma unroll
  for (int k = 0; k < 15; ++k) out[k] = c[k];
  if (a.gout) __builtin_nontemporal_store(ge, a.gout + e);
}


constexpr int kGatherStage = 4096;

struct GatherArgs
{
  int           ndof;
  const int*    inc_off;
  const int*    inc_pos;
  const double* cbuf;
  const double* lam;
  const int*    bc_mask;
  double*       jt;
};

// Phase 2 (one lane per DOF): sum the element contributions of the DOF's
// column in ascending element order, then pass lam through on constrained
// DOFs.  The block's incidence range is contiguous and is staged in LDS with
// coalesced loads when it fits.
__device__ __forceinline__ void vjp_gather_block(int blk, const GatherArgs& a, int* spos)
{
  const int  c0 = blk * 256, c1 = min(c0 + 256, a.ndof);
  const int  i0 = a.inc_off[c0], ni = a.inc_off[c1] - i0;
  const bool staged = ni <= kGatherStage;
  if (staged)
  {
    for (int i = threadIdx.x; i < ni; i += 256) spos[i] = a.inc_pos[i0 + i];
    __syncthreads();
  }
  const int c = c0 + threadIdx.x;
  if (c < a.ndof)
  {
    double    v  = 0.0;
    const int t1 = a.inc_off[c + 1];
    for (int t = a.inc_off[c]; t < t1; ++t) v = v + a.cbuf[staged ? spos[t - i0] : a.inc_pos[t]];
    if (a.bc_mask && a.bc_mask[c]) v = v + a.lam[c];
    __builtin_nontemporal_store(v, a.jt + c);
  }
  __syncthreads();
}


} // namespace fxk
