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
// Assembly design: a work item is the x and y rows of one velocity node (they
// share every column position and every quadrature-point value) or one
// single row; its lanes (as many as the block's lane map gives it) gather
// over the row's incident elements in ascending element order.  A 64-lane
// block owns a contiguous range of rows; their CSR values are accumulated in
// an LDS tile and written back with coalesced non-temporal stores.  Each CSR
// value is a sum over its elements in ascending element order, term after
// term, so repeated launches give bitwise equal results.  Element geometry is
// recomputed from the three vertex coordinates; the element-row tensors are
// staged once per workgroup in LDS.  Arithmetic uses fused multiply-adds.
//
// VJP design: two phases.  One lane per element contracts every column of
// its element matrix with the boundary-masked adjoint through the adjoint's
// quadrature-point interpolant (lambda_h, grad lambda_h, pi_h) -- the
// quadrature-point state is formed once per element and J is never formed --
// and writes each column's value to its slot in the row-incidence order;
// then one lane per DOF sums its contiguous run of slots in ascending element
// order (a segmented reduction over a stream) instead of scattering with
// atomics.
#pragma once
#include "fx_dev.hpp"

namespace fxk
{


constexpr int kThLanes = 64;
constexpr int kThWaves = 4; // waves per assembly workgroup, sharing one copy of the tables


struct ThTables
{
  double N[7][6];     // P2 shape values
  double dNdr[7][12]; // reference gradients, (shape, r/s) row-major
  double psi[7][3];   // P1 pressure shapes
  double w[7];        // weights on the reference triangle (sum 1/2)
};

constexpr double kSqrt15 = 3.872983346207417; // nearest double to sqrt(15)

// 7-point rule: centroid and two 3-point orbits with barycentric
// a = (6 -+ sqrt 15)/21, weights 9/80 and (155 -+ sqrt 15)/2400 (Radon;
// Stroud T2:5-1).  P2 shapes l_i(2 l_i - 1) and 4 l_i l_j (edges 01, 12, 20).
constexpr ThTables make_th_tables()
{
  const double s15 = kSqrt15;
  const double a1 = (6.0 - s15) / 21.0, b1 = (9.0 + 2.0 * s15) / 21.0;
  const double a2 = (6.0 + s15) / 21.0, b2 = (9.0 - 2.0 * s15) / 21.0;
  const double w1 = (155.0 - s15) / 2400.0, w2 = (155.0 + s15) / 2400.0;
  const double qr[7] = {1.0 / 3.0, a1, b1, a1, a2, b2, a2};
  const double qs[7] = {1.0 / 3.0, a1, a1, b1, a2, a2, b2};
  const double qw[7] = {9.0 / 80.0, w1, w1, w1, w2, w2, w2};
  ThTables     T{};
  for (int q = 0; q < 7; ++q)
  {
    const double l0 = 1.0 - qr[q] - qs[q], l1 = qr[q], l2 = qs[q];
    T.w[q]    = qw[q];
    T.N[q][0] = l0 * (2.0 * l0 - 1.0);
    T.N[q][1] = l1 * (2.0 * l1 - 1.0);
    T.N[q][2] = l2 * (2.0 * l2 - 1.0);
    T.N[q][3] = 4.0 * l0 * l1;
    T.N[q][4] = 4.0 * l1 * l2;
    T.N[q][5] = 4.0 * l2 * l0;
    double* d = T.dNdr[q];
    d[0] = -(4.0 * l0 - 1.0), d[1] = -(4.0 * l0 - 1.0);
    d[2] = 4.0 * l1 - 1.0, d[3] = 0.0;
    d[4] = 0.0, d[5] = 4.0 * l2 - 1.0;
    d[6] = 4.0 * (l0 - l1), d[7] = -4.0 * l1;
    d[8] = 4.0 * l2, d[9] = 4.0 * l1;
    d[10] = -4.0 * l2, d[11] = 4.0 * (l0 - l2);
    T.psi[q][0] = l0;
    T.psi[q][1] = l1;
    T.psi[q][2] = l2;
  }
  return T;
}

// The tables are compile-time constants in device memory.  The VJP reads
// them at a quadrature point that is the same for every lane of a wave, so
// the reads are scalar loads: no LDS copy and no workgroup barrier.
__device__ const ThTables kThTables = make_th_tables();

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

// ---------------------------------------------------------------------------
// Assembly in reference-tensor form.  On an affine triangle every integrand
// of the Jacobian is a polynomial times products of the constant inverse
// map, so the 7-point sums factor into reference tensors (the rule applied
// on the reference triangle, once, at compile time) contracted with a few
// per-element coefficients.  With H = |det J| J^-1 (the signed adjugate):
//   stiffness  int grad phi_a . grad phi_b = sum_rs G_rs A^rs_ab,
//              G = H H^T / |det J|
//   convection int phi_a (u . grad phi_b)  = sum_rc V^r_c T^r_acb,
//              V^r_c = sum_d u_(c,d) H_rd (reference-frame nodal velocity)
//   reaction   int phi_a phi_b d_d u_c     = sum_v W^cd_v Q_abv,
//              W^cd_v = |det J| (d_d u_c)(vertex v) (grad u is linear)
//   pressure   -int psi_b' d_d phi_a       = -sum_r H_rd P^r_b'a
// with A^rs_ab = sum_q w_q dN_a/dr_r dN_b/dr_s, T^r_acb = sum_q w_q N_a N_c
// dN_b/dr_r, Q_abv = sum_q w_q N_a N_b lambda_v, P^r_b'a = sum_q w_q psi_b'
// dN_a/dr_r.  This is the same 7-point quadrature, summed in another order.

// Row a of the (a, b) tensors is read with a differing per lane; its stride
// of 122 doubles (244 dwords) puts the six rows' reads in distinct LDS banks.
constexpr int kThRowStride = 122;

struct ThTensor
{
  // Row a, block b at [20 b ..]: T^r_acb at [6 r + c] (12), Q_abv at [12 + v]
  // (3), pad, A^00_ab, A^11_ab, A^01_ab + A^10_ab at [16..18], pad.
  double ab[6][kThRowStride];
  double pa[6][6]; // P^r_b'a at [a][2 b' + r]
};
constexpr int kThTensorLen = static_cast<int>(sizeof(ThTensor) / sizeof(double));

constexpr double p2_shape(int a, double l0, double l1, double l2)
{
  switch (a)
  {
  case 0: return l0 * (2.0 * l0 - 1.0);
  case 1: return l1 * (2.0 * l1 - 1.0);
  case 2: return l2 * (2.0 * l2 - 1.0);
  case 3: return 4.0 * l0 * l1;
  case 4: return 4.0 * l1 * l2;
  default: return 4.0 * l2 * l0;
  }
}

// d N_a / d r (dir 0) or d N_a / d s (dir 1), with l0 = 1 - r - s, l1 = r, l2 = s.
constexpr double p2_dshape(int a, int dir, double l0, double l1, double l2)
{
  switch (a)
  {
  case 0: return -(4.0 * l0 - 1.0);
  case 1: return dir == 0 ? 4.0 * l1 - 1.0 : 0.0;
  case 2: return dir == 0 ? 0.0 : 4.0 * l2 - 1.0;
  case 3: return dir == 0 ? 4.0 * (l0 - l1) : -4.0 * l1;
  case 4: return dir == 0 ? 4.0 * l2 : 4.0 * l1;
  default: return dir == 0 ? -4.0 * l2 : 4.0 * (l0 - l2);
  }
}

constexpr ThTensor make_th_tensor()
{
  const double a1 = (6.0 - kSqrt15) / 21.0, b1 = (9.0 + 2.0 * kSqrt15) / 21.0;
  const double a2 = (6.0 + kSqrt15) / 21.0, b2 = (9.0 - 2.0 * kSqrt15) / 21.0;
  const double w1 = (155.0 - kSqrt15) / 2400.0, w2 = (155.0 + kSqrt15) / 2400.0;
  const double qr[7] = {1.0 / 3.0, a1, b1, a1, a2, b2, a2};
  const double qs[7] = {1.0 / 3.0, a1, a1, b1, a2, a2, b2};
  const double qw[7] = {9.0 / 80.0, w1, w1, w1, w2, w2, w2};
  ThTensor t{};
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b)
    {
      double* e = t.ab[a] + 20 * b;
      for (int q = 0; q < 7; ++q)
      {
        const double l1 = qr[q], l2 = qs[q], l0 = 1.0 - l1 - l2, w = qw[q];
        const double lam[3] = {l0, l1, l2};
        const double na = p2_shape(a, l0, l1, l2), nb = p2_shape(b, l0, l1, l2);
        const double ar = p2_dshape(a, 0, l0, l1, l2), as = p2_dshape(a, 1, l0, l1, l2);
        const double br = p2_dshape(b, 0, l0, l1, l2), bs = p2_dshape(b, 1, l0, l1, l2);
        for (int c = 0; c < 6; ++c)
        {
          const double wnn = w * na * p2_shape(c, l0, l1, l2);
          e[c]             = e[c] + wnn * br;
          e[6 + c]         = e[6 + c] + wnn * bs;
        }
        for (int v = 0; v < 3; ++v) e[12 + v] = e[12 + v] + w * na * nb * lam[v];
        e[16] = e[16] + w * ar * br;
        e[17] = e[17] + w * as * bs;
        e[18] = e[18] + (w * ar * bs + w * as * br);
      }
    }
  for (int a = 0; a < 6; ++a)
    for (int q = 0; q < 7; ++q)
    {
      const double l1 = qr[q], l2 = qs[q], l0 = 1.0 - l1 - l2, w = qw[q];
      const double lam[3] = {l0, l1, l2};
      for (int bp = 0; bp < 3; ++bp)
        for (int r = 0; r < 2; ++r)
          t.pa[a][2 * bp + r] = t.pa[a][2 * bp + r] + w * lam[bp] * p2_dshape(a, r, l0, l1, l2);
    }
  return t;
}

__device__ const ThTensor kThTensor = make_th_tensor();

// Stage the reference tensors into LDS (all lanes; caller synchronises).
__device__ __forceinline__ void th_stage_tensor(ThTensor& T, int nthreads)
{
  const double* src = &kThTensor.ab[0][0];
  double*       dst = &T.ab[0][0];
  for (int i = threadIdx.x; i < kThTensorLen; i += nthreads) dst[i] = src[i];
}

// Element data for assembly: the 15 local state values, H = |det J| J^-1,
// |det J| and the element viscosity.
struct ThAff
{
  double h00, h01, h10, h11, adet;
};

// th_geom from the three vertices' coordinates.
__device__ __forceinline__ ThGeom th_geom_xy(const double (&x)[3], const double (&y)[3])
{
  const double j00 = x[1] - x[0], j01 = x[2] - x[0], j10 = y[1] - y[0], j11 = y[2] - y[0];
  const double det = j00 * j11 - j01 * j10;
  const double id  = 1.0 / det;
  return {j11 * id, -j01 * id, -j10 * id, j00 * id, fabs(det)};
}

struct ThElemT
{
  double s[15];
  ThAff  g;
  double nu;
};

__device__ __forceinline__ void th_gather_t(ThElemT& E, int e, const double* __restrict__ xv,
                                            const double* __restrict__ yv,
                                            const int* __restrict__ vconn,
                                            const int* __restrict__ dofs,
                                            const double* __restrict__ state, double nu,
                                            const double* __restrict__ nu_e)
{
  const int* d = dofs + 15 * static_cast<long>(e);
#pragma unroll
  for (int k = 0; k < 15; ++k) E.s[k] = state[d[k]];
  const int*   v  = vconn + 3 * static_cast<long>(e);
  const double x0 = xv[v[0]], y0 = yv[v[0]];
  const double j00 = xv[v[1]] - x0, j01 = xv[v[2]] - x0, j10 = yv[v[1]] - y0, j11 = yv[v[2]] - y0;
  const double det = j00 * j11 - j01 * j10;
  const double sg  = det < 0.0 ? -1.0 : 1.0;
  E.g              = {sg * j11, -sg * j01, -sg * j10, sg * j00, fabs(det)};
  E.nu             = nu_e ? nu_e[e] : nu;
}

// The same element data loaded in three dependent stages -- incidence,
// element tables (DOFs, vertices), values -- so that a row's next rounds can
// be in flight while the current one computes.
struct ThIdx
{
  int   e;
  uint4 meta;
};

struct ThDofs
{
  int d[15];
  int v[3];
};

struct ThRaw
{
  double s[15];
  double x[3], y[3];
  double nu;
};

// Incidence t of the row-incidence tables; t must be a valid index (callers
// clamp to the row's last incidence and discard what they do not own).
__device__ __forceinline__ ThIdx th_load_idx(const int* __restrict__ inc_elem,
                                             const uint4* __restrict__ inc_meta, int t)
{
  return {gat32(inc_elem, t), gat32(inc_meta, t)};
}

__device__ __forceinline__ ThDofs th_load_dofs(const int* __restrict__ dofs,
                                               const int* __restrict__ vconn, int e)
{
  typedef int i4 __attribute__((ext_vector_type(4), aligned(4)));
  typedef int i3 __attribute__((ext_vector_type(3), aligned(4)));
  ThDofs     D;
  const int* d = dofs + 15 * static_cast<long>(e);
#pragma unroll
  for (int q = 0; q < 3; ++q)
  {
    const i4 w = *reinterpret_cast<const i4*>(d + 4 * q);
    D.d[4 * q] = w.x, D.d[4 * q + 1] = w.y, D.d[4 * q + 2] = w.z, D.d[4 * q + 3] = w.w;
  }
  const i3 t = *reinterpret_cast<const i3*>(d + 12);
  D.d[12] = t.x, D.d[13] = t.y, D.d[14] = t.z;
  const i3 v = *reinterpret_cast<const i3*>(vconn + 3 * static_cast<long>(e));
  D.v[0] = v.x, D.v[1] = v.y, D.v[2] = v.z;
  return D;
}

__device__ __forceinline__ void th_make_elem(ThElemT& E, const ThRaw& R)
{
#pragma unroll
  for (int k = 0; k < 15; ++k) E.s[k] = R.s[k];
  const double x0 = R.x[0], y0 = R.y[0];
  const double j00 = R.x[1] - x0, j01 = R.x[2] - x0, j10 = R.y[1] - y0, j11 = R.y[2] - y0;
  const double det = j00 * j11 - j01 * j10;
  const double sg  = det < 0.0 ? -1.0 : 1.0;
  E.g              = {sg * j11, -sg * j01, -sg * j10, sg * j00, fabs(det)};
  E.nu             = R.nu;
}

// Velocity rows (a,0) and (a,1) of the element Jacobian and residual:
//   K_(a,c),(b,d) = delta_cd (nu A_ab + C_ab) + M^cd_ab,
//   K_(a,c),b'    = -int psi_b' d_c phi_a,
//   R_(a,c)       = sum_b (nu A_ab + C_ab) u_(b,c) + sum_b' K_(a,c),b' p_b'.
__device__ __forceinline__ void th_velocity_rows(const ThTensor& T, const ThElemT& E, int a,
                                                 double* Kx, double* Ky, double& Rx, double& Ry)
{
  const ThAff&  g   = E.g;
  const double* s   = E.s;
  const double  nui = E.nu / g.adet;
  const double  G00 = nui * fma(g.h00, g.h00, g.h01 * g.h01);
  const double  G11 = nui * fma(g.h10, g.h10, g.h11 * g.h11);
  const double  G01 = nui * fma(g.h00, g.h10, g.h01 * g.h11);
  double        V[12];
#pragma unroll
  for (int c = 0; c < 6; ++c)
  {
    V[c]     = fma(s[2 * c], g.h00, s[2 * c + 1] * g.h01);
    V[6 + c] = fma(s[2 * c], g.h10, s[2 * c + 1] * g.h11);
  }
  // Reference gradients of u_c at the vertices (exact for P2), then W^cd_v.
  double W[4][3];
#pragma unroll
  for (int c = 0; c < 2; ++c)
  {
    const double u0 = s[c], u1 = s[2 + c], u2 = s[4 + c], u3 = s[6 + c], u4 = s[8 + c],
                 u5 = s[10 + c];
    const double er[3] = {4.0 * u3 - 3.0 * u0 - u1, (u0 + 3.0 * u1) - 4.0 * u3,
                          (u0 - u1) + 4.0 * (u4 - u5)};
    const double es[3] = {4.0 * u5 - 3.0 * u0 - u2, (u0 - u2) + 4.0 * (u4 - u3),
                          (u0 + 3.0 * u2) - 4.0 * u5};
#pragma unroll
    for (int v = 0; v < 3; ++v)
    {
      W[2 * c][v]     = fma(er[v], g.h00, es[v] * g.h10);
      W[2 * c + 1][v] = fma(er[v], g.h01, es[v] * g.h11);
    }
  }
  Rx = 0.0, Ry = 0.0;
  const double* ta = T.ab[a];
#pragma unroll
  for (int b = 0; b < 6; ++b)
  {
    const double* e = ta + 20 * b;
    double        L = fma(G00, e[16], fma(G11, e[17], G01 * e[18]));
#pragma unroll
    for (int k = 0; k < 12; ++k) L = fma(V[k], e[k], L);
    double M[4];
#pragma unroll
    for (int cd = 0; cd < 4; ++cd) M[cd] = fma(W[cd][0], e[12], fma(W[cd][1], e[13], W[cd][2] * e[14]));
    Kx[2 * b]     = L + M[0];
    Kx[2 * b + 1] = M[1];
    Ky[2 * b]     = M[2];
    Ky[2 * b + 1] = L + M[3];
    Rx            = fma(L, s[2 * b], Rx);
    Ry            = fma(L, s[2 * b + 1], Ry);
  }
  const double* pa = T.pa[a];
#pragma unroll
  for (int bp = 0; bp < 3; ++bp)
  {
    Kx[12 + bp] = -fma(g.h00, pa[2 * bp], g.h10 * pa[2 * bp + 1]);
    Ky[12 + bp] = -fma(g.h01, pa[2 * bp], g.h11 * pa[2 * bp + 1]);
    Rx          = fma(Kx[12 + bp], s[12 + bp], Rx);
    Ry          = fma(Ky[12 + bp], s[12 + bp], Ry);
  }
}

// Pressure row b': K_b',(b,d) = -int psi_b' d_d phi_b, K_b',b'' = 0,
// R_b' = -int psi_b' div u = sum_(b,d) K_b',(b,d) u_(b,d).
__device__ __forceinline__ void th_pressure_row(const ThTensor& T, const ThElemT& E, int bp,
                                                double* Kp, double& Rp)
{
  const ThAff& g = E.g;
  Rp             = 0.0;
#pragma unroll
  for (int b = 0; b < 6; ++b)
  {
    const double pr = T.pa[b][2 * bp], ps = T.pa[b][2 * bp + 1];
    Kp[2 * b]       = -fma(g.h00, pr, g.h10 * ps);
    Kp[2 * b + 1]   = -fma(g.h01, pr, g.h11 * ps);
    Rp              = fma(Kp[2 * b], E.s[2 * b], fma(Kp[2 * b + 1], E.s[2 * b + 1], Rp));
  }
#pragma unroll
  for (int j = 12; j < 15; ++j) Kp[j] = 0.0;
}

// Add one element row -- and, for a velocity pair, the partner row, which
// has the same column positions -- into the LDS tile.  Each update is an
// LDS floating-point atomic add issued without waiting for a result: the 15
// positions of an element row are distinct CSR entries and the lanes that
// add at the same time own distinct rows, so no two updates in one
// instruction meet, and one wave's LDS operations execute in issue order.
// Every entry therefore still becomes old + K, in program order, with one
// LDS instruction instead of a read, a wait, an add and a write.
__device__ __forceinline__ void th_add_rows(double* ax, double* ay, uint4 meta, const double* Kx,
                                            const double* Ky, bool pairs)
{
  const unsigned w[4] = {meta.x, meta.y, meta.z, meta.w};
  int            pos[15];
#pragma unroll
  for (int j = 0; j < 15; ++j) pos[j] = (w[j >> 2] >> (8 * (j & 3))) & 0xff;
#pragma unroll
  for (int j = 0; j < 15; ++j) lds_add(ax + pos[j], Kx[j]);
  if (pairs)
  {
#pragma unroll
    for (int j = 0; j < 15; ++j) lds_add(ay + pos[j], Ky[j]);
  }
}


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
  const int*    blk_lanes; // 64 per block (fx_map_th_lanes), or null: two lanes per item
  const int*    row_ptr;
  double*       vals;
  double*       res;
  // round tables (fx_map_th_rounds), or null; a batch's walks over them
  // take blocks first, first + rt_stride, ... < rt_blocks
  const int*    rt_off;
  const int*    rt_elem;
  const uint4*  rt_meta;
  int           rt_blocks, rt_stride;
};

__device__ __forceinline__ ThRaw th_load_raw(const ThAsmArgs& a, const ThDofs& D, int e)
{
  ThRaw R;
#pragma unroll
  for (int k = 0; k < 15; ++k) R.s[k] = gat32(a.state, D.d[k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) R.x[k] = gat32(a.xv, D.v[k]), R.y[k] = gat32(a.yv, D.v[k]);
  R.nu = a.nu_e ? gat32(a.nu_e, e) : a.nu;
  return R;
}

// One row block per wave.  A block's items share its 64 lanes by the lane
// map (fx_map_th_lanes): item i has k_i consecutive lanes, and its lane of
// role j takes the item's incidences t0+j, t0+j+k_i, ...  Each round the
// roles add their element rows into the tile in sub-steps, role 0 first,
// then role 1, ..., so every CSR value receives its element contributions in
// ascending element order, in program order (the adds are LDS atomics only
// to skip the read round trip; no two lanes add to one entry at once).  The
// residual sums take the other roles' values by cross-lane reads, in the
// same order.  Without a lane map an item has two lanes.  `acc` is the
// wave's own row tile; the waves of a workgroup share only the tables.
//
// Loads are unconditional: a lane whose next incidence lies past the row's
// end loads the row's last incidence instead (always valid) and discards the
// result, so no load is predicated and nothing is zero-filled.
__device__ __forceinline__ void th_assemble_block(int b, const ThAsmArgs& a, const ThTensor& T,
                                                  double* acc)
{
  const int  lane  = threadIdx.x & (kThLanes - 1);
  const int  rs    = a.blk_row[b], re = a.blk_row[b + 1];
  const bool pairs = a.blk_flags[b] != 0;
  const int  base  = a.row_ptr[rs];
  const int  bn    = a.vals ? a.row_ptr[re] - base : 0;
  for (int i = lane; i < bn; i += kThLanes) acc[i] = 0.0;
  wave_sync();

  const int nitems = pairs ? (re - rs) >> 1 : re - rs;
  int       item, role, k, step;
  if (a.blk_lanes)
  {
    const int lm = a.blk_lanes[kThLanes * b + lane];
    item = lm < 0 ? nitems : lm & 0xff;
    role = (lm >> 8) & 0xff;
    k    = lm < 0 ? 1 : max(1, lm >> 16); // a stride of at least 1: every round loop ends
    step = nitems;                        // one item per lane
  }
  else
  {
    item = lane >> 1, role = lane & 1, k = 2, step = kThLanes / 2;
  }
  int kmax = k; // roles of the widest item: the sub-steps of a round
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
  for (; item < nitems; item += step)
  {
    const int r  = rs + (pairs ? 2 * item : item);
    double*   ax = acc + (a.row_ptr[r] - base);
    double*   ay = pairs ? acc + (a.row_ptr[r + 1] - base) : nullptr;
    double    sx = 0.0, sy = 0.0;
    const int t0 = a.inc_off[r], t1 = a.inc_off[r + 1], tl = t1 - 1;
    if (t0 < t1)
    {
      // software pipeline over this lane's incidences t0+role, t0+role+k, ...:
      // values of the next round and tables of the one after are in flight
      // while the current round computes
      ThIdx  ic = th_load_idx(a.inc_elem, a.inc_meta, min(t0 + role, tl));
      ThIdx  in = th_load_idx(a.inc_elem, a.inc_meta, min(t0 + role + k, tl));
      ThRaw  ra = th_load_raw(a, th_load_dofs(a.dofs, a.vconn, ic.e), ic.e), rb;
      ThDofs dn = th_load_dofs(a.dofs, a.vconn, in.e);
      // one round (incidences t .. t+k-1 of the item): compute this lane's
      // from `cur` while the next round's values land in `nxt` (the loop
      // alternates the two buffers, so no values are copied between rounds)
      auto round = [&](int t, const ThRaw& cur, ThRaw& nxt) {
        const int   tm   = t + role;
        const bool  have = tm < t1;
        const ThIdx in2  = th_load_idx(a.inc_elem, a.inc_meta, min(tm + 2 * k, tl));
        nxt              = th_load_raw(a, dn, in.e);
        double      K0[15], K1[15], R0, R1 = 0.0;
        const uint4 meta = ic.meta;
        {
          const int li = meta.w >> 24;
          ThElemT   E;
          th_make_elem(E, cur);
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
        // Apply in element order: role 0 (incidence t), then role 1 (t+1),
        // ..., with the wave's LDS accesses kept in program order.
        for (int j = 0; j < kmax; ++j)
        {
          if (role == j && have && a.vals) th_add_rows(ax, ay, meta, K0, K1, pairs);
          wave_sync();
        }
        if (role == 0)
        {
          sx = sx + R0;
          if (pairs) sy = sy + R1;
        }
        for (int j = 1; j < kmax; ++j)
        {
          const double pR0 = __shfl(R0, lane + j), pR1 = __shfl(R1, lane + j);
          if (role == 0 && j < k && t + j < t1)
          {
            sx = sx + pR0;
            if (pairs) sy = sy + pR1;
          }
        }
        dn = th_load_dofs(a.dofs, a.vconn, in2.e);
        ic = in;
        in = in2;
      };
      for (int t = t0;;)
      {
        round(t, ra, rb);
        if ((t += k) >= t1) break;
        round(t, rb, ra);
        if ((t += k) >= t1) break;
      }
    }
    if (role == 0)
    {
      __builtin_nontemporal_store(sx, a.res + r);
      if (pairs) __builtin_nontemporal_store(sy, a.res + r + 1);
    }
  }
  wave_sync();
  for (int i = lane; i < bn; i += kThLanes) __builtin_nontemporal_store(acc[i], a.vals + base + i);
  wave_sync();
}

// The velocity rows (or single row) of one incidence.
__device__ __forceinline__ void th_incidence_rows(const ThTensor& T, const ThRaw& cur, unsigned li,
                                                  bool pairs, double* K0, double* K1, double& R0,
                                                  double& R1)
{
  ThElemT E;
  th_make_elem(E, cur);
  R1 = 0.0;
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

// A position in a wave's walk over the round tables (fx_map_th_rounds):
// block b (the wave takes blocks first, first + stride, ... < end), round m of
// its R, whose entries start at q.  Every field is the same on all lanes.
struct ThCur
{
  int b, m, R, q;
};

__device__ __forceinline__ ThCur th_cur_at(const ThAsmArgs& a, int b)
{
  const int q = a.rt_off[b];
  return {b, 0, (a.rt_off[b + 1] - q) / kThLanes, q};
}

// The position after c; past the walk's last one it stays there (its loads
// stay valid and are never used).
__device__ __forceinline__ ThCur th_cur_next(const ThAsmArgs& a, const ThCur& c, int end,
                                             int stride)
{
  if (c.m + 1 < c.R) return {c.b, c.m + 1, c.R, c.q};
  if (c.b + stride < end) return th_cur_at(a, c.b + stride);
  return c;
}

__device__ __forceinline__ ThIdx th_cur_idx(const ThAsmArgs& a, const ThCur& c, int lane)
{
  const int i = c.q + c.m * kThLanes + lane;
  return {a.rt_elem[i], a.rt_meta[i]};
}

// A lane's part in one block: its item's row(s) in the tile, its role and
// its item's lane count k (the lane map), and the block's widest k.
struct ThLane
{
  double* ax;
  double* ay;
  int     r, role, k, kmax, base, bn;
  bool    own, pairs;
};

// What a lane reads of block b before it can add into the block's tile: the
// block's row range and kind, its lane-map entry, and its item's row offsets.
// The walk loads a block's part one block ahead (th_pre_lanes at the start
// of the block before it, th_pre_rows after that block's first round).
struct ThPre
{
  int rs, re, pairs, lm, rx, ry, base, bend;
};

// PreCsr: the block's CSR range is loaded here too (one block ahead in a
// walk); the batch's walk kernel, short of scalar registers, loads it when
// the block starts instead.
template <bool PreCsr = true>
__device__ __forceinline__ ThPre th_pre_lanes(const ThAsmArgs& a, int b, int lane)
{
  ThPre p;
  p.rs    = a.blk_row[b];
  p.re    = a.blk_row[b + 1];
  p.pairs = a.blk_flags[b] != 0;
  p.lm    = a.blk_lanes[kThLanes * b + lane];
  p.base  = PreCsr ? a.row_ptr[p.rs] : 0;
  p.bend  = PreCsr ? a.row_ptr[p.re] : 0;
  p.rx = p.ry = 0;
  return p;
}

// The lane's item row (or the block's first row for a lane without one).
__device__ __forceinline__ int th_pre_row(const ThPre& p)
{
  const int nitems = p.pairs ? (p.re - p.rs) >> 1 : p.re - p.rs;
  const bool own   = p.lm >= 0 && (p.lm & 0xff) < nitems;
  return own ? p.rs + (p.pairs ? 2 * (p.lm & 0xff) : p.lm & 0xff) : p.rs;
}

__device__ __forceinline__ void th_pre_rows(const ThAsmArgs& a, ThPre& p)
{
  const int r = th_pre_row(p);
  p.rx        = a.row_ptr[r];
  p.ry        = p.pairs ? a.row_ptr[r + 1] : 0;
}

// Set up a block from its preloaded part: the lane's item, and (unless the
// tile is known to be zero already) the zeroed tile.
template <bool PreCsr = true>
__device__ __forceinline__ ThLane th_lane_begin(const ThAsmArgs& a, const ThPre& p, double* acc,
                                                int lane, bool zero = true)
{
  ThLane    L;
  L.pairs          = p.pairs != 0;
  L.base           = PreCsr ? p.base : a.row_ptr[p.rs];
  L.bn             = a.vals ? (PreCsr ? p.bend : a.row_ptr[p.re]) - L.base : 0;
  const int nitems = L.pairs ? (p.re - p.rs) >> 1 : p.re - p.rs;
  L.own            = p.lm >= 0 && (p.lm & 0xff) < nitems;
  L.role           = (p.lm >> 8) & 0xff;
  L.k              = L.own ? max(1, p.lm >> 16) : 1;
  L.r              = th_pre_row(p);
  int kmax         = L.k;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
  L.kmax = kmax;
  L.ax   = acc + (p.rx - L.base);
  L.ay   = L.pairs ? acc + (p.ry - L.base) : nullptr;
  if (zero)
  {
    for (int i = lane; i < L.bn; i += kThLanes) acc[i] = 0.0;
    wave_sync();
  }
  return L;
}

__device__ __forceinline__ ThLane th_lane_begin(const ThAsmArgs& a, int b, double* acc, int lane)
{
  ThPre p = th_pre_lanes(a, b, lane);
  th_pre_rows(a, p);
  return th_lane_begin(a, p, acc, lane);
}

// Finish a block: residuals of the lane's item, then the tile's write-back,
// which leaves the tile's used range zero again.
__device__ __forceinline__ void th_lane_end(const ThAsmArgs& a, const ThLane& L, double sx, double sy,
                                            double* acc, int lane)
{
  if (L.own && L.role == 0)
  {
    __builtin_nontemporal_store(sx, a.res + L.r);
    if (L.pairs) __builtin_nontemporal_store(sy, a.res + L.r + 1);
  }
  wave_sync();
  for (int i = lane; i < L.bn; i += kThLanes)
  {
    __builtin_nontemporal_store(acc[i], a.vals + L.base + i);
    acc[i] = 0.0;
  }
  wave_sync();
}

// th_assemble_block over the round tables, for a wave that walks blocks
// first, first + stride, ... < end.  Lane l's incidence of round m of block b
// is entry 64 m + l of the block's range, so the incidences of the next
// rounds are read at addresses known from the walk alone, and the pipeline
// (incidence two rounds ahead, element tables and values one round ahead)
// runs on across block boundaries: the next block's first element loads are
// in flight while this block's last round computes.  Every lane runs each
// block's rounds, adding only what its entries do not mark idle.  The adds,
// residual sums and write-back are th_assemble_block's, in the same order:
// the results are bitwise equal to it.
// The wave's whole tile cleared (once per kernel: a walk takes a zero tile
// and leaves it zero, every block's write-back clearing what it used).
__device__ __forceinline__ void th_tile_clear(double* acc, int tile)
{
  for (int i = threadIdx.x & (kThLanes - 1); i < tile; i += kThLanes) acc[i] = 0.0;
  wave_sync();
}

// The tile must be zero on entry (th_tile_clear); it is zero again on exit.
template <bool PreCsr = true>
__device__ __forceinline__ void th_assemble_walk_rt(const ThAsmArgs& a, const ThTensor& T,
                                                    double* acc, int first, int end, int stride)
{
  const int lane = threadIdx.x & (kThLanes - 1);
  if (first >= end) return;
  ThCur c0 = th_cur_at(a, first);
  if (c0.R == 0)
  {
    // no incidences in the mesh (otherwise every block has a round)
    for (int b = first; b < end; b += stride)
      th_lane_end(a, th_lane_begin(a, b, acc, lane), 0.0, 0.0, acc, lane);
    return;
  }
  ThCur  c1 = th_cur_next(a, c0, end, stride), c2 = th_cur_next(a, c1, end, stride);
  ThIdx  ic = th_cur_idx(a, c0, lane), in = th_cur_idx(a, c1, lane);
  ThRaw  ra = th_load_raw(a, th_load_dofs(a.dofs, a.vconn, ic.e), ic.e), rb;
  ThDofs dn = th_load_dofs(a.dofs, a.vconn, in.e);
  ThPre  p0 = th_pre_lanes<PreCsr>(a, c0.b, lane);
  th_pre_rows(a, p0);
  ThLane L  = th_lane_begin<PreCsr>(a, p0, acc, lane, false);
  // the next block's part, one block ahead
  const bool more = first + stride < end;
  ThPre      pn   = th_pre_lanes<PreCsr>(a, more ? first + stride : first, lane);
  double sx = 0.0, sy = 0.0;
  bool   done  = false;
  auto   round = [&](const ThRaw& cur, ThRaw& nxt) {
    const ThIdx in2 = th_cur_idx(a, c2, lane);
    nxt             = th_load_raw(a, dn, in.e);
    const uint4 meta = ic.meta;
    const bool  have = L.own && !(meta.w >> 31);
    double      K0[15], K1[15], R0, R1;
    th_incidence_rows(T, cur, (meta.w >> 24) & 0x7f, L.pairs, K0, K1, R0, R1);
    for (int j = 0; j < L.kmax; ++j)
    {
      if (L.role == j && have && a.vals) th_add_rows(L.ax, L.ay, meta, K0, K1, L.pairs);
      wave_sync();
    }
    if (L.role == 0 && have)
    {
      sx = sx + R0;
      if (L.pairs) sy = sy + R1;
    }
    for (int j = 1; j < L.kmax; ++j)
    {
      const double pR0 = __shfl(R0, lane + j), pR1 = __shfl(R1, lane + j);
      const int    ph  = __shfl(static_cast<int>(have), lane + j);
      if (L.role == 0 && j < L.k && ph)
      {
        sx = sx + pR0;
        if (L.pairs) sy = sy + pR1;
      }
    }
    dn = th_load_dofs(a.dofs, a.vconn, in2.e);
    ic = in;
    in = in2;
    if (c0.m == 0) th_pre_rows(a, pn); // its lane-map entry has arrived by now
    if (c0.m + 1 >= c0.R)
    {
      th_lane_end(a, L, sx, sy, acc, lane);
      if (c0.b + stride >= end)
      {
        done = true;
        return;
      }
      sx = 0.0, sy = 0.0;
      L  = th_lane_begin<PreCsr>(a, pn, acc, lane, false);
      const int nb = c0.b + 2 * stride;
      pn           = th_pre_lanes<PreCsr>(a, nb < end ? nb : c0.b + stride, lane);
    }
    c0 = c1;
    c1 = c2;
    c2 = th_cur_next(a, c2, end, stride);
  };
  for (;;)
  {
    round(ra, rb);
    if (done) break;
    round(rb, ra);
    if (done) break;
  }
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
  const int*    elem_inc; // (e, li) -> row-incidence index (fx_map_export_rowinc)
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
// Element data of the VJP: the velocity state, the boundary-masked adjoint
// (velocity, then pressure), the affine geometry and the viscosity.
struct ThVjpElem
{
  double s[12];
  double l[15];
  ThGeom g;
  double nu;
};

__device__ __forceinline__ void th_vjp_math(const ThVjpElem& E, const ThTables& T, double (&c)[15],
                                            double& ge)
{
  const double(&l)[15] = E.l;
#pragma unroll
  for (int k = 0; k < 15; ++k) c[k] = 0.0;
  ge = 0.0;
  const ThGeom& G  = E.g;
  const double  nu = E.nu;
  // At each point the columns share three coefficients per component d:
  //   c_(b,d) += w (N_b Y_d + d_x phi_b X_xd + d_y phi_b X_yd),
  //   Y_d  = du_0/dx_d L_0 + du_1/dx_d L_1            (reaction)
  //   X_ed = nu dL_d/dx_e + L_d u_e - P delta_ed      (viscous, convective,
  //                                                    pressure)
  // and with d_e phi_b = sum_r dN_b/dr_r Jinv_re the last two terms are
  // dN_b/dr R_d + dN_b/ds S_d, R_d and S_d the rows of Jinv X: the shape
  // tables enter only as coefficients, and every gradient is formed from
  // the reference gradients once per point.
#pragma unroll 1
  for (int q = 0; q < 7; ++q)
  {
    const double* N  = T.N[q];
    const double* dN = T.dNdr[q];
    double u0 = 0.0, u1 = 0.0, L0 = 0.0, L1 = 0.0;
    double ur0 = 0.0, us0 = 0.0, ur1 = 0.0, us1 = 0.0, lr0 = 0.0, ls0 = 0.0, lr1 = 0.0, ls1 = 0.0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
    {
      const double na = N[a], dr = dN[2 * a], ds = dN[2 * a + 1];
      const double sx = E.s[2 * a], sy = E.s[2 * a + 1], lx = l[2 * a], ly = l[2 * a + 1];
      u0 = fma(na, sx, u0), u1 = fma(na, sy, u1), L0 = fma(na, lx, L0), L1 = fma(na, ly, L1);
      ur0 = fma(dr, sx, ur0), us0 = fma(ds, sx, us0), ur1 = fma(dr, sy, ur1), us1 = fma(ds, sy, us1);
      lr0 = fma(dr, lx, lr0), ls0 = fma(ds, lx, ls0), lr1 = fma(dr, ly, lr1), ls1 = fma(ds, ly, ls1);
    }
    double P = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) P = fma(T.psi[q][a], l[12 + a], P);
    // physical gradients: d_cx = du_c/dx, d_cy = du_c/dy; g likewise for L
    const double d00 = fma(G.i00, ur0, G.i10 * us0), d01 = fma(G.i01, ur0, G.i11 * us0);
    const double d10 = fma(G.i00, ur1, G.i10 * us1), d11 = fma(G.i01, ur1, G.i11 * us1);
    const double g00 = fma(G.i00, lr0, G.i10 * ls0), g01 = fma(G.i01, lr0, G.i11 * ls0);
    const double g10 = fma(G.i00, lr1, G.i10 * ls1), g11 = fma(G.i01, lr1, G.i11 * ls1);
    const double wq  = G.adet * T.w[q];
    const double Y0 = wq * fma(d00, L0, d10 * L1), Y1 = wq * fma(d01, L0, d11 * L1);
    const double Xx0 = fma(nu, g00, fma(L0, u0, -P)), Xy0 = fma(nu, g01, L0 * u1);
    const double Xx1 = fma(nu, g10, L1 * u0), Xy1 = fma(nu, g11, fma(L1, u1, -P));
    const double R0 = wq * fma(G.i00, Xx0, G.i01 * Xy0), S0 = wq * fma(G.i10, Xx0, G.i11 * Xy0);
    const double R1 = wq * fma(G.i00, Xx1, G.i01 * Xy1), S1 = wq * fma(G.i10, Xx1, G.i11 * Xy1);
#pragma unroll
    for (int b = 0; b < 6; ++b)
    {
      const double nb = N[b], dr = dN[2 * b], ds = dN[2 * b + 1];
      c[2 * b]        = fma(nb, Y0, fma(dr, R0, fma(ds, S0, c[2 * b])));
      c[2 * b + 1]    = fma(nb, Y1, fma(dr, R1, fma(ds, S1, c[2 * b + 1])));
    }
    const double dv = -wq * (g00 + g11); // -psi_b' div L
#pragma unroll
    for (int b = 0; b < 3; ++b) c[12 + b] = fma(T.psi[q][b], dv, c[12 + b]);
    ge = fma(fma(g00, d00, g01 * d01) + fma(g10, d10, g11 * d11), wq, ge);
  }
}

__device__ __forceinline__ void th_vjp_elem_cols(int e, const ThVjpArgs& a, const ThTables& T,
                                                 double (&c)[15], double& ge)
{
  // the DOF row and vertices in vector loads, then one 32-bit-offset gather
  // per value (the pressure state does not enter J^T lam or dR/dnu)
  const ThDofs D = th_load_dofs(a.dofs, a.vconn, e);
  ThVjpElem    E;
#pragma unroll
  for (int k = 0; k < 12; ++k) E.s[k] = gat32(a.state, D.d[k]);
  E.g  = th_geom(a.xv, a.yv, D.v[0], D.v[1], D.v[2]);
  E.nu = a.nu_e ? a.nu_e[e] : a.nu;
#pragma unroll
  for (int k = 0; k < 15; ++k)
  {
    const int dk = D.d[k];
    // both gathers issued together (dk < ndof for both arrays), then a select
    const double lv = gat32(a.lam, dk);
    const int    mk = a.bc_mask ? gat32(a.bc_mask, dk) : 0;
    E.l[k]          = mk ? 0.0 : lv;
  }
  th_vjp_math(E, T, c, ge);
}

__device__ __forceinline__ void th_vjp_elem_block(int blk, const ThVjpArgs& a, const ThTables& T)
{
  const int e = blk * kThLanes + threadIdx.x;
  if (e >= a.num_elems) return;
  double c[15], ge;
  th_vjp_elem_cols(e, a, T, c, ge);
  // each column's value goes to its slot in the row-incidence order, where
  // phase 2 finds a DOF's contributions contiguous
  const int* slot = a.elem_inc + 15 * static_cast<long>(e);
#pragma unroll
  for (int k = 0; k < 15; ++k) a.cbuf[slot[k]] = c[k];
  if (a.gout) __builtin_nontemporal_store(ge, a.gout + e);
}

// ---------------------------------------------------------------------------
// Single-pass VJP over element patches (fx_th_patches_create).  A workgroup
// of kPatchLanes lanes takes one patch: it loads the patch's local DOF values
// (state, masked adjoint) into LDS, one gather per local DOF, then, one
// element per lane, forms the element's columns as above from them and
// stores each into its slot of the patch's LDS slot array, where every local
// DOF owns a run of slots in element order.  After a barrier the lanes sum the runs, each in slot order; a DOF
// held by this patch alone is finished here (jt, with lam passed through on
// constrained DOFs), any other gets this patch's partial in its seam slot.
// LDS of one patch workgroup, in bytes: the staged local DOF values (later
// overwritten by the slot array), each run's start (u16, one past the last
// local DOF too) and each local DOF's constraint flag (u8).
__host__ __device__ constexpr int th_patch_lds_main(int max_patch_ldofs)
{
  return 8 * (2 * max_patch_ldofs > 15 * kPatchLanes ? 2 * max_patch_ldofs : 15 * kPatchLanes);
}
__host__ __device__ constexpr int th_patch_lds(int max_patch_ldofs)
{
  return (th_patch_lds_main(max_patch_ldofs) + 3 * max_patch_ldofs + 2 + 15) / 16 * 16;
}

__device__ __forceinline__ void th_vjp_patch_block(int p, const ThVjpArgs& a, const PatchArgs& q,
                                                   const ThTables& T, unsigned char* lds,
                                                   int max_patch_ldofs)
{
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  typedef int      i3 __attribute__((ext_vector_type(3), aligned(4)));
  constexpr int    U = kPatchUnroll;
  const int        t  = threadIdx.x;
  const int        i0 = q.patch_off[p], ne = q.patch_off[p + 1] - i0;
  const int        l0 = q.ldof_off[p], nl = q.ldof_off[p + 1] - l0;
  const bool       mine = t < ne;
  const int        i    = i0 + t;
  double*          sv   = reinterpret_cast<double*>(lds); // state, then the slot array
  double*          lv   = sv + nl;                        // masked adjoint
  unsigned short*  rs   = reinterpret_cast<unsigned short*>(lds + th_patch_lds_main(max_patch_ldofs));
  unsigned char*   mk   = reinterpret_cast<unsigned char*>(rs + max_patch_ldofs + 1);
  const u4*        lp   = reinterpret_cast<const u4*>(q.elem_ldof + 16 * static_cast<long>(i));
  // the element's tables and geometry are requested first, so that their
  // latency overlaps the staging below
  int      e = 0;
  unsigned w[8];
  double   xg[3], yg[3], nu = a.nu;
  if (mine)
  {
    e           = q.patch_elem[i];
    const u4 x0 = lp[0], x1 = lp[1];
    w[0] = x0.x, w[1] = x0.y, w[2] = x0.z, w[3] = x0.w, w[4] = x1.x, w[5] = x1.y, w[6] = x1.z,
    w[7]       = x1.w;
    const i3 v = *reinterpret_cast<const i3*>(q.elem_vert + 3 * static_cast<long>(i));
    xg[0] = gat32(a.xv, v.x), xg[1] = gat32(a.xv, v.y), xg[2] = gat32(a.xv, v.z);
    yg[0] = gat32(a.yv, v.x), yg[1] = gat32(a.yv, v.y), yg[2] = gat32(a.yv, v.z);
    if (a.nu_e) nu = a.nu_e[e];
  }
  // the patch's local DOF values, each gathered once (state, masked
  // adjoint), their constraint flags and the runs' starts
  if (t == 0) rs[0] = 0;
  for (int j0 = t; j0 < nl; j0 += U * kPatchLanes)
  {
    int  d[U], re[U];
    bool on[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      on[u]        = j0 + u * kPatchLanes < nl;
      const int  j = l0 + j0 + u * kPatchLanes;
      d[u]         = on[u] ? q.ldof_dof[j] : 0;
      re[u]        = on[u] ? q.ldof_end[j] : 0;
    }
    double sd[U], ld[U];
    int    m[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      sd[u] = on[u] ? gat32(a.state, d[u]) : 0.0;
      ld[u] = on[u] ? gat32(a.lam, d[u]) : 0.0;
      m[u]  = on[u] && a.bc_mask ? gat32(a.bc_mask, d[u]) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j0 + u * kPatchLanes < nl)
      {
        const int j = j0 + u * kPatchLanes;
        sv[j]       = sd[u];
        lv[j]       = m[u] ? 0.0 : ld[u];
        rs[j + 1]   = static_cast<unsigned short>(re[u]);
        mk[j]       = m[u] ? 1 : 0;
      }
  }
  __syncthreads();
  double c[15], ge = 0.0;
  if (mine)
  {
    ThVjpElem E;
#pragma unroll
    for (int k = 0; k < 15; ++k)
    {
      const unsigned j = (w[k >> 1] >> (16 * (k & 1))) & 0xfffu;
      if (k < 12) E.s[k] = sv[j];
      E.l[k] = lv[j];
    }
    E.g  = th_geom_xy(xg, yg);
    E.nu = nu;
    th_vjp_math(E, T, c, ge);
  }
  __syncthreads(); // every lane has read the staged values: the slots overwrite them
  double* cs = sv;
  if (mine)
  {
    const u4       x0   = lp[0], x1 = lp[1];
    const unsigned o[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (int k = 0; k < 15; ++k)
    {
      const unsigned h = (o[k >> 1] >> (16 * (k & 1))) & 0xffffu;
      cs[rs[h & 0xfffu] + (h >> 12)] = c[k];
    }
    if (a.gout) __builtin_nontemporal_store(ge, a.gout + e);
  }
  __syncthreads();
  for (int j0 = t; j0 < nl; j0 += U * kPatchLanes)
  {
    int d[U], part[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      const bool on = j0 + u * kPatchLanes < nl;
      const int  j  = l0 + j0 + u * kPatchLanes;
      d[u]          = on ? q.ldof_dof[j] : 0;
      part[u]       = on ? q.ldof_part[j] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      const int j = j0 + u * kPatchLanes;
      if (j < nl) patch_finish(cs, rs[j], rs[j + 1], d[u], part[u], mk[j] != 0, a.lam, q.jt, q.parts);
    }
  }
  __syncthreads(); // the LDS is reused by the next patch of a persistent walk
}

constexpr int kGatherStage = 4096;

struct GatherArgs
{
  int           ndof;
  const int*    inc_off;
  const double* cbuf; // element contributions in row-incidence order
  const double* lam;
  const int*    bc_mask;
  double*       jt;
};

// Phase 2 (one lane per DOF): sum the DOF's element contributions -- a
// contiguous run of cbuf, in ascending element order -- then pass lam through
// on constrained DOFs.  A block's 256 runs are one contiguous range, staged in
// LDS with coalesced non-temporal loads when it fits (a segmented sum over a
// stream).
__device__ __forceinline__ void vjp_gather_block(int blk, const GatherArgs& a, double* sval)
{
  const int  c0 = blk * 256, c1 = min(c0 + 256, a.ndof);
  const int  i0 = a.inc_off[c0], ni = a.inc_off[c1] - i0;
  const bool staged = ni <= kGatherStage;
  if (staged)
  {
    stage_nt<256>(sval, a.cbuf + i0, ni);
    __syncthreads();
  }
  const int c = c0 + threadIdx.x;
  if (c < a.ndof)
  {
    double    v  = 0.0;
    const int t1 = a.inc_off[c + 1];
    for (int t = a.inc_off[c]; t < t1; ++t) v = v + (staged ? sval[t - i0] : a.cbuf[t]);
    if (a.bc_mask && a.bc_mask[c]) v = v + a.lam[c];
    __builtin_nontemporal_store(v, a.jt + c);
  }
  __syncthreads();
}


} // namespace fxk
