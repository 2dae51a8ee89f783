// P1 Poisson assembly and log-kappa VJP for gfx950: device code shared by
// the single-instance entry points (fx_p1.hip) and the batched ones
// (fx_batch.hip).
//
// Design: one lane per global row (node), gathering over the row's incident
// elements in ascending element order (fx_map_export_rowinc3).  Each lane
// recomputes the one element-matrix row it needs from the three vertex
// coordinates in registers (affine triangles: constant gradients), so no
// per-element quadrature data is streamed from HBM.  A workgroup owns 256
// consecutive rows: their incidence entries are one contiguous range, staged
// in LDS with coalesced loads, and their CSR values are one contiguous range,
// accumulated in LDS and written back with coalesced stores.  Streams read or
// written once (incidences, outputs) use non-temporal accesses so that the
// L2 keeps the node and element data that neighbouring rows share.  Every CSR value and residual
// entry is a sum over its elements in ascending element order with no
// atomics: results are deterministic and equal, operation for operation, to
// femx's host assemble loop (femx/assembly/Assembly.hpp:184-225) evaluating
// PoissonComponents::evalRow (examples/poisson/PoissonComponents.hpp:35-64)
// on ElementValues' physical gradients (femx/fem/ElementValues.cpp:138-181).
// The file is compiled without floating-point contraction so that each
// product is rounded as in femx's host code.
#pragma once
#include "fx_common.hpp"

namespace fxk
{


constexpr int kP1Block = 256;
constexpr int kP1MaxRowLen = 32; // LDS bound: 256 rows x 32 x 8 B = 64 KiB (+16 KiB staging)

// Physical P1 gradients and |det J|*wt of an affine triangle, evaluated as
// ElementValues::calcPhysicalValues/invJacobian do with the P1 reference
// gradients (LagrangeTriangleP1::calcdNdr) and weight 1/6 of the 3-point rule
// (GaussQuadrature::triangle(2)).
struct P1Geom
{
  double gx[3], gy[3];
  double jxw;
};

__device__ __forceinline__ P1Geom p1_geom(double x0, double y0, double x1, double y1, double x2,
                                          double y2)
{
  // J = sum_in x(in) dNdr(in,:) with dNdr = {(-1,-1),(1,0),(0,1)}.
  const double j00 = x1 - x0, j01 = x2 - x0, j10 = y1 - y0, j11 = y2 - y0;
  const double det     = j00 * j11 - j01 * j10;
  const double inv_det = 1.0 / det;
  const double i00 = j11 * inv_det, i01 = -j01 * inv_det, i10 = -j10 * inv_det,
               i11 = j00 * inv_det;
  P1Geom g;
  g.gx[0] = -i00 - i10;
  g.gy[0] = -i01 - i11;
  g.gx[1] = i00;
  g.gy[1] = i01;
  g.gx[2] = i10;
  g.gy[2] = i11;
  g.jxw   = fabs(det) * (1.0 / 6.0);
  return g;
}

// Stiffness entry K_rc = kappa * sum_q (grad phi_r . grad phi_c) JxW_q over
// the three (identical) quadrature points, in PoissonComponents' order.
// Select v[i] without dynamic register-array indexing (which would spill).
__device__ __forceinline__ double sel3(const double (&v)[3], int i)
{
  return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

__device__ __forceinline__ double p1_k(const P1Geom& g, int r, int c, double kappa)
{
  const double gd = sel3(g.gx, r) * sel3(g.gx, c) + sel3(g.gy, r) * sel3(g.gy, c);
  double       s  = gd * g.jxw;
  s               = s + gd * g.jxw;
  s               = s + gd * g.jxw;
  return kappa * s;
}

// Consistent load F_r = sum_q N_r(q) f_h(q) JxW_q on the 3-point rule.
__device__ __forceinline__ double p1_load(const P1Geom& g, int r, double f0, double f1,
                                          double f2)
{
  const double a = 1.0 / 6.0, b = 2.0 / 3.0;
  const double qr[3] = {a, b, a}, qs[3] = {a, a, b};
  double       F     = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q)
  {
    const double n0 = 1.0 - qr[q] - qs[q], n1 = qr[q], n2 = qs[q];
    const double fq = n0 * f0 + n1 * f1 + n2 * f2;
    const double nr = r == 0 ? n0 : (r == 1 ? n1 : n2);
    F               = F + nr * fq * g.jxw;
  }
  return F;
}

// Incidences of one block staged in LDS when they fit (coalesced loads of a
// contiguous range); longer blocks read them straight from global memory.
constexpr int kIncStage = 2048;

__device__ __forceinline__ int2 inc_at(const long long* sinc, const int2* inc, bool staged, int t,
                                       int ib0)
{
  if (staged)
  {
    const long long v = sinc[t - ib0];
    return make_int2(static_cast<int>(v & 0xffffffffll), static_cast<int>(v >> 32));
  }
  return inc[t];
}

struct P1AsmArgs
{
  int           num_rows;
  const double* x;
  const double* y;
  const int*    conn;
  const double* kappa;
  int           kappa_is_log;
  const double* f;
  const double* u;
  const int*    inc_off;
  const int2*   inc;
  const int*    row_ptr;
  double*       vals;
  double*       res;
  int           acc_cap;
};

// One row block of the P1 assembly (rows [blk*256, blk*256+256)); smem holds
// acc_cap doubles of row tile followed by the incidence stage.
__device__ __forceinline__ void p1_assemble_block(int blk, const P1AsmArgs& a, double* smem)
{
  double*     acc  = smem;
  long long*  sinc = reinterpret_cast<long long*>(smem + a.acc_cap);
  const int   r0   = blk * kP1Block;
  const int   r1   = min(r0 + kP1Block, a.num_rows);
  const int   base = a.row_ptr[r0];
  const int   bn   = a.vals ? a.row_ptr[r1] - base : 0;
  const int   ib0 = a.inc_off[r0], ib1 = a.inc_off[r1];
  const bool  staged = ib1 - ib0 <= kIncStage;
  for (int i = threadIdx.x; i < bn; i += kP1Block) acc[i] = 0.0;
  if (staged)
    for (int i = threadIdx.x; i < ib1 - ib0; i += kP1Block)
      sinc[i] = __builtin_nontemporal_load(reinterpret_cast<const long long*>(a.inc) + ib0 + i);
  __syncthreads();

  const int r = r0 + threadIdx.x;
  if (r < a.num_rows)
  {
    double*   racc = acc + (a.row_ptr[r] - base);
    double    rsum = 0.0;
    const int t1   = a.inc_off[r + 1];
    for (int t = a.inc_off[r]; t < t1; ++t)
    {
      const int2     it   = inc_at(sinc, a.inc, staged, t, ib0);
      const int      e    = it.x;
      const unsigned meta = static_cast<unsigned>(it.y);
      const int      li   = meta & 0xff;
      const int      n0 = a.conn[3 * e], n1 = a.conn[3 * e + 1], n2 = a.conn[3 * e + 2];
      const P1Geom   g  = p1_geom(a.x[n0], a.y[n0], a.x[n1], a.y[n1], a.x[n2], a.y[n2]);
      const double   ke = a.kappa ? (a.kappa_is_log ? exp(a.kappa[e]) : a.kappa[e]) : 1.0;
      const double   k0 = p1_k(g, li, 0, ke), k1 = p1_k(g, li, 1, ke), k2 = p1_k(g, li, 2, ke);
      double         re = 0.0;
      if (a.u) re = k0 * a.u[n0] + k1 * a.u[n1] + k2 * a.u[n2];
      if (a.f) re = re - p1_load(g, li, a.f[n0], a.f[n1], a.f[n2]);
      else re = re - 0.0;
      rsum = rsum + re;
      if (a.vals)
      {
        // distinct columns: read all three, then write (no dependent round trips)
        const int    p0 = (meta >> 8) & 0xff, p1 = (meta >> 16) & 0xff, p2 = meta >> 24;
        const double o0 = racc[p0], o1 = racc[p1], o2 = racc[p2];
        racc[p0]        = o0 + k0;
        racc[p1]        = o1 + k1;
        racc[p2]        = o2 + k2;
      }
    }
    __builtin_nontemporal_store(rsum, a.res + r);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < bn; i += kP1Block)
    __builtin_nontemporal_store(acc[i], a.vals + base + i);
  __syncthreads();
}

struct P1VjpArgs
{
  int           num_rows;
  const double* x;
  const double* y;
  const int*    conn;
  const double* m;
  const double* u;
  const double* lam;
  const int*    bc_mask;
  const int*    inc_off;
  const int2*   inc;
  double*       jt;
  double*       gout;
};

// One pass: each lane (node j) sums column j of its elements' matrices against
// the boundary-masked adjoint; the lane holding an element's local row 0
// also writes that element's control sensitivity g_e = lam~_e^T K^e u_e.
__device__ __forceinline__ void p1_vjp_block(int blk, const P1VjpArgs& a, long long* sinc)
{
  const int  r0 = blk * kP1Block, r1 = min(r0 + kP1Block, a.num_rows);
  const int  ib0 = a.inc_off[r0], ib1 = a.inc_off[r1];
  const bool staged = ib1 - ib0 <= kIncStage;
  if (staged)
    for (int i = threadIdx.x; i < ib1 - ib0; i += kP1Block)
      sinc[i] = __builtin_nontemporal_load(reinterpret_cast<const long long*>(a.inc) + ib0 + i);
  __syncthreads();
  const int j = r0 + threadIdx.x;
  if (j < a.num_rows)
  {
    double    acc = 0.0;
    const int t1  = a.inc_off[j + 1];
    for (int t = a.inc_off[j]; t < t1; ++t)
    {
      const int2   it = inc_at(sinc, a.inc, staged, t, ib0);
      const int    e  = it.x;
      const int    li = static_cast<unsigned>(it.y) & 0xff;
      const int    n[3] = {a.conn[3 * e], a.conn[3 * e + 1], a.conn[3 * e + 2]};
      const P1Geom g    = p1_geom(a.x[n[0]], a.y[n[0]], a.x[n[1]], a.y[n[1]], a.x[n[2]], a.y[n[2]]);
      const double ke   = exp(a.m[e]);
      double       le[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) le[r] = (a.bc_mask && a.bc_mask[n[r]]) ? 0.0 : a.lam[n[r]];
      const double c0 = p1_k(g, 0, li, ke), c1 = p1_k(g, 1, li, ke), c2 = p1_k(g, 2, li, ke);
      acc = acc + (c0 * le[0] + c1 * le[1] + c2 * le[2]);
      if (li == 0)
      {
        const double u0 = a.u[n[0]], u1 = a.u[n[1]], u2 = a.u[n[2]];
        double       ge = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
          const double re = p1_k(g, r, 0, ke) * u0 + p1_k(g, r, 1, ke) * u1 + p1_k(g, r, 2, ke) * u2;
          ge              = ge + le[r] * (re - 0.0);
        }
        a.gout[e] = ge;
      }
    }
    if (a.bc_mask && a.bc_mask[j]) acc = acc + a.lam[j];
    __builtin_nontemporal_store(acc, a.jt + j);
  }
  __syncthreads();
}

} // namespace fxk
