// GLS-stabilised P1/P1 Navier-Stokes time-step assembly for gfx950: device
// code of fx_gls_assemble.
//
// The operator is femx's NavierOperator (femx/model/ns/Components.hpp:
// 230-577) on P1 triangles with the order-2 rule (three points) that
// NavierStokesModel uses, assembled with respect to the next state as the
// host time assemble does (femx/assembly/Assembly.hpp:235-297): every row's
// residual is -F + K nxt and its Jacobian row is K, where K and F depend on
// the history states (lag 0 = current, lag 1 = previous) and the geometry.
//
// Design: one lane per mesh node.  A node owns three rows -- its two velocity
// components and its pressure -- and every element touching the node
// contributes to exactly those three rows through the node's local index in
// the element.  The lane walks the node's incident elements in ascending
// element order; for each it recomputes the element's three quadrature-point
// states (velocity, gradient, extrapolated advection velocity, GLS tau) once
// and derives the three element rows from them, so each element's point
// states are evaluated three times (once per vertex) instead of nine (once per
// row).  Geometry is recomputed from the vertex coordinates in registers: P1
// gradients are constant on an affine triangle, so no per-element quadrature
// data is read from HBM.
//
// A wave owns 64 consecutive nodes.  Their rows form at most kGlsRuns runs of
// consecutive rows (two for femx's field-major mixed layout: the velocity
// rows 2n, 2n+1 and the pressure rows 2N+n), whose CSR ranges are contiguous;
// the wave accumulates them in an LDS tile and writes each run back with
// coalesced stores.  Each CSR value and residual entry is a sum over its
// elements in ascending element order, with every element term computed in
// femx's operation order and without contraction: the results equal femx's
// host assembly bit for bit.
#pragma once
#include "fx_common.hpp"

namespace fxk
{

constexpr int kGlsBlock = FX_GLS_BLOCK_NODES; // nodes per wave
constexpr int kGlsRuns  = FX_GLS_MAX_RUNS;    // row runs per block

struct GlsAsmArgs
{
  int           num_nodes;
  int           num_states;
  const double* x;
  const double* y;
  const int*    conn;
  const int*    node_rows; // [3 N]: velocity x, velocity y, pressure row of each node
  int           step;
  int           num_hist;
  const double* hist; // lag-major, num_hist x num_states
  const double* nxt;
  double        rho, mu, dt;
  const int*    inc_off;   // [N+1]
  const int*    inc_elem;  // [3E], ascending element per node
  const uint4*  inc_meta;  // [2 per incidence]: 27 CSR offsets + local node
  const int4*   blk_runs;  // [kGlsRuns per block]: csr begin, length, tile offset
  const int*    node_tile; // [3 N]: tile offset of each node row
  double*       vals;      // null: residual only
  double*       res;
  int           tile_cap;  // doubles of LDS tile per wave
};

__device__ __forceinline__ double gsel3(double a, double b, double c, int i)
{
  return i == 0 ? a : (i == 1 ? b : c);
}

__device__ __forceinline__ unsigned meta_byte(const uint4& m0, const uint4& m1, int k)
{
  // k is a compile-time constant after unrolling
  const unsigned w = k < 16 ? (k < 4 ? m0.x : k < 8 ? m0.y : k < 12 ? m0.z : m0.w)
                            : (k < 20 ? m1.x : k < 24 ? m1.y : k < 28 ? m1.z : m1.w);
  return (w >> (8 * (k & 3))) & 0xffu;
}

// One quadrature-point state of evalQp (Components.hpp:329-396) for dim 2.
struct GlsQp
{
  double u0, u1, a0, a1, ag0, ag1, g00, g01, g10, g11, tau;
};

__device__ __forceinline__ GlsQp gls_qp(const double (&N)[3], const double (&gx)[3],
                                        const double (&gy)[3], const double (&cur)[6],
                                        const double (&adv)[6], double nu, double time)
{
  GlsQp q;
  double u0 = 0.0, u1 = 0.0, g00 = 0.0, g01 = 0.0, g10 = 0.0, g11 = 0.0, a0 = 0.0, a1 = 0.0;
#pragma unroll
  for (int in = 0; in < 3; ++in)
  {
    const double vx = cur[2 * in], vy = cur[2 * in + 1];
    u0  = u0 + N[in] * vx;
    g00 = g00 + gx[in] * vx;
    g01 = g01 + gy[in] * vx;
    a0  = a0 + N[in] * adv[2 * in];
    u1  = u1 + N[in] * vy;
    g10 = g10 + gx[in] * vy;
    g11 = g11 + gy[in] * vy;
    a1  = a1 + N[in] * adv[2 * in + 1];
  }
  double ag0 = 0.0, ag1 = 0.0;
  ag0                 = ag0 + g00 * a0;
  ag0                 = ag0 + g01 * a1;
  ag1                 = ag1 + g10 * a0;
  ag1                 = ag1 + g11 * a1;
  const double speed2 = (0.0 + u0 * u0) + u1 * u1;
  const double speed  = sqrt(speed2);
  // elemLength (Components.hpp:286-327)
  double dx, dy;
  if (speed > 1.0e-10)
  {
    dx = u0 / speed;
    dy = u1 / speed;
  }
  else
  {
    dx = dy = 1.0 / sqrt(2.0);
  }
  double sum = 0.0;
#pragma unroll
  for (int in = 0; in < 3; ++in)
  {
    const double g = (0.0 + dx * gx[in]) + dy * gy[in];
    sum            = sum + (g >= 0.0 ? g : -g);
  }
  const double h    = sum > 1.0e-14 ? 2.0 / sum : 0.0;
  double       flow = 0.0, diff = 0.0;
  if (h > 0.0)
  {
    const double f = 2.0 * speed / h;
    const double v = 4.0 * nu / (h * h);
    flow           = f * f;
    diff           = v * v;
  }
  q.tau = 1.0 / sqrt(time + flow + diff);
  q.u0  = u0;
  q.u1  = u1;
  q.a0  = a0;
  q.a1  = a1;
  q.ag0 = ag0;
  q.ag1 = ag1;
  q.g00 = g00;
  q.g01 = g01;
  q.g10 = g10;
  q.g11 = g11;
  return q;
}

// Quadrature points of the order-2 triangle rule (femx/fem/GaussQuadrature.hpp:
// 175-204) and LagrangeTriangleP1::calcN there.
__device__ __forceinline__ void gls_shape(int q, double (&N)[3])
{
  const double a = 1.0 / 6.0, b = 2.0 / 3.0;
  const double r = q == 1 ? b : a, s = q == 2 ? b : a;
  N[0] = 1.0 - r - s;
  N[1] = r;
  N[2] = s;
}

// The three rows of local node li of one element: velocity x (Kx: columns
// u0 u1 u2 then p0 p1 p2), velocity y (Ky, same), pressure (Kp: columns
// u0 v0 u1 v1 u2 v2 p0 p1 p2), and their residuals -F + K nxt.
struct GlsRows
{
  double Kx[6], Ky[6], Kp[9];
  double rx, ry, rp;
};

__device__ __forceinline__ GlsRows gls_rows(int li, const double (&gx)[3], const double (&gy)[3],
                                            double jw, const double (&cur)[6],
                                            const double (&adv)[6], const double (&nx)[9],
                                            double rho, double mu, double dt)
{
  GlsRows o;
#pragma unroll
  for (int k = 0; k < 6; ++k) o.Kx[k] = o.Ky[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) o.Kp[k] = 0.0;
  double       Fx = 0.0, Fy = 0.0, Fp = 0.0;
  const double nu   = mu / rho;
  const double time = (2.0 / dt) * (2.0 / dt);
  const double rdt  = rho / dt;
  const double h5r  = 0.5 * rho;
  const double h5m  = 0.5 * mu;
  const double gix  = gsel3(gx[0], gx[1], gx[2], li);
  const double giy  = gsel3(gy[0], gy[1], gy[2], li);
#pragma unroll
  for (int iq = 0; iq < 3; ++iq)
  {
    double N[3];
    gls_shape(iq, N);
    const GlsQp  q    = gls_qp(N, gx, gy, cur, adv, nu, time);
    const double Ni   = gsel3(N[0], N[1], N[2], li);
    const double dvi  = (0.0 + gix * q.a0) + giy * q.a1;
    const double trd  = q.tau * rho / dt;
    const double h5tr = 0.5 * q.tau * rho;
    const double tdt  = q.tau / dt;
    const double h5t  = 0.5 * q.tau;
    const double tr   = q.tau / rho;
    const double tdv  = q.tau * dvi;
    // velocity rows (Components.hpp:507-536)
    const double gux = (0.0 + gix * q.g00) + giy * q.g01;
    const double guy = (0.0 + gix * q.g10) + giy * q.g11;
    Fx = Fx + ((((rdt * Ni * q.u0 - h5r * Ni * q.ag0) - h5m * gux) + trd * dvi * q.u0)
               - h5tr * dvi * q.ag0)
                  * jw;
    Fy = Fy + ((((rdt * Ni * q.u1 - h5r * Ni * q.ag1) - h5m * guy) + trd * dvi * q.u1)
               - h5tr * dvi * q.ag1)
                  * jw;
    // pressure row (Components.hpp:537-570)
    const double div_u   = (0.0 + gix * q.u0) + giy * q.u1;
    const double div_adv = (0.0 + gix * q.ag0) + giy * q.ag1;
    Fp                   = Fp + (tdt * div_u - h5t * div_adv) * jw;
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
      const double Nj  = N[j];
      const double dvj = (0.0 + gx[j] * q.a0) + gy[j] * q.a1;
      const double gd  = (0.0 + gix * gx[j]) + giy * gy[j];
      const double kuu =
          (((rdt * Ni * Nj + h5r * Ni * dvj) + h5m * gd) + trd * dvi * Nj) + h5tr * dvi * dvj;
      o.Kx[j]     = o.Kx[j] + kuu * jw;
      o.Ky[j]     = o.Ky[j] + kuu * jw;
      o.Kx[3 + j] = o.Kx[3 + j] + (-gix * Nj + tdv * gx[j]) * jw;
      o.Ky[3 + j] = o.Ky[3 + j] + (-giy * Nj + tdv * gy[j]) * jw;
      o.Kp[2 * j] = o.Kp[2 * j] + ((Ni * gx[j] + tdt * gix * Nj) + h5t * gix * dvj) * jw;
      o.Kp[2 * j + 1] =
          o.Kp[2 * j + 1] + ((Ni * gy[j] + tdt * giy * Nj) + h5t * giy * dvj) * jw;
      o.Kp[6 + j] = o.Kp[6 + j] + tr * gd * jw;
    }
  }
  // out = -F + sum_col K[col] nxt[col] over the row's columns in local order;
  // the structurally zero columns (the other velocity component) add nothing.
  double rx = -Fx, ry = -Fy, rp = -Fp;
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    rx = rx + o.Kx[j] * nx[2 * j];
    ry = ry + o.Ky[j] * nx[2 * j + 1];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    rx = rx + o.Kx[3 + j] * nx[6 + j];
    ry = ry + o.Ky[3 + j] * nx[6 + j];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) rp = rp + o.Kp[k] * nx[k];
  o.rx = rx;
  o.ry = ry;
  o.rp = rp;
  return o;
}

// One block of 64 nodes, run by one wave; tile holds tile_cap doubles.
__device__ __forceinline__ void gls_assemble_block(int blk, const GlsAsmArgs& a, double* tile)
{
  const int lane = threadIdx.x & 63;
  int4      runs[kGlsRuns];
  int       tile_len = 0;
#pragma unroll
  for (int k = 0; k < kGlsRuns; ++k)
  {
    runs[k]  = a.blk_runs[blk * kGlsRuns + k];
    tile_len = max(tile_len, runs[k].z + runs[k].y);
  }
  if (a.vals)
    for (int i = lane; i < tile_len; i += 64) tile[i] = 0.0;
  __syncthreads();

  // a node of no element owns no rows (node_rows -1) and is skipped
  const int n = blk * kGlsBlock + lane;
  if (n < a.num_nodes && a.node_rows[3 * n] >= 0)
  {
    const int tx = a.node_tile[3 * n], ty = a.node_tile[3 * n + 1], tp = a.node_tile[3 * n + 2];
    double    sx = 0.0, sy = 0.0, sp = 0.0;
    const bool extrap = a.step > 0 && a.num_hist > 1;
    const int  t1     = a.inc_off[n + 1];
    for (int t = a.inc_off[n]; t < t1; ++t)
    {
      const int   e  = a.inc_elem[t];
      const uint4 m0 = a.inc_meta[2 * t], m1 = a.inc_meta[2 * t + 1];
      const int   li = static_cast<int>(meta_byte(m0, m1, 27));
      int         v[3], rows[9];
      double      xe[3], ye[3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
      {
        v[k]           = a.conn[3 * e + k];
        xe[k]          = a.x[v[k]];
        ye[k]          = a.y[v[k]];
        rows[2 * k]     = a.node_rows[3 * v[k]];
        rows[2 * k + 1] = a.node_rows[3 * v[k] + 1];
        rows[6 + k]     = a.node_rows[3 * v[k] + 2];
      }
      double cur[6], adv[6], nx[9];
#pragma unroll
      for (int k = 0; k < 6; ++k) cur[k] = a.hist[rows[k]];
      if (extrap)
      {
        const double* prev = a.hist + a.num_states;
#pragma unroll
        for (int k = 0; k < 6; ++k) adv[k] = 1.5 * cur[k] - 0.5 * prev[rows[k]];
      }
      else
      {
#pragma unroll
        for (int k = 0; k < 6; ++k) adv[k] = cur[k];
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) nx[k] = a.nxt[rows[k]];

      // ElementValues on the affine triangle (femx/fem/ElementValues.cpp:138-181)
      const double j00 = xe[1] - xe[0], j01 = xe[2] - xe[0], j10 = ye[1] - ye[0],
                   j11 = ye[2] - ye[0];
      const double det = j00 * j11 - j01 * j10, inv_det = 1.0 / det;
      const double i00 = j11 * inv_det, i01 = -j01 * inv_det, i10 = -j10 * inv_det,
                   i11 = j00 * inv_det;
      const double gx[3] = {-i00 - i10, i00, i10};
      const double gy[3] = {-i01 - i11, i01, i11};
      const double jw    = fabs(det) * (1.0 / 6.0);

      const GlsRows o = gls_rows(li, gx, gy, jw, cur, adv, nx, a.rho, a.mu, a.dt);
      sx              = sx + o.rx;
      sy              = sy + o.ry;
      sp              = sp + o.rp;
      if (a.vals)
      {
        // meta byte 9k + j: position of local column j in the node's row k
        double* px = tile + tx;
        double* py = tile + ty;
        double* pp = tile + tp;
        // read all, then write: the positions within one row are distinct
        double ox[6], oy[6], op[9];
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
          ox[j]     = px[meta_byte(m0, m1, 2 * j)];
          ox[3 + j] = px[meta_byte(m0, m1, 6 + j)];
          oy[j]     = py[meta_byte(m0, m1, 9 + 2 * j + 1)];
          oy[3 + j] = py[meta_byte(m0, m1, 9 + 6 + j)];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) op[k] = pp[meta_byte(m0, m1, 18 + k)];
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
          px[meta_byte(m0, m1, 2 * j)]          = ox[j] + o.Kx[j];
          px[meta_byte(m0, m1, 6 + j)]          = ox[3 + j] + o.Kx[3 + j];
          py[meta_byte(m0, m1, 9 + 2 * j + 1)]  = oy[j] + o.Ky[j];
          py[meta_byte(m0, m1, 9 + 6 + j)]      = oy[3 + j] + o.Ky[3 + j];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) pp[meta_byte(m0, m1, 18 + k)] = op[k] + o.Kp[k];
      }
    }
    __builtin_nontemporal_store(sx, a.res + a.node_rows[3 * n]);
    __builtin_nontemporal_store(sy, a.res + a.node_rows[3 * n + 1]);
    __builtin_nontemporal_store(sp, a.res + a.node_rows[3 * n + 2]);
  }
  __syncthreads();
  if (a.vals)
  {
#pragma unroll
    for (int k = 0; k < kGlsRuns; ++k)
      for (int i = lane; i < runs[k].y; i += 64)
        __builtin_nontemporal_store(tile[runs[k].z + i], a.vals + runs[k].x + i);
  }
}

} // namespace fxk
