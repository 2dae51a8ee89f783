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
// A wave owns 32 consecutive nodes, two lanes per node.  Their rows form at most kGlsRuns runs of
// consecutive rows (two for femx's field-major mixed layout: the velocity
// rows 2n, 2n+1 and the pressure rows 2N+n), whose CSR ranges are contiguous;
// the wave accumulates them in an LDS tile and writes each run back with
// coalesced stores.  Each CSR value and residual entry is a sum over its
// elements in ascending element order, with every element term computed in
// femx's operation order and without contraction: the results equal femx's
// host assembly bit for bit.  femx starts each inner sum at 0.0; here a sum
// starts at its first term.  That changes at most the sign of an exact-zero
// partial sum, and nothing observes it: no such sum is divided by or square-
// rooted (the divisions and roots act on squares, guarded lengths and the
// positive tau argument), the comparisons are sign-blind at zero, and the
// outputs are accumulated onto +0.0 as in femx.
#pragma once
#include "fx_dev.hpp"

namespace fxk
{

constexpr int kGlsBlock = FX_GLS_BLOCK_NODES; // nodes per wave
constexpr int kGlsRuns  = FX_GLS_MAX_RUNS;    // row runs per block
constexpr int kGlsMaxTile = 8192;             // doubles of LDS row tile per wave (64 KiB)

// Loop-invariant factors of NavierOperator's expressions, formed once (on the
// host, passed as kernel arguments) with the same operations femx applies at
// every use, so the values are equal.
struct GlsConst
{
  double rho, mu, dt, nu, time, rdt, h5r, h5m;
  double idt, irho; // 1/dt and 1/rho: the VJP multiplies where the residual divides
};

__host__ __device__ inline GlsConst gls_const(double rho, double mu, double dt)
{
  GlsConst c;
  c.rho          = rho;
  c.mu           = mu;
  c.dt           = dt;
  c.nu           = mu / rho;
  const double t = 2.0 / dt;
  c.time         = t * t;
  c.rdt          = rho / dt;
  c.h5r          = 0.5 * rho;
  c.h5m          = 0.5 * mu;
  c.idt          = 1.0 / dt;
  c.irho         = 1.0 / rho;
  return c;
}

struct GlsAsmArgs
{
  int           num_nodes;
  int           num_states;
  const double* x;
  const double* y;
  const int*    conn;
  const int*    dofs;      // [9 E]: element DOF table (local order u0 v0 u1 v1 u2 v2 p0 p1 p2)
  const int*    node_rows; // [3 N]: velocity x, velocity y, pressure row of each node
  int           step;
  int           num_hist;
  const double* hist; // lag-major, num_hist x num_states
  const double* nxt;
  GlsConst      c;
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
  GlsQp  q;
  double u0 = N[0] * cur[0], g00 = gx[0] * cur[0], g01 = gy[0] * cur[0], a0 = N[0] * adv[0];
  double u1 = N[0] * cur[1], g10 = gx[0] * cur[1], g11 = gy[0] * cur[1], a1 = N[0] * adv[1];
#pragma unroll
  for (int in = 1; in < 3; ++in)
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
  const double ag0    = g00 * a0 + g01 * a1;
  const double ag1    = g10 * a0 + g11 * a1;
  const double speed2 = u0 * u0 + u1 * u1;
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
  double sum;
#pragma unroll
  for (int in = 0; in < 3; ++in)
  {
    const double g = dx * gx[in] + dy * gy[in];
    const double m = g >= 0.0 ? g : -g;
    sum            = in == 0 ? m : sum + m;
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

// The three rows of local node li of one element.  Velocity rows: the
// same-component velocity columns (Kuu, equal for both components) and the
// pressure columns (Kxp, Kyp); the other component's columns are zero.
// Pressure row: all nine columns, u0 v0 u1 v1 u2 v2 p0 p1 p2 (Kp).  r* are
// the row residuals -F + K nxt.
struct GlsRows
{
  double Kuu[3], Kxp[3], Kyp[3], Kp[9];
  double rx, ry, rp;
};

__device__ __forceinline__ void gls_rows(GlsRows& o, int li, const double (&gx)[3],
                                         const double (&gy)[3], double jw,
                                         const double (&cur)[6], const double (&adv)[6],
                                         const GlsConst& c)
{
  double       Fx, Fy, Fp;
  const double gix = gsel3(gx[0], gx[1], gx[2], li);
  const double giy = gsel3(gy[0], gy[1], gy[2], li);
#pragma unroll
  for (int iq = 0; iq < 3; ++iq)
  {
    double N[3];
    gls_shape(iq, N);
    const GlsQp  q    = gls_qp(N, gx, gy, cur, adv, c.nu, c.time);
    const double Ni   = gsel3(N[0], N[1], N[2], li);
    const double dvi  = gix * q.a0 + giy * q.a1;
    const double trd  = q.tau * c.rho / c.dt;
    const double h5tr = 0.5 * q.tau * c.rho;
    const double tdt  = q.tau / c.dt;
    const double h5t  = 0.5 * q.tau;
    const double tr   = q.tau / c.rho;
    const double tdv  = q.tau * dvi;
    // velocity rows (Components.hpp:507-536)
    const double gux = gix * q.g00 + giy * q.g01;
    const double guy = gix * q.g10 + giy * q.g11;
    const double fx = ((((c.rdt * Ni * q.u0 - c.h5r * Ni * q.ag0) - c.h5m * gux) + trd * dvi * q.u0)
               - h5tr * dvi * q.ag0)
                  * jw;
    const double fy = ((((c.rdt * Ni * q.u1 - c.h5r * Ni * q.ag1) - c.h5m * guy) + trd * dvi * q.u1)
               - h5tr * dvi * q.ag1)
                  * jw;
    // pressure row (Components.hpp:537-570)
    const double div_u   = gix * q.u0 + giy * q.u1;
    const double div_adv = gix * q.ag0 + giy * q.ag1;
    const double fp      = (tdt * div_u - h5t * div_adv) * jw;
    Fx                   = iq == 0 ? fx : Fx + fx;
    Fy                   = iq == 0 ? fy : Fy + fy;
    Fp                   = iq == 0 ? fp : Fp + fp;
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
      const double Nj  = N[j];
      const double dvj = gx[j] * q.a0 + gy[j] * q.a1;
      const double gd  = gix * gx[j] + giy * gy[j];
      const double kuu = (((c.rdt * Ni * Nj + c.h5r * Ni * dvj) + c.h5m * gd) + trd * dvi * Nj)
                         + h5tr * dvi * dvj;
      const double t[6] = {kuu * jw,
                           (-gix * Nj + tdv * gx[j]) * jw,
                           (-giy * Nj + tdv * gy[j]) * jw,
                           ((Ni * gx[j] + tdt * gix * Nj) + h5t * gix * dvj) * jw,
                           ((Ni * gy[j] + tdt * giy * Nj) + h5t * giy * dvj) * jw,
                           tr * gd * jw};
      o.Kuu[j]        = iq == 0 ? t[0] : o.Kuu[j] + t[0];
      o.Kxp[j]        = iq == 0 ? t[1] : o.Kxp[j] + t[1];
      o.Kyp[j]        = iq == 0 ? t[2] : o.Kyp[j] + t[2];
      o.Kp[2 * j]     = iq == 0 ? t[3] : o.Kp[2 * j] + t[3];
      o.Kp[2 * j + 1] = iq == 0 ? t[4] : o.Kp[2 * j + 1] + t[4];
      o.Kp[6 + j]     = iq == 0 ? t[5] : o.Kp[6 + j] + t[5];
    }
  }
  o.rx = -Fx;
  o.ry = -Fy;
  o.rp = -Fp;
}

// out = -F + sum_col K[col] nxt[col] over the row's columns in local order;
// the structurally zero columns (the other velocity component) add nothing.
__device__ __forceinline__ void gls_residuals(GlsRows& o, const double (&nx)[9])
{
  double rx = o.rx, ry = o.ry, rp = o.rp;
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    rx = rx + o.Kuu[j] * nx[2 * j];
    ry = ry + o.Kuu[j] * nx[2 * j + 1];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    rx = rx + o.Kxp[j] * nx[6 + j];
    ry = ry + o.Kyp[j] * nx[6 + j];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) rp = rp + o.Kp[k] * nx[k];
  o.rx = rx;
  o.ry = ry;
  o.rp = rp;
}

// Adds one element's rows into the node's three tile rows; meta byte 9k+j is
// the position of local column j in row k.  Row by row: read the row's
// entries, then write them (positions within a row are distinct).
__device__ __forceinline__ void gls_add_rows(double* tile, int tx, int ty, int tp, const uint4& m0,
                                             const uint4& m1, const GlsRows& o)
{
  double* px = tile + tx;
  double* py = tile + ty;
  double* pp = tile + tp;
  double  v[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    v[j]     = px[meta_byte(m0, m1, 2 * j)];
    v[3 + j] = px[meta_byte(m0, m1, 6 + j)];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    px[meta_byte(m0, m1, 2 * j)] = v[j] + o.Kuu[j];
    px[meta_byte(m0, m1, 6 + j)] = v[3 + j] + o.Kxp[j];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    v[j]     = py[meta_byte(m0, m1, 9 + 2 * j + 1)];
    v[3 + j] = py[meta_byte(m0, m1, 9 + 6 + j)];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
  {
    py[meta_byte(m0, m1, 9 + 2 * j + 1)] = v[j] + o.Kuu[j];
    py[meta_byte(m0, m1, 9 + 6 + j)]     = v[3 + j] + o.Kyp[j];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) v[k] = pp[meta_byte(m0, m1, 18 + k)];
#pragma unroll
  for (int k = 0; k < 9; ++k) pp[meta_byte(m0, m1, 18 + k)] = v[k] + o.Kp[k];
}

// Element inputs of one incidence: geometry and the local history and next
// states gathered through the element DOF table.
struct GlsElem
{
  double gx[3], gy[3], jw;
  double cur[6], adv[6], nx[9];
};

__device__ __forceinline__ void gls_load_elem(GlsElem& E, const GlsAsmArgs& a, int e, bool extrap)
{
  int    v[3], d[9];
  double xe[3], ye[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = a.conn[3 * e + k];
#pragma unroll
  for (int k = 0; k < 9; ++k) d[k] = a.dofs[9 * e + k];
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    xe[k] = a.x[v[k]];
    ye[k] = a.y[v[k]];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) E.cur[k] = a.hist[d[k]];
  if (extrap)
  {
    const double* prev = a.hist + a.num_states;
#pragma unroll
    for (int k = 0; k < 6; ++k) E.adv[k] = 1.5 * E.cur[k] - 0.5 * prev[d[k]];
  }
  else
  {
#pragma unroll
    for (int k = 0; k < 6; ++k) E.adv[k] = E.cur[k];
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) E.nx[k] = a.nxt[d[k]];
  // ElementValues on the affine triangle (femx/fem/ElementValues.cpp:138-181)
  const double j00 = xe[1] - xe[0], j01 = xe[2] - xe[0], j10 = ye[1] - ye[0], j11 = ye[2] - ye[0];
  const double det = j00 * j11 - j01 * j10, inv_det = 1.0 / det;
  const double i00 = j11 * inv_det, i01 = -j01 * inv_det, i10 = -j10 * inv_det,
               i11 = j00 * inv_det;
  E.gx[0] = -i00 - i10;
  E.gx[1] = i00;
  E.gx[2] = i10;
  E.gy[0] = -i01 - i11;
  E.gy[1] = i01;
  E.gy[2] = i11;
  E.jw    = fabs(det) * (1.0 / 6.0);
}

// One block of kGlsBlock nodes, run by one wave with two lanes per node: role
// 0 takes the node's incidences t0, t0+2, ..., role 1 takes t0+1, t0+3, ...
// After each round role 0 adds its element rows into the tile, then role 1
// adds its own, and the residuals are summed in the same order (role 1's
// fetched by a cross-lane swap): every CSR value and residual entry still
// receives its element terms in ascending element order.
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
  wave_sync();

  const GlsConst& c     = a.c;
  const bool     extrap = a.step > 0 && a.num_hist > 1;
  const int      role   = lane & 1;
  // a node of no element owns no rows (node_rows -1) and is skipped
  const int  n    = blk * kGlsBlock + (lane >> 1);
  const bool live = n < a.num_nodes && a.node_rows[3 * n] >= 0;
  int        t0 = 0, t1 = 0, tx = 0, ty = 0, tp = 0;
  if (live)
  {
    t0 = a.inc_off[n];
    t1 = a.inc_off[n + 1];
    tx = a.node_tile[3 * n];
    ty = a.node_tile[3 * n + 1];
    tp = a.node_tile[3 * n + 2];
  }
  double sx = 0.0, sy = 0.0, sp = 0.0;
  // the wave runs until its longest node is done; lanes past their node's
  // incidences idle through the rounds
  int rounds = (t1 - t0 + 1) >> 1;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) rounds = max(rounds, __shfl_xor(rounds, off));
  for (int r = 0; r < rounds; ++r)
  {
    const int  t    = t0 + 2 * r + role;
    const bool have = t < t1;
    GlsRows    o;
    uint4      m0 = make_uint4(0, 0, 0, 0), m1 = m0;
    if (have)
    {
      const int e = a.inc_elem[t];
      m0          = a.inc_meta[2 * t];
      m1          = a.inc_meta[2 * t + 1];
      const int li = static_cast<int>(meta_byte(m0, m1, 27));
      GlsElem   E;
      gls_load_elem(E, a, e, extrap);
      gls_rows(o, li, E.gx, E.gy, E.jw, E.cur, E.adv, c);
      gls_residuals(o, E.nx);
    }
    else
    {
      o.rx = o.ry = o.rp = 0.0;
    }
    if (a.vals)
    {
      if (role == 0 && have) gls_add_rows(tile, tx, ty, tp, m0, m1, o);
      wave_sync();
      if (role == 1 && have) gls_add_rows(tile, tx, ty, tp, m0, m1, o);
      wave_sync();
    }
    const double px = __shfl_xor(o.rx, 1), py = __shfl_xor(o.ry, 1), pz = __shfl_xor(o.rp, 1);
    if (role == 0 && have)
    {
      sx = sx + o.rx;
      sy = sy + o.ry;
      sp = sp + o.rp;
      if (t + 1 < t1)
      {
        sx = sx + px;
        sy = sy + py;
        sp = sp + pz;
      }
    }
  }
  if (live && role == 0)
  {
    __builtin_nontemporal_store(sx, a.res + a.node_rows[3 * n]);
    __builtin_nontemporal_store(sy, a.res + a.node_rows[3 * n + 1]);
    __builtin_nontemporal_store(sp, a.res + a.node_rows[3 * n + 2]);
  }
  wave_sync();
  if (a.vals)
  {
#pragma unroll
    for (int k = 0; k < kGlsRuns; ++k)
      for (int i = lane; i < runs[k].y; i += 64)
        __builtin_nontemporal_store(tile[runs[k].z + i], a.vals + runs[k].x + i);
  }
}


// ---------------------------------------------------------------------------
// History VJP: out = (dR/dhist_lag)^T adj, the derivative femx takes with
// Enzyme (applyNavierHistJacT, femx/model/ns/NavierStokesModel.cpp:321-363;
// device femx/model/ns/Components.cu:175-238).  Written out by hand: with the
// adjoint and next-state interpolants at a quadrature point,
//   adj^T R_e = sum_q Jw l(u, G, adv, tau(u)),
// and the column of node k is N_k dl/du + grad phi_k . dl/dG + alpha N_k dl/dadv
// for the current state (lag 0) and beta N_k dl/dadv for the previous one,
// with (alpha, beta) = (1.5, -0.5) when the advection velocity is
// extrapolated and (1, 0) otherwise; pressure columns are zero.  tau depends on
// u through |u| and elemLength's streamline direction.  One lane per element
// forms its three nodes' velocity columns; one lane per node then sums them.

struct GlsVjpArgs
{
  int           num_nodes;
  int           num_states;
  const double* x;
  const double* y;
  const int*    conn;
  const int*    dofs;
  const int*    node_rows;
  int           step;
  int           num_hist;
  int           lag;
  const double* hist;
  const double* nxt;
  const double* adj;
  GlsConst      c;
  const int*    inc_off;
  const int*    inc_slot; // [3E]: slot of element e in its node k's list
  double*       out;
  double*       work; // 2 per incidence slot t: work[2 t + c]
  int           num_elems;
};

// Columns (k, 0..1) of node k for lag 0 (o0, o1) and lag 1 (p0, p1).  The
// partials at each quadrature point do not depend on the column's node:
// gls_hvjp_nodes forms them once and contracts them with each of the
// element's NK nodes ks[0..NK) (o: the current lag's columns, p: the
// previous lag's), each node's sum running over the quadrature points in
// order.
template <int NK>
__device__ __forceinline__ void gls_hvjp_nodes(const int (&ks)[NK], const double (&gx)[3],
                                               const double (&gy)[3], double jw,
                                               const double (&cur)[6], const double (&prev)[6],
                                               bool extrap, const double (&nx)[9],
                                               const double (&aj)[9], const GlsConst& C,
                                               double (&o0)[NK], double (&o1)[NK],
                                               double (&p0)[NK], double (&p1)[NK])
{
  const double al = extrap ? 1.5 : 1.0, be = extrap ? -0.5 : 0.0;
  const double g[6] = {gx[0], gy[0], gx[1], gy[1], gx[2], gy[2]};
  double       Dl[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, Dv[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  double       G[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, Dpi[2] = {0.0, 0.0}, Dq[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int d = 0; d < 2; ++d)
    {
#pragma unroll
      for (int c = 0; c < 2; ++c)
      {
        Dl[c][d] += g[2 * k + d] * aj[2 * k + c];
        Dv[c][d] += g[2 * k + d] * nx[2 * k + c];
        G[c][d] += g[2 * k + d] * cur[2 * k + c];
      }
      Dpi[d] += g[2 * k + d] * aj[6 + k];
      Dq[d] += g[2 * k + d] * nx[6 + k];
    }
  double gkx[NK], gky[NK];
#pragma unroll
  for (int m = 0; m < NK; ++m)
  {
    gkx[m] = gsel3(gx[0], gx[1], gx[2], ks[m]);
    gky[m] = gsel3(gy[0], gy[1], gy[2], ks[m]);
    o0[m] = o1[m] = p0[m] = p1[m] = 0.0;
  }
#pragma unroll
  for (int iq = 0; iq < 3; ++iq)
  {
    double Nq[3];
    gls_shape(iq, Nq);
    double u[2] = {0.0, 0.0}, av[2] = {0.0, 0.0}, lam[2] = {0.0, 0.0}, v[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int c = 0; c < 2; ++c)
      {
        u[c] += Nq[k] * cur[2 * k + c];
        const double a = extrap ? 1.5 * cur[2 * k + c] - 0.5 * prev[2 * k + c] : cur[2 * k + c];
        av[c] += Nq[k] * a;
        lam[c] += Nq[k] * aj[2 * k + c];
        v[c] += Nq[k] * nx[2 * k + c];
      }
    double ag[2], wv[2], z[2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
    {
      ag[c] = G[c][0] * av[0] + G[c][1] * av[1];
      wv[c] = av[0] * Dl[c][0] + av[1] * Dl[c][1];
      z[c]  = av[0] * Dv[c][0] + av[1] * Dv[c][1];
    }
    const double speed   = sqrt(u[0] * u[0] + u[1] * u[1]);
    double       dtau[2] = {0.0, 0.0}, tau;
    {
      double     dir[2], sgn[3], sum = 0.0;
      const bool moving = speed > 1.0e-10;
      if (moving)
      {
        const double isp = 1.0 / speed;
        dir[0]           = u[0] * isp;
        dir[1]           = u[1] * isp;
      }
      else
        dir[0] = dir[1] = 1.0 / sqrt(2.0);
#pragma unroll
      for (int k = 0; k < 3; ++k)
      {
        const double gk = dir[0] * g[2 * k] + dir[1] * g[2 * k + 1];
        sgn[k]          = gk >= 0.0 ? 1.0 : -1.0;
        sum += sgn[k] * gk;
      }
      double flow = 0.0, diff = 0.0, dflow[2] = {0.0, 0.0}, ddiff[2] = {0.0, 0.0};
      if (sum > 1.0e-14)
      {
        // with h = 2 / sum: 2 speed / h = speed sum and 4 nu / h^2 = nu sum^2
        const double fs = speed * sum, dv = C.nu * sum * sum;
        flow            = fs * fs;
        diff            = dv * dv;
        if (moving)
        {
          const double S   = speed * sum;
          const double s4  = speed * speed * speed * speed;
          const double is4 = 1.0 / s4, is6 = 1.0 / (s4 * speed * speed);
#pragma unroll
          for (int c = 0; c < 2; ++c)
          {
            double dS = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) dS += sgn[k] * g[2 * k + c];
            dflow[c] = 2.0 * S * dS;
            ddiff[c] = C.nu * C.nu * (4.0 * S * S * S * dS * is4 - 4.0 * S * S * S * S * u[c] * is6);
          }
        }
        else
        {
#pragma unroll
          for (int c = 0; c < 2; ++c) dflow[c] = 2.0 * sum * sum * u[c]; // 8 / h^2
        }
      }
      tau = 1.0 / sqrt(C.time + flow + diff);
#pragma unroll
      for (int c = 0; c < 2; ++c) dtau[c] = -0.5 * tau * tau * tau * (dflow[c] + ddiff[c]);
    }
    double B[2], l1 = 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c)
    {
      B[c] = -C.rdt * u[c] + 0.5 * C.rho * ag[c] + C.rdt * v[c] + 0.5 * C.rho * z[c] + Dq[c];
      l1 += wv[c] * B[c];
    }
#pragma unroll
    for (int d = 0; d < 2; ++d)
      l1 += Dpi[d] * (-u[d] * C.idt + 0.5 * ag[d] + v[d] * C.idt + 0.5 * z[d] + Dq[d] * C.irho);
    double Pu[2], Pag[2], Padv[2], PG[2][2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
    {
      Pu[c]  = -C.rdt * lam[c] + tau * (-C.rdt * wv[c] - Dpi[c] * C.idt) + l1 * dtau[c];
      Pag[c] = 0.5 * C.rho * lam[c] + tau * (0.5 * C.rho * wv[c] + 0.5 * Dpi[c]);
    }
#pragma unroll
    for (int d = 0; d < 2; ++d)
    {
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < 2; ++c)
      {
        s += 0.5 * C.rho * lam[c] * Dv[c][d];
        s += tau * (Dl[c][d] * B[c] + 0.5 * C.rho * wv[c] * Dv[c][d] + 0.5 * Dpi[c] * Dv[c][d]);
        s += Pag[c] * G[c][d];
      }
      Padv[d] = s;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int d = 0; d < 2; ++d) PG[c][d] = 0.5 * C.mu * Dl[c][d] + Pag[c] * av[d];
#pragma unroll
    for (int m = 0; m < NK; ++m)
    {
      const double Nk = gsel3(Nq[0], Nq[1], Nq[2], ks[m]);
      const double d0 = Nk * Pu[0] + gkx[m] * PG[0][0] + gky[m] * PG[0][1] + al * Nk * Padv[0];
      const double d1 = Nk * Pu[1] + gkx[m] * PG[1][0] + gky[m] * PG[1][1] + al * Nk * Padv[1];
      o0[m] += jw * d0;
      o1[m] += jw * d1;
      p0[m] += jw * be * Nk * Padv[0];
      p1[m] += jw * be * Nk * Padv[1];
    }
  }
}

// Element inputs of the history VJP: geometry, the current, previous and
// next states and the adjoint, gathered through the element DOF table.
struct GlsVjpElem
{
  int    v[3];
  double gx[3], gy[3], jw, cur[6], prev[6], nx[9], aj[9];
};

__device__ __forceinline__ void gls_vjp_load(GlsVjpElem& E, const GlsVjpArgs& a, int e, bool extrap)
{
  int    d[9];
  double xe[3], ye[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) E.v[k] = a.conn[3 * e + k];
#pragma unroll
  for (int k = 0; k < 9; ++k) d[k] = a.dofs[9 * e + k];
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    xe[k] = a.x[E.v[k]];
    ye[k] = a.y[E.v[k]];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k)
  {
    E.cur[k]  = a.hist[d[k]];
    E.prev[k] = extrap ? a.hist[a.num_states + d[k]] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < 9; ++k)
  {
    E.nx[k] = a.nxt[d[k]];
    E.aj[k] = a.adj[d[k]];
  }
  const double j00 = xe[1] - xe[0], j01 = xe[2] - xe[0], j10 = ye[1] - ye[0], j11 = ye[2] - ye[0];
  const double det = j00 * j11 - j01 * j10, inv_det = 1.0 / det;
  const double i00 = j11 * inv_det, i01 = -j01 * inv_det, i10 = -j10 * inv_det, i11 = j00 * inv_det;
  E.gx[0] = -i00 - i10, E.gx[1] = i00, E.gx[2] = i10;
  E.gy[0] = -i01 - i11, E.gy[1] = i01, E.gy[2] = i11;
  E.jw    = fabs(det) * (1.0 / 6.0);
}

// Two-phase form, phase 1 (one lane per element): the partials once per
// quadrature point, contracted with each of the element's three nodes; node
// k's two columns for the requested lag go to the element's slot t =
// inc_slot[3 e + k] in that node's incidence list, work[2 t + c], so that each
// node's terms lie contiguous and in ascending element order.  (Evaluated
// per node incidence instead, the same partials would be formed three times
// per element.)
// LAG: the requested lag, 0 or 1, fixed at compile time so that the other
// lag's columns (and for lag 1 the state and gradient partials) are not
// formed; 2 stands for any older lag, whose columns are zero.
template <int LAG>
__device__ __forceinline__ void gls_hist_vjp_elem(int e, const GlsVjpArgs& a)
{
  if (e >= a.num_elems) return;
  const bool extrap = a.step > 0 && a.num_hist > 1;
  double     c[6]   = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int        slot[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) slot[k] = a.inc_slot[3 * e + k];
  if constexpr (LAG < 2)
  {
    GlsVjpElem E;
    gls_vjp_load(E, a, e, extrap);
    const int ks[3] = {0, 1, 2};
    double    o0[3], o1[3], p0[3], p1[3];
    gls_hvjp_nodes<3>(ks, E.gx, E.gy, E.jw, E.cur, E.prev, extrap, E.nx, E.aj, a.c, o0, o1, p0, p1);
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
      c[2 * k]     = LAG == 0 ? o0[k] : p0[k];
      c[2 * k + 1] = LAG == 0 ? o1[k] : p1[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    *reinterpret_cast<double2*>(a.work + 2 * static_cast<long>(slot[k])) =
        double2{c[2 * k], c[2 * k + 1]};
}

// Phase 2 (one lane per node): the node's two columns summed over its
// incidence slots -- its elements in ascending order -- from zero, term after
// term.  The pressure row's column is zero.  A block's 256 nodes own one
// contiguous slot range, staged in LDS with coalesced non-temporal loads when
// it fits (kGlsSumStage doubles, 8 slots per node on average).
constexpr int kGlsSumNodes = 256;
constexpr int kGlsSumStage = 4096;

__device__ __forceinline__ void gls_hist_vjp_sum_block(int blk, const GlsVjpArgs& a, double* sw)
{
  const int  n0 = blk * kGlsSumNodes, n1 = min(n0 + kGlsSumNodes, a.num_nodes);
  const int  i0 = a.inc_off[n0], ni = a.inc_off[n1] - i0;
  const bool staged = 2 * static_cast<long>(ni) <= kGlsSumStage;
  if (staged)
  {
    stage_nt<kGlsSumNodes>(sw, a.work + 2 * static_cast<long>(i0), 2 * ni);
    __syncthreads();
  }
  const int n = n0 + threadIdx.x;
  if (n >= a.num_nodes || a.node_rows[3 * n] < 0) return;
  double         s0 = 0.0, s1 = 0.0;
  const double2* w  = reinterpret_cast<const double2*>(a.work);
  const int      t1 = a.inc_off[n + 1];
  for (int t = a.inc_off[n]; t < t1; ++t)
  {
    if (staged)
    {
      s0 += sw[2 * (t - i0)];
      s1 += sw[2 * (t - i0) + 1];
    }
    else
    {
      const double2 wt = w[t];
      s0 += wt.x;
      s1 += wt.y;
    }
  }
  __builtin_nontemporal_store(s0, a.out + a.node_rows[3 * n]);
  __builtin_nontemporal_store(s1, a.out + a.node_rows[3 * n + 1]);
  __builtin_nontemporal_store(0.0, a.out + a.node_rows[3 * n + 2]);
}


// ---------------------------------------------------------------------------
// Both lags' history VJPs in one pass over element patches
// (fx_patches_create with the nodes as DOFs).  A workgroup stages its
// patch's local node values -- coordinates, the two velocity entries of
// lags 0 and 1, the three entries of the next state and of the adjoint --
// in LDS with one gather each; one lane per element evaluates the partials
// once (gls_hvjp_nodes, as the two-phase form) and stores its three nodes'
// columns of both lags into the nodes' runs (the staged values' LDS, after a
// barrier); the runs are summed in element order.  A node inside one patch
// gets the two-phase VJP's sums bit for bit; a seam node's per-patch
// partials (four values: lag 0 x, y, lag 1 x, y) are summed by
// gls_vjp_seam_block in patch order.
struct GlsPatchOut
{
  double* out0; // lag 0, or null
  double* out1; // lag 1, or null
};

constexpr int kGlsNodeVals = 12; // x, y, cur (2), prev (2), nxt (3), adj (3)

__host__ __device__ constexpr int gls_patch_lds_main(int max_patch_ldofs)
{
  return 8 * (kGlsNodeVals * max_patch_ldofs > 12 * kPatchLanes ? kGlsNodeVals * max_patch_ldofs
                                                                 : 12 * kPatchLanes);
}
__host__ __device__ constexpr int gls_patch_lds(int max_patch_ldofs)
{
  return (gls_patch_lds_main(max_patch_ldofs) + 2 * max_patch_ldofs + 2 + 15) / 16 * 16;
}

__device__ __forceinline__ void gls_vjp_write(const GlsVjpArgs& a, const GlsPatchOut& o, int n,
                                              const double (&v)[4])
{
  const int r0 = a.node_rows[3 * n], r1 = a.node_rows[3 * n + 1], r2 = a.node_rows[3 * n + 2];
  if (o.out0)
  {
    __builtin_nontemporal_store(v[0], o.out0 + r0);
    __builtin_nontemporal_store(v[1], o.out0 + r1);
    __builtin_nontemporal_store(0.0, o.out0 + r2);
  }
  if (o.out1)
  {
    __builtin_nontemporal_store(v[2], o.out1 + r0);
    __builtin_nontemporal_store(v[3], o.out1 + r1);
    __builtin_nontemporal_store(0.0, o.out1 + r2);
  }
}

__device__ __forceinline__ void gls_vjp_patch_block(int p, const GlsVjpArgs& a, const PatchArgs& q,
                                                    const GlsPatchOut& o, unsigned char* lds,
                                                    int max_patch_ldofs)
{
  constexpr int   U  = 3; // a patch has at most 3 * kPatchLanes local nodes
  const int       t  = threadIdx.x;
  const int       i0 = q.patch_off[p], ne = q.patch_off[p + 1] - i0;
  const int       l0 = q.ldof_off[p], nl = q.ldof_off[p + 1] - l0;
  const bool      mine   = t < ne;
  const int       i      = i0 + t;
  const bool      extrap = a.step > 0 && a.num_hist > 1;
  double*         sv     = reinterpret_cast<double*>(lds); // node values, then the slots
  unsigned short* rs = reinterpret_cast<unsigned short*>(lds + gls_patch_lds_main(max_patch_ldofs));
  unsigned        w0 = 0, w1 = 0;
  if (mine)
  {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    const u2 lw = __builtin_nontemporal_load(reinterpret_cast<const u2*>(q.elem_ldof + 4 * static_cast<long>(i)));
    w0 = lw.x, w1 = lw.y;
  }
  if (t == 0) rs[0] = 0;
  for (int j0 = t; j0 < nl; j0 += U * kPatchLanes)
  {
    int  n[U], re[U], r[U][3];
    bool on[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      on[u]       = j0 + u * kPatchLanes < nl;
      const int j = l0 + j0 + u * kPatchLanes;
      n[u]        = on[u] ? q.ldof_dof[j] : 0;
      re[u]       = on[u] ? q.ldof_end[j] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < 3; ++k) r[u][k] = on[u] ? a.node_rows[3 * n[u] + k] : 0;
    double v[U][kGlsNodeVals];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      v[u][0] = on[u] ? gat32(a.x, n[u]) : 0.0;
      v[u][1] = on[u] ? gat32(a.y, n[u]) : 0.0;
#pragma unroll
      for (int c = 0; c < 2; ++c)
      {
        v[u][2 + c] = on[u] ? gat32(a.hist, r[u][c]) : 0.0;
        v[u][4 + c] = on[u] && extrap ? gat32(a.hist + a.num_states, r[u][c]) : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k)
      {
        v[u][6 + k] = on[u] ? gat32(a.nxt, r[u][k]) : 0.0;
        v[u][9 + k] = on[u] ? gat32(a.adj, r[u][k]) : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (on[u])
      {
        const int j = j0 + u * kPatchLanes;
#pragma unroll
        for (int k = 0; k < kGlsNodeVals; ++k) sv[kGlsNodeVals * j + k] = v[u][k];
        rs[j + 1] = static_cast<unsigned short>(re[u]);
      }
  }
  __syncthreads();
  double o0[3], o1[3], p0[3], p1[3];
  if (mine)
  {
    const unsigned h[3] = {w0 & 0xffffu, w0 >> 16, w1 & 0xffffu};
    double         xe[3], ye[3], cur[6], prev[6], nx[9], aj[9];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
      const double* nv = sv + kGlsNodeVals * (h[k] & 0xfffu);
      xe[k] = nv[0], ye[k] = nv[1];
      cur[2 * k] = nv[2], cur[2 * k + 1] = nv[3];
      prev[2 * k] = nv[4], prev[2 * k + 1] = nv[5];
      // next state and adjoint in the element DOF order: velocity (2k, 2k+1),
      // pressure 6+k
      nx[2 * k] = nv[6], nx[2 * k + 1] = nv[7], nx[6 + k] = nv[8];
      aj[2 * k] = nv[9], aj[2 * k + 1] = nv[10], aj[6 + k] = nv[11];
    }
    const double j00 = xe[1] - xe[0], j01 = xe[2] - xe[0], j10 = ye[1] - ye[0], j11 = ye[2] - ye[0];
    const double det = j00 * j11 - j01 * j10, inv_det = 1.0 / det;
    const double i00 = j11 * inv_det, i01 = -j01 * inv_det, i10 = -j10 * inv_det, i11 = j00 * inv_det;
    const double gx[3] = {-i00 - i10, i00, i10}, gy[3] = {-i01 - i11, i01, i11};
    const double jw    = fabs(det) * (1.0 / 6.0);
    const int    ks[3] = {0, 1, 2};
    gls_hvjp_nodes<3>(ks, gx, gy, jw, cur, prev, extrap, nx, aj, a.c, o0, o1, p0, p1);
  }
  __syncthreads(); // every lane has read the staged values: the slots overwrite them
  double* cs = sv;
  if (mine)
  {
    const unsigned h[3] = {w0 & 0xffffu, w0 >> 16, w1 & 0xffffu};
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
      double* sl = cs + 4 * (rs[h[k] & 0xfffu] + (h[k] >> 12));
      sl[0] = o0[k], sl[1] = o1[k], sl[2] = p0[k], sl[3] = p1[k];
    }
  }
  __syncthreads();
  for (int j = t; j < nl; j += kPatchLanes)
  {
    const int n = q.ldof_dof[l0 + j], part = q.ldof_part[l0 + j];
    double    v[4] = {0.0, 0.0, 0.0, 0.0};
    const int r1   = rs[j + 1];
    for (int s = rs[j]; s < r1; ++s)
    {
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] += cs[4 * s + c];
    }
    if (part < 0)
      gls_vjp_write(a, o, n, v);
    else
    {
      double* pp = q.parts + 4 * static_cast<long>(part & ~FX_TH_PART_FIRST);
#pragma unroll
      for (int c = 0; c < 4; ++c) __builtin_nontemporal_store(v[c], pp + c);
    }
  }
  __syncthreads(); // the LDS is reused by the next patch of a persistent walk
}

// Seam pass of the GLS history VJP (one lane per seam node): the node's
// per-patch partials (4 values each) in patch order.  A node in no element
// has no rows and is skipped.
__device__ __forceinline__ void gls_vjp_seam_block(int blk, const GlsVjpArgs& a, const SeamArgs& s,
                                                   const GlsPatchOut& o)
{
  const int k = blk * 256 + threadIdx.x;
  if (k >= s.num_seam) return;
  const int n = s.seam_dof[k];
  if (a.node_rows[3 * n] < 0) return;
  double    v[4] = {0.0, 0.0, 0.0, 0.0};
  const int t1   = s.seam_off[k + 1];
  for (int t = s.seam_off[k]; t < t1; ++t)
  {
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] += s.parts[4 * static_cast<long>(t) + c];
  }
  gls_vjp_write(a, o, n, v);
}

} // namespace fxk
