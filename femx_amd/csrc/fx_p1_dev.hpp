// P1 Poisson assembly and log-kappa VJP for gfx950: device code shared by
// the single-instance entry points (fx_p1.hip) and the batched ones
// (fx_batch.hip).
//
// Design: one lane per global row (node), gathering over the row's incident
// elements in ascending element order (fx_map_export_rowinc3).  Each lane
// recomputes the one element-matrix row it needs from the three vertex
// coordinates in registers (affine triangles: constant gradients), so no
// per-element quadrature data is streamed from HBM.  A workgroup owns 128
// consecutive rows: their incidence entries are one contiguous range, staged
// in LDS with coalesced loads, and their CSR values are one contiguous range,
// accumulated in LDS and written back with coalesced stores.  Streams read or
// written once (incidences, outputs) use non-temporal accesses so that the
// L2 keeps the node and element data that neighbouring rows share.  Every
// CSR value and residual entry is a sum over its elements in ascending
// element order, added by the one lane that owns the row (tile updates are
// LDS atomic adds only to skip the read round trip; no two lanes ever add to
// one entry): results are deterministic and equal, operation for operation, to
// femx's host assemble loop (femx/assembly/Assembly.hpp:184-225) evaluating
// PoissonComponents::evalRow (examples/poisson/PoissonComponents.hpp:35-64)
// on ElementValues' physical gradients (femx/fem/ElementValues.cpp:138-181).
// The file is compiled without floating-point contraction so that each
// product is rounded as in femx's host code.
#pragma once
#include "fx_dev.hpp"

namespace fxk
{


// 128 rows per workgroup: finer blocks keep more of the chip busy as the
// launch drains (2% faster than 256 at 2048^2; 64 is slower again).
constexpr int kP1Block = 128;
// LDS row tile: 128 rows x 32 x 8 B = 32 KiB (+8 KiB staging).  A block whose
// CSR range is longer than its tile (rows of more than 32 entries) adds into
// the values in global memory instead, each lane in its own row, in the same
// order: rows of any length the 8-bit column positions allow (255) assemble,
// bitwise as through the tile.
constexpr int kP1MaxRowLen = 32;

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
constexpr int kIncStage = 8 * kP1Block;

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

// The data one incidence reads: its element's three vertex coordinates,
// state and load values, and the element's diffusion value.
struct P1Raw
{
  double x[3], y[3], u[3], f[3], kv;
};

// The row's own node is vertex li of every element it meets, so its values
// (o) are loaded once per row and only the element's two other vertices are
// gathered per incidence; the three are then put back in element order.
__device__ __forceinline__ double p1_place(int li, int k, double own, double oa, double ob)
{
  // vertices other than li, in element order: li=0 -> (1,2), 1 -> (0,2), 2 -> (0,1)
  return k == 0 ? (li == 0 ? own : oa) : (k == 1 ? (li == 0 ? oa : (li == 1 ? own : ob)) : (li == 2 ? own : ob));
}

struct P1Own
{
  double x, y, u, f;
};

// The element's two vertices other than li, in element order:
// li=0 -> (1,2), 1 -> (0,2), 2 -> (0,1).
typedef int fx_i3 __attribute__((ext_vector_type(3), aligned(4)));

__device__ __forceinline__ int2 p1_others(const int* __restrict__ conn, int e, int li)
{
  const fx_i3 c = *reinterpret_cast<const fx_i3*>(conn + 3 * static_cast<long>(e));
  return make_int2(li == 0 ? c.y : c.x, li == 2 ? c.y : c.z);
}

__device__ __forceinline__ P1Raw p1_load_raw(const P1AsmArgs& a, int e, int2 n, int li, const P1Own& o)
{
  const int ia = n.x, ib = n.y;
  const double xa = gat32(a.x, ia), xb = gat32(a.x, ib), ya = gat32(a.y, ia), yb = gat32(a.y, ib);
  const double ua = a.u ? gat32(a.u, ia) : 0.0, ub = a.u ? gat32(a.u, ib) : 0.0;
  const double fa = a.f ? gat32(a.f, ia) : 0.0, fb = a.f ? gat32(a.f, ib) : 0.0;
  P1Raw R;
#pragma unroll
  for (int k = 0; k < 3; ++k)
  {
    R.x[k] = p1_place(li, k, o.x, xa, xb);
    R.y[k] = p1_place(li, k, o.y, ya, yb);
    R.u[k] = p1_place(li, k, o.u, ua, ub);
    R.f[k] = p1_place(li, k, o.f, fa, fb);
  }
  R.kv = a.kappa ? gat32(a.kappa, e) : 1.0;
  return R;
}

// One row of the P1 assembly: the lane walks the row's incidences with the
// next incidence's vertex indices in flight while the current one computes
// (the lookahead past the row's last incidence re-reads that incidence,
// always valid, and is discarded).  The row's own node is loaded once; per
// incidence only the element's two other vertices and its diffusion value
// are gathered.  racc is the row's first value: in an LDS tile (tiled; the
// three positions of an incidence are distinct entries of the lane's own
// row, so each LDS atomic add makes it old + K in program order, as the host
// loop adds them) or in a.vals (read, add, write, in order).  The residual
// is stored; the values stay in racc.
__device__ __forceinline__ void p1_assemble_row(const P1AsmArgs& a, int r, double* racc, bool tiled,
                                                const long long* sinc, bool staged, int ib0)
{
  double    rsum = 0.0;
  const int ta = a.inc_off[r], t1 = a.inc_off[r + 1], tl = t1 - 1;
  if (!tiled && a.vals)
    for (int k = 0; k < a.row_ptr[r + 1] - a.row_ptr[r]; ++k) racc[k] = 0.0;
  if (ta < t1)
  {
    const P1Own o{a.x[r], a.y[r], a.u ? a.u[r] : 0.0, a.f ? a.f[r] : 0.0};
    int2        ic = inc_at(sinc, a.inc, staged, ta, ib0);
    int2        cc = p1_others(a.conn, ic.x, ic.y & 0xff);
    for (int t = ta; t < t1; ++t)
    {
      const int2     in   = inc_at(sinc, a.inc, staged, min(t + 1, tl), ib0);
      const unsigned meta = static_cast<unsigned>(ic.y);
      const int      li   = meta & 0xff;
      const P1Raw    R    = p1_load_raw(a, ic.x, cc, li, o);
      cc                  = p1_others(a.conn, in.x, in.y & 0xff);
      const P1Geom   g    = p1_geom(R.x[0], R.y[0], R.x[1], R.y[1], R.x[2], R.y[2]);
      const double   ke   = (a.kappa && a.kappa_is_log) ? exp(R.kv) : R.kv;
      const double   k0 = p1_k(g, li, 0, ke), k1 = p1_k(g, li, 1, ke), k2 = p1_k(g, li, 2, ke);
      double         re = 0.0;
      if (a.u) re = k0 * R.u[0] + k1 * R.u[1] + k2 * R.u[2];
      if (a.f) re = re - p1_load(g, li, R.f[0], R.f[1], R.f[2]);
      else re = re - 0.0;
      rsum = rsum + re;
      if (a.vals)
      {
        double* p0 = racc + ((meta >> 8) & 0xff);
        double* p1 = racc + ((meta >> 16) & 0xff);
        double* p2 = racc + (meta >> 24);
        if (tiled)
        {
          lds_add(p0, k0);
          lds_add(p1, k1);
          lds_add(p2, k2);
        }
        else
        {
          // the lane's own row in global memory: read, add, write, in order
          *p0 = *p0 + k0;
          *p1 = *p1 + k1;
          *p2 = *p2 + k2;
        }
      }
      ic = in;
    }
  }
  __builtin_nontemporal_store(rsum, a.res + r);
}

// One row block of the P1 assembly (rows [blk*kP1Block, (blk+1)*kP1Block)); smem holds
// acc_cap doubles of row tile followed by the incidence stage.
// kLong: rows may be longer than the tile width (max_row_len > kP1MaxRowLen),
// so a block whose CSR range does not fit the tile adds in global memory;
// false compiles the tile path alone.
template <bool kLong>
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
  const bool  tiled  = !kLong || bn <= a.acc_cap; // else: rows add straight into a.vals
  if (tiled)
    for (int i = threadIdx.x; i < bn; i += kP1Block) acc[i] = 0.0;
  if (staged)
    stage_nt<kP1Block>(sinc, reinterpret_cast<const long long*>(a.inc) + ib0, ib1 - ib0);
  __syncthreads();

  const int r = r0 + threadIdx.x;
  if (r < a.num_rows)
  {
    // two calls, so that each inlined copy knows its row's address space
    if (tiled)
      p1_assemble_row(a, r, acc + (a.row_ptr[r] - base), true, sinc, staged, ib0);
    else
      p1_assemble_row(a, r, a.vals + a.row_ptr[r], false, sinc, staged, ib0);
  }
  __syncthreads();
  if (tiled)
    for (int i = threadIdx.x; i < bn; i += kP1Block)
      __builtin_nontemporal_store(acc[i], a.vals + base + i);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// P1 assembly over stencil blocks (fx_p1_stencil_plan): one wave per block
// of 64 rows.  In a uniform block every row has the same column offsets and
// the same incidences -- element A r + c_k, local row li_k, column positions
// p0..p2 -- so the lane of row r finds vertex m of its k-th element at
// column slot p_m.  The vertex data (x, y, u, f) of the block's columns are
// loaded once into the wave's LDS band buffer, each band of consecutive
// offsets as one coalesced run of 63 + w values per array (slot q of lane l
// is then band value sidx_q + l), and each row accumulates its S values in
// registers.  The arithmetic, and the order of every sum, are
// p1_assemble_row's, so the results are bitwise equal; the block's 64 S
// values leave through the same LDS with coalesced stores.  Other blocks
// take p1_assemble_row, adding in global memory.
constexpr int kP1StMax   = FX_P1_STENCIL_MAX;
constexpr int kP1StBand  = FX_P1_STENCIL_BAND;
constexpr int kP1StWaves = 4;
static_assert(4 * kP1StBand >= 64 * kP1StMax, "the band buffer holds the write-back");

// Row li of an element's matrix and its residual entry, as p1_assemble_row
// forms them (same operations, same order), for a compile-time li.
template <int LI>
__device__ __forceinline__ void p1_row_fixed(const P1AsmArgs& a, const P1Geom& g, double ke,
                                             double u0, double u1, double u2, double f0,
                                             double f1, double f2, double& k0, double& k1,
                                             double& k2, double& re)
{
  k0 = p1_k(g, LI, 0, ke), k1 = p1_k(g, LI, 1, ke), k2 = p1_k(g, LI, 2, ke);
  re = 0.0;
  if (a.u) re = k0 * u0 + k1 * u1 + k2 * u2;
  if (a.f) re = re - p1_load(g, LI, f0, f1, f2);
  else re = re - 0.0;
}

__device__ __forceinline__ void p1_assemble_stencil_block(int b, const P1AsmArgs& a,
                                                          const int* __restrict__ blk_code,
                                                          const int* __restrict__ words,
                                                          double* bb)
{
  const int lane = threadIdx.x & 63;
  const int r0   = 64 * b;
  if (r0 >= a.num_rows) return;
  const int  r    = r0 + lane;
  const int  code = blk_code[b];
  const int* rec  = words + (code >= 0 ? code : 0);
  if (code < 0 || rec[0] > kP1StMax || rec[1] > kP1StMax)
  {
    if (r < a.num_rows)
      p1_assemble_row(a, r, a.vals ? a.vals + a.row_ptr[r] : nullptr, false, nullptr, false, 0);
    return;
  }
  const int      S = rec[0], T = rec[1];
  const unsigned A    = static_cast<unsigned>(rec[2]);
  const int*     off  = rec + 3;
  const int*     sidx = rec + 3 + S;
  const int*     inc  = rec + 3 + 2 * S;
  // the record's band buffer must fit this kernel's (a plan built with a
  // larger max_band takes the row path)
  int used = 0;
  for (int q = 0; q < S; ++q)
    if (q == S - 1 || off[q + 1] != off[q] + 1) used = sidx[q] + 64;
  if (used > kP1StBand)
  {
    p1_assemble_row(a, r, a.vals ? a.vals + a.row_ptr[r] : nullptr, false, nullptr, false, 0);
    return;
  }
  double* bx = bb, *by = bb + kP1StBand, *bu = bb + 2 * kP1StBand, *bf = bb + 3 * kP1StBand;
  // Buffer position v = lane + 64 j holds value i = v - sidx[q] of the band
  // q it falls in: row r0 + i's column at offset off[q] (i < 64), or row
  // r0 + 63's at offset off[q] + i - 63.  Every position's column is found
  // first and all loads are issued before the first is waited on (one memory
  // round trip per record, not one per band).
  constexpr int kSlots = (kP1StBand + 63) / 64;
  int           col[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) col[j] = -1;
  for (int q = 0; q < S;)
  {
    const int o = off[q];
    int       w = 1;
    while (q + w < S && off[q + w] == o + w) ++w;
    const int base = sidx[q], len = 63 + w;
#pragma unroll
    for (int j = 0; j < kSlots; ++j)
    {
      const int i = lane + 64 * j - base;
      if (i >= 0 && i < len) col[j] = r0 + o + i;
    }
    q += w;
  }
  // positions outside every band load row r0's values (a valid node, one
  // cache line for the wave) and are not stored; an absent u or f loads x
  // and stores zero
  const double* pu = a.u ? a.u : a.x;
  const double* pf = a.f ? a.f : a.x;
  double        xv[kSlots], yv[kSlots], uv[kSlots], fv[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j)
  {
    const int c = col[j] >= 0 ? col[j] : r0;
    xv[j] = gat32(a.x, c), yv[j] = gat32(a.y, c), uv[j] = gat32(pu, c), fv[j] = gat32(pf, c);
  }
#pragma unroll
  for (int j = 0; j < kSlots; ++j)
    if (col[j] >= 0)
    {
      const int v = lane + 64 * j;
      bx[v] = xv[j], by[v] = yv[j], bu[v] = a.u ? uv[j] : 0.0, bf[v] = a.f ? fv[j] : 0.0;
    }
  wave_sync();
  double acc[kP1StMax];
#pragma unroll
  for (int q = 0; q < kP1StMax; ++q) acc[q] = 0.0;
  double rsum = 0.0;
  // kappa one incidence ahead: the next element's value is in flight while
  // the current one computes
  auto kappa_at = [&](int k) {
    return a.kappa ? gat32(a.kappa, static_cast<int>(A * static_cast<unsigned>(r)
                                                     + static_cast<unsigned>(inc[2 * k])))
                   : 1.0;
  };
  double kv_next = kappa_at(0);
  for (int k = 0; k < T; ++k)
  {
    const unsigned meta = static_cast<unsigned>(inc[2 * k + 1]);
    const int      li = meta & 0xff, p0 = (meta >> 8) & 0xff, p1 = (meta >> 16) & 0xff, p2 = meta >> 24;
    const int      s0 = sidx[p0] + lane, s1 = sidx[p1] + lane, s2 = sidx[p2] + lane;
    const double   kv = kv_next;
    if (k + 1 < T) kv_next = kappa_at(k + 1);
    const P1Geom   g  = p1_geom(bx[s0], by[s0], bx[s1], by[s1], bx[s2], by[s2]);
    const double   ke = (a.kappa && a.kappa_is_log) ? exp(kv) : kv;
    const double   u0 = a.u ? bu[s0] : 0.0, u1 = a.u ? bu[s1] : 0.0, u2 = a.u ? bu[s2] : 0.0;
    const double   f0 = a.f ? bf[s0] : 0.0, f1 = a.f ? bf[s1] : 0.0, f2 = a.f ? bf[s2] : 0.0;
    double         k0, k1, k2, re;
    // li is the same for every lane: one branch into a copy of the row
    // arithmetic with li a constant (no per-lane selects)
    switch (li)
    {
    case 0: p1_row_fixed<0>(a, g, ke, u0, u1, u2, f0, f1, f2, k0, k1, k2, re); break;
    case 1: p1_row_fixed<1>(a, g, ke, u0, u1, u2, f0, f1, f2, k0, k1, k2, re); break;
    default: p1_row_fixed<2>(a, g, ke, u0, u1, u2, f0, f1, f2, k0, k1, k2, re); break;
    }
    rsum    = rsum + re;
    acc[p0] = acc[p0] + k0;
    acc[p1] = acc[p1] + k1;
    acc[p2] = acc[p2] + k2;
  }
  __builtin_nontemporal_store(rsum, a.res + r);
  if (a.vals)
  {
    const int base = a.row_ptr[r0], n = 64 * S;
    wave_sync(); // every lane has read the band buffer: it takes the row values now
#pragma unroll
    for (int q = 0; q < kP1StMax; ++q)
      if (q < S) bb[lane * S + q] = acc[q];
    wave_sync();
    for (int i = lane; i < n; i += 64) __builtin_nontemporal_store(bb[i], a.vals + base + i);
  }
  wave_sync(); // the buffer is free for the next block
}

// Rows from a list (the rows an element-patch pass leaves: nodes on patch
// seams), kP1Block per block, one lane per row, each lane with its own LDS
// row of kP1ListRow values (rows longer than that add in global memory).
constexpr int kP1ListRow = 16, kP1ListStride = 17; // odd stride: lanes' rows in distinct banks

__device__ __forceinline__ void p1_assemble_list_block(int blk, const P1AsmArgs& a,
                                                       const int* __restrict__ rows, int nrows,
                                                       double* smem)
{
  const int k = blk * kP1Block + threadIdx.x;
  if (k >= nrows) return;
  const int r  = rows[k];
  const int b0 = a.row_ptr[r], n = a.row_ptr[r + 1] - b0;
  // two calls, so that each inlined copy knows its row's address space (LDS
  // atomics for the tile, not flat ones)
  if (n <= kP1ListRow)
  {
    double* racc = smem + threadIdx.x * kP1ListStride;
    if (a.vals)
      for (int i = 0; i < n; ++i) racc[i] = 0.0;
    p1_assemble_row(a, r, racc, true, nullptr, false, 0);
    if (a.vals)
      for (int i = 0; i < n; ++i) a.vals[b0 + i] = racc[i];
  }
  else
    p1_assemble_row(a, r, a.vals + b0, false, nullptr, false, 0);
}

// ---------------------------------------------------------------------------
// P1 assembly over element patches (the VJP's Morton-ordered patches, their
// elements in ascending order): one lane per element computes the element's
// three matrix rows and residual entries once, from node values staged in
// LDS once per patch.  A node all of whose elements lie in the patch
// ("finished" here) has its CSR row in the patch's LDS tile and its residual
// entry in an LDS slot; the lanes add into them in rank steps -- step k adds
// the rows of every element that is its node's k-th element in ascending
// element order -- so each value is still old + K in ascending element order,
// as in p1_assemble_row, and the results are bitwise equal.  The tile holds
// the finished rows in ascending node order, so consecutive node ids form
// runs that are contiguous in the CSR values too: the write-back walks the
// tile with consecutive lanes on consecutive values.  Rows on patch seams
// (and any a patch cannot tile) are left to p1_assemble_list_block.
struct P1AsmPatch
{
  const unsigned short* elem_pos;   // [4 per patch-ordered element]: word k holds the
                                    // CSR positions (4 bits each) of the element's
                                    // three columns in the row of its vertex k
  const unsigned short* ldof_tile;  // [per local node]: tile offset of a finished row, or kP1NoTile
  const int*            patch_tile; // [per patch]: tile length | rank steps << 20
  const int*            run_off;    // [patches + 1]: the patch's runs in `runs`
  const int2*           runs;       // tile offset and CSR offset of each run
};

constexpr unsigned kP1NoTile = 0xffffu;

// LDS: the tile, x, y, u, f and the residual slot of each local node, the
// runs (and an end marker), then each local node's tile offset.
__host__ __device__ constexpr int p1_asm_patch_lds(int max_patch_ldofs, int max_tile, int max_runs)
{
  return (8 * max_tile + 40 * max_patch_ldofs + 8 * (max_runs + 1) + 2 * max_patch_ldofs + 15) / 16
         * 16;
}

__device__ __forceinline__ void p1_asm_patch_block(int p, const P1AsmArgs& a, const PatchArgs& q,
                                                   const P1AsmPatch& s, unsigned char* lds,
                                                   int max_patch_ldofs, int max_tile, int max_runs)
{
  constexpr int   U  = 3; // a P1 patch has at most 3 * kPatchLanes local nodes
  const int       t  = threadIdx.x;
  const int       i0 = q.patch_off[p], ne = q.patch_off[p + 1] - i0;
  const int       l0 = q.ldof_off[p], nl = q.ldof_off[p + 1] - l0;
  const int       pt = s.patch_tile[p], tlen = pt & 0xfffff, nrank = pt >> 20;
  const int       u0 = s.run_off[p], nr = s.run_off[p + 1] - u0;
  const bool      mine = t < ne;
  const int       i    = i0 + t;
  double*         tile = reinterpret_cast<double*>(lds);
  double*         xs   = tile + max_tile;
  double*         ys   = xs + max_patch_ldofs;
  double*         us   = ys + max_patch_ldofs;
  double*         fs   = us + max_patch_ldofs;
  double*         rt   = fs + max_patch_ldofs;
  int2*           ru   = reinterpret_cast<int2*>(rt + max_patch_ldofs);
  unsigned short* ts   = reinterpret_cast<unsigned short*>(ru + max_runs + 1);
  // the element's tables and coefficient are requested first, so that their
  // latency overlaps the staging below
  unsigned w0 = 0, w1 = 0, c0 = 0, c1 = 0;
  double   kv = 1.0;
  if (mine)
  {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    const u2 lw = __builtin_nontemporal_load(reinterpret_cast<const u2*>(q.elem_ldof + 4 * static_cast<long>(i)));
    const u2 pw = __builtin_nontemporal_load(reinterpret_cast<const u2*>(s.elem_pos + 4 * static_cast<long>(i)));
    w0 = lw.x, w1 = lw.y, c0 = pw.x, c1 = pw.y;
    if (a.kappa) kv = gat32(a.kappa, q.patch_elem[i]);
  }
  for (int k = t; k < tlen; k += kPatchLanes) tile[k] = 0.0;
  for (int k = t; k < nr; k += kPatchLanes) ru[k] = s.runs[u0 + k];
  if (t == 0) ru[nr] = make_int2(tlen, 0);
  for (int j0 = t; j0 < nl; j0 += U * kPatchLanes)
  {
    int  d[U], tf[U];
    bool on[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      on[u]       = j0 + u * kPatchLanes < nl;
      const int j = l0 + j0 + u * kPatchLanes;
      d[u]        = on[u] ? q.ldof_dof[j] : 0;
      tf[u]       = on[u] ? s.ldof_tile[j] : kP1NoTile;
    }
    double xv[U], yv[U], uv[U], fv[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      xv[u] = on[u] ? gat32(a.x, d[u]) : 0.0;
      yv[u] = on[u] ? gat32(a.y, d[u]) : 0.0;
      uv[u] = on[u] && a.u ? gat32(a.u, d[u]) : 0.0;
      fv[u] = on[u] && a.f ? gat32(a.f, d[u]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (on[u])
      {
        const int j = j0 + u * kPatchLanes;
        xs[j] = xv[u], ys[j] = yv[u], us[j] = uv[u], fs[j] = fv[u], rt[j] = 0.0;
        ts[j] = static_cast<unsigned short>(tf[u]);
      }
  }
  __syncthreads();
  const unsigned h[3]  = {w0 & 0xffffu, w0 >> 16, w1 & 0xffffu};
  const unsigned cp[3] = {c0 & 0xffffu, c0 >> 16, c1 & 0xffffu};
  const unsigned j[3]  = {h[0] & 0xfffu, h[1] & 0xfffu, h[2] & 0xfffu};
  double         K[3][3], R[3];
  unsigned       tof[3] = {kP1NoTile, kP1NoTile, kP1NoTile};
  if (mine)
  {
    const P1Geom g  = p1_geom(xs[j[0]], ys[j[0]], xs[j[1]], ys[j[1]], xs[j[2]], ys[j[2]]);
    const double ke = (a.kappa && a.kappa_is_log) ? exp(kv) : kv;
    const double uu[3] = {us[j[0]], us[j[1]], us[j[2]]};
    const double ff[3] = {fs[j[0]], fs[j[1]], fs[j[2]]};
#pragma unroll
    for (int li = 0; li < 3; ++li)
    {
#pragma unroll
      for (int c = 0; c < 3; ++c) K[li][c] = p1_k(g, li, c, ke);
      double re = 0.0;
      if (a.u) re = K[li][0] * uu[0] + K[li][1] * uu[1] + K[li][2] * uu[2];
      if (a.f) re = re - p1_load(g, li, ff[0], ff[1], ff[2]);
      else re = re - 0.0;
      R[li]   = re;
      tof[li] = ts[j[li]];
    }
  }
  for (int rk = 0; rk < nrank; ++rk)
  {
    if (mine)
    {
#pragma unroll
      for (int li = 0; li < 3; ++li)
        if ((h[li] >> 12) == static_cast<unsigned>(rk) && tof[li] != kP1NoTile)
        {
          lds_add(rt + j[li], R[li]);
          if (a.vals)
          {
            double* row = tile + tof[li];
#pragma unroll
            for (int c = 0; c < 3; ++c) lds_add(row + ((cp[li] >> (4 * c)) & 0xfu), K[li][c]);
          }
        }
    }
    __syncthreads();
  }
  for (int k = t; k < nl; k += kPatchLanes)
    if (ts[k] != kP1NoTile) __builtin_nontemporal_store(rt[k], a.res + q.ldof_dof[l0 + k]);
  if (a.vals)
  {
    // value k of the tile lies in run m, the last run starting at or before k
    for (int k = t; k < tlen; k += kPatchLanes)
    {
      int lo = 0, hi = nr - 1;
      while (lo < hi)
      {
        const int mid = (lo + hi + 1) >> 1;
        if (ru[mid].x <= k) lo = mid;
        else hi = mid - 1;
      }
      __builtin_nontemporal_store(tile[k], a.vals + ru[lo].y + (k - ru[lo].x));
    }
  }
  __syncthreads(); // the LDS is reused by the next patch of a persistent walk
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
// The VJP of row (node) j: its incidences from the block's LDS stage or,
// unstaged, from memory.
__device__ __forceinline__ void p1_vjp_row(const P1VjpArgs& a, int j, const long long* sinc,
                                           bool staged, int ib0)
{
  double    acc = 0.0;
  const int ta = a.inc_off[j], t1 = a.inc_off[j + 1], tl = t1 - 1;
  if (ta < t1)
  {
    // node j is vertex li of each of its elements: its values are loaded
    // once, the element's two other vertices are gathered per incidence
    // (their indices one incidence ahead)
    // both gathers issued together (n < num_rows for both arrays), then a select
    auto masked = [&](int n) {
      const double l = gat32(a.lam, n);
      const int    m = a.bc_mask ? gat32(a.bc_mask, n) : 0;
      return m ? 0.0 : l;
    };
    const double xo = a.x[j], yo = a.y[j], lo = masked(j), uo = a.u[j];
    int2         ic = inc_at(sinc, a.inc, staged, ta, ib0);
    int2         cc = p1_others(a.conn, ic.x, ic.y & 0xff);
    for (int t = ta; t < t1; ++t)
    {
      const int2   in = inc_at(sinc, a.inc, staged, min(t + 1, tl), ib0);
      const int    e  = ic.x;
      const int    li = static_cast<unsigned>(ic.y) & 0xff;
      const double xa = gat32(a.x, cc.x), xb = gat32(a.x, cc.y);
      const double ya = gat32(a.y, cc.x), yb = gat32(a.y, cc.y);
      const double la = masked(cc.x), lb = masked(cc.y);
      const double me = gat32(a.m, e);
      const int2   nx = cc;
      cc              = p1_others(a.conn, in.x, in.y & 0xff);
      const P1Geom g  = p1_geom(p1_place(li, 0, xo, xa, xb), p1_place(li, 0, yo, ya, yb),
                                p1_place(li, 1, xo, xa, xb), p1_place(li, 1, yo, ya, yb),
                                p1_place(li, 2, xo, xa, xb), p1_place(li, 2, yo, ya, yb));
      const double ke = exp(me);
      const double le[3] = {p1_place(li, 0, lo, la, lb), p1_place(li, 1, lo, la, lb),
                            p1_place(li, 2, lo, la, lb)};
      const double c0 = p1_k(g, 0, li, ke), c1 = p1_k(g, 1, li, ke), c2 = p1_k(g, 2, li, ke);
      acc = acc + (c0 * le[0] + c1 * le[1] + c2 * le[2]);
      if (li == 0)
      {
        // the element's vertex 0 is node j; vertices 1 and 2 are its others
        const double u0 = uo, u1 = gat32(a.u, nx.x), u2 = gat32(a.u, nx.y);
        double       ge = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
          const double re = p1_k(g, r, 0, ke) * u0 + p1_k(g, r, 1, ke) * u1 + p1_k(g, r, 2, ke) * u2;
          ge              = ge + le[r] * (re - 0.0);
        }
        a.gout[e] = ge;
      }
      ic = in;
    }
  }
  if (a.bc_mask && a.bc_mask[j]) acc = acc + a.lam[j];
  __builtin_nontemporal_store(acc, a.jt + j);
}

__device__ __forceinline__ void p1_vjp_block(int blk, const P1VjpArgs& a, long long* sinc)
{
  const int  r0 = blk * kP1Block, r1 = min(r0 + kP1Block, a.num_rows);
  const int  ib0 = a.inc_off[r0], ib1 = a.inc_off[r1];
  const bool staged = ib1 - ib0 <= kIncStage;
  if (staged)
    stage_nt<kP1Block>(sinc, reinterpret_cast<const long long*>(a.inc) + ib0, ib1 - ib0);
  __syncthreads();
  const int j = r0 + threadIdx.x;
  if (j < a.num_rows) p1_vjp_row(a, j, sinc, staged, ib0);
  __syncthreads();
}


// ---------------------------------------------------------------------------
// The same VJP in one pass over element patches (fx_patches_create with the
// nodes as DOFs): a workgroup stages its patch's local node values (x, y, u,
// masked lam, constraint flag) in LDS with one gather each, then one lane per
// element forms the element's three columns of K^T lam~ -- each exactly as
// the row-gather lane above forms it -- and g_e, stores the columns into the
// nodes' runs, and the runs are summed in element order.  A node inside one
// patch gets the row-gather VJP's sum bit for bit; a seam node's partials
// are summed by the seam pass.

// LDS of one P1 patch workgroup, in bytes: the slots (3 per element), four
// doubles per local node, the runs' starts (u16) and constraint flags (u8).
__host__ __device__ constexpr int p1_patch_lds(int max_patch_ldofs)
{
  return (8 * 3 * kPatchLanes + 32 * max_patch_ldofs + 3 * max_patch_ldofs + 2 + 15) / 16 * 16;
}

__device__ __forceinline__ void p1_vjp_patch_block(int p, const P1VjpArgs& a, const PatchArgs& q,
                                                   unsigned char* lds, int max_patch_ldofs)
{
  constexpr int   U  = 3; // a P1 patch has at most 3 * kPatchLanes local nodes
  const int       t  = threadIdx.x;
  const int       i0 = q.patch_off[p], ne = q.patch_off[p + 1] - i0;
  const int       l0 = q.ldof_off[p], nl = q.ldof_off[p + 1] - l0;
  const bool      mine = t < ne;
  const int       i    = i0 + t;
  double*         cs   = reinterpret_cast<double*>(lds);
  double*         xs   = cs + 3 * kPatchLanes;
  double*         ys   = xs + max_patch_ldofs;
  double*         us   = ys + max_patch_ldofs;
  double*         ls   = us + max_patch_ldofs;
  unsigned short* rs   = reinterpret_cast<unsigned short*>(ls + max_patch_ldofs);
  unsigned char*  mk   = reinterpret_cast<unsigned char*>(rs + max_patch_ldofs + 1);
  // the element's tables and log-coefficient are requested first, so that
  // their latency overlaps the staging below
  int      e  = 0;
  unsigned w0 = 0, w1 = 0;
  double   me = 0.0;
  if (mine)
  {
    e              = q.patch_elem[i];
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    const u2 lw    = __builtin_nontemporal_load(reinterpret_cast<const u2*>(q.elem_ldof + 4 * static_cast<long>(i)));
    w0 = lw.x, w1 = lw.y;
    me = gat32(a.m, e);
  }
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
    double xv[U], yv[U], uv[U], lv[U];
    int    m[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      xv[u] = on[u] ? gat32(a.x, d[u]) : 0.0;
      yv[u] = on[u] ? gat32(a.y, d[u]) : 0.0;
      uv[u] = on[u] ? gat32(a.u, d[u]) : 0.0;
      lv[u] = on[u] ? gat32(a.lam, d[u]) : 0.0;
      m[u]  = on[u] && a.bc_mask ? gat32(a.bc_mask, d[u]) : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j0 + u * kPatchLanes < nl)
      {
        const int j = j0 + u * kPatchLanes;
        xs[j] = xv[u], ys[j] = yv[u], us[j] = uv[u];
        ls[j]     = m[u] ? 0.0 : lv[u];
        rs[j + 1] = static_cast<unsigned short>(re[u]);
        mk[j]     = m[u] ? 1 : 0;
      }
  }
  __syncthreads();
  if (mine)
  {
    const unsigned h[3] = {w0 & 0xffffu, w0 >> 16, w1 & 0xffffu};
    const unsigned j[3] = {h[0] & 0xfffu, h[1] & 0xfffu, h[2] & 0xfffu};
    const P1Geom   g  = p1_geom(xs[j[0]], ys[j[0]], xs[j[1]], ys[j[1]], xs[j[2]], ys[j[2]]);
    const double   ke = exp(me);
    const double   le[3] = {ls[j[0]], ls[j[1]], ls[j[2]]};
#pragma unroll
    for (int li = 0; li < 3; ++li)
    {
      const double c = p1_k(g, 0, li, ke) * le[0] + p1_k(g, 1, li, ke) * le[1]
                       + p1_k(g, 2, li, ke) * le[2];
      cs[rs[j[li]] + (h[li] >> 12)] = c;
    }
    if (a.gout)
    {
      const double u0 = us[j[0]], u1 = us[j[1]], u2 = us[j[2]];
      double       ge = 0.0;
#pragma unroll
      for (int r = 0; r < 3; ++r)
      {
        const double re = p1_k(g, r, 0, ke) * u0 + p1_k(g, r, 1, ke) * u1 + p1_k(g, r, 2, ke) * u2;
        ge              = ge + le[r] * (re - 0.0);
      }
      __builtin_nontemporal_store(ge, a.gout + e);
    }
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

} // namespace fxk
