/*
 * femx_oracle.c -- CPU restatement of femx's data-parallel FE core.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * timed CPU baseline of bench.py.  Nothing in the product links this file.
 *
 * Built with -ffp-contract=off so that every product of a sum is rounded on
 * its own, as femx's host code is on x86-64 without FMA.  Each function cites
 * the femx source it restates (paths relative to the femx repository root).
 * Extensions that femx does not have (kappa, load, VJP, P2, 7-point rule,
 * Taylor-Hood Navier-Stokes) say so and are pinned in tests/ by closed forms,
 * finite differences, manufactured solutions and invariants.
 */
#include "femx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void or_free(void* p) { free(p); }

int or_omp_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Meshes.  femx has only Mesh::makeStructuredQuad (femx/fem/Mesh.cpp:10-51);
 * the triangle split and P2 node order are this build's conventions: cell
 * (i,j) with corners a=(i,j) b=(i+1,j) c=(i+1,j+1) d=(i,j+1) is split along
 * a-c into (a,b,c) and (a,c,d), both counter-clockwise; P2 nodes follow Gmsh
 * order: vertices, then edge midpoints (0,1), (1,2), (2,0).  Coordinates use
 * x = x_min + i*dx as Mesh.cpp:17-28 does. */
void or_mesh_p1(int n, double* x, double* y, or_index* conn)
{
  const double dx = 1.0 / (double)n;
  const int    np = n + 1;
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i)
    {
      x[j * np + i] = 0.0 + (double)i * dx;
      y[j * np + i] = 0.0 + (double)j * dx;
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
    {
      const or_index a = j * np + i, b = a + 1, d = a + np, c = d + 1;
      or_index*      e = conn + 6 * (j * n + i);
      e[0] = a; e[1] = b; e[2] = c;
      e[3] = a; e[4] = c; e[5] = d;
    }
}

void or_mesh_p2(int n, double* x, double* y, or_index* conn6, or_index* vconn3)
{
  const int    m  = 2 * n, mp = m + 1, np = n + 1;
  const double dx = 1.0 / (double)m;
  for (int J = 0; J <= m; ++J)
    for (int I = 0; I <= m; ++I)
    {
      x[J * mp + I] = 0.0 + (double)I * dx;
      y[J * mp + I] = 0.0 + (double)J * dx;
    }
#define L(I, J) ((J) * mp + (I))
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
    {
      const int I = 2 * i, J = 2 * j;
      or_index* e = conn6 + 12 * (j * n + i);
      /* (a,b,c) */
      e[0] = L(I, J); e[1] = L(I + 2, J); e[2] = L(I + 2, J + 2);
      e[3] = L(I + 1, J); e[4] = L(I + 2, J + 1); e[5] = L(I + 1, J + 1);
      /* (a,c,d) */
      e[6] = L(I, J); e[7] = L(I + 2, J + 2); e[8] = L(I, J + 2);
      e[9] = L(I + 1, J + 1); e[10] = L(I + 1, J + 2); e[11] = L(I, J + 1);
      or_index*      v = vconn3 + 6 * (j * n + i);
      const or_index a = j * np + i, b = a + 1, d = a + np, c = d + 1;
      v[0] = a; v[1] = b; v[2] = c;
      v[3] = a; v[4] = c; v[5] = d;
    }
#undef L
}

/* Taylor-Hood DOFs: FESpace::localDof/globalDof (femx/fem/FESpace.cpp:96-106:
 * local = comps*shape + c, global = comps*node + c) for the velocity field and
 * MixedFESpace::elemDofs (femx/fem/MixedFESpace.cpp:121-137: fields in order,
 * each offset by the preceding fields' DOF count) for the pressure field. */
void or_th_dofs(int n, or_index num_elems, const or_index* conn6, const or_index* vconn3,
                or_index* dofs15)
{
  const or_index nv = (2 * n + 1) * (2 * n + 1);
  for (or_index e = 0; e < num_elems; ++e)
  {
    or_index* d = dofs15 + 15 * e;
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < 2; ++c) d[2 * a + c] = 2 * conn6[6 * e + a] + c;
    for (int b = 0; b < 3; ++b) d[12 + b] = 2 * nv + vconn3[3 * e + b];
  }
}

/* ------------------------------------------------------------------------ */
/* makeAssemblyMap: femx/assembly/AssemblyMap.cpp:54-176.  COO of all local
 * (row, col) pairs in row-major local order (:108-121), sorted by (row, col)
 * (:123-132), deduplicated into CSR with ascending columns (:138-156), and
 * jac_map[k] = CSR index of local pair k (:150). */
typedef struct
{
  or_index row, col, idx;
} coo_t;

static int coo_cmp(const void* a, const void* b)
{
  const coo_t* l = (const coo_t*)a;
  const coo_t* r = (const coo_t*)b;
  if (l->row != r->row) return l->row < r->row ? -1 : 1;
  if (l->col != r->col) return l->col < r->col ? -1 : 1;
  return (l->idx > r->idx) - (l->idx < r->idx);
}

int or_map_build(or_index num_res, or_index num_states, or_index num_elems,
                 const or_index* res_offsets, const or_index* res_dofs,
                 const or_index* state_offsets, const or_index* state_dofs,
                 or_index** row_ptr_out, or_index** col_ind_out, or_index** jac_off_out,
                 or_index** jac_map_out, or_index* nnz_out)
{
  if (num_res < 0 || num_states < 0) return 1;
  or_index* jac_off = (or_index*)malloc(sizeof(or_index) * (num_elems + 1));
  int64_t   total   = 0;
  jac_off[0]        = 0;
  for (or_index e = 0; e < num_elems; ++e)
  {
    for (or_index p = res_offsets[e]; p < res_offsets[e + 1]; ++p)
      if (res_dofs[p] < 0 || res_dofs[p] >= num_res) { free(jac_off); return 2; }
    for (or_index p = state_offsets[e]; p < state_offsets[e + 1]; ++p)
      if (state_dofs[p] < 0 || state_dofs[p] >= num_states) { free(jac_off); return 2; }
    total += (int64_t)(res_offsets[e + 1] - res_offsets[e])
             * (state_offsets[e + 1] - state_offsets[e]);
    if (total > 2147483647LL) { free(jac_off); return 3; }
    jac_off[e + 1] = (or_index)total;
  }
  const or_index n   = (or_index)total;
  coo_t*         coo = (coo_t*)malloc(sizeof(coo_t) * (n > 0 ? n : 1));
  or_index       k   = 0;
  for (or_index e = 0; e < num_elems; ++e)
    for (or_index p = res_offsets[e]; p < res_offsets[e + 1]; ++p)
      for (or_index q = state_offsets[e]; q < state_offsets[e + 1]; ++q)
      {
        coo[k].row = res_dofs[p];
        coo[k].col = state_dofs[q];
        coo[k].idx = k;
        ++k;
      }
  qsort(coo, n, sizeof(coo_t), coo_cmp);
  or_index* row_ptr = (or_index*)calloc(num_res + 1, sizeof(or_index));
  or_index* cols    = (or_index*)malloc(sizeof(or_index) * (n > 0 ? n : 1));
  or_index* jac_map = (or_index*)malloc(sizeof(or_index) * (n > 0 ? n : 1));
  or_index  csr_i   = -1;
  for (or_index i = 0; i < n; ++i)
  {
    const int is_new = i == 0 || coo[i].row != coo[i - 1].row || coo[i].col != coo[i - 1].col;
    if (is_new)
    {
      ++csr_i;
      cols[csr_i] = coo[i].col;
      ++row_ptr[coo[i].row + 1];
    }
    jac_map[coo[i].idx] = csr_i;
  }
  for (or_index r = 0; r < num_res; ++r) row_ptr[r + 1] += row_ptr[r];
  free(coo);
  *row_ptr_out = row_ptr;
  *col_ind_out = cols;
  *jac_off_out = jac_off;
  *jac_map_out = jac_map;
  *nnz_out     = csr_i + 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* makeBoundaryMap: femx/assembly/BoundaryMap.cpp:59-139.  Returns 1 when a
 * DOF is out of range, 2 when DOFs repeat, 3 when a constrained row lacks its
 * diagonal, 4 for a duplicate diagonal. */
int or_bc_build(or_index rows, const or_index* row_ptr, const or_index* col_ind,
                or_index num_bcs, const or_index* dofs, or_index* diag,
                or_index* col_offsets, or_index** col_entries_out, or_index** col_rows_out,
                or_index* bc_mask)
{
  or_index* bc_by_col = (or_index*)malloc(sizeof(or_index) * (rows > 0 ? rows : 1));
  for (or_index i = 0; i < rows; ++i) { bc_by_col[i] = -1; bc_mask[i] = 0; }
  for (or_index ib = 0; ib <= num_bcs; ++ib) col_offsets[ib] = 0;
  for (or_index ib = 0; ib < num_bcs; ++ib)
  {
    const or_index dof = dofs[ib];
    if (dof < 0 || dof >= rows) { free(bc_by_col); return 1; }
    if (bc_mask[dof] != 0) { free(bc_by_col); return 2; }
    bc_mask[dof]   = 1;
    bc_by_col[dof] = ib;
    diag[ib]       = -1;
  }
  for (or_index r = 0; r < rows; ++r)
    for (or_index k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    {
      const or_index ib = bc_by_col[col_ind[k]];
      if (ib >= 0)
      {
        ++col_offsets[ib + 1];
        if (r == col_ind[k])
        {
          if (diag[ib] >= 0) { free(bc_by_col); return 4; }
          diag[ib] = k;
        }
      }
    }
  for (or_index ib = 0; ib < num_bcs; ++ib)
  {
    if (diag[ib] < 0) { free(bc_by_col); return 3; }
    col_offsets[ib + 1] += col_offsets[ib];
  }
  const or_index nce  = col_offsets[num_bcs];
  or_index*      ent  = (or_index*)malloc(sizeof(or_index) * (nce > 0 ? nce : 1));
  or_index*      crow = (or_index*)malloc(sizeof(or_index) * (nce > 0 ? nce : 1));
  or_index*      next = (or_index*)malloc(sizeof(or_index) * (num_bcs + 1));
  memcpy(next, col_offsets, sizeof(or_index) * (num_bcs + 1));
  for (or_index r = 0; r < rows; ++r)
    for (or_index k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    {
      const or_index ib = bc_by_col[col_ind[k]];
      if (ib >= 0)
      {
        const or_index dst = next[ib]++;
        ent[dst]           = k;
        crow[dst]          = r;
      }
    }
  free(next);
  free(bc_by_col);
  *col_entries_out = ent;
  *col_rows_out    = crow;
  return 0;
}

/* replaceRowsRaw: femx/assembly/BoundaryMap.cpp:38-55. */
void or_replace_rows(const or_index* row_ptr, or_index num_bcs, const or_index* bc_rows,
                     const or_index* diag, double* vals, double diag_val)
{
  for (or_index ib = 0; ib < num_bcs; ++ib)
  {
    const or_index row = bc_rows[ib];
    for (or_index k = row_ptr[row]; k < row_ptr[row + 1]; ++k) vals[k] = 0.0;
    vals[diag[ib]] = diag_val;
  }
}

/* replaceRes: femx/assembly/BoundaryMap.cpp:180-197. */
void or_replace_res(or_index num_bcs, const or_index* bc_rows, const double* state,
                    const double* bc_vals, double* res)
{
  for (or_index ib = 0; ib < num_bcs; ++ib)
    res[bc_rows[ib]] = state[bc_rows[ib]] - bc_vals[ib];
}

/* zeroBoundary: femx/assembly/BoundaryMap.cpp:207-216. */
void or_zero_boundary(or_index num_bcs, const or_index* bc_rows, double* v)
{
  for (or_index ib = 0; ib < num_bcs; ++ib) v[bc_rows[ib]] = 0.0;
}

/* prepareForwardSolve: femx/assembly/BoundaryMap.cpp:218-247. */
void or_prepare_forward(const or_index* row_ptr, or_index num_bcs, const or_index* bc_rows,
                        const or_index* diag, const or_index* col_offsets,
                        const or_index* col_entries, const or_index* col_rows,
                        const or_index* bc_mask, double* vals, double* rhs,
                        const double* bc_vals)
{
  for (or_index ib = 0; ib < num_bcs; ++ib)
  {
    const double bc = bc_vals[ib];
    for (or_index i = col_offsets[ib]; i < col_offsets[ib + 1]; ++i)
    {
      const or_index row = col_rows[i];
      const or_index k   = col_entries[i];
      if (bc_mask[row] == 0) rhs[row] -= vals[k] * bc;
      vals[k] = 0.0;
    }
  }
  or_replace_rows(row_ptr, num_bcs, bc_rows, diag, vals, 1.0);
  for (or_index ib = 0; ib < num_bcs; ++ib) rhs[bc_rows[ib]] = bc_vals[ib];
}

/* ------------------------------------------------------------------------ */
/* MatrixHandler<HostCsrBackend>::matvec / matvecT:
 * femx/linalg/handler/MatrixHandler.cpp:157-198. */
void or_matvec(or_index rows, const or_index* row_ptr, const or_index* col_ind,
               const double* vals, const double* x, double* y, double alpha, double beta)
{
  for (or_index r = 0; r < rows; ++r)
  {
    double v = 0.0;
    for (or_index k = row_ptr[r]; k < row_ptr[r + 1]; ++k) v += vals[k] * x[col_ind[k]];
    y[r] = alpha * v + beta * y[r];
  }
}

void or_matvecT(or_index rows, or_index cols, const or_index* row_ptr, const or_index* col_ind,
                const double* vals, const double* x, double* y, double alpha, double beta)
{
  for (or_index c = 0; c < cols; ++c) y[c] *= beta;
  for (or_index r = 0; r < rows; ++r)
  {
    const double v = alpha * x[r];
    for (or_index k = row_ptr[r]; k < row_ptr[r + 1]; ++k) y[col_ind[k]] += vals[k] * v;
  }
}

/* ------------------------------------------------------------------------ */
/* Quadrature.  Orders 1 and 2 restate GaussQuadrature::triangle
 * (femx/fem/GaussQuadrature.hpp:175-204).  Order 5 is the 7-point degree-5
 * rule of Radon (1948) as tabulated by Stroud (1971, T2:5-1) and Dunavant
 * (1985): the centroid plus two 3-point orbits with barycentric a=(6-+sqrt15)/21,
 * weights (on the reference area 1/2) 9/80 and (155-+sqrt15)/2400. */
int or_quad_tri(int order, double* rs, double* w)
{
  if (order == 1)
  {
    rs[0] = 1.0 / 3.0; rs[1] = 1.0 / 3.0; w[0] = 0.5;
    return 1;
  }
  if (order == 2)
  {
    const double p[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}};
    for (int q = 0; q < 3; ++q) { rs[2 * q] = p[q][0]; rs[2 * q + 1] = p[q][1]; w[q] = 1.0 / 6.0; }
    return 3;
  }
  if (order == 5)
  {
    const double s15 = sqrt(15.0);
    const double a1 = (6.0 - s15) / 21.0, b1 = (9.0 + 2.0 * s15) / 21.0;
    const double a2 = (6.0 + s15) / 21.0, b2 = (9.0 - 2.0 * s15) / 21.0;
    const double w1 = (155.0 - s15) / 2400.0, w2 = (155.0 + s15) / 2400.0;
    const double p[7][2] = {{1.0 / 3.0, 1.0 / 3.0}, {a1, a1}, {b1, a1}, {a1, b1},
                            {a2, a2}, {b2, a2}, {a2, b2}};
    const double ww[7] = {9.0 / 80.0, w1, w1, w1, w2, w2, w2};
    for (int q = 0; q < 7; ++q) { rs[2 * q] = p[q][0]; rs[2 * q + 1] = p[q][1]; w[q] = ww[q]; }
    return 7;
  }
  return -1;
}

/* P2 Lagrange triangle (not in femx): phi_i = l_i(2 l_i - 1) at the
 * vertices, 4 l_i l_j at the edge midpoints (0,1), (1,2), (2,0), with
 * l0 = 1-r-s, l1 = r, l2 = s.  dNdr is row-major (shape, d/dr or d/ds) as in
 * FiniteElement::calcdNdr. */
void or_p2_shape(double r, double s, double* N, double* dNdr)
{
  const double l0 = 1.0 - r - s, l1 = r, l2 = s;
  N[0] = l0 * (2.0 * l0 - 1.0);
  N[1] = l1 * (2.0 * l1 - 1.0);
  N[2] = l2 * (2.0 * l2 - 1.0);
  N[3] = 4.0 * l0 * l1;
  N[4] = 4.0 * l1 * l2;
  N[5] = 4.0 * l2 * l0;
  dNdr[0]  = -(4.0 * l0 - 1.0); dNdr[1]  = -(4.0 * l0 - 1.0);
  dNdr[2]  = 4.0 * l1 - 1.0;    dNdr[3]  = 0.0;
  dNdr[4]  = 0.0;               dNdr[5]  = 4.0 * l2 - 1.0;
  dNdr[6]  = 4.0 * (l0 - l1);   dNdr[7]  = -4.0 * l1;
  dNdr[8]  = 4.0 * l2;          dNdr[9]  = 4.0 * l1;
  dNdr[10] = -4.0 * l2;         dNdr[11] = 4.0 * (l0 - l2);
}

/* ElementValues::calcPhysicalValues + invJacobian for an affine triangle
 * (femx/fem/ElementValues.cpp:138-181,183-215), with the P1 geometry map of
 * LagrangeTriangleP1::calcdNdr (femx/fem/elements/LagrangeTriangleP1.hpp:60-70).
 * Returns |det J| * wt and invJ (row-major). */
static double tri_geometry(const double* xe, const double* ye, double wt, double* invJ)
{
  static const double dNdr[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
  double J[4] = {0.0, 0.0, 0.0, 0.0};
  for (int in = 0; in < 3; ++in)
    for (int a = 0; a < 2; ++a)
    {
      const double xa = a == 0 ? xe[in] : ye[in];
      for (int b = 0; b < 2; ++b) J[a * 2 + b] += xa * dNdr[in][b];
    }
  const double det     = J[0] * J[3] - J[1] * J[2];
  const double inv_det = 1.0 / det;
  invJ[0] = J[3] * inv_det;
  invJ[1] = -J[1] * inv_det;
  invJ[2] = -J[2] * inv_det;
  invJ[3] = J[0] * inv_det;
  return fabs(det) * wt;
}

/* dNdx(i,a) = sum_b dNdr(i,b) invJ(b,a) (femx/fem/ElementValues.cpp:166-179). */
static void phys_grad(int ns, const double* dNdr, const double* invJ, double* dNdx)
{
  for (int i = 0; i < ns; ++i)
    for (int a = 0; a < 2; ++a)
    {
      double v = 0.0;
      for (int b = 0; b < 2; ++b) v += dNdr[i * 2 + b] * invJ[b * 2 + a];
      dNdx[i * 2 + a] = v;
    }
}

/* ------------------------------------------------------------------------ */
/* P1 Poisson element.  The stiffness part restates PoissonComponents::evalRow
 * (examples/poisson/PoissonComponents.hpp:35-64) on the order-2 rule that
 * PoissonForwardProblem uses (examples/poisson/PoissonForward.cpp:181-182).
 * Extension (not in femx): the row is scaled by the element diffusion kappa
 * and the consistent P1 load F_i = sum_q N_i (sum_c N_c f_c) JxW_q is
 * subtracted: Re = kappa*K u - F.  With kappa = 1 and f = 0 this is femx's
 * operator exactly. */
void or_p1_element(const double* xe, const double* ye, double kappa, const double* fe,
                   const double* ue, double* Ke, double* Re)
{
  double rs[6], w[3], dNdx[3][6], JxW[3], N[3][3];
  static const double dNdr[6] = {-1.0, -1.0, 1.0, 0.0, 0.0, 1.0};
  or_quad_tri(2, rs, w);
  for (int q = 0; q < 3; ++q)
  {
    double invJ[4];
    JxW[q] = tri_geometry(xe, ye, w[q], invJ);
    phys_grad(3, dNdr, invJ, dNdx[q]);
    N[q][0] = 1.0 - rs[2 * q] - rs[2 * q + 1];
    N[q][1] = rs[2 * q];
    N[q][2] = rs[2 * q + 1];
  }
  for (int row = 0; row < 3; ++row)
  {
    double jac[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 3; ++q)
      for (int col = 0; col < 3; ++col)
      {
        double gd = 0.0;
        for (int d = 0; d < 2; ++d) gd += dNdx[q][row * 2 + d] * dNdx[q][col * 2 + d];
        jac[col] += gd * JxW[q];
      }
    double res = 0.0;
    for (int col = 0; col < 3; ++col)
    {
      jac[col] = kappa * jac[col];
      res += jac[col] * ue[col];
      Ke[row * 3 + col] = jac[col];
    }
    double F = 0.0;
    for (int q = 0; q < 3; ++q)
    {
      double fq = 0.0;
      for (int c = 0; c < 3; ++c) fq += N[q][c] * fe[c];
      F += N[q][row] * fq * JxW[q];
    }
    Re[row] = res - F;
  }
}

static void p1_gather(const double* x, const double* y, const or_index* c, const double* f,
                      const double* u, double* xe, double* ye, double* fe, double* ue)
{
  for (int i = 0; i < 3; ++i)
  {
    xe[i] = x[c[i]];
    ye[i] = y[c[i]];
    fe[i] = f ? f[c[i]] : 0.0;
    ue[i] = u ? u[c[i]] : 0.0;
  }
}

/* assembly::assemble host driver (femx/assembly/Assembly.hpp:153-226):
 * outputs replaced, elements visited in ascending order, res[resDof] += and
 * vals[jacIndex] += per element. */
void or_p1_assemble(or_index num_nodes, const double* x, const double* y, or_index num_elems,
                    const or_index* conn, const double* kappa, const double* f,
                    const double* u, const or_index* jac_map, or_index nnz, double* vals,
                    double* res)
{
  memset(res, 0, sizeof(double) * num_nodes);
  memset(vals, 0, sizeof(double) * nnz);
  for (or_index e = 0; e < num_elems; ++e)
  {
    double xe[3], ye[3], fe[3], ue[3], Ke[9], Re[3];
    p1_gather(x, y, conn + 3 * e, f, u, xe, ye, fe, ue);
    or_p1_element(xe, ye, kappa ? kappa[e] : 1.0, fe, ue, Ke, Re);
    for (int r = 0; r < 3; ++r) res[conn[3 * e + r]] += Re[r];
    for (int i = 0; i < 9; ++i) vals[jac_map[9 * e + i]] += Ke[i];
  }
}

/* OpenMP shape of femx's CPU assembly (femx/model/ns/NavierStokesModel.cpp:
 * 269-318): parallel for over elements with thread-private scratch and an
 * atomic scatter as in assembly::addElem (femx/assembly/Assembly.cpp:31-34). */
void or_p1_assemble_omp(or_index num_nodes, const double* x, const double* y,
                        or_index num_elems, const or_index* conn, const double* kappa,
                        const double* f, const double* u, const or_index* jac_map,
                        or_index nnz, double* vals, double* res)
{
#pragma omp parallel for schedule(static)
  for (or_index i = 0; i < num_nodes; ++i) res[i] = 0.0;
#pragma omp parallel for schedule(static)
  for (or_index k = 0; k < nnz; ++k) vals[k] = 0.0;
#pragma omp parallel
  {
    double xe[3], ye[3], fe[3], ue[3], Ke[9], Re[3];
#pragma omp for schedule(static)
    for (or_index e = 0; e < num_elems; ++e)
    {
      p1_gather(x, y, conn + 3 * e, f, u, xe, ye, fe, ue);
      or_p1_element(xe, ye, kappa ? kappa[e] : 1.0, fe, ue, Ke, Re);
      for (int r = 0; r < 3; ++r)
      {
#pragma omp atomic update
        res[conn[3 * e + r]] += Re[r];
      }
      for (int i = 0; i < 9; ++i)
      {
#pragma omp atomic update
        vals[jac_map[9 * e + i]] += Ke[i];
      }
    }
  }
}

/* P1 log-kappa VJP (not in femx; it fills the Residual::applyParamJacT
 * contract of femx/state/Residual.hpp:55-59 and the transposed state
 * Jacobian apply of the adjoint solve).  With m_e = log kappa_e,
 * dR_e/dm_e = K^e u_e, so g_e = lam~_e^T K^e u_e.  J is the row-replaced
 * Jacobian (femx/assembly/BoundaryMap.hpp:198-208): constrained rows are e_r,
 * so lam is zeroed on constrained rows before the physics transpose apply, as
 * ConstrainedTimeResidual::applyJacT does (femx/assembly/
 * ConstrainedTimeResidual.cpp:379-405), and J^T lam gains lam_r on those rows.
 * bc_mask may be NULL (no constraints). */
void or_p1_vjp(or_index num_nodes, const double* x, const double* y, or_index num_elems,
               const or_index* conn, const double* m, const double* u, const double* lam,
               const or_index* bc_mask, double* jt_lam, double* g)
{
  memset(jt_lam, 0, sizeof(double) * num_nodes);
  for (or_index e = 0; e < num_elems; ++e)
  {
    double xe[3], ye[3], fe[3], ue[3], Ke[9], Re[3], le[3];
    p1_gather(x, y, conn + 3 * e, NULL, u, xe, ye, fe, ue);
    or_p1_element(xe, ye, exp(m[e]), fe, ue, Ke, Re);
    for (int r = 0; r < 3; ++r)
    {
      const or_index d = conn[3 * e + r];
      le[r]            = (bc_mask && bc_mask[d]) ? 0.0 : lam[d];
    }
    double ge = 0.0;
    for (int r = 0; r < 3; ++r) ge += le[r] * Re[r];
    g[e] = ge;
    for (int c = 0; c < 3; ++c)
    {
      double v = 0.0;
      for (int r = 0; r < 3; ++r) v += Ke[r * 3 + c] * le[r];
      jt_lam[conn[3 * e + c]] += v;
    }
  }
  if (bc_mask)
    for (or_index i = 0; i < num_nodes; ++i)
      if (bc_mask[i]) jt_lam[i] += lam[i];
}

void or_p1_vjp_omp(or_index num_nodes, const double* x, const double* y, or_index num_elems,
                   const or_index* conn, const double* m, const double* u, const double* lam,
                   const or_index* bc_mask, double* jt_lam, double* g)
{
#pragma omp parallel for schedule(static)
  for (or_index i = 0; i < num_nodes; ++i) jt_lam[i] = 0.0;
#pragma omp parallel
  {
    double xe[3], ye[3], fe[3], ue[3], Ke[9], Re[3], le[3];
#pragma omp for schedule(static)
    for (or_index e = 0; e < num_elems; ++e)
    {
      p1_gather(x, y, conn + 3 * e, NULL, u, xe, ye, fe, ue);
      or_p1_element(xe, ye, exp(m[e]), fe, ue, Ke, Re);
      for (int r = 0; r < 3; ++r)
      {
        const or_index d = conn[3 * e + r];
        le[r]            = (bc_mask && bc_mask[d]) ? 0.0 : lam[d];
      }
      double ge = 0.0;
      for (int r = 0; r < 3; ++r) ge += le[r] * Re[r];
      g[e] = ge;
      for (int c = 0; c < 3; ++c)
      {
        double v = 0.0;
        for (int r = 0; r < 3; ++r) v += Ke[r * 3 + c] * le[r];
#pragma omp atomic update
        jt_lam[conn[3 * e + c]] += v;
      }
    }
  }
  if (bc_mask)
  {
#pragma omp parallel for schedule(static)
    for (or_index i = 0; i < num_nodes; ++i)
      if (bc_mask[i]) jt_lam[i] += lam[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Taylor-Hood P2/P1 steady incompressible Navier-Stokes element (not in
 * femx; femx's NS is GLS P1/P1 transient).  Galerkin residual with
 * nu = 1/Re, on the 7-point rule:
 *   R_(a,c) = int nu grad u_c . grad phi_a + (u . grad u_c) phi_a - p d_c phi_a
 *   R_b     = -int psi_b div u
 * and its exact Newton Jacobian
 *   J_(a,c),(b,d) = int nu delta_cd grad phi_a . grad phi_b
 *                   + phi_a (phi_b d_d u_c + delta_cd u . grad phi_b)
 *   J_(a,c),b'    = -int psi_b' d_c phi_a,   J_b',(b,d) = -int psi_b' d_d phi_b,
 *   J_b',b''      = 0 (kept in the pattern, as femx keeps dense element
 *   blocks: femx/assembly/AssemblyMap.cpp:108-121).
 * Local order: 2a+c for velocity (a = P2 node, c = component), 12+b for
 * pressure.  Geometry from the three vertices (affine map). */
void or_th_element(const double* xv, const double* yv, double nu, const double* se,
                   double* Ke, double* Re)
{
  double rs[14], w[7];
  or_quad_tri(5, rs, w);
  memset(Ke, 0, sizeof(double) * 225);
  memset(Re, 0, sizeof(double) * 15);
  for (int q = 0; q < 7; ++q)
  {
    const double r = rs[2 * q], s = rs[2 * q + 1];
    double       invJ[4], N[6], dNdr[12], dN[12];
    const double wq = tri_geometry(xv, yv, w[q], invJ);
    or_p2_shape(r, s, N, dNdr);
    phys_grad(6, dNdr, invJ, dN);
    const double psi[3] = {1.0 - r - s, r, s};
    double       u[2] = {0.0, 0.0}, du[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, p = 0.0;
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < 2; ++c)
      {
        u[c] += N[a] * se[2 * a + c];
        du[c][0] += dN[2 * a] * se[2 * a + c];
        du[c][1] += dN[2 * a + 1] * se[2 * a + c];
      }
    for (int b = 0; b < 3; ++b) p += psi[b] * se[12 + b];
    const double div = du[0][0] + du[1][1];
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < 2; ++c)
      {
        const double visc = nu * (dN[2 * a] * du[c][0] + dN[2 * a + 1] * du[c][1]);
        const double conv = N[a] * (u[0] * du[c][0] + u[1] * du[c][1]);
        Re[2 * a + c] += (visc + conv - p * dN[2 * a + c]) * wq;
      }
    for (int b = 0; b < 3; ++b) Re[12 + b] += (-psi[b] * div) * wq;
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < 2; ++c)
      {
        double* row = Ke + (2 * a + c) * 15;
        for (int b = 0; b < 6; ++b)
        {
          const double lap  = nu * (dN[2 * a] * dN[2 * b] + dN[2 * a + 1] * dN[2 * b + 1]);
          const double adv  = N[a] * (u[0] * dN[2 * b] + u[1] * dN[2 * b + 1]);
          for (int d = 0; d < 2; ++d)
          {
            double v = N[a] * N[b] * du[c][d];
            if (c == d) v += lap + adv;
            row[2 * b + d] += v * wq;
          }
        }
        for (int b = 0; b < 3; ++b) row[12 + b] += (-psi[b] * dN[2 * a + c]) * wq;
      }
    for (int bp = 0; bp < 3; ++bp)
    {
      double* row = Ke + (12 + bp) * 15;
      for (int b = 0; b < 6; ++b)
        for (int d = 0; d < 2; ++d) row[2 * b + d] += (-psi[bp] * dN[2 * b + d]) * wq;
    }
  }
}

static void th_gather(const double* xv, const double* yv, const or_index* v3,
                      const or_index* d15, const double* state, double* xe, double* ye,
                      double* se)
{
  for (int i = 0; i < 3; ++i) { xe[i] = xv[v3[i]]; ye[i] = yv[v3[i]]; }
  for (int i = 0; i < 15; ++i) se[i] = state[d15[i]];
}

/* assembly::assemble host driver order (femx/assembly/Assembly.hpp:184-225). */
void or_th_assemble(const double* xv, const double* yv, or_index num_elems,
                    const or_index* vconn3, const or_index* dofs15, or_index ndof,
                    double nu, const double* nu_e, const double* state, const or_index* jac_map,
                    or_index nnz, double* vals, double* res)
{
  memset(res, 0, sizeof(double) * ndof);
  memset(vals, 0, sizeof(double) * nnz);
  for (or_index e = 0; e < num_elems; ++e)
  {
    double xe[3], ye[3], se[15], Ke[225], Re[15];
    th_gather(xv, yv, vconn3 + 3 * e, dofs15 + 15 * e, state, xe, ye, se);
    or_th_element(xe, ye, nu_e ? nu_e[e] : nu, se, Ke, Re);
    for (int r = 0; r < 15; ++r) res[dofs15[15 * e + r]] += Re[r];
    for (int i = 0; i < 225; ++i) vals[jac_map[225 * (int64_t)e + i]] += Ke[i];
  }
}

void or_th_assemble_omp(const double* xv, const double* yv, or_index num_elems,
                        const or_index* vconn3, const or_index* dofs15, or_index ndof,
                        double nu, const double* nu_e, const double* state,
                        const or_index* jac_map, or_index nnz, double* vals, double* res)
{
#pragma omp parallel for schedule(static)
  for (or_index i = 0; i < ndof; ++i) res[i] = 0.0;
#pragma omp parallel for schedule(static)
  for (or_index k = 0; k < nnz; ++k) vals[k] = 0.0;
#pragma omp parallel
  {
    double xe[3], ye[3], se[15], Ke[225], Re[15];
#pragma omp for schedule(static)
    for (or_index e = 0; e < num_elems; ++e)
    {
      th_gather(xv, yv, vconn3 + 3 * e, dofs15 + 15 * e, state, xe, ye, se);
      or_th_element(xe, ye, nu_e ? nu_e[e] : nu, se, Ke, Re);
      for (int r = 0; r < 15; ++r)
      {
#pragma omp atomic update
        res[dofs15[15 * e + r]] += Re[r];
      }
      for (int i = 0; i < 225; ++i)
      {
#pragma omp atomic update
        vals[jac_map[225 * (int64_t)e + i]] += Ke[i];
      }
    }
  }
}

/* Viscous part of the residual, int grad u_c . grad phi_a (dR/dnu). */
static void th_visc_res(const double* xv, const double* yv, const double* se, double* Rv)
{
  double rs[14], w[7];
  or_quad_tri(5, rs, w);
  memset(Rv, 0, sizeof(double) * 15);
  for (int q = 0; q < 7; ++q)
  {
    double       invJ[4], N[6], dNdr[12], dN[12];
    const double wq = tri_geometry(xv, yv, w[q], invJ);
    or_p2_shape(rs[2 * q], rs[2 * q + 1], N, dNdr);
    phys_grad(6, dNdr, invJ, dN);
    double du[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < 2; ++c)
      {
        du[c][0] += dN[2 * a] * se[2 * a + c];
        du[c][1] += dN[2 * a + 1] * se[2 * a + c];
      }
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < 2; ++c)
        Rv[2 * a + c] += (dN[2 * a] * du[c][0] + dN[2 * a + 1] * du[c][1]) * wq;
  }
}

/* Navier-Stokes adjoint VJP (not in femx): J^T lam~ + mask*lam over the
 * row-replaced Newton Jacobian (same boundary convention as or_p1_vjp), and
 * g[e] = lam~_e^T dR_e/dnu_e, the viscous part of the element residual
 * (nu_e[e] per element when nu_e is non-null, else nu). */
void or_th_vjp(const double* xv, const double* yv, or_index num_elems,
               const or_index* vconn3, const or_index* dofs15, or_index ndof, double nu,
               const double* nu_e, const double* state, const double* lam,
               const or_index* bc_mask, double* jt_lam, double* g)
{
  memset(jt_lam, 0, sizeof(double) * ndof);
  for (or_index e = 0; e < num_elems; ++e)
  {
    double xe[3], ye[3], se[15], Ke[225], Re[15], Rv[15], le[15];
    th_gather(xv, yv, vconn3 + 3 * e, dofs15 + 15 * e, state, xe, ye, se);
    or_th_element(xe, ye, nu_e ? nu_e[e] : nu, se, Ke, Re);
    th_visc_res(xe, ye, se, Rv);
    for (int r = 0; r < 15; ++r)
    {
      const or_index d = dofs15[15 * e + r];
      le[r]            = (bc_mask && bc_mask[d]) ? 0.0 : lam[d];
    }
    double ge = 0.0;
    for (int r = 0; r < 15; ++r) ge += le[r] * Rv[r];
    g[e] = ge;
    for (int c = 0; c < 15; ++c)
    {
      double v = 0.0;
      for (int r = 0; r < 15; ++r) v += Ke[r * 15 + c] * le[r];
      jt_lam[dofs15[15 * e + c]] += v;
    }
  }
  if (bc_mask)
    for (or_index i = 0; i < ndof; ++i)
      if (bc_mask[i]) jt_lam[i] += lam[i];
}

/* ------------------------------------------------------------------------ */
/* GLS-stabilised P1/P1 Navier-Stokes time step on triangles: the femx
 * NavierOperator (femx/model/ns/Components.hpp:230-577) restated for dim 2,
 * three nodes and the order-2 rule NavierStokesModel uses
 * (femx/model/ns/NavierStokesModel.cpp:33), with element data as
 * makeNavierData stores it (femx/model/ns/Components.cpp:19-62).  Local DOF
 * order: velocity interleaved by node (vdof = 2*in + c), then pressure
 * (pdof = 6 + in) (Components.hpp:247-256).  Every expression keeps femx's
 * operation order so that, built without contraction, the results equal the
 * compiled reference's bit for bit. */
typedef struct
{
  double u[2], adv[2], grad[2][2], adv_grad[2], tau;
} gls_qp;

/* elemLength (Components.hpp:286-327). */
static double gls_elem_length(const double* dNdx, const gls_qp* q)
{
  double speed2 = 0.0;
  for (int d = 0; d < 2; ++d) speed2 += q->u[d] * q->u[d];
  const double speed = sqrt(speed2);
  double       dir[2];
  if (speed > 1.0e-10)
    for (int d = 0; d < 2; ++d) dir[d] = q->u[d] / speed;
  else
  {
    const double val = 1.0 / sqrt(2.0);
    for (int d = 0; d < 2; ++d) dir[d] = val;
  }
  double sum = 0.0;
  for (int in = 0; in < 3; ++in)
  {
    double g = 0.0;
    for (int d = 0; d < 2; ++d) g += dir[d] * dNdx[in * 2 + d];
    sum += g >= 0.0 ? g : -g;
  }
  return sum > 1.0e-14 ? 2.0 / sum : 0.0;
}

/* evalQp (Components.hpp:329-396): cur = history lag 0, prev = lag 1. */
static void gls_eval_qp(const double* N, const double* dNdx, int step, int num_hist,
                        const double* cur, const double* prev, double rho, double mu, double dt,
                        gls_qp* q)
{
  memset(q, 0, sizeof(*q));
  for (int in = 0; in < 3; ++in)
    for (int c = 0; c < 2; ++c)
    {
      const int    id  = 2 * in + c;
      const double val = cur[id];
      q->u[c] += N[in] * val;
      for (int d = 0; d < 2; ++d) q->grad[c][d] += dNdx[in * 2 + d] * val;
      double adv = val;
      if (step > 0 && num_hist > 1) adv = 1.5 * val - 0.5 * prev[id];
      q->adv[c] += N[in] * adv;
    }
  double speed2 = 0.0;
  for (int c = 0; c < 2; ++c)
  {
    speed2 += q->u[c] * q->u[c];
    for (int d = 0; d < 2; ++d) q->adv_grad[c] += q->grad[c][d] * q->adv[d];
  }
  const double speed = sqrt(speed2);
  const double h     = gls_elem_length(dNdx, q);
  const double nu    = mu / rho;
  const double time  = (2.0 / dt) * (2.0 / dt);
  double       flow = 0.0, diff = 0.0;
  if (h > 0.0)
  {
    flow = (2.0 * speed / h) * (2.0 * speed / h);
    diff = (4.0 * nu / (h * h)) * (4.0 * nu / (h * h));
  }
  q->tau = 1.0 / sqrt(time + flow + diff);
}

/* Element matrix (row-major 9x9) and residual of NavierOperator::evalRow with
 * respect to the next state (Components.hpp:412-428,473-577):
 * Ke = dR/d(next), Re = -F + Ke * nxt_e.  hist_e is lag-major (num_hist x 9). */
void or_gls_element(const double* xe, const double* ye, double rho, double mu, double dt,
                    int step, int num_hist, const double* hist_e, const double* nxt_e,
                    double* Ke, double* Re)
{
  double rs[6], w[3], invJ[4], jxw[3], N[3][3], dNdx[3][6];
  static const double dNdr[6] = {-1.0, -1.0, 1.0, 0.0, 0.0, 1.0};
  or_quad_tri(2, rs, w);
  for (int q = 0; q < 3; ++q)
  {
    jxw[q] = tri_geometry(xe, ye, w[q], invJ);
    phys_grad(3, dNdr, invJ, dNdx[q]);
    N[q][0] = 1.0 - rs[2 * q] - rs[2 * q + 1];
    N[q][1] = rs[2 * q];
    N[q][2] = rs[2 * q + 1];
  }
  const double* cur  = hist_e;
  const double* prev = num_hist > 1 ? hist_e + 9 : hist_e;
  for (int row = 0; row < 9; ++row)
  {
    double*      K      = Ke + 9 * row;
    const int    vel    = row < 6;
    const int    i      = vel ? row / 2 : row - 6;
    const int    c      = vel ? row - i * 2 : 0;
    double       F      = 0.0;
    for (int col = 0; col < 9; ++col) K[col] = 0.0;
    for (int iq = 0; iq < 3; ++iq)
    {
      gls_qp qp;
      gls_eval_qp(N[iq], dNdx[iq], step, num_hist, cur, prev, rho, mu, dt, &qp);
      const double* g   = dNdx[iq];
      const double  Jw  = jxw[iq];
      const double  Ni  = N[iq][i];
      const double  dvi = 0.0 + g[2 * i] * qp.adv[0] + g[2 * i + 1] * qp.adv[1];
      if (vel)
      {
        const double grad_u = 0.0 + g[2 * i] * qp.grad[c][0] + g[2 * i + 1] * qp.grad[c][1];
        F += (rho / dt * Ni * qp.u[c] - 0.5 * rho * Ni * qp.adv_grad[c] - 0.5 * mu * grad_u
              + qp.tau * rho / dt * dvi * qp.u[c] - 0.5 * qp.tau * rho * dvi * qp.adv_grad[c])
             * Jw;
        for (int j = 0; j < 3; ++j)
        {
          const double Nj  = N[iq][j];
          const double dvj = 0.0 + g[2 * j] * qp.adv[0] + g[2 * j + 1] * qp.adv[1];
          const double gd  = 0.0 + g[2 * i] * g[2 * j] + g[2 * i + 1] * g[2 * j + 1];
          K[2 * j + c] += (rho / dt * Ni * Nj + 0.5 * rho * Ni * dvj + 0.5 * mu * gd
                           + qp.tau * rho / dt * dvi * Nj + 0.5 * qp.tau * rho * dvi * dvj)
                          * Jw;
          K[6 + j] += (-g[2 * i + c] * Nj + qp.tau * dvi * g[2 * j + c]) * Jw;
        }
      }
      else
      {
        double div_u = 0.0, div_adv = 0.0;
        for (int d = 0; d < 2; ++d)
        {
          div_u += g[2 * i + d] * qp.u[d];
          div_adv += g[2 * i + d] * qp.adv_grad[d];
        }
        F += (qp.tau / dt * div_u - 0.5 * qp.tau * div_adv) * Jw;
        for (int j = 0; j < 3; ++j)
        {
          const double Nj  = N[iq][j];
          const double dvj = 0.0 + g[2 * j] * qp.adv[0] + g[2 * j + 1] * qp.adv[1];
          for (int d = 0; d < 2; ++d)
          {
            const double gi = g[2 * i + d];
            K[2 * j + d] += (Ni * g[2 * j + d] + qp.tau / dt * gi * Nj + 0.5 * qp.tau * gi * dvj)
                            * Jw;
          }
          const double gd = 0.0 + g[2 * i] * g[2 * j] + g[2 * i + 1] * g[2 * j + 1];
          K[6 + j] += qp.tau / rho * gd * Jw;
        }
      }
    }
    double out = -F;
    for (int col = 0; col < 9; ++col) out += K[col] * nxt_e[col];
    Re[row] = out;
  }
}

static void gls_gather(const double* x, const double* y, const or_index* c,
                       const or_index* d9, or_index num_states, int num_hist,
                       const double* hist, const double* nxt, double* xe, double* ye,
                       double* he, double* ne)
{
  for (int k = 0; k < 3; ++k)
  {
    xe[k] = x[c[k]];
    ye[k] = y[c[k]];
  }
  for (int lag = 0; lag < num_hist; ++lag)
    for (int col = 0; col < 9; ++col)
      he[lag * 9 + col] = hist[(size_t)lag * num_states + d9[col]];
  for (int col = 0; col < 9; ++col) ne[col] = nxt[d9[col]];
}

/* Host time assemble (femx/assembly/Assembly.hpp:235-297) of the GLS
 * operator: outputs replaced, elements in ascending order, res[resDof] +=
 * and vals[jacIndex] += per element.  num_hist <= 8.  vals may be NULL
 * (residual only, as assembleResidual). */
void or_gls_assemble(or_index num_states, const double* x, const double* y, or_index num_elems,
                     const or_index* conn, const or_index* dofs9, double rho, double mu,
                     double dt, int step, int num_hist, const double* hist, const double* nxt,
                     const or_index* jac_map, or_index nnz, double* vals, double* res)
{
  memset(res, 0, sizeof(double) * num_states);
  if (vals) memset(vals, 0, sizeof(double) * nnz);
  const int nh = num_hist < 2 ? num_hist : 2; /* only lags 0 and 1 are read */
  for (or_index e = 0; e < num_elems; ++e)
  {
    double xe[3], ye[3], he[18], ne[9], Ke[81], Re[9];
    gls_gather(x, y, conn + 3 * e, dofs9 + 9 * e, num_states, nh, hist, nxt, xe, ye, he, ne);
    or_gls_element(xe, ye, rho, mu, dt, step, num_hist, he, ne, Ke, Re);
    for (int r = 0; r < 9; ++r) res[dofs9[9 * e + r]] += Re[r];
    if (vals)
      for (int i = 0; i < 81; ++i) vals[jac_map[81 * e + i]] += Ke[i];
  }
}

/* The same with femx's OpenMP shape (NavierStokesModel.cpp:269-318): the CPU
 * baseline.  Atomic scatter, so the sums are not in element order. */
void or_gls_assemble_omp(or_index num_states, const double* x, const double* y,
                         or_index num_elems, const or_index* conn, const or_index* dofs9,
                         double rho, double mu, double dt, int step, int num_hist,
                         const double* hist, const double* nxt, const or_index* jac_map,
                         or_index nnz, double* vals, double* res)
{
  const int nh = num_hist < 2 ? num_hist : 2;
#pragma omp parallel for schedule(static)
  for (or_index i = 0; i < num_states; ++i) res[i] = 0.0;
#pragma omp parallel for schedule(static)
  for (or_index k = 0; k < nnz; ++k) vals[k] = 0.0;
#pragma omp parallel
  {
    double xe[3], ye[3], he[18], ne[9], Ke[81], Re[9];
#pragma omp for schedule(static)
    for (or_index e = 0; e < num_elems; ++e)
    {
      gls_gather(x, y, conn + 3 * e, dofs9 + 9 * e, num_states, nh, hist, nxt, xe, ye, he, ne);
      or_gls_element(xe, ye, rho, mu, dt, step, num_hist, he, ne, Ke, Re);
      for (int r = 0; r < 9; ++r)
      {
#pragma omp atomic update
        res[dofs9[9 * e + r]] += Re[r];
      }
      for (int i = 0; i < 81; ++i)
      {
#pragma omp atomic update
        vals[jac_map[81 * e + i]] += Ke[i];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* History derivatives of the GLS residual.  femx obtains them with Enzyme
 * (femx/model/ns/Components.cpp:65-121 histVjp over evalNavierAdj,
 * Components.hpp:604-643; device Components.cu:175-238), which is not
 * available here.  Two independent restatements:
 *   or_gls_element_jvp   forward mode: the residual code of or_gls_element
 *                        re-evaluated in dual numbers (value, tangent);
 *   or_gls_element_hvjp  reverse mode written out by hand from the
 *                        residual's closed form (see below).
 * They are pinned against each other by the adjoint identity
 * dir^T (dR/dh)^T adj = adj^T (dR/dh) dir, and the forward mode against
 * central differences of the residual. */
typedef struct
{
  double v, d;
} dual;

static dual dc(double v) { dual r = {v, 0.0}; return r; }
static dual dadd(dual a, dual b) { dual r = {a.v + b.v, a.d + b.d}; return r; }
static dual dsub(dual a, dual b) { dual r = {a.v - b.v, a.d - b.d}; return r; }
static dual dmul(dual a, dual b) { dual r = {a.v * b.v, a.d * b.v + a.v * b.d}; return r; }
static dual ddiv(dual a, dual b)
{
  dual r = {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
  return r;
}
static dual dsqrt(dual a)
{
  const double s = sqrt(a.v);
  dual         r = {s, s > 0.0 ? a.d / (2.0 * s) : 0.0};
  return r;
}
static dual dabs(dual a) { return a.v >= 0.0 ? a : dsub(dc(0.0), a); }
static dual dscale(double s, dual a) { dual r = {s * a.v, s * a.d}; return r; }

/* Residual rows of one element (NavierOperator::evalResidualRow,
 * Components.hpp:448-455,473-577) in dual numbers: hist_e and dhist_e are
 * lag-major (num_hist x 9), the tangent acting on the history only. */
void or_gls_element_jvp(const double* xe, const double* ye, double rho, double mu, double dt,
                        int step, int num_hist, const double* hist_e, const double* dhist_e,
                        const double* nxt_e, double* Re, double* dRe)
{
  double rs[6], w[3], invJ[4], jxw[3], N[3][3], g[3][6];
  static const double dNdr[6] = {-1.0, -1.0, 1.0, 0.0, 0.0, 1.0};
  or_quad_tri(2, rs, w);
  for (int q = 0; q < 3; ++q)
  {
    jxw[q] = tri_geometry(xe, ye, w[q], invJ);
    phys_grad(3, dNdr, invJ, g[q]);
    N[q][0] = 1.0 - rs[2 * q] - rs[2 * q + 1];
    N[q][1] = rs[2 * q];
    N[q][2] = rs[2 * q + 1];
  }
  const int extrap = step > 0 && num_hist > 1;
  dual      cur[6], adv[6];
  for (int k = 0; k < 6; ++k)
  {
    cur[k].v = hist_e[k];
    cur[k].d = dhist_e[k];
    adv[k]   = cur[k];
    if (extrap)
    {
      dual prev = {hist_e[9 + k], dhist_e[9 + k]};
      adv[k]    = dsub(dscale(1.5, cur[k]), dscale(0.5, prev));
    }
  }
  const double nu = mu / rho, time = (2.0 / dt) * (2.0 / dt);
  for (int row = 0; row < 9; ++row)
  {
    const int vel = row < 6;
    const int i   = vel ? row / 2 : row - 6;
    const int c   = vel ? row - 2 * i : 0;
    dual      F = dc(0.0), K[9];
    for (int col = 0; col < 9; ++col) K[col] = dc(0.0);
    for (int iq = 0; iq < 3; ++iq)
    {
      const double* gq = g[iq];
      dual          u[2] = {dc(0.0), dc(0.0)}, av[2] = {dc(0.0), dc(0.0)}, G[2][2], ag[2];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) G[a][b] = dc(0.0);
      for (int in = 0; in < 3; ++in)
        for (int cc = 0; cc < 2; ++cc)
        {
          u[cc]  = dadd(u[cc], dscale(N[iq][in], cur[2 * in + cc]));
          av[cc] = dadd(av[cc], dscale(N[iq][in], adv[2 * in + cc]));
          for (int d = 0; d < 2; ++d)
            G[cc][d] = dadd(G[cc][d], dscale(gq[2 * in + d], cur[2 * in + cc]));
        }
      for (int cc = 0; cc < 2; ++cc)
        ag[cc] = dadd(dmul(G[cc][0], av[0]), dmul(G[cc][1], av[1]));
      const dual speed = dsqrt(dadd(dmul(u[0], u[0]), dmul(u[1], u[1])));
      dual       dir[2];
      if (speed.v > 1.0e-10)
      {
        dir[0] = ddiv(u[0], speed);
        dir[1] = ddiv(u[1], speed);
      }
      else
        dir[0] = dir[1] = dc(1.0 / sqrt(2.0));
      dual sum = dc(0.0);
      for (int in = 0; in < 3; ++in)
        sum = dadd(sum, dabs(dadd(dscale(gq[2 * in], dir[0]), dscale(gq[2 * in + 1], dir[1]))));
      dual flow = dc(0.0), diff = dc(0.0);
      if (sum.v > 1.0e-14)
      {
        const dual h  = ddiv(dc(2.0), sum);
        const dual f  = ddiv(dscale(2.0, speed), h);
        const dual vv = ddiv(dc(4.0 * nu), dmul(h, h));
        flow          = dmul(f, f);
        diff          = dmul(vv, vv);
      }
      const dual   tau = ddiv(dc(1.0), dsqrt(dadd(dadd(dc(time), flow), diff)));
      const double Jw  = jxw[iq], Ni = N[iq][i];
      const dual   dvi = dadd(dscale(gq[2 * i], av[0]), dscale(gq[2 * i + 1], av[1]));
      if (vel)
      {
        const dual gu = dadd(dscale(gq[2 * i], G[c][0]), dscale(gq[2 * i + 1], G[c][1]));
        dual t = dscale(rho / dt * Ni, u[c]);
        t      = dsub(t, dscale(0.5 * rho * Ni, ag[c]));
        t      = dsub(t, dscale(0.5 * mu, gu));
        t      = dadd(t, dmul(dscale(rho / dt, tau), dmul(dvi, u[c])));
        t      = dsub(t, dmul(dscale(0.5 * rho, tau), dmul(dvi, ag[c])));
        F      = dadd(F, dscale(Jw, t));
        for (int j = 0; j < 3; ++j)
        {
          const double Nj  = N[iq][j];
          const dual   dvj = dadd(dscale(gq[2 * j], av[0]), dscale(gq[2 * j + 1], av[1]));
          const double gd  = gq[2 * i] * gq[2 * j] + gq[2 * i + 1] * gq[2 * j + 1];
          dual k = dc(rho / dt * Ni * Nj + 0.5 * mu * gd);
          k      = dadd(k, dscale(0.5 * rho * Ni, dvj));
          k      = dadd(k, dmul(dscale(rho / dt * Nj, tau), dvi));
          k      = dadd(k, dmul(dscale(0.5 * rho, tau), dmul(dvi, dvj)));
          K[2 * j + c] = dadd(K[2 * j + c], dscale(Jw, k));
          const dual kp = dadd(dc(-gq[2 * i + c] * Nj), dmul(tau, dscale(gq[2 * j + c], dvi)));
          K[6 + j]      = dadd(K[6 + j], dscale(Jw, kp));
        }
      }
      else
      {
        dual div_u = dc(0.0), div_adv = dc(0.0);
        for (int d = 0; d < 2; ++d)
        {
          div_u   = dadd(div_u, dscale(gq[2 * i + d], u[d]));
          div_adv = dadd(div_adv, dscale(gq[2 * i + d], ag[d]));
        }
        F = dadd(F, dscale(Jw, dsub(dmul(dscale(1.0 / dt, tau), div_u),
                                    dmul(dscale(0.5, tau), div_adv))));
        for (int j = 0; j < 3; ++j)
        {
          const double Nj  = N[iq][j];
          const dual   dvj = dadd(dscale(gq[2 * j], av[0]), dscale(gq[2 * j + 1], av[1]));
          for (int d = 0; d < 2; ++d)
          {
            const double gi = gq[2 * i + d];
            dual k = dc(Ni * gq[2 * j + d]);
            k      = dadd(k, dscale(gi * Nj / dt, tau));
            k      = dadd(k, dmul(dscale(0.5 * gi, tau), dvj));
            K[2 * j + d] = dadd(K[2 * j + d], dscale(Jw, k));
          }
          const double gd = gq[2 * i] * gq[2 * j] + gq[2 * i + 1] * gq[2 * j + 1];
          K[6 + j]        = dadd(K[6 + j], dscale(gd / rho * Jw, tau));
        }
      }
    }
    dual out = dsub(dc(0.0), F);
    for (int col = 0; col < 9; ++col) out = dadd(out, dscale(nxt_e[col], K[col]));
    Re[row]  = out.v;
    dRe[row] = out.d;
  }
}

/* Reverse mode by hand.  With adjoint a (9 per element) and next state n,
 * adj^T R_e = sum_q Jw_q l_q, where at a quadrature point, writing lam_c =
 * sum_i N_i a(i,c), Dlam_cd = sum_i dphi_i/dx_d a(i,c), pi = sum_i N_i a_p(i),
 * Dpi_d, v_c, Dv_cd, q, Dq_d likewise from n, w_c = adv . Dlam_c,
 * z_c = adv . Dv_c, ag = G adv with G the velocity gradient:
 *   l = -rho/dt lam.u + rho/2 lam.ag + mu/2 Dlam:G + rho/dt lam.v
 *       + rho/2 lam.z + mu/2 Dlam:Dv - q div lam + pi div v + tau l1,
 *   l1 = sum_c w_c B_c - Dpi.u/dt + Dpi.ag/2 + Dpi.v/dt + Dpi.z/2 + Dpi.Dq/rho,
 *   B_c = -rho/dt u_c + rho/2 ag_c + rho/dt v_c + rho/2 z_c + dq/dx_c.
 * u and G depend on lag 0, adv = 1.5 lag0 - 0.5 lag1 (step > 0, two lags) or
 * lag 0, tau on u through |u| and elemLength.  out is (9 per lag) the
 * element columns for both lags; pressure columns are zero.  The forward
 * values (direction u/|u|, h = 2/sum, (2|u|/h)^2, (4 nu/h^2)^2, tau) are
 * femx's own expressions (elemLength and evalQp, Components.hpp:286-394);
 * a device kernel that rearranges them is checked against this within a
 * tolerance, not bitwise. */
void or_gls_element_hvjp(const double* xe, const double* ye, double rho, double mu, double dt,
                         int step, int num_hist, const double* hist_e, const double* nxt_e,
                         const double* adj_e, double* out)
{
  double rs[6], w[3], invJ[4], jxw[3], N[3][3], g[6];
  static const double dNdr[6] = {-1.0, -1.0, 1.0, 0.0, 0.0, 1.0};
  or_quad_tri(2, rs, w);
  for (int q = 0; q < 3; ++q)
  {
    jxw[q] = tri_geometry(xe, ye, w[q], invJ);
    N[q][0] = 1.0 - rs[2 * q] - rs[2 * q + 1];
    N[q][1] = rs[2 * q];
    N[q][2] = rs[2 * q + 1];
  }
  phys_grad(3, dNdr, invJ, g);
  const int    extrap = step > 0 && num_hist > 1;
  const double al = extrap ? 1.5 : 1.0, be = extrap ? -0.5 : 0.0;
  const double nu = mu / rho, time = (2.0 / dt) * (2.0 / dt), rdt = rho / dt;
  double       Dl[2][2] = {{0, 0}, {0, 0}}, Dv[2][2] = {{0, 0}, {0, 0}}, G[2][2] = {{0, 0}, {0, 0}};
  double       Dpi[2] = {0, 0}, Dq[2] = {0, 0};
  for (int k = 0; k < 3; ++k)
    for (int d = 0; d < 2; ++d)
    {
      for (int c = 0; c < 2; ++c)
      {
        Dl[c][d] += g[2 * k + d] * adj_e[2 * k + c];
        Dv[c][d] += g[2 * k + d] * nxt_e[2 * k + c];
        G[c][d] += g[2 * k + d] * hist_e[2 * k + c];
      }
      Dpi[d] += g[2 * k + d] * adj_e[6 + k];
      Dq[d] += g[2 * k + d] * nxt_e[6 + k];
    }
  for (int k = 0; k < 18; ++k) out[k] = 0.0;
  for (int iq = 0; iq < 3; ++iq)
  {
    const double* Nq = N[iq];
    double        u[2] = {0, 0}, av[2] = {0, 0}, lam[2] = {0, 0}, v[2] = {0, 0};
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 2; ++c)
      {
        u[c] += Nq[k] * hist_e[2 * k + c];
        const double a = extrap ? 1.5 * hist_e[2 * k + c] - 0.5 * hist_e[9 + 2 * k + c]
                                : hist_e[2 * k + c];
        av[c] += Nq[k] * a;
        lam[c] += Nq[k] * adj_e[2 * k + c];
        v[c] += Nq[k] * nxt_e[2 * k + c];
      }
    double ag[2], wv[2], z[2];
    for (int c = 0; c < 2; ++c)
    {
      ag[c] = G[c][0] * av[0] + G[c][1] * av[1];
      wv[c] = av[0] * Dl[c][0] + av[1] * Dl[c][1];
      z[c]  = av[0] * Dv[c][0] + av[1] * Dv[c][1];
    }
    /* tau and its gradient in u */
    const double speed = sqrt(u[0] * u[0] + u[1] * u[1]);
    double       dtau[2] = {0.0, 0.0}, tau;
    {
      double dir[2], sgn[3], sum = 0.0;
      const int moving = speed > 1.0e-10;
      if (moving)
      {
        dir[0] = u[0] / speed;
        dir[1] = u[1] / speed;
      }
      else
        dir[0] = dir[1] = 1.0 / sqrt(2.0);
      for (int k = 0; k < 3; ++k)
      {
        const double gk = dir[0] * g[2 * k] + dir[1] * g[2 * k + 1];
        sgn[k]          = gk >= 0.0 ? 1.0 : -1.0;
        sum += sgn[k] * gk;
      }
      double flow = 0.0, diff = 0.0, dflow[2] = {0, 0}, ddiff[2] = {0, 0};
      if (sum > 1.0e-14)
      {
        const double h = 2.0 / sum;
        flow           = (2.0 * speed / h) * (2.0 * speed / h);
        diff           = (4.0 * nu / (h * h)) * (4.0 * nu / (h * h));
        if (moving)
        {
          /* S = sum_k |u.g_k| = speed*sum: flow = S^2, diff = nu^2 S^4 / speed^4 */
          const double S = speed * sum;
          for (int c = 0; c < 2; ++c)
          {
            double dS = 0.0;
            for (int k = 0; k < 3; ++k) dS += sgn[k] * g[2 * k + c];
            dflow[c] = 2.0 * S * dS;
            const double s4 = speed * speed * speed * speed;
            ddiff[c] = nu * nu * (4.0 * S * S * S * dS / s4 - 4.0 * S * S * S * S * u[c] / (s4 * speed * speed));
          }
        }
        else
          for (int c = 0; c < 2; ++c) dflow[c] = 8.0 / (h * h) * u[c];
      }
      tau = 1.0 / sqrt(time + flow + diff);
      for (int c = 0; c < 2; ++c) dtau[c] = -0.5 * tau * tau * tau * (dflow[c] + ddiff[c]);
    }
    double B[2], l1 = 0.0;
    for (int c = 0; c < 2; ++c)
    {
      B[c] = -rdt * u[c] + 0.5 * rho * ag[c] + rdt * v[c] + 0.5 * rho * z[c] + Dq[c];
      l1 += wv[c] * B[c];
    }
    for (int d = 0; d < 2; ++d)
      l1 += Dpi[d] * (-u[d] / dt + 0.5 * ag[d] + v[d] / dt + 0.5 * z[d] + Dq[d] / rho);
    double Pu[2], Pag[2], Padv[2], PG[2][2];
    for (int c = 0; c < 2; ++c)
    {
      Pu[c]  = -rdt * lam[c] + tau * (-rdt * wv[c] - Dpi[c] / dt) + l1 * dtau[c];
      Pag[c] = 0.5 * rho * lam[c] + tau * (0.5 * rho * wv[c] + 0.5 * Dpi[c]);
    }
    for (int d = 0; d < 2; ++d)
    {
      double s = 0.0;
      for (int c = 0; c < 2; ++c)
      {
        s += 0.5 * rho * lam[c] * Dv[c][d];
        s += tau * (Dl[c][d] * B[c] + 0.5 * rho * wv[c] * Dv[c][d] + 0.5 * Dpi[c] * Dv[c][d]);
        s += Pag[c] * G[c][d];
      }
      Padv[d] = s;
    }
    for (int c = 0; c < 2; ++c)
      for (int d = 0; d < 2; ++d) PG[c][d] = 0.5 * mu * Dl[c][d] + Pag[c] * av[d];
    const double Jw = jxw[iq];
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 2; ++c)
      {
        const double d0 = Nq[k] * Pu[c] + g[2 * k] * PG[c][0] + g[2 * k + 1] * PG[c][1]
                          + al * Nq[k] * Padv[c];
        out[2 * k + c] += Jw * d0;
        out[9 + 2 * k + c] += Jw * be * Nq[k] * Padv[c];
      }
  }
}

/* Operator-level history VJP (NavierStokesModel.cpp:321-363
 * applyNavierHistJacT): out = (dR/dhist_lag)^T adj over all elements,
 * summed in ascending element order. */
void or_gls_hist_vjp(or_index num_states, const double* x, const double* y, or_index num_elems,
                     const or_index* conn, const or_index* dofs9, double rho, double mu,
                     double dt, int step, int num_hist, int lag, const double* hist,
                     const double* nxt, const double* adj, double* out)
{
  const int nh = num_hist < 2 ? num_hist : 2;
  memset(out, 0, sizeof(double) * num_states);
  for (or_index e = 0; e < num_elems; ++e)
  {
    double xe[3], ye[3], he[18] = {0}, ne[9], ae[9], ve[18];
    gls_gather(x, y, conn + 3 * e, dofs9 + 9 * e, num_states, nh, hist, nxt, xe, ye, he, ne);
    for (int k = 0; k < 9; ++k) ae[k] = adj[dofs9[9 * e + k]];
    or_gls_element_hvjp(xe, ye, rho, mu, dt, step, num_hist, he, ne, ae, ve);
    if (lag < 2)
      for (int k = 0; k < 9; ++k) out[dofs9[9 * e + k]] += ve[9 * lag + k];
  }
}

/* Operator-level forward mode: out = (dR/dhist_lag) dir, assembled from the
 * dual-number element JVP in ascending element order (the independent check
 * of or_gls_hist_vjp through the adjoint identity). */
void or_gls_hist_jvp(or_index num_states, const double* x, const double* y, or_index num_elems,
                     const or_index* conn, const or_index* dofs9, double rho, double mu,
                     double dt, int step, int num_hist, int lag, const double* hist,
                     const double* nxt, const double* dir, double* out)
{
  const int nh = num_hist < 2 ? num_hist : 2;
  memset(out, 0, sizeof(double) * num_states);
  for (or_index e = 0; e < num_elems; ++e)
  {
    double xe[3], ye[3], he[18] = {0}, de[18] = {0}, ne[9], Re[9], dRe[9];
    gls_gather(x, y, conn + 3 * e, dofs9 + 9 * e, num_states, nh, hist, nxt, xe, ye, he, ne);
    if (lag < 2)
      for (int k = 0; k < 9; ++k) de[9 * lag + k] = dir[dofs9[9 * e + k]];
    or_gls_element_jvp(xe, ye, rho, mu, dt, step, num_hist, he, de, ne, Re, dRe);
    for (int k = 0; k < 9; ++k) out[dofs9[9 * e + k]] += dRe[k];
  }
}
