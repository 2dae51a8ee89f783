/*
 * femx_oracle.h -- CPU restatement of femx's data-parallel FE core.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * (or as the timed CPU baseline); the product never links it.
 *
 * Every function names the femx source it restates.  Functions that extend
 * femx (P1 diffusion coefficient and load, P1 log-kappa VJP, P2 element,
 * 7-point rule, Taylor-Hood Navier-Stokes) have no counterpart in femx; they
 * are pinned in tests/ by closed forms, finite differences, manufactured
 * solutions and invariants rather than by the reference.
 */
#ifndef FEMX_ORACLE_H
#define FEMX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t or_index;

void or_free(void* p);

/* ---- meshes ------------------------------------------------------------ */
/* Structured P1 triangles on [0,1]^2: N=(n+1)^2 nodes, E=2n^2 elements. */
void or_mesh_p1(int n, double* x, double* y, or_index* conn);
/* P2 lattice nodes (N_v=(2n+1)^2), 6-node connectivity (Gmsh order) and the
 * 3 vertex (pressure) node ids per element. */
void or_mesh_p2(int n, double* x, double* y, or_index* conn6, or_index* vconn3);
/* Taylor-Hood element DOF table (15 per element): 12 interleaved velocity
 * DOFs then 3 pressure DOFs offset by 2*N_v. */
void or_th_dofs(int n, or_index num_elems, const or_index* conn6, const or_index* vconn3,
                or_index* dofs15);

/* ---- AssemblyMap (femx/assembly/AssemblyMap.cpp:54-176) ---------------- */
int or_map_build(or_index num_res, or_index num_states, or_index num_elems,
                 const or_index* res_offsets, const or_index* res_dofs,
                 const or_index* state_offsets, const or_index* state_dofs,
                 or_index** row_ptr, or_index** col_ind, or_index** jac_offsets,
                 or_index** jac_map, or_index* nnz);

/* ---- BoundaryMap (femx/assembly/BoundaryMap.cpp:59-139) ---------------- */
int or_bc_build(or_index rows, const or_index* row_ptr, const or_index* col_ind,
                or_index num_bcs, const or_index* dofs, or_index* diag,
                or_index* col_offsets, or_index** col_entries, or_index** col_rows,
                or_index* bc_mask);
void or_replace_rows(const or_index* row_ptr, or_index num_bcs, const or_index* bc_rows,
                     const or_index* diag, double* vals, double diag_val);
void or_replace_res(or_index num_bcs, const or_index* bc_rows, const double* state,
                    const double* bc_vals, double* res);
void or_zero_boundary(or_index num_bcs, const or_index* bc_rows, double* v);
void or_prepare_forward(const or_index* row_ptr, or_index num_bcs, const or_index* bc_rows,
                        const or_index* diag, const or_index* col_offsets,
                        const or_index* col_entries, const or_index* col_rows,
                        const or_index* bc_mask, double* vals, double* rhs,
                        const double* bc_vals);

/* ---- CSR SpMV (femx/linalg/handler/MatrixHandler.cpp:157-198) ---------- */
void or_matvec(or_index rows, const or_index* row_ptr, const or_index* col_ind,
               const double* vals, const double* x, double* y, double alpha, double beta);
void or_matvecT(or_index rows, or_index cols, const or_index* row_ptr, const or_index* col_ind,
                const double* vals, const double* x, double* y, double alpha, double beta);

/* ---- quadrature and shapes --------------------------------------------- */
/* order 1, 2 (femx/fem/GaussQuadrature.hpp:175-204) or 5 (7-point rule). */
int or_quad_tri(int order, double* rs, double* w);
void or_p2_shape(double r, double s, double* N, double* dNdr);

/* ---- P1 Poisson (kappa, load) ------------------------------------------ */
void or_p1_element(const double* xe, const double* ye, double kappa, const double* fe,
                   const double* ue, double* Ke, double* Re);
void or_p1_assemble(or_index num_nodes, const double* x, const double* y, or_index num_elems,
                    const or_index* conn, const double* kappa, const double* f,
                    const double* u, const or_index* jac_map, or_index nnz, double* vals,
                    double* res);
void or_p1_assemble_omp(or_index num_nodes, const double* x, const double* y,
                        or_index num_elems, const or_index* conn, const double* kappa,
                        const double* f, const double* u, const or_index* jac_map,
                        or_index nnz, double* vals, double* res);
void or_p1_vjp(or_index num_nodes, const double* x, const double* y, or_index num_elems,
               const or_index* conn, const double* m, const double* u, const double* lam,
               const or_index* bc_mask, double* jt_lam, double* g);
void or_p1_vjp_omp(or_index num_nodes, const double* x, const double* y, or_index num_elems,
                   const or_index* conn, const double* m, const double* u, const double* lam,
                   const or_index* bc_mask, double* jt_lam, double* g);

/* ---- Taylor-Hood P2/P1 steady Navier-Stokes ---------------------------- */
void or_th_element(const double* xv, const double* yv, double nu, const double* se,
                   double* Ke, double* Re);
void or_th_assemble(const double* xv, const double* yv, or_index num_elems,
                    const or_index* vconn3, const or_index* dofs15, or_index ndof,
                    double nu, const double* nu_e, const double* state, const or_index* jac_map,
                    or_index nnz, double* vals, double* res);
void or_th_assemble_omp(const double* xv, const double* yv, or_index num_elems,
                        const or_index* vconn3, const or_index* dofs15, or_index ndof,
                        double nu, const double* nu_e, const double* state,
                        const or_index* jac_map, or_index nnz, double* vals, double* res);
void or_th_vjp(const double* xv, const double* yv, or_index num_elems,
               const or_index* vconn3, const or_index* dofs15, or_index ndof, double nu,
               const double* nu_e, const double* state, const double* lam,
               const or_index* bc_mask, double* jt_lam, double* g);

/* ---- GLS P1/P1 Navier-Stokes time step (femx/model/ns/Components.hpp) -- */
void or_gls_element(const double* xe, const double* ye, double rho, double mu, double dt,
                    int step, int num_hist, const double* hist_e, const double* nxt_e,
                    double* Ke, double* Re);
void or_gls_assemble(or_index num_states, const double* x, const double* y, or_index num_elems,
                     const or_index* conn, const or_index* dofs9, double rho, double mu,
                     double dt, int step, int num_hist, const double* hist, const double* nxt,
                     const or_index* jac_map, or_index nnz, double* vals, double* res);
void or_gls_assemble_omp(or_index num_states, const double* x, const double* y,
                         or_index num_elems, const or_index* conn, const or_index* dofs9,
                         double rho, double mu, double dt, int step, int num_hist,
                         const double* hist, const double* nxt, const or_index* jac_map,
                         or_index nnz, double* vals, double* res);

/* History derivatives (femx computes them with Enzyme): forward-mode JVP
 * of the element residual in dual numbers, hand-written element VJP (both
 * lags, 18 values), and the operator-level VJP for one lag. */
void or_gls_element_jvp(const double* xe, const double* ye, double rho, double mu, double dt,
                        int step, int num_hist, const double* hist_e, const double* dhist_e,
                        const double* nxt_e, double* Re, double* dRe);
void or_gls_element_hvjp(const double* xe, const double* ye, double rho, double mu, double dt,
                         int step, int num_hist, const double* hist_e, const double* nxt_e,
                         const double* adj_e, double* out);
void or_gls_hist_vjp(or_index num_states, const double* x, const double* y, or_index num_elems,
                     const or_index* conn, const or_index* dofs9, double rho, double mu,
                     double dt, int step, int num_hist, int lag, const double* hist,
                     const double* nxt, const double* adj, double* out);
void or_gls_hist_jvp(or_index num_states, const double* x, const double* y, or_index num_elems,
                     const or_index* conn, const or_index* dofs9, double rho, double mu,
                     double dt, int step, int num_hist, int lag, const double* hist,
                     const double* nxt, const double* dir, double* out);

int or_omp_threads(void);

#ifdef __cplusplus
}
#endif
#endif
