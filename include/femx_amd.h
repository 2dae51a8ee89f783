/*
 * femx_amd.h -- C-ABI of the MI355X-native femx FE core.
 *
 * All sizes and indices are int32 (femx `Index`, femx/common/Types.hpp:19);
 * all values are FP64 (femx `Real`, :17).  Pointers named d_* are device
 * pointers; h_* are host pointers.  `stream` is a hipStream_t passed as
 * void*, as femx's CudaContext hands its stream to kernels
 * (femx/common/Context.hpp:68-71).  Device entry points are asynchronous on
 * `stream`; callers synchronise explicitly, as with femx's CUDA path
 * (femx/assembly/CudaAssembly.hpp:302-318).
 *
 * Every function returns FX_OK (0) or an error code; fx_last_error() gives
 * the message of the last failure on the calling thread.  No C++ exception
 * crosses this boundary.  Outputs are replaced, never accumulated into,
 * unless a function says otherwise (femx/assembly/Assembly.hpp:168-169).
 *
 * Results are deterministic: no entry point uses floating-point atomics, so
 * repeated calls on the same inputs give bitwise equal outputs.
 */
#ifndef FEMX_AMD_H
#define FEMX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t fx_index;

enum fx_status
{
  FX_OK              = 0,
  FX_ERR_ARG         = 1, /* sizes, ranges or null pointers */
  FX_ERR_ALIAS       = 2, /* outputs overlap inputs or each other */
  FX_ERR_LAYOUT      = 3, /* map/matrix layout mismatch */
  FX_ERR_OVERFLOW    = 4, /* int32 index range exceeded */
  FX_ERR_HIP         = 5, /* HIP runtime error */
  FX_ERR_PATTERN     = 6, /* CSR pattern lacks a required entry */
  FX_ERR_UNSUPPORTED = 7, /* element/map shape not handled by a kernel */
  FX_ERR_ALLOC       = 8, /* host allocation failed */
  FX_ERR_INTERNAL    = 9
};

const char* fx_last_error(void);
int fx_version(void);

/* ======================================================================== */
/* Structured triangle meshes on [0,1]^2 (host).  femx has only
 * Mesh::makeStructuredQuad (femx/fem/Mesh.cpp:10-51); cell (i,j) is split
 * along its (i,j)-(i+1,j+1) diagonal into two counter-clockwise triangles.
 * P1: N=(n+1)^2 nodes, E=2n^2, h_conn[3E].  P2: N_v=(2n+1)^2 lattice nodes,
 * h_conn6[6E] in Gmsh order (vertices, edges (0,1),(1,2),(2,0)) and
 * h_vconn3[3E] vertex ids on the (n+1)^2 vertex lattice. */
int fx_mesh_p1(fx_index n, double* h_x, double* h_y, fx_index* h_conn);
int fx_mesh_p2(fx_index n, double* h_x, double* h_y, fx_index* h_conn6, fx_index* h_vconn3);
/* Taylor-Hood element DOF table h_dofs15[15E]: interleaved velocity DOFs
 * 2*node+c (FESpace::globalDof, femx/fem/FESpace.cpp:103-106) then pressure
 * DOFs offset by 2*N_v (MixedFESpace::elemDofs, femx/fem/MixedFESpace.cpp:
 * 121-137). */
int fx_th_dofs(fx_index n, const fx_index* h_conn6, const fx_index* h_vconn3,
               fx_index* h_dofs15);
/* Taylor-Hood tables of any conforming triangle mesh (host), as femx's P1
 * mesh (femx/fem/Mesh.hpp, e.g. from GmshReader) gives it: num_vertices
 * vertices and conn3[3E].  P2 node ids: vertex v keeps id v; edge nodes
 * follow, numbered num_vertices, num_vertices+1, ... in order of first
 * appearance over the elements in order, local edges (0,1), (1,2), (2,0)
 * (Gmsh 6-node order).  Writes *num_p2_nodes, and when the arrays are
 * non-null conn6[6E], edge_ends[2*(num_p2_nodes - num_vertices)] (each edge
 * node's two vertices, in first-appearance order) and dofs15[15E] (the
 * MixedFESpace order of fx_th_dofs: 2*p2node+c, then 2*num_p2_nodes +
 * vertex).  Call once with null arrays for the count.  FX_ERR_ARG for an
 * out-of-range vertex or a degenerate element (repeated vertex);
 * FX_ERR_UNSUPPORTED for an edge shared by more than two elements. */
int fx_th_mesh(fx_index num_vertices, fx_index num_elems, const fx_index* h_conn3,
               fx_index* num_p2_nodes, fx_index* h_conn6, fx_index* h_edge_ends,
               fx_index* h_dofs15);

/* ======================================================================== */
/* Assembly map (host).  Replaces femx::assembly::makeAssemblyMap
 * (femx/assembly/AssemblyMap.hpp:220-236, AssemblyMap.cpp:54-176): the CSR
 * pattern and jac_map are bit-identical to femx's.  It also derives the
 * row-incidence tables that the deterministic row-gather kernels read. */
typedef struct fx_map fx_map;

typedef struct
{
  fx_index num_elems, num_res, num_states;
  fx_index nnz;         /* CSR entries */
  fx_index jac_len;     /* jac_offsets[num_elems] (sum of local matrix sizes) */
  fx_index max_res, max_state, max_jac;
  fx_index max_row_len; /* longest CSR row */
  fx_index num_inc;     /* row incidences: sum over elements of local rows */
} fx_map_info;

int fx_map_create(fx_index num_res, fx_index num_states, fx_index num_elems,
                  const fx_index* h_res_offsets, const fx_index* h_res_dofs,
                  const fx_index* h_state_offsets, const fx_index* h_state_dofs,
                  fx_map** out);
int fx_map_info_get(const fx_map* map, fx_map_info* info);
/* row_ptr[num_res+1], col_ind[nnz] (femx CsrPattern, femx/linalg/CsrPattern.hpp). */
int fx_map_export_pattern(const fx_map* map, fx_index* h_row_ptr, fx_index* h_col_ind);
/* femx's jac_offsets[num_elems+1] and jac_map[jac_len] (AssemblyMap.hpp:22-68). */
int fx_map_export_jac(const fx_map* map, fx_index* h_jac_offsets, fx_index* h_jac_map);
/* Row incidence for elements with 3 local rows and 3 local columns (P1):
 * inc_off[num_res+1]; inc[2*num_inc] pairs (e, li | p0<<8 | p1<<16 | p2<<24)
 * listing, for each row in ascending element order, the element e, the local
 * row li and the positions p_j (within the row, from row_ptr[row]) of the
 * element's three columns.  FX_ERR_UNSUPPORTED for other shapes or rows
 * longer than 255. */
int fx_map_export_rowinc3(const fx_map* map, fx_index* h_inc_off, fx_index* h_inc);
/* Generic row incidence: inc_off[num_res+1]; inc_pos[num_inc] -- for each
 * row, in ascending element order, the position res_offsets[e] + li of its
 * (element, local row) in the flattened element DOF table; elem_inc, the
 * inverse (one entry per flattened element DOF): the incidence index of
 * (e, li).  inc_pos and elem_inc may be NULL. */
int fx_map_export_rowinc(const fx_map* map, fx_index* h_inc_off, fx_index* h_inc_pos,
                         fx_index* h_elem_inc);
/* Row incidence for elements with 15 local rows and columns (Taylor-Hood):
 * inc_off[num_res+1], inc_elem[num_inc] and inc_meta[4*num_inc] -- 16 bytes
 * per incidence: the positions (within the row, from row_ptr[row]) of the
 * element's 15 columns, then the local row.  Rows in ascending element
 * order.  FX_ERR_UNSUPPORTED for other shapes or rows longer than 255. */
int fx_map_export_rowinc15(const fx_map* map, fx_index* h_inc_off, fx_index* h_inc_elem,
                           fx_index* h_inc_meta);
/* Work blocks of the Taylor-Hood kernels: rows cut into contiguous blocks of
 * at most `lanes` work items and at most min(max_nnz, FX_TH_TILE_NNZ_MAX)
 * CSR entries, max_nnz = 0 meaning FX_TH_TILE_NNZ_MAX (a block always holds
 * at least one item).  An item is a pair of
 * interchangeable velocity rows (r, r+1) -- same elements, local rows li and
 * li+1 (li even, < 12), same column positions -- or a single row.
 * blk_row[num_blocks+1], blk_flags[num_blocks] (1 = pairs); max_block_nnz
 * sizes the per-wave LDS row tile.  Pass null arrays to query num_blocks and
 * max_block_nnz. */
#define FX_TH_BLOCK_ITEMS 32 /* work items per block the kernels are tuned for */
#define FX_TH_TILE_NNZ 0     /* default tile bound (0: the LDS capacity below) */
/* The largest row tile the assembly kernels hold per wave (LDS: four waves'
 * tiles and the reference tensors in 160 KiB).  max_nnz = 0 or above it
 * means this bound, so that every block fits; a block still always holds at
 * least one item. */
#define FX_TH_TILE_NNZ_MAX 4928
int fx_map_th_blocks(const fx_map* map, fx_index lanes, fx_index max_nnz, fx_index* num_blocks,
                     fx_index* max_block_nnz, fx_index* h_blk_row, fx_index* h_blk_flags);
/* Lane map of those blocks (host): each block's 64 lanes shared out so its
 * items finish together.  With d_i the incidence count of item i, the block
 * runs R rounds, the smallest R with sum_i ceil(d_i / R) <= 64, and item i
 * gets k_i = max(1, ceil(d_i / R)) consecutive lanes; its lane of role j
 * (0 <= j < k_i) takes incidences j, j + k_i, ...  h_blk_lanes[64 *
 * num_blocks]: item | role << 8 | k_i << 16 per lane, -1 for an idle lane.
 * (The map that fx_th_assemble uses when given none is two lanes per item.) */
int fx_map_th_lanes(const fx_map* map, fx_index num_blocks, const fx_index* h_blk_row,
                    const fx_index* h_blk_flags, fx_index* h_blk_lanes);
/* Round tables of those blocks (host): block b runs R_b rounds, the most
 * incidences any of its lanes takes under the lane map, and the incidence
 * its lane l takes in round m is entry h_rt_off[b] + 64 m + l of h_rt_elem
 * and h_rt_meta (four words per entry, as fx_map_export_rowinc15 writes
 * them, from its h_inc_off, h_inc_elem and h_inc_meta).  An entry past the
 * lane's last incidence, or on an idle lane, repeats a valid incidence of
 * the block with bit 31 of its last meta word set.  num_rows is the map's
 * residual row count (h_inc_off has num_rows + 1 entries).
 * h_rt_off[num_blocks+1]; pass null tables to get h_rt_off (the table length
 * is its last entry). */
int fx_map_th_rounds(fx_index num_rows, fx_index num_blocks, const fx_index* h_blk_row,
                     const fx_index* h_blk_flags, const fx_index* h_blk_lanes,
                     const fx_index* h_inc_off, const fx_index* h_inc_elem,
                     const fx_index* h_inc_meta, fx_index* h_rt_off, fx_index* h_rt_elem,
                     fx_index* h_rt_meta);
/* Node plan of the GLS P1/P1 Navier-Stokes kernel (fx_gls_assemble) over a
 * map whose elements have 9 local DOFs in NavierOperator order -- u0 v0 u1
 * v1 u2 v2 p0 p1 p2 (femx/model/ns/Components.hpp:247-256) -- for the mesh
 * conn3[3E] of num_nodes nodes.  Outputs (host):
 *   node_rows[3N]  velocity x, velocity y and pressure row of each node
 *                  (-1 for a node of no element);
 *   inc_off[N+1], inc_elem[3E]  each node's incident elements, ascending;
 *   inc_meta[8 per incidence]   32 bytes: byte 9k+j is the position of local
 *                  column j in the node's row k (from row_ptr[row]), byte 27
 *                  the node's local index;
 *   inc_slot[3E]   slot of element e in the list of its local node k:
 *                  inc_elem[inc_slot[3e+k]] == e, within node conn3[3e+k]'s
 *                  range (NULL: not written);
 *   blk_runs[4*FX_GLS_MAX_RUNS per block of FX_GLS_BLOCK_NODES nodes]: the
 *                  block's rows as runs of consecutive rows -- CSR begin,
 *                  CSR length, LDS tile offset, 0 (unused runs all zero);
 *   node_tile[3N]  tile offset of each node row's first entry;
 *   max_tile       the largest block tile, in doubles.
 * FX_ERR_UNSUPPORTED when the element tables disagree on a node's rows, a
 * row belongs to no node, a block's rows need more than FX_GLS_MAX_RUNS runs
 * or a row is longer than 255 entries. */
#define FX_GLS_BLOCK_NODES 32
#define FX_GLS_MAX_RUNS 4
int fx_map_gls_plan(const fx_map* map, fx_index num_nodes, const fx_index* h_conn3,
                    fx_index* h_node_rows, fx_index* h_inc_off, fx_index* h_inc_elem,
                    fx_index* h_inc_meta, fx_index* h_inc_slot, fx_index* h_blk_runs,
                    fx_index* h_node_tile, fx_index* max_tile);
void fx_map_destroy(fx_map* map);

/* CSR transpose permutation (host), for the deterministic transpose apply:
 * t_ptr[cols+1], t_row[nnz], t_perm[nnz] list column c's entries in
 * ascending row order with their source CSR index.  Replaces the cached
 * transpose map of MatrixHandler<CudaCsrBackend>::transpose
 * (femx/linalg/handler/MatrixHandler.cu:401-512). */
int fx_csr_transpose_map(fx_index rows, fx_index cols, const fx_index* h_row_ptr,
                         const fx_index* h_col_ind, fx_index* h_t_ptr, fx_index* h_t_row,
                         fx_index* h_t_perm);
/* Stencil blocks of fx_csr_matvec_multi_st (host; no counterpart in femx,
 * whose products go one vector at a time): rows in blocks of 64,
 * num_blocks = ceil(rows/64).  A full block whose rows all hold S entries
 * (1 <= S <= max_len <= 31) at the same column offsets col - row gets
 * blk_code = (start << 5) | S, its S offsets at st_off[start..]; equal
 * stencils share one copy; other blocks get -1.  Call with null blk_code and
 * st_off for num_blocks and num_offsets. */
#define FX_STENCIL_MAX_LEN 8 /* the longest stencil fx_csr_matvec_multi_st reads as one */
int fx_csr_stencil_plan(fx_index rows, fx_index cols, const fx_index* h_row_ptr,
                        const fx_index* h_col_ind, fx_index max_len, fx_index* num_blocks,
                        fx_index* num_offsets, fx_index* h_blk_code, fx_index* h_st_off);

/* ======================================================================== */
/* Boundary map (host).  Replaces femx::assembly::makeBoundaryMap
 * (femx/assembly/BoundaryMap.hpp:185-186, BoundaryMap.cpp:59-139), with the
 * same validation and the same arrays, plus the row-grouped elimination
 * lists used by fx_bc_prepare_forward in place of atomics. */
typedef struct fx_bcmap fx_bcmap;

typedef struct
{
  fx_index rows, num_bcs;
  fx_index num_col_entries; /* col_offsets[num_bcs] */
  fx_index num_elim_rows;   /* unconstrained rows touching constrained columns */
  fx_index num_elim_entries;
} fx_bcmap_info;

int fx_bcmap_create(fx_index rows, const fx_index* h_row_ptr, const fx_index* h_col_ind,
                    fx_index num_bcs, const fx_index* h_dofs, fx_bcmap** out);
int fx_bcmap_info_get(const fx_bcmap* map, fx_bcmap_info* info);
/* femx's BoundaryMapView arrays (BoundaryMap.hpp:17-62). */
int fx_bcmap_export(const fx_bcmap* map, fx_index* h_bc_rows, fx_index* h_diag,
                    fx_index* h_col_offsets, fx_index* h_col_entries, fx_index* h_col_rows,
                    fx_index* h_bc_mask);
/* elim_rows[num_elim_rows], elim_off[num_elim_rows+1], elim_k[num_elim_entries]
 * (CSR index) and elim_ib[num_elim_entries] (boundary index), each row's
 * entries in ascending boundary index -- the order femx's host loop applies
 * them (BoundaryMap.cpp:227-240). */
int fx_bcmap_export_elim(const fx_bcmap* map, fx_index* h_elim_rows, fx_index* h_elim_off,
                         fx_index* h_elim_k, fx_index* h_elim_ib);
void fx_bcmap_destroy(fx_bcmap* map);

/* ======================================================================== */
/* Boundary operations (device).  Replace the CUDA overloads of replaceRows,
 * replaceRes, zeroBoundary and prepareForwardSolve (femx/assembly/
 * BoundaryMap.hpp:220-294, BoundaryMap.cu:16-211). */
int fx_bc_replace_rows(fx_index num_bcs, const fx_index* d_bc_rows, const fx_index* d_diag,
                       const fx_index* d_row_ptr, double* d_vals, double diag, void* stream);
int fx_bc_replace_res(fx_index num_bcs, const fx_index* d_bc_rows, const double* d_state,
                      const double* d_bc_vals, double* d_res, void* stream);
int fx_bc_zero(fx_index num_bcs, const fx_index* d_bc_rows, double* d_v, void* stream);
int fx_bc_prepare_forward(fx_index num_bcs, const fx_index* d_bc_rows, const fx_index* d_diag,
                          const fx_index* d_row_ptr, fx_index num_elim_rows,
                          const fx_index* d_elim_rows, const fx_index* d_elim_off,
                          const fx_index* d_elim_k, const fx_index* d_elim_ib, double* d_vals,
                          double* d_rhs, const double* d_bc_vals, void* stream);

/* ======================================================================== */
/* CSR SpMV (device).  y = alpha*A*x + beta*y, sums in CSR order per row
 * (MatrixHandler<HostCsrBackend>::matvec, femx/linalg/handler/
 * MatrixHandler.cpp:157-175; replaces cusparseSpMV, MatrixHandler.cu:
 * 335-398,514-527).  With beta == 0, y is not read.  x and y must not
 * overlap each other or vals (MatrixHandler.cu:182-202). */
int fx_csr_matvec(fx_index rows, fx_index cols, const fx_index* d_row_ptr,
                  const fx_index* d_col_ind, const double* d_vals, const double* d_x,
                  double* d_y, double alpha, double beta, void* stream);
/* The same product with 16-bit column offsets: col_ind[k] = row + col_off[k].
 * Identical results (same order and arithmetic) from 2 fewer bytes per
 * entry.  col_off is a derived index of the pattern built once by
 * fx_csr_col_off16 and cached beside it, as femx caches its SpMV descriptors
 * per pattern (MatrixHandler.cu:103-115). */
int fx_csr_matvec16(fx_index rows, fx_index cols, const fx_index* d_row_ptr,
                    const int16_t* d_col_off, const double* d_vals, const double* d_x,
                    double* d_y, double alpha, double beta, void* stream);
/* Host: col_off[k] = col_ind[k] - row for every entry; FX_ERR_UNSUPPORTED when
 * an offset does not fit int16 (bandwidth of 32767 or more). */
int fx_csr_col_off16(fx_index rows, const fx_index* h_row_ptr, const fx_index* h_col_ind,
                     int16_t* h_col_off);
/* y = alpha*A^T*x + beta*y through the transpose map (matvecT,
 * MatrixHandler.cpp:177-198 and MatrixHandler.cu:529-542), each column summed
 * in ascending row order. */
int fx_csr_matvec_t(fx_index rows, fx_index cols, const fx_index* d_t_ptr,
                    const fx_index* d_t_row, const fx_index* d_t_perm, const double* d_vals,
                    const double* d_x, double* d_y, double alpha, double beta, void* stream);

/* Vector ops (VectorHandler<CudaCsrBackend>, femx/linalg/handler/
 * VectorHandler.cu:124-317).  axpby: y = a*x + b*y; with b == 0, y is not
 * read. */
int fx_vec_axpby(fx_index n, double a, const double* d_x, double b, double* d_y, void* stream);
/* nv products with one matrix: Y[j*ldy + :] = alpha A X[j*ldx + :] + beta
 * Y[j*ldy + :], j < nv, the columns given as col_ind or as 16-bit offsets
 * col_off (fx_csr_col_off16); exactly one of them non-null.  Each product
 * equals fx_csr_matvec's bit for bit; the matrix is read once per 16
 * vectors.  (cusparseSpMM's role; femx calls matvec per vector.) */
int fx_csr_matvec_multi(fx_index rows, fx_index cols, const fx_index* d_row_ptr,
                        const fx_index* d_col_ind, const int16_t* d_col_off, const double* d_vals,
                        fx_index nv, const double* d_X, fx_index ldx, double* d_Y, fx_index ldy,
                        double alpha, double beta, void* stream);
/* The same nv products over the stencil blocks of fx_csr_stencil_plan (its
 * blk_code and st_off on the device): in a stencil block each band of
 * consecutive offsets is one coalesced load per vector, its other offsets
 * the same values moved across lanes; other blocks (and codes with S above
 * FX_STENCIL_MAX_LEN) gather per entry.  Each product equals
 * fx_csr_matvec's bit for bit.  Vectors under 4 GiB (FX_ERR_OVERFLOW). */
int fx_csr_matvec_multi_st(fx_index rows, fx_index cols, const fx_index* d_row_ptr,
                           const fx_index* d_col_ind, const double* d_vals,
                           const fx_index* d_blk_code, const fx_index* d_st_off, fx_index nv,
                           const double* d_X, fx_index ldx, double* d_Y, fx_index ldy,
                           double alpha, double beta, void* stream);
/* gather: dst[i] = src[idx[i]]; scatter: dst[idx[i]] = src[i], i < n
 * (VectorHandler::gather/scatter, femx/linalg/handler/VectorHandler.cpp:
 * 46-80).  Indices must be in range (the host-side builders check them);
 * scatter indices must be distinct. */
int fx_vec_gather(fx_index n, const fx_index* d_idx, const double* d_src, fx_index src_len,
                  double* d_dst, void* stream);
int fx_vec_scatter(fx_index n, const fx_index* d_idx, const double* d_src, double* d_dst,
                   fx_index dst_len, void* stream);
/* y[i] = v, i < n (with v = 0, VectorHandler::zero and resizeOrZero,
 * femx/linalg/handler/VectorHandler.hpp:282-300,365-371): a kernel, not a
 * memset, so that it may be captured into a hipGraph. */
int fx_vec_fill(fx_index n, double v, double* d_y, void* stream);
/* z = x / d elementwise (the Jacobi step of the device iterative solver;
 * no counterpart in femx's VectorHandler).  z may alias x. */
int fx_vec_div(fx_index n, const double* d_x, const double* d_d, double* d_z, void* stream);
/* Deterministic dot product into d_out[0]; d_work holds fx_dot_work_len()
 * doubles. */
fx_index fx_dot_work_len(void);
int fx_vec_dot(fx_index n, const double* d_x, const double* d_y, double* d_work, double* d_out,
               void* stream);

/* ======================================================================== */
/* P1 Poisson assembly (device).  Drop-in for assembly::assemble with the
 * Poisson operator (femx/assembly/CudaAssembly.hpp:319-364 driving
 * examples/poisson/PoissonComponents.hpp:35-64), extended with a per-element
 * diffusion kappa[E] (or its logarithm m[E], kappa = exp(m), when
 * kappa_is_log is nonzero) and a nodal load f[N]:
 *   vals = CSR values of K (K^e = kappa_e * femx's Laplace element matrix),
 *   res  = K u - F,  F_i = int f_h phi_i  (f_h the P1 interpolant of f).
 * f or u may be NULL (zero); vals may be NULL for a residual-only pass
 * (assembleResidual, CudaAssembly.hpp:420-472).  Geometry is SoA x[N], y[N]; conn[3E] is the
 * element DOF table (res_dofs = state_dofs); inc_off/inc come from
 * fx_map_export_rowinc3, row_ptr from fx_map_export_pattern, max_row_len from
 * fx_map_info_get.  Each CSR value and residual entry is summed over its
 * elements in ascending element order, as femx's host assemble does
 * (Assembly.hpp:184-225). */
int fx_p1_assemble(fx_index num_nodes, fx_index num_elems, fx_index nnz,
                   fx_index max_row_len, const double* d_x, const double* d_y,
                   const fx_index* d_conn, const double* d_kappa, int kappa_is_log,
                   const double* d_f, const double* d_u, const fx_index* d_inc_off, const fx_index* d_inc,
                   const fx_index* d_row_ptr, double* d_vals, double* d_res, void* stream);

/* P1 log-kappa adjoint VJP (device), one pass over the row incidence:
 *   jt_lam = J^T lam   (J the row-replaced Jacobian: lam is zeroed on
 *                       constrained rows first and passes through there),
 *   g[e]   = (dR/dm_e)^T lam = lam~_e^T K^e u_e,  kappa_e = exp(m_e).
 * bc_mask[N] may be NULL (unconstrained).  Fills the Residual::applyParamJacT
 * contract (femx/state/Residual.hpp:55-59) used by ReducedFunctional::gradAt
 * (femx/inverse/ReducedFunctional.hpp:94-117), with the boundary convention
 * of ConstrainedTimeResidual::applyJacT (femx/assembly/
 * ConstrainedTimeResidual.cpp:379-405). */
int fx_p1_vjp(fx_index num_nodes, fx_index num_elems, const double* d_x, const double* d_y,
              const fx_index* d_conn, const double* d_m, const double* d_u,
              const double* d_lam, const fx_index* d_bc_mask, const fx_index* d_inc_off,
              const fx_index* d_inc, double* d_jt_lam, double* d_g, void* stream);
/* fx_p1_vjp in one pass over element patches: plan from fx_patches_create
 * with the nodes as DOFs (dofs = vconn3 = conn, coordinates x/y), on the
 * device as fx_patch_plan (declared below).  A node whose elements lie in one
 * patch gets fx_p1_vjp's value bit for bit (its elements in ascending order);
 * a node on a patch seam sums per-patch partials, so it may differ from
 * fx_p1_vjp in the last bits.  g is bitwise equal to fx_p1_vjp's. */
typedef struct fx_patch_plan fx_patch_plan;
int fx_p1_vjp_patch(fx_index num_nodes, fx_index num_elems, const double* d_x, const double* d_y,
                    const double* d_m, const double* d_u, const double* d_lam,
                    const fx_index* d_bc_mask, const fx_patch_plan* h_plan, double* d_jt_lam,
                    double* d_g, void* stream);

/* fx_p1_assemble over element patches in which every row is finished by
 * the one patch that owns its node, so no row is split between patches and
 * none needs a second pass.  Elements are visited in Morton order of their
 * centroids (vertex coordinates x/y); a node is owned by the patch that
 * meets it first, and that patch takes all of the node's elements -- its own
 * and, as a halo, elements of neighbouring patches, so an element near a
 * patch border is evaluated by each patch that owns one of its nodes.  A
 * workgroup loads its patch's node values into LDS once, evaluates each
 * element once (one lane per element), and adds the rows of its owned nodes
 * into an LDS tile in rank steps: step k adds every element that is its
 * node's k-th in ascending element order, so each CSR value and residual
 * entry is summed in ascending element order, as fx_p1_assemble sums it, and
 * the results are bitwise equal.  The tile holds the owned rows in ascending
 * node order and is written back along runs of consecutive node ids, which
 * are contiguous CSR ranges.  Nodes a patch cannot own (rows of more than 16
 * entries, more than 16 elements, nodes in no element) are left to
 * fx_p1_assemble's row-per-lane kernel (rows_left).
 * Plan (host, fx_p1_asm_create on the map's CSR pattern; max_patch <=
 * FX_TH_PATCH_MAX elements per patch, tiles of at most max_tile values):
 *   patch_off[P+1], patch_elem   the elements of each patch, ascending;
 *   elem_ldof[4 i + k]           node k of the patch's i-th element: its local
 *                                index | its rank among the node's elements
 *                                << 12 (owned nodes);
 *   elem_pos[4 i + k]            the CSR positions in node k's row of the
 *                                element's three columns, 4 bits each
 *                                (column j at bits 4j; owned nodes);
 *   ldof_off[P+1], ldof_dof      each patch's local nodes, ascending;
 *   ldof_tile                    tile offset of an owned node's row, 0xffff
 *                                for the others;
 *   patch_tile[P]                tile length | rank steps << 20;
 *   run_off[P+1], runs           (tile offset, CSR offset) of each run;
 *   rows_left                    the rows no patch owns, ascending. */
typedef struct fx_p1_asm fx_p1_asm;
typedef struct
{
  fx_index num_patches, num_entries, num_ldofs, num_runs, num_left;
  fx_index max_patch_ldofs, max_tile, max_runs;
} fx_p1_asm_info;
int fx_p1_asm_create(fx_index num_nodes, fx_index num_elems, const fx_index* h_conn,
                     const double* h_x, const double* h_y, const fx_index* h_row_ptr,
                     const fx_index* h_col_ind, fx_index max_patch, fx_index max_tile,
                     fx_p1_asm** out);
int fx_p1_asm_info_get(const fx_p1_asm* plan, fx_p1_asm_info* info);
int fx_p1_asm_export(const fx_p1_asm* plan, fx_index* h_patch_off, fx_index* h_patch_elem,
                     unsigned short* h_elem_ldof, unsigned short* h_elem_pos,
                     fx_index* h_ldof_off, fx_index* h_ldof_dof, unsigned short* h_ldof_tile,
                     fx_index* h_patch_tile, fx_index* h_run_off, fx_index* h_runs,
                     fx_index* h_rows_left);
void fx_p1_asm_destroy(fx_p1_asm* plan);
/* The plan's arrays on the device. */
typedef struct
{
  fx_index              num_patches, num_left, max_patch_ldofs, max_tile, max_runs;
  const fx_index*       patch_off;
  const fx_index*       patch_elem;
  const unsigned short* elem_ldof;
  const unsigned short* elem_pos;
  const fx_index*       ldof_off;
  const fx_index*       ldof_dof;
  const unsigned short* ldof_tile;
  const fx_index*       patch_tile;
  const fx_index*       run_off;
  const fx_index*       runs;
  const fx_index*       rows_left;
} fx_p1_asm_plan;
int fx_p1_assemble_patch(fx_index num_nodes, fx_index num_elems, fx_index nnz,
                         const double* d_x, const double* d_y, const fx_index* d_conn,
                         const double* d_kappa, int kappa_is_log, const double* d_f,
                         const double* d_u, const fx_index* d_inc_off, const fx_index* d_inc,
                         const fx_index* d_row_ptr, const fx_p1_asm_plan* h_plan, double* d_vals,
                         double* d_res, void* stream);

/* Stencil blocks of the P1 assembly (host; no counterpart in femx): rows in
 * blocks of 64, num_blocks = ceil(N/64).  A full block whose rows have the
 * same column offsets (S entries) and the same incidences -- element
 * A*row + c_k, local row and column positions as in fx_map_export_rowinc3,
 * T incidences, S and T at most max_len -- and whose bands of consecutive
 * offsets fit max_band values (63 + w for a band of w offsets) gets
 * blk_code = the index in words of its record [S, T, A, offsets[S],
 * sidx[S], (c_k, meta_k) for k < T], sidx being each offset's place in the
 * band buffer (equal blocks share a record); other blocks get -1.  Call
 * with null blk_code and words for num_blocks and num_words.
 * (Lexicographically numbered structured meshes: all but the blocks holding
 * boundary nodes.) */
#define FX_P1_STENCIL_MAX 8    /* the largest S and T fx_p1_assemble_stencil reads as one */
#define FX_P1_STENCIL_BAND 208 /* its band buffer, in values per vertex array */
int fx_p1_stencil_plan(fx_index num_nodes, const fx_index* h_row_ptr, const fx_index* h_col_ind,
                       const fx_index* h_inc_off, const fx_index* h_inc, fx_index max_len,
                       fx_index max_band, fx_index* num_blocks, fx_index* num_words,
                       fx_index* h_blk_code, fx_index* h_words);
/* fx_p1_assemble over the stencil blocks of fx_p1_stencil_plan (blk_code,
 * words on the device): a uniform block's lanes load the vertex data of
 * their rows' columns once (each band of consecutive offsets one coalesced
 * load per array, its other offsets by lane shifts) and sum their rows in
 * registers; other blocks (and records with S or T above FX_P1_STENCIL_MAX)
 * take the row-gather path.  The plan must have been built with max_len <=
 * FX_P1_STENCIL_MAX and max_band <= FX_P1_STENCIL_BAND.  Arguments
 * otherwise as fx_p1_assemble; the results are bitwise equal to it. */
int fx_p1_assemble_stencil(fx_index num_nodes, fx_index num_elems, fx_index nnz,
                           fx_index max_row_len, const double* d_x, const double* d_y,
                           const fx_index* d_conn, const double* d_kappa, int kappa_is_log,
                           const double* d_f, const double* d_u, const fx_index* d_inc_off,
                           const fx_index* d_inc, const fx_index* d_row_ptr,
                           const fx_index* d_blk_code, const fx_index* d_words, double* d_vals,
                           double* d_res, void* stream);

/* ======================================================================== */
/* Taylor-Hood P2/P1 steady incompressible Navier-Stokes (device).  Not in
 * femx (its NS operator is GLS P1/P1, femx/model/ns/Components.hpp:400-577);
 * drops in for assembly::assemble with an element operator
 * (femx/assembly/CudaAssembly.hpp:319-364).  Galerkin residual and exact
 * Newton Jacobian on the 7-point degree-5 rule, viscosity nu (or nu_e[E] per
 * element when non-null), vertex coordinates xv/yv, vconn3[3E] vertex ids,
 * dofs15[15E] element DOFs (fx_th_dofs order), tables from
 * fx_map_export_rowinc15 and fx_map_th_blocks (at most FX_TH_BLOCK_ITEMS
 * items per block); blk_lanes from fx_map_th_lanes, or NULL for two lanes
 * per item.  vals may be NULL (residual only).
 * Each CSR value and residual entry is a sum over its elements in ascending
 * element order (no global atomics); arithmetic uses fused multiply-adds. */
int fx_th_assemble(fx_index ndof, fx_index num_elems, fx_index num_blocks,
                   fx_index max_block_nnz, const double* d_xv, const double* d_yv,
                   const fx_index* d_vconn3, const fx_index* d_dofs15, const double* d_state,
                   double nu, const double* d_nu_e, const fx_index* d_inc_off,
                   const fx_index* d_inc_elem, const fx_index* d_inc_meta,
                   const fx_index* d_blk_row, const fx_index* d_blk_flags,
                   const fx_index* d_blk_lanes, const fx_index* d_row_ptr, double* d_vals,
                   double* d_res, void* stream);
/* fx_th_assemble over the blocks' round tables (fx_map_th_rounds, on the
 * device; the lane map is required): each round's incidences are read at
 * addresses known from the block alone, so a block's first element loads
 * do not wait on its lane map and row offsets.  Drops in for the same
 * assembly::assemble call (femx/assembly/CudaAssembly.hpp:319-364); the
 * results equal fx_th_assemble's bit for bit. */
int fx_th_assemble_rounds(fx_index ndof, fx_index num_elems, fx_index num_blocks,
                          fx_index max_block_nnz, const double* d_xv, const double* d_yv,
                          const fx_index* d_vconn3, const fx_index* d_dofs15,
                          const double* d_state, double nu, const double* d_nu_e,
                          const fx_index* d_blk_row, const fx_index* d_blk_flags,
                          const fx_index* d_blk_lanes, const fx_index* d_rt_off,
                          const fx_index* d_rt_elem, const fx_index* d_rt_meta,
                          const fx_index* d_row_ptr, double* d_vals, double* d_res,
                          void* stream);
/* Navier-Stokes adjoint VJP: jt = J^T lam over the row-replaced Newton
 * Jacobian (lam zeroed on constrained rows, passed through there; bc_mask may
 * be NULL) and, when g is non-null, g[e] = lam~_e^T dR_e/dnu_e, the
 * per-element viscosity sensitivity.  inc_off/elem_inc from
 * fx_map_export_rowinc; work holds 15*num_elems doubles: the element column
 * contributions in row-incidence order, summed per DOF in ascending element
 * order.  The contract of Residual::applyParamJacT (femx/state/
 * Residual.hpp:55-59). */
int fx_th_vjp(fx_index ndof, fx_index num_elems, const double* d_xv, const double* d_yv,
              const fx_index* d_vconn3, const fx_index* d_dofs15, const double* d_state,
              double nu, const double* d_nu_e, const double* d_lam, const fx_index* d_bc_mask,
              const fx_index* d_inc_off, const fx_index* d_elem_inc, double* d_jt, double* d_g,
              double* d_work, void* stream);
/* fx_th_vjp without the work argument, for callers that cannot allocate
 * device memory (code under hipGraph stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in femx.  The element
 * columns go to a fixed 64 MiB workspace that the library keeps in device
 * memory, one per GPU, until the process ends: num_elems may be at most
 * FX_TH_VJP_NOALLOC_MAX_ELEMS (FX_ERR_ARG beyond; the 512 x 512 cavity,
 * 524 288 elements, fits).  Calls this function makes allocate nothing.
 * Every call on one GPU shares that GPU's workspace: calls that overlap in
 * time on one GPU (other streams, other host threads) race, so serialise
 * them.  Results are bitwise equal to fx_th_vjp's. */
#define FX_TH_VJP_NOALLOC_MAX_ELEMS 559240
int fx_th_vjp_noalloc(fx_index ndof, fx_index num_elems, const double* d_xv, const double* d_yv,
                      const fx_index* d_vconn3, const fx_index* d_dofs15, const double* d_state,
                      double nu, const double* d_nu_e, const double* d_lam,
                      const fx_index* d_bc_mask, const fx_index* d_inc_off,
                      const fx_index* d_elem_inc, double* d_jt, double* d_g, void* stream);

/* Patch plan of the single-pass VJPs (host): fx_th_vjp_patch (Taylor-Hood,
 * 15 DOFs per element) and fx_p1_vjp_patch (P1, 3).  Elements are ordered
 * along a Morton curve of their centroids (vertex ids vconn3 into xv/yv) and
 * cut into patches of at most max_patch (<= FX_TH_PATCH_MAX) consecutive
 * elements -- a patch also closes before an element that would give one of
 * its DOFs more than FX_TH_PATCH_RUN_MAX elements in it -- ascending element
 * id inside a patch.  A patch's local DOFs (ascending id, at most 4095) each
 * own a run of slots, one per element of the patch holding the DOF, in
 * element order (ldof_end: each run's end within the patch).
 * elem_ldof[S i + k] = j | r << 12 for local DOF k of the i-th element in
 * patch order, S = ldof_stride (dofs_per_elem rounded up to a multiple of
 * 4; the tail entries are unused): j its index among the patch's local DOFs,
 * r the element's rank in j's run; elem_vert[3 i + k] is vertex k of that
 * element.  A DOF held by one patch only is finished by it (ldof_part -1);
 * the others (on seams between patches, or in no element) are the seam
 * DOFs, summed from per-patch partials in patch order: ldof_part is the
 * partial's index, with FX_TH_PART_FIRST set on the DOF's first patch
 * (which adds lam's pass-through on a constrained DOF).
 * Array sizes: patch_off num_patches+1, patch_elem num_elems, elem_vert
 * 3*num_elems, elem_ldof ldof_stride*num_elems, ldof_off num_patches+1,
 * ldof_dof/ldof_part/ldof_end num_ldofs, seam_off num_seam+1, seam_dof
 * num_seam. */
#define FX_TH_PATCH_MAX 256
#define FX_TH_PATCH_RUN_MAX 16
#define FX_TH_PART_FIRST (1 << 30)
typedef struct fx_patches fx_patches;
typedef struct
{
  fx_index num_patches, num_ldofs, num_seam, num_parts;
  fx_index max_patch_ldofs; /* the most local DOFs of one patch */
  fx_index ldof_stride;     /* elem_ldof entries per element */
} fx_patches_info;
int fx_patches_create(fx_index ndof, fx_index num_elems, fx_index dofs_per_elem,
                      const fx_index* h_dofs, const fx_index* h_vconn3, fx_index num_vertices,
                      const double* h_xv, const double* h_yv, fx_index max_patch,
                      fx_patches** out);
int fx_patches_info_get(const fx_patches* patches, fx_patches_info* info);
int fx_patches_export(const fx_patches* patches, fx_index* h_patch_off, fx_index* h_patch_elem,
                      fx_index* h_elem_vert, unsigned short* h_elem_ldof, fx_index* h_ldof_off,
                      fx_index* h_ldof_dof, fx_index* h_ldof_part, unsigned short* h_ldof_end,
                      fx_index* h_seam_off, fx_index* h_seam_dof);
void fx_patches_destroy(fx_patches* patches);

/* The patch plan's arrays on the device, and the seam partials workspace
 * (num_parts doubles). */
struct fx_patch_plan
{
  fx_index              num_patches, num_seam, num_parts, max_patch_ldofs;
  const fx_index*       patch_off;
  const fx_index*       patch_elem;
  const fx_index*       elem_vert;
  const unsigned short* elem_ldof;
  const fx_index*       ldof_off;
  const fx_index*       ldof_dof;
  const fx_index*       ldof_part;
  const unsigned short* ldof_end;
  const fx_index*       seam_off;
  const fx_index*       seam_dof;
  double*               parts;
};

/* fx_th_vjp in one pass over element patches: a workgroup loads its
 * patch's local DOF values (state, masked adjoint) into LDS once, forms the
 * element columns (as fx_th_vjp's first phase) from them, sums each local
 * DOF's run in LDS in element order and writes jt for the DOFs it finishes;
 * a second launch sums the seam DOFs' partials in patch order.  Only the
 * seam partials go through device memory; the element DOFs and vertices
 * come from the plan (elem_ldof, ldof_dof, elem_vert).  Results are deterministic; they differ from
 * fx_th_vjp's only in the order of the per-DOF sums. */
int fx_th_vjp_patch(fx_index ndof, fx_index num_elems, const double* d_xv, const double* d_yv,
                    const double* d_state, double nu,
                    const double* d_nu_e, const double* d_lam,
                    const fx_index* d_bc_mask, const fx_patch_plan* h_plan, double* d_jt,
                    double* d_g, void* stream);

/* ======================================================================== */
/* GLS-stabilised P1/P1 Navier-Stokes time step (device): femx's
 * NavierOperator (femx/model/ns/Components.hpp:400-577) assembled with
 * respect to the next state, as assembly::assemble with a time operator does
 * (femx/assembly/CudaAssembly.hpp:366-417 on the device, Assembly.hpp:235-297
 * on the host).  hist is lag-major (num_hist x num_states; lags 0 and 1 are
 * read), nxt the next state; rho, mu, dt as NavierStokesModel's FluidParams
 * and time step; dofs9[9E] the map's element DOF table (NavierOperator
 * local order), the other tables from fx_map_gls_plan.  Writes res[num_states] =
 * -F + K nxt and, when vals is non-null, the CSR values of K (vals NULL is
 * assembleResidual, CudaAssembly.hpp:419-472).  Outputs are replaced.  Sums
 * in ascending element order with femx's operation order and no
 * contraction: equal to femx's host assembly bit for bit. */
int fx_gls_assemble(fx_index num_nodes, fx_index num_states, fx_index num_elems, fx_index nnz,
                    fx_index max_tile, const double* d_x, const double* d_y,
                    const fx_index* d_conn3, const fx_index* d_dofs9, const fx_index* d_node_rows,
                    double rho, double mu,
                    double dt, fx_index step, fx_index num_hist, const double* d_hist,
                    const double* d_nxt, const fx_index* d_inc_off, const fx_index* d_inc_elem,
                    const fx_index* d_inc_meta, const fx_index* d_blk_runs,
                    const fx_index* d_node_tile, double* d_vals, double* d_res, void* stream);

/* History VJP of the same operator (device): out = (dR/dhist_lag)^T adj for
 * one history lag, the product femx takes with Enzyme
 * (applyNavierHistJacT, femx/model/ns/NavierStokesModel.cpp:321-363;
 * device femx/model/ns/Components.cu:175-238).  Hand-derived reverse mode of
 * NavierOperator's residual through u, grad u, the extrapolated advection
 * velocity and tau(u); pressure entries are zero.  adj has num_states
 * entries (the residual space).  Per state DOF, a sum over its elements in
 * ascending element order.  node_rows, inc_off and inc_slot from
 * fx_map_gls_plan.  work (6 * num_elems
 * doubles: two per incidence slot) receives each element's columns for its
 * three nodes, formed from one evaluation of the quadrature-point partials
 * per element and stored at the element's slot in each node's incidence
 * list (inc_slot); a second pass sums each node's contiguous slots.  Calls
 * sharing a work buffer must not overlap. */
int fx_gls_hist_vjp(fx_index num_nodes, fx_index num_states, fx_index num_elems,
                    const double* d_x, const double* d_y, const fx_index* d_conn3,
                    const fx_index* d_dofs9, const fx_index* d_node_rows, double rho, double mu,
                    double dt, fx_index step, fx_index num_hist, fx_index lag,
                    const double* d_hist, const double* d_nxt, const double* d_adj,
                    const fx_index* d_inc_off, const fx_index* d_inc_slot, double* d_out,
                    double* d_work, void* stream);
/* Both history lags' VJPs in one pass over element patches: out_lag0 =
 * (dR/dhist_0)^T adj and out_lag1 = (dR/dhist_1)^T adj (either may be NULL;
 * lag 1 needs num_hist > 1), from one evaluation of the element partials.
 * Plan from fx_patches_create with the mesh nodes as DOFs (dofs = vconn3 =
 * conn3, coordinates x/y) and a partials workspace of 4 * num_parts doubles
 * in plan->parts.  A node inside one patch gets fx_gls_hist_vjp's sums bit
 * for bit; a seam node sums per-patch partials, so it may differ from
 * fx_gls_hist_vjp in the last bits. */
int fx_gls_hist_vjp_patch(fx_index num_nodes, fx_index num_states, fx_index num_elems,
                          const double* d_x, const double* d_y, const fx_index* d_node_rows,
                          double rho, double mu, double dt, fx_index step, fx_index num_hist,
                          const double* d_hist, const double* d_nxt, const double* d_adj,
                          const fx_patch_plan* h_plan, double* d_out_lag0, double* d_out_lag1,
                          void* stream);

/* ======================================================================== */
/* Batches (device).  Many P1 Poisson, Taylor-Hood and GLS P1/P1
 * Navier-Stokes instances of uneven size assembled and differentiated by
 * persistent launches over one work list per phase, optionally captured in a
 * hipGraph and replayed with one launch.  femx has no batching; each
 * instance's results are bitwise equal to its single-instance calls
 * (fx_p1_assemble with kappa_is_log = 1 and fx_p1_vjp_patch; fx_th_assemble
 * and fx_th_vjp_patch; fx_gls_assemble and fx_gls_hist_vjp_patch).  Instance
 * arrays are caller-owned device memory and must outlive the batch.  VJP
 * pointers (lam or adj, the outputs, the patch plan) may be NULL (an empty
 * plan) to batch assembly only. */
typedef struct
{
  fx_index      num_nodes, num_elems, nnz, max_row_len;
  const double* x;
  const double* y;
  const fx_index* conn;
  const double* m; /* log kappa per element */
  const double* f;
  const double* u;
  const fx_index* inc_off;
  const fx_index* inc; /* fx_map_export_rowinc3 */
  const fx_index* row_ptr;
  double*       vals;
  double*       res;
  const double* lam;
  const fx_index* bc_mask;
  double*       jt;
  double*       g;
  fx_patch_plan vjp; /* the VJP's patch plan (fx_p1_vjp_patch) */
} fx_p1_instance;

typedef struct
{
  fx_index      ndof, num_elems, num_blocks, max_block_nnz;
  const double* xv;
  const double* yv;
  const fx_index* vconn3;
  const fx_index* dofs15;
  const double* state;
  double        nu;
  const double* nu_e;
  const fx_index* inc_off;
  const fx_index* inc_elem;
  const fx_index* inc_meta;
  const fx_index* blk_row;
  const fx_index* blk_flags;
  const fx_index* row_ptr;
  double*       vals;
  double*       res;
  const double* lam;
  const fx_index* bc_mask;
  double*       jt;
  double*       g;
  const fx_index* blk_lanes; /* fx_map_th_lanes, or NULL for two lanes per item */
  fx_patch_plan vjp; /* the VJP's patch plan (fx_th_vjp_patch) */
  /* fx_map_th_rounds' tables (with blk_lanes), or NULL: the assembly phase
   * then runs fx_th_assemble_rounds' walks instead of fx_th_assemble's blocks */
  const fx_index* rt_off;
  const fx_index* rt_elem;
  const fx_index* rt_meta;
} fx_th_instance;

/* One GLS time step (fx_gls_assemble's arguments; vals NULL assembles the
 * residual only) and its history VJP over both lags (fx_gls_hist_vjp_patch's
 * arguments; out_lag1 needs num_hist > 1). */
typedef struct
{
  fx_index      num_nodes, num_states, num_elems, max_tile;
  const double* x;
  const double* y;
  const fx_index* conn3;
  const fx_index* dofs9;
  const fx_index* node_rows;
  double        rho, mu, dt;
  fx_index      step, num_hist;
  const double* hist;
  const double* nxt;
  const fx_index* inc_off;
  const fx_index* inc_elem;
  const fx_index* inc_meta;
  const fx_index* blk_runs;
  const fx_index* node_tile;
  double*       vals;
  double*       res;
  const double* adj;
  double*       out_lag0;
  double*       out_lag1;
  fx_patch_plan vjp; /* the VJP's patch plan (fx_patches_create with 9 DOFs per element) */
} fx_gls_instance;

typedef struct fx_batch fx_batch;
enum
{
  FX_BATCH_ASSEMBLE = 1,
  FX_BATCH_VJP      = 2,
  FX_BATCH_SERIAL   = 4 /* with either: launch the chains one after another on one stream */
};

/* grid_per_cu: persistent workgroups per compute unit (the launch grid is
 * compute units x grid_per_cu, at least one full residency), or 0 for one
 * workgroup per work item.  Copies the descriptors to the device. */
int fx_batch_create(fx_index num_p1, const fx_p1_instance* h_p1, fx_index num_th,
                    const fx_th_instance* h_th, fx_index num_gls, const fx_gls_instance* h_gls,
                    fx_index grid_per_cu, fx_batch** out);
/* Work items per phase (P1 assembly, TH assembly, P1 VJP patches, TH VJP
 * patches, TH VJP seam blocks, P1 VJP seam blocks, GLS assembly blocks, GLS
 * VJP patches, GLS VJP seam blocks) and the persistent grid size. */
int fx_batch_work(const fx_batch* batch, fx_index* h_counts9, fx_index* h_grid);
/* Launch the phases in `what` (FX_BATCH_ASSEMBLE | FX_BATCH_VJP) ordered
 * after earlier work on stream, and make later work on stream wait for
 * them.  The independent chains (P1 assembly, TH assembly, P1 VJP, TH VJP,
 * GLS assembly, GLS VJP) run concurrently on the batch's private streams
 * (default priority), or in that order on stream with FX_BATCH_SERIAL. */
int fx_batch_run(fx_batch* batch, int what, void* stream);
/* Capture the phases in `what` into a hipGraph (on a private stream), one
 * graph branch per chain (a single chain of nodes with FX_BATCH_SERIAL). */
int fx_batch_capture(fx_batch* batch, int what);
/* Replay the captured graph on stream (one launch). */
int fx_batch_launch(fx_batch* batch, void* stream);
void fx_batch_destroy(fx_batch* batch);
/* One hipGraph over several batches (e.g. one per instance-size class, each
 * with its own grid and Taylor-Hood row tile): every chain of every batch is
 * a branch of the graph, or, with FX_BATCH_SERIAL, every phase in turn over
 * the batches.  Replays on one graph, or on graphs sharing a
 * batch, must not overlap.  The batches must outlive the graph. */
typedef struct fx_batch_graph fx_batch_graph;
int fx_batch_graph_create(fx_index count, fx_batch* const* batches, int what,
                          fx_batch_graph** out);
int fx_batch_graph_launch(fx_batch_graph* graph, void* stream);
void fx_batch_graph_destroy(fx_batch_graph* graph);

#ifdef __cplusplus
}
#endif
#endif
