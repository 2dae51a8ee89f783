// Checker-only driver over the compiled femx reference sources.
//
// This file is test infrastructure.  It is compiled together with the
// reference's own hot-path sources (listed in oracle/Makefile) into
// oracle/_ref/libfemx_ref.so and is used only to write golden fixtures and to
// cross-check the C restatement in oracle/femx_oracle.c.  Nothing in the
// product links or loads it.
//
// Every entry point takes plain host arrays, builds the reference's own
// objects from them and calls the reference routine named in its comment.

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

#include <femx/assembly/Assembly.hpp>
#include <femx/assembly/AssemblyMap.hpp>
#include <femx/assembly/BoundaryMap.hpp>
#include <femx/fem/DofLayout.hpp>
#include <femx/fem/ElementQuadratureData.hpp>
#include <femx/fem/DirichletBC.hpp>
#include <femx/fem/FESpace.hpp>
#include <femx/fem/GmshReader.hpp>
#include <femx/fem/GaussQuadrature.hpp>
#include <femx/fem/Geometry.hpp>
#include <femx/fem/Mesh.hpp>
#include <femx/fem/elements/LagrangeTriangleP1.hpp>
#include <femx/linalg/CsrMatrix.hpp>
#include <femx/linalg/CsrPattern.hpp>
#include <femx/linalg/handler/MatrixHandler.hpp>
#include <femx/fem/MixedFESpace.hpp>
#include <femx/model/ns/Components.hpp>
#include <femx/model/ns/NavierStokesModel.hpp>
#include <femx/assembly/ConstrainedTimeResidual.hpp>
#include <femx/fem/ControlMap.hpp>
#include <femx/fem/DirichletControl.hpp>
#include <femx/state/TimeResidual.hpp>

#include "PoissonComponents.hpp"

using femx::Array;
using femx::Index;
using femx::Real;

namespace
{
thread_local std::string g_err;

template <class F>
int guarded(F&& fn)
{
  try
  {
    fn();
    return 0;
  }
  catch (const std::exception& e)
  {
    g_err = e.what();
    return 1;
  }
}

struct MapHolder
{
  femx::assembly::HostAssemblyMap map;
};

struct BcHolder
{
  femx::assembly::HostBoundaryMap map;
};

femx::HostCsrPattern makePattern(Index rows, Index cols, const Index* row_ptr,
                                 const Index* col_ind)
{
  femx::HostIndexVector rp(rows + 1);
  for (Index i = 0; i <= rows; ++i) rp[i] = row_ptr[i];
  const Index           nnz = row_ptr[rows];
  femx::HostIndexVector ci(nnz);
  for (Index k = 0; k < nnz; ++k) ci[k] = col_ind[k];
  return femx::HostCsrPattern(rows, cols, std::move(rp), std::move(ci));
}

} // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// femx::assembly::makeAssemblyMap (femx/assembly/AssemblyMap.cpp:54-176).
int ref_map_build(Index num_res, Index num_states, Index num_elems,
                  const Index* res_offsets, const Index* res_dofs,
                  const Index* state_offsets, const Index* state_dofs,
                  void** out)
{
  return guarded([&] {
    Array<Array<Index>> rows(num_elems), cols(num_elems);
    for (Index e = 0; e < num_elems; ++e)
    {
      for (Index p = res_offsets[e]; p < res_offsets[e + 1]; ++p) rows[e].push_back(res_dofs[p]);
      for (Index p = state_offsets[e]; p < state_offsets[e + 1]; ++p) cols[e].push_back(state_dofs[p]);
    }
    auto* h = new MapHolder{femx::assembly::makeAssemblyMap(num_res, num_states, rows, cols)};
    *out    = h;
  });
}

int ref_map_sizes(void* h, Index* nnz, Index* jac_len, Index* rows)
{
  auto& m  = static_cast<MapHolder*>(h)->map;
  *nnz     = m.pattern().nnz();
  *jac_len = m.view().jac_offsets[m.numElems()];
  *rows    = m.numRes();
  return 0;
}

int ref_map_export(void* h, Index* row_ptr, Index* col_ind, Index* jac_offsets, Index* jac_map)
{
  auto&       m = static_cast<MapHolder*>(h)->map;
  const auto  v = m.view();
  const auto& p = m.pattern();
  std::memcpy(row_ptr, p.rowPtrData(), sizeof(Index) * (p.rows() + 1));
  std::memcpy(col_ind, p.colIndData(), sizeof(Index) * p.nnz());
  std::memcpy(jac_offsets, v.jac_offsets, sizeof(Index) * (m.numElems() + 1));
  std::memcpy(jac_map, v.jac_map, sizeof(Index) * v.jac_offsets[m.numElems()]);
  return 0;
}

void ref_map_free(void* h) { delete static_cast<MapHolder*>(h); }

// femx::assembly::makeBoundaryMap (femx/assembly/BoundaryMap.cpp:59-139).
int ref_bc_build(Index rows, const Index* row_ptr, const Index* col_ind, Index num_bcs,
                 const Index* dofs, void** out)
{
  return guarded([&] {
    auto         pat = makePattern(rows, rows, row_ptr, col_ind);
    Array<Index> d(num_bcs);
    for (Index i = 0; i < num_bcs; ++i) d[i] = dofs[i];
    *out = new BcHolder{femx::assembly::makeBoundaryMap(d, pat)};
  });
}

Index ref_bc_num_col_entries(void* h)
{
  const auto v = static_cast<BcHolder*>(h)->map.view();
  return v.col_offsets[v.num_bcs];
}

int ref_bc_export(void* h, Index* bc_rows, Index* diag, Index* col_offsets, Index* col_entries,
                  Index* col_rows, Index* bc_mask)
{
  const auto v = static_cast<BcHolder*>(h)->map.view();
  std::memcpy(bc_rows, v.bc_rows, sizeof(Index) * v.num_bcs);
  std::memcpy(diag, v.diag_entries, sizeof(Index) * v.num_bcs);
  std::memcpy(col_offsets, v.col_offsets, sizeof(Index) * (v.num_bcs + 1));
  const Index nce = v.col_offsets[v.num_bcs];
  std::memcpy(col_entries, v.col_entries, sizeof(Index) * nce);
  std::memcpy(col_rows, v.col_rows, sizeof(Index) * nce);
  std::memcpy(bc_mask, v.bc_mask, sizeof(Index) * v.num_rows);
  return 0;
}

void ref_bc_free(void* h) { delete static_cast<BcHolder*>(h); }

// femx::assembly::replaceRows / replaceRes / zeroBoundary / prepareForwardSolve
// (femx/assembly/BoundaryMap.cpp:172-247).  `op`: 0 replaceRows(diag),
// 1 prepareForwardSolve, 2 replaceRes, 3 zeroBoundary.
int ref_bc_apply(Index op, Index rows, const Index* row_ptr, const Index* col_ind, Index num_bcs,
                 const Index* dofs, Real* vals, Real* rhs, const Real* bc_vals, const Real* state,
                 Real diag)
{
  return guarded([&] {
    auto         pat = makePattern(rows, rows, row_ptr, col_ind);
    Array<Index> d(num_bcs);
    for (Index i = 0; i < num_bcs; ++i) d[i] = dofs[i];
    auto               map = femx::assembly::makeBoundaryMap(d, pat);
    femx::HostCsrMatrix mat(pat);
    const Index        nnz = pat.nnz();
    for (Index k = 0; k < nnz; ++k) mat.valsData()[k] = vals ? vals[k] : 0.0;
    femx::HostVector r(rows), g(num_bcs), s(rows);
    for (Index i = 0; i < rows; ++i) r[i] = rhs ? rhs[i] : 0.0;
    for (Index i = 0; i < num_bcs; ++i) g[i] = bc_vals ? bc_vals[i] : 0.0;
    for (Index i = 0; i < rows; ++i) s[i] = state ? state[i] : 0.0;
    if (op == 0) femx::assembly::replaceRows(map, mat, diag);
    else if (op == 1) femx::assembly::prepareForwardSolve(map, mat, r, g);
    else if (op == 2) femx::assembly::replaceRes(map, s, g, r);
    else if (op == 3) femx::assembly::zeroBoundary(map, r.view());
    if (vals)
      for (Index k = 0; k < nnz; ++k) vals[k] = mat.valsData()[k];
    if (rhs)
      for (Index i = 0; i < rows; ++i) rhs[i] = r[i];
  });
}

// femx::linalg::MatrixHandler<HostCsrBackend>::matvec / matvecT
// (femx/linalg/handler/MatrixHandler.cpp:157-198).
int ref_spmv(Index transpose, Index rows, Index cols, const Index* row_ptr, const Index* col_ind,
             const Real* vals, const Real* x, Real* y, Real alpha, Real beta)
{
  return guarded([&] {
    auto                pat = makePattern(rows, cols, row_ptr, col_ind);
    femx::HostCsrMatrix mat(pat);
    for (Index k = 0; k < pat.nnz(); ++k) mat.valsData()[k] = vals[k];
    const Index      nin = transpose ? rows : cols;
    const Index      nout = transpose ? cols : rows;
    femx::HostVector xv(nin), yv(nout);
    for (Index i = 0; i < nin; ++i) xv[i] = x[i];
    for (Index i = 0; i < nout; ++i) yv[i] = y[i];
    femx::CpuContext                ctx;
    femx::linalg::HostMatrixHandler mh(ctx);
    if (transpose) mh.matvecT(mat, xv.view(), yv.view(), alpha, beta);
    else mh.matvec(mat, xv.view(), yv.view(), alpha, beta);
    for (Index i = 0; i < nout; ++i) y[i] = yv[i];
  });
}

// Explicit transpose through femx's host MatrixHandler::transpose
// (femx/linalg/handler/MatrixHandler.cpp:37-70 counting sort, values through
// the cached source-to-transpose map).  Writes t_row_ptr[cols+1],
// t_col_ind[nnz], t_vals[nnz].
int ref_transpose(Index rows, Index cols, const Index* row_ptr, const Index* col_ind,
                  const Real* vals, Index* t_row_ptr, Index* t_col_ind, Real* t_vals)
{
  return guarded([&] {
    auto                pat = makePattern(rows, cols, row_ptr, col_ind);
    femx::HostCsrMatrix mat(pat), out;
    for (Index k = 0; k < pat.nnz(); ++k) mat.valsData()[k] = vals[k];
    femx::CpuContext                ctx;
    femx::linalg::HostMatrixHandler mh(ctx);
    mh.transpose(mat, out);
    std::memcpy(t_row_ptr, out.pattern().rowPtrData(), sizeof(Index) * (cols + 1));
    std::memcpy(t_col_ind, out.pattern().colIndData(), sizeof(Index) * pat.nnz());
    std::memcpy(t_vals, out.valsData(), sizeof(Real) * pat.nnz());
  });
}

// Poisson (Laplace) assembly on a P1 triangle mesh through the reference's own
// pipeline: Mesh/FESpace/makeGeometry (femx/fem), makeElementQuadratureData
// with the order-2 triangle rule, makeAssemblyMap(DofLayout), and
// assembly::assemble with examples/poisson/PoissonComponents.hpp:35-64.
// Writes the CSR pattern (row_ptr size N+1, col_ind size nnz_cap), the CSR
// values and the residual K*state.
int ref_poisson_p1(Index num_nodes, const Real* x, const Real* y, Index num_elems,
                   const Index* conn, const Real* state, Index nnz_cap, Index* row_ptr,
                   Index* col_ind, Real* vals, Real* res)
{
  return guarded([&] {
    femx::fem::Mesh mesh(2);
    for (Index i = 0; i < num_nodes; ++i) mesh.addNode({x[i], y[i], 0.0});
    for (Index e = 0; e < num_elems; ++e)
      mesh.addElem({conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]},
                   femx::fem::Element::Shape::Triangle, 2, 0, 0, {});
    femx::fem::LagrangeTriangleP1 fe;
    femx::fem::FESpace            space(&mesh, &fe);
    space.setup();
    const auto geom = femx::fem::makeGeometry(mesh);
    const auto data = femx::fem::makeElementQuadratureData(
        space, femx::fem::GaussQuadrature::make(fe.referenceElement(), 2));
    const auto map = femx::assembly::makeAssemblyMap(femx::fem::DofLayout(space));
    femx::require(map.pattern().nnz() <= nnz_cap, "nnz capacity too small");
    femx::HostCsrMatrix jac(map.pattern());
    femx::HostVector    st(num_nodes), r;
    for (Index i = 0; i < num_nodes; ++i) st[i] = state[i];
    femx::CpuContext ctx;
    femx::assembly::assemble(
        femx::examples::poisson::PoissonComponents<femx::MemorySpace::Host>(data.view()), geom,
        map, st, r, jac, ctx);
    const auto& p = map.pattern();
    std::memcpy(row_ptr, p.rowPtrData(), sizeof(Index) * (p.rows() + 1));
    std::memcpy(col_ind, p.colIndData(), sizeof(Index) * p.nnz());
    std::memcpy(vals, jac.valsData(), sizeof(Real) * p.nnz());
    for (Index i = 0; i < num_nodes; ++i) res[i] = r[i];
  });
}


// GLS-stabilised P1/P1 Navier-Stokes time assembly on a triangle mesh through
// the reference's own objects: the mixed space of NavierStokes' makeSpace
// (femx/model/ns/Helper.cpp:40-50: vector P1 velocity, then scalar P1
// pressure), makeNavierData with the order-2 rule
// (femx/model/ns/Components.cpp:19-62, NavierStokesModel.cpp:33), the map of
// DofLayout(space) (NavierStokesModel.cpp constructor) and the host time
// assemble (femx/assembly/Assembly.hpp:235-297) of NavierOperator<Host>
// (femx/model/ns/Components.hpp:400-577) with respect to the next state.
// hist is lag-major (num_hist x num_states).  Writes the element DOF table
// (9 per element, local order), the CSR pattern and values, and the residual.
int ref_ns_gls(Index num_nodes, const Real* x, const Real* y, Index num_elems, const Index* conn,
               Real rho, Real mu, Real dt, Index step, Index num_hist, const Real* hist,
               const Real* nxt, Index nnz_cap, Index* elem_dofs, Index* row_ptr, Index* col_ind,
               Real* vals, Real* res)
{
  return guarded([&] {
    femx::fem::Mesh mesh(2);
    for (Index i = 0; i < num_nodes; ++i) mesh.addNode({x[i], y[i], 0.0});
    for (Index e = 0; e < num_elems; ++e)
      mesh.addElem({conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]},
                   femx::fem::Element::Shape::Triangle, 2, 0, 0, {});
    femx::fem::LagrangeTriangleP1 fe;
    femx::fem::FESpace            u_space(&mesh, &fe, mesh.dim());
    femx::fem::FESpace            p_space(&mesh, &fe);
    femx::fem::MixedFESpace       space;
    space.addField(u_space);
    space.addField(p_space);
    space.setup();
    const auto quad = femx::fem::GaussQuadrature::make(
        space.field(0).space().finiteElement().referenceElement(), 2);
    const auto data = femx::model::ns::makeNavierData(space.field(0).space(), quad);
    const auto map  = femx::assembly::makeAssemblyMap(femx::fem::DofLayout(space));
    femx::require(map.pattern().nnz() <= nnz_cap, "nnz capacity too small");
    const Index ns = map.numStates();
    femx::HostVector h(num_hist * ns), nx(ns), r;
    for (Index i = 0; i < num_hist * ns; ++i) h[i] = hist[i];
    for (Index i = 0; i < ns; ++i) nx[i] = nxt[i];
    femx::HostCsrMatrix jac(map.pattern());
    femx::CpuContext    ctx;
    const femx::model::ns::NavierOperator<femx::MemorySpace::Host> op(data.view(), {rho, mu}, dt);
    femx::assembly::assemble(op, step, num_hist, femx::state::VariableBlock::NextState, map, h,
                             nx, r, jac, ctx);
    const auto mv = map.view();
    for (Index e = 0; e < num_elems; ++e)
    {
      femx::require(mv.numStateDofs(e) == 9, "expected 9 DOFs per element");
      for (Index k = 0; k < 9; ++k) elem_dofs[9 * e + k] = mv.stateDof(e, k);
    }
    const auto& p = map.pattern();
    std::memcpy(row_ptr, p.rowPtrData(), sizeof(Index) * (p.rows() + 1));
    std::memcpy(col_ind, p.colIndData(), sizeof(Index) * p.nnz());
    std::memcpy(vals, jac.valsData(), sizeof(Real) * p.nnz());
    for (Index i = 0; i < ns; ++i) res[i] = r[i];
  });
}


// femx's Navier-Stokes model (femx/model/ns/NavierStokesModel.cpp) on a
// triangle mesh, decorated with Dirichlet constraints by
// assembly::ConstrainedTimeResidual (femx/assembly/ConstrainedTimeResidual.cpp)
// over a ControlMap (femx/fem/ControlMap.cpp:36-137) built from a
// DirichletControl (dofs, P entries), fixed DOFs with per-step values and a
// per-step linear time interpolation of the control levels.  At one step it
// evaluates res, assembleNext, applyJacT with respect to the parameters, and
// the pre-solve sequence of TimeIntegrator::solveStep
// (femx/state/TimeIntegrator.hpp:272-289): rhs = J nxt - res, then
// prepareLinearSolve.  Writes the dims (num_states, num_hist, num_param,
// nnz) first; arrays are sized by the caller from a first call with
// query = 1.
int ref_ns_constrained(int query, Index num_nodes, const Real* x, const Real* y, Index num_elems,
                       const Index* conn, Index nstep, Real dt, Real rho, Real mu,
                       Index num_ctr_dofs, const Index* ctr_dofs, Index num_ctr_params,
                       Index num_entries, const Index* ent_row, const Index* ent_col,
                       const Real* ent_wt, Index num_fixed, const Index* fixed_dofs,
                       const Real* fixed_vals, const Index* t_lower, const Index* t_upper,
                       const Real* t_upper_wt, Index ctr_off, Index num_prm, Index step,
                       const Real* hist, const Real* nxt, const Real* prm, const Real* adj,
                       Index* dims, Real* res, Real* asm_res, Real* asm_vals, Real* grad,
                       Real* prep_vals, Real* prep_rhs)
{
  return guarded([&] {
    femx::fem::Mesh mesh(2);
    for (Index i = 0; i < num_nodes; ++i) mesh.addNode({x[i], y[i], 0.0});
    for (Index e = 0; e < num_elems; ++e)
      mesh.addElem({conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]},
                   femx::fem::Element::Shape::Triangle, 2, 0, 0, {});
    femx::model::ns::FluidParams fluid;
    fluid.rho = rho;
    fluid.mu  = mu;
    femx::model::ns::NavierStokesModel model(std::move(mesh), nstep, dt, fluid);
    Array<Index> cd(num_ctr_dofs);
    for (Index k = 0; k < num_ctr_dofs; ++k) cd[k] = ctr_dofs[k];
    Array<femx::fem::DirichletControlMapEntry> ents;
    for (Index k = 0; k < num_entries; ++k) ents.push_back({ent_row[k], ent_col[k], ent_wt[k]});
    femx::fem::DirichletControl ctr(cd, num_ctr_params, ents);
    Array<Index>                fd(num_fixed);
    for (Index k = 0; k < num_fixed; ++k) fd[k] = fixed_dofs[k];
    femx::HostVector            fv(nstep * num_fixed);
    for (Index k = 0; k < nstep * num_fixed; ++k) fv[k] = fixed_vals[k];
    Array<femx::LinearInterpolation> tm(nstep);
    for (Index s = 0; s < nstep; ++s) tm[s] = {t_lower[s], t_upper[s], t_upper_wt[s]};
    auto control = femx::fem::makeControlMap(nstep, model.numStates(), ctr, fd, std::move(fv),
                                             tm, ctr_off, num_prm);
    femx::assembly::HostConstrainedTimeResidual cres(model.residual(), std::move(control));
    const auto dm = cres.dims();
    dims[0]       = dm.num_states;
    dims[1]       = dm.num_hist;
    dims[2]       = dm.num_param;
    dims[3]       = model.map().pattern().nnz();
    if (query) return;
    const Index ns = dm.num_states;
    femx::HostVector h(dm.num_hist * ns), nx(ns), pr(dm.num_param), aj(ns);
    for (Index i = 0; i < dm.num_hist * ns; ++i) h[i] = hist[i];
    for (Index i = 0; i < ns; ++i) nx[i] = nxt[i], aj[i] = adj[i];
    for (Index i = 0; i < dm.num_param; ++i) pr[i] = prm[i];
    const femx::state::HostTimeContext time{
        step, nx.view(), pr.view(), {h.data(), dm.num_hist, ns}};
    femx::CpuContext ctx;
    femx::HostVector out;
    cres.res(time, out, ctx);
    for (Index i = 0; i < ns; ++i) res[i] = out[i];
    femx::HostCsrMatrix jac(model.map().pattern());
    femx::HostVector    r2;
    cres.assembleNext(time, r2, jac, ctx);
    for (Index i = 0; i < ns; ++i) asm_res[i] = r2[i];
    std::memcpy(asm_vals, jac.valsData(), sizeof(Real) * dims[3]);
    femx::HostVector g;
    cres.applyJacT(time, femx::state::VariableBlock::Param, aj.view(), g, ctx);
    for (Index i = 0; i < dm.num_param; ++i) grad[i] = g[i];
    femx::HostVector rhs(ns);
    femx::linalg::HostMatrixHandler mh(ctx);
    femx::linalg::HostVectorHandler vh(ctx);
    mh.matvec(jac, nx.view(), rhs);
    vh.axpby(-1.0, r2.view(), 1.0, rhs.view());
    cres.prepareLinearSolve(time, jac, rhs, ctx);
    std::memcpy(prep_vals, jac.valsData(), sizeof(Real) * dims[3]);
    for (Index i = 0; i < ns; ++i) prep_rhs[i] = rhs[i];
  });
}


// femx::fem::GmshReader::read (femx/fem/GmshReader.cpp:463-538) of a 2-D
// triangle mesh.  With query != 0 it writes dims = {num_nodes,
// num_triangles, num_facets} only; otherwise it also writes the node
// coordinates x, y, the triangles' node ids conn (3 per triangle, in the
// mesh's element order) and, per boundary facet (2 nodes each), its
// physical tag and node ids.
int ref_gmsh_read(const char* path, int query, Index* dims, Real* x, Real* y, Index* conn,
                  Index* facet_ptag, Index* facet_nids)
{
  return guarded([&] {
    const femx::fem::Mesh mesh = femx::fem::GmshReader::read(path);
    femx::require(mesh.dim() == 2, "expected a 2-D mesh");
    dims[0] = mesh.numNodes();
    dims[1] = mesh.numElems();
    dims[2] = mesh.boundaryFacets().size();
    if (query) return;
    for (Index i = 0; i < mesh.numNodes(); ++i)
    {
      x[i] = mesh.node(i)[0];
      y[i] = mesh.node(i)[1];
    }
    for (Index e = 0; e < mesh.numElems(); ++e)
    {
      femx::require(mesh.elem(e).numNodes() == 3, "expected triangles only");
      const Index* n = mesh.elemNodeIds(e);
      for (int k = 0; k < 3; ++k) conn[3 * e + k] = n[k];
    }
    const auto& facets = mesh.boundaryFacets();
    for (Index f = 0; f < static_cast<Index>(facets.size()); ++f)
    {
      femx::require(facets[f].nids.size() == 2, "expected 2-node boundary facets");
      facet_ptag[f]       = facets[f].ptag;
      facet_nids[2 * f]     = facets[f].nids[0];
      facet_nids[2 * f + 1] = facets[f].nids[1];
    }
  });
}

// femx::fem::DirichletBC::addBoundary(space, ptag, value, time, comp)
// (femx/fem/DirichletBC.cpp:34-65) on the P1 space of the mesh read from
// `path`, with ncomp components per node: writes the constrained DOFs in the
// order femx adds them (count first; dofs may be null to query).
int ref_gmsh_bc_dofs(const char* path, Index ptag, Index ncomp, Index comp, Index* count,
                     Index* dofs)
{
  return guarded([&] {
    const femx::fem::Mesh         mesh = femx::fem::GmshReader::read(path);
    femx::fem::LagrangeTriangleP1 fe;
    femx::fem::FESpace            space(&mesh, &fe, ncomp);
    space.setup();
    femx::fem::DirichletBC bc;
    bc.addBoundary(space, ptag, 0.0, 0.0, comp);
    const auto& d = bc.dofs();
    *count        = static_cast<Index>(d.size());
    if (dofs)
      for (Index i = 0; i < *count; ++i) dofs[i] = d[i];
  });
}

} // extern "C"
