// C++ tests of the femx-shaped host interface (femx_amd/include/femx_hip),
// restating the reference's own unit tests against the HIP path.  Device
// tests report SKIP when no HIP device is present, as femx's CUDA tests do
// (tests/unit/backend/runCudaAssemblyTests.cu:139-143).  Exit status is the
// number of failures.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include <fstream>

#include "femx_hip/femx_hip.hpp"
#include "femx_hip/solver.hpp"
#include "femx_hip/time.hpp"

using namespace femx;
using IV = std::vector<Index>;
using RV = std::vector<Real>;

namespace
{
int g_fail = 0, g_pass = 0, g_skip = 0;

struct Skip
{
};

void run(const char* name, const std::function<void()>& fn)
{
  try
  {
    fn();
    std::printf("PASS %s\n", name);
    ++g_pass;
  }
  catch (const Skip&)
  {
    std::printf("SKIP %s\n", name);
    ++g_skip;
  }
  catch (const std::exception& e)
  {
    std::printf("FAIL %s: %s\n", name, e.what());
    ++g_fail;
  }
}

void expect(bool c, const std::string& what)
{
  if (!c) throw std::runtime_error("check failed: " + what);
}

bool near(const RV& a, const RV& b, Real tol = 1e-12)
{
  if (a.size() != b.size()) return false;
  for (std::size_t i = 0; i < a.size(); ++i)
    if (std::abs(a[i] - b[i]) > tol * std::max({1.0, std::abs(a[i]), std::abs(b[i])})) return false;
  return true;
}

void needDevice()
{
  if (!HipContext::available()) throw Skip{};
}

std::shared_ptr<DeviceCsrPattern> devPattern(const HostCsrPattern& h, HipContext& ctx)
{
  return std::make_shared<DeviceCsrPattern>(h, ctx);
}

// Fixture arrays written by the Python harness as raw little-endian files
// <dir>/<name>.f64 and <name>.i32.
std::string g_fixture_dir;

template <class T>
std::vector<T> readRaw(const std::string& name, const char* ext)
{
  if (g_fixture_dir.empty()) throw Skip{};
  std::ifstream in(g_fixture_dir + "/" + name + ext, std::ios::binary | std::ios::ate);
  if (!in) throw std::runtime_error("missing fixture array " + name);
  const std::streamsize n = in.tellg();
  std::vector<T>        v(static_cast<std::size_t>(n) / sizeof(T));
  in.seekg(0);
  in.read(reinterpret_cast<char*>(v.data()), n);
  return v;
}
RV readF64(const std::string& name) { return readRaw<Real>(name, ".f64"); }
IV readI32(const std::string& name) { return readRaw<Index>(name, ".i32"); }

// femx's LinearStateSolver::solve (femx/state/StateSolver.hpp:90-107) over
// the device contracts: R(0, m) and dR/du at u = 0, one linear solve.
struct LinearStateSolver
{
  const state::Residual& res;
  DeviceCsrMatrix&       jac;
  linalg::LinearSolver&  solver;
  HipContext&            ctx;
  state::Dimensions      dims;
  DeviceVector           zero, res_vec, rhs;

  LinearStateSolver(const state::Residual& r, DeviceCsrMatrix& j, linalg::LinearSolver& s,
                    HipContext& c)
      : res(r), jac(j), solver(s), ctx(c), dims(r.dims()), zero(dims.num_states),
        res_vec(dims.num_res), rhs(dims.num_res)
  {
    expect(dims.num_res == dims.num_states, "LinearStateSolver requires square residual dimensions");
    linalg::HipVectorHandler(ctx).zero(view(zero));
  }
  void solve(const DeviceVector& prm, DeviceVector& state)
  {
    linalg::HipVectorHandler vh(ctx);
    res.res(zero, prm, res_vec, ctx);
    vh.axpby(-1.0, view(res_vec), 0.0, view(rhs));
    res.assembleStateJac(zero, prm, jac, ctx);
    linalg::HipMatrixHandler(ctx).finalize(jac);
    solver.solve(jac, rhs, state, ctx);
    ctx.sync();
  }
};

// femx's ReducedFunctional (femx/inverse/ReducedFunctional.hpp:60-117) with
// a tracking objective J(u, m) = |u - u_d|^2 / 2 + beta |m|^2 / 2 whose
// stateGrad and paramGrad are formed on the device: gradAt assembles the
// state Jacobian, solves the adjoint with solveT, and subtracts
// applyParamJacT(u, m, lambda) from dJ/dm.
struct ReducedFunctional
{
  LinearStateSolver&    state_solver;
  const state::Residual& res;
  DeviceCsrMatrix&      adj_jac;
  linalg::LinearSolver& adj_solver;
  HipContext&           ctx;
  const DeviceVector&   ud;
  Real                  beta;
  state::Dimensions     dims;
  DeviceVector          state, state_grad, adj, prm_grad, res_prm_adj;

  ReducedFunctional(LinearStateSolver& ss, DeviceCsrMatrix& aj, linalg::LinearSolver& as,
                    const DeviceVector& target, Real b)
      : state_solver(ss), res(ss.res), adj_jac(aj), adj_solver(as), ctx(ss.ctx), ud(target),
        beta(b), dims(res.dims()), state(dims.num_states), state_grad(dims.num_states),
        adj(dims.num_res), prm_grad(dims.num_param), res_prm_adj(dims.num_param)
  {
  }
  Real value(const DeviceVector& prm)
  {
    state_solver.solve(prm, state);
    const RV u = state.toHost(ctx), d = ud.toHost(ctx), m = prm.toHost(ctx);
    Real     j = 0.0;
    for (std::size_t i = 0; i < u.size(); ++i) j += 0.5 * (u[i] - d[i]) * (u[i] - d[i]);
    for (Real v : m) j += 0.5 * beta * v * v;
    return j;
  }
  void grad(const DeviceVector& prm, DeviceVector& out)
  {
    state_solver.solve(prm, state);
    linalg::HipVectorHandler vh(ctx);
    vh.copy(view(state), view(state_grad)); // stateGrad: u - u_d
    vh.axpby(-1.0, view(ud), 1.0, view(state_grad));
    res.assembleStateJac(state, prm, adj_jac, ctx);
    linalg::HipMatrixHandler(ctx).finalize(adj_jac);
    adj_solver.solveT(adj_jac, state_grad, adj, ctx);
    vh.axpby(beta, view(prm), 0.0, view(prm_grad)); // paramGrad: beta m
    res.applyParamJacT(state, prm, adj, res_prm_adj, ctx);
    vh.axpby(-1.0, view(res_prm_adj), 1.0, view(prm_grad));
    vh.copy(view(prm_grad), out);
    ctx.sync();
  }
};

HostCsrPattern dense3()
{
  HostCsrPattern p;
  p.rows = p.cols = 3;
  p.row_ptr       = {0, 3, 6, 9};
  p.col_ind       = {0, 1, 2, 0, 1, 2, 0, 1, 2};
  return p;
}
} // namespace

int main(int argc, char** argv)
{
  if (argc > 1) g_fixture_dir = argv[1];
  // tests/unit/backend/runBackendAssemblyTests.cpp:162-188
  run("rectangularMapBuildsExactCsrMapping", [] {
    auto map = assembly::makeAssemblyMap(2, 2, {{0, 1}, {1}}, {{0}, {0, 1}});
    IV   jo, jm;
    map.jacMap(jo, jm);
    expect(map.numElems() == 2 && map.pattern().nnz() == 3, "sizes");
    expect(map.pattern().row_ptr == IV({0, 1, 3}) && map.pattern().col_ind == IV({0, 0, 1}), "pattern");
    expect(jo == IV({0, 2, 4}) && jm == IV({0, 1, 1, 2}), "jac_map");
  });

  // tests/unit/linalg/runLinalgTests.cpp:235-270
  run("assemblyMapBuildsSharedElementSparsity", [] {
    auto map = assembly::makeAssemblyMap(3, 3, {{0, 1}, {1, 2}}, {{0, 1}, {1, 2}});
    IV   jo, jm;
    map.jacMap(jo, jm);
    expect(map.pattern().row_ptr == IV({0, 2, 5, 7}), "row_ptr");
    expect(map.pattern().col_ind == IV({0, 1, 0, 1, 2, 1, 2}), "col_ind");
    expect(jo == IV({0, 4, 8}) && jm == IV({0, 1, 2, 3, 3, 4, 5, 6}), "jac_map");
  });

  run("assemblyMapRejectsOutOfRangeDofs", [] {
    bool threw = false;
    try
    {
      assembly::makeAssemblyMap(2, 2, {{0, 2}}, {{0, 1}});
    }
    catch (const std::runtime_error& e)
    {
      threw = std::string(e.what()).find("out of range") != std::string::npos;
    }
    expect(threw, "residual DOF out of range throws");
  });

  // tests/unit/backend/runBackendStateTests.cpp:80-114
  run("boundaryRowsAndForwardEliminationStayDistinct", [] {
    needDevice();
    HipContext ctx;
    auto       hp  = dense3();
    auto       bm  = assembly::makeBoundaryMap({1}, hp);
    assembly::DeviceBoundaryMap map;
    assembly::copy(bm, map, ctx);
    auto       dp  = devPattern(hp, ctx);
    const RV   base{4.0, 1.0, 2.0, 3.0, 5.0, 6.0, 7.0, 8.0, 9.0};
    DeviceCsrMatrix auth(dp);
    auth.vals().assign(base, ctx);
    assembly::replaceRows(map, auth, 1.0, ctx);
    expect(near(auth.vals().toHost(ctx), {4, 1, 2, 0, 1, 0, 7, 8, 9}), "replaceRows(1)");
    DeviceCsrMatrix hist(dp);
    hist.vals().assign(base, ctx);
    assembly::replaceRows(map, hist, 0.0, ctx);
    expect(near(hist.vals().toHost(ctx), {4, 1, 2, 0, 0, 0, 7, 8, 9}), "replaceRows(0)");
    DeviceVector res(RV{10, 20, 30}, ctx), st(RV{4, 7, 9}, ctx), g(RV{2.5}, ctx);
    assembly::replaceRes(map, st, g, res, ctx);
    expect(near(res.toHost(ctx), {10, 4.5, 30}), "replaceRes");
    DeviceCsrMatrix solve(dp);
    solve.vals().assign(base, ctx);
    DeviceVector rhs(RV{10, 20, 30}, ctx), g2(RV{2.0}, ctx);
    assembly::prepareForwardSolve(map, solve, rhs, g2, ctx);
    expect(near(solve.vals().toHost(ctx), {4, 0, 2, 0, 1, 0, 7, 0, 9}), "forward values");
    expect(near(rhs.toHost(ctx), {8, 2, 14}), "forward rhs");
  });

  // tests/unit/backend/runCudaLinalgTests.cu:70-230
  run("persistentHipCsrOps", [] {
    needDevice();
    HipContext     ctx;
    HostCsrPattern hp;
    hp.rows = 3, hp.cols = 4, hp.row_ptr = {0, 2, 4, 7}, hp.col_ind = {0, 2, 1, 3, 0, 2, 3};
    DeviceCsrMatrix A(devPattern(hp, ctx));
    A.vals().assign({2.0, -1.0, 3.0, 4.0, -2.0, 5.0, 1.0}, ctx);
    linalg::HipMatrixHandler mh(ctx);
    linalg::HipVectorHandler vh(ctx);
    DeviceVector x(RV{1, 2, 3, 4}, ctx), y(3), nrm(1);
    mh.matvec(A, x, y);
    expect(near(y.toHost(ctx), {-1, 22, 17}), "rectangular CSR apply");
    vh.dot(x, x, nrm);
    expect(near(nrm.toHost(ctx), {30}), "squared norm");
    DeviceVector a(RV{1, 2, 3}, ctx);
    vh.axpby(-2.0, a, 0.5, y);
    expect(near(y.toHost(ctx), {-2.5, 7, 2.5}), "axpby");
    DeviceVector tin(RV{2, -1, 0.5}, ctx), tout(4);
    mh.matvecT(A, tin, view(tout));
    expect(near(tout.toHost(ctx), {3, -3, 0.5, -3.5}), "transposed CSR apply");
    A.vals().assign({-1.0, 2.0, 0.5, -3.0, 4.0, 1.0, -2.0}, ctx);
    mh.matvec(A, x, y);
    expect(near(y.toHost(ctx), {5, -11, -1}), "updated CSR values");
    bool rejected = false;
    try
    {
      mh.matvec(A, x, DeviceVectorView(x.data() + 1, 3));
    }
    catch (const std::runtime_error&)
    {
      rejected = true;
    }
    expect(rejected, "CSR apply rejects overlapping views");
  });

  // tests/unit/examples/runPoissonAssemblyTests.cpp:16-60 on P1 triangles
  run("poissonUsesMappedGraphAndBoundaryRows", [] {
    needDevice();
    HipContext ctx;
    const int  n = 4, N = (n + 1) * (n + 1), E = 2 * n * n;
    RV         x(N), y(N);
    IV         conn(3 * E);
    check(fx_mesh_p1(n, x.data(), y.data(), conn.data()));
    std::vector<IV> dofs(E);
    for (int e = 0; e < E; ++e) dofs[e] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
    auto hmap = assembly::makeAssemblyMap(N, N, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    IV bcd;
    for (int i = 0; i < N; ++i)
      if (x[i] == 0.0 || x[i] == 1.0 || y[i] == 0.0 || y[i] == 1.0) bcd.push_back(i);
    auto hb = assembly::makeBoundaryMap(bcd, hmap.pattern());
    assembly::DeviceBoundaryMap bmap;
    assembly::copy(hb, bmap, ctx);
    assembly::DeviceGeometry geom;
    geom.num_nodes = N, geom.num_elems = E, geom.nodes_per_elem = 3;
    geom.x.assign(x, ctx), geom.y.assign(y, ctx), geom.conn.assign(conn, ctx);
    DeviceVector f(RV(N, 1.0), ctx), res(N), rhs(N), bcv(RV(bcd.size(), 0.25), ctx);
    DeviceCsrMatrix mat(map.pattern);
    assembly::PoissonP1 op;
    op.f = f.data();
    assembly::assemble(op, geom, map, nullptr, res, &mat, ctx);
    linalg::HipVectorHandler vh(ctx);
    vh.axpby(-1.0, res, 0.0, rhs);
    assembly::prepareForwardSolve(bmap, mat, rhs, bcv, ctx);
    const RV v  = mat.vals().toHost(ctx);
    const RV r  = rhs.toHost(ctx);
    const auto& rp = hmap.pattern().row_ptr;
    const auto& ci = hmap.pattern().col_ind;
    bool interior = false;
    for (Index row : bcd)
    {
      expect(std::abs(r[row] - 0.25) < 1e-12, "constrained rhs");
      for (Index k = rp[row]; k < rp[row + 1]; ++k)
        expect(v[k] == (ci[k] == row ? 1.0 : 0.0), "identity row");
    }
    for (int i = 0; i < N; ++i)
    {
      expect(std::isfinite(r[i]), "finite rhs");
      if (!hb.bc_mask[i] && r[i] > 0.0) interior = true;
    }
    expect(interior, "interior forcing");
  });

  // the stencil-block kernel that assemble() picks on a structured mesh
  // against the row-per-lane entry point, bitwise
  run("poissonStencilAssemblyEqualsRowKernel", [] {
    needDevice();
    HipContext ctx;
    const int  n = 300, N = (n + 1) * (n + 1), E = 2 * n * n;
    RV         x(N), y(N);
    IV         conn(3 * E);
    check(fx_mesh_p1(n, x.data(), y.data(), conn.data()));
    std::vector<IV> dofs(E);
    for (int e = 0; e < E; ++e) dofs[e] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
    auto hmap = assembly::makeAssemblyMap(N, N, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    expect(map.p1_stencil_frac > 0.5 && map.p1_stencil_frac < 1.0, "stencil blocks found");
    assembly::DeviceGeometry geom;
    geom.num_nodes = N, geom.num_elems = E, geom.nodes_per_elem = 3;
    geom.x.assign(x, ctx), geom.y.assign(y, ctx), geom.conn.assign(conn, ctx);
    RV kh(E), fh(N), uh(N);
    for (int e = 0; e < E; ++e) kh[e] = 0.3 * std::sin(0.61 * e);
    for (int i = 0; i < N; ++i) fh[i] = std::cos(0.17 * i), uh[i] = std::sin(0.29 * i);
    DeviceVector kap(kh, ctx), f(fh, ctx), u(uh, ctx), r1(N), r2(N);
    DeviceCsrMatrix m1(map.pattern), m2(map.pattern);
    assembly::PoissonP1 op;
    op.kappa = kap.data(), op.log_kappa = true, op.f = f.data();
    assembly::assemble(op, geom, map, &u, r1, &m1, ctx);
    check(fx_p1_assemble(N, E, map.info.nnz, map.info.max_row_len, geom.x.data(), geom.y.data(),
                         geom.conn.data(), kap.data(), 1, f.data(), u.data(), map.inc_off.data(),
                         map.inc.data(), map.pattern->rowPtrData(), m2.valsData(), r2.data(),
                         ctx.raw()));
    expect(m1.vals().toHost(ctx) == m2.vals().toHost(ctx), "bitwise equal values");
    expect(r1.toHost(ctx) == r2.toHost(ctx), "bitwise equal residuals");
  });

  run("taylorHoodAssemblyIsDeterministicAndEqualsLaneMapKernel", [] {
    needDevice();
    HipContext ctx;
    const int  n = 9, E = 2 * n * n, Nv = (2 * n + 1) * (2 * n + 1), Np = (n + 1) * (n + 1);
    RV         x2(Nv), y2(Nv);
    IV         c6(6 * E), v3(3 * E), d15(15 * E);
    check(fx_mesh_p2(n, x2.data(), y2.data(), c6.data(), v3.data()));
    check(fx_th_dofs(n, c6.data(), v3.data(), d15.data()));
    RV xv(Np), yv(Np);
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i)
      {
        xv[j * (n + 1) + i] = x2[2 * j * (2 * n + 1) + 2 * i];
        yv[j * (n + 1) + i] = y2[2 * j * (2 * n + 1) + 2 * i];
      }
    const Index     ndof = 2 * Nv + Np;
    std::vector<IV> dofs(E);
    for (int e = 0; e < E; ++e) dofs[e] = IV(d15.begin() + 15 * e, d15.begin() + 15 * e + 15);
    auto hmap = assembly::makeAssemblyMap(ndof, ndof, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    assembly::DeviceTHGeometry geom;
    geom.num_elems = E, geom.ndof = ndof;
    geom.xv.assign(xv, ctx), geom.yv.assign(yv, ctx), geom.vconn3.assign(v3, ctx);
    geom.dofs15.assign(d15, ctx);
    RV s(ndof);
    for (Index i = 0; i < ndof; ++i) s[i] = std::sin(0.37 * i);
    DeviceVector st(s, ctx), r1(ndof), r2(ndof), r3(ndof);
    DeviceCsrMatrix m1(map.pattern), m2(map.pattern);
    assembly::TaylorHoodNS op;
    assembly::assemble(op, geom, map, st, r1, &m1, ctx);
    assembly::assemble(op, geom, map, st, r2, &m2, ctx);
    assembly::assemble(op, geom, map, st, r3, nullptr, ctx);
    const RV a = m1.vals().toHost(ctx), b = m2.vals().toHost(ctx);
    expect(a == b, "bitwise equal Jacobians");
    expect(r1.toHost(ctx) == r2.toHost(ctx) && r1.toHost(ctx) == r3.toHost(ctx), "bitwise equal residuals");
    // assembly::assemble runs over the round tables; fx_th_assemble over the
    // lane map gives the same bits
    DeviceCsrMatrix m4(map.pattern);
    DeviceVector    r4(ndof);
    check(fx_th_assemble(geom.ndof, geom.num_elems, map.num_blocks, map.max_block_nnz,
                         geom.xv.data(), geom.yv.data(), geom.vconn3.data(), geom.dofs15.data(),
                         st.data(), op.nu, op.nu_e, map.inc_off.data(), map.inc_elem.data(),
                         map.inc_meta.data(), map.blk_row.data(), map.blk_flags.data(),
                         map.blk_lanes.data(), map.pattern->rowPtrData(), m4.valsData(), r4.data(),
                         ctx.raw()));
    expect(m4.vals().toHost(ctx) == a && r4.toHost(ctx) == r1.toHost(ctx),
           "round tables bitwise equal to the lane-map kernel");
  });

  // tests/unit/model/ns/runNavierOperatorTests.cpp:223-290 (row assembly
  // invariants) and :292-382 (history VJP against central differences) on
  // the GLS operator, with the reference tests' state pattern (fillStates)
  run("navierRowAssemblyAndHistoryVjp", [] {
    needDevice();
    HipContext  ctx;
    const int   nx = 4, ny = 2, N = (nx + 1) * (ny + 1), E = 2 * nx * ny;
    const Index ns = 3 * N;
    RV          x(N), y(N);
    IV          conn;
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i)
      {
        x[j * (nx + 1) + i] = 2.0 * i / nx;
        y[j * (nx + 1) + i] = 1.0 * j / ny;
      }
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
      {
        const int a = j * (nx + 1) + i, b = a + 1, c = b + nx + 1, d = a + nx + 1;
        conn.insert(conn.end(), {a, b, c, a, c, d});
      }
    std::vector<IV> dofs(E);
    for (int e = 0; e < E; ++e)
    {
      const int* v = &conn[3 * e];
      dofs[e]      = {2 * v[0], 2 * v[0] + 1, 2 * v[1], 2 * v[1] + 1, 2 * v[2], 2 * v[2] + 1,
                      2 * N + v[0], 2 * N + v[1], 2 * N + v[2]};
    }
    auto hmap = assembly::makeAssemblyMap(ns, ns, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    assembly::DeviceNavierGeometry geom;
    assembly::makeNavierGeometry(x, y, conn, hmap, geom, ctx);
    model::ns::NavierOperator op{&geom, {1.2, 0.03}, 0.05};
    RV h(2 * ns), nx_(ns), zero(ns, 0.0);
    for (Index i = 0; i < ns; ++i)
    {
      h[i]      = 0.01 * (i + 1);
      h[ns + i] = -0.005 * (i + 1);
      nx_[i]    = 0.02 - 0.003 * i;
    }
    DeviceVector    hist(h, ctx), nxt(nx_, ctx), nz(zero, ctx), res(ns), res0(ns), resh(ns);
    DeviceCsrMatrix jac(map.pattern), jac0(map.pattern), jach(map.pattern);
    const auto      next = state::VariableBlock::NextState;
    assembly::assemble(op, 1, 2, next, map, hist, nxt, res, &jac, ctx);
    assembly::assemble(op, 1, 2, next, map, hist, nz, res0, &jac0, ctx);
    assembly::assemble(op, 1, 2, state::VariableBlock::hist(1), map, hist, nxt, resh, &jach, ctx);
    const RV K = jac.vals().toHost(ctx), r = res.toHost(ctx), r0 = res0.toHost(ctx);
    expect(K == jac0.vals().toHost(ctx), "the Jacobian does not depend on the next state");
    RV Kx(ns, 0.0);
    const auto& pat = hmap.pattern();
    for (Index i = 0; i < ns; ++i)
      for (Index k = pat.row_ptr[i]; k < pat.row_ptr[i + 1]; ++k) Kx[i] += K[k] * nx_[pat.col_ind[k]];
    RV delta(ns);
    for (Index i = 0; i < ns; ++i) delta[i] = Kx[i] + r0[i];
    expect(near(r, delta, 1e-11), "res(nxt) = K nxt + res(0)");
    const RV zh = jach.vals().toHost(ctx);
    expect(resh.toHost(ctx) == r, "history block: same residual");
    for (Real v : zh) expect(v == 0.0, "history block: zero Jacobian");
    bool threw = false;
    try
    {
      assembly::assemble(op, 1, 2, state::VariableBlock::Param, map, hist, nxt, res, nullptr, ctx);
    }
    catch (const std::runtime_error&)
    {
      threw = true;
    }
    expect(threw, "the parameter block is rejected");
    RV dir(ns), adj(ns);
    for (Index i = 0; i < ns; ++i)
    {
      dir[i] = 0.2 - 0.01 * i;
      adj[i] = -0.15 + 0.013 * i;
    }
    DeviceVector a(adj, ctx), jt;
    for (Index lag = 0; lag < 2; ++lag)
    {
      const Real eps = 1.0e-6;
      Real       fdv[2];
      for (int sgn = 0; sgn < 2; ++sgn)
      {
        RV hp = h;
        for (Index i = 0; i < ns; ++i) hp[lag * ns + i] += (sgn == 0 ? eps : -eps) * dir[i];
        DeviceVector hd(hp, ctx), rr(ns);
        assembly::assemble(op, 1, 2, next, map, hd, nxt, rr, nullptr, ctx);
        const RV rv = rr.toHost(ctx);
        Real     d  = 0.0;
        for (Index i = 0; i < ns; ++i) d += rv[i] * adj[i];
        fdv[sgn] = d;
      }
      const Real fd = (fdv[0] - fdv[1]) / (2.0 * eps);
      model::ns::applyNavierHistJacT(op, 1, 2, lag, hist, nxt, a, jt, ctx);
      const RV jv  = jt.toHost(ctx);
      Real     vjp = 0.0;
      for (Index i = 0; i < ns; ++i) vjp += dir[i] * jv[i];
      expect(std::abs(vjp - fd) <= 2.0e-6 * std::max({1.0, std::abs(vjp), std::abs(fd)}),
             "history VJP against central differences");
    }
  });

  // the finite-difference procedure of tests/unit/inverse/runInverseTests.cpp:192-234
  // on the Residual contract, and the adjoint identity against the assembled Jacobian
  run("poissonLogKappaResidualGradients", [] {
    needDevice();
    HipContext ctx;
    const int  n = 4, N = (n + 1) * (n + 1), E = 2 * n * n;
    RV         x(N), y(N);
    IV         conn(3 * E);
    check(fx_mesh_p1(n, x.data(), y.data(), conn.data()));
    std::vector<IV> dofs(E);
    for (int e = 0; e < E; ++e) dofs[e] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
    auto hmap = assembly::makeAssemblyMap(N, N, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    IV bcd;
    for (int i = 0; i < N; ++i)
      if (x[i] == 0.0 || y[i] == 0.0) bcd.push_back(i);
    auto hb = assembly::makeBoundaryMap(bcd, hmap.pattern());
    assembly::DeviceBoundaryMap bmap;
    assembly::copy(hb, bmap, ctx);
    assembly::DeviceGeometry geom;
    geom.num_nodes = N, geom.num_elems = E, geom.nodes_per_elem = 3;
    geom.x.assign(x, ctx), geom.y.assign(y, ctx), geom.conn.assign(conn, ctx);
    RV hm(E), hu(N), hl(N), hf(N);
    for (int e = 0; e < E; ++e) hm[e] = 0.4 * std::sin(1.3 * e);
    for (int i = 0; i < N; ++i) hu[i] = std::cos(0.7 * i), hl[i] = std::sin(0.9 * i + 0.2), hf[i] = 1.0 + 0.1 * i;
    DeviceVector m(hm, ctx), u(hu, ctx), lam(hl, ctx), f(hf, ctx), g(RV(bcd.size(), 0.5), ctx);
    state::PoissonLogKappaResidual R(geom, map, &bmap, &g, f.data());
    DeviceVector grad(E), jt(N), r(N);
    R.applyParamJacT(u, m, lam, grad, ctx);
    R.applyStateJacT(u, m, lam, jt, ctx);
    const RV hg = grad.toHost(ctx), hjt = jt.toHost(ctx);
    auto lamDotR = [&](const RV& mm) {
      DeviceVector dm(mm, ctx);
      R.res(u, dm, r, ctx);
      const RV hr = r.toHost(ctx);
      double s = 0.0;
      for (int i = 0; i < N; ++i) s += hl[i] * hr[i];
      return s;
    };
    const double h = 1e-6;
    for (int e : {0, 5, 17, E - 1})
    {
      RV mp = hm, mm = hm;
      mp[e] += h, mm[e] -= h;
      const double fd = (lamDotR(mp) - lamDotR(mm)) / (2 * h);
      expect(std::abs(hg[e] - fd) <= 1e-6 * std::max(1.0, std::abs(fd)), "d(lam.R)/dm_e by central differences");
    }
    DeviceCsrMatrix J(map.pattern);
    R.assembleStateJac(u, m, J, ctx);
    DeviceVector ref(N);
    linalg::HipMatrixHandler(ctx).matvecT(J, lam, view(ref));
    const RV href = ref.toHost(ctx);
    double err = 0.0, nrm = 0.0;
    for (int i = 0; i < N; ++i) err = std::max(err, std::abs(hjt[i] - href[i])), nrm = std::max(nrm, std::abs(href[i]));
    expect(err <= 1e-12 * nrm, "J^T lam equals the assembled row-replaced Jacobian's transpose apply");
  });


  // Residual contract: dims, hostPattern, pattern; and femx's
  // ReducedFunctional::gradAt over it (ReducedFunctional.hpp:94-117) against
  // central differences of the reduced functional, on the GPU
  run("reducedFunctionalGradientOverTheResidualContract", [] {
    needDevice();
    HipContext ctx;
    const int  n = 6, N = (n + 1) * (n + 1), E = 2 * n * n;
    RV         x(N), y(N);
    IV         conn(3 * E);
    check(fx_mesh_p1(n, x.data(), y.data(), conn.data()));
    std::vector<IV> dofs(E);
    for (int e = 0; e < E; ++e) dofs[e] = {conn[3 * e], conn[3 * e + 1], conn[3 * e + 2]};
    auto hmap = assembly::makeAssemblyMap(N, N, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    IV bcd;
    for (int i = 0; i < N; ++i)
      if (x[i] == 0.0 || x[i] == 1.0 || y[i] == 0.0 || y[i] == 1.0) bcd.push_back(i);
    auto hb = assembly::makeBoundaryMap(bcd, hmap.pattern());
    assembly::DeviceBoundaryMap bmap;
    assembly::copy(hb, bmap, ctx);
    assembly::DeviceGeometry geom;
    geom.num_nodes = N, geom.num_elems = E, geom.nodes_per_elem = 3;
    geom.x.assign(x, ctx), geom.y.assign(y, ctx), geom.conn.assign(conn, ctx);
    RV hf(N), hd(N), hm(E);
    for (int i = 0; i < N; ++i) hf[i] = 5.0 + std::sin(2.1 * i), hd[i] = 0.02 * std::cos(0.8 * i);
    for (int e = 0; e < E; ++e) hm[e] = 0.3 * std::sin(0.9 * e + 0.4);
    DeviceVector f(hf, ctx), ud(hd, ctx), m(hm, ctx), g(RV(bcd.size(), 0.0), ctx);
    state::PoissonLogKappaResidual R(geom, map, &bmap, &g, f.data());
    const state::Residual& base = R;
    const auto d = base.dims();
    expect(d.num_states == N && d.num_res == N && d.num_param == E, "dims");
    expect(base.hostPattern().layout_id == hmap.pattern().layout_id, "hostPattern is the map's");
    expect(&base.pattern() == map.pattern.get(), "pattern is the map's device pattern");
    DeviceCsrMatrix         jac(map.pattern), adj_jac(map.pattern);
    linalg::BiCGStabSolver  solver(1e-14), adj_solver(1e-14);
    LinearStateSolver       ss(R, jac, solver, ctx);
    ReducedFunctional       J(ss, adj_jac, adj_solver, ud, 1e-3);
    DeviceVector            grad(E);
    J.grad(m, grad);
    const RV hg = grad.toHost(ctx);
    const Real h = 1e-5;
    for (int e : {0, 7, 31, E - 1})
    {
      RV mp = hm, mm = hm;
      mp[e] += h, mm[e] -= h;
      DeviceVector dmp(mp, ctx), dmm(mm, ctx);
      const Real fd = (J.value(dmp) - J.value(dmm)) / (2 * h);
      expect(std::abs(hg[e] - fd) <= 1e-6 * std::max(std::abs(fd), 1e-3 * std::abs(hg[0]) + 1e-12),
             "reduced gradient against central differences (element " + std::to_string(e) + ")");
    }
  });

  run("transposeGatherScatterAndViews", [] {
    needDevice();
    HipContext     ctx;
    HostCsrPattern hp; // tests/unit/backend/runCudaLinalgTests.cu:70-110
    hp.rows = 3, hp.cols = 4, hp.row_ptr = {0, 2, 4, 7}, hp.col_ind = {0, 2, 1, 3, 0, 2, 3};
    DeviceCsrMatrix A(devPattern(hp, ctx)), At;
    A.vals().assign({2.0, -1.0, 3.0, 4.0, -2.0, 5.0, 1.0}, ctx);
    linalg::HipMatrixHandler mh(ctx);
    mh.transpose(A, At);
    expect(At.rows() == 4 && At.cols() == 3, "transposed shape");
    expect(At.pattern().host().row_ptr == IV({0, 2, 3, 5, 7}), "transposed row_ptr");
    expect(At.pattern().host().col_ind == IV({0, 2, 1, 0, 2, 1, 2}), "transposed col_ind");
    expect(At.vals().toHost(ctx) == RV({2.0, -2.0, 3.0, -1.0, 5.0, 4.0, 1.0}), "transposed values");
    const auto layout = At.pattern().layoutId();
    A.vals().assign({1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0}, ctx);
    mh.transpose(A, At);
    expect(At.pattern().layoutId() == layout && At.vals().toHost(ctx) == RV(7, 1.0),
           "an unchanged layout keeps the transposed pattern");
    DeviceVector x(RV{1, 2, 3}, ctx), y1(4), y2(4);
    A.vals().assign({2.0, -1.0, 3.0, 4.0, -2.0, 5.0, 1.0}, ctx);
    mh.transpose(A, At);
    mh.matvecT(A, x, y1);
    mh.matvec(At, x, y2);
    expect(y1.toHost(ctx) == y2.toHost(ctx), "A^T x: matvecT equals the explicit transpose");
    linalg::HipVectorHandler vh(ctx);
    DeviceVector      src(RV{10, 20, 30, 40}, ctx), dst(2), full(RV(4, 0.0), ctx);
    DeviceIndexVector idx(IV{3, 1}, ctx);
    vh.gather(src, idx, dst);
    expect(dst.toHost(ctx) == RV({40, 20}), "gather");
    vh.scatter(dst, idx, full);
    expect(full.toHost(ctx) == RV({0, 20, 0, 40}), "scatter");
    vh.axpby(2.0, DeviceConstVectorView(src).block(2, 2), 0.0, DeviceVectorView(full).block(0, 2));
    expect(full.toHost(ctx) == RV({60, 80, 0, 40}), "axpby on sub-views");
  });

  // the Taylor-Hood Residual: parameter VJP against central differences and
  // the state VJP against the assembled row-replaced Jacobian
  run("navierStokesResidualContract", [] {
    needDevice();
    HipContext ctx;
    const int  n = 3, E = 2 * n * n, Nv = (2 * n + 1) * (2 * n + 1), Np = (n + 1) * (n + 1);
    RV         x2(Nv), y2(Nv);
    IV         c6(6 * E), v3(3 * E), d15(15 * E);
    check(fx_mesh_p2(n, x2.data(), y2.data(), c6.data(), v3.data()));
    check(fx_th_dofs(n, c6.data(), v3.data(), d15.data()));
    RV xv(Np), yv(Np);
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i)
        xv[j * (n + 1) + i] = x2[2 * j * (2 * n + 1) + 2 * i], yv[j * (n + 1) + i] = y2[2 * j * (2 * n + 1) + 2 * i];
    const Index     ndof = 2 * Nv + Np;
    std::vector<IV> dofs(E);
    for (int e = 0; e < E; ++e) dofs[e] = IV(d15.begin() + 15 * e, d15.begin() + 15 * e + 15);
    auto hmap = assembly::makeAssemblyMap(ndof, ndof, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    IV bcd;
    for (int i = 0; i < Nv; ++i)
      if (x2[i] == 0.0 || x2[i] == 1.0 || y2[i] == 0.0 || y2[i] == 1.0) bcd.push_back(2 * i), bcd.push_back(2 * i + 1);
    auto hb = assembly::makeBoundaryMap(bcd, hmap.pattern());
    assembly::DeviceBoundaryMap bmap;
    assembly::copy(hb, bmap, ctx);
    assembly::DeviceTHGeometry geom;
    geom.num_elems = E, geom.ndof = ndof;
    geom.xv.assign(xv, ctx), geom.yv.assign(yv, ctx), geom.vconn3.assign(v3, ctx), geom.dofs15.assign(d15, ctx);
    RV hs(ndof), hl(ndof), hn(E);
    for (Index i = 0; i < ndof; ++i) hs[i] = std::sin(0.37 * i), hl[i] = std::cos(0.23 * i + 0.1);
    for (int e = 0; e < E; ++e) hn[e] = 0.01 + 0.005 * std::sin(1.7 * e);
    DeviceVector s(hs, ctx), lam(hl, ctx), nu(hn, ctx), g(RV(bcd.size(), 0.0), ctx), gr(E), jt, r(ndof);
    state::NavierStokesResidual R(geom, map, &bmap, &g);
    expect(R.dims().num_param == E && R.dims().num_states == ndof, "dims");
    R.applyParamJacT(s, nu, lam, gr, ctx);
    R.applyStateJacT(s, nu, lam, jt, ctx);
    const RV hg = gr.toHost(ctx), hjt = jt.toHost(ctx);
    auto lamDotR = [&](const RV& nn) {
      DeviceVector dn(nn, ctx);
      R.res(s, dn, r, ctx);
      const RV hr = r.toHost(ctx);
      Real     v  = 0.0;
      for (Index i = 0; i < ndof; ++i) v += hl[i] * hr[i];
      return v;
    };
    for (int e : {0, 5, E - 1})
    {
      RV np = hn, nm = hn;
      np[e] += 1e-6, nm[e] -= 1e-6;
      const Real fd = (lamDotR(np) - lamDotR(nm)) / 2e-6;
      expect(std::abs(hg[e] - fd) <= 1e-6 * std::max(1.0, std::abs(fd)), "d(lam.R)/dnu_e by central differences");
    }
    DeviceCsrMatrix J(map.pattern);
    R.assembleStateJac(s, nu, J, ctx);
    DeviceVector ref(ndof);
    linalg::HipMatrixHandler(ctx).matvecT(J, lam, view(ref));
    const RV href = ref.toHost(ctx);
    Real     err = 0.0, nrm = 0.0;
    for (Index i = 0; i < ndof; ++i) err = std::max(err, std::abs(hjt[i] - href[i])), nrm = std::max(nrm, std::abs(href[i]));
    expect(err <= 1e-12 * nrm, "J^T lam equals the assembled row-replaced Jacobian's transpose apply");
  });

  // femx's NavierStokesModel residual under ConstrainedTimeResidual against
  // femx's own outputs (the gls_constrained_n4 fixture): res, assembleNext,
  // applyJacT with respect to the parameters, and solveStep's system, bit for
  // bit; the history transpose zeroes the constrained adjoint entries
  run("constrainedTimeResidualEqualsFemx", [] {
    needDevice();
    const RV    x = readF64("x"), y = readF64("y"), dt = readF64("dt"), rho = readF64("rho"),
             mu = readF64("mu"), ent = readF64("entries"), fv = readF64("fixed_vals"),
             tm = readF64("time");
    const IV conn = readI32("conn"), ctr_dofs = readI32("ctr_dofs"), fixed = readI32("fixed_dofs"),
             nstep = readI32("nstep"), nctr = readI32("num_ctr_params"), ctr_off = readI32("ctr_off"),
             num_prm = readI32("num_prm");
    HipContext  ctx;
    const Index N = static_cast<Index>(x.size()), E = static_cast<Index>(conn.size() / 3), ns = 3 * N;
    std::vector<IV> dofs(E);
    for (Index e = 0; e < E; ++e)
    {
      const Index* v = &conn[3 * e];
      dofs[e]        = {2 * v[0], 2 * v[0] + 1, 2 * v[1], 2 * v[1] + 1, 2 * v[2], 2 * v[2] + 1,
                        2 * N + v[0], 2 * N + v[1], 2 * N + v[2]};
    }
    auto hmap = assembly::makeAssemblyMap(ns, ns, dofs, dofs);
    assembly::DeviceAssemblyMap map;
    assembly::copy(hmap, map, ctx);
    assembly::DeviceNavierGeometry geom;
    assembly::makeNavierGeometry(x, y, conn, hmap, geom, ctx);
    std::vector<fem::DirichletControlMapEntry> entries;
    for (std::size_t k = 0; k + 2 < ent.size(); k += 3)
      entries.push_back({static_cast<Index>(ent[k]), static_cast<Index>(ent[k + 1]), ent[k + 2]});
    fem::DirichletControl ctr(ctr_dofs, nctr[0], entries);
    std::vector<fem::LinearInterpolation> time;
    for (std::size_t k = 0; k + 2 < tm.size(); k += 3)
      time.push_back({static_cast<Index>(tm[k]), static_cast<Index>(tm[k + 1]), tm[k + 2]});
    auto cm    = fem::makeControlMap(nstep[0], ns, ctr, fixed, fv, time, ctr_off[0], num_prm[0]);
    auto model = std::make_unique<model::ns::NavierResidual>(geom, map, nstep[0], dt[0],
                                                             model::ns::FluidParams{rho[0], mu[0]});
    const model::ns::NavierOperator op = model->op();
    auto R = model::ns::makeDeviceTimeResidual(std::move(model), cm, ctx);
    expect(R->dims().num_param == num_prm[0] && R->dims().num_hist == 2, "dims");
    expect(R->hostPattern().layout_id == hmap.pattern().layout_id, "hostPattern");
    for (int k = 0; k < 2; ++k)
    {
      const std::string c = std::to_string(k);
      const IV          step = readI32("step" + c);
      DeviceVector hist(readF64("hist" + c), ctx), nxt(readF64("nxt" + c), ctx),
          prm(readF64("prm" + c), ctx), adj(readF64("adj" + c), ctx), res, rhs(ns), grad;
      state::TimeContext t{step[0], nxt, prm, state::TimeHistoryView(hist.data(), 2, ns)};
      R->res(t, res, ctx);
      expect(res.toHost(ctx) == readF64("res" + c), "res, case " + c);
      DeviceCsrMatrix jac(map.pattern);
      R->assembleNext(t, res, jac, ctx);
      expect(res.toHost(ctx) == readF64("asm_res" + c), "assembleNext residual, case " + c);
      expect(jac.vals().toHost(ctx) == readF64("asm_vals" + c), "assembleNext Jacobian, case " + c);
      R->applyJacT(t, state::VariableBlock::Param, adj, grad, ctx);
      expect(grad.toHost(ctx) == readF64("grad" + c), "applyJacT(param), case " + c);
      state::prepareStepSystem(*R, t, res, jac, rhs, ctx);
      expect(jac.vals().toHost(ctx) == readF64("prep_vals" + c), "solveStep matrix, case " + c);
      expect(rhs.toHost(ctx) == readF64("prep_rhs" + c), "solveStep rhs, case " + c);
      // history transpose: the base VJP of the adjoint with constrained rows zeroed
      RV masked = readF64("adj" + c);
      for (Index d : cm.dofs) masked[d] = 0.0;
      DeviceVector dm(masked, ctx), out, ref;
      R->applyJacT(t, state::VariableBlock::hist(1), adj, out, ctx);
      model::ns::applyNavierHistJacT(op, step[0], 2, 1, hist, nxt, dm, ref, ctx);
      expect(out.toHost(ctx) == ref.toHost(ctx), "applyJacT(history) masks the adjoint, case " + c);
    }
  });

  std::printf("%d passed, %d failed, %d skipped\n", g_pass, g_fail, g_skip);
  return g_fail;
}
