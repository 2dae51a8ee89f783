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

#include "femx_hip/femx_hip.hpp"

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

HostCsrPattern dense3()
{
  HostCsrPattern p;
  p.rows = p.cols = 3;
  p.row_ptr       = {0, 3, 6, 9};
  p.col_ind       = {0, 1, 2, 0, 1, 2, 0, 1, 2};
  return p;
}
} // namespace

int main()
{
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
    mh.matvec(A, x.data(), 4, y.data(), 3);
    expect(near(y.toHost(ctx), {-1, 22, 17}), "rectangular CSR apply");
    vh.dot(x, x, nrm);
    expect(near(nrm.toHost(ctx), {30}), "squared norm");
    DeviceVector a(RV{1, 2, 3}, ctx);
    vh.axpby(-2.0, a, 0.5, y);
    expect(near(y.toHost(ctx), {-2.5, 7, 2.5}), "axpby");
    DeviceVector tin(RV{2, -1, 0.5}, ctx), tout(4);
    mh.matvecT(A, tin.data(), 3, tout.data(), 4);
    expect(near(tout.toHost(ctx), {3, -3, 0.5, -3.5}), "transposed CSR apply");
    A.vals().assign({-1.0, 2.0, 0.5, -3.0, 4.0, 1.0, -2.0}, ctx);
    mh.matvec(A, x.data(), 4, y.data(), 3);
    expect(near(y.toHost(ctx), {5, -11, -1}), "updated CSR values");
    bool rejected = false;
    try
    {
      mh.matvec(A, x.data(), 4, x.data() + 1, 3);
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

  run("taylorHoodAssemblyIsDeterministic", [] {
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
    linalg::HipMatrixHandler(ctx).matvecT(J, lam.data(), N, ref.data(), N);
    const RV href = ref.toHost(ctx);
    double err = 0.0, nrm = 0.0;
    for (int i = 0; i < N; ++i) err = std::max(err, std::abs(hjt[i] - href[i])), nrm = std::max(nrm, std::abs(href[i]));
    expect(err <= 1e-12 * nrm, "J^T lam equals the assembled row-replaced Jacobian's transpose apply");
  });

  std::printf("%d passed, %d failed, %d skipped\n", g_pass, g_fail, g_skip);
  return g_fail;
}
