"""GLS P1/P1 Navier-Stokes time step on the GPU.  The kernel keeps femx's
operation order without contraction, so the bar is bitwise equality: with
femx's own assembly (the golden fixtures written by the compiled reference)
and with the oracle on perturbed and renumbered meshes, at the 1024^2 size,
and run to run.  Also the reference tests' invariants (Jacobian independent
of the next state, res(nxt) = K nxt + res(0)) through the CSR SpMV."""
import numpy as np
import pytest
import torch

from conftest import close, golden
from femx_amd import gls
from femx_amd.linalg import CsrMatrix, MatrixHandler
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def run(geom, amap, fluid, dt, step, nh, hist, nxt, jac=True, wrt=gls.NEXT_STATE):
    op = gls.NavierOperator(geom, fluid, dt)
    A = CsrMatrix(amap.pattern, DEV) if jac else None
    if A is not None:
        A.vals.fill_(np.nan)
    res = torch.full((geom.num_states,), np.nan, dtype=torch.float64, device=DEV)
    gls.assemble(op, step, nh, wrt, amap, dev(hist), dev(nxt), res, A)
    return (host(A.vals) if jac else None), host(res)


@pytest.mark.parametrize("name", ["gls_tri1.npz", "gls_p1_n4_perturbed.npz", "gls_p1_n6.npz"])
def test_gls_equals_reference_fixtures_bitwise(name):
    g = golden(name)
    print("case", name, flush=True)
    geom = gls.GLSGeometry(g["x"], g["y"], g["conn"], DEV)
    amap = gls.make_navier_map(geom)
    fluid = gls.FluidParams(float(g["rho"]), float(g["mu"]))
    for k in range(int(g["num_cases"])):
        v, r = run(geom, amap, fluid, float(g["dt"]), int(g[f"step{k}"]), int(g[f"num_hist{k}"]),
                   g[f"hist{k}"], g[f"nxt{k}"])
        assert np.array_equal(v, g[f"vals{k}"]), (name, k)
        assert np.array_equal(r, g[f"res{k}"]), (name, k)


def perturbed(n, seed, renumber=False):
    x, y, conn = O.mesh_p1(n)
    rng = np.random.default_rng(seed)
    inner = (x > 0) & (x < 1) & (y > 0) & (y < 1)
    x[inner] += 0.25 / n * rng.uniform(-1, 1, inner.sum())
    y[inner] += 0.25 / n * rng.uniform(-1, 1, inner.sum())
    if renumber:
        perm = rng.permutation(x.size)
        xs, ys = np.empty_like(x), np.empty_like(y)
        xs[perm], ys[perm] = x, y
        x, y, conn = xs, ys, perm[conn].astype(np.int32)
    return x, y, conn


@pytest.mark.parametrize("n,renumber", [(3, False), (40, False), (40, True)])
@pytest.mark.parametrize("step,nh", [(0, 1), (1, 2), (3, 3)])
def test_gls_matches_oracle_bitwise(n, renumber, step, nh):
    print("case", n, renumber, step, nh, flush=True)
    x, y, conn = perturbed(n, 7 + n, renumber)
    geom = gls.GLSGeometry(x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    rng = np.random.default_rng(step)
    nd = geom.num_states
    hist = rng.standard_normal(nh * nd)
    nxt = rng.standard_normal(nd)
    v, r = run(geom, amap, gls.FluidParams(1.2, 0.03), 0.05, step, nh, hist, nxt)
    ov, orr = O.gls_assemble(x, y, conn, geom.h_dofs9.reshape(-1), nd, 1.2, 0.03, 0.05, step, nh,
                             hist, nxt, amap.jac()[1], amap.pattern.nnz)
    assert np.array_equal(v, ov)
    assert np.array_equal(r, orr)


def test_gls_residual_only_and_history_block():
    x, y, conn = perturbed(12, 3)
    geom = gls.GLSGeometry(x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    rng = np.random.default_rng(4)
    nd = geom.num_states
    hist, nxt = rng.standard_normal(2 * nd), rng.standard_normal(nd)
    fluid = gls.FluidParams(1.0, 0.01)
    v, r = run(geom, amap, fluid, 0.1, 1, 2, hist, nxt)
    _, r2 = run(geom, amap, fluid, 0.1, 1, 2, hist, nxt, jac=False)
    assert np.array_equal(r, r2)
    # with respect to a history state the row operator leaves the Jacobian zero
    vh, rh = run(geom, amap, fluid, 0.1, 1, 2, hist, nxt, wrt=gls.history_state(1))
    assert np.array_equal(rh, r) and not vh.any()
    with pytest.raises(RuntimeError, match="variable block is invalid"):
        run(geom, amap, fluid, 0.1, 1, 2, hist, nxt, wrt=gls.PARAM)
    with pytest.raises(RuntimeError, match="variable block is invalid"):
        run(geom, amap, fluid, 0.1, 1, 2, hist, nxt, wrt=gls.history_state(2))


def test_gls_alias_is_rejected():
    x, y, conn = O.mesh_p1(2)
    geom = gls.GLSGeometry(x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    op = gls.NavierOperator(geom, gls.FluidParams(), 0.1)
    s = torch.zeros(geom.num_states, dtype=torch.float64, device=DEV)
    with pytest.raises(RuntimeError, match="must not alias"):
        gls.assemble(op, 0, 1, gls.NEXT_STATE, amap, s, s.clone(), s)


def test_gls_1024_matches_oracle_and_invariants():
    """1024^2 cells (2,097,152 elements, 3,151,875 states): bitwise equal to
    the oracle and run to run; the Jacobian does not depend on the next
    state and res(nxt) - res(0) = K nxt (the SpMV, 1e-12)."""
    n = 1024
    x, y, conn = perturbed(n, 1)
    geom = gls.GLSGeometry(x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    nd = geom.num_states
    assert nd == 3 * (n + 1) ** 2 and amap.pattern.nnz == 9 * (7 * n * n + 6 * n + 1)
    rng = np.random.default_rng(11)
    hist = rng.standard_normal(2 * nd)
    nxt = rng.standard_normal(nd)
    fluid = gls.FluidParams(1.0, 1e-3)
    op = gls.NavierOperator(geom, fluid, 0.01)
    A = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(nd, dtype=torch.float64, device=DEV)
    h_d, n_d = dev(hist), dev(nxt)
    gls.assemble(op, 1, 2, gls.NEXT_STATE, amap, h_d, n_d, res, A)
    v1, r1 = host(A.vals), host(res)
    ov, orr = O.gls_assemble(x, y, conn, geom.h_dofs9.reshape(-1), nd, 1.0, 1e-3, 0.01, 1, 2,
                             hist, nxt, amap.jac()[1], amap.pattern.nnz)
    assert np.array_equal(v1, ov) and np.array_equal(r1, orr)
    gls.assemble(op, 1, 2, gls.NEXT_STATE, amap, h_d, n_d, res, A)
    assert np.array_equal(host(A.vals), v1) and np.array_equal(host(res), r1)
    zero = torch.zeros(nd, dtype=torch.float64, device=DEV)
    r0 = torch.empty_like(res)
    gls.assemble(op, 1, 2, gls.NEXT_STATE, amap, h_d, zero, r0, A)
    assert np.array_equal(host(A.vals), v1)
    kx = torch.empty_like(res)
    MatrixHandler().matvec(A, n_d, kx)
    assert close(host(res) - host(r0), host(kx))


@pytest.mark.parametrize("n,renumber", [(3, False), (40, True)])
@pytest.mark.parametrize("step,nh", [(1, 2), (0, 1), (2, 3)])
def test_gls_hist_vjp_matches_oracle(n, renumber, step, nh):
    """The history VJP (femx's Enzyme product, hand-derived here) against the
    oracle's restatement, every lag, within 1e-12: the oracle forms tau with
    femx's own expressions (elemLength's h = 2/sum, (2|u|/h)^2, (4 nu/h^2)^2,
    u/|u|), the kernel with products that avoid the divisions.  Every lag
    also has a few whole elements at rest (both velocity components of all
    three nodes zero in every history lag), so that their quadrature points
    take the at-rest branch (speed 0: fixed direction, no diffusive tau
    derivative)."""
    print("case", n, renumber, step, nh, flush=True)
    x, y, conn = perturbed(n, 3 + n, renumber)
    geom = gls.GLSGeometry(x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    nd = geom.num_states
    rng = np.random.default_rng(20 + step)
    hist = rng.standard_normal(nh * nd)
    d9 = geom.h_dofs9.reshape(-1, 9)
    rest = rng.choice(geom.num_elems, max(2, geom.num_elems // 16), replace=False)
    vel = d9[rest, :6].reshape(-1)
    for lag in range(nh):
        hist[lag * nd + vel] = 0.0
    nxt, adj = rng.standard_normal(nd), rng.standard_normal(nd)
    op = gls.NavierOperator(geom, gls.FluidParams(1.2, 0.03), 0.05)
    for lag in range(nh):
        out = host(gls.hist_vjp(op, step, nh, lag, amap, dev(hist), dev(nxt), dev(adj)))
        ref = O.gls_hist_vjp(x, y, conn, geom.h_dofs9.reshape(-1), nd, 1.2, 0.03, 0.05, step, nh,
                             lag, hist, nxt, adj)
        assert close(out, ref), lag
        assert not out[2 * geom.num_nodes:].any()


def test_gls_hist_vjp_1024_adjoint_identity():
    """1024^2: within 1e-12 of the oracle, bitwise equal run to run; and the adjoint
    identity dir^T VJP_lag(adj) = adj^T JVP_lag(dir) against the oracle's
    dual-number forward mode, an independent restatement (1e-12)."""
    n = 1024
    x, y, conn = perturbed(n, 2)
    geom = gls.GLSGeometry(x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    nd = geom.num_states
    rng = np.random.default_rng(13)
    hist = rng.standard_normal(2 * nd)
    nxt, adj, dirv = (rng.standard_normal(nd) for _ in range(3))
    op = gls.NavierOperator(geom, gls.FluidParams(1.0, 1e-3), 0.01)
    h_d, n_d, a_d = dev(hist), dev(nxt), dev(adj)
    d9 = geom.h_dofs9.reshape(-1)
    for lag in (0, 1):
        out = host(gls.hist_vjp(op, 1, 2, lag, amap, h_d, n_d, a_d))
        assert np.array_equal(out, host(gls.hist_vjp(op, 1, 2, lag, amap, h_d, n_d, a_d)))
        ref = O.gls_hist_vjp(x, y, conn, d9, nd, 1.0, 1e-3, 0.01, 1, 2, lag, hist, nxt, adj)
        assert close(out, ref)
        jv = O.gls_hist_jvp(x, y, conn, d9, nd, 1.0, 1e-3, 0.01, 1, 2, lag, hist, nxt, dirv)
        lhs, rhs = dirv @ out, adj @ jv
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, np.abs(dirv * out).sum()), (lag, lhs, rhs)


@pytest.mark.parametrize("lag", [0, 1])
def test_gls_hist_vjp_finite_difference(lag):
    """femx's method (runNavierOperatorTests.cpp:292-382) on the GPU: the
    reference test's state pattern on a 32^2 mesh, central differences of
    the GPU residual with eps 1e-6, near(vjp, fd, 2e-6)."""
    x, y, conn = O.mesh_p1(32)
    geom = gls.GLSGeometry(2.0 * x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    nd = geom.num_states
    i = np.arange(nd)
    hist = np.concatenate([0.01 * (i + 1), -0.005 * (i + 1)]) / nd
    nxt = (0.02 - 0.003 * i) / nd
    dirv = 0.2 - 0.01 * i / nd
    adj = -0.15 + 0.013 * i / nd
    op = gls.NavierOperator(geom, gls.FluidParams(1.2, 0.03), 0.05)
    out = host(gls.hist_vjp(op, 1, 2, lag, amap, dev(hist), dev(nxt), dev(adj)))
    res = torch.empty(nd, dtype=torch.float64, device=DEV)
    eps, vals = 1e-6, []
    for sgn in (1.0, -1.0):
        hp = hist.copy()
        hp[lag * nd:(lag + 1) * nd] += sgn * eps * dirv
        gls.assemble_residual(op, 1, 2, amap, dev(hp), dev(nxt), res)
        vals.append(host(res) @ adj)
    fd = (vals[0] - vals[1]) / (2 * eps)
    vjp = dirv @ out
    assert abs(vjp - fd) <= 2e-6 * max(1.0, abs(vjp), abs(fd)), (vjp, fd)


@pytest.mark.parametrize("step,nh", [(1, 2), (0, 1)])
def test_gls_hist_vjp_both_lags_one_pass(step, nh):
    """hist_vjp_lags (one pass over element patches, both lags from one
    evaluation of the partials) against the two-phase fx_gls_hist_vjp per
    lag: rows of nodes inside one patch bitwise equal, seam nodes within
    1e-12, repeat launches bitwise; with one history lag only lag 0 exists."""
    x, y, conn = perturbed(40, 7, True)
    geom = gls.GLSGeometry(x, y, conn, DEV)
    amap = gls.make_navier_map(geom)
    nd = geom.num_states
    rng = np.random.default_rng(31 + step)
    hist = dev(rng.standard_normal(nh * nd))
    nxt, adj = dev(rng.standard_normal(nd)), dev(rng.standard_normal(nd))
    op = gls.NavierOperator(geom, gls.FluidParams(1.1, 0.02), 0.04)
    o0, o1 = gls.hist_vjp_lags(op, step, nh, amap, hist, nxt, adj)
    assert (o1 is None) == (nh == 1)
    r0, r1 = gls.hist_vjp_lags(op, step, nh, amap, hist, nxt, adj)
    pp = gls._patches(amap, geom)
    seam = pp.h["seam_dof"][:pp.num_seam]
    rows = gls._plan(amap, geom).h["node_rows"].reshape(-1, 3)
    inner = np.ones(nd, bool)
    inner[rows[seam].reshape(-1)] = False
    for lag, out, again in ((0, o0, r0), (1, o1, r1)):
        if lag >= nh:
            continue
        out, again = host(out), host(again)
        assert np.array_equal(out, again)
        two = gls.hist_vjp_two_phase(op, step, nh, lag, amap, hist, nxt, adj)
        two = host(two)
        ref = O.gls_hist_vjp(x, y, conn, geom.h_dofs9.reshape(-1), nd, 1.1, 0.02, 0.04, step, nh,
                             lag, host(hist), host(nxt), host(adj))
        assert close(out, ref), lag
        assert np.array_equal(out[inner], two[inner]), lag
        assert close(out, two), lag
