"""GLS P1/P1 Navier-Stokes on the CPU: the oracle against femx's own
assembly (golden fixtures written by the compiled reference), the oracle
against closed forms and the reference tests' invariants, and the node plan
of the HIP kernel (fx_map_gls_plan) replayed in numpy."""
import numpy as np
import pytest

from conftest import golden
from femx_amd import gls
from femx_amd.assembly import make_assembly_map
from oracle import oracle as O

FIXTURES = ["gls_tri1.npz", "gls_p1_n4_perturbed.npz", "gls_p1_n6.npz"]


def _map(g):
    nd = 3 * g["x"].size
    amap = make_assembly_map(nd, nd, g["dofs9"].reshape(-1, 9))
    return amap, amap.jac()[1]


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_equals_reference_bitwise(name):
    """femx's NavierOperator through the host time assemble, compiled from
    femx's sources, against the C restatement: identical bits."""
    g = golden(name)
    x, y, conn = g["x"], g["y"], g["conn"]
    N = x.size
    assert np.array_equal(g["dofs9"], O.gls_dofs(N, conn))
    assert np.array_equal(g["dofs9"].reshape(-1, 9), gls.navier_dofs(N, conn))
    amap, jm = _map(g)
    assert np.array_equal(amap.pattern.row_ptr, g["row_ptr"])
    assert np.array_equal(amap.pattern.col_ind, g["col_ind"])
    for k in range(int(g["num_cases"])):
        vals, res = O.gls_assemble(x, y, conn, g["dofs9"], 3 * N, float(g["rho"]),
                                   float(g["mu"]), float(g["dt"]), int(g[f"step{k}"]),
                                   int(g[f"num_hist{k}"]), g[f"hist{k}"], g[f"nxt{k}"], jm,
                                   g["col_ind"].size)
        assert np.array_equal(vals, g[f"vals{k}"]), (name, k)
        assert np.array_equal(res, g[f"res{k}"]), (name, k)


def _tri_grads(xe, ye):
    A2 = (xe[1] - xe[0]) * (ye[2] - ye[0]) - (xe[2] - xe[0]) * (ye[1] - ye[0])
    gx = np.array([ye[1] - ye[2], ye[2] - ye[0], ye[0] - ye[1]]) / A2
    gy = np.array([xe[2] - xe[1], xe[0] - xe[2], xe[1] - xe[0]]) / A2
    return gx, gy, abs(A2) / 2


def test_oracle_element_closed_form_at_rest():
    """Zero velocity in both lags: no advection, tau = 1/sqrt((2/dt)^2 +
    (4 nu/h^2)^2) with h from elemLength's fallback direction (1,1)/sqrt 2.
    The blocks are then textbook P1 integrals: K_uu = rho/dt M + mu/2 L
    (per component), K_up(i,c;j) = -A/3 dphi_i/dx_c, K_pu(i;j,d) = A/3
    dphi_j/dx_d + tau/dt A/3 dphi_i/dx_d, K_pp = tau/rho L."""
    xe = np.array([0.1, 1.3, 0.4])
    ye = np.array([-0.2, 0.3, 1.1])
    rho, mu, dt = 1.3, 0.07, 0.02
    K, R = O.gls_element(xe, ye, rho, mu, dt, 1, 2, np.zeros(18), np.zeros(9))
    gx, gy, A = _tri_grads(xe, ye)
    M = A / 12 * (np.ones((3, 3)) + np.eye(3))
    L = A * (np.outer(gx, gx) + np.outer(gy, gy))
    s = 1 / np.sqrt(2)
    h = 2 / np.abs(s * gx + s * gy).sum()
    tau = 1 / np.sqrt((2 / dt) ** 2 + (4 * mu / rho / h ** 2) ** 2)
    G = [gx, gy]
    ref = np.zeros((9, 9))
    for i in range(3):
        for j in range(3):
            for c in range(2):
                ref[2 * i + c, 2 * j + c] = rho / dt * M[i, j] + 0.5 * mu * L[i, j]
                ref[2 * i + c, 6 + j] = -A / 3 * G[c][i]
                ref[6 + i, 2 * j + c] = A / 3 * G[c][j] + tau / dt * A / 3 * G[c][i]
            ref[6 + i, 6 + j] = tau / rho * L[i, j]
    assert np.allclose(K, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert np.all(R == 0.0)


def test_oracle_element_uniform_flow():
    """A uniform velocity U in both lags: grad u = 0, the advective terms
    vanish, h = 2/sum|U.grad phi_i|/|U|, tau from the closed form, and
    F_u(i,c) = rho/dt U_c A/3 + tau rho/dt (grad phi_i.U) U_c A,
    F_p(i) = tau/dt (grad phi_i.U) A; the residual at nxt = 0 is -F."""
    xe = np.array([0.0, 0.9, 0.2])
    ye = np.array([0.0, 0.1, 0.7])
    rho, mu, dt = 0.9, 0.01, 0.05
    U = np.array([0.6, -0.35])
    cur = np.tile(U, 3)
    hist = np.concatenate([cur, np.zeros(3), cur, np.zeros(3)])
    _, R = O.gls_element(xe, ye, rho, mu, dt, 2, 2, hist, np.zeros(9))
    gx, gy, A = _tri_grads(xe, ye)
    sp = np.linalg.norm(U)
    gu = gx * U[0] + gy * U[1]
    h = 2 / (np.abs(gu).sum() / sp)
    tau = 1 / np.sqrt((2 / dt) ** 2 + (2 * sp / h) ** 2 + (4 * mu / rho / h ** 2) ** 2)
    F = np.zeros(9)
    for i in range(3):
        for c in range(2):
            F[2 * i + c] = rho / dt * U[c] * A / 3 + tau * rho / dt * gu[i] * U[c] * A
        F[6 + i] = tau / dt * gu[i] * A
    assert np.allclose(R, -F, rtol=1e-12, atol=1e-13 * np.abs(F).max())


def test_oracle_extrapolation_and_linearity():
    """runNavierOperatorTests.cpp:223-290 invariants on a perturbed mesh: the
    Jacobian does not depend on the next state and res(nxt) = K nxt + res(0).
    Extrapolation: step > 0 with two equal lags equals step 0 with one lag
    (1.5 v - 0.5 v = v up to rounding), and step 0 ignores the second lag."""
    g = golden("gls_p1_n4_perturbed.npz")
    x, y, conn, d9 = g["x"], g["y"], g["conn"], g["dofs9"]
    nd = 3 * x.size
    amap, jm = _map(g)
    rp, ci = amap.pattern.row_ptr, amap.pattern.col_ind
    nnz = ci.size
    rng = np.random.default_rng(5)
    h = rng.standard_normal(2 * nd)
    nx = rng.standard_normal(nd)
    args = (x, y, conn, d9, nd, 1.2, 0.03, 0.05)
    v1, r1 = O.gls_assemble(*args, 1, 2, h, nx, jm, nnz)
    v0, r0 = O.gls_assemble(*args, 1, 2, h, np.zeros(nd), jm, nnz)
    assert np.array_equal(v1, v0)
    Kx = np.array([v1[rp[r]:rp[r + 1]] @ nx[ci[rp[r]:rp[r + 1]]] for r in range(nd)])
    assert np.allclose(r1, Kx + r0, rtol=1e-12, atol=1e-12 * np.abs(r1).max())
    same = np.concatenate([h[:nd], h[:nd]])
    va, ra = O.gls_assemble(*args, 1, 2, same, nx, jm, nnz)
    vb, rb = O.gls_assemble(*args, 0, 1, h[:nd], nx, jm, nnz)
    assert np.allclose(va, vb, rtol=1e-13, atol=1e-13 * np.abs(vb).max())
    vc, rc = O.gls_assemble(*args, 0, 2, h, nx, jm, nnz)
    assert np.array_equal(vc, vb) and np.array_equal(rc, rb)


# ---------------------------------------------------------------------------
# node plan

def _plan(n, perturb=False):
    x, y, conn = O.mesh_p1(n)
    geom = gls.GLSGeometry(x, y, conn, device="cpu")
    amap = gls.make_navier_map(geom)
    return geom, amap, gls.GLSPlan(amap, geom, "cpu")


@pytest.mark.parametrize("n", [1, 4, 9])
def test_plan_tables(n):
    geom, amap, p = _plan(n)
    N, E = geom.num_nodes, geom.num_elems
    h = p.h
    nr = h["node_rows"].reshape(N, 3)
    assert np.array_equal(nr[:, 0], 2 * np.arange(N)) and np.array_equal(nr[:, 2], 2 * N + np.arange(N))
    off, elem = h["inc_off"], h["inc_elem"][:3 * E]
    meta = h["inc_meta"][:24 * E].view(np.uint8).reshape(-1, 32)
    jo, jm = amap.jac()
    rp = amap.pattern.row_ptr
    conn = geom.h_conn
    for v in range(N):
        es = elem[off[v]:off[v + 1]]
        assert np.all(np.diff(es) > 0)
        assert sorted(es.tolist()) == sorted(np.nonzero((conn == v).any(1))[0].tolist())
        for t in range(off[v], off[v + 1]):
            e, ln = elem[t], meta[t, 27]
            assert conn[e, ln] == v
            assert h["inc_slot"][3 * e + ln] == t
            for kk, lr in enumerate((2 * ln, 2 * ln + 1, 6 + ln)):
                got = rp[nr[v, kk]] + meta[t, 9 * kk: 9 * kk + 9].astype(np.int64)
                assert np.array_equal(got, jm[81 * e + 9 * lr: 81 * e + 9 * lr + 9])
    # runs: two per block (velocity rows, pressure rows) laid end to end
    runs = h["blk_runs"].reshape(p.num_blocks, gls.MAX_RUNS, 4)
    for b in range(p.num_blocks):
        n0, n1 = gls.BLOCK_NODES * b, min(N, gls.BLOCK_NODES * (b + 1))
        r = runs[b]
        if n0 == 0 and n1 == N:  # velocity rows run straight into the pressure rows
            assert r[0].tolist() == [0, rp[3 * N], 0, 0] and not r[1:].any()
            continue
        assert r[0, 0] == rp[2 * n0] and r[0, 1] == rp[2 * n1] - rp[2 * n0] and r[0, 2] == 0
        assert r[1, 0] == rp[2 * N + n0] and r[1, 1] == rp[2 * N + n1] - rp[2 * N + n0]
        assert r[1, 2] == r[0, 1] and not r[2:].any()
        assert r[0, 1] + r[1, 1] <= p.max_tile


def _emulate(geom, amap, p, rho, mu, dt, step, nh, hist, nxt):
    """The kernel's data path in numpy: per block a zeroed tile, each node's
    incidences in order, element rows from the oracle added at tile +
    node_tile + meta, then the runs copied out.  Asserts every tile index
    stays inside the block's tile."""
    h = p.h
    N = geom.num_nodes
    nnz = amap.pattern.nnz
    vals = np.full(nnz, np.nan)
    res = np.full(geom.num_states, np.nan)
    runs = h["blk_runs"].reshape(p.num_blocks, gls.MAX_RUNS, 4)
    meta = h["inc_meta"].view(np.uint8).reshape(-1, 32)
    for b in range(p.num_blocks):
        tl = int((runs[b, :, 2] + runs[b, :, 1]).max())
        assert tl <= p.max_tile
        tile = np.zeros(tl)
        for v in range(gls.BLOCK_NODES * b, min(N, gls.BLOCK_NODES * (b + 1))):
            rows = h["node_rows"][3 * v: 3 * v + 3]
            s = [0.0, 0.0, 0.0]
            for t in range(h["inc_off"][v], h["inc_off"][v + 1]):
                e, ln = h["inc_elem"][t], meta[t, 27]
                c = geom.h_conn[e]
                d9 = geom.h_dofs9[e]
                he = np.concatenate([hist[lag * geom.num_states + d9] for lag in range(min(nh, 2))])
                K, R = O.gls_element(geom.h_x[c], geom.h_y[c], rho, mu, dt, step, nh, he, nxt[d9])
                for kk, lr in enumerate((2 * ln, 2 * ln + 1, 6 + ln)):
                    s[kk] += R[lr]
                    for j in range(9):
                        i = h["node_tile"][3 * v + kk] + meta[t, 9 * kk + j]
                        assert 0 <= i < tl
                        tile[i] += K[lr, j]
            res[rows] = s
        for r in runs[b]:
            vals[r[0]: r[0] + r[1]] = tile[r[2]: r[2] + r[1]]
    return vals, res


def test_plan_replay_reproduces_oracle():
    """Replaying the plan gives the oracle's assembly bit for bit (same
    element order per entry), and covers every CSR entry and row once."""
    x, y, conn = O.mesh_p1(5)
    inner = (x > 0) & (x < 1) & (y > 0) & (y < 1)
    rng = np.random.default_rng(8)
    x[inner] += 0.04 * rng.uniform(-1, 1, inner.sum())
    geom = gls.GLSGeometry(x, y, conn, device="cpu")
    amap = gls.make_navier_map(geom)
    p = gls.GLSPlan(amap, geom, "cpu")
    nd = geom.num_states
    hist = rng.standard_normal(2 * nd)
    nxt = rng.standard_normal(nd)
    v, r = _emulate(geom, amap, p, 1.1, 0.02, 0.03, 1, 2, hist, nxt)
    ov, orr = O.gls_assemble(x, y, conn, geom.h_dofs9.reshape(-1), nd, 1.1, 0.02, 0.03, 1, 2,
                             hist, nxt, amap.jac()[1], amap.pattern.nnz)
    assert np.array_equal(v, ov) and np.array_equal(r, orr)


def test_plan_node_order_and_isolated_node():
    """Nodes numbered against the element order, and a node of no element
    (which owns no rows and is skipped)."""
    x, y, conn = O.mesh_p1(3)
    N = x.size
    perm = np.random.default_rng(2).permutation(N)
    conn = perm[conn].astype(np.int32)
    xs, ys = np.empty(N), np.empty(N)
    xs[perm], ys[perm] = x, y
    xs, ys = np.append(xs, 5.0), np.append(ys, 5.0)  # isolated node N
    geom = gls.GLSGeometry(xs, ys, conn, device="cpu")
    # the isolated node's DOFs are not in any element: the map's rows stop at
    # the element DOFs' range, so build the map over the element DOFs only
    d9 = geom.h_dofs9
    amap = make_assembly_map(geom.num_states, geom.num_states, d9)
    with pytest.raises(RuntimeError, match="belongs to no node"):
        gls.GLSPlan(amap, geom, "cpu")
    geom = gls.GLSGeometry(xs[:N], ys[:N], conn, device="cpu")
    amap = gls.make_navier_map(geom)
    p = gls.GLSPlan(amap, geom, "cpu")
    rng = np.random.default_rng(3)
    hist = rng.standard_normal(geom.num_states)
    nxt = rng.standard_normal(geom.num_states)
    v, r = _emulate(geom, amap, p, 1.0, 0.01, 0.1, 0, 1, hist, nxt)
    ov, orr = O.gls_assemble(xs[:N], ys[:N], conn, geom.h_dofs9.reshape(-1), geom.num_states,
                             1.0, 0.01,
                             0.1, 0, 1, hist, nxt, amap.jac()[1], amap.pattern.nnz)
    assert np.array_equal(v, ov) and np.array_equal(r, orr)


def test_plan_rejects_inconsistent_tables():
    x, y, conn = O.mesh_p1(2)
    N = x.size
    d9 = gls.navier_dofs(N, conn)
    bad = d9.copy()
    bad[0, 6] = bad[0, 7]  # pressure DOF of node 0 replaced within element 0
    amap = make_assembly_map(3 * N, 3 * N, bad)
    geom = gls.GLSGeometry(x, y, conn, device="cpu")
    with pytest.raises(RuntimeError, match="disagree|two nodes|no node"):
        gls.GLSPlan(amap, geom, "cpu")
    amap = make_assembly_map(3 * N, 3 * N, [[0, 1, 2]])
    with pytest.raises(RuntimeError, match="9x9"):
        _lib_plan(amap, N, conn[:3])


def _lib_plan(amap, N, conn):
    import ctypes as C
    from femx_amd import _lib
    from femx_amd.linalg import ptr
    buf = [np.empty(64, np.int32) for _ in range(8)]
    mx = C.c_int32()
    _lib.call("fx_map_gls_plan", amap._handle(), N, ptr(np.ascontiguousarray(conn, np.int32)),
              *[ptr(b) for b in buf], C.byref(mx))


def test_operator_parameter_checks():
    x, y, conn = O.mesh_p1(1)
    geom = gls.GLSGeometry(x, y, conn, device="cpu")
    with pytest.raises(RuntimeError, match="time step"):
        gls.NavierOperator(geom, gls.FluidParams(1.0, 1.0), 0.0)
    with pytest.raises(RuntimeError, match="density"):
        gls.NavierOperator(geom, gls.FluidParams(-1.0, 1.0), 0.1)
    with pytest.raises(RuntimeError, match="viscosity"):
        gls.NavierOperator(geom, gls.FluidParams(1.0, float("nan")), 0.1)


# ---------------------------------------------------------------------------
# history VJP (femx takes it with Enzyme; here hand-derived, checked against
# an independent dual-number forward mode and against finite differences)

def _near(a, b, tol):
    """femx's near() (runNavierOperatorTests.cpp:44-48)."""
    return abs(a - b) <= tol * max(1.0, abs(a), abs(b))


ELEM = (np.array([0.1, 1.2, 0.3]), np.array([-0.1, 0.2, 0.9]))


@pytest.mark.parametrize("step,nh", [(1, 2), (0, 1), (0, 2), (3, 2), (2, 3)])
@pytest.mark.parametrize("still", [False, True])
def test_hist_vjp_element_adjoint_identity(step, nh, still):
    """dir^T (dR/dh)^T a from the hand-written reverse mode equals a^T (dR/dh) dir
    from the dual-number forward mode (1e-13), for both lags at once; the
    dual value part equals the residual; pressure columns are zero."""
    rng = np.random.default_rng(10 * step + nh + (100 if still else 0))
    h = rng.standard_normal(27)
    if still:  # zero velocity in both lags: elemLength's fallback direction
        h[:6] = 0.0
        h[9:15] = 0.0
    n, a = rng.standard_normal(9), rng.standard_normal(9)
    d = rng.standard_normal(18)
    args = ELEM + (1.2, 0.03, 0.05, step, nh)
    nhe = min(nh, 2)
    R, dR = O.gls_element_jvp(*args, h[:18], np.r_[d, np.zeros(0)], n)
    _, R0 = O.gls_element(*args, h[:9 * nhe], n)
    assert np.allclose(R, R0, rtol=1e-14, atol=1e-14 * np.abs(R0).max())
    V = O.gls_element_hvjp(*args, h[:9 * nhe], n, a)
    assert not V[:, 6:].any()
    if not (step > 0 and nh > 1):
        assert not V[1].any()  # no extrapolation: the previous state does not enter
    lhs = a @ dR
    rhs = V[0] @ d[:9] + (V[1] @ d[9:] if nh > 1 else 0.0)
    assert abs(lhs - rhs) <= 1e-13 * max(1.0, abs(lhs))


def test_hist_vjp_element_finite_difference():
    """Central differences of the residual (h = 1e-6) against the forward mode."""
    rng = np.random.default_rng(4)
    h, n, a, d = (rng.standard_normal(k) for k in (18, 9, 9, 18))
    args = ELEM + (1.2, 0.03, 0.05, 1, 2)
    _, dR = O.gls_element_jvp(*args, h, d, n)
    eps = 1e-6
    _, Rp = O.gls_element(*args, h + eps * d, n)
    _, Rm = O.gls_element(*args, h - eps * d, n)
    fd = (Rp - Rm) / (2 * eps)
    assert np.allclose(dR, fd, rtol=1e-7, atol=1e-7 * np.abs(fd).max())


@pytest.mark.parametrize("lag", [0, 1])
def test_hist_vjp_femx_finite_difference_test(lag):
    """runNavierOperatorTests.cpp:292-382 restated on a triangle mesh: rho
    1.2, mu 0.03, dt 0.05, step 1, two lags filled as fillStates does, dir_i =
    0.2 - 0.01 i, adj_i = -0.15 + 0.013 i, eps 1e-6, near(vjp, fd, 2e-6)."""
    x, y, conn = O.mesh_p1(2)
    x = 2.0 * x  # the reference test's 2 x 1 domain
    N = x.size
    ns_ = 3 * N
    d9 = O.gls_dofs(N, conn)
    off = O.uniform_offsets(conn.size // 3, 9)
    _, ci, _, jm = O.map_build(ns_, ns_, off, d9, off, d9)
    i = np.arange(ns_)
    hist = np.concatenate([0.01 * (i + 1), -0.005 * (i + 1)])
    nxt = 0.02 - 0.003 * i
    dirv = 0.2 - 0.01 * i
    adj = -0.15 + 0.013 * i
    args = (x, y, conn, d9, ns_, 1.2, 0.03, 0.05, 1, 2)
    eps = 1e-6
    hp, hm = hist.copy(), hist.copy()
    hp[lag * ns_:(lag + 1) * ns_] += eps * dirv
    hm[lag * ns_:(lag + 1) * ns_] -= eps * dirv
    _, rp = O.gls_assemble(*args, hp, nxt, jm, ci.size, jac=False)
    _, rm = O.gls_assemble(*args, hm, nxt, jm, ci.size, jac=False)
    fd = (rp @ adj - rm @ adj) / (2 * eps)
    jtw = O.gls_hist_vjp(x, y, conn, d9, ns_, 1.2, 0.03, 0.05, 1, 2, lag, hist, nxt, adj)
    assert _near(dirv @ jtw, fd, 2e-6)
    assert not jtw[2 * N:].any()


def test_hist_vjp_operator_adjoint_identity():
    """Assembled over a perturbed mesh: dir^T VJP_lag(adj) = adj^T JVP(dir in
    lag), the JVP assembled element by element from the dual-number mode."""
    x, y, conn = O.mesh_p1(4)
    rng = np.random.default_rng(6)
    inner = (x > 0) & (x < 1) & (y > 0) & (y < 1)
    x[inner] += 0.05 * rng.uniform(-1, 1, inner.sum())
    N = x.size
    ns_ = 3 * N
    d9 = O.gls_dofs(N, conn).reshape(-1, 9)
    hist = rng.standard_normal(2 * ns_)
    nxt, adj = rng.standard_normal(ns_), rng.standard_normal(ns_)
    for lag in (0, 1):
        dirv = rng.standard_normal(ns_)
        jtw = O.gls_hist_vjp(x, y, conn, d9.reshape(-1), ns_, 1.0, 0.01, 0.02, 2, 2, lag, hist,
                             nxt, adj)
        jv = np.zeros(ns_)
        c3 = conn.reshape(-1, 3)
        for e in range(c3.shape[0]):
            he = np.concatenate([hist[d9[e]], hist[ns_ + d9[e]]])
            de = np.zeros(18)
            de[9 * lag:9 * lag + 9] = dirv[d9[e]]
            _, dR = O.gls_element_jvp(x[c3[e]], y[c3[e]], 1.0, 0.01, 0.02, 2, 2, he, de, nxt[d9[e]])
            np.add.at(jv, d9[e], dR)
        assert abs(dirv @ jtw - adj @ jv) <= 1e-12 * max(1.0, abs(adj @ jv))
        jc = O.gls_hist_jvp(x, y, conn, d9.reshape(-1), ns_, 1.0, 0.01, 0.02, 2, 2, lag, hist, nxt,
                            dirv)
        assert np.allclose(jc, jv, rtol=1e-14, atol=1e-14 * np.abs(jv).max())


# ---------------------------------------------------------------------------
# constrained time residual (femx's NavierStokesModel under
# ConstrainedTimeResidual; fixture from the compiled reference)

def constrained_spec(g):
    ent = [(int(r), int(c), float(w)) for r, c, w in g["entries"]]
    nc = g["ctr_dofs"].size
    return dict(P=O.control_matrix(nc, int(g["num_ctr_params"]), ent), ctr_dofs=g["ctr_dofs"],
                fixed_dofs=g["fixed_dofs"], fixed_vals=g["fixed_vals"], time=g["time"],
                num_ctr_params=int(g["num_ctr_params"]), ctr_off=int(g["ctr_off"]))


def test_constrained_residual_oracle_equals_reference_bitwise():
    """The oracle's composition -- GLS assembly, controlVals, replaceRes,
    replaceRows, addControlJacT, J nxt - res, prepareForwardSolve -- against
    femx's ConstrainedTimeResidual and TimeIntegrator::solveStep sequence."""
    g = golden("gls_constrained_n4.npz")
    x, y, conn = g["x"], g["y"], g["conn"]
    N = x.size
    ns_ = 3 * N
    d9 = O.gls_dofs(N, conn)
    off = O.uniform_offsets(conn.size // 3, 9)
    rp, ci, _, jm = O.map_build(ns_, ns_, off, d9, off, d9)
    cm = constrained_spec(g)
    dofs = np.concatenate([g["ctr_dofs"], g["fixed_dofs"]]).astype(np.int32)
    bc = O.bc_build(rp, ci, dofs)
    args = (x, y, conn, d9, ns_, float(g["rho"]), float(g["mu"]), float(g["dt"]))
    for k in range(int(g["num_cases"])):
        step = int(g[f"step{k}"])
        hist, nxt, prm, adj = (g[f"{a}{k}"] for a in ("hist", "nxt", "prm", "adj"))
        vals, res = O.gls_assemble(*args, step, 2, hist, nxt, jm, ci.size)
        bv = O.control_vals(cm, step, prm)
        O.replace_res(bc, nxt, bv, res)
        assert np.array_equal(res, g[f"res{k}"]) and np.array_equal(res, g[f"asm_res{k}"])
        O.replace_rows(rp, bc, vals, 1.0)
        assert np.array_equal(vals, g[f"asm_vals{k}"])
        grad = O.add_control_jac_t(cm, step, adj, np.zeros(int(g["num_prm"])))
        assert np.array_equal(grad, g[f"grad{k}"])
        rhs = np.empty(ns_)
        O.matvec(rp, ci, vals, nxt, rhs)
        rhs = -1.0 * res + 1.0 * rhs
        O.prepare_forward(rp, bc, vals, rhs, bv)
        assert np.array_equal(vals, g[f"prep_vals{k}"]) and np.array_equal(rhs, g[f"prep_rhs{k}"])


def test_control_map_builders_follow_reference():
    from femx_amd import control as Ct
    with pytest.raises(RuntimeError, match="duplicate map entry"):
        Ct.DirichletControl([1, 2], 2, [(0, 1, 1.0), (0, 1, 2.0)])
    with pytest.raises(RuntimeError, match="duplicate state id"):
        Ct.DirichletControl([1, 1])
    with pytest.raises(RuntimeError, match="control column is out of range"):
        Ct.DirichletControl([1], 1, [(0, 3, 1.0)])
    ctr = Ct.DirichletControl([4, 2], 2, [(1, 0, 0.5), (0, 1, 2.0), (0, 0, 3.0)])
    assert ctr.pattern.row_ptr.tolist() == [0, 2, 3]
    assert ctr.pattern.col_ind.tolist() == [1, 0, 0] and ctr.vals.tolist() == [2.0, 3.0, 0.5]
    kw = dict(device="cpu")
    with pytest.raises(RuntimeError, match="controlled DOF is invalid"):
        Ct.ControlMap(3, 4, Ct.DirichletControl([4]), **kw)
    with pytest.raises(RuntimeError, match="fixed DOF is invalid"):
        Ct.ControlMap(3, 6, Ct.DirichletControl([4]), fixed_dofs=[4], **kw)
    with pytest.raises(RuntimeError, match="time stencil is invalid"):
        Ct.ControlMap(2, 6, ctr, time=[(0, 0, 0.5), (0, 1, 0.5)], **kw)
    with pytest.raises(RuntimeError, match="parameter count is too small"):
        Ct.ControlMap(2, 6, ctr, time=[(0, 1, 0.5), (1, 1, 0.0)], num_prm=3, **kw)
    cm = Ct.ControlMap(2, 6, ctr, fixed_dofs=[0, 5], fixed_vals=[1.0, 2.0],
                       time=[(0, 1, 0.5), (1, 1, 0.0)], ctr_off=1, **kw)
    assert cm.num_prm == 1 + 2 * 2 and cm.dofs.tolist() == [4, 2, 0, 5]
    assert cm.fixed_vals.tolist() == [1.0, 2.0, 1.0, 2.0]
