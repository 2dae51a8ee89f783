"""femx's own unstructured Gmsh meshes (data/meshes, read by femx's
GmshReader into the gmsh_*.npz fixtures) and high-valence fan meshes on the
GPU.

* P1 Poisson assembly (kappa = 1) and prepareForwardSolve on femx's boundary
  groups: bitwise equal to femx's own outputs; random kappa, f, u bitwise
  equal to the oracle; the log-kappa VJP within 1e-12 of the oracle.
* GLS NavierOperator time assembly and prepareForwardSolve of the Navier-
  Stokes system: bitwise equal to femx's outputs; the history VJP of both
  lags within 1e-12 of the oracle.
* Taylor-Hood on the P2 mesh built from femx's triangles (fx_th_mesh):
  assembly and VJP within 1e-12 of the oracle.
* Long rows: a node with 60 neighbours (a 61-entry P1 row), and row blocks
  whose CSR range is past the 32-entry-per-row LDS tile (the global-memory
  path), bitwise equal to the oracle; a Taylor-Hood fan (a 213-entry row)
  against the oracle, and a tight tile bound (many small blocks) bitwise
  equal to the default blocking."""
import numpy as np
import pytest
import torch

from conftest import close, golden
from femx_amd import assembly as A
from femx_amd import gls, ns
from femx_amd.assembly import make_assembly_map
from femx_amd.inverse import PoissonLogKappaResidual
from femx_amd.linalg import CsrMatrix
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
MESHES = ["2d_square_res_low", "2d_rectangle", "2d_square_res_very_low"]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", MESHES)
def test_gmsh_p1_bitwise_vs_femx_and_oracle(name):
    print("mesh", name, flush=True)
    g = golden(f"gmsh_{name}.npz")
    geom = A.Geometry(g["x"], g["y"], g["conn"].reshape(-1, 3), DEV)
    N, E = geom.num_nodes, geom.num_elems
    amap = A.make_assembly_map(N, N, geom.h_conn)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(N, dtype=torch.float64, device=DEV)
    A.assemble(A.PoissonP1(), geom, amap, dev(g["state"]), res, jac)
    assert np.array_equal(host(jac.vals), g["vals"]) and np.array_equal(host(res), g["res"])
    # prepareForwardSolve on femx's boundary groups
    bmap = A.make_boundary_map(g["bc_dofs"], amap.pattern, DEV)
    m = CsrMatrix(amap.pattern, DEV, dev(g["vals"]))
    rhs = dev(g["fwd_rhs0"])
    A.prepare_forward_solve(bmap, m, rhs, dev(g["fwd_g"]))
    assert np.array_equal(host(m.vals), g["fwd_vals"]) and np.array_equal(host(rhs), g["fwd_rhs"])
    # random coefficients against the oracle, and the VJP
    rng = np.random.default_rng(7)
    kappa, f, u = np.exp(0.5 * rng.standard_normal(E)), rng.standard_normal(N), rng.standard_normal(N)
    A.assemble(A.PoissonP1(dev(kappa), dev(f)), geom, amap, dev(u), res, jac)
    _, jm = amap.jac()
    ov, orr = O.p1_assemble(geom.h_x, geom.h_y, geom.h_conn.reshape(-1), kappa, f, u, jm,
                            amap.pattern.nnz)
    assert np.array_equal(host(jac.vals), ov) and np.array_equal(host(res), orr)
    R = PoissonLogKappaResidual(geom, g["bc_dofs"], amap=amap)
    mlog, lam = 0.5 * rng.standard_normal(E), rng.standard_normal(N)
    jt, ge = R.vjp(dev(u), dev(mlog), dev(lam))
    ojt, og = O.p1_vjp(geom.h_x, geom.h_y, geom.h_conn.reshape(-1), mlog, u, lam, R.bmap.bc_mask)
    assert close(host(jt), ojt) and close(host(ge), og)


@pytest.mark.parametrize("name", ["2d_rectangle", "2d_square_res_very_low"])
def test_gmsh_gls_bitwise_vs_femx(name):
    print("mesh", name, flush=True)
    g = golden(f"gmsh_{name}.npz")
    geom = gls.GLSGeometry(g["x"], g["y"], g["conn"], DEV)
    amap = gls.make_navier_map(geom)
    op = gls.NavierOperator(geom, gls.FluidParams(1.0, 1e-3), 0.01)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(geom.num_states, dtype=torch.float64, device=DEV)
    gls.assemble(op, 1, 2, gls.NEXT_STATE, amap, dev(g["ns_hist"]), dev(g["ns_nxt"]), res, jac)
    assert np.array_equal(host(jac.vals), g["ns_vals"]) and np.array_equal(host(res), g["ns_res"])
    bmap = A.make_boundary_map(g["ns_bc_dofs"], amap.pattern, DEV)
    rhs = dev(g["ns_fwd_rhs0"])
    A.prepare_forward_solve(bmap, jac, rhs, dev(g["ns_fwd_g"]))
    assert np.array_equal(host(jac.vals), g["ns_fwd_vals"])
    assert np.array_equal(host(rhs), g["ns_fwd_rhs"])
    rng = np.random.default_rng(11)
    adj = rng.standard_normal(geom.num_states)
    for lag in (0, 1):
        out = host(gls.hist_vjp(op, 1, 2, lag, amap, dev(g["ns_hist"]), dev(g["ns_nxt"]), dev(adj)))
        ref = O.gls_hist_vjp(g["x"], g["y"], g["conn"], g["ns_dofs9"], geom.num_states, 1.0, 1e-3,
                             0.01, 1, 2, lag, g["ns_hist"], g["ns_nxt"], adj)
        assert close(out, ref), lag


def th_check(gh, seed, plan_tile=None):
    """Assembly and VJP of a Taylor-Hood mesh (host geometry gh) against the
    oracle; returns the GPU CSR values and residual."""
    g = ns.THGeometry(gh.h_xv, gh.h_yv, gh.h_vconn3, gh.h_conn6, gh.h_dofs15, gh.num_vel_nodes, DEV)
    rng = np.random.default_rng(seed)
    s, lam = rng.standard_normal(g.ndof), rng.standard_normal(g.ndof)
    nu_e = rng.uniform(0.005, 0.05, g.num_elems)
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
    ns.assemble_ns(ns.TaylorHoodNS(0.0, dev(nu_e)), g, amap, dev(s), res, jac)
    _, jm = amap.jac()
    ov, orr = O.th_assemble(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1),
                            g.ndof, 0.0, s, jm, amap.pattern.nnz, nu_e=nu_e)
    v, r = host(jac.vals), host(res)
    assert close(v, ov) and close(r, orr)
    R = ns.NavierStokesResidual(g, None, amap=amap)
    jt, ge = R.vjp(dev(s), dev(nu_e), dev(lam))
    ojt, og = O.th_vjp(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1), g.ndof,
                       0.0, s, lam, None, nu_e=nu_e)
    assert close(host(jt), ojt) and close(host(ge), og)
    return g, amap, s, nu_e, v, r


@pytest.mark.parametrize("name", ["2d_rectangle", "2d_square_res_low"])
def test_gmsh_taylor_hood_vs_oracle(name):
    print("mesh", name, flush=True)
    m = golden(f"gmsh_{name}.npz")
    th_check(ns.th_from_p1(m["x"], m["y"], m["conn"], "cpu"), 3)


def fan(m, rings=2):
    """A disc: centre node 0, `rings` rings of m nodes, the inner ring joined
    to the centre (valence m), ring-to-ring quads split into two triangles."""
    x, y = [0.0], [0.0]
    for k in range(1, rings + 1):
        t = 2 * np.pi * (np.arange(m) + 0.5 * (k % 2)) / m
        x += list(k * np.cos(t))
        y += list(k * np.sin(t))
    conn = [[0, 1 + i, 1 + (i + 1) % m] for i in range(m)]
    for k in range(1, rings):
        a, b = 1 + (k - 1) * m, 1 + k * m
        for i in range(m):
            j = (i + 1) % m
            conn += [[a + i, b + i, b + j], [a + i, b + j, a + j]]
    return np.array(x), np.array(y), np.array(conn, np.int32)


def hubs(nh=160, m=40):
    """nh hubs (nodes 0..nh-1), each the centre of its own wheel of m rim
    nodes: the first row block's 128 rows hold 41 entries each, so its CSR
    range (5 248 entries) is past the 4 096-entry LDS tile."""
    x, y, conn = list(3.0 * np.arange(nh)), [0.0] * nh, []
    t = 2 * np.pi * np.arange(m) / m
    for h in range(nh):
        base = len(x)
        x += list(3.0 * h + np.cos(t))
        y += list(np.sin(t))
        conn += [[h, base + i, base + (i + 1) % m] for i in range(m)]
    return np.array(x), np.array(y), np.array(conn, np.int32)


@pytest.mark.parametrize("mesh", ["fan", "hubs"])
def test_p1_long_rows_bitwise(mesh):
    """A 61-entry row inside the tile (fan), and whole row blocks past the
    tile, accumulated in global memory (hubs)."""
    x, y, conn = fan(60) if mesh == "fan" else hubs()
    geom = A.Geometry(x, y, conn, DEV)
    N, E = geom.num_nodes, geom.num_elems
    amap = A.make_assembly_map(N, N, conn)
    assert amap.max_row_len == (61 if mesh == "fan" else 41)
    if mesh == "hubs":
        assert amap.pattern.row_ptr[128] > 128 * 32
    rng = np.random.default_rng(4)
    kappa, f, u = np.exp(0.5 * rng.standard_normal(E)), rng.standard_normal(N), rng.standard_normal(N)
    jac = CsrMatrix(amap.pattern, DEV)
    jac.vals.fill_(np.nan)
    res = torch.empty(N, dtype=torch.float64, device=DEV)
    A.assemble(A.PoissonP1(dev(kappa), dev(f)), geom, amap, dev(u), res, jac)
    _, jm = amap.jac()
    ov, orr = O.p1_assemble(x, y, conn.reshape(-1), kappa, f, u, jm, amap.pattern.nnz)
    assert np.array_equal(host(jac.vals), ov) and np.array_equal(host(res), orr)
    R = PoissonLogKappaResidual(geom, None, amap=amap)
    mlog, lam = 0.5 * rng.standard_normal(E), rng.standard_normal(N)
    jt, ge = R.vjp(dev(u), dev(mlog), dev(lam))
    ojt, og = O.p1_vjp(x, y, conn.reshape(-1), mlog, u, lam, None)
    assert close(host(jt), ojt) and close(host(ge), og)


def test_taylor_hood_fan_and_tile_bound():
    from femx_amd import _lib
    from femx_amd.linalg import ptr, stream_handle
    x, y, conn = fan(30, rings=3)
    g, amap, s, nu_e, v, r = th_check(ns.th_from_p1(x, y, conn, "cpu"), 5)
    assert amap.max_row_len == 7 * 30 + 3
    p = ns.THPlan(amap, DEV, tile_nnz=600)  # many small blocks
    assert p.num_blocks > ns._plan(amap, DEV).num_blocks
    rp, _ = amap.pattern.device_arrays(DEV)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
    d_s, d_nu = dev(s), dev(nu_e)  # held until the launch has run
    _lib.call("fx_th_assemble", g.ndof, g.num_elems, p.num_blocks, p.max_block_nnz, ptr(g.xv),
              ptr(g.yv), ptr(g.vconn3), ptr(g.dofs15), ptr(d_s), 0.0, ptr(d_nu),
              ptr(p.inc_off), ptr(p.inc_elem), ptr(p.inc_meta), ptr(p.blk_row), ptr(p.blk_flags),
              ptr(p.blk_lanes), ptr(rp), ptr(jac.vals), ptr(res), stream_handle())
    assert np.array_equal(host(jac.vals), v) and np.array_equal(host(res), r)
    # the round tables of the same small-block layout (rows of up to 213 entries, vertices of
    # 30 elements: rounds with many idle lanes)
    jac2 = CsrMatrix(amap.pattern, DEV)
    res2 = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
    _lib.call("fx_th_assemble_rounds", g.ndof, g.num_elems, p.num_blocks, p.max_block_nnz,
              ptr(g.xv), ptr(g.yv), ptr(g.vconn3), ptr(g.dofs15), ptr(d_s), 0.0, ptr(d_nu),
              ptr(p.blk_row), ptr(p.blk_flags), ptr(p.blk_lanes), ptr(p.rt_off), ptr(p.rt_elem),
              ptr(p.rt_meta), ptr(rp), ptr(jac2.vals), ptr(res2), stream_handle())
    assert np.array_equal(host(jac2.vals), v) and np.array_equal(host(res2), r)
