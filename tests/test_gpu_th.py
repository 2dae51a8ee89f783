"""Taylor-Hood P2/P1 Navier-Stokes on the GPU against the oracle (1e-12
mixed norm: the kernels use fused multiply-adds), finite differences of the
GPU residual, autograd gradcheck, run-to-run determinism, and the config-3
size (512^2 cells, 524,288 elements)."""
import numpy as np
import pytest
import torch

from conftest import close
from femx_amd import ns
from femx_amd.assembly import make_assembly_map
from femx_amd.linalg import CsrMatrix, MatrixHandler
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
RNG = np.random.default_rng(9)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def setup(n):
    g = ns.structured_th(n, DEV)
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    return g, amap


@pytest.mark.parametrize("n,per_elem", [(1, False), (2, True), (7, False)])
def test_th_assemble_matches_oracle(n, per_elem):
    g, amap = setup(n)
    s = RNG.standard_normal(g.ndof)
    nu = 0.01
    nu_e = RNG.uniform(0.0025, 0.1, g.num_elems) if per_elem else None
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
    ns.assemble_ns(ns.TaylorHoodNS(nu, dev(nu_e) if per_elem else None), g, amap, dev(s), res, jac)
    _, jm = amap.jac()
    ov, orr = O.th_assemble(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1),
                            g.ndof, nu, s, jm, amap.pattern.nnz, nu_e=nu_e)
    assert close(host(jac.vals), ov)
    assert close(host(res), orr)
    res2 = torch.full((g.ndof,), np.nan, dtype=torch.float64, device=DEV)
    ns.assemble_ns(ns.TaylorHoodNS(nu, dev(nu_e) if per_elem else None), g, amap, dev(s), res2)
    assert np.array_equal(host(res2), host(res))


def test_th_lane_map_and_two_lane_default_agree():
    """fx_th_assemble with the block lane map (fx_map_th_lanes) and without
    one (two lanes per item) add every entry's terms in the same order: the
    results are bitwise equal."""
    from femx_amd import _lib
    from femx_amd.linalg import ptr, stream_handle
    g, amap = setup(9)
    s = dev(RNG.standard_normal(g.ndof))
    p = ns.THPlan(amap, DEV)
    rp, _ = amap.pattern.device_arrays(DEV)
    out = []
    for lanes in (p.blk_lanes, None):
        jac = CsrMatrix(amap.pattern, DEV)
        res = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
        _lib.call("fx_th_assemble", g.ndof, g.num_elems, p.num_blocks, p.max_block_nnz, ptr(g.xv),
                  ptr(g.yv), ptr(g.vconn3), ptr(g.dofs15), ptr(s), 0.01, None, ptr(p.inc_off),
                  ptr(p.inc_elem), ptr(p.inc_meta), ptr(p.blk_row), ptr(p.blk_flags), ptr(lanes),
                  ptr(rp), ptr(jac.vals), ptr(res), stream_handle())
        out.append((host(jac.vals), host(res)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("n,items,tile,relabel", [(9, 32, 0, None), (9, 16, 0, None),
                                                  (9, 24, 0, None), (9, 32, 1500, None),
                                                  (7, 32, 0, True), (1, 32, 0, None)])
def test_th_round_tables_equal_lane_map_kernel(n, items, tile, relabel):
    """fx_th_assemble_rounds (each round's incidences from the block's round
    table) and fx_th_assemble (through the lane map and row offsets) add the
    same terms in the same order: Jacobian and residual bitwise equal, with
    and without the Jacobian, for several block layouts and a perturbed,
    relabelled mesh."""
    from femx_amd import _lib
    from femx_amd.linalg import ptr, stream_handle
    if relabel is None:
        g, amap = setup(n)
    else:
        g = perturbed(n, 4, relabel)
        amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    s = dev(RNG.standard_normal(g.ndof))
    p = ns.THPlan(amap, DEV, items=items, tile_nnz=tile)
    rp, _ = amap.pattern.device_arrays(DEV)
    for with_jac in (True, False):
        out = []
        for rounds in (False, True):
            jac = CsrMatrix(amap.pattern, DEV)
            res = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
            head = (g.ndof, g.num_elems, p.num_blocks, p.max_block_nnz, ptr(g.xv), ptr(g.yv),
                    ptr(g.vconn3), ptr(g.dofs15), ptr(s), 0.01, None)
            tail = (ptr(rp), ptr(jac.vals) if with_jac else None, ptr(res), stream_handle())
            if rounds:
                _lib.call("fx_th_assemble_rounds", *head, ptr(p.blk_row), ptr(p.blk_flags),
                          ptr(p.blk_lanes), ptr(p.rt_off), ptr(p.rt_elem), ptr(p.rt_meta), *tail)
            else:
                _lib.call("fx_th_assemble", *head, ptr(p.inc_off), ptr(p.inc_elem),
                          ptr(p.inc_meta), ptr(p.blk_row), ptr(p.blk_flags), ptr(p.blk_lanes),
                          *tail)
            out.append((host(jac.vals), host(res)))
        assert np.array_equal(out[0][1], out[1][1])
        if with_jac:
            assert np.array_equal(out[0][0], out[1][0])


def test_th_full_size_matches_oracle_and_is_deterministic():
    """Config 3: 512^2 cavity, Re = 100, Fourier state with lid.  The
    round-table form (waves walking four blocks at this size) and the
    lane-map form give the same bits; both match the oracle."""
    g, amap = setup(512)
    assert g.ndof == 2364419 and amap.pattern.nnz == 70048777
    s = ns.fourier_state(g, 5)
    jac = CsrMatrix(amap.pattern, DEV)
    res = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
    op = ns.TaylorHoodNS(1.0 / 100.0)
    ns.assemble_ns(op, g, amap, dev(s), res, jac)
    v1, r1 = host(jac.vals), host(res)
    ns.assemble_ns(op, g, amap, dev(s), res, jac)
    assert np.array_equal(host(jac.vals), v1) and np.array_equal(host(res), r1)
    ns.assemble_ns(op, g, amap, dev(s), res, jac, rounds=False)
    assert np.array_equal(host(jac.vals), v1) and np.array_equal(host(res), r1)
    _, jm = amap.jac()
    ov, orr = O.th_assemble(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1),
                            g.ndof, 0.01, s, jm, amap.pattern.nnz, omp=True)
    assert close(v1, ov) and close(r1, orr)


def perturbed(n, seed, relabel):
    """Structured mesh with interior vertices moved by up to 0.3 h; with
    `relabel`, every other element is listed clockwise (vertices 0,2,1, edges
    (0,2),(2,1),(1,0)), so its det J is negative."""
    g0 = ns.structured_th(n, DEV)
    rng = np.random.default_rng(seed)
    xv, yv = g0.h_xv.copy(), g0.h_yv.copy()
    inner = (xv > 0) & (xv < 1) & (yv > 0) & (yv < 1)
    xv[inner] += rng.uniform(-0.3, 0.3, inner.sum()) / n
    yv[inner] += rng.uniform(-0.3, 0.3, inner.sum()) / n
    v3, c6, d = g0.h_vconn3.copy(), g0.h_conn6.copy(), g0.h_dofs15.copy()
    if relabel:
        f = np.arange(g0.num_elems) % 2 == 1
        P = np.array([0, 2, 1, 5, 4, 3])
        v3[f] = v3[f][:, [0, 2, 1]]
        c6[f] = c6[f][:, P]
        perm = np.concatenate([np.stack([2 * P, 2 * P + 1], 1).reshape(-1), 12 + np.array([0, 2, 1])])
        d[f] = d[f][:, perm]
    return ns.THGeometry(xv, yv, v3, c6, d, g0.num_vel_nodes, DEV)


def test_th_general_affine_and_relabelling_invariance():
    """Skewed elements of both orientations against the oracle, and the
    assembled operator independent of the elements' local numbering."""
    n, nu = 5, 0.02
    ga, gb = perturbed(n, 4, False), perturbed(n, 4, True)
    s = RNG.standard_normal(ga.ndof)
    out = {}
    for key, g in (("a", ga), ("b", gb)):
        amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
        jac = CsrMatrix(amap.pattern, DEV)
        res = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
        ns.assemble_ns(ns.TaylorHoodNS(nu), g, amap, dev(s), res, jac)
        _, jm = amap.jac()
        ov, orr = O.th_assemble(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1),
                                g.ndof, nu, s, jm, amap.pattern.nnz)
        assert close(host(jac.vals), ov) and close(host(res), orr)
        out[key] = (amap.pattern, host(jac.vals), host(res))
    (pa, va, ra), (pb, vb, rb) = out["a"], out["b"]
    assert np.array_equal(pa.row_ptr, pb.row_ptr) and np.array_equal(pa.col_ind, pb.col_ind)
    assert close(vb, va) and close(rb, ra)
    lam = RNG.standard_normal(gb.ndof)
    R = ns.NavierStokesResidual(gb, None)
    jt, ge = R.vjp(dev(s), dev(np.full(gb.num_elems, nu)), dev(lam))
    ojt, og = O.th_vjp(gb.h_xv, gb.h_yv, gb.h_vconn3.reshape(-1), gb.h_dofs15.reshape(-1), gb.ndof,
                       nu, s, lam, None, nu_e=np.full(gb.num_elems, nu))
    assert close(host(jt), ojt) and close(host(ge), og)


@pytest.mark.parametrize("n,with_bc", [(2, False), (6, True)])
def test_th_vjp_matches_oracle(n, with_bc):
    g, amap = setup(n)
    s = RNG.standard_normal(g.ndof)
    lam = RNG.standard_normal(g.ndof)
    nu_e = RNG.uniform(0.005, 0.05, g.num_elems)
    R = ns.NavierStokesResidual(g, ns.cavity_bc(g)[0] if with_bc else None, amap=amap)
    jt, ge = R.vjp(dev(s), dev(nu_e), dev(lam))
    mask = R.bmap.bc_mask if with_bc else None
    ojt, og = O.th_vjp(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1), g.ndof,
                       0.0, s, lam, mask, nu_e=nu_e)
    assert close(host(jt), ojt) and close(host(ge), og)


def test_th_residual_finite_differences():
    g, amap = setup(3)
    dofs, vals = ns.cavity_bc(g)
    R = ns.NavierStokesResidual(g, dofs, dev(vals), amap=amap)
    s = dev(ns.fourier_state(g, 3))
    nu_e = dev(RNG.uniform(0.005, 0.05, g.num_elems))
    lam = dev(RNG.standard_normal(g.ndof))
    jt, ge = (host(t) for t in R.vjp(s, nu_e, lam))
    out = torch.empty(g.ndof, dtype=torch.float64, device=DEV)

    def lr(ss, nn):
        R.res(ss, nn, out)
        return float(host(torch.dot(lam, out)))

    h = 1e-6
    for k in list(RNG.choice(g.ndof, 10, replace=False)) + [int(dofs[0])]:
        e = torch.zeros(g.ndof, dtype=torch.float64, device=DEV)
        e[k] = h
        fd = (lr(s + e, nu_e) - lr(s - e, nu_e)) / (2 * h)
        assert abs(jt[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, jt[k], fd)
    for k in RNG.choice(g.num_elems, 6, replace=False):
        e = torch.zeros(g.num_elems, dtype=torch.float64, device=DEV)
        e[k] = h
        fd = (lr(s, nu_e + e) - lr(s, nu_e - e)) / (2 * h)
        assert abs(ge[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, ge[k], fd)
    # assembled Jacobian times a direction against a central difference
    jac = CsrMatrix(amap.pattern, DEV)
    R.assemble_state_jac(s, nu_e, jac)
    dirv = dev(RNG.standard_normal(g.ndof))
    y = torch.empty(g.ndof, dtype=torch.float64, device=DEV)
    MatrixHandler().matvec(jac, dirv, y)
    rp, rm = torch.empty_like(out), torch.empty_like(out)
    R.res(s + 1e-6 * dirv, nu_e, rp)
    R.res(s - 1e-6 * dirv, nu_e, rm)
    fd = host((rp - rm) / 2e-6)
    assert np.abs(host(y) - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())


def test_th_autograd_gradcheck():
    g, amap = setup(1)
    R = ns.NavierStokesResidual(g, ns.cavity_bc(g)[0][:6], amap=amap)
    u = dev(RNG.standard_normal(g.ndof)).requires_grad_()
    nu_e = dev(RNG.uniform(0.01, 0.1, g.num_elems)).requires_grad_()
    assert torch.autograd.gradcheck(lambda a, b: ns.NSResidualVJP.apply(a, b, R), (u, nu_e),
                                    eps=1e-6, atol=1e-7, rtol=1e-6)
