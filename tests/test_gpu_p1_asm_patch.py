"""The P1 assembly over owner-computes element patches
(fx_p1_assemble_patch) and over 64-row stencil blocks
(fx_p1_assemble_stencil) through the C-ABI against the row-per-lane
fx_p1_assemble: CSR values and residuals bitwise equal (all sum every entry
in ascending element order) on structured, scrambled, unused-node and hub
meshes, small patches and small tiles (rows left to the row kernel), log and
plain kappa, without load or state, residual only, repeated launches, and
the 2048^2 mesh, which is also checked against the oracle."""
import numpy as np
import pytest
import torch

from femx_amd import _lib
from femx_amd import assembly as A
from femx_amd.linalg import CsrMatrix, ptr
from femx_amd.patches import P1AsmPlan
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def both(g, amap, op, u, jac=True, paths=((False, False), (True, False))):
    out = []
    for patches, stencil in paths:
        res = torch.full((g.num_nodes,), float("nan"), dtype=torch.float64, device=DEV)
        J = CsrMatrix(amap.pattern, DEV) if jac else None
        if J is not None:
            J.vals.fill_(float("nan"))
        A.assemble(op, g, amap, u, res, J, patches=patches, stencil=stencil)
        out.append((host(J.vals) if jac else None, host(res)))
    return out


def check(x, y, conn, seed, max_patch=256, max_tile=P1AsmPlan.MAX_TILE):
    print("case", x.size, len(conn), seed, max_patch, max_tile, flush=True)
    g = A.Geometry(x, y, conn, DEV)
    amap = A.make_assembly_map(g.num_nodes, g.num_nodes, g.h_conn)
    if max_patch != 256 or max_tile != P1AsmPlan.MAX_TILE:
        setattr(amap, "_p1_asm_" + str(g.device),
                (g, P1AsmPlan(g.h_conn, g.h_x, g.h_y, amap.pattern, DEV, max_patch, max_tile)))
    rng = np.random.default_rng(seed)
    m = dev(0.5 * rng.standard_normal(g.num_elems))
    kap = dev(np.exp(0.5 * rng.standard_normal(g.num_elems)))
    f = dev(rng.standard_normal(g.num_nodes))
    u = dev(rng.standard_normal(g.num_nodes))
    cases = [(A.PoissonP1(f=f, log_kappa=m), u, True), (A.PoissonP1(f=f, kappa=kap), u, True),
             (A.PoissonP1(log_kappa=m), None, True), (A.PoissonP1(f=f, log_kappa=m), u, False)]
    for op, uu, jac in cases:
        (rv, rr), (pv, pr) = both(g, amap, op, uu, jac)
        assert np.array_equal(rr, pr)
        if jac:
            assert np.array_equal(rv, pv)
    # repeated launches bitwise equal
    (_, _), (pv2, pr2) = both(g, amap, cases[0][0], u)
    (_, _), (pv3, pr3) = both(g, amap, cases[0][0], u)
    assert np.array_equal(pv2, pv3) and np.array_equal(pr2, pr3)
    return g, amap


def test_patch_assembly_structured_and_small_patches():
    g = A.structured_p1(20, "cpu")
    check(g.h_x, g.h_y, g.h_conn, 1)
    check(g.h_x, g.h_y, g.h_conn, 2, max_patch=13)
    check(g.h_x, g.h_y, g.h_conn, 3, max_tile=30)
    # patches of 5 elements own no interior node: rows left to the row kernel
    check(g.h_x, g.h_y, g.h_conn, 4, max_patch=5)


def test_patch_assembly_scrambled_unused_node_and_hub():
    rng = np.random.default_rng(5)
    x, y, c3 = O.mesh_p1(12)
    c3 = c3.reshape(-1, 3)[rng.permutation(2 * 144)]
    check(np.append(x, 2.0), np.append(y, 2.0), c3, 4, max_patch=50)
    m = 30
    ang = 2 * np.pi * np.arange(m) / m
    xh = np.concatenate([[0.0], np.cos(ang)])
    yh = np.concatenate([[0.0], np.sin(ang)])
    ch = np.array([[0, 1 + i, 1 + (i + 1) % m] for i in range(m)], np.int32)
    check(xh, yh, ch, 5)


STENCIL = ((False, False), (False, True))  # the row kernel, then the stencil-block kernel


def check_stencil(n, seed):
    """fx_p1_assemble_stencil against fx_p1_assemble on the structured n x n
    mesh: uniform 7- and 5-point blocks beside row-gather blocks, log and
    plain kappa, no load or state, residual only."""
    print("stencil case", n, seed, flush=True)
    g = A.structured_p1(n, DEV)
    amap = A.make_assembly_map(g.num_nodes, g.num_nodes, g.h_conn)
    sp = A.p1_stencil_plan(g, amap)
    rng = np.random.default_rng(seed)
    m = dev(0.5 * rng.standard_normal(g.num_elems))
    kap = dev(np.exp(0.5 * rng.standard_normal(g.num_elems)))
    f = dev(rng.standard_normal(g.num_nodes))
    u = dev(rng.standard_normal(g.num_nodes))
    for op, uu, jac in ((A.PoissonP1(f=f, log_kappa=m), u, True),
                        (A.PoissonP1(f=f, kappa=kap), u, True),
                        (A.PoissonP1(kappa=kap), None, True),
                        (A.PoissonP1(f=f, log_kappa=m), u, False)):
        (rv, rr), (sv, sr) = both(g, amap, op, uu, jac, STENCIL)
        assert np.array_equal(rr, sr)
        if jac:
            assert np.array_equal(rv, sv)
    return sp


def test_stencil_assembly_matches_row_kernel():
    # 300 and 129: interior and bottom/top-row stencil blocks, partial last
    # block; 40: no uniform block (every block holds a boundary node)
    for n, seed in ((300, 1), (129, 2)):
        sp = check_stencil(n, seed)
        assert 0.3 < sp.frac < 1.0
    assert check_stencil(40, 3).frac == 0.0


def test_stencil_assembly_record_over_band_buffer_takes_row_path():
    """A record whose band buffer would not fit the kernel's (here: every
    record's band slots moved up by 100) sends its blocks down the row path:
    the results are still bitwise equal to fx_p1_assemble."""
    n = 129
    g = A.structured_p1(n, DEV)
    amap = A.make_assembly_map(g.num_nodes, g.num_nodes, g.h_conn)
    sp = A.p1_stencil_plan(g, amap)
    words = sp.h_words.copy()
    for c in np.unique(sp.h_code[sp.h_code >= 0]):
        S = words[c]
        words[c + 3 + S:c + 3 + 2 * S] += 100
    rng = np.random.default_rng(9)
    kap = dev(np.exp(0.5 * rng.standard_normal(g.num_elems)))
    f, u = dev(rng.standard_normal(g.num_nodes)), dev(rng.standard_normal(g.num_nodes))
    rp, _ = amap.pattern.device_arrays(DEV)
    out = []
    for name, extra in (("fx_p1_assemble_stencil", (ptr(sp.blk_code), ptr(dev(words)))),
                        ("fx_p1_assemble", ())):
        J = CsrMatrix(amap.pattern, DEV)
        J.vals.fill_(float("nan"))
        res = torch.full((g.num_nodes,), float("nan"), dtype=torch.float64, device=DEV)
        _lib.call(name, g.num_nodes, g.num_elems, amap.pattern.nnz, amap.max_row_len, ptr(g.x),
                  ptr(g.y), ptr(g.conn), ptr(kap), 0, ptr(f), ptr(u),
                  ptr(amap.device("inc_off", DEV)), ptr(amap.device("inc", DEV)), ptr(rp), *extra,
                  ptr(J.vals), ptr(res), None)
        out.append((host(J.vals), host(res)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_patch_assembly_full_size_matches_row_kernel_and_oracle():
    """2048^2 (8.4 M elements): bitwise equal to fx_p1_assemble and to the
    oracle's element loop."""
    n = 2048
    g = A.structured_p1(n, DEV)
    amap = A.make_assembly_map(g.num_nodes, g.num_nodes, g.h_conn)
    rng = np.random.default_rng(11)
    hm = 0.5 * rng.standard_normal(g.num_elems)
    hk = np.exp(0.5 * rng.standard_normal(g.num_elems))
    hf = rng.standard_normal(g.num_nodes)
    hu = rng.standard_normal(g.num_nodes)
    (rv, rr), (pv, pr) = both(g, amap, A.PoissonP1(f=dev(hf), log_kappa=dev(hm)), dev(hu))
    assert np.array_equal(rv, pv) and np.array_equal(rr, pr)
    # plain kappa against the oracle (a host exp could differ in the last bit)
    (rv, rr), (pv, pr) = both(g, amap, A.PoissonP1(f=dev(hf), kappa=dev(hk)), dev(hu))
    assert np.array_equal(rv, pv) and np.array_equal(rr, pr)
    ov, orr = O.p1_assemble(g.h_x, g.h_y, g.h_conn.reshape(-1), hk, hf, hu, amap.jac()[1],
                            amap.pattern.nnz)
    assert np.array_equal(pv, ov) and np.array_equal(pr, orr)
    # the stencil-block kernel (97% of the rows in uniform blocks here)
    assert A.p1_stencil_plan(g, amap).frac > 0.9
    (_, _), (sv, sr) = both(g, amap, A.PoissonP1(f=dev(hf), kappa=dev(hk)), dev(hu),
                            paths=STENCIL)
    assert np.array_equal(sv, ov) and np.array_equal(sr, orr)
    (_, _), (sv, sr) = both(g, amap, A.PoissonP1(f=dev(hf), log_kappa=dev(hm)), dev(hu),
                            paths=STENCIL)
    (rv, rr), (_, _) = both(g, amap, A.PoissonP1(f=dev(hf), log_kappa=dev(hm)), dev(hu))
    assert np.array_equal(sv, rv) and np.array_equal(sr, rr)
