"""The P1 log-kappa VJP over element patches (fx_p1_vjp_patch) through the
C-ABI, against the row-gather fx_p1_vjp and the oracle: a node whose
elements lie in one patch and every g_e bitwise equal to fx_p1_vjp, seam
nodes within the 1e-12 mixed norm; Dirichlet nodes, scrambled element
order, a hub vertex with more elements than one patch may hold, patch sizes
from 1 to 256, repeated launches bitwise equal, and the 1024^2 mesh."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import close
from femx_amd import _lib
from femx_amd.assembly import Geometry, make_assembly_map, structured_p1, unit_square_edge_nodes
from femx_amd.linalg import ptr
from femx_amd.patches import Patches
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def empties(N, E):
    return (torch.full((N,), float("nan"), dtype=torch.float64, device=DEV),
            torch.full((E,), float("nan"), dtype=torch.float64, device=DEV))


def row_gather(g, m, u, lam, mask):
    amap = make_assembly_map(g.num_nodes, g.num_nodes, g.h_conn)
    jt, ge = empties(g.num_nodes, g.num_elems)
    _lib.call("fx_p1_vjp", g.num_nodes, g.num_elems, ptr(g.x), ptr(g.y), ptr(g.conn), ptr(m),
              ptr(u), ptr(lam), ptr(mask), ptr(amap.device("inc_off", DEV)),
              ptr(amap.device("inc", DEV)), ptr(jt), ptr(ge), None)
    return host(jt), host(ge)


def patched(g, m, u, lam, mask, pp):
    jt, ge = empties(g.num_nodes, g.num_elems)
    _lib.call("fx_p1_vjp_patch", g.num_nodes, g.num_elems, ptr(g.x), ptr(g.y), ptr(m), ptr(u),
              ptr(lam), ptr(mask), C.byref(pp.plan), ptr(jt), ptr(ge), None)
    return host(jt), host(ge)


def check(g, mask_nodes, patch_sizes, seed, oracle=True):
    rng = np.random.default_rng(seed)
    N, E = g.num_nodes, g.num_elems
    m = 0.5 * rng.standard_normal(E)
    u = rng.standard_normal(N)
    lam = rng.standard_normal(N)
    hmask = None
    if mask_nodes is not None:
        hmask = np.zeros(N, np.int32)
        hmask[mask_nodes] = 1
    d = (dev(m), dev(u), dev(lam), dev(hmask) if hmask is not None else None)
    rjt, rg = row_gather(g, *d)
    if oracle:
        ojt, og = O.p1_vjp(g.h_x, g.h_y, g.h_conn.reshape(-1), m, u, lam, hmask)
        assert close(rjt, ojt) and close(rg, og)
    for mp in patch_sizes:
        pp = Patches(N, E, 3, g.h_conn, g.h_conn, g.h_x, g.h_y, DEV, mp)
        jt, ge = patched(g, *d, pp)
        assert np.array_equal(ge, rg), f"max_patch {mp}"
        assert close(jt, rjt), f"max_patch {mp}"
        inner = np.ones(N, bool)
        inner[pp.h["seam_dof"][:pp.num_seam]] = False
        assert np.array_equal(jt[inner], rjt[inner]), f"max_patch {mp}"
        jt2, ge2 = patched(g, *d, pp)
        assert np.array_equal(jt2, jt) and np.array_equal(ge2, ge)


def test_p1_patch_vjp_structured_with_dirichlet_nodes():
    g = structured_p1(12, DEV)
    check(g, unit_square_edge_nodes(g.h_x, g.h_y), (1, 5, 64, 256), 1)


def test_p1_patch_vjp_scrambled_and_hub():
    rng = np.random.default_rng(2)
    x, y, c3 = O.mesh_p1(10)
    c3 = c3.reshape(-1, 3)[rng.permutation(200)]
    check(Geometry(x, y, c3, DEV), np.array([0, 7, 50]), (16, 256), 2)
    # a hub vertex in 30 elements: its runs are capped at 16 per patch
    k = 30
    ang = 2 * np.pi * np.arange(k) / k
    xh = np.concatenate([[0.0], np.cos(ang)])
    yh = np.concatenate([[0.0], np.sin(ang)])
    ch = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32)
    gh = Geometry(xh, yh, ch, DEV)
    pp = Patches(gh.num_nodes, gh.num_elems, 3, gh.h_conn, gh.h_conn, gh.h_x, gh.h_y, DEV)
    assert pp.num_patches >= 2 and 0 in pp.h["seam_dof"][:pp.num_seam]
    check(gh, np.array([0, 3]), (256,), 3)


def test_p1_patch_vjp_full_size():
    """The 1024^2 mesh of the benchmark's VJP (edges constrained)."""
    g = structured_p1(1024, DEV)
    check(g, unit_square_edge_nodes(g.h_x, g.h_y), (256,), 4, oracle=False)
