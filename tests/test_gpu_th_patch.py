"""The single-pass Taylor-Hood VJP over element patches (fx_th_vjp_patch)
through the C-ABI, against the oracle (1e-12 mixed norm) and against the
two-phase fx_th_vjp (same element columns, other per-DOF summation order;
g_e bitwise equal): structured with Dirichlet rows, skewed with half the
elements listed clockwise, scrambled element order with an unused vertex,
patch sizes from 1 to 256; repeated launches bitwise equal; the 512^2
cavity."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import close
from femx_amd import _lib, ns
from femx_amd.assembly import make_assembly_map
from femx_amd.linalg import ptr
from oracle import oracle as O
from test_gpu_th import perturbed

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def empties(g):
    return (torch.full((g.ndof,), float("nan"), dtype=torch.float64, device=DEV),
            torch.full((g.num_elems,), float("nan"), dtype=torch.float64, device=DEV))


def two_phase(g, s, nu_e, lam, mask):
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    p = ns.THPlan(amap, DEV)
    jt, ge = empties(g)
    _lib.call("fx_th_vjp", g.ndof, g.num_elems, ptr(g.xv), ptr(g.yv), ptr(g.vconn3),
              ptr(g.dofs15), ptr(s), 0.0, ptr(nu_e), ptr(lam), ptr(mask), ptr(p.inc_off),
              ptr(p.elem_inc), ptr(jt), ptr(ge), ptr(p.work), None)
    return host(jt), host(ge)


def patch_vjp(g, s, nu_e, lam, mask, pp):
    jt, ge = empties(g)
    _lib.call("fx_th_vjp_patch", g.ndof, g.num_elems, ptr(g.xv), ptr(g.yv), ptr(s), 0.0, ptr(nu_e), ptr(lam), ptr(mask), C.byref(pp.plan),
              ptr(jt), ptr(ge), None)
    return host(jt), host(ge)


def check(g, mask_dofs, patch_sizes, seed):
    rng = np.random.default_rng(seed)
    s = rng.standard_normal(g.ndof)
    lam = rng.standard_normal(g.ndof)
    nu_e = rng.uniform(0.005, 0.05, g.num_elems)
    hmask = None
    if mask_dofs is not None:
        hmask = np.zeros(g.ndof, np.int32)
        hmask[mask_dofs] = 1
    d_s, d_lam, d_nu = dev(s), dev(lam), dev(nu_e)
    d_mask = dev(hmask) if hmask is not None else None
    ojt, og = O.th_vjp(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1), g.ndof,
                       0.0, s, lam, hmask, nu_e=nu_e)
    tjt, tg = two_phase(g, d_s, d_nu, d_lam, d_mask)
    assert close(tjt, ojt) and close(tg, og)
    for mp in patch_sizes:
        pp = ns.THPatches(g, DEV, mp)
        jt, ge = patch_vjp(g, d_s, d_nu, d_lam, d_mask, pp)
        assert close(jt, ojt) and close(jt, tjt), f"max_patch {mp}"
        assert np.array_equal(ge, tg), f"max_patch {mp}"
        jt2, ge2 = patch_vjp(g, d_s, d_nu, d_lam, d_mask, pp)
        assert np.array_equal(jt2, jt) and np.array_equal(ge2, ge)


def test_patch_vjp_structured_with_dirichlet_rows():
    g = ns.structured_th(6, DEV)
    check(g, ns.cavity_bc(g)[0], (1, 7, 64, 256), 1)


def test_patch_vjp_skewed_and_clockwise_elements():
    check(perturbed(5, 4, True), None, (13, 256), 2)


def test_patch_vjp_scrambled_order_and_unused_vertex():
    rng = np.random.default_rng(3)
    x, y, c3 = O.mesh_p1(8)
    c3 = c3.reshape(-1, 3)[rng.permutation(128)]
    g = ns.th_from_p1(np.append(x, 2.0), np.append(y, 2.0), c3, DEV)
    check(g, np.array([0, 5, g.ndof - 1]), (32, 256), 3)


def test_patch_vjp_full_size_cavity():
    """512^2 cavity (524,288 elements, Re = 100, lid rows constrained):
    against the oracle, and two launches bitwise equal."""
    g = ns.structured_th(512, DEV)
    s = ns.fourier_state(g, 5)
    lam = np.random.default_rng(7).standard_normal(g.ndof)
    nu_e = np.full(g.num_elems, 0.01)
    hmask = np.zeros(g.ndof, np.int32)
    hmask[ns.cavity_bc(g)[0]] = 1
    pp = ns.THPatches(g, DEV)
    args = (dev(s), dev(nu_e), dev(lam), dev(hmask))
    jt, ge = patch_vjp(g, *args, pp)
    jt2, ge2 = patch_vjp(g, *args, pp)
    assert np.array_equal(jt, jt2) and np.array_equal(ge, ge2)
    ojt, og = O.th_vjp(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1), g.ndof,
                       0.0, s, lam, hmask, nu_e=nu_e)
    assert close(jt, ojt) and close(ge, og)


def test_patch_vjp_fan_hub_cut_into_patches():
    """24 elements around one hub vertex (runs capped at 16 per patch): the
    hub's DOFs are seam DOFs over two or more patches, two of them
    constrained.  (40 elements would give the hub a CSR row of more than 255
    entries, beyond the two-phase VJP's row tables; the patch VJP has no row
    limit.)"""
    m = 24
    ang = 2 * np.pi * np.arange(m) / m
    x = np.concatenate([[0.0], np.cos(ang)])
    y = np.concatenate([[0.0], np.sin(ang)])
    c3 = np.array([[0, 1 + i, 1 + (i + 1) % m] for i in range(m)], np.int32)
    g = ns.th_from_p1(x, y, c3, DEV)
    hub = g.h_dofs15[0, [0, 12]]  # velocity x and pressure of vertex 0
    check(g, hub, (256, 8), 4)
