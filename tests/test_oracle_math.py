"""Pins for the oracle's extensions that femx does not have (kappa, load,
P1 log-kappa VJP, 7-point rule, P2 element, Taylor-Hood Navier-Stokes).
Each check is an independent computation -- closed forms from textbook
formulas, monomial integrals, central finite differences, a manufactured
solution -- chosen so that a dropped term, a wrong sign or a transposed
index fails."""
from math import factorial

import numpy as np
import pytest

from conftest import close
from oracle import oracle as O

RNG = np.random.default_rng(7)


@pytest.mark.parametrize("order,degree", [(1, 1), (2, 2), (5, 5)])
def test_quadrature_exact_on_monomials(order, degree):
    rs, w = O.quad_tri(order)
    assert abs(w.sum() - 0.5) < 1e-15
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            got = np.sum(w * rs[:, 0] ** a * rs[:, 1] ** b)
            assert abs(got - exact) < 1e-15, (a, b)
    # degree + 1 is not integrated exactly (guards against a wrong table)
    errs = [abs(np.sum(w * rs[:, 0] ** a * rs[:, 1] ** (degree + 1 - a))
                - factorial(a) * factorial(degree + 1 - a) / factorial(degree + 3))
            for a in range(degree + 2)]
    assert max(errs) > 1e-6


def test_p2_shapes():
    nodes = [(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)]
    for i, (r, s) in enumerate(nodes):
        N, _ = O.p2_shape(r, s)
        assert np.allclose(N, np.eye(6)[i], atol=1e-15)
    for _ in range(20):
        r, s = RNG.uniform(0, 0.5, 2)
        N, d = O.p2_shape(r, s)
        assert abs(N.sum() - 1) < 1e-14
        assert np.abs(d.sum(axis=0)).max() < 1e-13
        h = 1e-6
        fd_r = (O.p2_shape(r + h, s)[0] - O.p2_shape(r - h, s)[0]) / (2 * h)
        fd_s = (O.p2_shape(r, s + h)[0] - O.p2_shape(r, s - h)[0]) / (2 * h)
        assert np.abs(fd_r - d[:, 0]).max() < 1e-8 and np.abs(fd_s - d[:, 1]).max() < 1e-8


def _closed_form_p1(xe, ye, kappa, fe, ue):
    """Textbook P1 element: grad phi_i = (y_j - y_k, x_k - x_j) / (2 A_signed)
    with (i, j, k) cyclic; K = kappa |A| G G^T; F = |A|/12 [2 1 1;1 2 1;1 1 2] f."""
    A2 = (xe[1] - xe[0]) * (ye[2] - ye[0]) - (xe[2] - xe[0]) * (ye[1] - ye[0])
    G = np.array([[ye[(i + 1) % 3] - ye[(i + 2) % 3], xe[(i + 2) % 3] - xe[(i + 1) % 3]]
                  for i in range(3)]) / A2
    A = abs(A2) / 2
    K = kappa * A * G @ G.T
    F = A / 12 * np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]]) @ fe
    return K, K @ ue - F


def test_p1_element_closed_form():
    for orient in (1, -1):
        for _ in range(25):
            xe, ye = RNG.uniform(-1, 1, 3), RNG.uniform(-1, 1, 3)
            if orient < 0:
                xe, ye = xe[::-1].copy(), ye[::-1].copy()
            kappa = np.exp(RNG.standard_normal())
            fe, ue = RNG.standard_normal(3), RNG.standard_normal(3)
            K, R = O.p1_element(xe, ye, kappa, fe, ue)
            Kc, Rc = _closed_form_p1(xe, ye, kappa, fe, ue)
            assert close(K, Kc, 1e-12) and close(R, Rc, 1e-12)


def _p1_problem(n, seed=0):
    x, y, conn = O.mesh_p1(n)
    N, E = x.size, conn.size // 3
    off = O.uniform_offsets(E, 3)
    rp, ci, jo, jm = O.map_build(N, N, off, conn, off, conn)
    tol = 1e-12  # x_min + i*dx can miss 1.0 by an ulp
    on = (abs(x) <= tol) | (abs(x - 1) <= tol) | (abs(y) <= tol) | (abs(y - 1) <= tol)
    dofs = np.nonzero(on)[0].astype(np.int32)
    return x, y, conn, rp, ci, jm, dofs


def _dense(rp, ci, v, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    return A


def test_p1_global_invariants():
    x, y, conn, rp, ci, jm, _ = _p1_problem(6)
    N, E = x.size, conn.size // 3
    kappa = np.exp(0.5 * RNG.standard_normal(E))
    vals, _ = O.p1_assemble(x, y, conn, kappa, np.zeros(N), np.zeros(N), jm, ci.size)
    K = _dense(rp, ci, vals, N)
    assert np.abs(K.sum(axis=1)).max() < 1e-13 * np.abs(K).max()
    assert np.abs(K - K.T).max() < 1e-13 * np.abs(K).max()
    # load of f = 1 integrates to the domain area
    _, res = O.p1_assemble(x, y, conn, kappa, np.ones(N), np.zeros(N), jm, ci.size)
    assert abs(-res.sum() - 1.0) < 1e-13


def test_p1_manufactured_solution_converges():
    """Config-1 problem: -div grad u = 2 pi^2 sin(pi x) sin(pi y), u = 0 on the
    boundary; the nodal error must fall like h^2."""
    errs = []
    for n in (8, 16, 32):
        x, y, conn, rp, ci, jm, dofs = _p1_problem(n)
        N, E = x.size, conn.size // 3
        f = 2 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y)
        vals, res = O.p1_assemble(x, y, conn, np.ones(E), f, np.zeros(N), jm, ci.size)
        rhs = -res
        bc = O.bc_build(rp, ci, dofs)
        O.prepare_forward(rp, bc, vals, rhs, np.zeros(dofs.size))
        u = np.linalg.solve(_dense(rp, ci, vals, N), rhs)
        errs.append(np.abs(u - np.sin(np.pi * x) * np.sin(np.pi * y)).max())
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5, errs


def _p1_residual(x, y, conn, jm, nnz, bc, m, u, f, g):
    N = x.size
    _, res = O.p1_assemble(x, y, conn, np.exp(m), f, u, jm, nnz)
    O.replace_res(bc, u, g, res)
    return res


def test_p1_vjp_finite_differences_and_adjoint_identity():
    x, y, conn, rp, ci, jm, dofs = _p1_problem(5)
    N, E = x.size, conn.size // 3
    bc = O.bc_build(rp, ci, dofs)
    m = 0.5 * RNG.standard_normal(E)
    u = RNG.standard_normal(N)
    f = RNG.standard_normal(N)
    g = RNG.standard_normal(dofs.size)
    lam = RNG.standard_normal(N)
    jt, ge = O.p1_vjp(x, y, conn, m, u, lam, bc["bc_mask"])
    h = 1e-6
    for k in RNG.choice(E, 8, replace=False):
        mp, mm = m.copy(), m.copy()
        mp[k] += h
        mm[k] -= h
        fd = lam @ (_p1_residual(x, y, conn, jm, ci.size, bc, mp, u, f, g)
                    - _p1_residual(x, y, conn, jm, ci.size, bc, mm, u, f, g)) / (2 * h)
        assert abs(ge[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, ge[k], fd)
    for k in list(RNG.choice(N, 6, replace=False)) + list(dofs[:2]):
        up, um = u.copy(), u.copy()
        up[k] += h
        um[k] -= h
        fd = lam @ (_p1_residual(x, y, conn, jm, ci.size, bc, m, up, f, g)
                    - _p1_residual(x, y, conn, jm, ci.size, bc, m, um, f, g)) / (2 * h)
        assert abs(jt[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, jt[k], fd)
    # J^T lam against the assembled row-replaced Jacobian's transpose apply
    vals, _ = O.p1_assemble(x, y, conn, np.exp(m), f, u, jm, ci.size)
    O.replace_rows(rp, bc, vals, 1.0)
    ref = np.zeros(N)
    O.matvecT(rp, ci, vals, N, lam, ref)
    assert close(jt, ref, 1e-13)


def _tri(rng):
    while True:
        xe, ye = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
        if abs((xe[1] - xe[0]) * (ye[2] - ye[0]) - (xe[2] - xe[0]) * (ye[1] - ye[0])) > 0.1:
            return xe, ye


def _p2_nodes(xe, ye):
    px = np.r_[xe, (xe[0] + xe[1]) / 2, (xe[1] + xe[2]) / 2, (xe[2] + xe[0]) / 2]
    py = np.r_[ye, (ye[0] + ye[1]) / 2, (ye[1] + ye[2]) / 2, (ye[2] + ye[0]) / 2]
    return px, py


def test_th_jacobian_matches_finite_differences():
    for _ in range(5):
        xe, ye = _tri(RNG)
        nu = RNG.uniform(0.002, 0.1)
        s = RNG.standard_normal(15)
        K, _ = O.th_element(xe, ye, nu, s)
        h = 1e-6
        for j in range(15):
            sp, sm = s.copy(), s.copy()
            sp[j] += h
            sm[j] -= h
            col = (O.th_element(xe, ye, nu, sp)[1] - O.th_element(xe, ye, nu, sm)[1]) / (2 * h)
            assert np.abs(col - K[:, j]).max() <= 1e-7 * max(1.0, np.abs(K[:, j]).max()), j


def test_th_stokes_blocks():
    xe, ye = _tri(RNG)
    K, R = O.th_element(xe, ye, 0.05, np.zeros(15))
    assert np.abs(R).max() == 0.0
    Kuu, Kup, Kpu, Kpp = K[:12, :12], K[:12, 12:], K[12:, :12], K[12:, 12:]
    assert np.abs(Kuu - Kuu.T).max() < 1e-13 * np.abs(Kuu).max()
    assert np.abs(Kup - Kpu.T).max() < 1e-14 * np.abs(Kup).max()
    assert np.abs(Kpp).max() == 0.0
    # viscous block annihilates constant velocities (sum of P2 gradients = 0)
    assert np.abs(Kuu @ np.tile([1.0, 0.0], 6)).max() < 1e-12 * np.abs(Kuu).max()


def test_th_residual_closed_forms():
    xe, ye = _tri(RNG)
    A = abs((xe[1] - xe[0]) * (ye[2] - ye[0]) - (xe[2] - xe[0]) * (ye[1] - ye[0])) / 2
    xc, yc = xe.mean(), ye.mean()
    px, py = _p2_nodes(xe, ye)
    # constant velocity, zero pressure: R = 0
    s = np.zeros(15)
    s[0:12:2] = 1.0
    _, R = O.th_element(xe, ye, 0.3, s)
    assert np.abs(R).max() < 1e-14
    # u = (x, -y), p = 0: div u = 0, sum_a R_(a,c) = int u.grad u_c = (A xc, A yc)
    s = np.zeros(15)
    s[0:12:2], s[1:12:2] = px, -py
    _, R = O.th_element(xe, ye, 0.3, s)
    assert np.abs(R[12:]).max() < 1e-14
    assert abs(R[0:12:2].sum() - A * xc) < 1e-13 and abs(R[1:12:2].sum() - A * yc) < 1e-13
    # u = 0, p = x: sum_a v_a R_(a,0) for v = x equals -int p dv/dx = -A xc
    s = np.zeros(15)
    s[12:] = xe
    _, R = O.th_element(xe, ye, 0.3, s)
    assert abs(px @ R[0:12:2] + A * xc) < 1e-13
    assert abs(R[0:12:2].sum()) < 1e-14
    # u = (y, 0): div u = 0; u = (x, 0): R_p = -int psi_b div u = -A/3 each
    s = np.zeros(15)
    s[0:12:2] = px
    _, R = O.th_element(xe, ye, 0.3, s)
    assert np.allclose(R[12:], -A / 3, atol=1e-14)
    # viscous scale: at u = 0 the Jacobian's velocity block is nu * stiffness,
    # and v^T K w for the P2-exact v = x, w = x^2 is nu int grad x . grad x^2
    # = 2 nu A xc (a dropped or doubled nu fails this; FD and symmetry do not)
    nu = 0.3
    K, _ = O.th_element(xe, ye, nu, np.zeros(15))
    assert abs(px @ K[0:12:2, 0:12:2] @ px ** 2 - 2 * nu * A * xc) < 1e-13
    assert abs(py @ K[1:12:2, 1:12:2] @ py ** 2 - 2 * nu * A * yc) < 1e-13
    assert np.abs(K[0:12:2, 1:12:2]).max() == 0.0


def test_th_vjp_oracle_finite_differences():
    n = 2
    xv, yv, c6, v3 = O.mesh_p2(n)
    d = O.th_dofs(n, c6, v3)
    x1, y1, _ = O.mesh_p1(n)
    ndof = 2 * xv.size + x1.size
    off = O.uniform_offsets(2 * n * n, 15)
    rp, ci, jo, jm = O.map_build(ndof, ndof, off, d, off, d)
    s = RNG.standard_normal(ndof)
    lam = RNG.standard_normal(ndof)
    nu = 0.02
    mask = np.zeros(ndof, np.int32)
    mask[:10] = 1
    jt, g = O.th_vjp(x1, y1, v3, d, ndof, nu, s, lam, mask)
    dnu = g.sum()
    vals, _ = O.th_assemble(x1, y1, v3, d, ndof, nu, s, jm, ci.size)
    bc = O.bc_build(rp, ci, np.nonzero(mask)[0].astype(np.int32))
    O.replace_rows(rp, bc, vals, 1.0)
    ref = np.zeros(ndof)
    O.matvecT(rp, ci, vals, ndof, lam, ref)
    assert close(jt, ref, 1e-13)
    lt = np.where(mask == 1, 0.0, lam)
    h = 1e-6
    rpl = O.th_assemble(x1, y1, v3, d, ndof, nu + h, s, jm, ci.size)[1]
    rmi = O.th_assemble(x1, y1, v3, d, ndof, nu - h, s, jm, ci.size)[1]
    fd = lt @ (rpl - rmi) / (2 * h)
    assert abs(dnu - fd) <= 1e-7 * max(1.0, abs(fd))
