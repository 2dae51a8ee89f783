"""The CPU oracle against the compiled femx reference's golden fixtures and
the reference tests' known-answer arrays.  Integer outputs are bit-exact;
FP64 outputs are bitwise equal (same operation order, no contraction)."""
import numpy as np
import pytest

from conftest import golden
from oracle import oracle as O


@pytest.mark.parametrize("n", [1, 2, 4, 8])
def test_map_p1_matches_reference(n):
    g = golden(f"map_p1_n{n}.npz")
    x, y, conn = O.mesh_p1(n)
    assert np.array_equal(conn, g["conn"])
    N, E = x.size, conn.size // 3
    off = O.uniform_offsets(E, 3)
    rp, ci, jo, jm = O.map_build(N, N, off, conn, off, conn)
    for k, v in dict(row_ptr=rp, col_ind=ci, jac_offsets=jo, jac_map=jm).items():
        assert np.array_equal(v, g[k]), k
    assert ci.size == 7 * n * n + 6 * n + 1


@pytest.mark.parametrize("n", [1, 2])
def test_map_th_matches_reference(n):
    g = golden(f"map_th_n{n}.npz")
    xv, yv, c6, v3 = O.mesh_p2(n)
    d = O.th_dofs(n, c6, v3)
    assert np.array_equal(d, g["dofs15"])
    ndof = 2 * (2 * n + 1) ** 2 + (n + 1) ** 2
    off = O.uniform_offsets(2 * n * n, 15)
    rp, ci, jo, jm = O.map_build(ndof, ndof, off, d, off, d)
    for k, v in dict(row_ptr=rp, col_ind=ci, jac_offsets=jo, jac_map=jm).items():
        assert np.array_equal(v, g[k]), k
    assert ci.size == 267 * n * n + 110 * n + 9


def test_map_kats():
    # tests/unit/backend/runBackendAssemblyTests.cpp:162-188 (rectangular map)
    rp, ci, jo, jm = O.map_build(2, 2, [0, 2, 3], [0, 1, 1], [0, 1, 3], [0, 0, 1])
    assert rp.tolist() == [0, 1, 3] and ci.tolist() == [0, 0, 1]
    assert jo.tolist() == [0, 2, 4] and jm.tolist() == [0, 1, 1, 2]
    # tests/unit/linalg/runLinalgTests.cpp:235-270 (shared-element map)
    rp, ci, jo, jm = O.map_build(3, 3, [0, 2, 4], [0, 1, 1, 2], [0, 2, 4], [0, 1, 1, 2])
    assert rp.tolist() == [0, 2, 5, 7] and ci.tolist() == [0, 1, 0, 1, 2, 1, 2]
    assert jo.tolist() == [0, 4, 8] and jm.tolist() == [0, 1, 2, 3, 3, 4, 5, 6]


def test_bcmap_matches_reference():
    g = golden("bc_p1_n4.npz")
    bc = O.bc_build(g["row_ptr"], g["col_ind"], g["dofs"])
    for k in ("bc_rows", "diag", "col_offsets", "col_entries", "col_rows", "bc_mask"):
        assert np.array_equal(bc[k], g[k]), k


def test_bc_ops_match_reference():
    g = golden("bcops_p1_n4.npz")
    rp, ci = g["row_ptr"], g["col_ind"]
    bc = O.bc_build(rp, ci, g["dofs"])
    v = g["vals0"].copy()
    O.replace_rows(rp, bc, v, 1.0)
    assert np.array_equal(v, g["rows1_vals"])
    v = g["vals0"].copy()
    O.replace_rows(rp, bc, v, 0.0)
    assert np.array_equal(v, g["rows0_vals"])
    v, r = g["vals0"].copy(), g["rhs0"].copy()
    O.prepare_forward(rp, bc, v, r, g["bc_vals"])
    assert np.array_equal(v, g["forward_vals"]) and np.array_equal(r, g["forward_rhs"])
    r = g["rhs0"].copy()
    O.replace_res(bc, g["state"], g["bc_vals"], r)
    assert np.array_equal(r, g["res_rhs"])
    r = g["rhs0"].copy()
    O.zero_boundary(bc, r)
    assert np.array_equal(r, g["zero_rhs"])


def test_bc_kat():
    # tests/unit/backend/runBackendStateTests.cpp:80-114
    rp = np.array([0, 3, 6, 9], np.int32)
    ci = np.array([0, 1, 2] * 3, np.int32)
    bc = O.bc_build(rp, ci, [1])
    base = np.array([4.0, 1.0, 2.0, 3.0, 5.0, 6.0, 7.0, 8.0, 9.0])
    v = base.copy()
    O.replace_rows(rp, bc, v, 1.0)
    assert v.tolist() == [4.0, 1.0, 2.0, 0.0, 1.0, 0.0, 7.0, 8.0, 9.0]
    v = base.copy()
    O.replace_rows(rp, bc, v, 0.0)
    assert v.tolist() == [4.0, 1.0, 2.0, 0.0, 0.0, 0.0, 7.0, 8.0, 9.0]
    r = np.array([10.0, 20.0, 30.0])
    O.replace_res(bc, np.array([4.0, 7.0, 9.0]), np.array([2.5]), r)
    assert r.tolist() == [10.0, 4.5, 30.0]
    v, r = base.copy(), np.array([10.0, 20.0, 30.0])
    O.prepare_forward(rp, bc, v, r, np.array([2.0]))
    assert v.tolist() == [4.0, 0.0, 2.0, 0.0, 1.0, 0.0, 7.0, 0.0, 9.0]
    assert r.tolist() == [8.0, 2.0, 14.0]


@pytest.mark.parametrize("n", [2, 8])
def test_poisson_matches_reference(n):
    g = golden(f"poisson_p1_n{n}.npz")
    x, y, conn = g["x"], g["y"], g["conn"]
    N, E = x.size, conn.size // 3
    off = O.uniform_offsets(E, 3)
    rp, ci, jo, jm = O.map_build(N, N, off, conn, off, conn)
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(ci, g["col_ind"])
    vals, res = O.p1_assemble(x, y, conn, np.ones(E), np.zeros(N), g["state"], jm, ci.size)
    assert np.array_equal(vals, g["vals"])
    assert np.array_equal(res, g["res"])


def test_spmv_matches_reference():
    g = golden("spmv_p1_n4.npz")
    rp, ci, v = g["row_ptr"], g["col_ind"], g["vals"]
    N = rp.size - 1
    for tr in (0, 1):
        for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
            yy = g["y0"].copy()
            if tr:
                O.matvecT(rp, ci, v, N, g["x"], yy, alpha, beta)
            else:
                O.matvec(rp, ci, v, g["x"], yy, alpha, beta)
            assert np.array_equal(yy, g[f"y_t{tr}_a{alpha}_b{beta}"])


def test_spmv_kat():
    # tests/unit/backend/runCudaLinalgTests.cu:70-80,176-181,224-230
    rp = np.array([0, 2, 4, 7], np.int32)
    ci = np.array([0, 2, 1, 3, 0, 2, 3], np.int32)
    v = np.array([2.0, -1.0, 3.0, 4.0, -2.0, 5.0, 1.0])
    y = np.zeros(3)
    O.matvec(rp, ci, v, np.array([1.0, 2.0, 3.0, 4.0]), y)
    assert y.tolist() == [-1.0, 22.0, 17.0]
    v2 = np.array([-1.0, 2.0, 0.5, -3.0, 4.0, 1.0, -2.0])
    O.matvec(rp, ci, v2, np.array([1.0, 2.0, 3.0, 4.0]), y)
    assert y.tolist() == [5.0, -11.0, -1.0]
