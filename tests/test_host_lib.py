"""libfemx_amd.so on the CPU: it loads, exports every C-ABI symbol the
header declares, and its host builders match the compiled reference's
fixtures and the oracle bit for bit, with femx's validation errors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden
from femx_amd import _lib
from femx_amd.assembly import make_assembly_map
from femx_amd.linalg import CsrPattern, ptr
from oracle import oracle as O

RNG = np.random.default_rng(11)


def declared_symbols():
    with open(os.path.join(ROOT, "include", "femx_amd.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) <= set(_lib.exported_symbols())
    assert lib.fx_version() >= 1


@pytest.mark.parametrize("n", [1, 2, 4, 8])
def test_map_builder_matches_reference_p1(n):
    g = golden(f"map_p1_n{n}.npz")
    E = 2 * n * n
    amap = make_assembly_map((n + 1) ** 2, (n + 1) ** 2, g["conn"].reshape(E, 3))
    assert np.array_equal(amap.pattern.row_ptr, g["row_ptr"])
    assert np.array_equal(amap.pattern.col_ind, g["col_ind"])
    jo, jm = amap.jac()
    assert np.array_equal(jo, g["jac_offsets"]) and np.array_equal(jm, g["jac_map"])
    assert amap.max_row_len == (7 if n > 1 else 4)


@pytest.mark.parametrize("n", [1, 2])
def test_map_builder_matches_reference_th(n):
    g = golden(f"map_th_n{n}.npz")
    ndof = 2 * (2 * n + 1) ** 2 + (n + 1) ** 2
    amap = make_assembly_map(ndof, ndof, g["dofs15"].reshape(-1, 15))
    assert np.array_equal(amap.pattern.row_ptr, g["row_ptr"])
    assert np.array_equal(amap.pattern.col_ind, g["col_ind"])
    jo, jm = amap.jac()
    assert np.array_equal(jo, g["jac_offsets"]) and np.array_equal(jm, g["jac_map"])


def test_map_builder_matches_oracle_on_irregular_tables():
    # ragged elements, repeated DOFs inside an element, empty elements and
    # rows with no entries -- femx's makeAssemblyMap accepts all of them
    for trial in range(5):
        nr, ns, E = 40, 35, 60
        res = [RNG.integers(0, nr, RNG.integers(0, 5)).tolist() for _ in range(E)]
        st = [RNG.integers(0, ns, RNG.integers(0, 5)).tolist() for _ in range(E)]
        amap = make_assembly_map(nr, ns, res, st)
        ro = np.r_[0, np.cumsum([len(r) for r in res])].astype(np.int32)
        so = np.r_[0, np.cumsum([len(s) for s in st])].astype(np.int32)
        rp, ci, jo, jm = O.map_build(nr, ns, ro, np.concatenate(res + [[]]).astype(np.int32),
                                     so, np.concatenate(st + [[]]).astype(np.int32))
        assert np.array_equal(amap.pattern.row_ptr, rp) and np.array_equal(amap.pattern.col_ind, ci)
        ajo, ajm = amap.jac()
        assert np.array_equal(ajo, jo) and np.array_equal(ajm, jm)


def test_map_kats():
    # tests/unit/backend/runBackendAssemblyTests.cpp:162-188
    amap = make_assembly_map(2, 2, [[0, 1], [1]], [[0], [0, 1]])
    assert amap.pattern.row_ptr.tolist() == [0, 1, 3]
    assert amap.pattern.col_ind.tolist() == [0, 0, 1]
    jo, jm = amap.jac()
    assert jo.tolist() == [0, 2, 4] and jm.tolist() == [0, 1, 1, 2]
    # tests/unit/linalg/runLinalgTests.cpp:235-270
    amap = make_assembly_map(3, 3, [[0, 1], [1, 2]])
    assert amap.pattern.row_ptr.tolist() == [0, 2, 5, 7]
    assert amap.pattern.col_ind.tolist() == [0, 1, 0, 1, 2, 1, 2]
    assert amap.jac()[1].tolist() == [0, 1, 2, 3, 3, 4, 5, 6]


def test_map_errors_follow_reference():
    with pytest.raises(RuntimeError, match="residual DOF is out of range"):
        make_assembly_map(2, 2, [[0, 2]], [[0, 1]])
    with pytest.raises(RuntimeError, match="state DOF is out of range"):
        make_assembly_map(2, 2, [[0, 1]], [[0, -1]])
    with pytest.raises(RuntimeError, match="element counts differ"):
        make_assembly_map(2, 2, [[0, 1]], [[0], [1]])
    with pytest.raises(RuntimeError, match="non-negative"):
        make_assembly_map(-1, 2, [[0]], [[0]])


def test_rowinc3_positions_reproduce_jac_map():
    n = 5
    x, y, conn = O.mesh_p1(n)
    E = 2 * n * n
    amap = make_assembly_map(x.size, x.size, conn.reshape(E, 3))
    inc_off, inc = amap.rowinc3()
    jo, jm = amap.jac()
    rp = amap.pattern.row_ptr
    seen = 0
    for r in range(x.size):
        prev = -1
        for t in range(inc_off[r], inc_off[r + 1]):
            e, meta = int(inc[2 * t]), int(inc[2 * t + 1]) & 0xFFFFFFFF
            li = meta & 0xFF
            assert e > prev  # ascending element order
            prev = e
            assert conn[3 * e + li] == r
            for j in range(3):
                assert rp[r] + ((meta >> (8 * (j + 1))) & 0xFF) == jm[9 * e + 3 * li + j]
            seen += 1
    assert seen == 3 * E


def test_rowinc3_rejects_other_shapes():
    amap = make_assembly_map(4, 4, [[0, 1, 2, 3]])
    with pytest.raises(RuntimeError, match="3x3"):
        amap.rowinc3()


def test_boundary_map_matches_reference():
    from femx_amd.assembly import BoundaryMap
    g = golden("bc_p1_n4.npz")
    pat = CsrPattern(g["row_ptr"].size - 1, g["row_ptr"].size - 1, g["row_ptr"], g["col_ind"])
    b = BoundaryMap(g["dofs"], pat, device="cpu")
    for k in ("bc_rows", "diag", "col_offsets", "col_entries", "col_rows", "bc_mask"):
        assert np.array_equal(getattr(b, k), g[k]), k
    # elimination lists: exactly the unconstrained-row entries of constrained
    # columns, grouped by row, ascending boundary index within a row
    expect = {}
    for ib in range(b.num_bcs):
        for i in range(b.col_offsets[ib], b.col_offsets[ib + 1]):
            r = b.col_rows[i]
            if not b.bc_mask[r]:
                expect.setdefault(r, []).append((b.col_entries[i], ib))
    assert b.elim_rows.tolist() == sorted(expect)
    for t, r in enumerate(b.elim_rows):
        got = list(zip(b.elim_k[b.elim_off[t]:b.elim_off[t + 1]],
                       b.elim_ib[b.elim_off[t]:b.elim_off[t + 1]]))
        assert got == expect[r]


def test_boundary_map_errors_follow_reference():
    from femx_amd.assembly import BoundaryMap
    rp = np.array([0, 2, 3, 5], np.int32)
    ci = np.array([0, 1, 0, 1, 2], np.int32)  # row 1 lacks its diagonal
    pat = CsrPattern(3, 3, rp, ci)
    with pytest.raises(RuntimeError, match="no diagonal entry"):
        BoundaryMap([1], pat, device="cpu")
    with pytest.raises(RuntimeError, match="must be unique"):
        BoundaryMap([0, 0], pat, device="cpu")
    with pytest.raises(RuntimeError, match="out of range"):
        BoundaryMap([3], pat, device="cpu")


def test_transpose_map():
    # tests/unit/backend/runCudaLinalgTests.cu:70-110 pattern and transpose
    pat = CsrPattern(3, 4, [0, 2, 4, 7], [0, 2, 1, 3, 0, 2, 3])
    t_ptr = np.empty(5, np.int32)
    t_row = np.empty(7, np.int32)
    t_perm = np.empty(7, np.int32)
    _lib.call("fx_csr_transpose_map", 3, 4, ptr(pat.row_ptr), ptr(pat.col_ind), ptr(t_ptr),
              ptr(t_row), ptr(t_perm))
    assert t_ptr.tolist() == [0, 2, 3, 5, 7]
    assert t_row.tolist() == [0, 2, 1, 0, 2, 1, 2]
    vals = np.array([2.0, -1.0, 3.0, 4.0, -2.0, 5.0, 1.0])
    assert vals[t_perm].tolist() == [2.0, -2.0, 3.0, -1.0, 5.0, 4.0, 1.0]


def test_col_off16():
    pat = CsrPattern(3, 4, [0, 2, 4, 7], [0, 2, 1, 3, 0, 2, 3])
    off = np.empty(7, np.int16)
    _lib.call("fx_csr_col_off16", 3, ptr(pat.row_ptr), ptr(pat.col_ind), ptr(off))
    assert off.tolist() == [0, 2, 0, 2, -2, 0, 1]
    assert pat.col_off16("cpu").tolist() == off.tolist()
    # an offset of 32768 does not fit: the builder refuses, the pattern keeps int32 only
    wide = CsrPattern(2, 40000, [0, 1, 2], [32768, 1])
    with pytest.raises(RuntimeError, match="16 bits"):
        _lib.call("fx_csr_col_off16", 2, ptr(wide.row_ptr), ptr(wide.col_ind), ptr(off))
    assert wide.col_off16("cpu") is None
    edge = CsrPattern(2, 40000, [0, 1, 2], [32767, 1 - 32768 + 32768])
    assert edge.col_off16("cpu").tolist() == [32767, 0]


def test_stencil_plan_codes_reproduce_columns():
    """fx_csr_stencil_plan on the P1 mesh (300 x 300 cells, 90601 rows): a
    block gets a code exactly when its 64 rows share one offset stencil, and
    the code's offsets rebuild every column; blocks inside a grid row share
    the one 7-point stencil; a random pattern has no stencil blocks."""
    n = 300
    x = np.empty((n + 1) ** 2)
    y = np.empty_like(x)
    conn = np.empty(6 * n * n, np.int32)
    _lib.call("fx_mesh_p1", n, ptr(x), ptr(y), ptr(conn))
    pat = make_assembly_map(x.size, x.size, conn.reshape(-1, 3)).pattern
    nb, no = C.c_int32(), C.c_int32()
    args = (pat.rows, pat.cols, ptr(pat.row_ptr), ptr(pat.col_ind), 8)
    _lib.call("fx_csr_stencil_plan", *args, C.byref(nb), C.byref(no), None, None)
    code = np.empty(nb.value, np.int32)
    off = np.empty(max(no.value, 1), np.int32)
    _lib.call("fx_csr_stencil_plan", *args, C.byref(nb), C.byref(no), ptr(code), ptr(off))
    assert nb.value == -(-pat.rows // 64)
    rp, ci = pat.row_ptr, pat.col_ind
    for b in range(nb.value):
        r0 = 64 * b
        rows = range(r0, min(r0 + 64, pat.rows))
        stencils = {tuple(ci[rp[r]:rp[r + 1]] - r) for r in rows}
        uniform = len(rows) == 64 and len(stencils) == 1 and 1 <= len(next(iter(stencils))) <= 8
        assert (code[b] >= 0) == uniform, b
        if code[b] >= 0:
            S, start = code[b] & 31, code[b] >> 5
            assert tuple(off[start:start + S]) == next(iter(stencils))
    # the interior 7-point stencil holds most blocks; the bottom and top grid
    # rows have 5-point stencils of their own
    used = code[code >= 0]
    vals_, counts_ = np.unique(used, return_counts=True)
    top = vals_[np.argmax(counts_)]
    assert top & 31 == 7
    assert tuple(off[(top >> 5):(top >> 5) + 7]) == (-(n + 2), -(n + 1), -1, 0, 1, n + 1, n + 2)
    assert {c & 31 for c in vals_} == {5, 7}
    assert used.size >= nb.value // 2
    # distinct stencils are stored once
    starts = sorted({c >> 5 for c in code if c >= 0})
    assert no.value == sum(int(code[np.nonzero((code >> 5) == s_)[0][0]] & 31) for s_ in starts)
    rnd = _xs_rand_pattern()
    _lib.call("fx_csr_stencil_plan", rnd.rows, rnd.cols, ptr(rnd.row_ptr), ptr(rnd.col_ind), 8,
              C.byref(nb), C.byref(no), None, None)
    code = np.empty(nb.value, np.int32)
    _lib.call("fx_csr_stencil_plan", rnd.rows, rnd.cols, ptr(rnd.row_ptr), ptr(rnd.col_ind), 8,
              C.byref(nb), C.byref(no), ptr(code), ptr(off))
    assert np.all(code == -1) and no.value == 0


def _xs_rand_pattern():
    rng = np.random.default_rng(5)
    rows, cols = 3000, 3100
    counts = rng.integers(1, 9, rows)
    ci = np.concatenate([np.sort(rng.choice(np.arange(max(0, r - 600), min(cols, r + 600)), c,
                                            replace=False)) for r, c in enumerate(counts)])
    rp = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return CsrPattern(rows, cols, rp, ci.astype(np.int32))


@pytest.mark.parametrize("n", [1, 3, 16])
def test_meshes_match_oracle(n):
    N, E = (n + 1) ** 2, 2 * n * n
    x, y, conn = np.empty(N), np.empty(N), np.empty(3 * E, np.int32)
    _lib.call("fx_mesh_p1", n, ptr(x), ptr(y), ptr(conn))
    ox, oy, oc = O.mesh_p1(n)
    assert np.array_equal(x, ox) and np.array_equal(y, oy) and np.array_equal(conn, oc)
    Nv = (2 * n + 1) ** 2
    x, y = np.empty(Nv), np.empty(Nv)
    c6, v3 = np.empty(6 * E, np.int32), np.empty(3 * E, np.int32)
    _lib.call("fx_mesh_p2", n, ptr(x), ptr(y), ptr(c6), ptr(v3))
    ox, oy, oc6, ov3 = O.mesh_p2(n)
    assert np.array_equal(x, ox) and np.array_equal(c6, oc6) and np.array_equal(v3, ov3)
    d = np.empty(15 * E, np.int32)
    _lib.call("fx_th_dofs", n, ptr(c6), ptr(v3), ptr(d))
    assert np.array_equal(d, O.th_dofs(n, oc6, ov3))


def test_mesh_overflow_is_reported():
    with pytest.raises(RuntimeError, match="int32"):
        # sizes are checked before any output is written
        _lib.call("fx_mesh_p1", 40000, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8))


def test_gather_offsets_are_checked_before_launch():
    """The kernels gather through 32-bit byte offsets; arrays past 4 GiB are
    refused with FX_ERR_OVERFLOW before anything is launched (the pointers
    below are never dereferenced)."""
    p = C.c_void_p(8)
    big = 1 << 29  # 2^29 doubles = 4 GiB
    with pytest.raises(RuntimeError, match="beyond 4 GiB"):
        _lib.call("fx_p1_assemble", big, 8, 8, 7, p, p, p, p, 0, p, p, p, p, p, p, p, None)
    with pytest.raises(RuntimeError, match="beyond 4 GiB"):
        _lib.call("fx_p1_vjp", 8, big, p, p, p, p, p, p, None, p, p, p, p, None)
    with pytest.raises(RuntimeError, match="beyond 4 GiB"):
        _lib.call("fx_th_assemble", 8, (1 << 28) // 15 + 1, 1, 8, p, p, p, p, p, 0.01, None, p, p, p,
                  p, p, None, p, p, p, None)
    with pytest.raises(RuntimeError, match="beyond 4 GiB"):
        _lib.call("fx_th_vjp", big, 8, p, p, p, p, p, 0.01, None, p, None, p, p, p, None, p, None)
    with pytest.raises(RuntimeError, match="beyond 4 GiB"):
        _lib.call("fx_th_vjp", 8, (1 << 29) // 15 + 1, p, p, p, p, p, 0.01, None, p, None, p, p, p,
                  None, p, None)


@pytest.mark.parametrize("name", ["rect", "p1_n3"])
def test_transpose_pattern_matches_reference(name):
    """MatrixHandler::transpose (compiled femx): the transposed pattern from the
    cached transpose map, bit-exact, and values through its permutation."""
    g = golden("transpose.npz")
    nr, nc = (int(v) for v in g[f"{name}_shape"])
    pat = CsrPattern(nr, nc, g[f"{name}_row_ptr"], g[f"{name}_col_ind"])
    tp = pat.transposed()
    assert (tp.rows, tp.cols) == (nc, nr)
    assert np.array_equal(tp.row_ptr, g[f"{name}_t_row_ptr"])
    assert np.array_equal(tp.col_ind, g[f"{name}_t_col_ind"])
    _, _, perm = pat._host_transpose()
    assert np.array_equal(g[f"{name}_vals"][perm], g[f"{name}_t_vals"])
    assert pat.transposed() is tp  # cached


def test_noalloc_vjp_rejects_sizes_beyond_its_workspace():
    """fx_th_vjp_noalloc refuses more elements than its fixed workspace
    holds, before touching the device (the pointers are never used)."""
    p = C.c_void_p(8)
    with pytest.raises(RuntimeError, match="FX_TH_VJP_NOALLOC_MAX_ELEMS"):
        _lib.call("fx_th_vjp_noalloc", 8, 559241, p, p, p, p, p, 0.01, None, p, None, p, p, p,
                  None, None)
    with pytest.raises(RuntimeError, match="FX_TH_VJP_NOALLOC_MAX_ELEMS"):
        _lib.call("fx_th_vjp_noalloc", 8, -1, p, p, p, p, p, 0.01, None, p, None, p, p, p, None,
                  None)
