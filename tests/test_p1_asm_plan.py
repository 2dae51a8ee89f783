"""Host checks of the P1 assembly's owner-computes patch plan
(fx_p1_asm_create): every row owned by one patch that holds all of the
node's elements, or left to the row-per-lane kernel; and the reduction the
plan describes -- rank steps adding each element's rows into its patch's LDS
tile, written back along runs -- applied on the CPU to integer-valued
element matrices (exact in any order) against a plain scatter-add, with
every CSR value and residual entry receiving its elements in ascending
element order, as fx_p1_assemble adds them."""
import numpy as np
import pytest

from femx_amd.assembly import Geometry, make_assembly_map, structured_p1
from femx_amd.patches import P1AsmPlan
from oracle import oracle as O

NO_TILE = 0xFFFF


def _check(x, y, conn, max_patch=256, max_tile=P1AsmPlan.MAX_TILE, seed=0):
    conn = np.ascontiguousarray(conn, np.int32).reshape(-1, 3)
    g = Geometry(x, y, conn, "cpu")
    N, E = g.num_nodes, g.num_elems
    amap = make_assembly_map(N, N, conn)
    pat = amap.pattern
    rp, ci = pat.row_ptr.astype(np.int64), pat.col_ind
    ap = P1AsmPlan(conn, g.h_x, g.h_y, pat, "cpu", max_patch, max_tile)
    a = ap.h
    po, pe = a["patch_off"], a["patch_elem"][: ap.num_entries]
    lo, ld = a["ldof_off"], a["ldof_dof"][: ap.num_ldofs]
    words = a["elem_ldof"].reshape(-1, 4)[: ap.num_entries, :3].astype(np.int64)
    jloc, rank = words & 0xFFF, words >> 12
    pos = a["elem_pos"].reshape(-1, 4)[: ap.num_entries, :3].astype(np.int64)
    tiles = a["ldof_tile"][: ap.num_ldofs].astype(np.int64)
    left = a["rows_left"][: ap.num_left]
    node_elems = [np.flatnonzero((conn == n).any(axis=1)) for n in range(N)]
    rng = np.random.default_rng(seed)
    K = rng.integers(-1000, 1000, size=(E, 3, 3)).astype(np.float64)
    R = rng.integers(-1000, 1000, size=(E, 3)).astype(np.float64)
    vals = np.full(pat.nnz, np.nan)
    res = np.full(N, np.nan)
    order = [[] for _ in range(pat.nnz)]  # elements added into each CSR value
    owned_by = np.full(N, -1)
    used = 0
    assert po[0] == 0 and np.all(np.diff(po) >= 1) and np.all(np.diff(po) <= max_patch)
    for p in range(ap.num_patches):
        els = pe[po[p]:po[p + 1]]
        assert np.all(np.diff(els) > 0)
        dofs = ld[lo[p]:lo[p + 1]]
        assert np.all(np.diff(dofs) > 0) and dofs.size <= ap.max_patch_ldofs
        tl = tiles[lo[p]:lo[p + 1]]
        jl, rk, ps = jloc[po[p]:po[p + 1]], rank[po[p]:po[p + 1]], pos[po[p]:po[p + 1]]
        assert np.array_equal(dofs[jl], conn[els])
        assert np.array_equal(np.unique(conn[els]), dofs)
        pt = int(a["patch_tile"][p])
        tlen, steps = pt & 0xFFFFF, pt >> 20
        assert tlen <= max_tile
        used = max(used, tlen)
        # owned rows: every element of the node is in the patch, at most 16 entries,
        # rows one after another in ascending node order
        on = tl != NO_TILE
        for n in dofs[on]:
            assert owned_by[n] == -1
            owned_by[n] = p
            assert np.all(np.isin(node_elems[n], els))
        lens = (rp[dofs + 1] - rp[dofs])[on]
        assert np.all(lens <= 16)
        assert np.array_equal(tl[on], np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64))
        assert tlen == lens.sum()
        tile = np.zeros(tlen)
        rslot = np.zeros(dofs.size)
        tile_order = [[] for _ in range(tlen)]
        rslot_order = [[] for _ in range(dofs.size)]
        ranks_seen = []
        for step in range(steps):
            for li, e in enumerate(els):
                for k in range(3):
                    j = jl[li, k]
                    if tl[j] == NO_TILE or rk[li, k] != step:
                        continue
                    r = dofs[j]
                    assert rk[li, k] == np.searchsorted(node_elems[r], e)  # rank among all
                    ranks_seen.append(rk[li, k])
                    base = tl[j]
                    rslot[j] += R[e, k]
                    rslot_order[j].append(e)
                    for c in range(3):
                        q = (ps[li, k] >> (4 * c)) & 0xF
                        assert ci[rp[r] + q] == conn[e, c]
                        tile[base + q] += K[e, k, c]
                        tile_order[base + q].append(e)
        assert steps == (max(ranks_seen) + 1 if ranks_seen else 0)
        for j in np.flatnonzero(on):
            res[dofs[j]] = rslot[j]
            assert rslot_order[j] == list(node_elems[dofs[j]])
        # write-back along the runs: value k of the tile goes to run m's CSR
        # offset + (k - run m's tile offset), m the last run starting at or before k
        runs = a["runs"].reshape(-1, 2)[a["run_off"][p]:a["run_off"][p + 1]]
        fin = dofs[on]
        starts = np.flatnonzero(np.diff(np.concatenate([[-2], fin])) != 1)
        assert np.array_equal(runs[:, 0], tl[on][starts])
        assert np.array_equal(runs[:, 1], rp[fin[starts]])
        for k in range(tlen):
            m = np.searchsorted(runs[:, 0], k, side="right") - 1
            vals[runs[m, 1] + k - runs[m, 0]] = tile[k]
            order[runs[m, 1] + k - runs[m, 0]] = tile_order[k]
    assert ap.max_tile == used
    assert ap.max_runs == np.diff(a["run_off"]).max(initial=0)
    # every row is owned by exactly one patch or left to the row kernel
    assert np.array_equal(left, np.flatnonzero(owned_by < 0))
    # the row kernel: each left row over its elements in ascending order
    for r in left:
        res[r] = 0.0
        vals[rp[r]:rp[r + 1]] = 0.0
        for e in node_elems[r]:
            k = int(np.flatnonzero(conn[e] == r)[0])
            res[r] += R[e, k]
            for c in range(3):
                q = int(np.searchsorted(ci[rp[r]:rp[r + 1]], conn[e, c]))
                vals[rp[r] + q] += K[e, k, c]
                order[rp[r] + q].append(e)
    assert all(o == sorted(o) for o in order)
    rref = np.zeros(N)
    np.add.at(rref, conn.reshape(-1), R.reshape(-1))
    vref = np.zeros(pat.nnz)
    for e in range(E):
        for k in range(3):
            r = conn[e, k]
            for c in range(3):
                vref[rp[r] + np.searchsorted(ci[rp[r]:rp[r + 1]], conn[e, c])] += K[e, k, c]
    assert np.array_equal(res, rref) and np.array_equal(vals, vref)
    return ap


@pytest.mark.parametrize("n,max_patch", [(1, 256), (6, 17), (16, 64), (24, 256)])
def test_p1_asm_plan_structured(n, max_patch):
    g = structured_p1(n, "cpu")
    ap = _check(g.h_x, g.h_y, g.h_conn, max_patch, seed=n)
    # every node of a structured mesh is owned (rows of at most 7 entries)
    assert ap.num_left == 0
    if n == 24:
        # compact patches: the halo adds well under half of the elements again
        assert ap.num_entries < 1.5 * g.num_elems


def test_p1_asm_plan_scrambled_with_unused_node_and_small_tiles():
    rng = np.random.default_rng(4)
    x, y, c3 = O.mesh_p1(9)
    c3 = c3.reshape(-1, 3)[rng.permutation(2 * 81)]
    # an unused node (no element: an empty row, left to the row kernel)
    xs, ys = np.append(x, 2.0), np.append(y, 2.0)
    ap = _check(xs, ys, c3, 40, seed=1)
    assert np.array_equal(ap.h["rows_left"][: ap.num_left], [xs.size - 1])
    # tiles of 20 values: the patches close early, and still own every row
    small = _check(xs, ys, c3, 40, max_tile=20, seed=2)
    assert small.max_tile <= 20 and small.num_patches > ap.num_patches
    # patches of 5 elements cannot take a node of 6: such nodes are left
    tiny = _check(xs, ys, c3, 5, seed=3)
    assert tiny.num_left > 1


def test_p1_asm_plan_hub_row_over_sixteen_entries():
    """A hub vertex in 30 elements: its row has 31 entries, beyond the 4-bit
    positions, so it is left to the row kernel; the rim nodes are tiled."""
    m = 30
    ang = 2 * np.pi * np.arange(m) / m
    x = np.concatenate([[0.0], np.cos(ang)])
    y = np.concatenate([[0.0], np.sin(ang)])
    c3 = np.array([[0, 1 + i, 1 + (i + 1) % m] for i in range(m)], np.int32)
    ap = _check(x, y, c3, 256, seed=3)
    assert np.array_equal(ap.h["rows_left"][: ap.num_left], [0])


@pytest.mark.parametrize("n,lo", [(70, 0.05), (129, 0.3)])
def test_stencil_plan_records_reproduce_incidences(n, lo):
    """fx_p1_stencil_plan: a 64-row block gets a record exactly when its rows
    share column offsets and incidences (element A r + c_k, same local row and
    positions) within the bounds; the record's offsets, band-buffer indices
    (bands of consecutive offsets packed in order, 63 + w values each) and
    (c_k, meta_k) rebuild every row's columns and incidences."""
    from femx_amd.assembly import P1StencilPlan
    g = structured_p1(n, "cpu")
    amap = make_assembly_map(g.num_nodes, g.num_nodes, g.h_conn)
    sp = P1StencilPlan(amap, "cpu")
    rp, ci = amap.pattern.row_ptr, amap.pattern.col_ind
    inc_off, inc = amap.rowinc3()
    inc = inc.reshape(-1, 2)
    N = g.num_nodes
    code, w = sp.h_code, sp.h_words
    assert code.size == -(-N // 64)
    for b in range(code.size):
        r0 = 64 * b
        rows = range(r0, min(r0 + 64, N))
        offs = {tuple(ci[rp[r]:rp[r + 1]] - r) for r in rows}
        incs = [inc[inc_off[r]:inc_off[r + 1]] for r in rows]
        A_ = int(incs[1][0, 0] - incs[0][0, 0]) if len(rows) > 1 and len(incs[1]) else 0
        aff = {tuple((x[:, 0] - A_ * r).tolist() + x[:, 1].tolist()) for r, x in zip(rows, incs)}
        S = len(next(iter(offs)))
        T = len(incs[0])
        if len(offs) == 1:
            o = np.array(next(iter(offs)))
            nb_ = 1 + np.count_nonzero(np.diff(o) != 1)
            band = 63 * nb_ + S
        uniform = (len(rows) == 64 and len(offs) == 1 and len(aff) == 1 and 1 <= S <= 8
                   and 1 <= T <= 8 and band <= 208)
        assert (code[b] >= 0) == uniform, b
        if code[b] < 0:
            continue
        rec = w[code[b]:]
        assert rec[0] == S and rec[1] == T and rec[2] == A_
        off = rec[3:3 + S]
        assert tuple(off) == next(iter(offs))
        sidx = rec[3 + S:3 + 2 * S]
        used, q = 0, 0
        while q < S:
            wd = 1
            while q + wd < S and off[q + wd] == off[q] + wd:
                wd += 1
            assert list(sidx[q:q + wd]) == list(range(used, used + wd))
            used += 63 + wd
            q += wd
        pairs = rec[3 + 2 * S:3 + 2 * S + 2 * T].reshape(-1, 2)
        for r, x in zip(rows, incs):
            assert np.array_equal(A_ * r + pairs[:, 0], x[:, 0])
            assert np.array_equal(pairs[:, 1], x[:, 1])
    assert lo < sp.frac < 1.0
