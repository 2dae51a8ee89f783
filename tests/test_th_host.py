"""Taylor-Hood host tables on the CPU: mesh/DOF conventions, the 15x15
row-incidence export and the work-block plan."""
import numpy as np
import pytest

from femx_amd import _lib, ns
from femx_amd.assembly import make_assembly_map
from femx_amd.linalg import ptr
from oracle import oracle as O


@pytest.mark.parametrize("n", [1, 2, 5])
def test_structured_th_tables(n):
    g = ns.structured_th(n, "cpu")
    x2, y2, c6, v3 = O.mesh_p2(n)
    assert np.array_equal(g.h_conn6.reshape(-1), c6) and np.array_equal(g.h_vconn3.reshape(-1), v3)
    assert np.array_equal(g.h_dofs15.reshape(-1), O.th_dofs(n, c6, v3))
    x1, y1, _ = O.mesh_p1(n)
    assert np.array_equal(g.h_xv, x1) and np.array_equal(g.h_yv, y1)
    # vertex k of each element sits at P2 node k of that element
    c6 = c6.reshape(-1, 6)
    v3 = v3.reshape(-1, 3)
    assert np.array_equal(x2[c6[:, :3]], x1[v3]) and np.array_equal(y2[c6[:, :3]], y1[v3])
    assert g.ndof == 2 * (2 * n + 1) ** 2 + (n + 1) ** 2


@pytest.mark.parametrize("n", [1, 3])
def test_rowinc15_and_blocks(n):
    g = ns.structured_th(n, "cpu")
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    plan = ns.THPlan(amap, "cpu")
    off = plan.inc_off.numpy()
    elem = plan.inc_elem.numpy()
    meta = plan.inc_meta.numpy().view(np.uint8).reshape(-1, 16)
    jo, jm = amap.jac()
    rp = amap.pattern.row_ptr
    d = g.h_dofs15
    for r in range(g.ndof):
        es = elem[off[r]:off[r + 1]]
        assert np.all(np.diff(es) > 0)
        for t in range(off[r], off[r + 1]):
            e, li = elem[t], meta[t, 15]
            assert d[e, li] == r
            assert np.array_equal(rp[r] + meta[t, :15].astype(np.int64), jm[225 * e + 15 * li:
                                                                           225 * e + 15 * li + 15])
    br, bf = plan.h_blk_row, plan.h_blk_flags
    assert br[0] == 0 and br[-1] == g.ndof and np.all(np.diff(br) > 0)
    nv2 = 2 * g.num_vel_nodes
    for b in range(plan.num_blocks):
        items = (br[b + 1] - br[b]) // (2 if bf[b] else 1)
        assert items <= ns.TH_ITEMS
        if bf[b]:
            assert br[b + 1] <= nv2 and br[b] % 2 == 0
        assert rp[br[b + 1]] - rp[br[b]] <= plan.max_block_nnz
        assert plan.tile_nnz == 0 or items == 1 or rp[br[b + 1]] - rp[br[b]] <= plan.tile_nnz
    # every velocity row is paired, every pressure row single
    paired = np.zeros(g.ndof, bool)
    for b in range(plan.num_blocks):
        paired[br[b]:br[b + 1]] = bool(bf[b])
    assert paired[:nv2].all() and not paired[nv2:].any()


def test_elem_inc_inverts_the_row_incidence():
    """elem_inc[15 e + li] is the incidence index of (e, li): reading a row's
    run of incidences back through it gives the row's elements, ascending."""
    g = ns.structured_th(3, "cpu")
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    plan = ns.THPlan(amap, "cpu")
    off = plan.inc_off.numpy()
    slots = plan.elem_inc.numpy()
    assert np.array_equal(np.sort(slots), np.arange(off[-1]))  # a permutation of the incidences
    d = g.h_dofs15.reshape(-1)
    row_of_slot = np.empty(off[-1], np.int64)
    row_of_slot[slots] = d  # the DOF (row) each (e, li) belongs to
    assert np.array_equal(row_of_slot, np.repeat(np.arange(g.ndof), np.diff(off)))
    elem_of_slot = np.empty(off[-1], np.int64)
    elem_of_slot[slots] = np.repeat(np.arange(g.num_elems), 15)
    for r in range(g.ndof):
        assert np.all(np.diff(elem_of_slot[off[r]:off[r + 1]]) > 0)


def test_rowinc15_rejects_other_shapes():
    amap = make_assembly_map(3, 3, [[0, 1, 2]])
    with pytest.raises(RuntimeError, match="15x15"):
        ns.THPlan(amap, "cpu")


def test_fourier_state_and_cavity_bc():
    g = ns.structured_th(4, "cpu")
    s = ns.fourier_state(g, 5)
    assert s.shape == (g.ndof,)
    top = np.nonzero(np.abs(g.h_y2 - 1.0) <= 1e-12)[0]
    assert top.size == 2 * 4 + 1 and np.all(s[2 * top] == 1.0)
    s2 = ns.fourier_state(g, 5)
    assert np.array_equal(s, s2)
    dofs, vals = ns.cavity_bc(g)
    assert np.all(np.diff(dofs) > 0) and dofs.size == 2 * 4 * 8
    assert vals.sum() == (2 * 4 + 1)


@pytest.mark.parametrize("n", [1, 3, 8])
def test_th_lane_map(n):
    """fx_map_th_lanes: every item of a block gets k_i >= 1 consecutive lanes
    with roles 0..k_i-1; R = max_i ceil(d_i / k_i) is the fewest rounds whose
    lane demand fits 64; unused lanes are -1."""
    g = ns.structured_th(n, "cpu")
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    plan = ns.THPlan(amap, "cpu")
    off = plan.inc_off.numpy()
    lanes = plan.h_blk_lanes.reshape(-1, 64)
    br, bf = plan.h_blk_row, plan.h_blk_flags
    saved = 0
    for b in range(plan.num_blocks):
        step = 2 if bf[b] else 1
        deg = np.diff(off)[br[b]:br[b + 1]:step]
        lm = lanes[b]
        live = lm[lm >= 0]
        item, role, k = live & 0xFF, (live >> 8) & 0xFF, live >> 16
        assert np.all(lm[live.size:] == -1)  # live lanes come first
        ks = []
        pos = 0
        for i, d in enumerate(deg):
            ki = int(k[pos])
            assert ki >= 1 and np.all(item[pos:pos + ki] == i)
            assert np.array_equal(role[pos:pos + ki], np.arange(ki)) and np.all(k[pos:pos + ki] == ki)
            ks.append(ki)
            pos += ki
        assert pos == live.size <= 64
        ks = np.array(ks)
        R = int(np.max((deg + ks - 1) // ks))
        need = lambda r: int(np.sum(np.maximum(1, (deg + r - 1) // r)))  # noqa: E731
        assert need(R) <= 64 and (R == 1 or need(R - 1) > 64 or R >= deg.max())
        assert np.array_equal(ks, np.maximum(1, (deg + R - 1) // R))
        saved += int(np.max((deg + 1) // 2)) - R  # rounds saved over two lanes per item
    assert saved > 0 or n == 1
    bad = np.array([0, 5, 3], np.int32)
    with pytest.raises(RuntimeError, match="block partition"):
        _lib.call("fx_map_th_lanes", amap._handle(), 2, ptr(bad), ptr(np.zeros(2, np.int32)),
                  ptr(np.empty(128, np.int32)))


@pytest.mark.parametrize("n,items,tile", [(1, 32, 0), (3, 32, 0), (8, 32, 0), (8, 16, 0),
                                           (8, 32, 600)])
def test_th_round_tables(n, items, tile):
    """fx_map_th_rounds: block b's lane l takes in round m the incidence
    t0 + role + m k of its lane-map item's row; past the lane's last
    incidence, and on idle lanes, an entry repeats a valid incidence (the
    lane's last, else the block's first) with bit 31 of its last meta word
    set; every incidence of every row appears exactly once unflagged, and
    R_b is the most incidences of any lane."""
    g = ns.structured_th(n, "cpu")
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    plan = ns.THPlan(amap, "cpu", items=items, tile_nnz=tile)
    off, elem = plan.inc_off.numpy(), plan.inc_elem.numpy()
    meta = plan.inc_meta.numpy().reshape(-1, 4).view(np.uint32)
    rto, rte = plan.rt_off.numpy(), plan.rt_elem.numpy()
    rtm = plan.rt_meta.numpy().reshape(-1, 4).view(np.uint32)
    lanes = plan.h_blk_lanes.reshape(-1, 64)
    br, bf = plan.h_blk_row, plan.h_blk_flags
    assert rto[0] == 0 and np.all(np.diff(rto) % 64 == 0) and rto[-1] == rte.size
    flag = np.uint32(1 << 31)
    seen = np.zeros(off[-1], np.int64)
    for b in range(plan.num_blocks):
        step = 2 if bf[b] else 1
        R = (rto[b + 1] - rto[b]) // 64
        lists = {}
        for lane, lm in enumerate(lanes[b]):
            if lm < 0:
                continue
            item, role, k = lm & 0xFF, (lm >> 8) & 0xFF, max(1, lm >> 16)
            r = br[b] + step * item
            lists[lane] = list(range(off[r] + role, off[r + 1], k))
        assert R == max([len(v) for v in lists.values()] + [1])  # at least one round
        fill = next((v[0] for _, v in sorted(lists.items()) if v), off[br[0]])
        for m in range(R):
            for lane in range(64):
                q = rto[b] + 64 * m + lane
                ts = lists.get(lane, [])
                if m < len(ts):
                    t = ts[m]
                    assert rte[q] == elem[t] and np.array_equal(rtm[q], meta[t])
                    seen[t] += 1
                else:
                    t = ts[-1] if ts else fill
                    assert rtm[q, 3] & flag and rte[q] == elem[t]
                    assert np.array_equal(rtm[q, :3], meta[t, :3]) and rtm[q, 3] == meta[t, 3] | flag
    # every row's incidences, each once (rows of a pair block: both rows' lists
    # are the same elements, taken once through the item's first row)
    for b in range(plan.num_blocks):
        rows = range(br[b], br[b + 1], 2 if bf[b] else 1)
        for r in rows:
            assert np.all(seen[off[r]:off[r + 1]] == 1)
    bad = np.array([0, 5, 3], np.int32)
    with pytest.raises(RuntimeError, match="block partition"):
        _lib.call("fx_map_th_rounds", off.size - 1, 2, ptr(bad), ptr(np.zeros(2, np.int32)),
                  ptr(np.full(128, -1, np.int32)), ptr(off), ptr(elem), ptr(plan.inc_meta.numpy()),
                  ptr(np.empty(3, np.int32)), None, None)
    past = np.array([0, off.size], np.int32)  # a block beyond the last row
    with pytest.raises(RuntimeError, match="block partition"):
        _lib.call("fx_map_th_rounds", off.size - 1, 1, ptr(past), ptr(np.zeros(1, np.int32)),
                  ptr(np.full(64, -1, np.int32)), ptr(off), ptr(elem), ptr(plan.inc_meta.numpy()),
                  ptr(np.empty(2, np.int32)), None, None)
    with pytest.raises(RuntimeError, match="invalid arguments"):
        _lib.call("fx_map_th_rounds", off.size - 1, plan.num_blocks, ptr(br), ptr(bf),
                  ptr(plan.h_blk_lanes), ptr(off), None, None, ptr(np.empty(plan.num_blocks + 1,
                                                                         np.int32)),
                  ptr(np.empty(64, np.int32)), ptr(np.empty(256, np.int32)))


@pytest.mark.parametrize("n", [49, 499])
def test_lid_and_walls_found_when_the_last_coordinate_misses_one(n):
    """femx's structured coordinates are x_min + i*dx (Mesh.cpp:19-29); for
    these n the last one is 0.9999999999999999, and the walls at x = 1 and
    y = 1 must still be found (they were dropped by exact comparison)."""
    g = ns.structured_th(n, "cpu")
    assert g.h_y2.max() < 1.0  # the case this test is about
    dofs, vals = ns.cavity_bc(g)
    assert dofs.size == 2 * 4 * (2 * n)  # both components of every edge node
    assert np.count_nonzero(vals) == 2 * n + 1  # u_x = 1 on the whole lid, corners included
    s = ns.fourier_state(g, 3)
    top = np.argsort(g.h_y2)[-(2 * n + 1):]
    assert np.all(s[2 * top] == 1.0)
    from femx_amd.assembly import structured_p1, unit_square_edge_nodes
    p = structured_p1(n, "cpu")
    assert unit_square_edge_nodes(p.h_x, p.h_y).size == 4 * n


def _check_patches(g, max_patch, rng):
    """Invariants of the VJP patch plan, and the reduction it describes
    applied to integer-valued element columns (exact in any order) against a
    plain scatter-add, with lam's pass-through on constrained DOFs."""
    pp = ns.THPatches(g, "cpu", max_patch)
    h = pp.h
    E, P = g.num_elems, pp.num_patches
    d15 = g.h_dofs15
    po, pe = h["patch_off"], h["patch_elem"][:E]
    assert po[0] == 0 and po[-1] == E and np.all(np.diff(po) >= 1) and np.all(np.diff(po) <= max_patch)
    assert np.array_equal(np.sort(pe), np.arange(E))
    lo, ld, lp, le = h["ldof_off"], h["ldof_dof"], h["ldof_part"], h["ldof_end"].astype(np.int64)
    so, sd = h["seam_off"], h["seam_dof"][:pp.num_seam]
    packed = h["elem_ldof"].reshape(-1, 16)[:E, :15].astype(np.int64)
    jloc, rank = packed & 0xfff, packed >> 12
    FIRST = 1 << 30
    npatch = np.zeros(g.ndof, np.int64)
    for p in range(P):
        npatch[np.unique(d15[pe[po[p]:po[p + 1]]])] += 1
    seam = np.flatnonzero(npatch != 1)
    assert np.array_equal(sd, seam) and np.array_equal(np.diff(so), npatch[seam])
    c = rng.integers(-1000, 1000, size=(E, 15)).astype(np.float64)
    lam = rng.integers(-1000, 1000, size=g.ndof).astype(np.float64)
    mask = rng.random(g.ndof) < 0.2
    parts = np.full(pp.num_parts, np.nan)
    jt = np.full(g.ndof, np.nan)
    firsts = 0
    for p in range(P):
        els = pe[po[p]:po[p + 1]]
        assert np.all(np.diff(els) > 0)
        assert np.array_equal(h["elem_vert"].reshape(-1, 3)[po[p]:po[p + 1]], g.h_vconn3[els])
        dofs = ld[lo[p]:lo[p + 1]]
        assert np.all(np.diff(dofs) > 0) and dofs.size <= pp.max_patch_ldofs
        ends = le[lo[p]:lo[p + 1]]
        begs = np.concatenate([[0], ends[:-1]])
        jl, rk = jloc[po[p]:po[p + 1]], rank[po[p]:po[p + 1]]
        assert np.array_equal(dofs[jl], d15[els])
        assert np.all(ends - begs <= 16)
        # ranks number each run's elements 0, 1, ... in element order
        for j in range(dofs.size):
            r = np.sort(rk[jl == j])
            assert np.array_equal(r, np.arange(ends[j] - begs[j]))
            assert np.all(np.diff(rk[jl == j]) > 0)
        slot = begs[jl] + rk
        assert np.array_equal(np.sort(slot.reshape(-1)), np.arange(ends[-1]))
        cs = np.empty(ends[-1])
        cs[slot.reshape(-1)] = c[els].reshape(-1)
        for j in range(dofs.size):
            v = cs[begs[j]:ends[j]].sum()
            d, part = dofs[j], lp[lo[p] + j]
            if part < 0:
                assert npatch[d] == 1
                jt[d] = v + (lam[d] if mask[d] else 0.0)
            else:
                first = bool(part & FIRST)
                firsts += first
                part &= ~FIRST
                assert first == (part == so[np.searchsorted(sd, d)])
                assert np.isnan(parts[part])
                parts[part] = v + (lam[d] if first and mask[d] else 0.0)
    assert firsts == np.count_nonzero(npatch[seam] > 0)
    # a DOF in no element has no partial: the seam pass gives it lam's
    # pass-through alone
    empty = seam[npatch[seam] == 0]
    assert not np.isnan(parts).any()
    for k, dof in enumerate(sd):
        jt[dof] = parts[so[k]:so[k + 1]].sum() if so[k + 1] > so[k] else (lam[dof] if mask[dof] else 0.0)
    ref = np.where(mask, lam, 0.0)
    np.add.at(ref, d15.reshape(-1), c.reshape(-1))
    assert np.array_equal(jt, ref)
    return pp, empty


@pytest.mark.parametrize("n,max_patch", [(1, 256), (4, 7), (12, 64), (30, 256)])
def test_vjp_patch_plan_structured(n, max_patch):
    pp, _ = _check_patches(ns.structured_th(n, "cpu"), max_patch, np.random.default_rng(n))
    if n == 30:
        # Morton patches are compact: most DOFs are finished inside a patch
        assert pp.num_seam < 0.35 * ns.structured_th(n, "cpu").ndof


def test_vjp_patch_plan_scrambled_and_unstructured():
    rng = np.random.default_rng(5)
    x, y, c3 = O.mesh_p1(9)
    c3 = c3.reshape(-1, 3)[rng.permutation(2 * 81)]
    g = ns.th_from_p1(x, y, c3, "cpu")
    _check_patches(g, 32, rng)
    # a mesh with an unused vertex: its pressure DOF is in no element
    xs = np.append(x, 2.0)
    ys = np.append(y, 2.0)
    g2 = ns.th_from_p1(xs, ys, c3, "cpu")
    pp, empty = _check_patches(g2, 256, rng)
    unused = np.setdiff1d(np.arange(g2.ndof), g2.h_dofs15)
    assert unused.size == 3 and g2.ndof - 1 in unused and np.array_equal(empty, unused)


def test_vjp_patch_plan_rejects_bad_sizes():
    g = ns.structured_th(2, "cpu")
    for mp in (0, 257):
        with pytest.raises(RuntimeError, match="max_patch"):
            ns.THPatches(g, "cpu", mp)


def test_vjp_patch_plan_cuts_runs_at_sixteen_elements():
    """A hub vertex shared by 40 elements: its DOFs' runs are capped at 16
    elements per patch, so the fan is cut into at least three patches."""
    m = 40
    ang = 2 * np.pi * np.arange(m) / m
    x = np.concatenate([[0.0], np.cos(ang)])
    y = np.concatenate([[0.0], np.sin(ang)])
    c3 = np.array([[0, 1 + i, 1 + (i + 1) % m] for i in range(m)], np.int32)
    g = ns.th_from_p1(x, y, c3, "cpu")
    pp, empty = _check_patches(g, 256, np.random.default_rng(1))
    assert pp.num_patches >= 3 and empty.size == 0
