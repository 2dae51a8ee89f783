"""femx's own unstructured Gmsh meshes (data/meshes, read by femx's
GmshReader into the gmsh_*.npz fixtures) on the host: the map builder, the
boundary map and the oracle bit for bit against femx's outputs, the GLS node
plan, and the Taylor-Hood mesh builder (fx_th_mesh) against brute force."""
import numpy as np
import pytest

from conftest import golden
from femx_amd import gls, ns
from femx_amd.assembly import BoundaryMap, make_assembly_map
from femx_amd.linalg import CsrPattern
from oracle import oracle as O

MESHES = ["2d_square_res_low", "2d_rectangle", "2d_square_res_very_low"]


@pytest.mark.parametrize("name", MESHES)
def test_gmsh_map_and_boundary_map_match_reference(name):
    g = golden(f"gmsh_{name}.npz")
    N, conn = g["x"].size, g["conn"]
    amap = make_assembly_map(N, N, conn.reshape(-1, 3))
    assert np.array_equal(amap.pattern.row_ptr, g["row_ptr"])
    assert np.array_equal(amap.pattern.col_ind, g["col_ind"])
    jo, jm = amap.jac()
    assert np.array_equal(jo, g["jac_offsets"]) and np.array_equal(jm, g["jac_map"])
    b = BoundaryMap(g["bc_dofs"], amap.pattern, device="cpu")
    for k in ("bc_rows", "diag", "col_offsets", "col_entries", "col_rows", "bc_mask"):
        assert np.array_equal(getattr(b, k), g["bc_" + k]), k
    # the facets of the groups are the femx nodes the DOFs came from
    assert set(g["bc_dofs"].tolist()) <= set(g["facet_nids"].tolist())


@pytest.mark.parametrize("name", MESHES)
def test_gmsh_oracle_matches_reference(name):
    """The C restatement on femx's unstructured meshes: Poisson K and K u
    (kappa = 1, f = 0) and prepareForwardSolve bit for bit."""
    g = golden(f"gmsh_{name}.npz")
    x, y, conn = g["x"], g["y"], g["conn"]
    N, E = x.size, conn.size // 3
    rp, ci, jo, jm = O.map_build(N, N, O.uniform_offsets(E, 3), conn, O.uniform_offsets(E, 3),
                                 conn)
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(jm, g["jac_map"])
    vals, res = O.p1_assemble(x, y, conn, np.ones(E), None, g["state"], jm, ci.size)
    assert np.array_equal(vals, g["vals"]) and np.array_equal(res, g["res"])
    bc = O.bc_build(rp, ci, g["bc_dofs"])
    v, r = g["vals"].copy(), g["fwd_rhs0"].copy()
    O.prepare_forward(rp, bc, v, r, g["fwd_g"])
    assert np.array_equal(v, g["fwd_vals"]) and np.array_equal(r, g["fwd_rhs"])


@pytest.mark.parametrize("name", ["2d_rectangle", "2d_square_res_very_low"])
def test_gmsh_gls_oracle_and_plan(name):
    g = golden(f"gmsh_{name}.npz")
    x, y, conn = g["x"], g["y"], g["conn"]
    geom = gls.GLSGeometry(x, y, conn, "cpu")
    assert np.array_equal(geom.h_dofs9.reshape(-1), g["ns_dofs9"])
    amap = gls.make_navier_map(geom)
    assert np.array_equal(amap.pattern.row_ptr, g["ns_row_ptr"])
    assert np.array_equal(amap.pattern.col_ind, g["ns_col_ind"])
    gls.GLSPlan(amap, geom, "cpu")  # the node plan accepts the unstructured mesh
    _, jm = amap.jac()
    nd = geom.num_states
    vals, res = O.gls_assemble(x, y, conn, g["ns_dofs9"], nd, 1.0, 1e-3, 0.01, 1, 2, g["ns_hist"],
                               g["ns_nxt"], jm, amap.pattern.nnz)
    assert np.array_equal(vals, g["ns_vals"]) and np.array_equal(res, g["ns_res"])
    b = BoundaryMap(g["ns_bc_dofs"], amap.pattern, device="cpu")
    assert np.array_equal(b.bc_mask, g["ns_bc_mask"]) and np.array_equal(b.diag, g["ns_bc_diag"])


def brute_th_mesh(x, conn):
    """Edge nodes numbered by first appearance, local edges (0,1),(1,2),(2,0)."""
    ids, c6 = {}, []
    for tri in conn.reshape(-1, 3):
        row = list(tri)
        for k in range(3):
            key = tuple(sorted((int(tri[k]), int(tri[(k + 1) % 3]))))
            if key not in ids:
                ids[key] = x.size + len(ids)
            row.append(ids[key])
        c6.append(row)
    return np.array(c6, np.int32), ids


@pytest.mark.parametrize("name", ["2d_rectangle", "2d_square_res_very_low"])
def test_th_mesh_builder_on_gmsh_meshes(name):
    g = golden(f"gmsh_{name}.npz")
    x, y, conn = g["x"], g["y"], g["conn"]
    geom = ns.th_from_p1(x, y, conn, "cpu")
    c6, ids = brute_th_mesh(x, conn)
    assert np.array_equal(geom.h_conn6, c6)
    np2 = x.size + len(ids)
    assert geom.num_vel_nodes == np2 and geom.ndof == 2 * np2 + x.size
    # a disc: V - E_edges + F = 1; boundary edges are exactly the facets
    E = conn.size // 3
    assert len(ids) == x.size + E - 1
    count = np.bincount(c6[:, 3:].reshape(-1) - x.size, minlength=len(ids))
    bnd = {k for k, v in ids.items() if count[v - x.size] == 1}
    facets = {tuple(sorted(map(int, f))) for f in g["facet_nids"].reshape(-1, 2)}
    assert bnd == facets and set(count.tolist()) == {1, 2}
    # midpoints and the MixedFESpace DOF order
    for (a, b), v in list(ids.items())[:50]:
        assert geom.h_x2[v] == 0.5 * (x[a] + x[b]) and geom.h_y2[v] == 0.5 * (y[a] + y[b])
    d = geom.h_dofs15
    assert np.array_equal(d[:, 0:12:2], 2 * c6) and np.array_equal(d[:, 1:12:2], 2 * c6 + 1)
    assert np.array_equal(d[:, 12:], 2 * np2 + conn.reshape(-1, 3))


def test_th_mesh_builder_rejects_bad_tables():
    from femx_amd import _lib
    from femx_amd.linalg import ptr
    import ctypes as C
    n = C.c_int32()
    bad = np.array([0, 1, 1], np.int32)
    with pytest.raises(RuntimeError, match="repeated vertex"):
        _lib.call("fx_th_mesh", 3, 1, ptr(bad), C.byref(n), None, None, None)
    bad = np.array([0, 1, 3], np.int32)
    with pytest.raises(RuntimeError, match="out of range"):
        _lib.call("fx_th_mesh", 3, 1, ptr(bad), C.byref(n), None, None, None)
    three = np.array([0, 1, 2, 1, 0, 3, 0, 1, 4], np.int32)  # edge (0,1) in three triangles
    with pytest.raises(RuntimeError, match="more than two"):
        _lib.call("fx_th_mesh", 5, 3, ptr(three), C.byref(n), None, None, None)


def test_th_blocks_fit_the_lds_tile():
    """With no bound asked for, blocks are cut at the LDS tile capacity; an
    explicit smaller bound is honoured; a single item always fits."""
    g = ns.structured_th(6, "cpu")
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    p = ns.THPlan(amap, "cpu")
    assert p.max_block_nnz <= 4928
    q = ns.THPlan(amap, "cpu", tile_nnz=300)
    assert q.max_block_nnz <= max(300, 2 * amap.max_row_len) and q.num_blocks > p.num_blocks
