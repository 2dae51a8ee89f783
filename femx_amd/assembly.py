"""Assembly maps, geometry, boundary maps and device assembly, mirroring
femx::assembly and femx::fem on the HIP kernels of libfemx_amd.so.

Reference interfaces replaced:
  make_assembly_map      femx/assembly/AssemblyMap.hpp:220-236 (AssemblyMap.cpp:54-176)
  make_boundary_map      femx/assembly/BoundaryMap.hpp:185-186 (BoundaryMap.cpp:59-139)
  replace_rows / replace_res / zero_boundary / prepare_forward_solve
                         femx/assembly/BoundaryMap.hpp:220-294 (device overloads)
  assemble               femx/assembly/CudaAssembly.hpp:319-364
Element operators are a closed set of precompiled kernels (PoissonP1 here,
TaylorHoodNS in femx_amd.ns): femx's templated evalRow cannot cross a C-ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .linalg import CsrMatrix, CsrPattern, _check_dev_vec, _require, ptr, stream_handle


def _flatten_dofs(dofs):
    """Element DOF tables as (offsets, flat): a 2-D array (uniform elements)
    or a sequence of per-element sequences (femx's Array<Array<Index>>)."""
    if isinstance(dofs, np.ndarray) and dofs.ndim == 2:
        E, k = dofs.shape
        return (np.arange(E + 1, dtype=np.int64) * k).astype(np.int32), \
            np.ascontiguousarray(dofs.reshape(-1), np.int32)
    lens = [len(d) for d in dofs]
    off = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    _require(off[-1] < 2 ** 31, "AssemblyMap flattened data exceeds the Index range")
    flat = (np.concatenate([np.asarray(d, np.int64) for d in dofs]) if lens and off[-1] > 0
            else np.zeros(0, np.int64))
    return off.astype(np.int32), flat.astype(np.int32)


class AssemblyMap:
    """Element-to-global DOF map and the CSR layout it induces.

    Holds femx's arrays (res/state offsets and DOFs, jac_offsets, jac_map,
    pattern) on the host and derives the row-incidence tables of the
    row-gather kernels; device copies are made on first use per device."""

    def __init__(self, num_res, num_states, res_dofs, state_dofs):
        self.res_offsets, self.res_dofs = _flatten_dofs(res_dofs)
        self.state_offsets, self.state_dofs = _flatten_dofs(state_dofs)
        _require(self.res_offsets.size == self.state_offsets.size,
                 "AssemblyMap residual/state element counts differ")
        E = self.res_offsets.size - 1
        h = C.c_void_p()
        _lib.call("fx_map_create", int(num_res), int(num_states), E, ptr(self.res_offsets),
                  ptr(self.res_dofs), ptr(self.state_offsets), ptr(self.state_dofs), C.byref(h))
        self._h = h
        self._destroy = _lib.lib().fx_map_destroy
        info = _lib.MapInfo()
        _lib.call("fx_map_info_get", h, C.byref(info))
        self.info = info
        self.num_elems, self.num_res, self.num_states = info.num_elems, info.num_res, info.num_states
        self.max_res, self.max_state, self.max_jac = info.max_res, info.max_state, info.max_jac
        self.max_row_len = info.max_row_len
        rp = np.empty(self.num_res + 1, np.int32)
        ci = np.empty(max(info.nnz, 1), np.int32)
        _lib.call("fx_map_export_pattern", h, ptr(rp), ptr(ci))
        self.pattern = CsrPattern(self.num_res, self.num_states, rp, ci[: info.nnz])
        self._jac = None
        self._rowinc3 = None
        self._dev = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def release_host(self):
        """Free the host builder (jac_map and incidence tables) once every
        device table has been exported; the pattern stays."""
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def _handle(self):
        _require(self._h is not None, "AssemblyMap host tables were released")
        return self._h

    def jac(self):
        """femx's (jac_offsets, jac_map) host arrays."""
        if self._jac is None:
            jo = np.empty(self.num_elems + 1, np.int32)
            jm = np.empty(max(self.info.jac_len, 1), np.int32)
            _lib.call("fx_map_export_jac", self._handle(), ptr(jo), ptr(jm))
            self._jac = (jo, jm[: self.info.jac_len])
        return self._jac

    def rowinc3(self):
        """(inc_off, inc) host arrays for 3x3 elements (fx_map_export_rowinc3)."""
        if self._rowinc3 is None:
            off = np.empty(self.num_res + 1, np.int32)
            inc = np.empty(max(2 * self.info.num_inc, 2), np.int32)
            _lib.call("fx_map_export_rowinc3", self._handle(), ptr(off), ptr(inc))
            self._rowinc3 = (off, inc[: 2 * self.info.num_inc])
        return self._rowinc3

    def device(self, name, device="cuda"):
        """Device copy of a named host table, cached per device."""
        dev = torch.device(device)
        key = (name, str(dev))
        if key not in self._dev:
            if name in ("inc_off", "inc"):
                arr = self.rowinc3()[0 if name == "inc_off" else 1]
            elif name == "state_dofs":
                arr = self.state_dofs
            elif name == "res_dofs":
                arr = self.res_dofs
            else:
                raise KeyError(name)
            self._dev[key] = torch.from_numpy(np.ascontiguousarray(arr)).to(dev)
        return self._dev[key]


def make_assembly_map(num_res, num_states, res_dofs, state_dofs=None):
    """Build a map from element DOF tables (femx makeAssemblyMap)."""
    return AssemblyMap(num_res, num_states, res_dofs, res_dofs if state_dofs is None else state_dofs)


class Geometry:
    """Flattened 2-D geometry (femx/fem/Geometry.hpp:32-107) stored SoA:
    x[N], y[N] on the device and element-major connectivity conn[E, k]."""

    def __init__(self, x, y, conn, device="cuda"):
        self.device = torch.device(device)
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        conn = np.ascontiguousarray(conn, np.int32)
        _require(x.shape == y.shape and x.ndim == 1, "Geometry coordinates must be 1-D x and y")
        _require(conn.ndim == 2, "Geometry connectivity must be [elements, nodes]")
        _require(conn.size == 0 or (conn.min() >= 0 and conn.max() < x.size),
                 "Geometry node index is out of range")
        self.num_nodes = x.size
        self.num_elems, self.nodes_per_elem = conn.shape
        self.dim = 2
        self.h_x, self.h_y, self.h_conn = x, y, conn
        self.x = torch.from_numpy(x).to(self.device)
        self.y = torch.from_numpy(y).to(self.device)
        self.conn = torch.from_numpy(conn.reshape(-1)).to(self.device)


def structured_p1(n, device="cuda"):
    """P1 triangle mesh of [0,1]^2 with n x n cells (fx_mesh_p1)."""
    N, E = (n + 1) ** 2, 2 * n * n
    x = np.empty(N)
    y = np.empty(N)
    conn = np.empty(3 * E, np.int32)
    _lib.call("fx_mesh_p1", n, ptr(x), ptr(y), ptr(conn))
    return Geometry(x, y, conn.reshape(E, 3), device)


def on_unit_square_edge(x, y, tol=1e-12):
    """Mask of the nodes on the boundary of [0,1]^2.  Structured coordinates
    are x_min + i*dx as femx's Mesh::makeStructuredQuad forms them
    (femx/fem/Mesh.cpp:19-29), so the last lattice coordinate can miss 1.0
    by an ulp (n = 49: 49 * (1/49) = 0.9999999999999999); exact comparison
    would drop the x = 1 and y = 1 edges for such n."""
    x = np.asarray(x)
    y = np.asarray(y)
    return ((np.abs(x) <= tol) | (np.abs(x - 1.0) <= tol) | (np.abs(y) <= tol)
            | (np.abs(y - 1.0) <= tol))


def unit_square_edge_nodes(x, y):
    """Ascending ids of the nodes on the boundary of [0,1]^2 (the order
    DirichletBC::addBoundary adds them, femx/fem/DirichletBC.cpp:138-145)."""
    return np.nonzero(on_unit_square_edge(x, y))[0].astype(np.int32)


# ---------------------------------------------------------------------------
# Boundary map


class BoundaryMap:
    """Constrained DOFs of one CSR layout (femx BoundaryMap)."""

    def __init__(self, dofs, pattern: CsrPattern, device="cuda"):
        _require(pattern.rows == pattern.cols, "BoundaryMap requires a square CSR pattern")
        dofs = np.ascontiguousarray(dofs, np.int32)
        h = C.c_void_p()
        _lib.call("fx_bcmap_create", pattern.rows, ptr(pattern.row_ptr), ptr(pattern.col_ind),
                  dofs.size, ptr(dofs), C.byref(h))
        try:
            info = _lib.BcMapInfo()
            _lib.call("fx_bcmap_info_get", h, C.byref(info))
            nb, nce = info.num_bcs, info.num_col_entries
            self.bc_rows = np.empty(nb, np.int32)
            self.diag = np.empty(nb, np.int32)
            self.col_offsets = np.empty(nb + 1, np.int32)
            self.col_entries = np.empty(nce, np.int32)
            self.col_rows = np.empty(nce, np.int32)
            self.bc_mask = np.empty(pattern.rows, np.int32)
            _lib.call("fx_bcmap_export", h, ptr(self.bc_rows), ptr(self.diag),
                      ptr(self.col_offsets), ptr(self.col_entries), ptr(self.col_rows),
                      ptr(self.bc_mask))
            ne, nen = info.num_elim_rows, info.num_elim_entries
            self.elim_rows = np.empty(ne, np.int32)
            self.elim_off = np.empty(ne + 1, np.int32)
            self.elim_k = np.empty(nen, np.int32)
            self.elim_ib = np.empty(nen, np.int32)
            _lib.call("fx_bcmap_export_elim", h, ptr(self.elim_rows), ptr(self.elim_off),
                      ptr(self.elim_k), ptr(self.elim_ib))
        finally:
            _lib.lib().fx_bcmap_destroy(h)
        self.rows = pattern.rows
        self.layout_id = pattern.layout_id
        self.num_bcs = nb
        self.device = torch.device(device)
        t = lambda a: torch.from_numpy(a).to(self.device)  # noqa: E731
        self.d_bc_rows, self.d_diag, self.d_mask = t(self.bc_rows), t(self.diag), t(self.bc_mask)
        self.d_elim_rows, self.d_elim_off = t(self.elim_rows), t(self.elim_off)
        self.d_elim_k, self.d_elim_ib = t(self.elim_k), t(self.elim_ib)


def make_boundary_map(dofs, pattern, device="cuda"):
    return BoundaryMap(dofs, pattern, device)


def _check_layout(bmap, mat):
    _require(mat.pattern.layout_id == bmap.layout_id,
             "BoundaryMap matrix does not match the mapped CSR layout")


def replace_rows(bmap: BoundaryMap, jac: CsrMatrix, diag, stream=None):
    """Zero constrained rows and set their diagonal; columns are kept."""
    _check_layout(bmap, jac)
    rp, _ = jac.pattern.device_arrays(jac.device)
    _lib.call("fx_bc_replace_rows", bmap.num_bcs, ptr(bmap.d_bc_rows), ptr(bmap.d_diag), ptr(rp),
              ptr(jac.vals), float(diag), stream_handle(stream))


def replace_res(bmap: BoundaryMap, state, bc_vals, res, stream=None):
    """res[row] = state[row] - bc_vals[ib] on constrained rows."""
    _check_dev_vec(state, bmap.rows, torch.float64, "BoundaryMap state")
    _check_dev_vec(res, bmap.rows, torch.float64, "BoundaryMap residual")
    _check_dev_vec(bc_vals, bmap.num_bcs, torch.float64, "BoundaryMap prescribed values")
    _lib.call("fx_bc_replace_res", bmap.num_bcs, ptr(bmap.d_bc_rows), ptr(state), ptr(bc_vals),
              ptr(res), stream_handle(stream))


def zero_boundary(bmap: BoundaryMap, vals, stream=None):
    _check_dev_vec(vals, bmap.rows, torch.float64, "BoundaryMap vector")
    _lib.call("fx_bc_zero", bmap.num_bcs, ptr(bmap.d_bc_rows), ptr(vals), stream_handle(stream))


def prepare_forward_solve(bmap: BoundaryMap, solve_mat: CsrMatrix, rhs, bc_vals, stream=None):
    """Eliminate constrained columns into rhs, set identity rows, rhs[bc]=g."""
    _check_layout(bmap, solve_mat)
    _check_dev_vec(rhs, bmap.rows, torch.float64, "BoundaryMap forward RHS")
    _check_dev_vec(bc_vals, bmap.num_bcs, torch.float64, "BoundaryMap prescribed values")
    rp, _ = solve_mat.pattern.device_arrays(solve_mat.device)
    _lib.call("fx_bc_prepare_forward", bmap.num_bcs, ptr(bmap.d_bc_rows), ptr(bmap.d_diag),
              ptr(rp), int(bmap.elim_rows.size), ptr(bmap.d_elim_rows), ptr(bmap.d_elim_off),
              ptr(bmap.d_elim_k), ptr(bmap.d_elim_ib), ptr(solve_mat.vals), ptr(rhs),
              ptr(bc_vals), stream_handle(stream))


# ---------------------------------------------------------------------------
# Element operators and assemble()


class PoissonP1:
    """P1 Poisson operator: femx's PoissonComponents Laplace element
    (examples/poisson/PoissonComponents.hpp:35-64) scaled by a per-element
    diffusion kappa[E], minus the consistent load of a nodal f[N].
    kappa=None means 1, f=None means 0."""

    def __init__(self, kappa=None, f=None, log_kappa=None):
        if kappa is not None and log_kappa is not None:
            raise RuntimeError("PoissonP1 takes kappa or log_kappa, not both")
        self.kappa, self.f, self.log_kappa = kappa, f, log_kappa


def _check_p1(geom: Geometry, amap: AssemblyMap):
    key = "_p1_checked"
    if getattr(amap, key, None) is geom:
        return
    _require(geom.num_elems == amap.num_elems,
             "Geometry and AssemblyMap have different element counts")
    _require(geom.nodes_per_elem == 3 and amap.max_res == 3 and amap.max_state == 3
             and amap.num_res == geom.num_nodes and amap.num_states == geom.num_nodes,
             "PoissonP1 needs a scalar P1 map over the geometry's nodes")
    _require(np.array_equal(amap.res_dofs, geom.h_conn.reshape(-1))
             and np.array_equal(amap.state_dofs, geom.h_conn.reshape(-1)),
             "PoissonP1 needs element DOFs equal to the element nodes")
    setattr(amap, key, geom)


def p1_patches(geom: Geometry, amap: AssemblyMap):
    """The element-patch plan of the P1 kernels on this mesh (fx_patches_create
    with the nodes as DOFs), built on first use and kept with the map, per
    device.  Calls that share it (its seam partials) must not overlap on
    different streams."""
    key = "_p1_patches_" + str(geom.device)
    p = getattr(amap, key, None)
    if p is None or p[0] is not geom:
        from .patches import Patches
        p = (geom, Patches(geom.num_nodes, geom.num_elems, 3, geom.h_conn, geom.h_conn, geom.h_x,
                           geom.h_y, geom.device))
        setattr(amap, key, p)
    return p[1]


def p1_asm_plan(geom: Geometry, amap: AssemblyMap):
    """The P1 patch assembly's plan (fx_p1_asm_create), built on first use
    and kept with the map, per device."""
    key = "_p1_asm_" + str(geom.device)
    p = getattr(amap, key, None)
    if p is None or p[0] is not geom:
        from .patches import P1AsmPlan
        p = (geom, P1AsmPlan(geom.h_conn, geom.h_x, geom.h_y, amap.pattern, geom.device))
        setattr(amap, key, p)
    return p[1]


class P1StencilPlan:
    """The P1 assembly's 64-row stencil blocks (fx_p1_stencil_plan) on one
    device: `blk_code` and `words` device tensors, `frac` the share of rows in
    uniform blocks."""

    MAX, BAND = 8, 208  # FX_P1_STENCIL_MAX, FX_P1_STENCIL_BAND

    def __init__(self, amap: "AssemblyMap", device):
        inc_off, inc = amap.rowinc3()
        rp, ci = amap.pattern.row_ptr, amap.pattern.col_ind
        N = amap.pattern.rows
        nb, nw = C.c_int32(), C.c_int32()
        args = (N, ptr(rp), ptr(ci), ptr(inc_off), ptr(inc), self.MAX, self.BAND)
        _lib.call("fx_p1_stencil_plan", *args, C.byref(nb), C.byref(nw), None, None)
        code = np.empty(max(nb.value, 1), np.int32)
        words = np.empty(max(nw.value, 1), np.int32)
        _lib.call("fx_p1_stencil_plan", *args, C.byref(nb), C.byref(nw), ptr(code), ptr(words))
        self.h_code, self.h_words = code[: nb.value], words[: nw.value]
        self.frac = 64.0 * np.count_nonzero(self.h_code >= 0) / max(N, 1)
        dev = torch.device(device)
        self.blk_code = torch.from_numpy(code).to(dev)
        self.words = torch.from_numpy(words).to(dev)


def p1_stencil_plan(geom: Geometry, amap: AssemblyMap):
    """The P1 assembly's stencil-block plan, built on first use and kept with
    the map, per device."""
    key = "_p1_stencil_" + str(geom.device)
    p = getattr(amap, key, None)
    if p is None or p[0] is not geom:
        p = (geom, P1StencilPlan(amap, geom.device))
        setattr(amap, key, p)
    return p[1]


STENCIL_MIN_FRAC = 0.5  # the stencil-block kernel when at least this share of rows is in uniform blocks


def assemble(op, geom: Geometry, amap: AssemblyMap, state, res, jac: CsrMatrix = None,
             stream=None, patches=True, stencil=None):
    """Assemble the residual into `res` and, when `jac` is given, the CSR
    Jacobian values into `jac.vals`; both outputs are replaced (femx
    assembly::assemble, CudaAssembly.hpp:319-364).  Returns `res`.

    When at least half of the rows lie in 64-row blocks that share one
    incidence stencil (lexicographically numbered structured meshes), those
    blocks assemble from coalesced loads of their rows' vertex data with the
    rows summed in registers (fx_p1_assemble_stencil; `stencil` True/False
    forces it on or off).  Otherwise, with `patches` (the default), the
    mesh's element patches assemble the rows of the nodes they own, each
    element evaluated once per patch (fx_p1_assemble_patch); without, one
    lane per row evaluates the row's elements (fx_p1_assemble).  The plans
    are built on the first call; the results are bitwise equal."""
    if not isinstance(op, PoissonP1):
        raise RuntimeError(f"assemble: unsupported element operator {type(op).__name__}")
    _check_p1(geom, amap)
    N, E = geom.num_nodes, geom.num_elems
    if state is not None:
        _check_dev_vec(state, amap.num_states, torch.float64, "Assembly state")
    _check_dev_vec(res, amap.num_res, torch.float64, "Assembly residual")
    if jac is not None:
        _require(jac.pattern.layout_id == amap.pattern.layout_id,
                 "Assembly matrix must use the AssemblyMap CSR layout")
        _require(res.data_ptr() != jac.vals.data_ptr() and
                 (state is None or state.data_ptr() not in (res.data_ptr(), jac.vals.data_ptr())),
                 "Assembly state, residual, and matrix values must not alias")
    coef = op.kappa if op.kappa is not None else op.log_kappa
    if coef is not None:
        _check_dev_vec(coef, E, torch.float64, "PoissonP1 kappa")
    if op.f is not None:
        _check_dev_vec(op.f, N, torch.float64, "PoissonP1 load")
    dev = geom.device
    rp, _ = amap.pattern.device_arrays(dev)
    if stencil is not False and E > 0:
        sp = p1_stencil_plan(geom, amap)
        if stencil or sp.frac >= STENCIL_MIN_FRAC:
            _lib.call("fx_p1_assemble_stencil", N, E, amap.pattern.nnz, amap.max_row_len,
                      ptr(geom.x), ptr(geom.y), ptr(geom.conn), ptr(coef),
                      int(op.log_kappa is not None), ptr(op.f), ptr(state),
                      ptr(amap.device("inc_off", dev)), ptr(amap.device("inc", dev)), ptr(rp),
                      ptr(sp.blk_code), ptr(sp.words),
                      ptr(jac.vals) if jac is not None else None, ptr(res),
                      stream_handle(stream))
            return res
    if patches and E > 0:
        ap = p1_asm_plan(geom, amap)
        _lib.call("fx_p1_assemble_patch", N, E, amap.pattern.nnz, ptr(geom.x), ptr(geom.y),
                  ptr(geom.conn), ptr(coef), int(op.log_kappa is not None), ptr(op.f),
                  ptr(state), ptr(amap.device("inc_off", dev)), ptr(amap.device("inc", dev)),
                  ptr(rp), C.byref(ap.plan),
                  ptr(jac.vals) if jac is not None else None, ptr(res), stream_handle(stream))
        return res
    _lib.call("fx_p1_assemble", N, E, amap.pattern.nnz, amap.max_row_len, ptr(geom.x),
              ptr(geom.y), ptr(geom.conn), ptr(coef), int(op.log_kappa is not None), ptr(op.f),
              ptr(state),
              ptr(amap.device("inc_off", dev)), ptr(amap.device("inc", dev)), ptr(rp),
              ptr(jac.vals) if jac is not None else None, ptr(res), stream_handle(stream))
    return res
