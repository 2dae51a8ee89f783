"""Taylor-Hood P2/P1 steady Navier-Stokes on the HIP kernels.

`THGeometry` holds the vertex coordinates (the geometry of the affine P2
element), the vertex and P2 connectivity, and the element DOF table in the
FESpace/MixedFESpace numbering (velocity DOFs 2*node + c, then pressure DOFs
offset by 2*N_v; femx/fem/FESpace.cpp:96-106, femx/fem/MixedFESpace.cpp:
121-137).  `assemble_ns` drops in for assembly::assemble with this operator
(femx/assembly/CudaAssembly.hpp:319-364); `NavierStokesResidual` fills the
stationary Residual contract (femx/state/Residual.hpp:18-60) with the
per-element viscosity as control, and `NSResidualVJP` exposes it to torch
autograd.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .assembly import (AssemblyMap, make_assembly_map, make_boundary_map, replace_res, replace_rows,
                       unit_square_edge_nodes)
from .linalg import CsrMatrix, _check_dev_vec, _require, ptr, stream_handle
from .patches import Patches

TH_ITEMS = 32  # work items per block (FX_TH_BLOCK_ITEMS)
TH_TILE_NNZ = 0  # CSR entries per block row tile, 0 = bounded by items (FX_TH_TILE_NNZ)


class THGeometry:
    """Structured or user-supplied Taylor-Hood mesh data (device copies)."""

    def __init__(self, xv, yv, vconn3, conn6, dofs15, num_vel_nodes, device="cuda"):
        self.device = torch.device(device)
        self.h_xv = np.ascontiguousarray(xv, np.float64)
        self.h_yv = np.ascontiguousarray(yv, np.float64)
        self.h_vconn3 = np.ascontiguousarray(vconn3, np.int32).reshape(-1, 3)
        self.h_conn6 = np.ascontiguousarray(conn6, np.int32).reshape(-1, 6)
        self.h_dofs15 = np.ascontiguousarray(dofs15, np.int32).reshape(-1, 15)
        self.num_elems = self.h_dofs15.shape[0]
        self.num_vertices = self.h_xv.size
        self.num_vel_nodes = int(num_vel_nodes)
        self.ndof = 2 * self.num_vel_nodes + self.num_vertices
        _require(self.h_vconn3.shape[0] == self.num_elems == self.h_conn6.shape[0],
                 "Taylor-Hood tables have different element counts")
        t = lambda a: torch.from_numpy(a.reshape(-1)).to(self.device)  # noqa: E731
        self.xv, self.yv = t(self.h_xv), t(self.h_yv)
        self.vconn3, self.dofs15 = t(self.h_vconn3), t(self.h_dofs15)


def structured_th(n, device="cuda"):
    """Taylor-Hood mesh of [0,1]^2 with n x n cells (fx_mesh_p2, fx_th_dofs)."""
    Nv, E = (2 * n + 1) ** 2, 2 * n * n
    x = np.empty(Nv)
    y = np.empty(Nv)
    c6 = np.empty(6 * E, np.int32)
    v3 = np.empty(3 * E, np.int32)
    _lib.call("fx_mesh_p2", n, ptr(x), ptr(y), ptr(c6), ptr(v3))
    d = np.empty(15 * E, np.int32)
    _lib.call("fx_th_dofs", n, ptr(c6), ptr(v3), ptr(d))
    # vertex coordinates: lattice nodes (2i, 2j) of the P2 lattice
    m = 2 * n + 1
    idx = (np.arange(n + 1)[:, None] * 2 * m + np.arange(n + 1)[None, :] * 2).reshape(-1)
    geom = THGeometry(x[idx], y[idx], v3, c6, d, Nv, device)
    geom.h_x2, geom.h_y2 = x, y  # P2 node coordinates (for boundary marking, states)
    geom.n = n
    return geom


def th_from_p1(x, y, conn3, device="cuda"):
    """Taylor-Hood mesh over any conforming triangle mesh given as femx holds
    a P1 mesh (vertex coordinates and 3 vertex ids per triangle, e.g. from
    femx's GmshReader): P2 edge nodes at the edge midpoints, numbered after
    the vertices in order of first appearance (fx_th_mesh)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    conn3 = np.ascontiguousarray(conn3, np.int32).reshape(-1)
    E = conn3.size // 3
    np2 = C.c_int32()
    _lib.call("fx_th_mesh", x.size, E, ptr(conn3), C.byref(np2), None, None, None)
    ne = np2.value - x.size
    c6 = np.empty(max(6 * E, 1), np.int32)
    ends = np.empty(max(2 * ne, 1), np.int32)
    d = np.empty(max(15 * E, 1), np.int32)
    _lib.call("fx_th_mesh", x.size, E, ptr(conn3), C.byref(np2), ptr(c6), ptr(ends), ptr(d))
    ends = ends[: 2 * ne].reshape(-1, 2)
    geom = THGeometry(x, y, conn3, c6[: 6 * E], d[: 15 * E], np2.value, device)
    geom.h_x2 = np.concatenate([x, 0.5 * (x[ends[:, 0]] + x[ends[:, 1]])])
    geom.h_y2 = np.concatenate([y, 0.5 * (y[ends[:, 0]] + y[ends[:, 1]])])
    geom.h_edge_ends = ends
    return geom


class THBlocks:
    """One block layout of the Taylor-Hood assembly (fx_map_th_blocks with
    `items` work items per block and row tile bound `tile_nnz`, and its lane
    map), on one device."""

    def __init__(self, amap: AssemblyMap, device, items, tile_nnz, rowinc=None):
        self.items, self.tile_nnz = int(items), int(tile_nnz)
        nb, mx = C.c_int32(), C.c_int32()
        args = (amap._handle(), self.items, self.tile_nnz)
        _lib.call("fx_map_th_blocks", *args, C.byref(nb), C.byref(mx), None, None)
        blk_row = np.empty(nb.value + 1, np.int32)
        blk_flags = np.empty(max(nb.value, 1), np.int32)
        _lib.call("fx_map_th_blocks", *args, C.byref(nb), C.byref(mx), ptr(blk_row), ptr(blk_flags))
        self.num_blocks, self.max_block_nnz = nb.value, mx.value
        self.h_blk_row, self.h_blk_flags = blk_row, blk_flags[: nb.value]
        # each block's 64 lanes shared out so its items finish together
        lanes = np.empty(max(64 * nb.value, 1), np.int32)
        _lib.call("fx_map_th_lanes", amap._handle(), nb.value, ptr(blk_row), ptr(blk_flags),
                  ptr(lanes))
        self.h_blk_lanes = lanes[: 64 * nb.value]
        t = lambda a: torch.from_numpy(a).to(device)  # noqa: E731
        self.blk_row, self.blk_flags = t(blk_row), t(self.h_blk_flags)
        self.blk_lanes = t(self.h_blk_lanes)
        # round tables (fx_map_th_rounds) from the exported row incidences
        self.rt_off = self.rt_elem = self.rt_meta = None
        if rowinc is not None:
            off, elem, meta = rowinc
            rt_off = np.empty(nb.value + 1, np.int32)
            args = (amap.num_res, nb.value, ptr(blk_row), ptr(self.h_blk_flags),
                    ptr(self.h_blk_lanes), ptr(off), ptr(elem), ptr(meta))
            _lib.call("fx_map_th_rounds", *args, ptr(rt_off), None, None)
            n = int(rt_off[-1])
            rt_elem = np.empty(max(n, 1), np.int32)
            rt_meta = np.empty(max(4 * n, 4), np.int32)
            _lib.call("fx_map_th_rounds", *args, ptr(rt_off), ptr(rt_elem), ptr(rt_meta))
            self.h_rt_off, self.h_rt_elem, self.h_rt_meta = rt_off, rt_elem[:n], rt_meta[: 4 * n]
            self.rt_off, self.rt_elem, self.rt_meta = t(rt_off), t(rt_elem), t(rt_meta)


class THPlan:
    """Row-incidence tables, work blocks and the two-phase VJP's workspace
    of the Taylor-Hood kernels for one assembly map, on one device.  Calls
    that share a plan share its workspace: VJPs of one map on different
    streams must not overlap.  `variants` lists extra (items, tile_nnz) block layouts to build
    while the map's host tables exist; `select` switches among them (results
    do not depend on the layout: every entry is summed in ascending element
    order whatever the blocks)."""

    def __init__(self, amap: AssemblyMap, device="cuda", items=TH_ITEMS, tile_nnz=TH_TILE_NNZ,
                 variants=()):
        dev = torch.device(device)
        info = amap.info
        off = np.empty(amap.num_res + 1, np.int32)
        elem = np.empty(max(info.num_inc, 1), np.int32)
        meta = np.empty(max(4 * info.num_inc, 4), np.int32)
        _lib.call("fx_map_export_rowinc15", amap._handle(), ptr(off), ptr(elem), ptr(meta))
        t = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
        self.inc_off, self.inc_elem = t(off), t(elem[: info.num_inc])
        self.inc_meta = t(meta[: 4 * info.num_inc])
        slots = np.empty(max(15 * amap.num_elems, 1), np.int32)
        _lib.call("fx_map_export_rowinc", amap._handle(), ptr(off), None, ptr(slots))
        self.elem_inc = t(slots[: 15 * amap.num_elems])
        self._work, self._dev, self._num_elems = None, dev, amap.num_elems
        self.layouts = {}
        rowinc = (off, elem, meta)
        for it, tl in [(items, tile_nnz)] + list(variants):
            if (int(it), int(tl)) not in self.layouts:
                self.layouts[(int(it), int(tl))] = THBlocks(amap, dev, it, tl, rowinc)
        self.select(items, tile_nnz)

    @property
    def work(self):
        """The two-phase VJP's workspace (fx_th_vjp: 15 doubles per element),
        allocated on first use."""
        if self._work is None:
            self._work = torch.empty(15 * self._num_elems, dtype=torch.float64, device=self._dev)
        return self._work

    def select(self, items, tile_nnz):
        """Use the block layout (items, tile_nnz), built at construction."""
        b = self.layouts[(int(items), int(tile_nnz))]
        self.items, self.tile_nnz = b.items, b.tile_nnz
        self.num_blocks, self.max_block_nnz = b.num_blocks, b.max_block_nnz
        self.h_blk_row, self.h_blk_flags, self.h_blk_lanes = b.h_blk_row, b.h_blk_flags, b.h_blk_lanes
        self.blk_row, self.blk_flags, self.blk_lanes = b.blk_row, b.blk_flags, b.blk_lanes
        self.rt_off, self.rt_elem, self.rt_meta = b.rt_off, b.rt_elem, b.rt_meta


TH_PATCH = 256  # elements per VJP patch (FX_TH_PATCH_MAX)


def THPatches(geom: "THGeometry", device, max_patch=TH_PATCH):
    """Patch plan of the single-pass Taylor-Hood VJP (15 DOFs per element)."""
    return Patches(geom.ndof, geom.num_elems, 15, geom.h_dofs15, geom.h_vconn3, geom.h_xv,
                   geom.h_yv, device, max_patch)


def _patches(geom, amap, device):
    key = "_th_patches_" + str(torch.device(device))
    p = getattr(amap, key, None)
    if p is None:
        p = THPatches(geom, device)
        setattr(amap, key, p)
    return p


class TaylorHoodNS:
    """Steady incompressible NS operator: viscosity nu (scalar) or nu_e[E]."""

    def __init__(self, nu, nu_e=None):
        self.nu, self.nu_e = float(nu), nu_e


def _plan(amap, device):
    key = "_th_plan_" + str(torch.device(device))
    p = getattr(amap, key, None)
    if p is None:
        p = THPlan(amap, device)
        setattr(amap, key, p)
    return p


def _check(geom: THGeometry, amap: AssemblyMap):
    if getattr(amap, "_th_checked", None) is geom:
        return
    _require(geom.num_elems == amap.num_elems,
             "Geometry and AssemblyMap have different element counts")
    _require(amap.num_res == geom.ndof and amap.num_states == geom.ndof
             and amap.max_res == 15 and amap.max_state == 15,
             "TaylorHoodNS needs a 15x15 map over the Taylor-Hood DOFs")
    _require(np.array_equal(amap.res_dofs, geom.h_dofs15.reshape(-1))
             and np.array_equal(amap.state_dofs, geom.h_dofs15.reshape(-1)),
             "TaylorHoodNS needs element DOFs equal to the geometry's DOF table")
    amap._th_checked = geom


def assemble_ns(op: TaylorHoodNS, geom: THGeometry, amap: AssemblyMap, state, res,
                jac: CsrMatrix = None, stream=None, rounds=True):
    """Residual into `res` and, when `jac` is given, the Newton Jacobian into
    `jac.vals`; outputs replaced.  Returns `res`.  `rounds` reads each block's
    incidences through its round table (fx_th_assemble_rounds) instead of
    through the lane map and row offsets (fx_th_assemble); the results are
    the same bit for bit."""
    _check(geom, amap)
    _check_dev_vec(state, geom.ndof, torch.float64, "Assembly state")
    _check_dev_vec(res, geom.ndof, torch.float64, "Assembly residual")
    if jac is not None:
        _require(jac.pattern.layout_id == amap.pattern.layout_id,
                 "Assembly matrix must use the AssemblyMap CSR layout")
        _require(len({state.data_ptr(), res.data_ptr(), jac.vals.data_ptr()}) == 3,
                 "Assembly state, residual, and matrix values must not alias")
    if op.nu_e is not None:
        _check_dev_vec(op.nu_e, geom.num_elems, torch.float64, "Per-element viscosity")
    p = _plan(amap, geom.device)
    rp, _ = amap.pattern.device_arrays(geom.device)
    out = (ptr(rp), ptr(jac.vals) if jac is not None else None, ptr(res), stream_handle(stream))
    head = (geom.ndof, geom.num_elems, p.num_blocks, p.max_block_nnz, ptr(geom.xv), ptr(geom.yv),
            ptr(geom.vconn3), ptr(geom.dofs15), ptr(state), op.nu, ptr(op.nu_e))
    if rounds and p.rt_off is not None:
        _lib.call("fx_th_assemble_rounds", *head, ptr(p.blk_row), ptr(p.blk_flags),
                  ptr(p.blk_lanes), ptr(p.rt_off), ptr(p.rt_elem), ptr(p.rt_meta), *out)
    else:
        _lib.call("fx_th_assemble", *head, ptr(p.inc_off), ptr(p.inc_elem), ptr(p.inc_meta),
                  ptr(p.blk_row), ptr(p.blk_flags), ptr(p.blk_lanes), *out)
    return res


def ns_vjp(op: TaylorHoodNS, geom: THGeometry, amap: AssemblyMap, state, lam, bc_mask=None,
           jt_out=None, g_out=None, stream=None):
    """(J^T lam, per-element dR/dnu_e^T lam) over the row-replaced Jacobian."""
    _check(geom, amap)
    dev = geom.device
    _check_dev_vec(state, geom.ndof, torch.float64, "VJP state")
    _check_dev_vec(lam, geom.ndof, torch.float64, "VJP adjoint")
    if jt_out is None:
        jt_out = torch.empty(geom.ndof, dtype=torch.float64, device=dev)
    if g_out is None:
        g_out = torch.empty(geom.num_elems, dtype=torch.float64, device=dev)
    pp = _patches(geom, amap, dev)
    _lib.call("fx_th_vjp_patch", geom.ndof, geom.num_elems, ptr(geom.xv), ptr(geom.yv),
              ptr(state), op.nu, ptr(op.nu_e), ptr(lam),
              ptr(bc_mask), C.byref(pp.plan), ptr(jt_out), ptr(g_out), stream_handle(stream))
    return jt_out, g_out


def cavity_bc(geom: THGeometry, lid=1.0):
    """Lid-driven cavity velocity Dirichlet DOFs in ascending node order
    (DirichletBC::addBoundary, femx/fem/DirichletBC.cpp:138-145): both
    components on every boundary P2 node; u_x = lid on y = 1 (corners
    included), 0 elsewhere."""
    x, y = geom.h_x2, geom.h_y2
    nodes = unit_square_edge_nodes(x, y)
    dofs = np.stack([2 * nodes, 2 * nodes + 1], axis=1).reshape(-1).astype(np.int32)
    vals = np.zeros(dofs.size)
    vals[0::2] = np.where(np.abs(y[nodes] - 1.0) <= 1e-12, lid, 0.0)
    return dofs, vals


def fourier_state(geom: THGeometry, seed, lid=1.0):
    """Config-3 state: each of u_x, u_y, p = sum over the 16 modes
    (kx, ky) in {1..4}^2 of a_k/|k|^2 sin(pi kx x) sin(pi ky y), a_k ~ N(0,1)
    (numpy Philox seeded with `seed`, drawn in the order u_x, u_y, p), then
    u_x = lid on y = 1."""
    rng = np.random.Generator(np.random.Philox(seed))
    kx, ky = np.meshgrid(np.arange(1, 5), np.arange(1, 5), indexing="ij")
    kx, ky = kx.reshape(-1), ky.reshape(-1)
    amp = rng.standard_normal((3, 16)) / (kx ** 2 + ky ** 2)

    def field(a, x, y):
        return (a[None, :] * np.sin(np.pi * kx[None, :] * x[:, None])
                * np.sin(np.pi * ky[None, :] * y[:, None])).sum(axis=1)

    x2, y2 = geom.h_x2, geom.h_y2
    s = np.empty(geom.ndof)
    s[0:2 * geom.num_vel_nodes:2] = field(amp[0], x2, y2)
    s[1:2 * geom.num_vel_nodes:2] = field(amp[1], x2, y2)
    s[2 * geom.num_vel_nodes:] = field(amp[2], geom.h_xv, geom.h_yv)
    top = np.nonzero(np.abs(y2 - 1.0) <= 1e-12)[0]
    s[2 * top] = lid
    return s


class NavierStokesResidual:
    """Residual contract for steady NS with Dirichlet rows replaced and the
    per-element viscosity nu_e as the control."""

    def __init__(self, geom: THGeometry, bc_dofs=None, bc_vals=None, amap=None):
        self.geom = geom
        self.map = amap if amap is not None else make_assembly_map(geom.ndof, geom.ndof,
                                                                   geom.h_dofs15)
        self.num_states = self.num_res = geom.ndof
        self.num_params = geom.num_elems
        self.bmap = (make_boundary_map(bc_dofs, self.map.pattern, geom.device)
                     if bc_dofs is not None else None)
        if self.bmap is not None:
            self.bc_vals = (torch.zeros(self.bmap.num_bcs, dtype=torch.float64, device=geom.device)
                            if bc_vals is None else bc_vals)

    def res(self, state, nu_e, out, stream=None):
        assemble_ns(TaylorHoodNS(0.0, nu_e), self.geom, self.map, state, out, None, stream)
        if self.bmap is not None:
            replace_res(self.bmap, state, self.bc_vals, out, stream)
        return out

    def assemble_state_jac(self, state, nu_e, out: CsrMatrix, stream=None):
        tmp = torch.empty(self.num_res, dtype=torch.float64, device=self.geom.device)
        assemble_ns(TaylorHoodNS(0.0, nu_e), self.geom, self.map, state, tmp, out, stream)
        if self.bmap is not None:
            replace_rows(self.bmap, out, 1.0, stream)
        return out

    def vjp(self, state, nu_e, adj, jt_out=None, g_out=None, stream=None):
        mask = self.bmap.d_mask if self.bmap is not None else None
        return ns_vjp(TaylorHoodNS(0.0, nu_e), self.geom, self.map, state, adj, mask, jt_out,
                      g_out, stream)

    def apply_param_jac_t(self, state, nu_e, adj, out=None, stream=None):
        return self.vjp(state, nu_e, adj, g_out=out, stream=stream)[1]


class NSResidualVJP(torch.autograd.Function):
    """R = residual(u, nu_e) with backward lam -> (J^T lam, (dR/dnu_e)^T lam)."""

    @staticmethod
    def forward(ctx, u, nu_e, residual):
        u, nu_e = u.contiguous(), nu_e.contiguous()
        out = torch.empty(residual.num_res, dtype=torch.float64, device=u.device)
        residual.res(u, nu_e, out)
        ctx.residual = residual
        ctx.save_for_backward(u, nu_e)
        return out

    @staticmethod
    def backward(ctx, lam):
        u, nu_e = ctx.saved_tensors
        jt, g = ctx.residual.vjp(u, nu_e, lam.contiguous())
        return jt, g, None
