"""GLS-stabilised P1/P1 Navier-Stokes time step on the HIP kernels.

femx's Navier-Stokes model (femx/model/ns/NavierStokesModel.cpp) discretises
velocity and pressure with P1 on triangles, stabilises with GLS and steps in
time with an extrapolated advection velocity.  Its row operator is
`NavierOperator` (femx/model/ns/Components.hpp:400-577) and its host and
device time assembly is `assembly::assemble(op, step, num_hist, wrt, map,
hist, nxt, res, jac, ctx)` (femx/assembly/Assembly.hpp:235-297,
CudaAssembly.hpp:366-417).  This module mirrors those names:

    geom = GLSGeometry(x, y, conn)                # the mesh (device copies)
    amap = make_navier_map(geom)                  # DofLayout(MixedFESpace) map
    op   = NavierOperator(geom, FluidParams(rho, mu), dt)
    assemble(op, step, num_hist, NEXT_STATE, amap, hist, nxt, res, jac)

The element values (N, dNdx, JxW) that femx streams from HBM are recomputed
from the vertex coordinates inside the kernel, and the results equal femx's
host assembly bit for bit.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .assembly import AssemblyMap, make_assembly_map
from .linalg import CsrMatrix, _check_dev_vec, _require, ptr, stream_handle
from .patches import Patches

BLOCK_NODES = 32  # FX_GLS_BLOCK_NODES
MAX_RUNS = 4  # FX_GLS_MAX_RUNS

#: state::VariableBlock::NextState (femx/state/TimeResidual.hpp:96-165)
NEXT_STATE = ("next", 0)


def history_state(lag):
    """state::VariableBlock::historyState(lag)."""
    return ("history", int(lag))


#: state::VariableBlock::Param
PARAM = ("param", 0)


@dataclass(frozen=True)
class FluidParams:
    """femx::model::ns::FluidParams (femx/model/ns/Config.hpp:10-14)."""

    rho: float = 1.0
    mu: float = 1.0


def navier_dofs(num_nodes, conn):
    """Element DOF table of femx's mixed P1/P1 space: the vector velocity
    field first (global DOF comps*node + c, femx/fem/FESpace.cpp:96-106), then
    the pressure field offset by the velocity DOFs (femx/fem/MixedFESpace.cpp:
    57-84,121-137), in NavierOperator's local order u0 v0 u1 v1 u2 v2 p0 p1 p2
    (femx/model/ns/Components.hpp:247-256).  Returns int32 [E, 9]."""
    c = np.ascontiguousarray(conn, np.int64).reshape(-1, 3)
    d = np.empty((c.shape[0], 9), np.int64)
    d[:, 0:6:2] = 2 * c
    d[:, 1:6:2] = 2 * c + 1
    d[:, 6:9] = 2 * int(num_nodes) + c
    return d.astype(np.int32)


class GLSGeometry:
    """P1 triangle mesh (node coordinates, 3 node ids per element) with its
    mixed velocity-pressure DOF table and device copies."""

    def __init__(self, x, y, conn, device="cuda"):
        self.device = torch.device(device)
        self.h_x = np.ascontiguousarray(x, np.float64)
        self.h_y = np.ascontiguousarray(y, np.float64)
        self.h_conn = np.ascontiguousarray(conn, np.int32).reshape(-1, 3)
        _require(self.h_x.size == self.h_y.size, "node coordinate arrays differ in length")
        self.num_nodes = self.h_x.size
        self.num_elems = self.h_conn.shape[0]
        _require(self.num_elems == 0 or (self.h_conn.min() >= 0
                                         and self.h_conn.max() < self.num_nodes),
                 "element node id out of range")
        self.num_states = 3 * self.num_nodes
        self.h_dofs9 = navier_dofs(self.num_nodes, self.h_conn)
        t = lambda a: torch.from_numpy(a.reshape(-1)).to(self.device)  # noqa: E731
        self.x, self.y, self.conn = t(self.h_x), t(self.h_y), t(self.h_conn)
        self.dofs9 = t(self.h_dofs9)


def make_navier_map(geom: GLSGeometry) -> AssemblyMap:
    """makeAssemblyMap(DofLayout(space)) of the model's mixed space."""
    return make_assembly_map(geom.num_states, geom.num_states, geom.h_dofs9)


class GLSPlan:
    """Node incidence, CSR positions and LDS row-tile runs of the GLS kernel
    (fx_map_gls_plan) for one map and mesh, on one device."""

    def __init__(self, amap: AssemblyMap, geom: GLSGeometry, device="cuda"):
        dev = torch.device(device)
        N, E = geom.num_nodes, geom.num_elems
        nb = (N + BLOCK_NODES - 1) // BLOCK_NODES
        self.num_blocks = nb
        h = dict(node_rows=np.empty(max(3 * N, 1), np.int32),
                 inc_off=np.empty(N + 1, np.int32),
                 inc_elem=np.empty(max(3 * E, 1), np.int32),
                 inc_meta=np.empty(max(24 * E, 8), np.int32),
                 inc_slot=np.empty(max(3 * E, 1), np.int32),
                 blk_runs=np.empty(max(4 * MAX_RUNS * nb, 4), np.int32),
                 node_tile=np.empty(max(3 * N, 1), np.int32))
        mx = C.c_int32()
        _lib.call("fx_map_gls_plan", amap._handle(), N, ptr(geom.h_conn), ptr(h["node_rows"]),
                  ptr(h["inc_off"]), ptr(h["inc_elem"]), ptr(h["inc_meta"]), ptr(h["inc_slot"]),
                  ptr(h["blk_runs"]), ptr(h["node_tile"]), C.byref(mx))
        self.max_tile = mx.value
        self.h = h
        for k, v in h.items():
            setattr(self, k, torch.from_numpy(v).to(dev))


class NavierOperator:
    """femx::model::ns::NavierOperator(data, fluid, dt): the GLS row operator
    of one time step on a mesh."""

    def __init__(self, geom: GLSGeometry, fluid: FluidParams, dt: float):
        # NavierStokesModel's parameter checks (NavierStokesModel.cpp:35-46)
        _require(dt > 0.0 and np.isfinite(dt),
                 "NavierStokesModel requires a positive finite time step")
        _require(np.isfinite(fluid.rho) and fluid.rho > 0.0,
                 "NavierStokesModel requires positive finite density")
        _require(np.isfinite(fluid.mu) and fluid.mu > 0.0,
                 "NavierStokesModel requires positive finite viscosity")
        self.geom, self.fluid, self.dt = geom, fluid, float(dt)


def _plan(amap, geom):
    key = "_gls_plan_" + str(geom.device)
    p = getattr(amap, key, None)
    if p is None or p[0] is not geom:
        _require(geom.num_elems == amap.num_elems,
                 "Geometry and AssemblyMap have different element counts")
        _require(amap.num_res == geom.num_states and amap.num_states == geom.num_states
                 and np.array_equal(amap.res_dofs, geom.h_dofs9.reshape(-1))
                 and np.array_equal(amap.state_dofs, geom.h_dofs9.reshape(-1)),
                 "NavierOperator needs the map of the mesh's mixed P1/P1 space")
        p = (geom, GLSPlan(amap, geom, geom.device))
        setattr(amap, key, p)
    return p[1]


def assemble(op: NavierOperator, step, num_hist, wrt, amap: AssemblyMap, hist, nxt, res,
             jac: CsrMatrix = None, stream=None):
    """One time residual and its Jacobian block (femx/assembly/Assembly.hpp:
    235-297).  hist is lag-major (num_hist x num_states).  With respect to the
    next state the Jacobian is K; for a history block NavierOperator::evalRow
    leaves the row Jacobian zero (Components.hpp:412-435), as here.  Outputs
    are replaced; returns res."""
    geom = op.geom
    kind, lag = wrt
    # checkTimeAssemblyInputs (Assembly.hpp:86-104)
    _require(num_hist > 0 and hist.numel() == num_hist * amap.num_states
             and nxt.numel() == amap.num_states,
             "Time assembly state dimensions do not match AssemblyMap")
    _require(kind == "next" or (kind == "history" and 0 <= lag < num_hist),
             "Time assembly variable block is invalid")
    _check_dev_vec(hist, num_hist * geom.num_states, torch.float64, "Time assembly history")
    _check_dev_vec(nxt, geom.num_states, torch.float64, "Time assembly next state")
    _check_dev_vec(res, geom.num_states, torch.float64, "Time assembly residual")
    vals = None
    if jac is not None:
        _require(jac.pattern.layout_id == amap.pattern.layout_id,
                 "Time assembly matrix must use the AssemblyMap CSR layout")
        vals = jac.vals
    p = _plan(amap, geom)
    _lib.call("fx_gls_assemble", geom.num_nodes, geom.num_states, geom.num_elems,
              amap.pattern.nnz, p.max_tile, ptr(geom.x), ptr(geom.y), ptr(geom.conn),
              ptr(geom.dofs9), ptr(p.node_rows), op.fluid.rho, op.fluid.mu, op.dt, int(step), int(num_hist),
              ptr(hist), ptr(nxt), ptr(p.inc_off), ptr(p.inc_elem), ptr(p.inc_meta),
              ptr(p.blk_runs), ptr(p.node_tile), ptr(vals) if kind == "next" else None,
              ptr(res), stream_handle(stream))
    if vals is not None and kind != "next":
        # a history-lag block of the Jacobian is zero; cleared on the launch
        # stream so that later work queued there sees it
        with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream()):
            vals.zero_()
    return res


def hist_vjp_two_phase(op: NavierOperator, step, num_hist, lag, amap: AssemblyMap, hist, nxt,
                       adj, out=None, stream=None):
    """hist_vjp through fx_gls_hist_vjp: element columns go to the plan's
    workspace (6 doubles per element, allocated on first use), then are
    summed per node in ascending element order."""
    geom = op.geom
    _require(num_hist > 0 and 0 <= lag < num_hist, "Time assembly variable block is invalid")
    _check_dev_vec(hist, num_hist * geom.num_states, torch.float64, "History VJP history")
    _check_dev_vec(nxt, geom.num_states, torch.float64, "History VJP next state")
    _check_dev_vec(adj, geom.num_states, torch.float64, "History VJP adjoint")
    if out is None:
        out = torch.empty(geom.num_states, dtype=torch.float64, device=geom.device)
    _check_dev_vec(out, geom.num_states, torch.float64, "History VJP output")
    p = _plan(amap, geom)
    if getattr(p, "vjp_work", None) is None:
        p.vjp_work = torch.empty(max(6 * geom.num_elems, 1), dtype=torch.float64,
                                 device=geom.device)
    work = p.vjp_work
    _lib.call("fx_gls_hist_vjp", geom.num_nodes, geom.num_states, geom.num_elems, ptr(geom.x),
              ptr(geom.y), ptr(geom.conn), ptr(geom.dofs9), ptr(p.node_rows), op.fluid.rho,
              op.fluid.mu, op.dt, int(step), int(num_hist), int(lag), ptr(hist), ptr(nxt),
              ptr(adj), ptr(p.inc_off), ptr(p.inc_slot), ptr(out), ptr(work),
              stream_handle(stream))
    return out


def _patches(amap, geom):
    """The history VJP's patch plan (mesh nodes as DOFs, four doubles per
    seam partial), cached with the map's GLS plan."""
    p = _plan(amap, geom)
    if getattr(p, "vjp_patches", None) is None:
        p.vjp_patches = Patches(geom.num_nodes, geom.num_elems, 3, geom.h_conn, geom.h_conn,
                                geom.h_x, geom.h_y, geom.device, part_width=4)
    return p.vjp_patches


def hist_vjp_lags(op: NavierOperator, step, num_hist, amap: AssemblyMap, hist, nxt, adj,
                  out_lag0=None, out_lag1=None, stream=None):
    """Both history lags' VJPs, (dR/dhist_0)^T adj and (dR/dhist_1)^T adj, in
    one pass over element patches (fx_gls_hist_vjp_patch): the element
    partials are evaluated once for both.  Lag 1 needs num_hist > 1; pass
    out_lag1=False to skip it.  Returns (out_lag0, out_lag1).  Calls sharing
    a map share its partials workspace: they must not overlap on different
    streams."""
    geom = op.geom
    _require(num_hist > 0, "Time assembly variable block is invalid")
    for v, n, what in ((hist, num_hist * geom.num_states, "History VJP history"),
                       (nxt, geom.num_states, "History VJP next state"),
                       (adj, geom.num_states, "History VJP adjoint")):
        _check_dev_vec(v, n, torch.float64, what)
    want1 = out_lag1 is not False and num_hist > 1
    outs = []
    for o, want in ((out_lag0, out_lag0 is not False), (out_lag1 if want1 else None, want1)):
        if want and (o is None or o is True):
            o = torch.empty(geom.num_states, dtype=torch.float64, device=geom.device)
        if want:
            _check_dev_vec(o, geom.num_states, torch.float64, "History VJP output")
        outs.append(o if want else None)
    pp = _patches(amap, geom)
    p = _plan(amap, geom)  # every state row belongs to a mesh node (fx_map_gls_plan)
    _lib.call("fx_gls_hist_vjp_patch", geom.num_nodes, geom.num_states, geom.num_elems,
              ptr(geom.x), ptr(geom.y), ptr(p.node_rows), op.fluid.rho, op.fluid.mu, op.dt,
              int(step), int(num_hist), ptr(hist), ptr(nxt), ptr(adj), C.byref(pp.plan),
              ptr(outs[0]), ptr(outs[1]), stream_handle(stream))
    return outs[0], outs[1]


def hist_vjp(op: NavierOperator, step, num_hist, lag, amap: AssemblyMap, hist, nxt, adj,
             out=None, stream=None):
    """(dR/dhist_lag)^T adj: TimeResidual::applyJacT with respect to a history
    block (femx/model/ns/NavierStokesModel.cpp:321-363), the product femx
    takes with Enzyme, here hand-derived.  Returns out (allocated when None).
    Lags 0 and 1 go through the single-pass patch VJP (hist_vjp_lags with the
    other lag skipped); an older lag does not enter the operator and gets
    zeros."""
    geom = op.geom
    _require(num_hist > 0 and 0 <= lag < num_hist, "Time assembly variable block is invalid")
    if lag >= 2:
        if out is None:
            out = torch.empty(geom.num_states, dtype=torch.float64, device=geom.device)
        _check_dev_vec(out, geom.num_states, torch.float64, "History VJP output")
        with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream()):
            out.zero_()
        return out
    if lag == 0:
        return hist_vjp_lags(op, step, num_hist, amap, hist, nxt, adj, out, False, stream)[0]
    return hist_vjp_lags(op, step, num_hist, amap, hist, nxt, adj, False, out, stream)[1]


def assemble_residual(op: NavierOperator, step, num_hist, amap: AssemblyMap, hist, nxt, res,
                      stream=None):
    """assembleResidual (femx/assembly/CudaAssembly.hpp:419-472): the time
    residual only."""
    return assemble(op, step, num_hist, NEXT_STATE, amap, hist, nxt, res, None, stream)


class NavierResidual:
    """The parameter-free time residual of femx's Navier-Stokes model
    (NavierResidual, femx/model/ns/NavierStokesModel.cpp:447-545) on this
    mesh: two history lags, res / assemble_next through the GLS kernel,
    apply_jac_t with respect to a history lag through the history VJP.
    `time` is a femx_amd.control.TimeContext."""

    NUM_HIST = 2

    def __init__(self, geom: GLSGeometry, fluid: FluidParams, dt: float, num_steps: int,
                 amap: AssemblyMap = None):
        _require(num_steps > 0, "NavierStokesModel requires a positive number of time steps")
        self.geom, self.num_steps = geom, int(num_steps)
        self.op = NavierOperator(geom, fluid, dt)
        self.map = amap if amap is not None else make_navier_map(geom)

    def dims(self):
        ns_ = self.geom.num_states
        return dict(num_steps=self.num_steps, num_states=ns_, num_param=0, num_res=ns_,
                    num_hist=self.NUM_HIST)

    @property
    def pattern(self):
        return self.map.pattern

    def _check(self, time):
        _require(0 <= time.step < self.num_steps, "Navier residual step is out of range")
        ns_ = self.geom.num_states
        _require(time.num_hist >= self.NUM_HIST and time.hist.numel() == time.num_hist * ns_
                 and time.nxt.numel() == ns_ and time.prm.numel() == 0,
                 "Navier residual vector size mismatch")

    def _hist(self, time):
        return time.hist[: self.NUM_HIST * self.geom.num_states]

    def res(self, time, out, stream=None):
        self._check(time)
        return assemble(self.op, time.step, self.NUM_HIST, NEXT_STATE, self.map, self._hist(time),
                        time.nxt, out, None, stream)

    def assemble_next(self, time, res, jac: CsrMatrix, stream=None):
        self._check(time)
        return assemble(self.op, time.step, self.NUM_HIST, NEXT_STATE, self.map, self._hist(time),
                        time.nxt, res, jac, stream)

    def apply_jac_t(self, time, wrt, adj, out, stream=None):
        self._check(time)
        kind, lag = wrt
        _require(kind != "next",
                 "Navier transpose apply supports only history and parameter blocks")
        _require(adj.numel() == self.geom.num_states, "Navier residual adjoint size mismatch")
        if kind == "param":
            return out[:0]
        _require(0 <= lag < self.NUM_HIST, "Navier residual history lag is out of range")
        return hist_vjp(self.op, time.step, self.NUM_HIST, lag, self.map, self._hist(time),
                        time.nxt, adj, out, stream)

    def apply_jac_t_lags(self, time, adj, out_lag0=None, out_lag1=None, stream=None):
        """apply_jac_t for both history lags at once (one pass over the
        mesh, hist_vjp_lags): (dR/dhist_0)^T adj, (dR/dhist_1)^T adj."""
        self._check(time)
        _require(adj.numel() == self.geom.num_states, "Navier residual adjoint size mismatch")
        return hist_vjp_lags(self.op, time.step, self.NUM_HIST, self.map, self._hist(time),
                             time.nxt, adj, out_lag0, out_lag1, stream)

    def prepare_linear_solve(self, time, jac, rhs, stream=None):
        """TimeResidual's default: nothing to do (TimeResidual.hpp:229-238)."""
        return None
