"""Dirichlet control and the constrained time residual on the device.

femx prescribes Dirichlet data of a time-dependent problem through a
`DirichletControl` (controlled state DOFs u_D = P q, femx/fem/
DirichletControl.cpp:21-209) and a `ControlMap` (femx/fem/ControlMap.cpp:
36-376: the controlled DOFs, fixed DOFs with per-step values, and a linear
time interpolation of the control levels).  `ConstrainedTimeResidual`
(femx/assembly/ConstrainedTimeResidual.cpp:318-405) decorates a physics
residual with them, and `TimeIntegrator::solveStep` (femx/state/
TimeIntegrator.hpp:272-289) forms the linear system of one step.  This module
mirrors those names over the HIP kernels: every device operation is a C-ABI
call (CSR matvec/matvecT, gather/scatter, axpby, the boundary operations and
the GLS kernels), in femx's order, so the results equal femx's host code bit
for bit.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .assembly import (make_boundary_map, prepare_forward_solve, replace_res, replace_rows,
                       zero_boundary)
from .linalg import CsrMatrix, CsrPattern, MatrixHandler, VectorHandler, _check_dev_vec, _require, ptr
from .linalg import stream_handle


@dataclass(frozen=True)
class LinearInterpolation:
    """femx::LinearInterpolation (femx/common/LinearInterpolation.hpp)."""

    lower: int = 0
    upper: int = 0
    upper_weight: float = 0.0

    def has_upper(self):
        return self.upper != self.lower and self.upper_weight != 0.0

    def is_valid(self):
        w = self.upper_weight
        return (self.lower >= 0 and self.upper >= self.lower and 0.0 <= w <= 1.0
                and np.isfinite(w) and (self.upper != self.lower or w == 0.0))


class DirichletControl:
    """u_D = P q on selected state DOFs.  `entries` are (state_row, ctr_col,
    weight); P is stored as CSR with each row's entries in input order, as
    makeControlMatrix builds it (DirichletControl.cpp:32-73)."""

    def __init__(self, state_dofs, num_ctr_params=None, entries=None):
        dofs = np.asarray(state_dofs, np.int64).reshape(-1)
        if num_ctr_params is None:  # the identity control of DirichletControl(dofs)
            num_ctr_params = dofs.size
            entries = [(i, i, 1.0) for i in range(dofs.size)]
        _require(num_ctr_params >= 0, "DirichletControl received negative control parameter count")
        rows = dofs.size
        seen = set()
        count = np.zeros(rows + 1, np.int64)
        for r, c, w in entries:
            _require(0 <= r < rows, "DirichletControl map state row is out of range")
            _require(0 <= c < num_ctr_params, "DirichletControl map control column is out of range")
            _require(np.isfinite(w), "DirichletControl received non-finite map weight")
            _require((r, c) not in seen, "DirichletControl received duplicate map entry")
            seen.add((r, c))
            count[r + 1] += 1
        row_ptr = np.cumsum(count)
        nxt = row_ptr[:-1].copy()
        col = np.empty(len(entries), np.int64)
        val = np.empty(len(entries), np.float64)
        for r, c, w in entries:
            k = nxt[r]
            nxt[r] += 1
            col[k], val[k] = c, w
        _require(np.all(dofs >= 0), "DirichletControl received negative state id")
        _require(np.unique(dofs).size == dofs.size, "DirichletControl received duplicate state id")
        self.dofs = dofs.astype(np.int32)
        self.num_ctr_params = int(num_ctr_params)
        self.pattern = CsrPattern(rows, self.num_ctr_params, row_ptr, col)
        self.vals = val


class ControlMap:
    """Fixed and controlled Dirichlet data (makeControlMap, ControlMap.cpp:
    36-137), with its device copies.  `dofs` lists the controlled DOFs then
    the fixed ones; `fixed_vals` has one row of values per step."""

    def __init__(self, num_steps, num_states, ctr: DirichletControl, fixed_dofs=(),
                 fixed_vals=None, time=None, ctr_off=0, num_prm=-1, device="cuda"):
        _require(num_steps > 0 and num_states > 0 and ctr_off >= 0,
                 "ControlMap received invalid dimensions")
        used = np.zeros(num_states, bool)
        for d in ctr.dofs:
            _require(0 <= d < num_states and not used[d], "ControlMap controlled DOF is invalid")
            used[d] = True
        fixed = np.asarray(fixed_dofs, np.int64).reshape(-1)
        for d in fixed:
            _require(0 <= d < num_states and not used[d], "ControlMap fixed DOF is invalid")
            used[d] = True
        if not time:
            time = [LinearInterpolation(s, s, 0.0) for s in range(num_steps)]
        time = [t if isinstance(t, LinearInterpolation)
                else LinearInterpolation(int(t[0]), int(t[1]), float(t[2])) for t in time]
        _require(len(time) == num_steps, "ControlMap time stencil count mismatch")
        for t in time:
            _require(t.is_valid(), "ControlMap time stencil is invalid")
        num_levels = max(t.upper + 1 for t in time)
        required = ctr_off + num_levels * ctr.num_ctr_params
        if num_prm < 0:
            num_prm = required
        _require(num_prm >= required, "ControlMap parameter count is too small")
        nf = fixed.size
        if fixed_vals is None or np.size(fixed_vals) == 0:
            fv = np.zeros(num_steps * nf)
        elif np.size(fixed_vals) == nf:
            fv = np.tile(np.asarray(fixed_vals, np.float64).reshape(-1), num_steps)
        else:
            _require(np.size(fixed_vals) == num_steps * nf, "ControlMap fixed value size mismatch")
            fv = np.asarray(fixed_vals, np.float64).reshape(-1).copy()
        self.num_steps, self.num_states, self.num_prm = int(num_steps), int(num_states), int(num_prm)
        self.num_fixed, self.ctr_off = int(nf), int(ctr_off)
        self.time = time
        self.dofs = np.concatenate([ctr.dofs, fixed]).astype(np.int32)
        self.num_bcs = self.dofs.size
        self.num_controlled = ctr.dofs.size
        self.block = ctr.num_ctr_params
        self.device = torch.device(device)
        self.pattern, self.p_vals, self.fixed_vals = ctr.pattern, ctr.vals, fv
        self._d = None

    def _dev(self):
        """Device copies (P as a CSR matrix, DOF list, fixed values, scratch),
        made on first use."""
        if self._d is None:
            dev = self.device
            P = CsrMatrix(self.pattern, dev, torch.from_numpy(self.p_vals).to(dev))
            t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
            self._d = dict(P=P, dofs=t(self.dofs), fixed_vals=t(self.fixed_vals),
                           compact=torch.empty(self.num_controlled, dtype=torch.float64,
                                               device=dev))
        return self._d

    def _step(self, step):
        _require(0 <= step < self.num_steps, "ControlMap vector size mismatch")
        t = self.time[step]
        return t.lower, t.upper, t.upper_weight, 1.0 - t.upper_weight

    def _prm_block(self, prm, level):
        off = self.ctr_off + level * self.block
        return prm[off:off + self.block]


def control_vals(cm: ControlMap, step, prm, out, stream=None):
    """Prescribed values of every constrained DOF at one step (controlVals,
    ControlMap.cpp:165-199): the interpolated P q, then the fixed values."""
    _check_dev_vec(prm, cm.num_prm, torch.float64, "ControlMap parameters")
    _check_dev_vec(out, cm.num_bcs, torch.float64, "ControlMap values")
    lo, hi, hi_wt, lo_wt = cm._step(step)
    d = cm._dev()
    mh = MatrixHandler(stream)
    ctl = out[:cm.num_controlled]
    mh.matvec(d["P"], cm._prm_block(prm, lo), ctl, lo_wt, 0.0)
    if hi != lo and hi_wt != 0.0:
        mh.matvec(d["P"], cm._prm_block(prm, hi), ctl, hi_wt, 1.0)
    if cm.num_fixed:
        src = d["fixed_vals"][step * cm.num_fixed:(step + 1) * cm.num_fixed]
        _lib.call("fx_vec_axpby", cm.num_fixed, 1.0, ptr(src), 0.0, ptr(out[cm.num_controlled:]),
                  stream_handle(stream))
    return out


def control_jac(cm: ControlMap, step, dirv, out, stream=None):
    """d(residual)/d(parameters) applied to a direction (controlJac,
    ControlMap.cpp:237-271): -P (interpolated direction) on the controlled
    DOFs, zero elsewhere."""
    _check_dev_vec(dirv, cm.num_prm, torch.float64, "ControlMap direction")
    _check_dev_vec(out, cm.num_states, torch.float64, "ControlMap Jacobian output")
    lo, hi, hi_wt, lo_wt = cm._step(step)
    d = cm._dev()
    out.zero_()
    mh = MatrixHandler(stream)
    mh.matvec(d["P"], cm._prm_block(dirv, lo), d["compact"], -lo_wt, 0.0)
    if hi != lo and hi_wt != 0.0:
        mh.matvec(d["P"], cm._prm_block(dirv, hi), d["compact"], -hi_wt, 1.0)
    _lib.call("fx_vec_scatter", cm.num_controlled, ptr(d["dofs"]), ptr(d["compact"]), ptr(out),
              cm.num_states, stream_handle(stream))
    return out


def add_control_jac_t(cm: ControlMap, step, adj, grad, stream=None):
    """grad += (d residual / d parameters)^T adj (addControlJacT,
    ControlMap.cpp:309-342)."""
    _check_dev_vec(adj, cm.num_states, torch.float64, "ControlMap adjoint")
    _check_dev_vec(grad, cm.num_prm, torch.float64, "ControlMap gradient")
    lo, hi, hi_wt, lo_wt = cm._step(step)
    d = cm._dev()
    _lib.call("fx_vec_gather", cm.num_controlled, ptr(d["dofs"]), ptr(adj), cm.num_states,
              ptr(d["compact"]), stream_handle(stream))
    mh = MatrixHandler(stream)
    mh.matvecT(d["P"], d["compact"], cm._prm_block(grad, lo), -lo_wt, 1.0)
    if hi != lo and hi_wt != 0.0:
        mh.matvecT(d["P"], d["compact"], cm._prm_block(grad, hi), -hi_wt, 1.0)
    return grad


@dataclass
class TimeContext:
    """state::TimeContext (femx/state/TimeResidual.hpp:80-86): the step, the
    next state, the parameters and the lag-major history window."""

    step: int
    nxt: torch.Tensor
    prm: torch.Tensor
    hist: torch.Tensor
    num_hist: int


class ConstrainedTimeResidual:
    """femx's ConstrainedTimeResidual over a parameter-free base residual
    (`femx_amd.gls.NavierResidual`): constrained rows of the residual become
    nxt - u_D, of the Jacobian identity rows, and the adjoint is zeroed there
    before history transposes (ConstrainedTimeResidual.cpp:318-405)."""

    def __init__(self, base, control: ControlMap):
        d = base.dims()
        _require(d["num_res"] == d["num_states"],
                 "ConstrainedTimeResidual requires square state residuals")
        _require(d["num_param"] == 0,
                 "ConstrainedTimeResidual requires a parameter-free base residual")
        _require(control.num_steps == d["num_steps"] and control.num_states == d["num_states"],
                 "ConstrainedTimeResidual control dimensions do not match")
        self.base, self.control = base, control
        self._dims = dict(d, num_param=control.num_prm)
        self.boundary = make_boundary_map(control.dofs, base.pattern, control.device)
        dev = control.device
        self.bc_vals = torch.empty(control.num_bcs, dtype=torch.float64, device=dev)
        self.base_adj = torch.empty(d["num_res"], dtype=torch.float64, device=dev)
        self.base_prm = torch.empty(d["num_param"], dtype=torch.float64, device=dev)

    def dims(self):
        return dict(self._dims)

    @property
    def pattern(self):
        return self.base.pattern

    def _check(self, time: TimeContext):
        d = self._dims
        _require(0 <= time.step < d["num_steps"] and time.num_hist >= d["num_hist"]
                 and time.hist.numel() == time.num_hist * d["num_states"]
                 and time.nxt.numel() == d["num_states"] and time.prm.numel() == d["num_param"],
                 "ConstrainedTimeResidual context dimensions do not match")

    def _base(self, time: TimeContext):
        """baseCtx (ConstrainedTimeResidual.cpp:407-413): the base residual sees
        its own (empty) parameter vector."""
        return TimeContext(time.step, time.nxt, self.base_prm, time.hist, time.num_hist)

    def res(self, time: TimeContext, out, stream=None):
        self._check(time)
        self.base.res(self._base(time), out, stream)
        control_vals(self.control, time.step, time.prm, self.bc_vals, stream)
        replace_res(self.boundary, time.nxt, self.bc_vals, out, stream)
        return out

    def assemble_next(self, time: TimeContext, res, jac: CsrMatrix, stream=None):
        self._check(time)
        self.base.assemble_next(self._base(time), res, jac, stream)
        control_vals(self.control, time.step, time.prm, self.bc_vals, stream)
        replace_res(self.boundary, time.nxt, self.bc_vals, res, stream)
        replace_rows(self.boundary, jac, 1.0, stream)
        return res

    def apply_jac_t(self, time: TimeContext, wrt, adj, out, stream=None):
        """wrt is ("history", lag) or ("param", 0)."""
        self._check(time)
        kind, lag = wrt
        _require(kind != "next",
                 "Constrained transpose apply supports only history and parameter blocks")
        _require(adj.numel() == self._dims["num_res"], "ConstrainedTimeResidual adjoint size mismatch")
        if kind == "param":
            _check_dev_vec(out, self._dims["num_param"], torch.float64, "Parameter gradient")
            out.zero_()
            return add_control_jac_t(self.control, time.step, adj, out, stream)
        _lib.call("fx_vec_axpby", adj.numel(), 1.0, ptr(adj), 0.0, ptr(self.base_adj),
                  stream_handle(stream))
        zero_boundary(self.boundary, self.base_adj, stream)
        return self.base.apply_jac_t(self._base(time), wrt, self.base_adj, out, stream)

    def prepare_linear_solve(self, time: TimeContext, jac: CsrMatrix, rhs, stream=None):
        self._check(time)
        self.base.prepare_linear_solve(self._base(time), jac, rhs, stream)
        control_vals(self.control, time.step, time.prm, self.bc_vals, stream)
        prepare_forward_solve(self.boundary, jac, rhs, self.bc_vals, stream)


def step_system(residual, time: TimeContext, res, jac: CsrMatrix, rhs, stream=None):
    """The assembly half of TimeIntegrator::solveStep (TimeIntegrator.hpp:
    272-289): assembleNext, rhs = J nxt - res, prepareLinearSolve.  The
    linear solve and the history shift are the caller's."""
    residual.assemble_next(time, res, jac, stream)
    MatrixHandler(stream).matvec(jac, time.nxt, rhs)
    VectorHandler(stream).axpby(-1.0, res, 1.0, rhs)
    residual.prepare_linear_solve(time, jac, rhs, stream)
    return rhs
