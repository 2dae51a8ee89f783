"""Device linear solves and the reduced-functional gradient loop.

femx's adjoint gradient (`ReducedFunctional::gradAt`, femx/inverse/
ReducedFunctional.hpp:94-117) needs a `LinearSolver` with `solve` and
`solveT` (femx/linalg/LinearSolver.hpp:10-48); femx's device solver is
ReSolve.  `BiCGStabSolver` is a small deterministic replacement built on this
library's kernels only -- CSR SpMV, the fixed-order dot product, axpby and a
Jacobi scaling -- so the forward solve, the adjoint solve with the transposed
Jacobian (through the cached explicit transpose, MatrixHandler::transpose)
and the parameter VJP all run on the GPU.  It is a correctness-grade solver
for the gradient loop, not a tuned preconditioned solver.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .linalg import CsrMatrix, MatrixHandler, VectorHandler, _check_dev_vec, _require, ptr
from .linalg import stream_handle


def _diag_positions(pattern):
    """CSR position of each row's diagonal, cached on the pattern."""
    if "diag" not in pattern._tmap:
        rp, ci = pattern.row_ptr, pattern.col_ind
        pos = np.full(pattern.rows, -1, np.int64)
        for r in range(pattern.rows):
            hit = np.nonzero(ci[rp[r]:rp[r + 1]] == r)[0]
            if hit.size:
                pos[r] = rp[r] + hit[0]
        _require(pattern.rows == pattern.cols and np.all(pos >= 0),
                 "BiCGStab's Jacobi scaling needs a diagonal entry in every row")
        pattern._tmap["diag"] = pos.astype(np.int32)
    return pattern._tmap["diag"]


class BiCGStabSolver:
    """Right-preconditioned (Jacobi) BiCGStab, every reduction in the fixed
    order of fx_vec_dot, so repeated solves are bitwise identical.  `sol`
    holds the initial guess on entry.  Stops when ||r|| <= tol ||b||."""

    def __init__(self, tol=1e-12, max_iter=5000, stream=None):
        self.tol, self.max_iter, self.stream = float(tol), int(max_iter), stream
        self.mh, self.vh = MatrixHandler(stream), VectorHandler(stream)
        self.iterations = 0
        self.rel_residual = float("nan")
        self._transposes = {}

    def _dot(self, a, b):
        return float(self.vh.dot(a, b).item())

    def _copy(self, src, dst):
        _lib.call("fx_vec_axpby", dst.numel(), 1.0, ptr(src), 0.0, ptr(dst),
                  stream_handle(self.stream))
        return dst

    def _div(self, x, d, z):
        _lib.call("fx_vec_div", x.numel(), ptr(x), ptr(d), ptr(z), stream_handle(self.stream))
        return z

    def solve(self, mat: CsrMatrix, rhs, sol):
        """sol <- mat^-1 rhs (LinearSolver::solve)."""
        n = mat.rows
        _require(mat.rows == mat.cols, "BiCGStab needs a square matrix")
        _check_dev_vec(rhs, n, torch.float64, "Linear solve right-hand side")
        _check_dev_vec(sol, n, torch.float64, "Linear solve solution")
        dev = mat.device
        new = lambda: torch.empty(n, dtype=torch.float64, device=dev)  # noqa: E731
        pos = torch.from_numpy(_diag_positions(mat.pattern)).to(dev)
        d = new()
        _lib.call("fx_vec_gather", n, ptr(pos), ptr(mat.vals), mat.nnz, ptr(d),
                  stream_handle(self.stream))
        r, rhat, p, v, ph, s, sh, t = (new() for _ in range(8))
        self.mh.matvec(mat, sol, r, -1.0, 0.0)
        self.vh.axpby(1.0, rhs, 1.0, r)
        self._copy(r, rhat)
        p.zero_()
        v.zero_()
        bnorm = np.sqrt(self._dot(rhs, rhs))
        self.iterations = 0
        if bnorm == 0.0:
            sol.zero_()
            self.rel_residual = 0.0
            return sol
        rho = alpha = omega = 1.0
        rel = np.sqrt(self._dot(r, r)) / bnorm
        while rel > self.tol:
            _require(self.iterations < self.max_iter, "BiCGStab did not converge")
            self.iterations += 1
            rho_new = self._dot(rhat, r)
            _require(rho_new != 0.0 and omega != 0.0, "BiCGStab broke down")
            beta = (rho_new / rho) * (alpha / omega)
            self.vh.axpby(-omega, v, 1.0, p)
            self.vh.axpby(1.0, r, beta, p)
            self._div(p, d, ph)
            self.mh.matvec(mat, ph, v)
            alpha = rho_new / self._dot(rhat, v)
            self._copy(r, s)
            self.vh.axpby(-alpha, v, 1.0, s)
            if np.sqrt(self._dot(s, s)) / bnorm <= self.tol:
                self.vh.axpby(alpha, ph, 1.0, sol)
                rel = np.sqrt(self._dot(s, s)) / bnorm
                break
            self._div(s, d, sh)
            self.mh.matvec(mat, sh, t)
            tt = self._dot(t, t)
            omega = self._dot(t, s) / tt if tt != 0.0 else 0.0
            self.vh.axpby(alpha, ph, 1.0, sol)
            self.vh.axpby(omega, sh, 1.0, sol)
            self._copy(s, r)
            self.vh.axpby(-omega, t, 1.0, r)
            rel = np.sqrt(self._dot(r, r)) / bnorm
            rho = rho_new
        self.rel_residual = rel
        return sol

    def solveT(self, mat: CsrMatrix, rhs, sol):  # noqa: N802 (femx's name)
        """sol <- mat^-T rhs (LinearSolver::solveT): a solve with the explicit
        transpose, whose pattern and permutation are cached per layout."""
        key = mat.pattern.layout_id
        mt = self._transposes.get(key)
        mt = self.mh.transpose(mat, mt)
        self._transposes[key] = mt
        return self.solve(mt, rhs, sol)


class NewtonStateSolver:
    """Solves R(u, m) = 0 for u by Newton's method from u = 0 with the
    residual's state Jacobian and a LinearSolver (femx's state solvers play
    this role for ReducedFunctional)."""

    def __init__(self, residual, linear_solver, tol=1e-12, max_iter=20):
        self.res, self.lin = residual, linear_solver
        self.tol, self.max_iter = tol, max_iter
        d = residual.dims()
        dev = residual.geom.device
        self.jac = CsrMatrix(residual.pattern(), dev)
        self.r = torch.empty(d["num_res"], dtype=torch.float64, device=dev)
        self.du = torch.empty(d["num_states"], dtype=torch.float64, device=dev)
        self.rhs = torch.empty(d["num_res"], dtype=torch.float64, device=dev)
        self.vh = VectorHandler()

    def solve(self, prm, state):
        state.zero_()
        for _ in range(self.max_iter):
            self.res.res(state, prm, self.r)
            rn = float(self.vh.dot(self.r, self.r).item()) ** 0.5
            if rn == 0.0:
                break
            self.res.assemble_state_jac(state, prm, self.jac)
            self.du.zero_()
            self.vh.axpby(-1.0, self.r, 0.0, self.rhs)
            self.lin.solve(self.jac, self.rhs, self.du)
            self.vh.axpby(1.0, self.du, 1.0, state)
            un = float(self.vh.dot(state, state).item()) ** 0.5
            dn = float(self.vh.dot(self.du, self.du).item()) ** 0.5
            if dn <= self.tol * max(un, 1.0):
                break
        return state


class StateMisfit:
    """J(u, m) = 1/2 ||u - u_obs||^2 + beta/2 ||m||^2 with its state and
    parameter gradients (femx's Objective contract)."""

    def __init__(self, u_obs, beta=0.0):
        self.u_obs, self.beta = u_obs, float(beta)
        self.vh = VectorHandler()

    def value(self, state, prm):
        diff = state.clone()
        self.vh.axpby(-1.0, self.u_obs, 1.0, diff)
        v = 0.5 * float(self.vh.dot(diff, diff).item())
        if self.beta:
            v += 0.5 * self.beta * float(self.vh.dot(prm, prm).item())
        return v

    def state_grad(self, state, prm, out):
        self.vh.axpby(1.0, state, 0.0, out)
        return self.vh.axpby(-1.0, self.u_obs, 1.0, out)

    def param_grad(self, state, prm, out):
        return self.vh.axpby(self.beta, prm, 0.0, out)


class ReducedFunctional:
    """m -> J(u(m), m) with its adjoint gradient, femx's ReducedFunctional
    (femx/inverse/ReducedFunctional.hpp:20-140): solve the state, form
    dJ/du, solve J_u^T lam = dJ/du, and return dJ/dm - (dR/dm)^T lam."""

    def __init__(self, state_solver, residual, adj_jac: CsrMatrix, adj_solver, objective):
        self.state_solver, self.res = state_solver, residual
        self.adj_jac, self.adj_solver, self.obj = adj_jac, adj_solver, objective
        d = residual.dims()
        dev = adj_jac.device
        new = lambda n: torch.empty(n, dtype=torch.float64, device=dev)  # noqa: E731
        self.num_params = d["num_params"]
        self.state, self.state_grad = new(d["num_states"]), new(d["num_states"])
        self.adj, self.prm_grad = new(d["num_res"]), new(self.num_params)
        self.res_prm_adj = new(self.num_params)
        self.vh = VectorHandler()

    def _check(self, prm):
        _require(prm.numel() == self.num_params, "ReducedFunctional parameter size mismatch")

    def value(self, prm):
        self._check(prm)
        self.state_solver.solve(prm, self.state)
        return self.obj.value(self.state, prm)

    def grad(self, prm, out):
        self._check(prm)
        self.state_solver.solve(prm, self.state)
        return self._grad_at(prm, out)

    def value_grad(self, prm, out):
        self._check(prm)
        self.state_solver.solve(prm, self.state)
        val = self.obj.value(self.state, prm)
        self._grad_at(prm, out)
        return val

    def _grad_at(self, prm, out):
        self.obj.state_grad(self.state, prm, self.state_grad)
        self.res.assemble_state_jac(self.state, prm, self.adj_jac)
        self.adj.zero_()
        self.adj_solver.solveT(self.adj_jac, self.state_grad, self.adj)
        self.obj.param_grad(self.state, prm, self.prm_grad)
        self.res.apply_param_jac_t(self.state, prm, self.adj, self.res_prm_adj)
        self.vh.axpby(-1.0, self.res_prm_adj, 1.0, self.prm_grad)
        return self.vh.axpby(1.0, self.prm_grad, 0.0, out)
