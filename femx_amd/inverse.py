"""Stationary residual contract and its adjoint VJP as a torch autograd
Function.

`PoissonLogKappaResidual` implements femx's stationary Residual contract
(femx/state/Residual.hpp:18-60) for the P1 Poisson residual with control
m_e = log kappa_e and Dirichlet rows replaced:
    R_i(u, m) = (K(exp m) u - F)_i   on unconstrained rows,
    R_r(u, m) = u_r - g_r            on constrained rows (replaceRes),
whose state Jacobian is the row-replaced stiffness (replaceRows with diagonal
one, femx/assembly/BoundaryMap.hpp:198-208).  `apply_param_jac_t` is the
(dR/dm)^T adj that ReducedFunctional::gradAt subtracts from dJ/dm
(femx/inverse/ReducedFunctional.hpp:94-117).

`ResidualVJP` exposes R(u, m) to torch autograd: its backward is one pass
of fx_p1_vjp_patch giving J^T lam and (dR/dm)^T lam together.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .assembly import (AssemblyMap, BoundaryMap, Geometry, PoissonP1, assemble, make_assembly_map,
                       make_boundary_map, p1_patches, replace_res, replace_rows)
from .linalg import CsrMatrix, _check_dev_vec, ptr, stream_handle


class PoissonLogKappaResidual:
    def __init__(self, geom: Geometry, bc_dofs=None, bc_vals=None, f=None,
                 amap: AssemblyMap = None, bmap: BoundaryMap = None):
        self.geom = geom
        conn = geom.h_conn
        self.map = amap if amap is not None else make_assembly_map(geom.num_nodes, geom.num_nodes,
                                                                   conn)
        self.num_states = self.num_res = geom.num_nodes
        self.num_params = geom.num_elems
        if bmap is None and bc_dofs is not None:
            bmap = make_boundary_map(bc_dofs, self.map.pattern, geom.device)
        self.bmap = bmap
        dev = geom.device
        if bmap is not None:
            self.bc_vals = (torch.zeros(bmap.num_bcs, dtype=torch.float64, device=dev)
                            if bc_vals is None else bc_vals)
        else:
            self.bc_vals = None
        self.f = f
        self._patches = None

    def dims(self):
        return dict(num_states=self.num_states, num_res=self.num_res, num_params=self.num_params)

    def pattern(self):
        return self.map.pattern

    def res(self, state, prm, out, stream=None):
        """out = R(state, prm)."""
        assemble(PoissonP1(f=self.f, log_kappa=prm), self.geom, self.map, state, out, None, stream)
        if self.bmap is not None:
            replace_res(self.bmap, state, self.bc_vals, out, stream)
        return out

    def assemble_state_jac(self, state, prm, out: CsrMatrix, stream=None):
        """out = dR/dstate (row-replaced stiffness)."""
        tmp = torch.empty(self.num_res, dtype=torch.float64, device=self.geom.device)
        assemble(PoissonP1(f=self.f, log_kappa=prm), self.geom, self.map, state, tmp, out, stream)
        if self.bmap is not None:
            replace_rows(self.bmap, out, 1.0, stream)
        return out

    def vjp(self, state, prm, adj, jt_out=None, g_out=None, stream=None):
        """(J^T adj, (dR/dm)^T adj) in one pass over element patches
        (fx_p1_vjp_patch)."""
        g = self.geom
        dev = g.device
        _check_dev_vec(state, self.num_states, torch.float64, "VJP state")
        _check_dev_vec(prm, self.num_params, torch.float64, "VJP parameters")
        _check_dev_vec(adj, self.num_res, torch.float64, "VJP adjoint")
        if jt_out is None:
            jt_out = torch.empty(self.num_states, dtype=torch.float64, device=dev)
        if g_out is None:
            g_out = torch.empty(self.num_params, dtype=torch.float64, device=dev)
        mask = self.bmap.d_mask if self.bmap is not None else None
        _lib.call("fx_p1_vjp_patch", g.num_nodes, g.num_elems, ptr(g.x), ptr(g.y), ptr(prm),
                  ptr(state), ptr(adj), ptr(mask), C.byref(self.patches().plan), ptr(jt_out),
                  ptr(g_out), stream_handle(stream))
        return jt_out, g_out

    def patches(self):
        """The VJP's patch plan on the geometry's device (built on first use;
        VJPs of this residual on different streams must not overlap)."""
        if self._patches is None:
            self._patches = p1_patches(self.geom, self.map)
        return self._patches

    def apply_param_jac_t(self, state, prm, adj, out=None, stream=None):
        """out = (dR/dprm)^T adj (Residual::applyParamJacT)."""
        _, g = self.vjp(state, prm, adj, g_out=out, stream=stream)
        return g

    def apply_state_jac_t(self, state, prm, adj, out=None, stream=None):
        """out = (dR/dstate)^T adj over the row-replaced Jacobian."""
        jt, _ = self.vjp(state, prm, adj, jt_out=out, stream=stream)
        return jt


class ResidualVJP(torch.autograd.Function):
    """R = residual(u, m) with backward (lam) -> (J^T lam, (dR/dm)^T lam)."""

    @staticmethod
    def forward(ctx, u, m, residual):
        u = u.contiguous()
        m = m.contiguous()
        out = torch.empty(residual.num_res, dtype=torch.float64, device=u.device)
        residual.res(u, m, out)
        ctx.residual = residual
        ctx.save_for_backward(u, m)
        return out

    @staticmethod
    def backward(ctx, lam):
        u, m = ctx.saved_tensors
        jt, g = ctx.residual.vjp(u, m, lam.contiguous())
        return jt, g, None


def residual_fn(u, m, residual):
    return ResidualVJP.apply(u, m, residual)
