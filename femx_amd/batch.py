"""Batched assembly + adjoint VJP of many P1 Poisson and Taylor-Hood
instances (fx_batch_* in include/femx_amd.h), the per-instance "plain" path
it is compared with, and the autotuner for its launch shape.

The batched path runs persistent launches over (instance, block) work lists
and can be captured into one hipGraph; the plain path is femx's call pattern
-- one assembly and one VJP call per instance, each with its own checks and
launches (femx/assembly/CudaAssembly.hpp:319-364).  Both write the same
per-instance output arrays, bitwise identically.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib, ns
from .assembly import make_assembly_map, make_boundary_map, structured_p1, unit_square_edge_nodes
from .inverse import PoissonLogKappaResidual
from .linalg import CsrMatrix, ptr, stream_handle


def instance_specs(count=256, seed=9, n_min=16, n_max=512):
    """Config-5 draws: n_i = round(n_min (n_max/n_min)^U_i), U_i ~ U(0,1)
    (numpy Philox seeded with `seed`); even i -> P1 Poisson, odd i ->
    Taylor-Hood NS with Re_i ~ U[10, 400] (Philox seed 200 + i)."""
    u = np.random.Generator(np.random.Philox(seed)).random(count)
    ns_ = np.rint(n_min * (n_max / n_min) ** u).astype(int)
    specs = []
    for i, n in enumerate(ns_):
        if i % 2 == 0:
            specs.append(dict(kind="p1", n=int(n), seed=100 + i))
        else:
            re = float(np.random.Generator(np.random.Philox(200 + i)).uniform(10.0, 400.0))
            specs.append(dict(kind="th", n=int(n), seed=200 + i, Re=re))
    return specs


class P1Case:
    """One P1 instance: log-kappa control, load, state, adjoint, BCs on all
    edges, and its output arrays."""

    def __init__(self, n, seed, device):
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
        self.geom = structured_p1(n, device)
        g = self.geom
        x, y = g.h_x, g.h_y
        dofs = unit_square_edge_nodes(x, y)
        rng = np.random.Generator(np.random.Philox(seed))
        self.m = t(0.5 * rng.standard_normal(g.num_elems))
        self.f = t(rng.standard_normal(g.num_nodes))
        self.u = t(rng.standard_normal(g.num_nodes))
        self.lam = t(rng.standard_normal(g.num_nodes))
        self.R = PoissonLogKappaResidual(g, dofs, f=self.f)
        self.map = self.R.map
        self.jac = CsrMatrix(self.map.pattern, device)
        self.res = torch.empty(g.num_nodes, dtype=torch.float64, device=device)
        self.jt = torch.empty(g.num_nodes, dtype=torch.float64, device=device)
        self.g = torch.empty(g.num_elems, dtype=torch.float64, device=device)
        self.elements = g.num_elems

    def descriptor(self):
        g, m = self.geom, self.map
        rp, _ = m.pattern.device_arrays(g.device)
        self._keep = rp
        return _lib.P1Instance(
            g.num_nodes, g.num_elems, m.pattern.nnz, m.max_row_len, ptr(g.x), ptr(g.y),
            ptr(g.conn), ptr(self.m), ptr(self.f), ptr(self.u), ptr(m.device("inc_off", g.device)),
            ptr(m.device("inc", g.device)), ptr(rp), ptr(self.jac.vals), ptr(self.res),
            ptr(self.lam), ptr(self.R.bmap.d_mask), ptr(self.jt), ptr(self.g))

    def plain(self, what):
        from .assembly import PoissonP1, assemble
        if what & 1:
            assemble(PoissonP1(f=self.f, log_kappa=self.m), self.geom, self.map, self.u, self.res,
                     self.jac)
        if what & 2:
            self.R.vjp(self.u, self.m, self.lam, self.jt, self.g)

    def outputs(self):
        return [self.jac.vals, self.res, self.jt, self.g]


class THCase:
    """One Taylor-Hood instance: lid-driven cavity, Re, Fourier state,
    adjoint, per-element viscosity sensitivities."""

    def __init__(self, n, seed, Re, device, items=ns.TH_ITEMS):
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
        self.geom = ns.structured_th(n, device)
        g = self.geom
        self.nu = 1.0 / Re
        self.state = t(ns.fourier_state(g, seed))
        self.lam = t(np.random.Generator(np.random.Philox(seed + 1)).standard_normal(g.ndof))
        self.map = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
        dofs, _ = ns.cavity_bc(g)
        self.bmap = make_boundary_map(dofs, self.map.pattern, device)
        self.plan = ns.THPlan(self.map, device, items)
        setattr(self.map, "_th_plan_" + str(torch.device(device)), self.plan)
        self.map.release_host()
        self.jac = CsrMatrix(self.map.pattern, device)
        self.res = torch.empty(g.ndof, dtype=torch.float64, device=device)
        self.jt = torch.empty(g.ndof, dtype=torch.float64, device=device)
        self.g = torch.empty(g.num_elems, dtype=torch.float64, device=device)
        self.elements = g.num_elems

    def descriptor(self):
        g, p = self.geom, self.plan
        rp, _ = self.map.pattern.device_arrays(g.device)
        self._keep = rp
        return _lib.THInstance(
            g.ndof, g.num_elems, p.num_blocks, p.max_block_nnz, ptr(g.xv), ptr(g.yv),
            ptr(g.vconn3), ptr(g.dofs15), ptr(self.state), self.nu, None, ptr(p.inc_off),
            ptr(p.inc_elem), ptr(p.inc_meta), ptr(p.blk_row), ptr(p.blk_flags), ptr(rp),
            ptr(self.jac.vals), ptr(self.res), ptr(p.elem_inc), ptr(self.lam), ptr(self.bmap.d_mask),
            ptr(self.jt), ptr(self.g), ptr(p.work), ptr(p.blk_lanes))

    def plain(self, what):
        op = ns.TaylorHoodNS(self.nu)
        if what & 1:
            ns.assemble_ns(op, self.geom, self.map, self.state, self.res, self.jac)
        if what & 2:
            ns.ns_vjp(op, self.geom, self.map, self.state, self.lam, self.bmap.d_mask, self.jt,
                      self.g)

    def outputs(self):
        return [self.jac.vals, self.res, self.jt, self.g]


def build_cases(specs, device="cuda", items=ns.TH_ITEMS, log=None):
    cases = []
    for k, s in enumerate(specs):
        if log and k % 32 == 0:
            log(f"[batch] building instance {k}/{len(specs)}")
        if s["kind"] == "p1":
            cases.append(P1Case(s["n"], s["seed"], device))
        else:
            cases.append(THCase(s["n"], s["seed"], s["Re"], device, items))
    return cases


class Batch:
    """fx_batch over a list of cases."""

    ASSEMBLE, VJP = 1, 2

    def __init__(self, cases, grid_per_cu=4):
        self.cases = cases
        p1 = [c.descriptor() for c in cases if isinstance(c, P1Case)]
        th = [c.descriptor() for c in cases if isinstance(c, THCase)]
        self._p1 = (_lib.P1Instance * max(len(p1), 1))(*p1)
        self._th = (_lib.THInstance * max(len(th), 1))(*th)
        h = C.c_void_p()
        _lib.call("fx_batch_create", len(p1), C.cast(self._p1, C.c_void_p), len(th),
                  C.cast(self._th, C.c_void_p), int(grid_per_cu), C.byref(h))
        self._h = h
        self._destroy = _lib.lib().fx_batch_destroy
        self.grid_per_cu = grid_per_cu

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def work(self):
        counts = (C.c_int32 * 5)()
        grid = C.c_int32()
        _lib.call("fx_batch_work", self._h, C.cast(counts, C.c_void_p), C.byref(grid))
        return list(counts), grid.value

    def run(self, what=3, stream=None):
        _lib.call("fx_batch_run", self._h, int(what), stream_handle(stream))

    def capture(self, what=3):
        _lib.call("fx_batch_capture", self._h, int(what))

    def launch(self, stream=None):
        _lib.call("fx_batch_launch", self._h, stream_handle(stream))


def plain(cases, what=3):
    for c in cases:
        c.plain(what)


def time_fn(fn, reps, warmup=1):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps * 1e-3, (time.perf_counter() - t0) / reps


def autotune(cases, grid_candidates=(1, 2, 4, 8), reps=3):
    """Pick the persistent grid multiplier (workgroups per compute unit) by
    timing the captured graph for each candidate; returns (best, timings)."""
    timings = {}
    for gpc in grid_candidates:
        b = Batch(cases, gpc)
        b.capture(3)
        t, _ = time_fn(b.launch, reps)
        timings[gpc] = t
        del b
    best = min(timings, key=timings.get)
    return best, timings
