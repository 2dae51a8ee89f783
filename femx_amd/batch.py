"""Batched assembly + adjoint VJP of many P1 Poisson, Taylor-Hood and GLS
P1/P1 Navier-Stokes instances (fx_batch_* in include/femx_amd.h), the
per-instance "plain" path it is compared with, and the autotuner for its
launch shape.

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

from . import _lib, gls, ns
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
            ptr(self.lam), ptr(self.R.bmap.d_mask), ptr(self.jt), ptr(self.g),
            self.R.patches().plan)

    def plain(self, what):
        from .assembly import PoissonP1, assemble
        if what & 1:
            assemble(PoissonP1(f=self.f, log_kappa=self.m), self.geom, self.map, self.u, self.res,
                     self.jac)
        if what & 2:
            self.R.vjp(self.u, self.m, self.lam, self.jt, self.g)

    def outputs(self):
        return [self.jac.vals, self.res, self.jt, self.g]


# Taylor-Hood block layouts the autotuner chooses from: (work items per
# block, row-tile bound in CSR entries; 0 = the LDS capacity)
TH_LAYOUTS = ((32, 0), (24, 0), (16, 0), (32, 1500))
# persistent workgroups per compute unit (0: one workgroup per work item)
GRID_CANDIDATES = (0, 2, 8)


class THCase:
    """One Taylor-Hood instance: lid-driven cavity, Re, Fourier state,
    adjoint, per-element viscosity sensitivities.  With `rounds` the batch's
    assembly phase walks the instance's round tables (fx_th_assemble_rounds'
    walks, twelve blocks per work item) instead of taking one block per item."""

    def __init__(self, n, seed, Re, device, items=ns.TH_ITEMS, layouts=TH_LAYOUTS, rounds=True):
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
        self.geom = ns.structured_th(n, device)
        g = self.geom
        self.nu = 1.0 / Re
        self.state = t(ns.fourier_state(g, seed))
        self.lam = t(np.random.Generator(np.random.Philox(seed + 1)).standard_normal(g.ndof))
        self.map = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
        dofs, _ = ns.cavity_bc(g)
        self.bmap = make_boundary_map(dofs, self.map.pattern, device)
        self.plan = ns.THPlan(self.map, device, items, variants=layouts)
        setattr(self.map, "_th_plan_" + str(torch.device(device)), self.plan)
        self.patches = ns.THPatches(g, device)
        setattr(self.map, "_th_patches_" + str(torch.device(device)), self.patches)
        self.map.release_host()
        self.jac = CsrMatrix(self.map.pattern, device)
        self.res = torch.empty(g.ndof, dtype=torch.float64, device=device)
        self.jt = torch.empty(g.ndof, dtype=torch.float64, device=device)
        self.g = torch.empty(g.num_elems, dtype=torch.float64, device=device)
        self.elements = g.num_elems
        self.rounds = rounds  # the assembly phase over the round tables

    def descriptor(self):
        g, p = self.geom, self.plan
        rp, _ = self.map.pattern.device_arrays(g.device)
        self._keep = rp
        return _lib.THInstance(
            g.ndof, g.num_elems, p.num_blocks, p.max_block_nnz, ptr(g.xv), ptr(g.yv),
            ptr(g.vconn3), ptr(g.dofs15), ptr(self.state), self.nu, None, ptr(p.inc_off),
            ptr(p.inc_elem), ptr(p.inc_meta), ptr(p.blk_row), ptr(p.blk_flags), ptr(rp),
            ptr(self.jac.vals), ptr(self.res), ptr(self.lam), ptr(self.bmap.d_mask),
            ptr(self.jt), ptr(self.g), ptr(p.blk_lanes), self.patches.plan,
            ptr(p.rt_off) if self.rounds else None, ptr(p.rt_elem) if self.rounds else None,
            ptr(p.rt_meta) if self.rounds else None)

    def plain(self, what):
        op = ns.TaylorHoodNS(self.nu)
        if what & 1:
            ns.assemble_ns(op, self.geom, self.map, self.state, self.res, self.jac)
        if what & 2:
            ns.ns_vjp(op, self.geom, self.map, self.state, self.lam, self.bmap.d_mask, self.jt,
                      self.g)

    def outputs(self):
        return [self.jac.vals, self.res, self.jt, self.g]


class GLSCase:
    """One GLS P1/P1 Navier-Stokes time step on the structured unit square
    (n x n cells, two triangles each): history, next state and adjoint drawn
    from Philox(seed); assembly of the residual and Jacobian and the history
    VJP of every lag (two lags: step 1, extrapolated advection; one lag:
    step 0)."""

    def __init__(self, n, seed, device, rho=1.0, mu=1e-3, dt=0.01, num_hist=2):
        from .assembly import structured_p1
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
        p = structured_p1(n, "cpu")
        self.geom = gls.GLSGeometry(p.h_x, p.h_y, p.h_conn, device)
        g = self.geom
        self.map = gls.make_navier_map(g)
        self.op = gls.NavierOperator(g, gls.FluidParams(rho, mu), dt)
        self.step, self.num_hist = (1, 2) if num_hist > 1 else (0, 1)
        rng = np.random.Generator(np.random.Philox(seed))
        S = g.num_states
        self.hist = t(0.3 * rng.standard_normal(self.num_hist * S))
        self.nxt = t(0.3 * rng.standard_normal(S))
        self.adj = t(rng.standard_normal(S))
        self.plan = gls._plan(self.map, g)
        self.patches = gls._patches(self.map, g)
        self.jac = CsrMatrix(self.map.pattern, device)
        self.res = torch.empty(S, dtype=torch.float64, device=device)
        self.out0 = torch.empty(S, dtype=torch.float64, device=device)
        self.out1 = (torch.empty(S, dtype=torch.float64, device=device) if self.num_hist > 1
                     else None)
        self.elements = g.num_elems

    def descriptor(self):
        g, p, f = self.geom, self.plan, self.op.fluid
        return _lib.GLSInstance(
            g.num_nodes, g.num_states, g.num_elems, p.max_tile, ptr(g.x), ptr(g.y), ptr(g.conn),
            ptr(g.dofs9), ptr(p.node_rows), f.rho, f.mu, self.op.dt, self.step, self.num_hist,
            ptr(self.hist), ptr(self.nxt), ptr(p.inc_off), ptr(p.inc_elem), ptr(p.inc_meta),
            ptr(p.blk_runs), ptr(p.node_tile), ptr(self.jac.vals), ptr(self.res), ptr(self.adj),
            ptr(self.out0), ptr(self.out1), self.patches.plan)

    def plain(self, what):
        if what & 1:
            gls.assemble(self.op, self.step, self.num_hist, gls.NEXT_STATE, self.map, self.hist,
                         self.nxt, self.res, self.jac)
        if what & 2:
            gls.hist_vjp_lags(self.op, self.step, self.num_hist, self.map, self.hist, self.nxt,
                              self.adj, self.out0,
                              self.out1 if self.out1 is not None else False)

    def outputs(self):
        return [t for t in (self.jac.vals, self.res, self.out0, self.out1) if t is not None]


def size_class(case, bounds=(8192, 131072)):
    """0, 1 or 2 by element count (n < 64, n < 256, larger on the unit
    square)."""
    return sum(case.elements >= b for b in bounds)


def build_cases(specs, device="cuda", items=ns.TH_ITEMS, log=None):
    cases = []
    for k, s in enumerate(specs):
        if log and k % 32 == 0:
            log(f"[batch] building instance {k}/{len(specs)}")
        if s["kind"] == "p1":
            cases.append(P1Case(s["n"], s["seed"], device))
        elif s["kind"] == "gls":
            cases.append(GLSCase(s["n"], s["seed"], device, num_hist=s.get("num_hist", 2)))
        else:
            cases.append(THCase(s["n"], s["seed"], s["Re"], device, items))
    return cases


class Batch:
    """fx_batch over a list of cases."""

    ASSEMBLE, VJP, SERIAL = 1, 2, 4

    def __init__(self, cases, grid_per_cu=4):
        self.cases = cases
        p1 = [c.descriptor() for c in cases if isinstance(c, P1Case)]
        th = [c.descriptor() for c in cases if isinstance(c, THCase)]
        gl = [c.descriptor() for c in cases if isinstance(c, GLSCase)]
        self._p1 = (_lib.P1Instance * max(len(p1), 1))(*p1)
        self._th = (_lib.THInstance * max(len(th), 1))(*th)
        self._gls = (_lib.GLSInstance * max(len(gl), 1))(*gl)
        h = C.c_void_p()
        _lib.call("fx_batch_create", len(p1), C.cast(self._p1, C.c_void_p), len(th),
                  C.cast(self._th, C.c_void_p), len(gl), C.cast(self._gls, C.c_void_p),
                  int(grid_per_cu), C.byref(h))
        self._h = h
        self._destroy = _lib.lib().fx_batch_destroy
        self.grid_per_cu = grid_per_cu

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def work(self):
        counts = (C.c_int32 * 9)()
        grid = C.c_int32()
        _lib.call("fx_batch_work", self._h, C.cast(counts, C.c_void_p), C.byref(grid))
        return list(counts), grid.value

    def run(self, what=3, stream=None):
        _lib.call("fx_batch_run", self._h, int(what), stream_handle(stream))

    def capture(self, what=3):
        _lib.call("fx_batch_capture", self._h, int(what))

    def launch(self, stream=None):
        _lib.call("fx_batch_launch", self._h, stream_handle(stream))


class BatchGraph:
    """One hipGraph over several batches (fx_batch_graph_create): every
    chain of every batch is a branch, or (what & SERIAL) every phase runs in
    turn over the batches."""

    def __init__(self, batches, what=3):
        self.batches = list(batches)
        arr = (C.c_void_p * len(self.batches))(*[b._h.value for b in self.batches])
        h = C.c_void_p()
        _lib.call("fx_batch_graph_create", len(self.batches), C.cast(arr, C.c_void_p), int(what),
                  C.byref(h))
        self._h = h
        self._destroy = _lib.lib().fx_batch_graph_destroy

    def __del__(self):
        if getattr(self, "_h", None):
            self._destroy(self._h)
            self._h = None

    def launch(self, stream=None):
        _lib.call("fx_batch_graph_launch", self._h, stream_handle(stream))


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


def _graph_time(cases, grid_per_cu, reps):
    """A group's phases in series, so that its shape is timed alone."""
    b = Batch(cases, grid_per_cu)
    b.capture(3 | Batch.SERIAL)
    t, _ = time_fn(b.launch, reps)
    return t


def autotune(cases, layouts=TH_LAYOUTS, grids=GRID_CANDIDATES, reps=3, log=None):
    """The launch shape of the batch, chosen by timing captured graphs.

    Instances are grouped by kind and size class; each group becomes a batch
    of its own (its own grid and Taylor-Hood row tile, so a small-instance
    group does not inherit a large group's LDS tile), and all groups' chains
    are branches of one graph.  Per group: for Taylor-Hood, the block layout
    (work items per block -- the elements a workgroup's rows gather -- and
    the row-tile bound) from `layouts`; then the persistent-grid multiplier
    (or one workgroup per work item, 0) from `grids`.  Last, the per-group
    batches and one batch over all instances (grid 8) are each captured with
    the chains in series and as concurrent branches, and the fastest of the
    four graphs is kept.  Returns (runner, report): runner.launch() replays
    the tuned graph; the report holds every timing and choice."""
    groups = {}
    for c in cases:
        kind = "th" if isinstance(c, THCase) else "gls" if isinstance(c, GLSCase) else "p1"
        groups.setdefault((kind, size_class(c)), []).append(c)
    report = {"groups": {}}
    batches = []
    for (kind, cls), gc in sorted(groups.items()):
        is_th = kind == "th"
        key = f"{kind}_class{cls}"
        rec = {"instances": len(gc), "elements": int(sum(c.elements for c in gc))}
        if is_th:
            tl = {}
            for lay in layouts:
                for c in gc:
                    c.plan.select(*lay)
                tl[lay] = _graph_time(gc, 8, reps)
            best_lay = min(tl, key=tl.get)
            for c in gc:
                c.plan.select(*best_lay)
            rec["layout_ms"] = {f"{a}x{b}": v * 1e3 for (a, b), v in tl.items()}
            rec["layout"] = {"items": best_lay[0], "tile_nnz": best_lay[1]}
        tg = {g: _graph_time(gc, g, reps) for g in grids}
        best_g = min(tg, key=tg.get)
        rec["grid_ms"] = {str(g): v * 1e3 for g, v in tg.items()}
        rec["grid_per_cu"] = best_g
        report["groups"][key] = rec
        if log:
            log(f"[autotune] {key}: {rec}")
        batches.append(Batch(gc, best_g))
    single = Batch(cases, 8)
    cands = {}
    for mode, what in (("serial", 3 | Batch.SERIAL), ("concurrent", 3)):
        cands[f"grouped_{mode}"] = BatchGraph(batches, what)
        cands[f"single_{mode}"] = BatchGraph([single], what)
    tf = {k: time_fn(g.launch, reps)[0] for k, g in cands.items()}
    best = min(tf, key=tf.get)
    report["graph_ms"] = {k: v * 1e3 for k, v in tf.items()}
    report["graph"] = best
    if log:
        log(f"[autotune] graphs {report['graph_ms']} -> {best}")
    runner = cands[best]
    runner.keep = batches + [single]
    return runner, report
