"""Benchmark of the FE core on one MI355X.

Step (the timed unit): on the P1 Poisson 2048x2048 structured triangle mesh
(8,388,608 elements, 4,198,401 nodes, 29,372,417 CSR entries), FP64,
kappa_e = exp(0.5 z_e) (seed 2), f ~ N(0,1) (seed 3), u ~ N(0,1) (seed 4):
  1. assemble the CSR Jacobian and the residual K u - F: 97% of the rows lie
     in 64-row blocks that share one incidence stencil, assembled from
     coalesced loads of their columns' vertex data (fx_p1_assemble_stencil);
     the rest gather per row
  2. rhs = -res; Dirichlet elimination on all edges, g = 0 (fx_bc_prepare_forward)
  3. 10 SpMVs y_k = A x_k with x_k ~ U(-1,1) (seeds 10..19), issued as one
     multi-vector product that streams A once, its 64-row stencil blocks
     reading each band of the vectors with one coalesced load
     (fx_csr_matvec_multi_st); each y_k equals the single-vector
     fx_csr_matvec result bit for bit
All inputs are resident in HBM before timing.  `value` is elements per
second over whole steps.  `roofline` is the step's longest kernel (the
10-vector SpMV or the assembly, by measured mean duration): its algorithmic
bytes per launch over its mean duration inside the timed steps (HIP events on
its stream around each launch), with the other kernel under `also`.  Inside
the step the SpMV's 1 GB stream has evicted the assembly's inputs from the
256 MB Infinity Cache, so these are the kernels' in-step durations;
`--back-to-back` adds each one's mean over back-to-back launches (inputs left
in the caches by the launch before) under `kernels`.  The general-mesh forms
of both kernels (element-patch and row-per-lane assembly, gathering SpMV) are
timed back to back beside them under `kernels`.  Assembly bytes: coords, conn, kappa, f, u,
femx's jac_map, CSR values, residual; SpMV bytes: CSR values, columns, row
pointers, 10 x and 10 y vectors.  The
`cpu_baseline` times the same step on the host with the oracle's OpenMP
restatement of femx's CPU assembly (element loop, thread-private scratch,
atomic scatter) plus femx's serial host elimination and matvec.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "elements/s assembled (Jac+res) & VJPs/s, %HBM roofline; batched vs plain GPU"


def p1_alg_bytes(N, E, nnz):
    """Algorithmic bytes of one P1 assembly launch: each input read once,
    each output written once, plus femx's element->CSR map (jac_map)."""
    return 2 * N * 8 + 3 * E * 4 + E * 8 + N * 8 + N * 8 + 9 * E * 4 + nnz * 8 + N * 8


def vjp_alg_bytes(N, E):
    """coords, conn, m, lam, u, node->element incidence + offsets, J^T lam, g."""
    return 2 * N * 8 + 3 * E * 4 + E * 8 + N * 8 + N * 8 + 3 * E * 4 + (N + 1) * 4 + N * 8 + E * 8


def th_alg_bytes(Np, E, ndof, nnz):
    """vertex coords, element DOF table, state, femx's jac_map, CSR values, residual."""
    return 2 * Np * 8 + 15 * E * 4 + ndof * 8 + 225 * E * 4 + nnz * 8 + ndof * 8


def th_vjp_alg_bytes(Np, E, ndof):
    """vertex coords, DOF table, state, adjoint, incidence + offsets, J^T lam, g."""
    return 2 * Np * 8 + 15 * E * 4 + 2 * ndof * 8 + 15 * E * 4 + (ndof + 1) * 4 + ndof * 8 + E * 8


def spmv_alg_bytes(N, nnz):
    return nnz * 8 + nnz * 4 + (N + 1) * 4 + N * 8 + N * 8


def spmv_multi_alg_bytes(N, nnz, nv):
    """The matrix once, then each x read and each y written."""
    return nnz * 8 + nnz * 4 + (N + 1) * 4 + nv * (N * 8 + N * 8)


def gls_alg_bytes(N, E, nnz):
    """GLS P1/P1 time step: coords, conn, femx's element DOF table, the
    velocity of both history lags, the next state, femx's jac_map, CSR values,
    residual."""
    return (2 * N * 8 + 3 * E * 4 + 9 * E * 4 + 2 * (2 * N) * 8 + 3 * N * 8 + 81 * E * 4
            + nnz * 8 + 3 * N * 8)


def gls_vjp_alg_bytes(N, E):
    """GLS history VJP: coords, conn, element DOF table, the velocity of both
    history lags, next state, adjoint, output."""
    return 2 * N * 8 + 3 * E * 4 + 9 * E * 4 + 2 * (2 * N) * 8 + 3 * N * 8 * 3


# FP64 operations (add, mul, div, sqrt each 1) of one node-incidence of the
# GLS kernel: geometry 15, extrapolation 18, velocity gradient 16, per
# quadrature point 242 (point state 55, row factors 11, F terms 41, 3 columns
# x 45), residual dot products 42.
GLS_FLOPS_PER_INCIDENCE = 15 + 18 + 16 + 3 * 242 + 42
FP64_PEAK_TFLOPS = 78.6


def build_gls(n, dev, t):
    """GLS P1/P1 Navier-Stokes time step on the P1 n x n mesh: rho 1, mu 1e-3,
    dt 0.01, step 1 with two history lags ~ N(0,1) (seed 12) and next state
    ~ N(0,1) (seed 13); returns launch closures."""
    import torch
    from femx_amd import gls
    from femx_amd.linalg import CsrMatrix, ptr
    x = np.empty((n + 1) ** 2)
    y = np.empty_like(x)
    conn = np.empty(6 * n * n, np.int32)
    from femx_amd import _lib
    _lib.call("fx_mesh_p1", n, ptr(x), ptr(y), ptr(conn))
    geom = gls.GLSGeometry(x, y, conn, dev)
    amap = gls.make_navier_map(geom)
    nd = geom.num_states
    hist = np.random.default_rng(12).standard_normal(2 * nd)
    nxt = np.random.default_rng(13).standard_normal(nd)
    d_hist, d_nxt = t(hist), t(nxt)
    op = gls.NavierOperator(geom, gls.FluidParams(1.0, 1e-3), 0.01)
    jac = CsrMatrix(amap.pattern, dev)
    res = torch.empty(nd, dtype=torch.float64, device=dev)
    gls.assemble(op, 1, 2, gls.NEXT_STATE, amap, d_hist, d_nxt, res, jac)
    d_adj = t(np.random.default_rng(14).standard_normal(nd))
    vjp_out = torch.empty(nd, dtype=torch.float64, device=dev)
    vjp_out1 = torch.empty(nd, dtype=torch.float64, device=dev)
    return {"hist_vjp": lambda: gls.hist_vjp(op, 1, 2, 0, amap, d_hist, d_nxt, d_adj, vjp_out),
            "hist_vjp_lags": lambda: gls.hist_vjp_lags(op, 1, 2, amap, d_hist, d_nxt, d_adj,
                                                       vjp_out, vjp_out1),
            "vjp_alg_bytes": gls_vjp_alg_bytes(geom.num_nodes, geom.num_elems),
            "assemble": lambda: gls.assemble(op, 1, 2, gls.NEXT_STATE, amap, d_hist, d_nxt, res,
                                             jac),
            "residual": lambda: gls.assemble_residual(op, 1, 2, amap, d_hist, d_nxt, res),
            "E": geom.num_elems, "N": geom.num_nodes, "states": nd, "nnz": amap.pattern.nnz,
            "alg_bytes": gls_alg_bytes(geom.num_nodes, geom.num_elems, amap.pattern.nnz)}


def build_th(n, dev, t):
    """Config-3 workload: Taylor-Hood n x n lid-driven cavity, Re = 100,
    Fourier state (seed 5) with lid u = 1; returns launch closures."""
    import torch
    from femx_amd import ns
    from femx_amd.assembly import make_assembly_map
    from femx_amd.linalg import CsrMatrix
    g = ns.structured_th(n, dev)
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    s = t(ns.fourier_state(g, 5))
    lam = t(np.random.default_rng(7).standard_normal(g.ndof))
    dofs, vals = ns.cavity_bc(g)
    R = ns.NavierStokesResidual(g, dofs, t(vals), amap=amap)
    jac = CsrMatrix(amap.pattern, dev)
    res = torch.empty(g.ndof, dtype=torch.float64, device=dev)
    jt = torch.empty(g.ndof, dtype=torch.float64, device=dev)
    ge = torch.empty(g.num_elems, dtype=torch.float64, device=dev)
    nu_e = torch.full((g.num_elems,), 0.01, dtype=torch.float64, device=dev)
    op = ns.TaylorHoodNS(0.01)
    return {"assemble": lambda: ns.assemble_ns(op, g, amap, s, res, jac),
            "assemble_lane_map": lambda: ns.assemble_ns(op, g, amap, s, res, jac, rounds=False),
            "vjp": lambda: R.vjp(s, nu_e, lam, jt, ge),
            "E": g.num_elems, "ndof": g.ndof, "nnz": amap.pattern.nnz,
            "alg_bytes": th_alg_bytes(g.num_vertices, g.num_elems, g.ndof, amap.pattern.nnz),
            "vjp_alg_bytes": th_vjp_alg_bytes(g.num_vertices, g.num_elems, g.ndof)}


def run_batched(args, log):
    """Config 5: 256 instances, n ~ logU[16,512] (seed 9), even P1 Poisson,
    odd Taylor-Hood NS (Re ~ U[10,400]); assembly + VJP of every instance.
    Plain = one assembly and one VJP call per instance (femx's pattern);
    batched = one hipGraph chosen by the autotuner: per-size-class batches
    (the Taylor-Hood block layout and the grid of each class tuned) or one
    batch over all instances, with the chains in series or as concurrent
    branches."""
    from femx_amd import batch as B
    t0 = time.perf_counter()
    specs = B.instance_specs(count=args.batch_count, seed=9)
    cases = B.build_cases(specs, "cuda", log=log)
    log(f"[bench] batch built in {time.perf_counter() - t0:.1f} s")
    E = sum(c.elements for c in cases)
    b_alg = 0
    for c in cases:
        if isinstance(c, B.P1Case):
            g = c.geom
            b_alg += p1_alg_bytes(g.num_nodes, g.num_elems, c.map.pattern.nnz)
            b_alg += vjp_alg_bytes(g.num_nodes, g.num_elems)
        else:
            g = c.geom
            b_alg += th_alg_bytes(g.num_vertices, g.num_elems, g.ndof, c.map.pattern.nnz)
            b_alg += th_vjp_alg_bytes(g.num_vertices, g.num_elems, g.ndof)
    runner, tune = B.autotune(cases, reps=3, log=log)
    t_graph, _ = B.time_fn(runner.launch, 10)
    single = B.Batch(cases, 8)  # one batch over all instances, persistent grids
    single.capture(3)
    t_single, _ = B.time_fn(single.launch, 10)
    t_run, _ = B.time_fn(lambda: single.run(3), 10)
    t_plain, wall_plain = B.time_fn(lambda: B.plain(cases, 3), 3)
    counts, grid = single.work()
    return {"instances": len(cases), "elements": E, "alg_bytes": b_alg,
            "graph_ms": t_graph * 1e3, "single_batch_graph_ms": t_single * 1e3,
            "persistent_ms": t_run * 1e3, "plain_ms": t_plain * 1e3,
            "plain_wall_ms": wall_plain * 1e3,
            "elements_per_s_graph": E / t_graph, "elements_per_s_plain": E / t_plain,
            "speedup_graph_vs_plain": t_plain / t_graph,
            "GBs_graph": b_alg / t_graph / 1e9, "frac_graph": b_alg / t_graph / 1e9 / HBM_PEAK_GBS,
            "autotune": tune, "work_items": counts}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_p1_2048_th_512_gls_2048.md")


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (read + written) on the profiled
    2048^2 workload, from the committed rocprofv3 counter summary (calibrated
    FETCH_SIZE and WRITE_SIZE medians); None when the summary is absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                f = line.split()
                if len(f) == 4 and f[0] == "traffic" and f[1] == kernel:
                    return int(f[2]) + int(f[3])
    except OSError:
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vjp", action="store_true")
    ap.add_argument("--no-th", action="store_true")
    ap.add_argument("--th-n", type=int, default=512)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-gls", action="store_true")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--back-to-back", action="store_true",
                    help="also time the step's assembly and 10-vector SpMV launched back to "
                         "back (their inputs left in the caches by the launch before); off by "
                         "default so that a profile of this command averages only the launches "
                         "inside the timed steps")
    ap.add_argument("--kernels-only", action="store_true",
                    help="for rocprofv3 counter runs: a 1 GiB device copy (counter "
                         "calibration), then 5 launches each of assembly, SpMV, elimination "
                         "and the VJPs (the GLS history VJP for lag 0, then for both lags); "
                         "prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "the multi-GPU partition is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch

    from femx_amd import _lib
    from femx_amd import assembly as A
    from femx_amd.inverse import PoissonLogKappaResidual
    from femx_amd.linalg import CsrMatrix, MatrixHandler, VectorHandler, ptr, stream_handle

    _lib.lib()
    dev = "cuda"
    n = args.n
    log(f"[bench] building P1 {n}x{n} mesh and maps")
    geom = A.structured_p1(n, dev)
    N, E = geom.num_nodes, geom.num_elems
    amap = A.make_assembly_map(N, N, geom.h_conn)
    nnz = amap.pattern.nnz
    x, y = geom.h_x, geom.h_y
    dofs = A.unit_square_edge_nodes(x, y)
    bmap = A.make_boundary_map(dofs, amap.pattern, dev)

    kappa = np.exp(0.5 * np.random.default_rng(2).standard_normal(E))
    f = np.random.default_rng(3).standard_normal(N)
    u = np.random.default_rng(4).standard_normal(N)
    xs = [np.random.default_rng(10 + k).uniform(-1.0, 1.0, N) for k in range(10)]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    d_kappa, d_f, d_u = t(kappa), t(f), t(u)
    d_X = t(np.stack(xs))
    d_Y = torch.empty((10, N), dtype=torch.float64, device=dev)
    d_xs = [d_X[k] for k in range(10)]
    d_ys = [d_Y[k] for k in range(10)]
    jac = CsrMatrix(amap.pattern, dev)
    res = torch.empty(N, dtype=torch.float64, device=dev)
    rhs = torch.empty(N, dtype=torch.float64, device=dev)
    g0 = torch.zeros(dofs.size, dtype=torch.float64, device=dev)
    op = A.PoissonP1(d_kappa, d_f)
    mh, vh = MatrixHandler(), VectorHandler()
    rp, ci = amap.pattern.device_arrays(dev)
    inc_off, inc = amap.device("inc_off", dev), amap.device("inc", dev)
    stream = stream_handle()

    def assemble_rows_launch():
        _lib.call("fx_p1_assemble", N, E, nnz, amap.max_row_len, ptr(geom.x), ptr(geom.y),
                  ptr(geom.conn), ptr(d_kappa), 0, ptr(d_f), ptr(d_u), ptr(inc_off), ptr(inc),
                  ptr(rp), ptr(jac.vals), ptr(res), stream)

    asm_plan = A.p1_asm_plan(geom, amap)
    st_plan = A.p1_stencil_plan(geom, amap)

    def assemble_patch_launch():
        _lib.call("fx_p1_assemble_patch", N, E, nnz, ptr(geom.x), ptr(geom.y), ptr(geom.conn),
                  ptr(d_kappa), 0, ptr(d_f), ptr(d_u), ptr(inc_off), ptr(inc), ptr(rp),
                  C.byref(asm_plan.plan), ptr(jac.vals), ptr(res), stream)

    def assemble_launch():
        _lib.call("fx_p1_assemble_stencil", N, E, nnz, amap.max_row_len, ptr(geom.x),
                  ptr(geom.y), ptr(geom.conn), ptr(d_kappa), 0, ptr(d_f), ptr(d_u),
                  ptr(inc_off), ptr(inc), ptr(rp), ptr(st_plan.blk_code), ptr(st_plan.words),
                  ptr(jac.vals), ptr(res), stream)

    def step():
        A.assemble(op, geom, amap, d_u, res, jac)
        vh.axpby(-1.0, res, 0.0, rhs)
        A.prepare_forward_solve(bmap, jac, rhs, g0)
        mh.matvec_multi(jac, d_X, d_Y)

    if args.kernels_only:
        src = torch.empty(2 ** 27, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        for _ in range(5):
            assemble_rows_launch()
        for _ in range(5):
            assemble_patch_launch()
        for _ in range(5):
            assemble_launch()
        for _ in range(5):
            mh.matvec(jac, d_xs[0], d_ys[0])
        for _ in range(5):
            mh.matvec_multi(jac, d_X, d_Y, stencil=False)
        for _ in range(5):
            mh.matvec_multi(jac, d_X, d_Y)
        for _ in range(5):
            A.prepare_forward_solve(bmap, jac, rhs, g0)
        gv = A.structured_p1(1024, dev)
        xv, yv = gv.h_x, gv.h_y
        vd = A.unit_square_edge_nodes(xv, yv)
        Rv = PoissonLogKappaResidual(gv, vd)
        m = t(0.5 * np.random.default_rng(6).standard_normal(gv.num_elems))
        lam = t(np.random.default_rng(7).standard_normal(gv.num_nodes))
        uv = t(np.random.default_rng(8).standard_normal(gv.num_nodes))
        for _ in range(5):
            Rv.vjp(uv, m, lam)
        torch.cuda.synchronize()
        if not args.no_th:
            th = build_th(args.th_n, dev, t)
            for _ in range(5):
                th["assemble_lane_map"]()
            for _ in range(5):
                th["assemble"]()
            for _ in range(5):
                th["vjp"]()
            torch.cuda.synchronize()
        if not args.no_gls:
            gl = build_gls(n, dev, t)
            for _ in range(5):
                gl["assemble"]()
            for _ in range(5):
                gl["hist_vjp"]()
            for _ in range(5):
                gl["hist_vjp_lags"]()
            torch.cuda.synchronize()
        log("[bench] kernels-only done")
        return

    log("[bench] warmup")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the two step kernels timed inside the timed steps (events on the launch
    # stream around each): their durations within the step, after the other
    # kernels of the step have passed through the caches
    mk = lambda: [torch.cuda.Event(enable_timing=True) for _ in range(2)]  # noqa: E731
    ev_asm = [mk() for _ in range(args.steps)]
    ev_mv = [mk() for _ in range(args.steps)]
    log(f"[bench] timing {args.steps} steps")
    ev0.record()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev_asm[k][0].record()
        A.assemble(op, geom, amap, d_u, res, jac)
        ev_asm[k][1].record()
        vh.axpby(-1.0, res, 0.0, rhs)
        A.prepare_forward_solve(bmap, jac, rhs, g0)
        ev_mv[k][0].record()
        mh.matvec_multi(jac, d_X, d_Y)
        ev_mv[k][1].record()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    t_asm_step = float(np.mean([a.elapsed_time(b) for a, b in ev_asm])) * 1e-3
    t_mv_step = float(np.mean([a.elapsed_time(b) for a, b in ev_mv])) * 1e-3

    # Dominant kernel alone: mean duration over back-to-back launches.
    def time_kernel(fn, reps):
        for _ in range(2):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps * 1e-3

    reps = max(args.steps, 10)
    # the step's own two kernels are timed inside the timed steps above; their
    # back-to-back means only on request (a profile of the default command then
    # averages the same launches the roofline uses)
    t_asm = time_kernel(assemble_launch, reps) if args.back_to_back else None
    t_asm_rows = time_kernel(assemble_rows_launch, reps)
    t_asm_patch = time_kernel(assemble_patch_launch, reps)
    t_spmv = time_kernel(lambda: mh.matvec(jac, d_xs[0], d_ys[0]), reps)
    t_spmv10 = time_kernel(lambda: mh.matvec_multi(jac, d_X, d_Y), reps) if args.back_to_back else None
    t_spmv10_g = time_kernel(lambda: mh.matvec_multi(jac, d_X, d_Y, stencil=False), reps)
    t_bc = time_kernel(lambda: A.prepare_forward_solve(bmap, jac, rhs, g0), reps)
    # the box's achievable HBM rate: a 1 GiB device-to-device copy (read + write)
    csrc = torch.empty(2 ** 27, dtype=torch.float64, device=dev).fill_(1.0)
    cdst = torch.empty_like(csrc)
    t_copy = time_kernel(lambda: cdst.copy_(csrc), 10)
    copy_GBs = 2 * csrc.numel() * 8 / t_copy / 1e9
    del csrc, cdst
    b_asm = p1_alg_bytes(N, E, nnz)
    b_spmv = spmv_alg_bytes(N, nnz)
    kernels = {
        "p1_assemble": {"us": t_asm_step * 1e6, "alg_bytes": b_asm,
                        "GBs": b_asm / t_asm_step / 1e9, "elements_per_s": E / t_asm_step,
                        "back_to_back_us": t_asm * 1e6 if t_asm else None,
                        "form": "64-row stencil blocks (the rest row-per-lane)",
                        "rows_in_stencil_blocks": st_plan.frac},
        "p1_assemble_patches": {"us": t_asm_patch * 1e6, "alg_bytes": b_asm,
                                "GBs": b_asm / t_asm_patch / 1e9,
                                "form": "owner-computes element patches (any mesh)",
                                "patches": asm_plan.num_patches, "rows_left": asm_plan.num_left},
        "p1_assemble_row_per_lane": {"us": t_asm_rows * 1e6, "alg_bytes": b_asm,
                                     "GBs": b_asm / t_asm_rows / 1e9},
        "csr_spmv": {"us": t_spmv * 1e6, "alg_bytes": b_spmv, "GBs": b_spmv / t_spmv / 1e9},
        "csr_spmv_x10": {"us": t_mv_step * 1e6, "alg_bytes": spmv_multi_alg_bytes(N, nnz, 10),
                         "GBs": spmv_multi_alg_bytes(N, nnz, 10) / t_mv_step / 1e9,
                         "back_to_back_us": t_spmv10 * 1e6 if t_spmv10 else None,
                         "form": "64-row stencil blocks, bands by lane shifts",
                         "rows_in_stencil_blocks": jac.pattern.stencil_plan(dev)[2]},
        "csr_spmv_x10_gathered": {"us": t_spmv10_g * 1e6,
                                  "GBs": spmv_multi_alg_bytes(N, nnz, 10) / t_spmv10_g / 1e9,
                                  "form": "every entry gathered (any pattern)"},
        "bc_prepare_forward": {"us": t_bc * 1e6},
    }

    extra = {}
    if not args.no_vjp:
        nv = 1024
        log(f"[bench] VJP P1 {nv}x{nv}")
        gv = A.structured_p1(nv, dev)
        xv, yv = gv.h_x, gv.h_y
        vd = A.unit_square_edge_nodes(xv, yv)
        Rv = PoissonLogKappaResidual(gv, vd)
        m = t(0.5 * np.random.default_rng(6).standard_normal(gv.num_elems))
        lam = t(np.random.default_rng(7).standard_normal(gv.num_nodes))
        uv = t(np.random.default_rng(8).standard_normal(gv.num_nodes))
        jt = torch.empty(gv.num_nodes, dtype=torch.float64, device=dev)
        gg = torch.empty(gv.num_elems, dtype=torch.float64, device=dev)
        t_vjp = time_kernel(lambda: Rv.vjp(uv, m, lam, jt, gg), reps)
        b_vjp = vjp_alg_bytes(gv.num_nodes, gv.num_elems)
        extra["vjp_p1_1024"] = {"VJPs_per_s": 1.0 / t_vjp, "us": t_vjp * 1e6,
                                "elements_per_s": gv.num_elems / t_vjp, "alg_bytes": b_vjp,
                                "GBs": b_vjp / t_vjp / 1e9,
                                "frac": b_vjp / t_vjp / 1e9 / HBM_PEAK_GBS}

    if not args.no_th:
        log(f"[bench] Taylor-Hood {args.th_n}x{args.th_n}")
        th = build_th(args.th_n, dev, t)
        t_th = time_kernel(th["assemble"], reps)
        t_th_lm = time_kernel(th["assemble_lane_map"], reps)
        t_thv = time_kernel(th["vjp"], reps)
        b_th, b_thv = th["alg_bytes"], th["vjp_alg_bytes"]
        extra["th_ns_%d" % args.th_n] = {
            "elements": th["E"], "ndof": th["ndof"], "nnz": th["nnz"],
            "assemble_us": t_th * 1e6, "elements_per_s": th["E"] / t_th, "alg_bytes": b_th,
            "traffic": profiled_traffic("k_th_assemble_rt") if args.th_n == 512 else None,
            "GBs": b_th / t_th / 1e9, "frac": b_th / t_th / 1e9 / HBM_PEAK_GBS,
            "form": "per-block round tables, four-block walks (fx_th_assemble_rounds)",
            "lane_map_us": t_th_lm * 1e6, "lane_map_frac": b_th / t_th_lm / 1e9 / HBM_PEAK_GBS,
            "vjp_us": t_thv * 1e6, "VJPs_per_s": 1.0 / t_thv, "vjp_alg_bytes": b_thv,
            "vjp_GBs": b_thv / t_thv / 1e9, "vjp_frac": b_thv / t_thv / 1e9 / HBM_PEAK_GBS}

    if not args.no_gls:
        log(f"[bench] GLS Navier-Stokes time step {n}x{n}")
        gl = build_gls(n, dev, t)
        t_gl = time_kernel(gl["assemble"], reps)
        t_glr = time_kernel(gl["residual"], reps)
        t_glv = time_kernel(gl["hist_vjp"], reps)
        t_glv2 = time_kernel(gl["hist_vjp_lags"], reps)
        inc = 3 * gl["E"]
        fl = GLS_FLOPS_PER_INCIDENCE * inc
        extra["gls_ns_%d" % n] = {
            "elements": gl["E"], "states": gl["states"], "nnz": gl["nnz"],
            "assemble_us": t_gl * 1e6, "elements_per_s": gl["E"] / t_gl,
            "residual_us": t_glr * 1e6, "residual_elements_per_s": gl["E"] / t_glr,
            "alg_bytes": gl["alg_bytes"], "GBs": gl["alg_bytes"] / t_gl / 1e9,
            "frac_hbm": gl["alg_bytes"] / t_gl / 1e9 / HBM_PEAK_GBS,
            "flops": fl, "TFLOPs": fl / t_gl / 1e12,
            "frac_fp64": fl / t_gl / 1e12 / FP64_PEAK_TFLOPS,
            "hist_vjp_us": t_glv * 1e6, "hist_vjp_elements_per_s": gl["E"] / t_glv,
            "hist_vjp_alg_bytes": gl["vjp_alg_bytes"],
            "hist_vjp_GBs": gl["vjp_alg_bytes"] / t_glv / 1e9,
            "hist_vjp_both_lags_us": t_glv2 * 1e6}
        if not args.no_cpu:
            from oracle import oracle as O
            ns_ = 256  # bounded sample: the 256^2 mesh of the same problem
            xs_ = np.empty((ns_ + 1) ** 2)
            ys_ = np.empty_like(xs_)
            cs_ = np.empty(6 * ns_ * ns_, np.int32)
            _lib.call("fx_mesh_p1", ns_, ptr(xs_), ptr(ys_), ptr(cs_))
            from femx_amd import gls as G
            gs_ = G.GLSGeometry(xs_, ys_, cs_, "cpu")
            ms_ = G.make_navier_map(gs_)
            _, jms = ms_.jac()
            nds = gs_.num_states
            hs_ = np.random.default_rng(12).standard_normal(2 * nds)
            nxs = np.random.default_rng(13).standard_normal(nds)
            args_ = (xs_, ys_, cs_, gs_.h_dofs9.reshape(-1), nds, 1.0, 1e-3, 0.01, 1, 2, hs_, nxs,
                     jms, ms_.pattern.nnz)
            O.gls_assemble(*args_, omp=True)
            ts = []
            t_start = time.perf_counter()
            while len(ts) < 5 and time.perf_counter() - t_start < 8.0:
                a = time.perf_counter()
                O.gls_assemble(*args_, omp=True)
                ts.append(time.perf_counter() - a)
            tcg = float(np.median(ts))
            extra["gls_ns_%d" % n]["cpu_baseline"] = {
                "value": gs_.num_elems / tcg, "unit": "elements/s", "cores": O.omp_threads(),
                "kind": "port",
                "sample": f"{len(ts)} assemblies of the same problem on the {ns_}^2 mesh "
                          f"({gs_.num_elems} elements), median; OpenMP element loop with "
                          "atomic scatter as femx's NavierStokesModel"}

    if not args.no_batch:
        extra["batched_%d" % args.batch_count] = run_batched(args, log)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        log("[bench] CPU baseline (oracle OpenMP assembly + host elimination + 10 matvec)")
        _, jm = amap.jac()
        bc = dict(bc_rows=bmap.bc_rows, diag=bmap.diag, col_offsets=bmap.col_offsets,
                  col_entries=bmap.col_entries, col_rows=bmap.col_rows, bc_mask=bmap.bc_mask)
        conn = geom.h_conn.reshape(-1)
        rp_h, ci_h = amap.pattern.row_ptr, amap.pattern.col_ind
        yk = np.empty(N)

        def cpu_step():
            vals, r = O.p1_assemble(x, y, conn, kappa, f, u, jm, nnz, omp=True)
            rr = -r
            O.prepare_forward(rp_h, bc, vals, rr, np.zeros(dofs.size))
            for k in range(10):
                O.matvec(rp_h, ci_h, vals, xs[k], yk)

        cpu_step()
        times = []
        t_start = time.perf_counter()
        while len(times) < 5 and (time.perf_counter() - t_start) < 12.0:
            a = time.perf_counter()
            cpu_step()
            times.append(time.perf_counter() - a)
        tc = float(np.median(times))
        cpu = {"value": E / tc, "unit": "elements/s", "cores": O.omp_threads(), "kind": "port",
               "sample": f"{len(times)} full steps of the same workload (P1 {n}^2, "
                         f"{E} elements) after 1 warm-up, median; OpenMP assembly on "
                         f"{O.omp_threads()} threads, serial elimination and matvec",
               "ms_per_step": tc * 1e3}

    # The headline roofline is the step's longest kernel; the other kernel of
    # the step and the Taylor-Hood and batched figures ride beside it.
    b_spmv10 = spmv_multi_alg_bytes(N, nnz, 10)
    step_kernels = {
        "k_p1_asm_stencil": (b_asm, t_asm_step),
        "k_spmv_multi_st<10>": (b_spmv10, t_mv_step),
    }

    def roof(name):
        b, tk = step_kernels[name]
        key = name.split("<")[0]
        return {"achieved": b / tk / 1e9, "frac": b / tk / 1e9 / HBM_PEAK_GBS, "us": tk * 1e6,
                "alg_bytes": b, "traffic": profiled_traffic(key) if n == 2048 else None}

    dom = max(step_kernels, key=lambda k: step_kernels[k][1])
    head = roof(dom)
    roofline = {"bound": "hbm", "achieved": head["achieved"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": head["frac"], "traffic": head["traffic"],
                "kernel": dom, "us": head["us"], "alg_bytes": head["alg_bytes"],
                "also": {k: roof(k) for k in step_kernels if k != dom},
                "copy_GBs": copy_GBs, "frac_of_copy": head["achieved"] / copy_GBs,
                "traffic_source": "profiles/hbm_traffic_p1_2048_th_512_gls_2048.md"}
    th_key = "th_ns_%d" % args.th_n
    if th_key in extra:
        roofline["th_assemble_frac"] = extra[th_key]["frac"]
        roofline["th_vjp_frac"] = extra[th_key]["vjp_frac"]
    b_key = "batched_%d" % args.batch_count
    if b_key in extra:
        roofline["speedup_graph_vs_plain"] = extra[b_key]["speedup_graph_vs_plain"]

    out = {
        "metric": METRIC,
        "value": E / (ms_step * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"P1 Poisson {n}x{n} structured triangles: Jacobian+residual+load "
                               "assembly, Dirichlet elimination, 10 SpMV",
                   "elements": E, "nodes": N, "nnz": nnz, "spmv_per_step": 10},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "extra": extra,
        "wall_s_timed": wall,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
