"""fx_th_vjp_noalloc (the library's fixed device workspace) against
fx_th_vjp (a caller workspace), through the C-ABI: bitwise equal J^T lam and
per-element viscosity sensitivities, on two calls in a row, the second on a
smaller mesh (the workspace keeps the first call's contents)."""
import numpy as np
import pytest
import torch

from femx_amd import _lib, ns
from femx_amd.assembly import make_assembly_map, make_boundary_map
from femx_amd.linalg import ptr, stream_handle

pytestmark = pytest.mark.gpu
DEV = "cuda"


def run_both(n, seed):
    print(f"noalloc TH VJP n={n} seed={seed}", flush=True)
    g = ns.structured_th(n, DEV)
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    dofs, _ = ns.cavity_bc(g)
    bmap = make_boundary_map(dofs, amap.pattern, DEV)
    p = ns.THPlan(amap, DEV)
    rng = np.random.default_rng(seed)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    s = t(ns.fourier_state(g, seed))
    lam = t(rng.standard_normal(g.ndof))
    nu_e = t(rng.uniform(0.005, 0.05, g.num_elems))
    assert g.num_elems <= 559240
    out = {}
    for name in ("fx_th_vjp", "fx_th_vjp_noalloc"):
        jt = torch.full((g.ndof,), float("nan"), dtype=torch.float64, device=DEV)
        ge = torch.full((g.num_elems,), float("nan"), dtype=torch.float64, device=DEV)
        args = [g.ndof, g.num_elems, ptr(g.xv), ptr(g.yv), ptr(g.vconn3), ptr(g.dofs15), ptr(s),
                0.0, ptr(nu_e), ptr(lam), ptr(bmap.d_mask), ptr(p.inc_off), ptr(p.elem_inc),
                ptr(jt), ptr(ge)]
        if name == "fx_th_vjp":
            work = torch.empty(15 * g.num_elems, dtype=torch.float64, device=DEV)
            args.append(ptr(work))
        args.append(stream_handle())
        _lib.call(name, *args)
        torch.cuda.synchronize()
        out[name] = (jt.cpu().numpy(), ge.cpu().numpy())
    a, b = out["fx_th_vjp"], out["fx_th_vjp_noalloc"]
    assert np.isfinite(a[0]).all() and np.isfinite(a[1]).all()
    assert np.array_equal(a[0].view(np.int64), b[0].view(np.int64)), f"J^T lam, n={n}"
    assert np.array_equal(a[1].view(np.int64), b[1].view(np.int64)), f"g, n={n}"


def test_noalloc_vjp_equals_workspace_vjp_on_two_calls():
    run_both(96, 3)
    run_both(17, 4)  # smaller: the workspace still holds the first call's columns
