"""The batched workload at its full size: 256 instances, n_i ~ logU[16, 512]
(seed 9), even instances P1 Poisson with log-kappa control, odd instances
Taylor-Hood cavity flow with Re_i ~ U[10, 400]; assembly + VJP of every
instance.

* The persistent launches and their hipGraph replays are bitwise equal to
  the plain per-instance calls, for all four outputs of every instance
  (compared on the device, bit pattern by bit pattern).
* The plain path matches the oracle on the largest P1 and the largest
  Taylor-Hood instance and on 16 more drawn with a fixed seed: P1 CSR
  values, residual and VJP within the 1e-12 mixed norm (kappa = exp(m) is
  taken on the device, so one ulp of exp may differ from the host's), TH
  within the same bar (fused multiply-adds)."""
import numpy as np
import pytest
import torch

from conftest import close
from femx_amd import batch as B
from femx_amd.assembly import make_assembly_map
from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def bits(t):
    return t.detach().contiguous().view(torch.int64)


def host(t):
    return t.detach().cpu().numpy()


def check_p1(case):
    g = case.geom
    amap = make_assembly_map(g.num_nodes, g.num_nodes, g.h_conn)
    _, jm = amap.jac()
    m, f, u, lam = (host(a) for a in (case.m, case.f, case.u, case.lam))
    ov, orr = O.p1_assemble(g.h_x, g.h_y, g.h_conn.reshape(-1), np.exp(m), f, u, jm,
                            amap.pattern.nnz, omp=True)
    ojt, og = O.p1_vjp(g.h_x, g.h_y, g.h_conn.reshape(-1), m, u, lam, case.R.bmap.bc_mask)
    vals, res, jt, ge = (host(t) for t in case.outputs())
    assert close(vals, ov) and close(res, orr), f"P1 n={g.num_elems} assembly"
    assert close(jt, ojt) and close(ge, og), f"P1 n={g.num_elems} VJP"


def check_th(case):
    g = case.geom
    amap = make_assembly_map(g.ndof, g.ndof, g.h_dofs15)
    _, jm = amap.jac()
    s, lam = host(case.state), host(case.lam)
    ov, orr = O.th_assemble(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1),
                            g.ndof, case.nu, s, jm, amap.pattern.nnz, omp=True)
    ojt, og = O.th_vjp(g.h_xv, g.h_yv, g.h_vconn3.reshape(-1), g.h_dofs15.reshape(-1), g.ndof,
                       case.nu, s, lam, case.bmap.bc_mask)
    vals, res, jt, ge = (host(t) for t in case.outputs())
    assert close(vals, ov) and close(res, orr), f"TH E={g.num_elems} assembly"
    assert close(jt, ojt) and close(ge, og), f"TH E={g.num_elems} VJP"


def test_config5_full_size_batch_bitwise_and_oracle():
    specs = B.instance_specs(count=256, seed=9)
    cases = B.build_cases(specs, "cuda")
    assert len(cases) == 256
    B.plain(cases, 3)
    torch.cuda.synchronize()
    ref = [[t.detach().clone() for t in c.outputs()] for c in cases]
    assert all(bool(torch.isfinite(t).all()) for r in ref for t in r)

    def same_as_plain(tag):
        torch.cuda.synchronize()
        for i, (c, r) in enumerate(zip(cases, ref)):
            for k, (a, b) in enumerate(zip(c.outputs(), r)):
                assert torch.equal(bits(a), bits(b)), f"{tag}: instance {i} output {k}"

    def clear():
        for c in cases:
            for t in c.outputs():
                t.fill_(float("nan"))

    b = B.Batch(cases, 8)
    clear()
    b.run(3)
    same_as_plain("persistent launches")
    b.capture(3)
    for rep in range(2):
        clear()
        b.launch()
        same_as_plain(f"graph replay {rep}")
        cases[0].plain(1)  # eager work between replays
    # the benchmark's form: autotuned per-size-class batches in one graph
    runner, _ = B.autotune(cases, reps=1)
    for rep in range(2):
        clear()
        runner.launch()
        same_as_plain(f"tuned graph replay {rep}")
        cases[0].plain(1)
    del ref

    p1 = [c for c in cases if isinstance(c, B.P1Case)]
    th = [c for c in cases if isinstance(c, B.THCase)]
    picks = {max(p1, key=lambda c: c.elements), max(th, key=lambda c: c.elements)}
    rng = np.random.default_rng(2024)
    for i in rng.choice(len(cases), 16, replace=False):
        picks.add(cases[int(i)])
    # restore the plain outputs of the checked instances (replays overwrote them
    # bitwise identically, but the eager work between replays touched case 0)
    for c in picks:
        c.plain(3)
    torch.cuda.synchronize()
    for c in sorted(picks, key=lambda c: c.elements):
        if isinstance(c, B.P1Case):
            check_p1(c)
        else:
            check_th(c)
