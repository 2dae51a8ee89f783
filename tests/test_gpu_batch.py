"""Batched persistent launches and their hipGraph replay against the plain
per-instance calls: every instance's outputs must be bitwise equal, on every
replay, with eager work between replays."""
import numpy as np
import pytest
import torch

from femx_amd import batch as B

pytestmark = pytest.mark.gpu


def snapshot(cases):
    torch.cuda.synchronize()
    return [[t.detach().cpu().numpy().copy() for t in c.outputs()] for c in cases]


def clear(cases):
    for c in cases:
        for t in c.outputs():
            t.fill_(float("nan"))


def same(a, b):
    return all(np.array_equal(x, y, equal_nan=False) for xa, ya in zip(a, b) for x, y in zip(xa, ya))


@pytest.mark.parametrize("grid_per_cu,rounds", [(1, True), (3, True), (1, False)])
def test_batch_matches_plain_bitwise(grid_per_cu, rounds):
    """The batch's outputs equal the plain calls' bit for bit, with its
    Taylor-Hood assembly phase walking the round tables or taking one block
    per work item."""
    specs = B.instance_specs(count=10, seed=9, n_min=3, n_max=40)
    cases = B.build_cases(specs, "cuda")
    for c in cases:
        if isinstance(c, B.THCase):
            c.rounds = rounds
    B.plain(cases, 3)
    ref = snapshot(cases)
    assert all(np.isfinite(x).all() for xa in ref for x in xa)
    b = B.Batch(cases, grid_per_cu)
    counts, grid = b.work()
    assert grid > 0 and all(c > 0 for c in counts[:6]) and counts[6:] == [0, 0, 0]
    clear(cases)
    b.run(3)
    assert same(snapshot(cases), ref)
    b.capture(3)
    for k in range(3):
        clear(cases)
        b.launch()
        assert same(snapshot(cases), ref), f"graph replay {k}"
        B.plain(cases[:2], 1)  # eager work between replays


def test_batch_assembly_only_and_autotune():
    specs = B.instance_specs(count=6, seed=3, n_min=4, n_max=20)
    cases = B.build_cases(specs, "cuda")
    B.plain(cases, 1)
    ref = [[t.detach().cpu().numpy().copy() for t in c.outputs()[:2]] for c in cases]
    b = B.Batch(cases, 2)
    for c in cases:
        for t in c.outputs()[:2]:
            t.fill_(float("nan"))
    b.capture(1)
    b.launch()
    torch.cuda.synchronize()
    got = [[t.detach().cpu().numpy() for t in c.outputs()[:2]] for c in cases]
    assert same(got, ref)
    runner, report = B.autotune(cases, layouts=B.TH_LAYOUTS[:2], grids=(1, 2), reps=2)
    for g in report["groups"].values():
        assert g["grid_per_cu"] in (1, 2) and all(t > 0 for t in g["grid_ms"].values())


def test_instance_specs_follow_config5():
    s = B.instance_specs()
    assert len(s) == 256
    assert all(x["kind"] == ("p1" if i % 2 == 0 else "th") for i, x in enumerate(s))
    ns = np.array([x["n"] for x in s])
    assert ns.min() >= 16 and ns.max() <= 512
    assert all(10.0 <= x["Re"] <= 400.0 for x in s if x["kind"] == "th")


def test_grouped_graph_and_autotune_bitwise():
    """The autotuner's product -- one hipGraph over per-size-class batches
    or one batch, chains in series or concurrent, chosen block layouts and
    grids (including one workgroup per work item) -- is bitwise equal to the
    plain calls, replay after replay; every shape it tries is too."""
    specs = B.instance_specs(count=12, seed=5, n_min=3, n_max=90)
    cases = B.build_cases(specs, "cuda")
    assert len({B.size_class(c) for c in cases}) >= 2
    B.plain(cases, 3)
    ref = snapshot(cases)
    runner, report = B.autotune(cases, grids=(0, 3), reps=1)
    assert report["groups"] and all("grid_per_cu" in g for g in report["groups"].values())
    assert report["graph"] in report["graph_ms"] and len(report["graph_ms"]) == 4
    for k in range(3):
        clear(cases)
        runner.launch()
        assert same(snapshot(cases), ref), f"tuned graph replay {k}"
        B.plain(cases[:1], 1)
    # every candidate graph shape: per-class batches and one batch, chains
    # in series and as concurrent branches
    th = [c for c in cases if isinstance(c, B.THCase)]
    p1 = [c for c in cases if isinstance(c, B.P1Case)]
    parts = [B.Batch(p1, 2), B.Batch(th, 0)]
    for what in (3, 3 | B.Batch.SERIAL):
        for bs in (parts, [B.Batch(cases, 1)]):
            g = B.BatchGraph(bs, what)
            for k in range(2):
                clear(cases)
                g.launch()
                assert same(snapshot(cases), ref), (what, len(bs), k)
    for lay in B.TH_LAYOUTS:
        for c in cases:
            if isinstance(c, B.THCase):
                c.plan.select(*lay)
        for g, what in ((0, 3), (2, 3 | B.Batch.SERIAL)):
            b = B.Batch(cases, g)
            clear(cases)
            b.run(what)
            assert same(snapshot(cases), ref), (lay, g, what)


def test_batch_with_gls_instances_bitwise():
    """GLS P1/P1 Navier-Stokes time steps (assembly and the both-lags history
    VJP; one instance with a single history lag, so step 0 and no lag-1
    output) batched beside P1 and Taylor-Hood instances: bitwise equal to
    fx_gls_assemble and fx_gls_hist_vjp_patch, through persistent launches,
    one workgroup per item, graph replays and the autotuner's graph."""
    specs = B.instance_specs(count=4, seed=9, n_min=3, n_max=20)
    specs += [dict(kind="gls", n=n, seed=300 + n, num_hist=h)
              for n, h in ((5, 2), (23, 2), (2, 1), (40, 2))]
    cases = B.build_cases(specs, "cuda")
    gl = [c for c in cases if isinstance(c, B.GLSCase)]
    assert len(gl) == 4 and gl[2].out1 is None
    B.plain(cases, 3)
    ref = snapshot(cases)
    assert all(np.isfinite(x).all() for xa in ref for x in xa)
    for grid in (0, 2):
        b = B.Batch(cases, grid)
        counts, _ = b.work()
        assert all(c > 0 for c in counts[6:8])
        clear(cases)
        b.run(3)
        assert same(snapshot(cases), ref), grid
        clear(cases)
        b.run(3 | B.Batch.SERIAL)
        assert same(snapshot(cases), ref), grid
    b = B.Batch(gl, 1)
    b.capture(3)
    for k in range(3):
        for c in gl:
            for t in c.outputs():
                t.fill_(float("nan"))
        b.launch()
        assert same(snapshot(gl), [ref[cases.index(c)] for c in gl]), f"graph replay {k}"
        B.plain(cases[:1], 1)
    runner, report = B.autotune(cases, layouts=B.TH_LAYOUTS[:1], grids=(0, 2), reps=1)
    assert any(k.startswith("gls_") for k in report["groups"])
    clear(cases)
    runner.launch()
    assert same(snapshot(cases), ref)
