"""The benchmarked workloads at full size against the oracle, bit for bit:

* the 64-forest batch exactly as bench.py builds it (seeds 2..65, arena
  layout, 131 field sets, 38,460 blocks), filled by one batched fill and by
  replays of its hipGraph;
* the large single forest in the headline's form (5209 blocks of 32^3, G=3,
  20 variables, PointValue(6,6), reflecting z-low face, arena layout, the
  graph-captured fill) with the default launch options and with forked
  launches, copies from a bounded grid and copy tiles of 16;
* the adjoint of the large single forest (2 variables) by the dot-product
  identity.

Host arrays are compared block range by block range to keep host memory
near one working array.
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import FieldSet, FillGraph, GhostSchedule, fill_ghosts, fill_ghosts_adjoint, fill_ghosts_batched
from treeamr_hip import autodiff as AD
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_expected(fs, sched, init):
    host = init.cpu().numpy()
    fac = fs.factors.cpu().numpy().copy() if fs.factors is not None else None
    oracle.fill(sched._descptr, fs.layout, host, fac, nthreads=0)
    return host


def _compare(got_dev, expected_host, chunk=1 << 27):
    """Bits of a device array against a host array, chunk by chunk: the
    number of values that differ."""
    bad = 0
    n = expected_host.size
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        g = got_dev[lo:hi].cpu().numpy().view(np.int64)
        bad += int(np.count_nonzero(g != expected_host[lo:hi].view(np.int64)))
    return bad


@pytest.mark.timeout(900)
def test_config3_batched_and_graph_against_oracle():
    pairs = W.config3(count=64, layout="arena")
    nvals = sum(s.points() * fs.nvars for fs, s in pairs)
    print(f"[fullsize] config 3: {len(pairs)} field sets, {nvals} values per fill", flush=True)
    assert len(pairs) == 131
    init = [fs.work.clone() for fs, _ in pairs]
    fill_ghosts_batched(pairs)
    torch.cuda.synchronize()
    res = [fs.work.clone() for fs, _ in pairs]
    bad = []
    for i, ((fs, s), i0) in enumerate(zip(pairs, init)):
        expected = _oracle_expected(fs, s, i0)
        nb = _compare(res[i], expected)
        if nb:
            bad.append((i, nb))
        if i % 16 == 0:
            print(f"[fullsize] config 3: field set {i} checked", flush=True)
    assert not bad, f"batched fill differs from the oracle: {bad}"
    graph = FillGraph(pairs)
    for rep in range(2):
        for (fs, _), i0 in zip(pairs, init):
            fs.work.copy_(i0)  # eager work between replays
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        for i, ((fs, _), r) in enumerate(zip(pairs, res)):
            assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64)), f"replay {rep}, field set {i}"
    print("[fullsize] config 3: graph replays equal the batched fill", flush=True)


def _large_forest(nvars):
    f, ops, kw, _ = W.config5(nvars=nvars)
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout="arena", fill=0.0, **kw)
    W.fill_gaussians(fs, 5)
    return fs, GhostSchedule(fs, ops)


@pytest.mark.timeout(900)
def test_large_forest_headline_form_against_oracle():
    fs, s = _large_forest(20)
    print(f"[fullsize] large forest: {fs.nblocks} blocks, {s.points() * fs.nvars} values per fill", flush=True)
    init = fs.work.clone()
    expected = _oracle_expected(fs, s, init)
    print("[fullsize] large forest: oracle done", flush=True)
    plan = s.plan(fs.layout)
    for opts in (dict(fork=0, merge=1, defer=1, copy_per_thread=8, copy_wgs=0, interp_pairs=1),
                 dict(fork=1, merge=1, defer=1, copy_per_thread=16, copy_wgs=512, interp_pairs=1),
                 dict(fork=1, merge=1, defer=0, copy_per_thread=16, copy_wgs=0, interp_pairs=1),
                 dict(fork=1, merge=1, defer=0, copy_per_thread=8, copy_wgs=0, interp_pairs=1, copy_order=1)):
        plan.set_options(**opts)
        graph = FillGraph([(fs, s)])
        for form, run in (("staged", lambda: fill_ghosts(fs, s)), ("graph", graph.launch)):
            fs.work.copy_(init)
            torch.cuda.synchronize()
            run()
            torch.cuda.synchronize()
            nb = _compare(fs.work, expected.reshape(-1))
            print(f"[fullsize] large forest {form} {opts}: {nb} values differ", flush=True)
            assert nb == 0, (form, opts)
        del graph


@pytest.mark.timeout(600)
def test_large_forest_adjoint_dot_product():
    fs, s = _large_forest(2)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    own = AD.own_mask(fs)
    x = torch.where(own, torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1,
                    torch.zeros_like(fs.work))
    y = torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    fx = x.clone()
    AD._fill_tensor(fs, s, fx)
    aty = y.clone()
    fill_ghosts_adjoint(fs, s, aty)
    torch.cuda.synchronize()
    lhs = float(torch.dot(fx, y))
    rhs = float(torch.dot(x, aty))
    print(f"[fullsize] adjoint dot: {lhs!r} {rhs!r}", flush=True)
    assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), 1.0), (lhs, rhs)
