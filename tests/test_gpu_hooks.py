"""The outer-face hook on the GPU (tah_fill_boundary through fill_ghosts(...,
boundary=...)): bit for bit the oracle's fill with the same hook values, on
meshes whose coarse-fine faces meet the outer faces, for the values form and
the relative form, cell and vertex centering, dense and arena layouts, with
the launch options that move transfers between stage groups (defer) or run
them on several streams (fork); one fill from NaN ghosts is the relative
condition's fixed point; a forest without outer faces runs as the plain
fill."""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import (AllVariables, CellBoundary, FieldSet, Forest, GhostSchedule, Operators, Relative,
                         boundary_by_coordinates, boundary_values, cellcentered, fill_ghosts, vertexcentered)
from treeamr_hip import workloads as W

from test_hooks_host import case_fn, nested_forest

pytestmark = pytest.mark.gpu


def _hooks(D, nvars):
    f = case_fn(D)
    if nvars == 1:
        rel = CellBoundary(Relative(f))
    else:
        rel = CellBoundary(Relative(AllVariables(lambda x: [f(x) * (v + 1) for v in range(nvars)])))
    return {"relative": rel,
            "values": CellBoundary(lambda x, v, d: torch.cos(x[0] * v) + d[-1] * 0.25),
            "by-coordinates": boundary_by_coordinates(lambda x, v: x[0] - 2.0 * v)}


def _against_oracle(fs, s, hook, what):
    init = fs.work.clone()
    kind, values, fr = boundary_values(fs, s, hook)
    host = init.cpu().numpy().copy()
    fac = fs.factors.cpu().numpy().copy() if fs.factors is not None else None
    oracle.fill_boundary(s._descptr, fs.layout, host, fac, kind, values.cpu().numpy(),
                         fr.cpu().numpy() if fr is not None else None, nthreads=8)
    fill_ghosts(fs, s, boundary=hook)
    torch.cuda.synchronize()
    got = fs.work.cpu().numpy()
    bad = np.nonzero(got.view(np.int64) != host.view(np.int64))[0]
    assert bad.size == 0, f"{what}: {bad.size} values differ, first at {bad[:5]}"
    fs.work.copy_(init)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("C", ["cell", "vertex"])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_hook_fill_equals_oracle(D, C, layout):
    forest = nested_forest(D)
    cen = cellcentered(D) if C == "cell" else vertexcentered(D)
    for nvars in (1, 2):
        fs = FieldSet(forest, nvars, G=2, centering=cen, device="cuda", layout=layout, fill=float("nan"))
        W.fill_uniform(fs, 7 + nvars)
        s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
        assert s.desc.nboundaries > 0
        print(f"[hooks] D={D} {C} {layout} nvars={nvars}: {s.desc.boundary_points} hook points", flush=True)
        for name, hook in _hooks(D, nvars).items():
            _against_oracle(fs, s, hook, f"{name} nvars={nvars}")


@pytest.mark.parametrize("C", ["cell", "vertex"])
def test_hook_fill_under_launch_options(C):
    """Deferral of phase-1 transfers into the last stage group must not move
    a transfer past the relative hook that reads what it writes; forked
    launches, unmerged classes and LDS-staged interpolation alike."""
    forest = nested_forest(3, N=16)
    cen = cellcentered(3) if C == "cell" else vertexcentered(3)
    fs = FieldSet(forest, 4, G=3, centering=cen, device="cuda", layout="arena", fill=float("nan"))
    W.fill_uniform(fs, 11)
    s = GhostSchedule(fs, Operators(prolongation=6, restriction=6))
    plan = s.plan(fs.layout)
    hook = _hooks(3, 4)["relative"]
    for opts in (dict(defer=1), dict(defer=0), dict(fork=1, defer=1), dict(merge=0, copy_per_thread=4),
                 dict(lds=3, defer=1), dict(lds=0, defer=1), dict(lds=4, fork=1, defer=0)):
        plan.set_options(**opts)
        print(f"[hooks options] {C} {opts}", flush=True)
        _against_oracle(fs, s, hook, str(opts))
    plan.set_options(fork=0, merge=1, defer=1, copy_per_thread=8, lds=4)


@pytest.mark.parametrize("C", ["cell", "vertex"])
def test_hook_fill_is_the_relative_fixed_point(C):
    forest = nested_forest(3)
    cen = cellcentered(3) if C == "cell" else vertexcentered(3)
    fs = FieldSet(forest, 1, G=2, centering=cen, device="cuda", fill=float("nan"))
    W.fill_uniform(fs, 5)
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    hook = CellBoundary(Relative(case_fn(3)))
    fill_ghosts(fs, s, boundary=hook)
    torch.cuda.synchronize()
    assert not torch.isnan(fs.array()).any()
    once = fs.work.clone()
    fill_ghosts(fs, s, boundary=hook)
    torch.cuda.synchronize()
    assert torch.equal(fs.work.view(torch.int64), once.view(torch.int64))


def test_no_outer_faces_runs_as_the_fill():
    forest = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    fs = FieldSet(forest, 2, G=2, device="cuda", fill=float("nan"))
    W.fill_uniform(fs, 2)
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    assert s.desc.nboundaries == 0
    init = fs.work.clone()
    fill_ghosts(fs, s)
    ref = fs.work.clone()
    fs.work.copy_(init)
    fill_ghosts(fs, s, boundary=CellBoundary(lambda x, v, d: x[0]))
    torch.cuda.synchronize()
    assert torch.equal(fs.work.view(torch.int64), ref.view(torch.int64))
