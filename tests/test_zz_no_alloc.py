"""tah_fill_adjoint_noalloc, the adjoint that allocates nothing (its passes
stage their values in a buffer compiled into the library), against
tah_fill_adjoint bit for bit, through the C ABI: two calls in a row, the
second on a smaller mesh, in both layouts; then a third call back on the
first plan from another stream. Also its refusals: a plan whose adjoint
tables are not built yet, and one whose adjoint stages more than the
buffer's 8 Mi doubles."""
import ctypes as C

import numpy as np
import pytest
import torch

from treeamr_hip import FieldSet, GhostSchedule, fill_ghosts_adjoint
from treeamr_hip import workloads as W
from treeamr_hip._abi import TAH_EINVAL, TAH_EUNSUPPORTED, lib
from treeamr_hip.core import _stream_ptr, default_context

pytestmark = pytest.mark.gpu

MAX_STAGED = 8 << 20


def _mesh(which, layout):
    if which == "config1":
        f, ops, kw, _ = W.config1()
    elif which == "small":
        # 2^3 periodic roots of 8^3 refined once, then one leaf again:
        # 71 blocks, cell, G=2, PointValue(4,4)
        from treeamr_hip import Forest, Operators
        f = Forest((2, 2, 2), N=8, periodic=(True,) * 3, extents=((0.0, 1.0),) * 3)
        f.refine(range(f.nleaves))
        f.refine([0])
        f.balance()
        ops = Operators(prolongation=4, restriction=4)
        kw = dict(nvars=4, G=2)
    else:  # a reflecting vertex mesh
        f, ops, kw, _ = W.config2(nvars=4, roots=2, N=8, radii=(0.5, 0.3))
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout=layout, fill=0.0, **kw)
    return fs, GhostSchedule(fs, ops)


def _adjoint_input(fs, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1


def _noalloc(fs, sched, adj, stream=None):
    plan = sched.plan(fs.layout)
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    s = C.c_void_p(stream.cuda_stream) if stream is not None else _stream_ptr(fs)
    return lib().tah_fill_adjoint_noalloc(plan._h, C.c_void_p(adj.data_ptr()), fptr, s)


def _staged_values(fs, sched):
    """The adjoint's staging size, as the library computes it: per stage, the
    outputs of the z, y and x passes of every transfer that is not a unit
    copy with steps of at most one row."""
    best = 0
    per_stage = {}
    for g in sched.groups():
        st = g["stencils"]
        n = [s["ntarget"] for s in st]
        if min(n) == 0 or g["ntransfers"] == 0:
            continue
        copy = all(s["order"] == 1 and np.all(s["weights"] == 1.0) for s in st)
        if copy:
            for s in st:
                ss = s["srcstart"]
                if len(ss) > 1:
                    step = ss[1] - ss[0]
                    copy = copy and abs(step) <= 1 and np.array_equal(ss, ss[0] + step * np.arange(len(ss)))
        if copy:
            continue
        e = [int(s["srcstart"].max()) + s["order"] - int(s["srcstart"].min()) for s in st]
        sz = n[0] * n[1] * e[2] + n[0] * e[1] * e[2] + e[0] * e[1] * e[2]
        per_stage[g["stage"]] = per_stage.get(g["stage"], 0) + sz * fs.nvars * g["ntransfers"]
    for v in per_stage.values():
        best = max(best, v)
    return best


@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_noalloc_adjoint_equals_adjoint(layout):
    cases = [_mesh("config1", layout), _mesh("small", layout)]
    assert cases[1][0].nelements < cases[0][0].nelements  # the second call's input is smaller
    for k, (fs, sched) in enumerate(cases):
        print(f"[noalloc] {layout} call {k + 1}: {fs.nblocks} blocks, {fs.nvars} vars", flush=True)
        assert _staged_values(fs, sched) <= MAX_STAGED
        y = _adjoint_input(fs, 10 + k)
        ref = y.clone()
        fill_ghosts_adjoint(fs, sched, ref)  # the existing entry point (builds the tables)
        got = y.clone()
        assert _noalloc(fs, sched, got) == 0
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.int64), ref.view(torch.int64)), f"call {k + 1}"
    # back on the first plan, from another stream: ordered after the calls before
    fs, sched = cases[0]
    y = _adjoint_input(fs, 30)
    ref = y.clone()
    fill_ghosts_adjoint(fs, sched, ref)
    got = y.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        assert _noalloc(fs, sched, got, stream=side) == 0
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int64), ref.view(torch.int64)), "call 3 (side stream)"


def test_noalloc_adjoint_reflecting_vertex():
    fs, sched = _mesh("vertex", "arena")
    y = _adjoint_input(fs, 7)
    ref = y.clone()
    fill_ghosts_adjoint(fs, sched, ref)
    got = y.clone()
    assert _noalloc(fs, sched, got) == 0
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int64), ref.view(torch.int64))


def test_noalloc_adjoint_refusals():
    # no tables yet: refused before any launch, the array untouched
    fs, sched = _mesh("small", "dense")
    y = _adjoint_input(fs, 3)
    got = y.clone()
    assert _noalloc(fs, sched, got) == TAH_EINVAL
    assert b"not built" in lib().tah_last_error()
    torch.cuda.synchronize()
    assert torch.equal(got, y)
    # a mesh whose adjoint stages more than the buffer holds
    f, ops, kw, _ = W.config5(nvars=20, roots=4, N=16)
    kw = dict(kw)
    nv = kw.pop("nvars")
    big = FieldSet(f, nv, device="cuda", layout="arena", fill=0.0, **kw)
    bs = GhostSchedule(big, ops)
    need = _staged_values(big, bs)
    assert need > MAX_STAGED, need
    print(f"[noalloc] refusal mesh: {big.nblocks} blocks, {need} staged values", flush=True)
    yb = _adjoint_input(big, 4)
    fill_ghosts_adjoint(big, bs, yb.clone())  # builds the tables
    got = yb.clone()
    assert _noalloc(big, bs, got) == TAH_EUNSUPPORTED
    assert b"64 MiB" in lib().tah_last_error()
    torch.cuda.synchronize()
    assert torch.equal(got, yb)
    default_context()  # the context that set the buffer up is still alive
