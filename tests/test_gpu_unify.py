"""unify! on the GPU (stages P and 0 alone, src/ghosts.jl:1836-1906) against
the oracle on the same descriptors, bit for bit: on the working array and on
a state vector, with a GhostSchedule and with a UnifySchedule (no ghosts), on
vertex sets with a reflecting wall (stage P) and on a face-centred set.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import (EvenParity, FieldSet, Forest, GhostSchedule, OddParity, Operators, UnifySchedule,
                         gather, ownextents, scatter, statelength, unify, unify_desc)
from treeamr_hip import _abi
from treeamr_hip.core import _own_first
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _bits(t):
    return t.detach().cpu().contiguous().view(torch.int64)


def _state_layout(desc, nvars):
    L = _abi.tah_layout()
    _abi.check(_abi.lib().tah_layout_dense(desc._descptr, nvars, C.byref(L)))
    return L


@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_unify_on_the_working_array_and_on_a_state(layout):
    """config 2 at a small size: nested vertex mesh, G=3, reflecting z-low
    with odd variables, PointValue(6,6)."""
    f, ops, kw, _ = W.config2(roots=2, N=8, radii=(0.5, 0.3))
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout=layout, fill=0.0, **kw)
    W.fill_uniform(fs, 7)
    s = GhostSchedule(fs, ops)
    fac = fs.factors.cpu().numpy()
    init = fs.work.clone()
    # the state form first, from the untouched array
    v = torch.zeros(statelength(fs), device="cuda", dtype=torch.float64)
    gather(v, fs)
    host_v = v.cpu().numpy().copy()
    st = unify_desc(s, state=True)
    oracle.fill(st._descptr, _state_layout(st, nv), host_v, fac)
    unify(v, fs, s)
    torch.cuda.synchronize()
    assert (_bits(v) == torch.from_numpy(host_v).view(torch.int64)).all()
    assert torch.equal(fs.work, init)             # the state form leaves the array alone
    # the working-array form
    host = fs.work.cpu().numpy().copy()
    oracle.fill(unify_desc(s)._descptr, fs.layout, host, fac)
    unify(fs, s)
    torch.cuda.synchronize()
    assert (_bits(fs.work) == torch.from_numpy(host).view(torch.int64)).all()
    # and the two agree through gather!
    w = torch.zeros_like(v)
    gather(w, fs)
    torch.cuda.synchronize()
    assert torch.equal(w, v)


def test_unify_schedule_without_ghosts_face_centred():
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    f.refine([0])
    f.balance()
    f.refine([4])
    f.balance()
    fs = FieldSet(f, 3, G=0, centering=("vertex", "cell", "cell"), device="cuda", layout="arena", fill=0.0)
    W.fill_uniform(fs, 8)
    us = UnifySchedule(fs)
    host = fs.work.cpu().numpy().copy()
    oracle.fill(unify_desc(us)._descptr, fs.layout, host, None)
    unify(fs, us)
    torch.cuda.synchronize()
    assert (_bits(fs.work) == torch.from_numpy(host).view(torch.int64)).all()
    # on a state vector, scatter! / gather! around it give the same numbers
    v = torch.rand(statelength(fs), device="cuda", dtype=torch.float64)
    scatter(fs, v)
    unify(v, fs, us)
    unify(fs, us)
    w = torch.zeros_like(v)
    gather(w, fs)
    torch.cuda.synchronize()
    assert torch.equal(w, v)


# The vertex-centred forests among seeds 20..39 of the batched configuration.
VERTEX_SEEDS = (20, 21, 24, 26, 29, 30, 32, 34, 35)


def test_unify_and_state_copies_on_random_forests():
    """The vertex-centred forests among seeds 20..39 of the batched
    configuration, dense and arena: unify! on the working array and on a
    state vector, gather! and scatter!, each bit for bit against the
    oracle."""
    failed = []
    done = 0
    for seed in VERTEX_SEEDS:
        f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=300)
        assert centering == "vertex", f"seed {seed} no longer draws a vertex-centred forest"

        for layout in ("dense", "arena"):
            nv = nvars[0]
            print(f"[unify] seed={seed} {layout} G={G} nvars={nv} blocks={f.nleaves}", flush=True)
            fs = FieldSet(f, nv, G=G, centering=centering, device="cuda", layout=layout, fill=0.0)
            W.fill_uniform(fs, seed)
            s = GhostSchedule(fs, ops)
            # gather! against the oracle
            v = torch.zeros(statelength(fs), device="cuda", dtype=torch.float64)
            gather(v, fs)
            want = np.zeros(statelength(fs))
            oracle.gather(fs.layout, _own_first(fs), ownextents(fs), want, fs.work.cpu().numpy().copy())
            torch.cuda.synchronize()
            if not (_bits(v) == torch.from_numpy(want).view(torch.int64)).all():
                failed.append((seed, layout, "gather"))
            # unify! on the state vector
            host_v = v.cpu().numpy().copy()
            oracle.fill(unify_desc(s, state=True)._descptr, _state_layout(unify_desc(s, state=True), nv), host_v,
                        fs.factors.cpu().numpy() if fs.factors is not None else None)
            unify(v, fs, s)
            torch.cuda.synchronize()
            if not (_bits(v) == torch.from_numpy(host_v).view(torch.int64)).all():
                failed.append((seed, layout, "unify state"))
            # scatter! of the unified state, against the oracle
            host = fs.work.cpu().numpy().copy()
            oracle.scatter(fs.layout, _own_first(fs), ownextents(fs), host, v.cpu().numpy().copy())
            scatter(fs, v)
            torch.cuda.synchronize()
            if not (_bits(fs.work) == torch.from_numpy(host).view(torch.int64)).all():
                failed.append((seed, layout, "scatter"))
            # unify! on the working array
            host = fs.work.cpu().numpy().copy()
            oracle.fill(unify_desc(s)._descptr, fs.layout, host,
                        fs.factors.cpu().numpy() if fs.factors is not None else None)
            unify(fs, s)
            torch.cuda.synchronize()
            if not (_bits(fs.work) == torch.from_numpy(host).view(torch.int64)).all():
                failed.append((seed, layout, "unify array"))
            done += 1
    assert done == 2 * len(VERTEX_SEEDS) and not failed, (done, failed)


def test_unify_refusals():
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    f.refine([0])
    f.balance()
    fs = FieldSet(f, 1, G=1, centering="vertex", device="cuda", fill=0.0)
    us = UnifySchedule(fs)
    other = FieldSet(f, 1, G=2, centering="vertex", device="cuda", fill=0.0)
    with pytest.raises(ValueError, match="another layout"):
        unify(other, us)
    with pytest.raises(ValueError, match="entries"):
        unify(torch.zeros(3, device="cuda", dtype=torch.float64), fs, us)
    f.refine([0])
    with pytest.raises(Exception, match="stale"):
        unify(fs, us)
