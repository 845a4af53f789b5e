"""unify! on the CPU: UnifySchedule's stages (src/unify.jl:76-170), the
descriptors unify! replays (stages P and 0 alone, in working-array or state
indices, unify_arrays!, src/ghosts.jl:1874-1906), and the oracle on them.

Pinned by what the reference fixes: unify! is the first two stages of the
fill, so on a set with a vertex-like dimension its result on the own points
is the full fill's, bit for bit; the state form is the working-array form
seen through gather!; and on an all-vertex set the restriction is injection
whatever the operators' order, so a UnifySchedule and a GhostSchedule agree.
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import (EvenParity, FieldSet, Forest, GhostSchedule, OddParity, Operators, UnifySchedule,
                         ownextents, statelength, unify_desc)
from treeamr_hip.core import _own_first
from treeamr_hip import workloads as W


def _forest(reflect=False):
    if reflect:
        f = Forest((2, 2, 2), N=8, reflecting=((False, False), (False, False), (True, False)),
                   outflow=((True, True), (True, True), (False, True)))
    else:
        f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    f.refine([0])
    f.balance()
    f.refine([3])
    f.balance()
    return f


def _parity(nvars, reflect):
    return tuple((EvenParity, EvenParity, EvenParity if v % 2 == 0 else OddParity)
                 for v in range(nvars)) if reflect else None


def test_unify_schedule_refuses_a_cell_set():
    fs = FieldSet(_forest(), 1, G=2, device="cpu")
    with pytest.raises(Exception, match="no shared point"):
        UnifySchedule(fs)


@pytest.mark.parametrize("reflect", [False, True])
def test_unify_desc_is_stages_p_and_0_and_the_state_form_shifts_by_g(reflect):
    f = _forest(reflect)
    fs = FieldSet(f, 2, G=3, centering="vertex", parity=_parity(2, reflect), device="cpu")
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    w = unify_desc(s)
    tags = [w.desc.stage_tags[i] for i in range(w.desc.nstages)]
    assert tags == ([-1, 0] if reflect else [0])
    st = unify_desc(s, state=True)
    assert tuple(st.desc.stored) == (9, 9, 9) and st.desc.ngroups == w.desc.ngroups
    for a, b in zip(w.groups(), st.groups()):
        for d in range(3):
            assert a["stencils"][d]["targetfirst"] - 3 == b["stencils"][d]["targetfirst"]
            assert (a["stencils"][d]["srcstart"] - 3 == b["stencils"][d]["srcstart"]).all()


@pytest.mark.parametrize("reflect", [False, True])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_oracle_unify_is_the_fill_on_own_points(reflect, layout):
    f = _forest(reflect)
    fs = FieldSet(f, 2, G=3, centering="vertex", parity=_parity(2, reflect), device="cpu", layout=layout)
    fs.work.copy_(torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, fs.nelements)))
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    fac = fs.factors.numpy() if fs.factors is not None else None
    full = fs.work.numpy().copy()
    oracle.fill(s._descptr, fs.layout, full, fac)
    uni = fs.work.numpy().copy()
    oracle.fill(unify_desc(s)._descptr, fs.layout, uni, fac)
    k, j, i = fs.own_slices()
    A = lambda x: torch.from_numpy(x).as_strided(fs.array().shape, fs.array().stride())[:, :, k, j, i]
    assert torch.equal(A(full).contiguous().view(torch.int64), A(uni).contiguous().view(torch.int64))
    # the state form: gather, unify on the state, equals gather of the unified array
    u = np.zeros(statelength(fs))
    oracle.gather(fs.layout, _own_first(fs), ownextents(fs), u, fs.work.numpy().copy())
    st = unify_desc(s, state=True)
    from treeamr_hip import _abi
    import ctypes as C
    L = _abi.tah_layout()
    _abi.check(_abi.lib().tah_layout_dense(st._descptr, fs.nvars, C.byref(L)))
    oracle.fill(st._descptr, L, u, fac)
    want = np.zeros_like(u)
    oracle.gather(fs.layout, _own_first(fs), ownextents(fs), want, uni)
    assert (u.view(np.int64) == want.view(np.int64)).all()


def test_unify_schedule_matches_the_ghost_schedule_on_a_vertex_set():
    """Injection along a vertex-like dimension, whatever the order."""
    f = _forest()
    fs = FieldSet(f, 1, G=0, centering="vertex", device="cpu")
    fs.work.copy_(torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, fs.nelements)))
    a = fs.work.numpy().copy()
    us = UnifySchedule(fs)
    oracle.fill(unify_desc(us)._descptr, fs.layout, a, None)
    g3 = FieldSet(f, 1, G=3, centering="vertex", device="cpu")
    s = GhostSchedule(g3, Operators(prolongation=6, restriction=6))
    b = np.zeros(statelength(g3))
    oracle.gather(fs.layout, _own_first(fs), ownextents(fs), b, fs.work.numpy().copy())
    st = unify_desc(s, state=True)
    from treeamr_hip import _abi
    import ctypes as C
    L = _abi.tah_layout()
    _abi.check(_abi.lib().tah_layout_dense(st._descptr, 1, C.byref(L)))
    oracle.fill(st._descptr, L, b, None)
    got = np.zeros_like(b)
    oracle.gather(fs.layout, _own_first(fs), ownextents(fs), got, a)
    assert (got.view(np.int64) == b.view(np.int64)).all()


def test_face_centred_unify_averages_across_the_face():
    """A face-centred set (vertex-like in x only): the coarse copy of a face
    value at a coarse-fine face becomes the exact mean of the four fine ones."""
    f = Forest((2, 1, 1), N=8, periodic=(True,) * 3, extents=((0.0, 2.0), (0.0, 1.0), (0.0, 1.0)))
    f.refine([1])
    f.balance()
    fs = FieldSet(f, 1, G=0, centering=("vertex", "cell", "cell"), device="cpu")
    rng = np.random.default_rng(6)
    fs.work.copy_(torch.from_numpy(rng.uniform(-1, 1, fs.nelements)))
    us = UnifySchedule(fs)
    a = fs.work.numpy().copy()
    oracle.fill(unify_desc(us)._descptr, fs.layout, a, None)
    arr = torch.from_numpy(a).as_strided(fs.array().shape, fs.array().stride())
    before = fs.array()
    keys = f.leaves
    coarse = [b for b, k in enumerate(keys) if k.level == 0][0]
    # the coarse block's high-x face (x = N) meets the fine blocks' low-x faces (x = 0)
    fine = [b for b, k in enumerate(keys) if k.level == 1 and k.coords[0] == 0]
    assert len(fine) == 4
    face = torch.zeros(16, 16, dtype=torch.float64)
    for b in fine:
        cy, cz = keys[b].coords[1], keys[b].coords[2]
        face[cz * 8:(cz + 1) * 8, cy * 8:(cy + 1) * 8] = before[b, 0, :, :, 0]
    want = face.view(8, 2, 8, 2).mean(dim=(1, 3))
    torch.testing.assert_close(arr[coarse, 0, :, :, 8], want, rtol=0, atol=1e-15)


def test_unify_gpu_seed_list_draws_vertex_forests():
    """The seed list the GPU unify test loops over is exactly the vertex-centred
    draws among seeds 20..39, so that its coverage cannot shrink unnoticed."""
    from treeamr_hip import workloads as W
    from test_gpu_unify import VERTEX_SEEDS
    assert VERTEX_SEEDS == tuple(s for s in range(20, 40) if W.random_forest(s, max_blocks=300)[4] == "vertex")
