"""The state vector on the CPU: the oracle's scatter/gather (scatter_kernel!
and gather_kernel!, src/state.jl:101-115) and the host logic of the state
API (statelength, staterange, statearray and their checks, src/state.jl).

The oracle is pinned against an independent formulation: the working array
viewed as (block, var, k, j, i) with the own range sliced out and flattened
in C order is the state vector (x fastest, then y, z, variable, block), as
the reference's statearray reshape defines it. The reference's own tests of
the pair are round trips (test/closed_tests.jl:798-912,
test/glue_ball_tests.jl:72-86): gather after scatter gives the state back,
and a scatter leaves the ghosts alone.
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import (FieldSet, Forest, ownextents, statearray, statelength, staterange,
                         statevector, stateview)
from treeamr_hip.core import _own_first
from treeamr_hip import workloads as W


def _forest(D=3):
    f = Forest((2,) * D, N=8, periodic=(True,) * D)
    f.refine([0])
    f.balance()
    return f


def _own_view(fs):
    """The own points of a CPU field set as (block, var, k, j, i), by slicing."""
    k, j, i = fs.own_slices()
    return fs.array()[:, :, k, j, i]


CASES = [(3, "cell", "dense", 4), (3, "vertex", "arena", 4), (3, "cell", "arena", 3),
         (2, "vertex", "dense", 2), (2, "cell", "arena", 5), (1, "cell", "dense", 3)]


@pytest.mark.parametrize("D,centering,layout,nvars", CASES)
def test_oracle_gather_is_the_sliced_own_range(D, centering, layout, nvars):
    fs = FieldSet(_forest(D), nvars, G=2, centering=centering, device="cpu", layout=layout)
    rng = np.random.default_rng(D + nvars)
    fs.work.copy_(torch.from_numpy(rng.uniform(-1, 1, fs.nelements)))
    u = np.zeros(statelength(fs))
    oracle.gather(fs.layout, _own_first(fs), ownextents(fs), u, fs.work)
    want = _own_view(fs).contiguous().view(-1).numpy()
    assert u.size == want.size
    assert (u.view(np.int64) == want.view(np.int64)).all()


@pytest.mark.parametrize("D,centering,layout,nvars", CASES)
def test_oracle_scatter_round_trip_and_ghosts_untouched(D, centering, layout, nvars):
    fs = FieldSet(_forest(D), nvars, G=2, centering=centering, device="cpu", layout=layout,
                  fill=float("nan"))
    rng = np.random.default_rng(7 * D + nvars)
    u = rng.uniform(-1, 1, statelength(fs))
    oracle.scatter(fs.layout, _own_first(fs), ownextents(fs), fs.work, u)
    own = _own_view(fs)
    assert not torch.isnan(own).any(), "an own point was not written"
    # every stored point outside the own range is still NaN
    nan = torch.isnan(fs.array())
    assert int(nan.sum()) == fs.nblocks * fs.nvars * (np.prod(fs.stored) - np.prod(ownextents(fs)))
    back = np.zeros_like(u)
    oracle.gather(fs.layout, _own_first(fs), ownextents(fs), back, fs.work)
    assert (back.view(np.int64) == u.view(np.int64)).all()


def test_statelength_counts_every_own_point():
    f = _forest()
    cell = FieldSet(f, 3, G=2, device="cpu")
    vert = FieldSet(f, 2, G=3, centering="vertex", device="cpu")
    assert ownextents(cell) == (8, 8, 8) and ownextents(vert) == (9, 9, 9)
    assert statelength(cell) == 8 ** 3 * 3 * f.nleaves
    assert statelength(vert) == 9 ** 3 * 2 * f.nleaves
    assert statelength((cell, vert)) == statelength(cell) + statelength(vert)
    assert staterange((cell, vert), 0) == range(0, statelength(cell))
    assert staterange((cell, vert), 1) == range(statelength(cell), statelength(cell) + statelength(vert))
    f2 = Forest((2, 2), N=4, periodic=(True, True))
    assert ownextents(FieldSet(f2, 1, G=1, centering="vertex", device="cpu")) == (5, 5, 1)


def test_state_views_share_memory_and_follow_the_reference_shape():
    f = _forest()
    a = FieldSet(f, 2, G=2, device="cpu")
    b = FieldSet(f, 1, G=2, centering="vertex", device="cpu")
    u = statevector((a, b))
    assert u.numel() == statelength((a, b)) and float(u.abs().sum()) == 0.0
    va = statearray(u, (a, b), 0)
    assert tuple(va.shape) == (f.nleaves, 2, 8, 8, 8)
    vb = statearray(u, (a, b), 1)
    assert tuple(vb.shape) == (f.nleaves, 1, 9, 9, 9)
    vb[0, 0, 0, 0, 1] = 5.0  # x = 1 of the first block: one past the segment's start
    assert float(u[statelength(a) + 1]) == 5.0
    assert float(stateview(u, (a, b), 1)[1]) == 5.0


def test_state_checks():
    f = _forest()
    a = FieldSet(f, 2, G=2, device="cpu")
    with pytest.raises(ValueError, match="state vector has 3 entries but this field set needs"):
        statearray(torch.zeros(3, dtype=torch.float64), a)
    with pytest.raises(ValueError, match="at least one set"):
        statelength(())
    other = FieldSet(_forest(), 2, G=2, device="cpu")
    with pytest.raises(ValueError, match="different forest"):
        statelength((a, other))
    with pytest.raises(ValueError, match="outside"):
        staterange((a,), 1)
    g = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    s = FieldSet(g, 1, G=2, device="cpu")
    g.refine([0])
    with pytest.raises(ValueError, match="not resized"):
        statelength((s,))


def test_state_entry_points_refuse_bad_ranges():
    """tah_scatter / tah_gather check the own range against the layout before
    any launch (no GPU is touched: the checks fail first)."""
    import ctypes as C
    from treeamr_hip import _abi
    from treeamr_hip.core import _i32
    fs = FieldSet(_forest(), 2, G=2, device="cpu")
    L = C.byref(fs.layout)
    st = _abi.lib().tah_scatter(L, _i32((2, 2, 2)), _i32((13, 8, 8)), None, None, None)
    assert st == _abi.TAH_ELAYOUT
    st = _abi.lib().tah_gather(L, _i32((2, 2, 2)), _i32((0, 8, 8)), None, None, None)
    assert st == _abi.TAH_EINVAL
    n = C.c_int64()
    assert _abi.lib().tah_state_length(L, _i32((8, 8, 8)), C.byref(n)) == 0
    assert n.value == statelength(fs)
