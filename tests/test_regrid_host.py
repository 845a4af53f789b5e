"""The regrid transfer on the CPU: forest coarsening, the regrid schedule's
classification (classify_regrid!, src/regrid.jl:620-652) and the oracle's
transfer (run_stages! in regrid!, src/regrid.jl:1337-1384).

The oracle is pinned by what the mathematics fixes, as the reference's own
regrid tests do (test/regrid_tests.jl, test/closed_tests.jl): a regrid onto
the same leaves copies every own point; prolongation and restriction of a
polynomial below the operators' order reproduce it; a conservative
coarsening keeps the mean of the children; and over a vertex-centred set
every created copy of a point shared with a kept block takes the kept
block's value.
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import Conservative, FieldSet, Forest, Operators, RegridSchedule


def _forest(D=3, roots=2, N=8, periodic=True):
    f = Forest((roots,) * D, N=N, periodic=(periodic,) * D, extents=((0.0, 1.0),) * D)
    f.refine([0])
    f.balance()
    return f


def _move(f, refine=(), coarsen=()):
    """Coarsen, refine (old indices) and balance; returns the old leaf arrays."""
    old = f.leaf_arrays()
    keys = f.leaves
    if coarsen:
        f.coarsen(coarsen)
    if refine:
        now = {k: i for i, k in enumerate(f.leaves)}
        f.refine([now[keys[i]] for i in refine])
    f.balance()
    return old


def test_coarsen_merges_whole_families_and_keeps_balance():
    f = _forest()
    before = f.leaves
    assert f.coarsen(range(8)) == 1          # the refined root's children
    assert f.nleaves == 8 and f.isbalanced()
    f.refine([0])
    f.balance()
    assert f.leaves == before
    assert f.coarsen(range(7)) == 0          # an incomplete family stays
    # children of a family next to a leaf two levels finer stay
    g = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    g.refine([0])
    g.balance()
    g.refine([7])                            # a child of root 0, at its far corner
    g.balance()
    lv = [k.level for k in g.leaves]
    assert max(lv) == 2
    fam = [i for i, k in enumerate(g.leaves) if k.level == 1 and k.root == 0]
    n = g.nleaves
    assert g.coarsen(fam) == 0 and g.nleaves == n and g.isbalanced()


def test_regrid_schedule_classifies_copy_prolong_restrict():
    f = _forest()
    old = _move(f, refine=[9], coarsen=range(8))
    rs = RegridSchedule(f, old, Operators(prolongation=4, restriction=4), G=2, centering="cell", nvars=1)
    assert rs.stage_tags() == [100] and rs.desc.nsources == len(old[0])
    counts = {}
    for g in rs.groups():
        counts[g["kind"]] = counts.get(g["kind"], 0) + g["ntransfers"]
    # 15 -> 15 leaves: 6 kept, one parent from 8 children, 8 children from a parent
    assert counts == {1: 6, 2: 8, 3: 8}
    with pytest.raises(Exception, match="at most one level"):
        g2 = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
        o2 = g2.leaf_arrays()
        g2.refine([0])
        g2.refine([0])                       # two levels in one step
        RegridSchedule(g2, o2, Operators(prolongation=4, restriction=4), G=2, centering="cell", nvars=1)


def _fieldset(f, centering, nvars=2, G=2, layout="dense"):
    return FieldSet(f, nvars, G=G, centering=centering, device="cpu", layout=layout, fill=0.0)


def _own(fs):
    k, j, i = fs.own_slices()
    return fs.array()[:, :, k, j, i]


@pytest.mark.parametrize("centering", ["cell", "vertex"])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_identity_regrid_copies_every_own_point(centering, layout):
    f = _forest()
    old = f.leaf_arrays()
    a = _fieldset(f, centering, layout=layout)
    a.work.copy_(torch.from_numpy(np.random.default_rng(1).uniform(-1, 1, a.nelements)))
    rs = RegridSchedule(f, old, Operators(prolongation=4, restriction=4), G=2, centering=centering, nvars=2)
    b = _fieldset(f, centering, layout=layout)
    oracle.regrid(rs._descptr, b.layout, a.work.numpy(), b.work.numpy())
    assert torch.equal(_own(b).contiguous().view(torch.int64), _own(a).contiguous().view(torch.int64))
    own_count = _own(b).numel()
    assert int((b.work != 0).sum()) == own_count    # ghosts untouched (zero)


def _poly(x, y, z, v):
    return (0.3 + v) + 0.7 * x - 1.1 * y * y + 0.5 * x * y * z + 0.2 * z ** 3


@pytest.mark.parametrize("centering", ["cell", "vertex"])
def test_regrid_reproduces_cubics(centering):
    """PointValue(4,4): prolongation and restriction are exact on cubics (a
    non-periodic brick: a cubic does not wrap)."""
    f = _forest(periodic=False)
    a = _fieldset(f, centering)
    X, Y, Z = a.coordinates()
    for v in range(a.nvars):
        a.array()[:, v] = torch.from_numpy(_poly(X, Y, Z, v))
    old = _move(f, refine=[9, 12], coarsen=range(8))
    rs = RegridSchedule(f, old, Operators(prolongation=4, restriction=4), G=2, centering=centering, nvars=2)
    b = _fieldset(f, centering)
    oracle.regrid(rs._descptr, b.layout, a.work.numpy(), b.work.numpy())
    X, Y, Z = b.coordinates()
    k, j, i = b.own_slices()
    for v in range(b.nvars):
        want = _poly(X, Y, Z, v)[:, k, j, i]
        got = b.array()[:, v, k, j, i].numpy()
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_conservative_coarsening_keeps_the_mean():
    f = _forest()
    a = _fieldset(f, "cell", nvars=1)
    a.work.copy_(torch.from_numpy(np.random.default_rng(2).uniform(-1, 1, a.nelements)))
    children = _own(a)[:8, 0].clone()        # the 8 children of root 0 (refined first)
    old = _move(f, coarsen=range(8))
    ops = Operators(prolongation=3, restriction=2, family=Conservative)
    rs = RegridSchedule(f, old, ops, G=2, centering="cell", nvars=1)
    b = _fieldset(f, "cell", nvars=1)
    oracle.regrid(rs._descptr, b.layout, a.work.numpy(), b.work.numpy())
    coarse = _own(b)[0, 0]
    assert abs(float(coarse.mean()) - float(children.mean())) < 1e-14
    # each coarse cell is the mean of the 8 fine cells it covers
    full = torch.zeros(16, 16, 16, dtype=torch.float64)
    co = old[2]
    for c in range(8):
        ox, oy, oz = (int(x) for x in co[c][:3])
        full[oz * 8:(oz + 1) * 8, oy * 8:(oy + 1) * 8, ox * 8:(ox + 1) * 8] = children[c]
    want = full.view(8, 2, 8, 2, 8, 2).mean(dim=(1, 3, 5))
    torch.testing.assert_close(coarse, want, rtol=0, atol=1e-15)


def test_vertex_created_copies_take_the_kept_value():
    """Over a vertex-centred set, after the transfer every point a created
    block shares with a same-level kept block holds the kept block's value
    (the kept-over-created pass); a prolongated plane alone would differ."""
    f = _forest()
    a = _fieldset(f, "vertex", nvars=1)
    a.work.copy_(torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, a.nelements)))
    keys_before = set(f.leaves)
    old = _move(f, refine=[9])
    rs = RegridSchedule(f, old, Operators(prolongation=4, restriction=4), G=2, centering="vertex", nvars=1)
    assert rs.stage_tags() == [100, 101]
    b = _fieldset(f, "vertex", nvars=1)
    oracle.regrid(rs._descptr, b.layout, a.work.numpy(), b.work.numpy())
    X, Y, Z = b.coordinates()
    k, j, i = b.own_slices()
    lows, h, level = f.block_geometry()
    keys = f.leaves
    vals = {}
    for blk in range(f.nleaves):
        kept = keys[blk] in keys_before
        pos = np.stack([X[blk][k, j, i], Y[blk][k, j, i], Z[blk][k, j, i]], -1).reshape(-1, 3)
        pos = np.mod(np.round(pos * 1024), 1024).astype(np.int64)
        val = b.array()[blk, 0, k, j, i].reshape(-1).numpy()
        for p, x in zip(map(tuple, pos), val):
            vals.setdefault(p, []).append((int(level[blk]), kept, x))
    checked = 0
    for p, copies in vals.items():
        top = max(c[0] for c in copies)
        keptv = [c[2] for c in copies if c[0] == top and c[1]]
        made = [c[2] for c in copies if c[0] == top and not c[1]]
        if keptv and made:
            assert all(m == keptv[0] for m in made), p
            checked += 1
    assert checked > 0
