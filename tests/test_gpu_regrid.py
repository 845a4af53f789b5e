"""The regrid transfer on the GPU (tah_regrid_transfer) against the oracle's
(run_stages! in regrid!, src/regrid.jl:1337-1384), bit for bit: copies,
prolongations and restrictions from the old array into the new one, and the
kept-over-created pass of a vertex-centred set; then regrid() end to end as
regrid! runs it (fill, coarsen/refine/balance, transfer), followed by a
fill on the new mesh.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import (Conservative, FieldSet, Forest, GhostSchedule, Operators, RegridSchedule, fill_ghosts,
                         regrid, regrid_transfer)
from treeamr_hip import _abi
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _forest(N=8):
    f = Forest((2, 2, 2), N=N, periodic=(True,) * 3, extents=((0.0, 1.0),) * 3)
    f.refine([0])
    f.balance()
    return f


def _move(f, refine=(), coarsen=()):
    old = f.leaf_arrays()
    keys = f.leaves
    if coarsen:
        f.coarsen(coarsen)
    if refine:
        now = {k: i for i, k in enumerate(f.leaves)}
        f.refine([now[keys[i]] for i in refine])
    f.balance()
    return old


def _bits(t):
    return t.detach().cpu().contiguous().view(torch.int64)


@pytest.mark.parametrize("centering", ["cell", "vertex"])
@pytest.mark.parametrize("layout,nvars", [("dense", 3), ("arena", 3), ("arena", 4)])
@pytest.mark.parametrize("family", ["pointvalue", "conservative"])
def test_transfer_matches_the_oracle(centering, layout, nvars, family):
    print(f"[regrid] {centering} {layout} nvars={nvars} {family}", flush=True)
    if family == "conservative" and centering == "vertex":
        pytest.skip("the conservative family is cell-centred")
    ops = (Operators(prolongation=4, restriction=4) if family == "pointvalue" else
           Operators(prolongation=3, restriction=2, family=Conservative))
    f = _forest()
    a = FieldSet(f, nvars, G=2, centering=centering, device="cuda", layout=layout, fill=0.0)
    g = torch.Generator(device="cuda")
    g.manual_seed(nvars + 10 * (layout == "arena"))
    a.work.copy_(torch.rand(a.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1)
    old = _move(f, refine=[9, 12], coarsen=range(8))
    rs = RegridSchedule(f, old, ops, G=2, centering=centering, nvars=nvars)
    old_layout = a.layout
    b = FieldSet(f, nvars, G=2, centering=centering, device="cuda", layout=layout, fill=0.0)
    new, plan = regrid_transfer(b, old_layout, a.work, rs)
    torch.cuda.synchronize()
    host = np.zeros(b.nelements)
    oracle.regrid(rs._descptr, b.layout, a.work.cpu().numpy().copy(), host)
    assert (_bits(new) == torch.from_numpy(host).view(torch.int64)).all()
    # the plan runs only as a regrid transfer
    st = _abi.lib().tah_fill(plan._h, C.c_void_p(new.data_ptr()), None, None)
    assert st == _abi.TAH_EINVAL and "tah_regrid_transfer" in _abi.lib().tah_last_error().decode()


@pytest.mark.parametrize("centering", ["cell", "vertex"])
def test_regrid_end_to_end_then_fill(centering):
    f = _forest()
    ops = Operators(prolongation=4, restriction=4)
    fs = FieldSet(f, 2, G=2, centering=centering, device="cuda", layout="arena", fill=0.0)
    W.fill_uniform(fs, 4)
    s = GhostSchedule(fs, ops)
    # the same steps by hand on the CPU: fill, then the oracle's transfer
    fill_ghosts(fs, s)
    torch.cuda.synchronize()
    filled = fs.work.cpu().numpy().copy()
    keys = f.leaves
    old = f.leaf_arrays()
    assert regrid(f, [(fs, s)], refine=[9], coarsen=range(8))
    assert s.isstale() and fs.nblocks == f.nleaves
    rs = RegridSchedule(f, old, ops, G=2, centering=centering, nvars=2)
    host = np.zeros(fs.nelements)
    oracle.regrid(rs._descptr, fs.layout, filled, host)
    assert (_bits(fs.work) == torch.from_numpy(host).view(torch.int64)).all()
    # a fill on the new mesh, against the oracle's
    s2 = GhostSchedule(fs, ops)
    want = fs.work.cpu().numpy().copy()
    oracle.fill(s2._descptr, fs.layout, want, None)
    fill_ghosts(fs, s2)
    torch.cuda.synchronize()
    assert (_bits(fs.work) == torch.from_numpy(want).view(torch.int64)).all()
    assert not regrid(f, [(fs, s2)])          # nothing flagged: the mesh stays
    assert f.leaves != keys


def test_transfers_on_random_forests():
    """Twelve forests of the batched configuration (seeds 20..31): the leaves
    of the finest level coarsened where whole families allow, a few coarse
    leaves refined, balanced; the transfer in the dense layout and in the
    arena layout with an even variable count, bit for bit against the
    oracle's."""
    import numpy as np
    failed = []
    for seed in range(20, 32):
        f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=300)
        rng = np.random.default_rng(seed)
        for layout in ("dense", "arena"):
            g2, _, _, _, _ = W.random_forest(seed, max_blocks=300)   # a fresh copy of the forest
            nv = nvars[0] + (nvars[0] % 2 if layout == "arena" else 0)
            a = FieldSet(g2, nv, G=G, centering=centering, device="cuda", layout=layout, fill=0.0)
            gen = torch.Generator(device="cuda")
            gen.manual_seed(seed)
            a.work.copy_(torch.rand(a.work.shape, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1)
            lv = np.array([k.level for k in g2.leaves])
            fine = [int(i) for i in np.nonzero(lv == lv.max())[0]] if lv.max() > 0 else []
            coarse = [int(i) for i in np.nonzero(lv == lv.min())[0]]
            refine = [int(i) for i in rng.choice(coarse, size=min(3, len(coarse)), replace=False)]
            refine = [i for i in refine if i not in fine]
            old = _move(g2, refine=refine, coarsen=fine)
            print(f"[regrid] seed={seed} {layout} {centering} G={G} nvars={nv} {len(old[0])} -> {g2.nleaves} "
                  f"blocks", flush=True)
            rs = RegridSchedule(g2, old, ops, G=G, centering=centering, nvars=nv)
            b = FieldSet(g2, nv, G=G, centering=centering, device="cuda", layout=layout, fill=0.0)
            new, _ = regrid_transfer(b, a.layout, a.work, rs)
            torch.cuda.synchronize()
            host = np.zeros(b.nelements)
            oracle.regrid(rs._descptr, b.layout, a.work.cpu().numpy().copy(), host)
            mism = int((_bits(new) != torch.from_numpy(host).view(torch.int64)).sum())
            if mism:
                failed.append((seed, layout, mism))
    assert not failed, failed


def test_identity_transfer_at_size():
    """The reference's own regrid benchmark (bench/gpu.jl:159-174): old
    leaves = new leaves, the two-level mesh of 960 blocks of 32^3."""
    f, ops, kw = W.two_level(N=32)
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout="arena", fill=0.0, **kw)
    W.fill_uniform(fs, 5)
    rs = RegridSchedule(f, f.leaf_arrays(), ops, G=fs.G, centering=fs.centering, nvars=nv)
    new, _ = regrid_transfer(fs, fs.layout, fs.work, rs)
    torch.cuda.synchronize()
    k, j, i = fs.own_slices()
    got = new.as_strided(fs.array().shape, fs.array().stride())[:, :, k, j, i]
    assert torch.equal(got, fs.array()[:, :, k, j, i])


def test_transfer_refusals():
    f = _forest()
    ops = Operators(prolongation=4, restriction=4)
    a = FieldSet(f, 2, G=2, device="cuda", fill=0.0)
    old = _move(f, refine=[9])
    rs = RegridSchedule(f, old, ops, G=2, centering="cell", nvars=2)
    b = FieldSet(f, 2, G=2, device="cuda", fill=0.0)
    with pytest.raises(Exception, match="blocks"):
        regrid_transfer(b, b.layout, b.work, rs)          # the new array as the old one
    s = GhostSchedule(b, ops)
    plan = s.plan(b.layout)
    st = _abi.lib().tah_regrid_transfer(plan._h, C.byref(a.layout), C.c_void_p(a.work.data_ptr()),
                                        C.c_void_p(b.work.data_ptr()), None)
    assert st == _abi.TAH_EINVAL
