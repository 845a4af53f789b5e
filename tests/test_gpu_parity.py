"""GPU parity: libtah's fills against the oracle (the CPU restatement of the
reference fill) on the same seeded inputs, bit for bit.

The reference checks its device fill against the serial CPU fill the same
way (test/gpu_tests.jl:476-499) and its flat launch against the shaped one
with the full sum, on awkward data, NaNs compared as one pattern
(test/ghost_tests.jl:507-560, test/helpers/ghost.jl:7-45).
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import FieldSet, GhostSchedule, fill_ghosts
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _bits(t):
    return t.detach().cpu().contiguous().view(torch.int64)


def _prepare(forest, ops, kw, layout="dense", seed=0, data="uniform"):
    kw = dict(kw)
    nvars = kw.pop("nvars")
    fs = FieldSet(forest, nvars, device="cuda", layout=layout, fill=float("nan"), **kw)
    if data == "uniform":
        W.fill_uniform(fs, seed)
    else:
        W.fill_gaussians(fs, seed)
    sched = GhostSchedule(fs, ops)
    return fs, sched


def _oracle_result(fs, sched, skip=True):
    host = fs.work.detach().cpu().numpy().copy()
    fac = fs.factors.cpu().numpy().copy() if fs.factors is not None else None
    oracle.fill(sched._descptr, fs.layout, host, fac, nthreads=0, skip=skip)
    return torch.from_numpy(host)


def _check(fs, sched, method):
    print(f"[parity] method={method} blocks={fs.nblocks} nvars={fs.nvars}", flush=True)
    expected = _oracle_result(fs, sched)
    fill_ghosts(fs, sched, method=method)
    torch.cuda.synchronize()
    got = fs.work.detach().cpu()
    assert not torch.isnan(fs.array()).any(), "a ghost was left unwritten"
    mism = (_bits(got) != expected.view(torch.int64)).sum().item()
    assert mism == 0, f"{mism} values differ from the oracle"


@pytest.mark.parametrize("method", ["plain", "staged"])
def test_config1_bit_exact(method):
    f, ops, kw, _ = W.config1()
    fs, sched = _prepare(f, ops, kw, seed=0)
    _check(fs, sched, method)


@pytest.mark.parametrize("method", ["plain", "staged"])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_config5_small_bit_exact(method, layout):
    f, ops, kw, _ = W.config5(nvars=3, roots=4, N=16)
    fs, sched = _prepare(f, ops, kw, layout=layout, seed=1, data="gaussian")
    _check(fs, sched, method)


@pytest.mark.parametrize("method", ["plain", "staged"])
@pytest.mark.parametrize("offset", [0, 1])
def test_arena_even_variables_bit_exact(method, offset):
    """Arena layout with an even variable count, where the kernels move two
    variables per 16-byte access; with offset 1 the array starts 8 bytes past
    a 16-byte boundary, so they fall back to single values."""
    f, ops, kw, _ = W.config5(nvars=4, roots=2, N=16)
    fs, sched = _prepare(f, ops, kw, layout="arena", seed=3, data="gaussian")
    buf = torch.empty(fs.nelements + 1, dtype=torch.float64, device="cuda")
    view = buf[offset:offset + fs.nelements]
    view.copy_(fs.work)
    fs.work = view
    assert fs.work.data_ptr() % 16 == 8 * offset
    _check(fs, sched, method)


@pytest.mark.parametrize("family,orders", [("pointvalue", (2, 2)), ("pointvalue", (4, 4)), ("pointvalue", (6, 6)),
                                            ("pointvalue", (6, 2)), ("conservative", (3, 2)),
                                            ("conservative", (5, 2))])
@pytest.mark.parametrize("centering", ["cell", "vertex"])
@pytest.mark.parametrize("offset", [0, 1])
def test_variable_pair_interpolation_every_order(family, orders, centering, offset):
    """Arena layout, 6 variables: the merged interpolations take two
    neighbouring variables per thread, with 16-byte accesses (offset 0) or
    single ones (the array 8 bytes past a 16-byte boundary), for every
    specialised order of prolongation and restriction."""
    from treeamr_hip import Conservative, Forest, Operators
    if family == "conservative" and centering == "vertex":
        pytest.skip("the conservative family is cell-centred")
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    f.refine([0])
    f.balance()
    f.refine([3])
    f.balance()
    ops = Operators(prolongation=orders[0], restriction=orders[1],
                    **({"family": Conservative} if family == "conservative" else {}))
    fs, sched = _prepare(f, ops, dict(nvars=6, G=3, centering=centering), layout="arena", seed=sum(orders))
    buf = torch.empty(fs.nelements + 1, dtype=torch.float64, device="cuda")
    view = buf[offset:offset + fs.nelements]
    view.copy_(fs.work)
    fs.work = view
    assert sched.plan(fs.layout).options["interp_pairs"] == 1
    _check(fs, sched, "staged")


@pytest.mark.parametrize("nvars", [2, 514])
def test_variable_pair_tiles_at_the_extremes(nvars):
    """Pair tiles with one thread per unit (2 variables: 256 units a tile)
    and with a variable count above one tile's 512 (514: a tile of 512
    variables and one of 2), cell-centred, arena layout."""
    from treeamr_hip import Forest, Operators
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    f.refine([0])
    f.balance()
    ops = Operators(prolongation=4, restriction=4)
    fs, sched = _prepare(f, ops, dict(nvars=nvars, G=2, centering="cell"), layout="arena", seed=nvars)
    assert sched.plan(fs.layout).options["interp_pairs"] == 1
    _check(fs, sched, "staged")


@pytest.mark.parametrize("nvars", [3, 5, 9, 515])
@pytest.mark.parametrize("centering", ["cell", "vertex"])
def test_variable_pairs_over_odd_counts(nvars, centering):
    """Arena layout with an odd variable count: prolongation over pairs of
    single loads, the last variable of each tile alone (515: a tile of 512
    variables and one of 3), at workgroups of 256 and 64 threads, and with
    the pairs off, each bit for bit the oracle."""
    from treeamr_hip import Forest, Operators
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    f.refine([0])
    f.balance()
    f.refine([2])
    f.balance()
    fs, sched = _prepare(f, Operators(prolongation=4, restriction=4), dict(nvars=nvars, G=2, centering=centering),
                         layout="arena", seed=nvars)
    plan = sched.plan(fs.layout)
    init = fs.work.clone()
    fam = next(x for x in plan.families if not x["restriction"])
    for threads, pairs in ((256, 1), (64, 1), (256, 0)):
        print(f"[odd pairs] nvars={nvars} {centering} threads={threads} pairs={pairs}", flush=True)
        plan.set_family(fam["order"], 0, threads=threads, pairs=pairs, merge=1)
        fs.work.copy_(init)
        _check(fs, sched, "staged")
    plan.set_family(fam["order"], 0, threads=256, pairs=1, merge=1)


@pytest.mark.parametrize("seed", [2, 3, 4, 5, 6, 12])
def test_random_forests_bit_exact(seed):
    """Forests of the batched configuration, cell (seeds 2, 4, 5) and vertex
    (3, 6, 12) centering."""
    f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=600)
    fs, sched = _prepare(f, ops, dict(nvars=nvars[0], G=G, centering=centering), seed=seed)
    _check(fs, sched, "staged")


@pytest.mark.parametrize("seed", [2, 3, 4, 5, 6, 12])
def test_random_forests_arena_even_variables(seed):
    """The same forests in the arena layout with an even variable count
    (the set's own count rounded up), so that prolongation runs over
    variable pairs, beside restriction with its rows loaded ahead."""
    f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=600)
    nv = nvars[0] + nvars[0] % 2
    fs, sched = _prepare(f, ops, dict(nvars=nv, G=G, centering=centering), layout="arena", seed=seed)
    assert sched.plan(fs.layout).options["interp_pairs"] == 1
    _check(fs, sched, "staged")


def test_forty_more_random_forests_both_layouts():
    """Seeds 20..59 of the batched configuration (up to 500 blocks each), in
    the arena layout with an even variable count and in the dense layout with
    the drawn count: 80 fills, every value against the oracle."""
    failed = []
    for seed in range(20, 60):
        f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=500)
        for layout in ("arena", "dense"):
            nv = nvars[0] + (nvars[0] % 2 if layout == "arena" else 0)
            print(f"[parity] seed={seed} {layout} {centering} G={G} nvars={nv} blocks={f.nleaves}", flush=True)
            fs, sched = _prepare(f, ops, dict(nvars=nv, G=G, centering=centering), layout=layout, seed=seed)
            expected = _oracle_result(fs, sched)
            fill_ghosts(fs, sched)
            torch.cuda.synchronize()
            mism = int((_bits(fs.work) != expected.view(torch.int64)).sum())
            if mism:
                failed.append((seed, layout, mism))
    assert not failed, failed


@pytest.mark.parametrize("method", ["plain", "staged"])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_config2_vertex_bit_exact(method, layout):
    """The nested vertex-centred mesh (232 blocks of 16^3, G=3, 8 variables,
    PointValue(6,6), reflecting z-low with odd variables 7-8, outflow
    elsewhere): stage P, stage 0, phase 1 and three prolongation levels."""
    f, ops, kw, _ = W.config2()
    kw = dict(kw)
    nvars = kw.pop("nvars")
    fs = FieldSet(f, nvars, device="cuda", layout=layout, fill=float("nan"), **kw)
    W.fill_waves(fs, 1)
    sched = GhostSchedule(fs, ops)
    _check(fs, sched, method)


def test_config5_full_mesh_one_variable():
    """The full-size large-forest mesh (about 5k blocks of 32^3), one
    variable: every value against the oracle."""
    f, ops, kw, _ = W.config5(nvars=1)
    fs, sched = _prepare(f, ops, kw, seed=7, data="gaussian")
    _check(fs, sched, "staged")


def test_awkward_data_full_sum():
    """Signed zeros, subnormals, huge values, NaN and Inf in own points: the
    device's full sum against the oracle's full sum (skip=False); finite
    values bit for bit, NaNs as one pattern."""
    f, ops, kw, _ = W.config5(nvars=2, roots=2, N=16)
    fs, sched = _prepare(f, ops, kw, seed=3)
    a = fs.array()
    own = (slice(None), slice(None)) + fs.own_slices()
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x = a[own]
    pick = torch.randint(0, 8, x.shape, generator=g, device="cuda")
    specials = torch.tensor([0.0, -0.0, 5e-324, -2.2e-308, 1e300, -1e300, float("inf"), float("nan")],
                            dtype=torch.float64, device="cuda")
    mask = torch.rand(x.shape, generator=g, device="cuda", dtype=torch.float64) < 0.05
    a[own] = torch.where(mask, specials[pick], x)
    expected = _oracle_result(fs, sched, skip=False)
    fill_ghosts(fs, sched, method="staged")
    torch.cuda.synchronize()
    got = fs.work.detach().cpu()
    nan_g, nan_e = torch.isnan(got), torch.isnan(expected)
    assert torch.equal(nan_g, nan_e)
    keep = ~nan_g
    assert torch.equal(_bits(got)[keep], expected.view(torch.int64)[keep])


def test_fill_is_idempotent_and_interiors_untouched():
    f, ops, kw, _ = W.config5(nvars=2, roots=4, N=16)
    fs, sched = _prepare(f, ops, kw, seed=4, data="gaussian")
    own = (slice(None), slice(None)) + fs.own_slices()
    before = fs.array()[own].clone()
    fill_ghosts(fs, sched)
    once = fs.work.clone()
    fill_ghosts(fs, sched)
    torch.cuda.synchronize()
    assert torch.equal(_bits(once), _bits(fs.work))
    assert torch.equal(_bits(before), _bits(fs.array()[own]))


def test_plain_equals_staged_full_config5():
    """At the full large-forest size (20 variables): the plain per-group fill
    and the staged fill agree bit for bit."""
    f, ops, kw, _ = W.config5(nvars=20)
    fs, sched = _prepare(f, ops, kw, seed=5, data="gaussian")
    init = fs.work.clone()
    fill_ghosts(fs, sched, method="plain")
    plain = fs.work.clone()
    fs.work.copy_(init)
    del init
    fill_ghosts(fs, sched, method="staged")
    torch.cuda.synchronize()
    assert not torch.isnan(fs.array()).any()
    assert torch.equal(plain.view(torch.int64), fs.work.view(torch.int64))
