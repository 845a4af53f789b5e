"""CPU tests: the C ABI's symbols, the operators and their weights, the
forest, the schedule (partition, polynomial exactness, boundary faces), and
the oracle itself. No GPU is touched here.

Modelled on the reference's own tests: test/ghost_tests.jl (weights KATs
:5-110, check_operators :112-212, partition :214-245, polynomial exactness
:247-301), test/ghost_oracles.jl (write_counts :228-255, exchange_error
:111-117), test/outflow_tests.jl (:123-193) and test/tree_tests.jl.
"""
import json
import os
import re

import numpy as np
import pytest

import oracle
from treeamr_hip import (Conservative, EvenParity, FieldSet, Forest, GhostSchedule, NoParity,
                         OddParity, Operators, StaleScheduleError, UnsupportedError)
from treeamr_hip import _abi
from treeamr_hip import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# --- the boundary ------------------------------------------------------------

def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "tah.h")).read()
    declared = set(re.findall(r"^\s*(?:tah_status|const char\*)\s+(tah_\w+)\s*\(", header, re.M))
    assert len(declared) > 30
    lib = _abi.lib()
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python bindings type every one of them
    assert declared - {"tah_last_error"} <= set(_abi.SIGNATURES)


# --- weights (test/ghost_tests.jl:27-53, 56-97) ------------------------------

def _lagrange(p, num, den):
    out = np.zeros(p)
    _abi.check(_abi.lib().tah_lagrange_weights(p, num, den, out.ctypes.data_as(_abi.DP)))
    return out


def _conservative(p):
    lo, hi = np.zeros(p), np.zeros(p)
    _abi.check(_abi.lib().tah_conservative_weights(p, lo.ctypes.data_as(_abi.DP),
                                                   hi.ctypes.data_as(_abi.DP)))
    return lo, hi


def test_lagrange_known_answers():
    # [1/4, 3/4] a quarter cell off centre, [1/2, 1/2] at a midpoint,
    # the order-4 midpoint [-1/16, 9/16, 9/16, -1/16], a node [0, 0, 0, 1].
    assert list(_lagrange(2, 3, 4)) == [0.25, 0.75]
    assert list(_lagrange(2, 1, 2)) == [0.5, 0.5]
    assert list(_lagrange(4, 3, 2)) == [-1 / 16, 9 / 16, 9 / 16, -1 / 16]
    assert list(_lagrange(4, 3, 1)) == [0.0, 0.0, 0.0, 1.0]


def test_weights_match_exact_rational_fixtures():
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "weights.json")))
    for e in fx["lagrange"]:
        got = _lagrange(e["p"], e["num"], e["den"])
        assert list(got) == e["float"], e
    for e in fx["conservative"]:
        lo, hi = _conservative(e["p"])
        assert list(lo) == e["low_float"] and list(hi) == e["high_float"], e


def test_conservative_known_answers():
    lo, hi = _conservative(3)
    assert list(lo) == [1 / 8, 1.0, -1 / 8] and list(hi) == [-1 / 8, 1.0, 1 / 8]
    lo, hi = _conservative(1)
    assert list(lo) == [1.0] and list(hi) == [1.0]


# --- operators and their refusals (test/ghost_tests.jl:5-25, 112-212) --------

@pytest.mark.parametrize("kw", [
    dict(prolongation=1, restriction=2), dict(prolongation=3, restriction=2),
    dict(prolongation=2, restriction=0), dict(prolongation=2, restriction=2, family=Conservative),
    dict(prolongation=3, restriction=4, family=Conservative)])
def test_operators_refused(kw):
    with pytest.raises(ValueError):
        Operators(**kw)


def test_operators_need_both_orders():
    with pytest.raises(ValueError, match="no default order"):
        Operators(prolongation=2)


def test_check_operators_against_the_layout():
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    with pytest.raises(ValueError, match="needs G >= 3"):
        GhostSchedule(f, Operators(prolongation=6, restriction=2), G=2)
    f16 = Forest((2, 2, 2), N=6, periodic=(True,) * 3)
    with pytest.raises(ValueError, match="N >="):
        GhostSchedule(f16, Operators(prolongation=6, restriction=6), G=3)
    with pytest.raises(ValueError, match="N must be >= 2G"):
        FieldSet(f, 1, G=5, device="cpu")


def test_bounded_refused_as_unsupported():
    with pytest.raises(UnsupportedError):
        Operators(prolongation=4, restriction=4, bounded=True)


# --- the forest (test/tree_tests.jl) ---------------------------------------

def test_forest_basics_and_balance():
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    assert f.nleaves == 8 and f.generation == 0
    f.refine([0])
    assert f.nleaves == 15 and f.generation == 1
    keys = f.leaves
    assert keys[0] == (0, 1, (0, 0, 0)) and keys[7] == (0, 1, (1, 1, 1))
    f.refine([0])  # a level-2 leaf next to level-0 roots: unbalanced
    assert not f.isbalanced()
    f.balance()
    assert f.isbalanced()
    with pytest.raises(ValueError):
        Forest((2, 2, 2), N=7)
    with pytest.raises(ValueError):
        Forest((2, 2, 2), N=8, periodic=(True, False, False), reflecting=((True, False),) * 3)


def test_config1_mesh_is_120_blocks():
    f = W.config1_forest()
    _, _, level = f.block_geometry()
    assert f.nleaves == 120 and list(np.bincount(level)) == [0, 56, 64]


def test_curve_order_is_sorted_and_unique():
    f = W.config5_forest(roots=2, N=16)
    keys = f.leaves
    assert len(set(keys)) == len(keys)
    # descendants follow their root: roots non-decreasing along the curve
    roots = [k.root for k in keys]
    assert roots == sorted(roots)


# --- the schedule --------------------------------------------------------------

def _ownmask(fs):
    counts_shape = (fs.nblocks,) + tuple(reversed(fs.stored))
    own = np.zeros(counts_shape, bool)
    own[(slice(None),) + fs.own_slices()] = True
    return own


@pytest.mark.parametrize("build", ["config1", "config5_small", "random"])
def test_every_ghost_written_exactly_once(build):
    """The partition oracle (test/ghost_oracles.jl:228-255): no own point is
    written, every ghost exactly once (periodic, reflecting and outflow faces
    leave no hook region)."""
    if build == "config1":
        f, ops, kw, _ = W.config1()
    elif build == "config5_small":
        f, ops, kw, _ = W.config5(nvars=2, roots=4, N=16)
    else:
        f, ops, G, nv, centering = W.random_forest(2, max_blocks=400)
        assert centering == "cell"
        kw = dict(nvars=nv[0], G=G, centering=centering)
    kw = dict(kw)
    fs = FieldSet(f, kw.pop("nvars"), device="cpu", **kw)
    s = GhostSchedule(fs, ops)
    wc = oracle.write_counts(s._descptr, fs.layout)
    own = _ownmask(fs)
    assert wc[own].max() == 0
    assert wc[~own].min() == 1 and wc[~own].max() == 1
    assert s.points() == int(wc.sum())


# --- closed blocks (vertex centering): stage 0, stage P, ghost-plane parts ----

def _positions(fs):
    """Integer position of every stored point, in units of the finest
    spacing, wrapped along periodic dimensions: (nblocks, S2, S1, S0, 3).
    The forest spans the unit cube."""
    lows, h, level = fs.forest.block_geometry()
    scale = 1.0 / h.min()
    top = int(round(scale))
    S, G = fs.stored, fs.G
    out = np.zeros((fs.nblocks, S[2], S[1], S[0], 3), np.int64)
    for d in range(3):
        step = np.rint(h * scale).astype(np.int64)
        base = np.rint(lows[:, d] * scale).astype(np.int64)
        idx = np.arange(S[d]) - G[d]
        x = base[:, None] + idx[None, :] * step[:, None]
        if fs.forest.periodic[d]:
            x = np.mod(x, top)
        shape = [fs.nblocks, 1, 1, 1]
        shape[3 - d] = S[d]
        out[..., d] = x.reshape(shape)
    return out


def _vertex_case(which):
    if which == "config2":
        f, ops, kw, _ = W.config2()
    elif which == "config2_small":
        f, ops, kw, _ = W.config2(roots=2, N=8, radii=(0.5, 0.3))
    else:
        f, ops, G, nv, centering = W.random_forest(int(which[6:]), max_blocks=300)
        assert centering == "vertex"
        kw = dict(nvars=nv[0], G=G, centering="vertex")
    kw = dict(kw)
    fs = FieldSet(f, kw.pop("nvars"), device="cpu", fill=float("nan"), **kw)
    return fs, GhostSchedule(fs, ops)


@pytest.mark.parametrize("which", ["config2", "random3", "random12"])
def test_vertex_ghosts_written_once_own_points_at_most_once(which):
    """write_counts (test/ghost_oracles.jl:228-255) on closed blocks: every
    ghost once; an own point at most once (stage 0 writes secondaries, stage
    P primaries on a reflecting wall, never both)."""
    fs, s = _vertex_case(which)
    wc = oracle.write_counts(s._descptr, fs.layout)
    own = _ownmask(fs)
    assert wc[~own].min() == 1 and wc[~own].max() == 1
    assert wc[own].max() == 1 and wc[own].sum() > 0
    assert s.points() == int(wc.sum())
    tags = s.stage_tags()
    assert 0 in tags and 1 in tags


@pytest.mark.parametrize("which", ["config2", "random3", "random12"])
def test_vertex_one_value_per_point_and_injection(which):
    """After a fill of per-copy noisy data (M22 rule 1): every stored copy of
    an own point holds one value, the value a copy on the finest level held
    before; and every ghost at a point some block of its level or a finer one
    owns holds that value bit for bit (injection_report,
    test/ghost_oracles.jl:186-220)."""
    fs, s = _vertex_case(which)
    rng = np.random.default_rng(5)
    a = fs.array().numpy()
    own = (slice(None), slice(None)) + fs.own_slices()
    a[own] = rng.uniform(-1, 1, size=a[own].shape)
    before = a.copy()
    fac = fs.factors.numpy().copy() if fs.factors is not None else None
    oracle.fill(s._descptr, fs.layout, fs.work.numpy(), fac, nthreads=8)
    _, _, level = fs.forest.block_geometry()
    pos = _positions(fs)
    nb, nv = fs.nblocks, fs.nvars
    ownm = _ownmask(fs)
    # own copies: group by position
    bidx = np.broadcast_to(np.arange(nb)[:, None, None, None], ownm.shape)
    P = pos[ownm]
    B = bidx[ownm]
    vals = np.moveaxis(a, 1, -1)[ownm]           # (npoints, nv)
    vals0 = np.moveaxis(before, 1, -1)[ownm]
    keys, inv = np.unique(P, axis=0, return_inverse=True)
    inv = inv.ravel()
    order = np.argsort(inv, kind="stable")
    first = np.zeros(len(keys), np.int64)
    first[inv[order[::-1]]] = order[::-1]      # some copy of each point
    ref = vals[first]
    assert np.array_equal(vals, ref[inv]), "copies of a point differ after the fill"
    # the value is one a finest copy held before the fill, projected (stage P)
    # onto its parity on a reflecting wall plane: times (1 + sigma)/2 there
    lv = level[B]
    maxlv = np.full(len(keys), -1)
    np.maximum.at(maxlv, inv, lv)
    fine = lv == maxlv[inv]
    proj = np.ones((len(P), nv))
    if fs.parity is not None:
        top = np.rint(1.0 / fs.forest.block_geometry()[1].min()).astype(np.int64)
        for d in range(3):
            for side, at in ((0, P[:, d] == 0), (1, P[:, d] == top)):
                if not fs.forest.reflecting[d][side]:
                    continue
                sig = np.array([-1.0 if fs.parity[v][d] == OddParity else 1.0 for v in range(nv)])
                proj[at] *= (1 + sig) / 2
    hit = np.zeros(len(keys), bool)
    np.logical_or.at(hit, inv, fine & np.all(vals0 * proj + 0.0 == ref[inv], axis=1))
    assert hit.all(), "a point's value is not one of its finest copies' values"
    # ghosts at a point owned on their level or a finer one
    lookup = {tuple(k): i for i, k in enumerate(keys)}
    ghost = ~ownm
    G_pos = pos[ghost]
    G_blk = bidx[ghost]
    G_val = np.moveaxis(a, 1, -1)[ghost]
    checked = 0
    for p_, b_, v_ in zip(G_pos, G_blk, G_val):
        i = lookup.get(tuple(p_))
        if i is None or maxlv[i] < level[b_]:
            continue
        checked += 1
        assert np.array_equal(v_, ref[i]), (b_, p_)
    assert checked > 0


def test_vertex_wall_point_even_kept_odd_zero():
    """The wall point of a reflecting vertex-like dimension is the state's own
    for an even variable, bit for bit, and zero for an odd one (stage P;
    test/reflect_tests.jl:132-154, in 3D with a refined corner)."""
    f = Forest((1, 1, 1), N=8, extents=((0, 1),) * 3, periodic=(False, True, True),
               reflecting=((True, True), (False, False), (False, False)))
    f.refine([0])
    f.balance()
    par = ((EvenParity, NoParity, NoParity), (OddParity, NoParity, NoParity), (EvenParity, NoParity, NoParity))
    fs = FieldSet(f, 3, G=1, centering="vertex", parity=par, device="cpu", fill=float("nan"))
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    X = fs.coordinates()
    a = fs.array().numpy()
    fns = [lambda x: 1 + (x[0] - 1) ** 2, lambda x: 1 + x[0], lambda x: 1 + (x[0] - 1) ** 4]
    for v, fn in enumerate(fns):
        a[(slice(None), v) + fs.own_slices()] = fn(X)[(slice(None),) + fs.own_slices()]
    before = a.copy()
    oracle.fill(s._descptr, fs.layout, fs.work.numpy(), fs.factors.numpy().copy())
    G, N = fs.G[0], f.N
    lows, h, _ = f.block_geometry()
    for b in range(f.nleaves):
        for wall, at_wall in ((G, lows[b, 0] == 0.0), (G + N, lows[b, 0] + N * h[b] == 1.0)):
            if not at_wall:
                continue
            sl = (b, slice(None), slice(G, G + N + 1), slice(G, G + N + 1), wall)
            np.testing.assert_array_equal(a[sl][0], before[sl][0])
            assert (a[sl][1] == 0.0).all()
            np.testing.assert_array_equal(a[sl][2], before[sl][2])


@pytest.mark.parametrize("p", [2, 4, 6])
def test_vertex_pointvalue_polynomial_exactness(p):
    """On closed blocks, order p reproduces degree p-1 everywhere (prolongation
    at coarse nodes and midpoints, injection, stage-0 copies) and not degree p."""
    f = _nested(N=8)
    ops = Operators(prolongation=p, restriction=p)
    G = max(1, p // 2)
    assert _exchange_error(f, ops, _poly(p - 1), G=G, centering="vertex") < 1e-12
    assert _exchange_error(f, ops, _poly(p), G=G, centering="vertex") > 1e-10


def _nested(N, refl=None, outf=None, periodic=False):
    f = Forest((2, 2, 2), N=N, periodic=(periodic,) * 3, extents=((0, 1),) * 3,
               reflecting=refl, outflow=outf)
    for it in range(2):
        c = f.block_centers()
        f.refine([i for i in range(f.nleaves) if np.all(c[i] < 0.5 - 0.1 * it)])
        f.balance()
    return f


def _poly(deg):
    return lambda x, v: sum(0.3 * v + 0.7 * d + 0.11 * (d + v) * x[d] ** e
                            for d in range(3) for e in range(deg + 1))


def _exchange_error(f, ops, fn, G, nvars=2, centering=None):
    """exchange_error (test/ghost_oracles.jl:111-117): own points and the
    outer-face regions hold fn, the fill computes every other ghost; the
    worst deviation from fn anywhere."""
    fs = FieldSet(f, nvars, G=G, centering=centering, device="cpu", fill=float("nan"))
    s = GhostSchedule(fs, ops)
    X = fs.coordinates()
    a = fs.array().numpy()
    for v in range(nvars):
        exact = fn(X, v + 1)
        a[(slice(None), v) + fs.own_slices()] = exact[(slice(None),) + fs.own_slices()]
        for (b, _, lo, hi) in s.boundaries():
            sl = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
            a[(b, v) + sl] = exact[(b,) + sl]
    oracle.fill(s._descptr, fs.layout, fs.work.numpy(), None, nthreads=4)
    assert not np.isnan(a).any(), "a ghost was left unwritten"
    return max(float(np.abs(a[:, v] - fn(X, v + 1)).max()) for v in range(nvars))


@pytest.mark.parametrize("p", [2, 4, 6])
def test_pointvalue_polynomial_exactness(p):
    """Order p reproduces degree p-1 and not degree p (test/ghost_tests.jl:247-281)."""
    f = _nested(N=8)
    ops = Operators(prolongation=p, restriction=p)
    assert _exchange_error(f, ops, _poly(p - 1), G=p // 2) < 1e-12
    assert _exchange_error(f, ops, _poly(p), G=p // 2) > 1e-10


GAUSS_X = [-0.906179845938664, -0.5384693101056831, 0.0, 0.5384693101056831, 0.906179845938664]
GAUSS_W = [0.23692688505618908, 0.47862867049936647, 0.5688888888888889,
           0.47862867049936647, 0.23692688505618908]


def _cell_averages(fn, X, h):
    tot = 0.0
    for i, (xi, wi) in enumerate(zip(GAUSS_X, GAUSS_W)):
        for j, (xj, wj) in enumerate(zip(GAUSS_X, GAUSS_W)):
            for k, (xk, wk) in enumerate(zip(GAUSS_X, GAUSS_W)):
                pt = [X[0] + h * xi / 2, X[1] + h * xj / 2, X[2] + h * xk / 2]
                tot = tot + wi * wj * wk / 8 * fn(pt)
    return tot


@pytest.mark.parametrize("p", [1, 3, 5])
def test_conservative_cell_average_exactness(p):
    """The conservative family reproduces cell averages of degree p-1, not p
    (test/ghost_tests.jl:99-110, test/ghost_oracles.jl:14-84)."""
    f = _nested(N=8)
    G = max(1, (p - 1) // 2)
    ops = Operators(prolongation=p, restriction=2, family=Conservative)

    def err(deg):
        fn = lambda x: sum(0.7 * d + 0.31 * (d + 1) * x[d] ** e for d in range(3) for e in range(deg + 1))
        fs = FieldSet(f, 1, G=G, device="cpu", fill=float("nan"))
        s = GhostSchedule(fs, ops)
        _, h, _ = f.block_geometry()
        X = fs.coordinates()
        hh = h[:, None, None, None]
        exact = _cell_averages(fn, X, hh)
        a = fs.array().numpy()
        a[(slice(None), 0) + fs.own_slices()] = exact[(slice(None),) + fs.own_slices()]
        for (b, _, lo, hi) in s.boundaries():
            sl = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
            a[(b, 0) + sl] = exact[(b,) + sl]
        oracle.fill(s._descptr, fs.layout, fs.work.numpy(), None, nthreads=4)
        return float(np.abs(a[:, 0] - exact).max())

    assert err(p - 1) < 1e-11
    if p > 1:
        assert err(p) > 1e-9


def test_outflow_constant_field_stays_constant():
    """A constant field is 1.0 everywhere after a fill over outflow faces
    (test/outflow_tests.jl:179-193)."""
    out = ((True, True),) * 3
    f = _nested(N=8, outf=out)
    fs = FieldSet(f, 2, G=3, device="cpu", fill=float("nan"))
    a = fs.array().numpy()
    a[(slice(None), slice(None)) + fs.own_slices()] = 1.0
    s = GhostSchedule(fs, Operators(prolongation=6, restriction=6))
    assert s.boundaries() == []
    oracle.fill(s._descptr, fs.layout, fs.work.numpy(), None)
    assert (a == 1.0).all()


def test_outflow_repeats_nearest_own_point_on_a_uniform_mesh():
    """Outflow ghosts repeat the nearest owned point (test/outflow_tests.jl:123-151)."""
    f = Forest((2, 2, 2), N=8, extents=((0, 1),) * 3, outflow=((True, True),) * 3)
    fs = FieldSet(f, 1, G=2, device="cpu", fill=float("nan"))
    rng = np.random.default_rng(3)
    a = fs.array().numpy()
    a[(slice(None), slice(None)) + fs.own_slices()] = rng.uniform(-1, 1, size=(8, 1, 8, 8, 8))
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    oracle.fill(s._descptr, fs.layout, fs.work.numpy(), None)
    # assemble the global 16^3 own field and compare every ghost with it clamped
    root, level, coords = f.leaf_arrays()
    glob = np.zeros((16, 16, 16))
    for b in range(8):
        r = root[b]
        px, py, pz = r % 2, (r // 2) % 2, r // 4
        glob[pz * 8:(pz + 1) * 8, py * 8:(py + 1) * 8, px * 8:(px + 1) * 8] = a[b, 0, 2:10, 2:10, 2:10]
    for b in range(8):
        r = root[b]
        px, py, pz = r % 2, (r // 2) % 2, r // 4
        for k in range(12):
            for j in range(12):
                for i in range(12):
                    gz = np.clip(pz * 8 + k - 2, 0, 15)
                    gy = np.clip(py * 8 + j - 2, 0, 15)
                    gx = np.clip(px * 8 + i - 2, 0, 15)
                    assert a[b, 0, k, j, i] == glob[gz, gy, gx]


def test_reflecting_parities_on_a_uniform_mesh():
    """Across a reflecting face a ghost is its mirror point times the
    variable's parity sign, bit for bit (an even variable mirrored, an odd one
    negated; the reference's factor rule, src/storage.jl:536-550)."""
    f = Forest((1, 1, 2), N=8, extents=((0, 1), (0, 1), (0, 2)), periodic=(True, True, False),
               reflecting=((False, False), (False, False), (True, True)))
    par = ((EvenParity,) * 3, (EvenParity, EvenParity, OddParity))
    fs = FieldSet(f, 2, G=2, device="cpu", fill=float("nan"), parity=par)
    rng = np.random.default_rng(4)
    a = fs.array().numpy()
    a[(slice(None), slice(None)) + fs.own_slices()] = rng.uniform(-1, 1, size=(2, 2, 8, 8, 8))
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    oracle.fill(s._descptr, fs.layout, fs.work.numpy(), fs.factors.numpy())
    # block 0 spans z in [0,1): its low z ghosts k=0,1 mirror own k=3,2
    for v, sign in ((0, 1.0), (1, -1.0)):
        for k, ks in ((0, 3), (1, 2)):
            np.testing.assert_array_equal(a[0, v, k, 2:10, 2:10], sign * a[0, v, ks, 2:10, 2:10] + 0.0)


def test_stale_schedule_refused():
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    fs = FieldSet(f, 1, G=1, device="cpu")
    s = GhostSchedule(fs, Operators(prolongation=2, restriction=2))
    f.refine([0])
    f.balance()
    from treeamr_hip.core import _check_fill
    with pytest.raises(StaleScheduleError):
        _check_fill(fs, s)


def test_check_fill_refusals():
    """check_fill (src/ghosts.jl:2058-2115): a schedule of another forest, ghost
    width, centering or parity, a field set of another block count or element
    type, each refused before any launch."""
    import torch
    from treeamr_hip.core import _check_fill
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3)
    fs = FieldSet(f, 2, G=2, device="cpu")
    ops = Operators(prolongation=4, restriction=4)
    s = GhostSchedule(fs, ops)
    _check_fill(fs, s)
    with pytest.raises(ValueError, match="different forests"):
        _check_fill(FieldSet(Forest((2, 2, 2), N=8, periodic=(True,) * 3), 2, G=2, device="cpu"), s)
    with pytest.raises(ValueError, match="G="):
        _check_fill(FieldSet(f, 2, G=3, device="cpu"), s)
    with pytest.raises(ValueError, match="centering"):
        _check_fill(FieldSet(f, 2, G=2, centering="vertex", device="cpu"), s)
    bad = FieldSet(f, 2, G=2, device="cpu")
    bad.work = bad.work.to(torch.float32)
    with pytest.raises(ValueError, match="Float64"):
        _check_fill(bad, s)
    bad = FieldSet(f, 2, G=2, device="cpu")
    bad.layout.nblocks = 7
    with pytest.raises(ValueError, match="blocks"):
        _check_fill(bad, s)


def test_parity_factor_table():
    par = np.array([[1, 1, 1], [1, 1, 2]], np.int8).ravel()
    out = np.zeros(2 * 27)
    _abi.check(_abi.lib().tah_parity_factors(3, 2, par.ctypes.data_as(_abi.C.POINTER(_abi.C.c_int8)),
                                             out.ctypes.data_as(_abi.DP)))
    f = out.reshape(27, 2)  # column-major: factors[v + nvars*(col-1)]
    assert (f[0] == 1).all()                 # no mirror
    assert list(f[9]) == [1.0, -1.0]         # mirrored along z (state 1): sigma
    assert list(f[18]) == [1.0, 0.0]         # wall projection along z (state 2): (1+sigma)/2


# --- the oracle itself ----------------------------------------------------------

def test_oracle_one_tap_skip_equals_full_sum_on_finite_data():
    """The CPU's one-tap skip and the full sum agree bit for bit on finite data
    (test/ghost_tests.jl:562-634)."""
    f, ops, kw, _ = W.config5(nvars=2, roots=2, N=16)
    kw = dict(kw)
    fs = FieldSet(f, kw.pop("nvars"), device="cpu", **kw)
    W.fill_uniform(fs, 9)
    s = GhostSchedule(fs, ops)
    w1 = fs.work.numpy().copy()
    w2 = w1.copy()
    fac = fs.factors.numpy()
    oracle.fill(s._descptr, fs.layout, w1, fac, nthreads=3, skip=True)
    oracle.fill(s._descptr, fs.layout, w2, fac, nthreads=1, skip=False)
    assert np.array_equal(w1.view(np.int64), w2.view(np.int64))


@pytest.mark.parametrize("which", ["config1", "vertex"])
def test_oracle_adjoint_dot_product(which):
    """<fill(x), y> = <x, fill^T(y)> for the CPU transpose, with x own-only;
    on closed blocks with a reflecting wall too (stage P scales a point onto
    itself, stage 0 copies own points)."""
    if which == "config1":
        f, ops, kw, _ = W.config1()
    else:
        f, ops, kw, _ = W.config2(roots=2, N=8, radii=(0.5, 0.3))
    kw = dict(kw)
    fs = FieldSet(f, kw.pop("nvars"), device="cpu", **kw)
    s = GhostSchedule(fs, ops)
    fac = fs.factors.numpy().copy() if fs.factors is not None else None
    if which == "vertex":
        assert s.stage_tags()[:2] == [-1, 0]
    rng = np.random.default_rng(0)
    own = np.zeros(fs.array().shape, bool)
    own[(slice(None), slice(None)) + fs.own_slices()] = True
    x = np.where(own, rng.uniform(-1, 1, own.shape), 0.0).ravel()
    y = rng.uniform(-1, 1, own.shape).ravel()
    fx = x.copy()
    oracle.fill(s._descptr, fs.layout, fx, fac)
    aty = y.copy()
    oracle.fill_adjoint(s._descptr, fs.layout, aty, fac)
    lhs = float(fx @ y)
    rhs = float(x @ aty)
    assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), 1.0)


def test_arena_layout_gives_the_same_values_as_the_dense_one():
    """The variable-innermost arena holds the same field; the oracle's fill
    over it equals the dense fill value for value."""
    f, ops, kw, _ = W.config5(nvars=3, roots=2, N=16)
    kw = dict(kw)
    nv = kw.pop("nvars")
    dense = FieldSet(f, nv, device="cpu", **kw)
    arena = FieldSet(f, nv, device="cpu", layout="arena", **kw)
    assert arena.layout.stride[3] == 1 and arena.layout.stride[0] == nv
    W.fill_uniform(dense, 5)
    arena.array()[...] = dense.array()
    s = GhostSchedule(dense, ops)
    oracle.fill(s._descptr, dense.layout, dense.work.numpy(), dense.factors.numpy())
    oracle.fill(s._descptr, arena.layout, arena.work.numpy(), arena.factors.numpy())
    assert np.array_equal(dense.array().numpy().view(np.int64), arena.array().numpy().view(np.int64))


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("nvars", [1, 2])
def test_layouts_of_every_dimension_validate(D, nvars):
    """Dense and arena layouts of 1, 2 and 3 dimensions and of one variable
    (axes of extent one, which add no offset) pass the stride check; a layout
    whose axes overlap is refused (tah_layout_size)."""
    import ctypes as C
    f = Forest((2,) * D, N=8)
    for layout in ("dense", "arena"):
        fs = FieldSet(f, nvars, G=2, device="cpu", layout=layout)
        assert fs.nelements >= fs.nblocks * nvars * int(np.prod(fs.stored))
    bad = _abi.tah_layout()
    for d in range(3):
        bad.stored[d] = 4
    bad.nvars, bad.nblocks = 2, 3
    for d, s in enumerate((1, 4, 16, 32, 128)):
        bad.stride[d] = s
    bad.stride[3] = 8  # variables over the z rows
    n = C.c_int64()
    with pytest.raises(_abi.LayoutError, match="overlap"):
        _abi.check(_abi.lib().tah_layout_size(C.byref(bad), C.byref(n)))
