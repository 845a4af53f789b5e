"""CPU tests that pin the schedule builder and the oracle by what the
reference's own definitional tests and documented rules fix, independently
of this library's restatement of them:

* node positions, one rounding (nodeposition, src/storage.jl:1008-1041);
* stage 0's primary copy of a shared point (piece_primary,
  src/schedule.jl:1914-1921): finest level, then the greatest class tuple in
  the copy's own frame;
* the tap order of stencil_sum (src/ghosts.jl:58-118), by known answers on
  data whose every other order rounds differently;
* periodic wraparound = an explicit tiling, bit for bit (periodic_vs_tiled,
  test/ghost_oracles.jl:358-399; test/ghost_tests.jl:455-466);
* a reflecting face = the mirrored doubled domain (reflecting_vs_doubled,
  test/ghost_oracles.jl:716-765; test/reflect_tests.jl:156-176), on meshes
  whose refinement meets the wall.

The fills here run the oracle (the restated reference CPU fill) on the
library's schedules; tests/test_gpu_pinning.py runs the same cases through
the HIP path.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from treeamr_hip import EvenParity, FieldSet, Forest, GhostSchedule, NoParity, OddParity, Operators
from treeamr_hip import workloads as W

from test_host import _ownmask, _positions, _vertex_case


def _fill(fs, s):
    """The oracle's fill of a CPU field set; the library's HIP fill of a
    device one."""
    if fs.work.device.type == "cuda":
        import torch

        from treeamr_hip import fill_ghosts
        fill_ghosts(fs, s)
        torch.cuda.synchronize()
        return
    fac = fs.factors.numpy().copy() if fs.factors is not None else None
    oracle.fill(s._descptr, fs.layout, fs.work.numpy(), fac, nthreads=8)


def _host(fs):
    """fs.array() as a host numpy array [block, var, k, j, i] (a copy for a
    device field set)."""
    return fs.array().cpu().numpy() if fs.work.device.type == "cuda" else fs.array().numpy()


def _set_own(fs, values):
    """Write host values [block, var, k, j, i] of the own ranges into fs."""
    import torch
    own = (slice(None), slice(None)) + fs.own_slices()
    fs.array()[own] = torch.as_tensor(values, dtype=torch.float64).to(fs.work.device)


# --- positions --------------------------------------------------------------

def test_positions_round_once():
    """fma(k, h, lo) with one rounding: against the exact rational lo + k*h
    correctly rounded, on a brick whose low corner is not a grid node."""
    f = Forest((3, 1, 1), N=8, extents=((0.1, 0.4), (0.1, 0.2), (0.0, 0.1)))
    f.refine([1])
    h0 = (0.4 - 0.1) / 3 / 8
    root, level, coords = f.leaf_arrays()
    for c in (0, 1):
        fs = FieldSet(f, 1, G=2, centering="vertex" if c else "cell", device="cpu")
        p = fs.positions(0)
        for b in range(f.nleaves):
            lvl = int(level[b])
            h = Fraction(h0) / 2 ** lvl
            B = ((int(root[b]) % 3 << lvl) + int(coords[b][0])) * 8
            for j in range(fs.stored[0]):
                k = B + j - 2
                exact = Fraction(0.1) + (k * h if c else (2 * k + 1) * h / 2)
                assert p[b, j] == float(exact), (c, b, j)


def test_positions_anchored_at_zero_are_antisymmetric():
    """Where zero is a node of the level-0 grid, positions count from zero
    (src/geometry.jl:76-91): a brick centred on zero gives x and -x bit for
    bit (a two-rounding lo + (I + 1/2)h from the low corner does not)."""
    f = Forest((6, 2, 2), N=8, extents=((-0.3, 0.3), (0.0, 0.2), (0.0, 0.2)))
    f.refine([0, 5])
    f.balance()
    fs = FieldSet(f, 1, G=3, device="cpu")
    xs = np.unique(fs.positions(0))
    assert np.array_equal(np.sort(-xs), xs)


def test_every_copy_of_a_point_has_one_position_config2():
    """Every stored copy of a point inside the domain (own copies on several
    blocks, ghosts of the same level or a coarser or finer one) is at one
    position, bit for bit, on config 2's nested vertex mesh: the property that
    makes a field filled by coordinates agree across copies."""
    fs, _ = _vertex_case("config2")
    X = fs.coordinates()
    key = _positions(fs).reshape(-1, 3)
    per = np.stack([x.reshape(-1) for x in X], axis=1)
    inside = np.all((per >= 0) & (per <= 1), axis=1)
    key, per = key[inside], per[inside]
    order = np.lexsort(key.T[::-1])
    k, p = key[order], per[order]
    same = np.all(k[1:] == k[:-1], axis=1)
    assert same.sum() > 1000
    assert np.array_equal(p[1:][same].view(np.int64), p[:-1][same].view(np.int64))


# --- stage 0's primary ------------------------------------------------------

def _classes(fs):
    """Per stored point and dimension: its class in the block's own closed box
    (0 on the low face, 2 on the high face, 1 between; -1 for a ghost)."""
    S, G, N = fs.stored, fs.G, fs.forest.N
    out = np.full((fs.nblocks, S[2], S[1], S[0], 3), -1, np.int64)
    for d in range(3):
        j = np.arange(S[d]) - G[d]
        c = np.where(j == 0, 0, np.where(j == N, 2, 1))
        c = np.where((j < 0) | (j > N), -1, c)
        shape = [1, 1, 1, 1]
        shape[3 - d] = S[d]
        out[..., d] = c.reshape(shape)
    return out


@pytest.mark.parametrize("which", ["config2", "config2_small", "random3", "random12"])
def test_stage0_value_is_the_primary_copys(which):
    """After a fill of per-copy noisy data, every copy of an own point holds
    the value its primary copy held before (times stage P's projection on a
    reflecting wall plane): among the copies on the finest level, the one
    with the greatest class tuple in its own frame, compared dimension 1
    first (piece_primary, src/schedule.jl:1914-1921; one patch and no seams,
    so positions compare equal, a periodic wrap included)."""
    check_primary(*_vertex_case(which))


def vertex_case(which, device, layout="dense"):
    fs, s = _vertex_case(which)
    if device == "cpu" and layout == "dense":
        return fs, s
    f, ops = fs.forest, s.operators
    fs = FieldSet(f, fs.nvars, G=fs.G, centering=fs.centering, parity=fs.parity, device=device, layout=layout,
                  fill=float("nan"))
    return fs, GhostSchedule(fs, ops)


def check_primary(fs, s):
    rng = np.random.default_rng(11)
    own = (slice(None), slice(None)) + fs.own_slices()
    shape = _host(fs)[own].shape
    _set_own(fs, rng.uniform(-1, 1, size=shape))
    before = _host(fs).copy()
    _fill(fs, s)
    a = _host(fs)
    _, _, level = fs.forest.block_geometry()
    ownm = _ownmask(fs)
    nb, nv = fs.nblocks, fs.nvars
    P = _positions(fs)[ownm]
    B = np.broadcast_to(np.arange(nb)[:, None, None, None], ownm.shape)[ownm]
    cls = _classes(fs)[ownm]
    vals = np.moveaxis(a, 1, -1)[ownm]
    vals0 = np.moveaxis(before, 1, -1)[ownm]
    keys, inv = np.unique(P, axis=0, return_inverse=True)
    inv = inv.ravel()
    # sort copies by (point, -level, -class1, -class2, -class3): the first of
    # each point's run is its primary
    lv = level[B]
    order = np.lexsort((-cls[:, 2], -cls[:, 1], -cls[:, 0], -lv, inv))
    first_of = np.full(len(keys), -1)
    runs = np.r_[True, inv[order][1:] != inv[order][:-1]]
    first_of[inv[order][runs]] = order[runs]
    proj = np.ones((len(P), nv))
    if fs.parity is not None:
        top = np.rint(1.0 / fs.forest.block_geometry()[1].min()).astype(np.int64)
        for d in range(3):
            for side, at in ((0, P[:, d] == 0), (1, P[:, d] == top)):
                if not fs.forest.reflecting[d][side]:
                    continue
                sig = np.array([-1.0 if fs.parity[v][d] == OddParity else 1.0 for v in range(nv)])
                proj[at] *= (1 + sig) / 2
    prim = first_of[inv]
    expected = vals0[prim] * proj[prim] + 0.0
    shared = np.bincount(inv) > 1
    assert shared.sum() > 100
    bad = ~np.all(vals.view(np.int64) == expected.view(np.int64), axis=1)
    assert not bad.any(), f"{bad.sum()} copies differ from their primary's value"
    # and the rule decides: some shared point's finest copies held different
    # values, so taking another of them would have failed
    lvmax = np.full(len(keys), -1)
    np.maximum.at(lvmax, inv, lv)
    fine = lv == lvmax[inv]
    other = fine & (np.arange(len(P)) != prim) & np.any(vals0 != vals0[prim], axis=1)
    assert other.sum() > 100


# --- the tap order ----------------------------------------------------------

def tap_case(device="cpu", layout="dense"):
    """Two periodic roots of 8^3 along x, the second refined: cell, G=3,
    PointValue(6,6), two variables; own points +-10^e with e uniform in
    [-6, 15], so that the sum of a stencil rounds differently in every
    order."""
    f = Forest((2, 1, 1), N=8, periodic=(True,) * 3, extents=((0.0, 2.0), (0.0, 1.0), (0.0, 1.0)))
    f.refine([1])
    f.balance()
    fs = FieldSet(f, 2, G=3, device=device, layout=layout, fill=float("nan"))
    rng = np.random.default_rng(2024)
    own = (slice(None), slice(None)) + fs.own_slices()
    shape = _host(fs)[own].shape
    _set_own(fs, rng.choice([-1.0, 1.0], size=shape) * 10.0 ** rng.uniform(-6, 15, size=shape))
    return fs, GhostSchedule(fs, Operators(prolongation=6, restriction=6))


# (kind, target block, variable, k, j, i) -> the bits the reference's sum
# forms: acc = +0, then for m3, m2, m1 ascending (m1 innermost) acc += w * u
# with w = ((w1[m1] * w2[m2]) * w3[m3]), every product and sum rounded on its
# own (src/ghosts.jl:58-118), over the transfer's window of the input data.
# Derived once from the schedule's weights and _tap_case's input by that
# loop nest.
TAP_KATS = {
    ("prolong", 1, 0, 4, 5, 2): 0x42D174B9EF1085EE,
    ("prolong", 1, 1, 0, 0, 0): 0xC293D97EA5FB1BEB,
    ("prolong", 3, 0, 9, 9, 1): 0xC29DA20C3E8B0EC5,
    ("restrict", 0, 0, 6, 4, 13): 0xC2F0BACFA398E399,
    ("restrict", 0, 1, 2, 12, 12): 0xC2D5C35DA9C9D144,
    ("restrict", 0, 0, 3, 3, 0): 0xC2D0B2968170B439,
}


def _writer(s, b, k, j, i):
    """The (group, transfer) writing stored point (i, j, k) of block b."""
    for g in s.groups():
        st = g["stencils"]
        for t in np.nonzero(g["targetblocks"] == b)[0]:
            if all(st[d]["targetfirst"] <= x < st[d]["targetfirst"] + st[d]["ntarget"]
                   for d, x in enumerate((i, j, k))):
                return g, int(t)
    raise AssertionError("no writer")


def _stencil_sum(a, g, t, v, k, j, i, order="reference"):
    st = g["stencils"]
    r = [x - st[d]["targetfirst"] for d, x in enumerate((i, j, k))]
    ss = [int(st[d]["srcstart"][r[d]]) for d in range(3)]
    w = [st[d]["weights"][r[d]] for d in range(3)]
    src = a[int(g["sourceblocks"][t]), v]
    p = [st[d]["order"] for d in range(3)]
    taps = [(m3, m2, m1) for m3 in range(p[2]) for m2 in range(p[1]) for m1 in range(p[0])]
    if order == "m3 innermost":
        taps = [(m3, m2, m1) for m1 in range(p[0]) for m2 in range(p[1]) for m3 in range(p[2])]
    elif order == "m2 innermost":
        taps = [(m3, m2, m1) for m3 in range(p[2]) for m1 in range(p[0]) for m2 in range(p[1])]
    acc = np.float64(0.0)
    for m3, m2, m1 in taps:
        if order == "w1*(w2*w3)":
            wt = np.float64(w[0][m1]) * (np.float64(w[1][m2]) * np.float64(w[2][m3]))
        else:
            wt = (np.float64(w[0][m1]) * np.float64(w[1][m2])) * np.float64(w[2][m3])
        u = np.float64(src[ss[2] + m3, ss[1] + m2, ss[0] + m1])
        acc = acc + wt * u
    if order == "descending":
        acc = np.float64(0.0)
        for m3, m2, m1 in reversed(taps):
            acc = acc + (np.float64(w[0][m1]) * np.float64(w[1][m2])) * np.float64(w[2][m3]) * \
                np.float64(src[ss[2] + m3, ss[1] + m2, ss[0] + m1])
    return acc


def test_tap_order_known_answers():
    """Six ghosts (prolongation into the fine blocks, restriction into the
    coarse one, faces, edges and corners) against the bits of the
    reference's sum. On this data every other summation order gives other
    bits; the weight products of these operators are dyadic with short
    numerators, exact in Float64, so their association cannot change a bit
    (checked too)."""
    fs, s = tap_case()
    _fill(fs, s)
    check_tap_kats(_host(fs), s)


def check_tap_kats(a, s):
    others = {"m3 innermost": set(), "m2 innermost": set(), "descending": set()}
    for (kind, b, v, k, j, i), bits in TAP_KATS.items():
        g, t = _writer(s, b, k, j, i)
        assert g["kind"] == (3 if kind == "prolong" else 2) and g["stencils"][0]["order"] == 6
        got = int(a[b, v, k, j, i].view(np.int64)) & 0xFFFFFFFFFFFFFFFF
        assert got == bits, (kind, b, v, k, j, i, hex(got))
        # the data decides the order (sources read after the fill: a
        # prolongation's coarse window holds ghosts an earlier stage wrote)
        for other, differs in others.items():
            alt = _stencil_sum(a, g, t, v, k, j, i, order=other)
            if int(np.float64(alt).view(np.int64)) & 0xFFFFFFFFFFFFFFFF != bits:
                differs.add(kind)
        w = [g["stencils"][d]["weights"].ravel() for d in range(3)]
        for x in w[0]:
            for y in w[1]:
                assert np.all((np.float64(x) * y) * w[2] == np.float64(x) * (y * w[2]))
                assert all(Fraction(float(x)) * Fraction(float(y)) * Fraction(float(z)) ==
                           Fraction(float((np.float64(x) * y) * z)) for z in w[2])
    # every other order is caught, for prolongation and for restriction
    for other, differs in others.items():
        assert differs == {"prolong", "restrict"}, (other, differs)


# --- periodic = tiled -------------------------------------------------------

def _refine_where(f, pred, passes):
    for _ in range(passes):
        c = f.block_centers()
        _, _, level = f.block_geometry()
        idx = [b for b in range(f.nleaves) if pred(c[b], int(level[b]))]
        if not idx:
            break
        f.refine(idx)
        f.balance()
    return f


def _by_coordinates(fs, fn):
    X = fs.coordinates()
    own = fs.own_slices()
    _set_own(fs, np.stack([fn(X, v + 1)[(slice(None),) + own] for v in range(fs.nvars)], axis=1))


def periodic_vs_tiled(M, *, N=4, G=1, ops=(2, 2), centering="cell", passes=2, nvars=2, device="cpu",
                      layout="dense"):
    """An M-root periodic domain and the middle tile of a 3M-root tiling of
    the same data (outflow faces outside, far from the middle tile): the
    pairs of fills, and the block pairs with one level and lower corner."""
    L = float(M)

    def seam(c, lvl):
        return lvl < passes and all(min(c[d] % L, L - c[d] % L) < 0.55 for d in range(3))

    per = _refine_where(Forest((M,) * 3, N=N, periodic=(True,) * 3, extents=((0.0, L),) * 3), seam, passes)
    til = _refine_where(Forest((3 * M,) * 3, N=N, outflow=((True, True),) * 3, extents=((-L, 2 * L),) * 3),
                        seam, passes)

    def data(X, v):
        return sum(np.sin(3.1 * np.mod(X[d], L) + 0.7 * v) * (1 + 0.3 * (d + 1)) for d in range(3))

    O = Operators(prolongation=ops[0], restriction=ops[1])
    out = []
    for f in (per, til):
        fs = FieldSet(f, nvars, G=G, centering=centering, device=device, layout=layout, fill=float("nan"))
        _by_coordinates(fs, data)
        _fill(fs, GhostSchedule(fs, O))
        out.append(fs)
    lows_p, _, lv_p = per.block_geometry()
    lows_t, _, lv_t = til.block_geometry()
    index = {(int(l), tuple(x)): b for b, (l, x) in enumerate(zip(lv_t, lows_t))}
    pairs = [(b, index[(int(l), tuple(x))]) for b, (l, x) in enumerate(zip(lv_p, lows_p))]
    return out[0], out[1], pairs


@pytest.mark.parametrize("M,N,G,ops,centering", [(1, 4, 1, (2, 2), "cell"), (2, 4, 1, (2, 2), "cell"),
                                                   (1, 8, 2, (4, 4), "cell"), (1, 8, 3, (6, 6), "cell"),
                                                   (1, 8, 2, (4, 4), "vertex")])
def test_periodic_equals_tiled(M, N, G, ops, centering):
    p, t, pairs = periodic_vs_tiled(M, N=N, G=G, ops=ops, centering=centering)
    ap, at = _host(p), _host(t)
    assert len(pairs) == p.nblocks and p.nblocks > 8
    assert not np.isnan(ap).any()
    for b, tb in pairs:
        assert np.array_equal(ap[b].view(np.int64), at[tb].view(np.int64)), b


# --- reflecting = doubled ---------------------------------------------------

def reflecting_vs_doubled(side, centering, *, N=8, p=4, device="cpu", layout="dense"):
    """A half domain with a reflecting face at x = 0 against the doubled
    domain holding the mirrored data: the two field sets and the block pairs
    with one level and lower corner. Outflow on the outer x faces (far from
    the wall), periodic elsewhere."""
    G = max(1, p // 2 - 1) if centering == "vertex" else max(1, math.ceil(p / 2))
    xext = (0.0, 2.0) if side == "lo" else (-2.0, 0.0)
    half = Forest((2, 2, 2), N=N, periodic=(False, True, True),
                  reflecting=((side == "lo", side == "hi"), (False, False), (False, False)),
                  outflow=((side == "hi", side == "lo"), (False, False), (False, False)),
                  extents=(xext, (0.0, 2.0), (0.0, 2.0)))
    full = Forest((4, 2, 2), N=N, periodic=(False, True, True), outflow=((True, True), (False, False), (False, False)),
                  extents=((-2.0, 2.0), (0.0, 2.0), (0.0, 2.0)))

    def pred(c, lvl):
        return (lvl == 0 and abs(c[0]) < 1 and c[1] < 1) or (lvl == 1 and abs(c[0]) < 0.5 and c[1] < 0.5)

    _refine_where(half, pred, 2)
    _refine_where(full, pred, 2)

    def shape(X, v):
        out = np.sin(3.1 * np.abs(X[0]) + 0.7 * v)
        for d in (1, 2):
            out = out * (1 + 0.3 * np.sin(np.pi * X[d] + 0.2 * (d + 1)))
        return out

    def data(X, v):
        return shape(X, v) if v == 1 else np.sign(X[0]) * shape(X, v)

    par = ((EvenParity, NoParity, NoParity), (OddParity, NoParity, NoParity))
    ops = Operators(prolongation=p, restriction=p)
    fsh = FieldSet(half, 2, G=G, centering=centering, parity=par, device=device, layout=layout, fill=float("nan"))
    fsf = FieldSet(full, 2, G=G, centering=centering, device=device, layout=layout, fill=float("nan"))
    for fs in (fsh, fsf):
        _by_coordinates(fs, data)
        _fill(fs, GhostSchedule(fs, ops))
    lows_h, _, lv_h = half.block_geometry()
    lows_f, _, lv_f = full.block_geometry()
    index = {(int(l), tuple(x)): b for b, (l, x) in enumerate(zip(lv_f, lows_f))}
    pairs = [(b, index[(int(l), tuple(x))]) for b, (l, x) in enumerate(zip(lv_h, lows_h))]
    return fsh, fsf, pairs


@pytest.mark.parametrize("side,centering", [("lo", "cell"), ("hi", "cell"), ("lo", "vertex"), ("hi", "vertex")])
def test_reflecting_equals_doubled(side, centering):
    """To roundoff (a mirrored stencil sums the same terms in the other
    order), every stored value of every half-domain block, on a mesh whose
    coarse-fine faces meet the wall."""
    h, f, pairs = reflecting_vs_doubled(side, centering)
    ah, af = _host(h), _host(f)
    _, _, lv = h.forest.block_geometry()
    assert len(set(lv)) == 3 and len(pairs) == h.nblocks
    assert not np.isnan(ah).any()
    worst = max(np.abs(ah[b] - af[fb]).max() for b, fb in pairs)
    assert worst < 1e-13, worst


# --- stage groups: any order within a group ---------------------------------

def restaged(s, mode):
    """The schedule's flattened descriptor with its stages rearranged: mode
    "reversed" runs the stages of every stage group in reverse order
    (fill_stagewise!(...; reverse=true), test/ghost_oracles.jl:314-329),
    "alone" makes every stage a stage group of its own. Returns (pointer to
    the new descriptor, objects it points into)."""
    import ctypes as C

    from treeamr_hip import _abi
    d = s.desc
    sgs = [(d.stagegroup_first[k], d.stagegroup_last[k]) for k in range(d.nstagegroups)]
    newidx = list(range(d.nstages))
    if mode == "reversed":
        for a, b in sgs:
            for st in range(a, b + 1):
                newidx[st] = a + b - st
    elif mode == "alone":
        sgs = [(k, k) for k in range(d.nstages)]
    tags = [0] * d.nstages
    for st in range(d.nstages):
        tags[newidx[st]] = d.stage_tags[st]
    groups = [d.groups[gi] for gi in range(d.ngroups)]
    order = sorted(range(len(groups)), key=lambda gi: newidx[groups[gi].stage])
    garr = (_abi.tah_group_desc * max(1, len(groups)))()
    for k, gi in enumerate(order):
        C.memmove(C.byref(garr[k]), C.byref(groups[gi]), C.sizeof(_abi.tah_group_desc))
        garr[k].stage = newidx[groups[gi].stage]
    tag_arr = (C.c_int32 * d.nstages)(*tags)
    first = (C.c_int32 * len(sgs))(*[a for a, _ in sgs])
    last = (C.c_int32 * len(sgs))(*[b for _, b in sgs])
    nd = _abi.tah_schedule_desc()
    C.memmove(C.byref(nd), C.byref(d), C.sizeof(nd))
    nd.stage_tags = C.cast(tag_arr, C.POINTER(C.c_int32))
    nd.nstagegroups = len(sgs)
    nd.stagegroup_first = C.cast(first, C.POINTER(C.c_int32))
    nd.stagegroup_last = C.cast(last, C.POINTER(C.c_int32))
    nd.groups = C.cast(garr, C.POINTER(_abi.tah_group_desc))
    return C.pointer(nd), (nd, garr, tag_arr, first, last, s)


def grouped_case(which, device="cpu", layout="dense"):
    if which == "config2":
        f, ops, kw, _ = W.config2()
        kw = dict(kw)
        fs = FieldSet(f, kw.pop("nvars"), device=device, layout=layout, fill=float("nan"), **kw)
        W.fill_waves(fs, 1)
    else:
        if which == "config5_small":
            f, ops, kw, _ = W.config5(nvars=2, roots=4, N=16)
        else:
            f, ops, G, nv, centering = W.random_forest(int(which[6:]), max_blocks=300)
            kw = dict(nvars=nv[0], G=G, centering=centering)
        kw = dict(kw)
        fs = FieldSet(f, kw.pop("nvars"), device=device, layout=layout, fill=float("nan"), **kw)
        W.fill_uniform(fs, 3)
    s = GhostSchedule(fs, ops)
    return fs, s


@pytest.mark.parametrize("which", ["config2", "config5_small", "random10"])
def test_grouped_fill_in_any_order_within_its_groups(which):
    """From NaN in every ghost: the fill, the fill with every stage group's
    stages in reverse order, and the fill one stage at a time agree in every
    bit and leave no NaN (test/ghost_tests.jl:399-453): a level grouped with
    one it depends on would read a NaN."""
    fs, s = grouped_case(which)
    assert any(b > a for a, b in s.stagegroups())
    init = fs.work.numpy().copy()
    fac = fs.factors.numpy().copy() if fs.factors is not None else None
    out = []
    for mode in ("grouped", "reversed", "alone"):
        w = init.copy()
        if mode == "grouped":
            oracle.fill(s._descptr, fs.layout, w, fac, nthreads=8)
        else:
            ptr, _keep = restaged(s, mode)
            oracle.fill(ptr, fs.layout, w, fac, nthreads=8)
        assert not np.isnan(w).any(), mode
        out.append(w.view(np.int64))
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
