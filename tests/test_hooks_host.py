"""The outer-face hook on the CPU: the schedule's hook regions (clamps and
slots), the oracle's fill with the hook between its stage groups, and the
Python forms of CellBoundary, against what the geometry and the mathematics
fix (test/relative_tests.jl, test/ghost_tests.jl:399-453 restated):

- the nearest evolved point of every hook ghost, from the geometry alone (a
  component leaves the domain where the block touches the domain's face),
  equals the one the schedule's clamps give, and the relative condition
  holds there bit for bit after a fill, corner rows included;
- on a single block the relative hook is the region form that clamps every
  outward component;
- a linear case is reproduced by the relative hook and by the values hook
  to rounding, from NaN ghosts (a hook run before the ghosts it reads are
  filled, or after the prolongations that read it, leaves NaN);
- one fill is the relative condition's fixed point;
- Relative(0) on outer faces is an outflow face, bit for bit (src/ghosts.jl:
  1431, "f = 0 is a zero-gradient outflow");
- stage groups that hold hook slots fill alike in any order within a group
  and one stage at a time;
- the forms the reference refuses are refused.
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import (AllVariables, CellBoundary, FieldSet, Forest, GhostSchedule, Operators, Relative,
                         UnsupportedError, WithPatch, boundary_by_coordinates, boundary_values,
                         cellcentered, fill_by_coordinates, vertexcentered)
from treeamr_hip import workloads as W
from treeamr_hip._abi import TAH_HOOK_RELATIVE, TAH_HOOK_VALUES

from test_pinning_host import restaged


def nested_forest(D, N=8, roots=4):
    """test/ghost_oracles.jl:275-284: roots^D of [0, roots]^D, every face an
    outer one, level 1 within 1.0 of (1.5, ...), level 2 within 0.3."""
    f = Forest((roots,) * D, N=N, extents=((0.0, float(roots)),) * D)
    for lvl, r in ((0, 1.0), (1, 0.3)):
        c = f.block_centers()
        _, _, level = f.block_geometry()
        near = np.all(np.abs(c - 1.5) <= r, axis=1) & (level == lvl)
        if near.any():
            f.refine(np.nonzero(near)[0])
            f.balance()
    return f


def case_fn(D):
    """relative_case (test/relative_tests.jl:41): sum over d of sin(1.3 x_d + d), d 1-based."""
    return lambda x: sum(torch.sin(1.3 * x[d] + (d + 1)) for d in range(D))


def _fs(forest, C, G=2, nvars=1, seed=3, nan=False):
    fs = FieldSet(forest, nvars, G=G, centering=C, device="cpu", fill=float("nan") if nan else 0.0)
    W.fill_uniform(fs, seed)
    return fs


def _oracle_hook(fs, s, hook, nthreads=4, desc=None):
    kind, values, fr = boundary_values(fs, s, hook)
    w = fs.work.numpy()
    fac = fs.factors.numpy().copy() if fs.factors is not None else None
    oracle.fill_boundary(desc if desc is not None else s._descptr, fs.layout, w, fac, kind,
                         values.numpy(), fr.numpy() if fr is not None else None, nthreads=nthreads)
    return kind, values, fr


def _geometric_r(fs, b, delta, idx, lows, h):
    """oracle_relative_point (test/relative_tests.jl:14-23), 0-based: a
    component leaves where the block touches the domain's face on that side
    and the dimension is not periodic."""
    f = fs.forest
    N, out = f.N, []
    for d in range(3):
        if d >= f.D or f.periodic[d] or delta[d] == 0:
            out.append(idx[d])
            continue
        lo = lows[b, d]
        hi = lo + h[b] * N
        if delta[d] > 0 and hi == f.extents[d][1]:
            out.append(fs.G[d] + N + fs.centering[d] - 1)
        elif delta[d] < 0 and lo == f.extents[d][0]:
            out.append(fs.G[d])
        else:
            out.append(idx[d])
    return tuple(out)


def _regions(s):
    d = s.desc
    return [d.boundaries[r] for r in range(d.nboundaries)]


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("C", ["cell", "vertex"])
def test_relative_condition_with_geometric_points(D, C):
    """test/relative_tests.jl:44-62: on a mesh whose coarse-fine faces meet
    the outer faces, every hook ghost is f(x_g) + (u(x_r) - f(x_r)) bit for
    bit, x_r found from the geometry (not from the schedule), some of them
    across a block face, and some hook regions run after a prolongation."""
    forest = nested_forest(D)
    fs = _fs(forest, cellcentered(D) if C == "cell" else vertexcentered(D))
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    f = case_fn(D)
    _oracle_hook(fs, s, CellBoundary(Relative(f)))
    a = fs.array().numpy()
    lows, h, _ = forest.block_geometry()
    pos = [fs.positions(d) for d in range(D)]
    n = ncorner = nslot = 0
    worst = 0.0
    for R in _regions(s):
        b, delta = R.block, tuple(R.direction)
        nslot += R.slot > 1
        for k in range(R.lo[2], R.hi[2]):
            for j in range(R.lo[1], R.hi[1]):
                for i in range(R.lo[0], R.hi[0]):
                    idx = (i, j, k)
                    r = _geometric_r(fs, b, delta, idx, lows, h)
                    assert r == tuple(R.clamps[d] if R.clamps[d] >= 0 else idx[d] for d in range(3))
                    xg = tuple(torch.tensor([pos[d][b, idx[d]]]) for d in range(D))
                    xr = tuple(torch.tensor([pos[d][b, r[d]]]) for d in range(D))
                    want = (f(xg) + (torch.tensor([a[b, 0, r[2], r[1], r[0]]]) - f(xr))).item()
                    got = a[b, 0, k, j, i]
                    worst = max(worst, 0.0 if got == want else abs(got - want) + 1.0)
                    n += 1
                    ncorner += any(delta[d] != 0 and r[d] == idx[d] for d in range(D))
    assert n > 0 and worst == 0.0
    assert ncorner > 0
    assert nslot > 0, "no hook region reads prolongated ghosts on this mesh"


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("C", ["cell", "vertex"])
def test_single_block_relative_is_the_region_form(D, C):
    """test/relative_tests.jl:64-95: on one block every region leaves the
    domain along every outward component, so clamping the whole region onto
    the block after the fill is the condition."""
    forest = Forest((1,) * D, N=8, extents=((0.0, 1.0),) * D)
    cen = cellcentered(D) if C == "cell" else vertexcentered(D)
    ops = Operators(prolongation=4, restriction=4)
    f = case_fn(D)
    a_fs = _fs(forest, cen)
    b_fs = _fs(forest, cen)
    s = GhostSchedule(a_fs, ops)
    _oracle_hook(a_fs, s, CellBoundary(Relative(f)))
    w = b_fs.work.numpy()
    oracle.fill(s._descptr, b_fs.layout, w, None, nthreads=1)
    a = b_fs.array()
    pos = [b_fs.positions(d) for d in range(D)]
    G, N, c = b_fs.G, forest.N, b_fs.centering
    for R in _regions(s):
        for k in range(R.lo[2], R.hi[2]):
            for j in range(R.lo[1], R.hi[1]):
                for i in range(R.lo[0], R.hi[0]):
                    idx = (i, j, k)
                    r = tuple(idx[d] if d >= D or R.direction[d] == 0 else
                              (G[d] + N + c[d] - 1 if R.direction[d] > 0 else G[d]) for d in range(3))
                    xg = tuple(torch.tensor([pos[d][0, idx[d]]]) for d in range(D))
                    xr = tuple(torch.tensor([pos[d][0, r[d]]]) for d in range(D))
                    a[0, 0, k, j, i] = f(xg) + (a[0, 0, r[2], r[1], r[0]] - f(xr))
    assert torch.equal(a_fs.work.view(torch.int64), b_fs.work.view(torch.int64))


def _linear(D):
    return lambda x: 0.3 + sum((0.2 + 0.1 * (d + 1)) * x[d] for d in range(D))


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("C", ["cell", "vertex"])
@pytest.mark.parametrize("form", ["relative", "values", "relative-all", "by-coordinates"])
def test_linear_case_reproduced_from_nan_ghosts(D, C, form):
    """test/relative_tests.jl:97-118 and the values form alike: own data f
    linear, every ghost NaN; after the fill with the hook every stored value
    is f to rounding -- a hook that ran before what it reads was filled, or
    after a prolongation read its region, would leave NaN."""
    forest = nested_forest(D)
    cen = cellcentered(D) if C == "cell" else vertexcentered(D)
    nv = 2 if form == "relative-all" else 1
    fs = FieldSet(forest, nv, G=2, centering=cen, device="cpu", fill=float("nan"))
    f = _linear(D)
    fill_by_coordinates(lambda x, v: f(x) * v, fs)
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    hook = {"relative": CellBoundary(Relative(f)),
            "values": CellBoundary(lambda x, v, d: f(x) * v),
            "relative-all": CellBoundary(Relative(AllVariables(lambda x: (f(x), 2 * f(x))))),
            "by-coordinates": boundary_by_coordinates(lambda x, v: f(x) * v)}[form]
    _oracle_hook(fs, s, hook)
    a = fs.array()
    assert not torch.isnan(a).any()
    x = [torch.tensor(fs.positions(d)) for d in range(D)]
    want = _linear(D)([x[d].reshape([fs.nblocks] + [1] * (2 - d) + [-1] + [1] * d)
                        for d in range(D)] if D == 3 else
                       [x[d].reshape([fs.nblocks, 1] + ([1, -1] if d == 0 else [-1, 1])) for d in range(D)])
    for v in range(nv):
        err = (a[:, v] - (v + 1) * want).abs().max().item()
        assert err < 1e-12, (v, err)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("C", ["cell", "vertex"])
def test_one_fill_is_the_relative_fixed_point(D, C):
    """test/relative_tests.jl:120-141: every value NaN but the own data; one
    fill leaves nothing undefined, a second changes nothing."""
    forest = nested_forest(D)
    fs = _fs(forest, cellcentered(D) if C == "cell" else vertexcentered(D), nan=True)
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    hook = CellBoundary(Relative(case_fn(D)))
    _oracle_hook(fs, s, hook)
    a = fs.array()
    assert not torch.isnan(a).any()
    once = fs.work.clone()
    _oracle_hook(fs, s, hook)
    assert torch.equal(fs.work.view(torch.int64), once.view(torch.int64))


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("C", ["cell", "vertex"])
def test_relative_zero_is_an_outflow_face(D, C):
    """Relative(f = 0) sets each hook ghost to its nearest evolved point's
    value, the zero-gradient outflow (src/ghosts.jl:1431): on a uniform
    forest every stored value of the hook fill over outer faces equals the
    plain fill over outflow faces, bit for bit."""
    cen = cellcentered(D) if C == "cell" else vertexcentered(D)
    outer = Forest((2,) * D, N=8)
    flow = Forest((2,) * D, N=8, outflow=((True, True),) * D)
    a, b = _fs(outer, cen), _fs(flow, cen)
    sa = GhostSchedule(a, Operators(prolongation=4, restriction=4))
    sb = GhostSchedule(b, Operators(prolongation=4, restriction=4))
    assert sa.desc.nboundaries > 0 and sb.desc.nboundaries == 0
    _oracle_hook(a, sa, CellBoundary(Relative(lambda x: torch.zeros_like(x[0]))))
    oracle.fill(sb._descptr, b.layout, b.work.numpy(), None, nthreads=4)
    assert torch.equal(a.work.view(torch.int64), b.work.view(torch.int64))


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("C", ["cell", "vertex"])
def test_hook_stage_groups_fill_alike_in_any_order(D, C):
    """test/ghost_tests.jl:399-453 with the relative hook: from NaN ghosts
    the fill, the fill with every stage group's stages reversed and the fill
    one stage at a time agree in every bit and leave no NaN; the grouping
    keeps a hook region's writes and reads apart from the stages grouped
    with its slot."""
    forest = nested_forest(D, roots=4)
    cen = cellcentered(D) if C == "cell" else vertexcentered(D)
    outs = []
    s = None
    for mode in ("grouped", "reversed", "alone"):
        fs = _fs(forest, cen, nan=True)
        s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
        desc = None
        keep = None
        if mode != "grouped":
            desc, keep = restaged(s, mode)
        _oracle_hook(fs, s, CellBoundary(Relative(case_fn(D))), desc=desc)
        assert not torch.isnan(fs.work).any(), mode
        outs.append(fs.work.clone())
        del keep
    assert any(R.slot > 1 for R in _regions(s))
    assert torch.equal(outs[0].view(torch.int64), outs[1].view(torch.int64))
    assert torch.equal(outs[0].view(torch.int64), outs[2].view(torch.int64))


def test_values_forms_agree():
    """The per-variable, AllVariables, WithPatch and boundary_by_coordinates
    forms of one hook give the same values (src/ghosts.jl:1355-1361: the
    forms write bit for bit the same numbers)."""
    forest = nested_forest(3)
    fs = _fs(forest, cellcentered(3), nvars=2)
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    g = lambda x, v, d: x[0] * v + d[0] * 0.5 - x[2]
    forms = [CellBoundary(g),
             CellBoundary(AllVariables(lambda x, d: (g(x, 1, d), g(x, 2, d)))),
             CellBoundary(WithPatch(lambda x, p, v, d: g(x, v, d) * p)),
             CellBoundary(WithPatch(AllVariables(lambda x, p, d: torch.stack([g(x, 1, d), g(x, 2, d)], 1))))]
    vals = [boundary_values(fs, s, h) for h in forms]
    for kind, v, fr in vals:
        assert kind == TAH_HOOK_VALUES and fr is None
        assert torch.equal(v, vals[0][1])
    kind, v, fr = boundary_values(fs, s, boundary_by_coordinates(lambda x, v: x[0] * v - x[2]))
    assert v.shape == (s.desc.boundary_points, 2)
    kind, fg, fr = boundary_values(fs, s, CellBoundary(Relative(AllVariables(lambda x: (x[0], x[1])))))
    assert kind == TAH_HOOK_RELATIVE and fg.shape == fr.shape == (s.desc.boundary_points, 2)


def test_hook_forms_refused():
    """test/relative_tests.jl:178-193, test/closed_tests.jl:1081 and the
    region form's refusal (src/ghosts.jl:1658-1671)."""
    forest = Forest((2, 2), N=8)
    two = FieldSet(forest, 2, G=1, device="cpu")
    s = GhostSchedule(two, Operators(prolongation=2, restriction=2))
    with pytest.raises(ValueError, match="wrap the function in AllVariables"):
        boundary_values(two, s, CellBoundary(Relative(lambda x: x[0] * 0)))
    with pytest.raises(ValueError, match="one value per variable"):
        boundary_values(two, s, CellBoundary(Relative(AllVariables(lambda x: (x[0] * 0,)))))
    with pytest.raises(ValueError, match="one value per variable"):
        boundary_values(two, s, CellBoundary(AllVariables(lambda x, d: (x[0],))))
    with pytest.raises(ValueError, match="wall must be"):
        CellBoundary(lambda x, v, d: 0.0, wall="symmetric")
    with pytest.raises(UnsupportedError):
        CellBoundary(lambda x, v, d: 0.0, wall="prescribed")
    with pytest.raises(ValueError, match="not a form the package has"):
        CellBoundary(Relative(WithPatch(lambda x, p: 0.0)))
    with pytest.raises(ValueError, match="host form"):
        boundary_values(two, s, lambda fs, b, key, d, region: None)


def test_hook_regions_numbered_and_refused_when_inconsistent():
    """The regions' points are numbered consecutively (first), their sum is
    boundary_points, and a plan refuses regions that do not add up or lie
    outside the stored extent (before any device work)."""
    forest = nested_forest(2)
    fs = _fs(forest, cellcentered(2))
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    first = 0
    for R in _regions(s):
        assert R.first == first
        first += int(np.prod([R.hi[d] - R.lo[d] for d in range(3)]))
        for d in range(3):
            assert -1 <= R.clamps[d] < fs.stored[d]
    assert first == s.desc.boundary_points
