"""Forest, Operators, FieldSet, GhostSchedule and fill_ghosts over libtah.

Names, argument meaning and refusals follow the reference's API for the
fill path: Forest (src/forest.jl:314-341), refine!/balance!
(src/forest.jl:1761,1837), Operators (src/operators.jl:314-415),
FieldSet (src/storage.jl:417-472), GhostSchedule (src/schedule.jl:3069-3292),
fill_ghosts! and check_fill (src/ghosts.jl:1783-1794, 2058-2115).
"""
import ctypes as C
from collections import namedtuple

import numpy as np

from . import _abi
from ._abi import check, lib

PointValue = 0
Conservative = 1
NoParity, EvenParity, OddParity = 0, 1, 2

MortonKey = namedtuple("MortonKey", ["root", "level", "coords"])


def cellcentered(D):
    return ("cell",) * D


def vertexcentered(D):
    return ("vertex",) * D


def _i32(seq):
    arr = (C.c_int32 * len(seq))(*[int(x) for x in seq])
    return arr


class Forest:
    """A brick of roots refined into a leaf-only octree (src/forest.jl:248-341)."""

    def __init__(self, roots, *, N, periodic=None, reflecting=None, outflow=None, extents=None):
        D = len(roots)
        self.D = D
        self.roots = tuple(int(r) for r in roots)
        self.N = int(N)
        self.periodic = tuple(bool(p) for p in (periodic or (False,) * D))
        self.reflecting = tuple(tuple(bool(x) for x in f) for f in (reflecting or ((False, False),) * D))
        self.outflow = tuple(tuple(bool(x) for x in f) for f in (outflow or ((False, False),) * D))
        self.extents = tuple((float(a), float(b)) for a, b in
                             (extents or tuple((0.0, float(r)) for r in self.roots)))
        refl = [int(x) for f in self.reflecting for x in f]
        outf = [int(x) for f in self.outflow for x in f]
        ext = (C.c_double * (2 * D))(*[x for e in self.extents for x in e])
        h = C.c_void_p()
        check(lib().tah_forest_create(D, _i32(self.roots), self.N, _i32(self.periodic), _i32(refl),
                                      _i32(outf), ext, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _abi is not None and _abi._lib is not None:
            _abi._lib.tah_forest_destroy(h)
            self._h = None

    @property
    def nleaves(self):
        n = C.c_int32()
        check(lib().tah_forest_info(self._h, C.byref(n), None))
        return n.value

    @property
    def generation(self):
        g = C.c_int64()
        check(lib().tah_forest_info(self._h, None, C.byref(g)))
        return g.value

    def leaf_arrays(self):
        """(root[n], level[n], coords[n,3]) of the sorted leaves."""
        n = self.nleaves
        root = np.zeros(n, np.int32)
        level = np.zeros(n, np.int32)
        coords = np.zeros((n, 3), np.uint32)
        check(lib().tah_forest_leaves(self._h, root.ctypes.data_as(_abi.I32P),
                                      level.ctypes.data_as(_abi.I32P),
                                      coords.ctypes.data_as(C.POINTER(C.c_uint32))))
        return root, level, coords

    @property
    def leaves(self):
        root, level, coords = self.leaf_arrays()
        return [MortonKey(int(r), int(l), tuple(int(c) for c in cc[:self.D]))
                for r, l, cc in zip(root, level, coords)]

    def refine(self, leaf_indices):
        """refine! of the leaves at the given (0-based) indices."""
        idx = np.ascontiguousarray(np.asarray(list(leaf_indices), dtype=np.int32))
        check(lib().tah_forest_refine(self._h, len(idx), idx.ctypes.data_as(_abi.I32P)))
        return self

    def coarsen(self, leaf_indices):
        """Merge every family whose children are all among the given (0-based)
        leaves into its parent, where 2:1 balance allows; returns the number
        of families merged (regrid!'s coarsening flags)."""
        idx = np.ascontiguousarray(np.asarray(list(leaf_indices), dtype=np.int32))
        m = C.c_int32()
        check(lib().tah_forest_coarsen(self._h, len(idx), idx.ctypes.data_as(_abi.I32P), C.byref(m)))
        return m.value

    def balance(self):
        check(lib().tah_forest_balance(self._h))
        return self

    def isbalanced(self):
        b = C.c_int32()
        check(lib().tah_forest_isbalanced(self._h, C.byref(b)))
        return bool(b.value)

    # --- geometry (src/geometry.jl) ---------------------------------------
    def root_spacing(self):
        """Block width at level 0 per dimension."""
        return np.array([(e[1] - e[0]) / r for e, r in zip(self.extents, self.roots)])

    def block_geometry(self):
        """Per leaf: the low corner of its block (n, D) and its cell size h (n,)."""
        root, level, coords = self.leaf_arrays()
        rp = np.zeros((len(root), self.D), np.int64)
        r = root.astype(np.int64).copy()
        for d in range(self.D):
            rp[:, d] = r % self.roots[d]
            r //= self.roots[d]
        H = self.root_spacing()
        scale = np.ldexp(1.0, -level.astype(np.int64))
        lows = np.zeros((len(root), self.D))
        for d in range(self.D):
            g = rp[:, d] * (1 << 0) * np.ldexp(1.0, level) + coords[:, d]
            lows[:, d] = self.extents[d][0] + g * H[d] * scale
        h = H[0] * scale / self.N
        return lows, h, level

    def positions(self, d, G, c, S):
        """tah_forest_positions: the positions along dimension d (0-based) of
        stored indices 0..S-1 of every leaf, for ghost width G and centering c
        (0 cell, 1 vertex), as nodeposition forms them, (nleaves, S)."""
        out = np.zeros((self.nleaves, int(S)), np.float64)
        check(lib().tah_forest_positions(self._h, int(d), int(G), int(c), int(S), out.ctypes.data_as(_abi.DP)))
        return out

    def block_centers(self):
        lows, h, level = self.block_geometry()
        return lows + 0.5 * (h * self.N)[:, None]


class Operators:
    """Operators(; prolongation, restriction, family, bounded, vertex_prolongation)."""

    def __init__(self, *, prolongation=None, restriction=None, family=PointValue, bounded=False,
                 vertex_prolongation=None):
        for name, p in (("prolongation", prolongation), ("restriction", restriction)):
            if p is None:
                raise ValueError(f"Operators has no default order: pass {name} explicitly")
        self.family = int(family)
        self.prolongation = int(prolongation)
        self.restriction = int(restriction)
        self.bounded = bool(bounded)
        self.vertex_prolongation = int(vertex_prolongation or 0)
        check(lib().tah_operators_check(C.byref(self._c())))

    def _c(self):
        return _abi.tah_operators(self.family, self.prolongation, self.restriction,
                                  int(self.bounded), self.vertex_prolongation)


def _centering_codes(centering, D):
    if isinstance(centering, str):
        centering = (centering,) * D
    codes = []
    for c in centering:
        if c in ("cell", 0):
            codes.append(0)
        elif c in ("vertex", 1):
            codes.append(1)
        else:
            raise ValueError(f"centering must be 'cell' or 'vertex', got {c!r}")
    if len(codes) != D:
        raise ValueError(f"centering needs one entry per dimension, got {centering}")
    return tuple(codes)


class Context:
    """One device (tah_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().tah_ctx_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _abi is not None and _abi._lib is not None:
            _abi._lib.tah_ctx_destroy(h)


_contexts = {}


def default_context(device=0):
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


class GhostSchedule:
    """The precomputed fill of one forest for one field-set layout."""

    def __init__(self, fs_or_forest, operators, *, G=None, centering=None, nvars=None, parity=None):
        if isinstance(fs_or_forest, FieldSet):
            fs = fs_or_forest
            forest, G, centering, nvars, parity = fs.forest, fs.G, fs.centering, fs.nvars, fs.parity
        else:
            forest = fs_or_forest
            if G is None:
                raise ValueError("GhostSchedule over a forest needs G")
            centering = centering or cellcentered(forest.D)
            nvars = nvars or 1
        D = forest.D
        self.forest = forest
        self.operators = operators
        self.G = tuple(int(g) for g in (G if hasattr(G, "__len__") else (G,) * D))
        self.centering = _centering_codes(centering, D)
        self.parity = None if parity is None else tuple(tuple(int(x) for x in p) for p in parity)
        self.nvars = int(nvars)
        par = None
        if self.parity is not None:
            flat = [x for p in self.parity for x in p]
            par = (C.c_int8 * len(flat))(*flat)
        h = C.c_void_p()
        check(lib().tah_schedule_create(forest._h, C.byref(operators._c()), _i32(self.G),
                                        _i32(self.centering), self.nvars, par, C.byref(h)))
        self._h = h
        self.generation = forest.generation
        dp = C.POINTER(_abi.tah_schedule_desc)()
        check(lib().tah_schedule_desc_get(h, C.byref(dp)))
        self.desc = dp.contents
        self._descptr = dp
        self._plans = {}

    def __del__(self):
        self._plans = {}
        h = getattr(self, "_h", None)
        if h and _abi is not None and _abi._lib is not None:
            _abi._lib.tah_schedule_destroy(h)

    def isstale(self):
        return self.forest.generation != self.generation

    @property
    def stored(self):
        return tuple(self.desc.stored[:3])

    def stage_tags(self):
        return [self.desc.stage_tags[i] for i in range(self.desc.nstages)]

    def stagegroups(self):
        return [(self.desc.stagegroup_first[i], self.desc.stagegroup_last[i])
                for i in range(self.desc.nstagegroups)]

    def groups(self):
        """Python view of the flattened groups (for tests and reporting)."""
        out = []
        for gi in range(self.desc.ngroups):
            g = self.desc.groups[gi]
            st = []
            for d in range(3):
                s = g.stencils[d]
                st.append(dict(targetfirst=s.targetfirst, ntarget=s.ntarget, order=s.order,
                               unit=bool(s.unit),
                               srcstart=np.ctypeslib.as_array(s.srcstart, (s.ntarget,)).copy()
                               if s.ntarget else np.zeros(0, np.int32),
                               weights=np.ctypeslib.as_array(s.weights, (s.ntarget * s.order,))
                               .reshape(s.ntarget, s.order).copy() if s.ntarget else
                               np.zeros((0, s.order))))
            n = g.ntransfers
            out.append(dict(kind=g.kind, stage=g.stage, factorcol=g.factorcol, ntransfers=n,
                            targetblocks=np.ctypeslib.as_array(g.targetblocks, (n,)).copy() if n else
                            np.zeros(0, np.int32),
                            sourceblocks=np.ctypeslib.as_array(g.sourceblocks, (n,)).copy() if n else
                            np.zeros(0, np.int32),
                            stencils=st))
        return out

    def boundaries(self):
        n = C.c_int32()
        check(lib().tah_schedule_boundaries(self._h, C.byref(n), None, None, None, None))
        k = n.value
        block = np.zeros(max(k, 1), np.int32)
        direction = np.zeros((max(k, 1), 3), np.int32)
        lo = np.zeros((max(k, 1), 3), np.int32)
        hi = np.zeros((max(k, 1), 3), np.int32)
        check(lib().tah_schedule_boundaries(self._h, C.byref(n), block.ctypes.data_as(_abi.I32P),
                                            direction.ctypes.data_as(_abi.I32P),
                                            lo.ctypes.data_as(_abi.I32P), hi.ctypes.data_as(_abi.I32P)))
        return [(int(block[i]), tuple(direction[i]), tuple(lo[i]), tuple(hi[i])) for i in range(k)]

    def points(self):
        """Spatial points one fill writes (times nvars: values)."""
        p = C.c_int64()
        check(lib().tah_schedule_points(self._h, C.byref(p)))
        return p.value

    def plan(self, layout, ctx=None):
        """The device plan of this schedule for a layout (uploaded once)."""
        ctx = ctx or default_context()
        key = (tuple(layout.stride), layout.nvars, ctx.device)
        if key not in self._plans:
            self._plans[key] = Plan(ctx, self, layout)
        return self._plans[key]


class Plan:
    """A schedule's tables in HBM for one layout (tah_plan)."""

    def __init__(self, ctx, schedule, layout):
        self.ctx = ctx
        self.schedule = schedule
        self.layout = layout
        h = C.c_void_p()
        check(lib().tah_plan_create(ctx._h, schedule._descptr, C.byref(layout), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _abi is not None and _abi._lib is not None:
            _abi._lib.tah_plan_destroy(h)

    def values(self):
        v = C.c_int64()
        check(lib().tah_plan_values(self._h, C.byref(v)))
        return v.value

    @property
    def options(self):
        """The launch options: dict(fork=, merge=, defer=, fuse=, interp_pairs=)
        each 0 or 1 (fuse always 0), lds 0 to 7 (bit 2, the default:
        restriction streamed along z; bits 0 and 1: restriction and
        prolongation from LDS chunks), copy_per_thread 4, 8 or 16, copy_wgs 0
        or a multiple of 8, queue 0 or 1, copy_order 0 (copy tiles by source
        block) or 1 (by target block)."""
        o = _abi.tah_plan_options()
        check(lib().tah_plan_get_options(self._h, C.byref(o)))
        return dict(fork=o.fork, merge=o.merge, defer=o.defer, copy_per_thread=o.copy_per_thread,
                    fuse=o.fuse, copy_wgs=o.copy_wgs, interp_pairs=o.interp_pairs, lds=o.lds,
                    queue=o.queue, copy_order=o.copy_order)

    def set_options(self, fork=None, merge=None, defer=None, copy_per_thread=None, fuse=None,
                    copy_wgs=None, interp_pairs=None, lds=None, queue=None, copy_order=None):
        cur = self.options
        pick = lambda name, v: cur[name] if v is None else int(v)
        o = _abi.tah_plan_options(pick("fork", fork), pick("merge", merge), pick("defer", defer),
                                  pick("copy_per_thread", copy_per_thread), pick("fuse", fuse),
                                  pick("copy_wgs", copy_wgs), pick("interp_pairs", interp_pairs),
                                  pick("lds", lds), pick("queue", queue), pick("copy_order", copy_order))
        check(lib().tah_plan_set_options(self._h, C.byref(o)))

    @property
    def families(self):
        """The interpolation families of the plan and their options: a list of
        dict(order=, restriction=, threads=, merge=, pairs=, lds=, vars=)."""
        n = C.c_int32()
        check(lib().tah_plan_get_family_options(self._h, C.byref(n), None))
        arr = (_abi.tah_family_options * max(1, n.value))()
        check(lib().tah_plan_get_family_options(self._h, C.byref(n), arr))
        return [dict(order=e.order, restriction=e.restriction, threads=e.threads, merge=e.merge, pairs=e.pairs,
                     lds=e.lds, vars=e.vars) for e in arr[:n.value]]

    def set_family(self, order, restriction, **kw):
        """Set one family's options (threads 64, 128 or 256; merge, pairs 0 or
        1; lds 0 or 1, or 2 for restriction streamed along z; vars 0 or a cap
        on a tile's variables); the others keep theirs."""
        cur = next((f for f in self.families if f["order"] == order and f["restriction"] == int(restriction)),
                   dict(order=order, restriction=int(restriction), threads=256, merge=1, pairs=0,
                        lds=2 if restriction else 0, vars=0))
        cur.update({k: int(v) for k, v in kw.items()})
        o = _abi.tah_family_options(int(order), int(restriction), cur["threads"], cur["merge"], cur["pairs"],
                                    cur["lds"], cur["vars"])
        check(lib().tah_plan_set_family_options(self._h, C.byref(o)))


def _layout_of(desc_ptr, nvars, kind):
    L = _abi.tah_layout()
    fn = lib().tah_layout_dense if kind == "dense" else lib().tah_layout_arena
    check(fn(desc_ptr, int(nvars), C.byref(L)))
    return L


class FieldSet:
    """Field data over a forest: work[i, j, k, var, block] with explicit strides.

    layout="dense" is the reference's array (src/storage.jl:196-200);
    layout="arena" pads rows to 32-byte sectors (one HBM arena, Morton order).
    """

    def __init__(self, forest, nvars, *, G, centering=None, parity=None, device="cuda",
                 layout="dense", fill=0.0):
        import torch
        D = forest.D
        self.forest = forest
        self.nvars = int(nvars)
        self.G = tuple(int(g) for g in (G if hasattr(G, "__len__") else (G,) * D))
        self.centering = _centering_codes(centering or cellcentered(D), D)
        for d in range(D):
            if forest.N < 2 * self.G[d]:
                raise ValueError(f"N must be >= 2G in every dimension, got N={forest.N}, G={self.G}")
        hasrefl = any(any(f) for f in forest.reflecting)
        if parity is not None:
            parity = tuple(tuple(int(x) for x in p) for p in parity)
            if len(parity) != self.nvars:
                raise ValueError("parity needs one entry per variable")
            for v, p in enumerate(parity):
                for d in range(D):
                    if any(forest.reflecting[d]) and p[d] == NoParity:
                        raise ValueError(f"variable {v + 1} has NoParity in dimension {d + 1}, which has "
                                         "a reflecting face. Give it EvenParity or OddParity.")
        elif hasrefl:
            raise ValueError("a field set over a forest with reflecting faces needs a parity per variable")
        self.parity = parity
        self.stored = tuple(forest.N + 2 * self.G[d] + self.centering[d] for d in range(D)) + (1,) * (3 - D)
        self.device = device
        self.layout_kind = layout
        self._relayout()
        self.work = torch.full((self.nelements,), float(fill), dtype=torch.float64, device=device)
        self.factors = None
        if parity is not None:
            ncols = 3 ** D
            f = np.zeros(self.nvars * ncols)
            flat = [x for p in parity for x in p]
            check(lib().tah_parity_factors(D, self.nvars, (C.c_int8 * len(flat))(*flat),
                                           f.ctypes.data_as(_abi.DP)))
            self.factors = torch.tensor(f, dtype=torch.float64, device=device)
        self.generation = forest.generation

    def _relayout(self):
        """The layout of the forest's current leaves (the stored size and
        the block count are all it needs)."""
        desc = _abi.tah_schedule_desc()
        desc.dim = self.forest.D
        desc.nblocks = self.forest.nleaves
        for d in range(3):
            desc.stored[d] = self.stored[d]
        self.layout = _layout_of(C.byref(desc), self.nvars, self.layout_kind)
        n = C.c_int64()
        check(lib().tah_layout_size(C.byref(self.layout), C.byref(n)))
        self.nelements = n.value
        self.generation = self.forest.generation

    @property
    def nblocks(self):
        return self.layout.nblocks

    def array(self):
        """A strided view indexed [block, var, k, j, i]."""
        s = self.layout.stride
        S = self.stored
        return self.work.as_strided((self.nblocks, self.nvars, S[2], S[1], S[0]),
                                    (s[4], s[3], s[2], s[1], s[0]))

    def own_slices(self):
        """Slices of the own range along (k, j, i)."""
        D = self.forest.D
        sl = []
        for d in (2, 1, 0):
            if d < D:
                sl.append(slice(self.G[d], self.G[d] + self.forest.N + self.centering[d]))
            else:
                sl.append(slice(0, 1))
        return tuple(sl)

    def positions(self, d):
        """Positions along dimension d of every stored index of every block,
        (nblocks, S[d]), as nodeposition forms them (one rounding;
        src/storage.jl:1008-1041)."""
        return self.forest.positions(d, self.G[d], self.centering[d], self.stored[d])

    def coordinates(self):
        """Positions of every stored point: three arrays (nblocks, S2, S1, S0)."""
        D = self.forest.D
        out = []
        S = self.stored
        for d in range(3):
            if d >= D:
                out.append(np.zeros((self.nblocks, S[2], S[1], S[0])))
                continue
            x = self.positions(d)
            shape = [self.nblocks, 1, 1, 1]
            shape[3 - d] = S[d]
            out.append(np.broadcast_to(x.reshape(shape), (self.nblocks, S[2], S[1], S[0])))
        return out


def _check_fill(fs, schedule):
    """check_fill (src/ghosts.jl:2058-2115), raised before any launch."""
    if fs.forest is not schedule.forest:
        raise ValueError("the field set and the schedule were built over different forests")
    if schedule.isstale():
        raise _abi.StaleScheduleError(_abi.TAH_ESTALE, "the schedule is stale: the forest has changed "
                                      "since it was built (refine, coarsen or regrid); rebuild it")
    check(lib().tah_schedule_check(schedule._h))
    if fs.nblocks != schedule.forest.nleaves:
        raise ValueError(f"the field set has {fs.nblocks} blocks but the schedule's forest has "
                         f"{schedule.forest.nleaves}; rebuild both")
    if str(fs.work.dtype) != "torch.float64":
        raise ValueError(f"the field set stores {fs.work.dtype} but the fill computes in Float64 "
                         "(the schedule's weights); store the field set as torch.float64")
    if fs.G != schedule.G:
        raise ValueError(f"the field set has G={fs.G}, the schedule was built for G={schedule.G}")
    if fs.centering != schedule.centering:
        raise ValueError("the field set and the schedule differ in centering")
    if schedule.parity is not None and fs.parity != schedule.parity:
        raise ValueError("the schedule was built for other parities than the field set's")


def _stream_ptr(fs):
    import torch
    return C.c_void_p(torch.cuda.current_stream(fs.work.device).cuda_stream)


# --- callbacks of position (src/storage.jl, src/ghosts.jl) ------------------
#
# The reference hands its callbacks one point at a time inside a kernel. Here
# a callback is called once per use over every point at once: a position x
# is a tuple of D float64 tensors (one entry per point, on the field set's
# device), a direction a tuple of D integer tensors, a patch an integer
# tensor (1 on every block of a brick forest), and a variable index v is
# 1-based, as in the reference. A callback returns a tensor over the points
# (or a number), or, wrapped in AllVariables, nvars of them (a sequence, or
# a tensor whose last axis is the variable). Written with torch operations
# it runs on the GPU.

class AllVariables:
    """AllVariables(f): the callback returns every variable's value at a point
    at once, instead of taking the variable index."""

    def __init__(self, f):
        self.f = f


class WithPatch:
    """WithPatch(f): the callback is handed the point's patch after the
    position."""

    def __init__(self, f):
        self.f = f


class Relative:
    """Relative(f) (src/ghosts.jl:1419-1465): a ghost beyond an outer face
    becomes f(x_g) + (u(x_r) - f(x_r)), x_r its nearest evolved point; f(x)
    gives one variable's value (a field set of one variable) or, wrapped in
    AllVariables, all of them."""

    def __init__(self, f):
        self.f = f


class CellBoundary:
    """CellBoundary(g; wall) (src/ghosts.jl:1321-1413): the outer-face hook
    per ghost point, g(x, v, delta) -> value. Forms: CellBoundary(g),
    CellBoundary(AllVariables(g)) with g(x, delta) -> values,
    CellBoundary(WithPatch(g)) with g(x, p, v, delta),
    CellBoundary(WithPatch(AllVariables(g))) with g(x, p, delta),
    CellBoundary(Relative(f)) and CellBoundary(Relative(AllVariables(f))).
    wall="evolved" (the default) leaves a vertex-like dimension's wall plane
    to the application; wall="prescribed" (stage W) is not provided."""

    def __init__(self, g, *, wall="evolved"):
        if wall not in ("evolved", "prescribed"):
            raise ValueError("wall must be 'evolved' (the wall plane of a vertex-like dimension is an own "
                             "point of the state, evolved by the application, and the hook fills the ghosts "
                             "beyond it) or 'prescribed' (the hook gives the wall plane its value too, a "
                             f"Dirichlet condition), got {wall!r}")
        if wall == "prescribed":
            raise _abi.UnsupportedError(_abi.TAH_EUNSUPPORTED, "CellBoundary(g, wall='prescribed') (stage "
                                        "W) is not provided: the wall plane stays the application's")
        if isinstance(g, Relative) and isinstance(g.f, WithPatch):
            raise ValueError("CellBoundary(Relative(WithPatch(f))) is not a form the package has: the relative "
                             "hook calls f at the ghost and at the nearest evolved point, by position alone, "
                             "and is not handed the patch")
        self.g = g
        self.wall = wall


def boundary_by_coordinates(f, *, wall="evolved"):
    """boundary_by_coordinates(f) (src/ghosts.jl:2189-2219): CellBoundary of
    a function of position and variable, f(x, v), or of its AllVariables and
    WithPatch forms, the direction dropped."""
    if isinstance(f, WithPatch):
        if isinstance(f.f, AllVariables):
            h = f.f.f
            return CellBoundary(WithPatch(AllVariables(lambda x, p, d: h(x, p))), wall=wall)
        h = f.f
        return CellBoundary(WithPatch(lambda x, p, v, d: h(x, p, v)), wall=wall)
    if isinstance(f, AllVariables):
        h = f.f
        return CellBoundary(AllVariables(lambda x, d: h(x)), wall=wall)
    return CellBoundary(lambda x, v, d: f(x, v), wall=wall)


def _values_of(out, npts, nvars, what, device):
    """A callback's result as a contiguous (npts, nvars) float64 tensor."""
    import torch
    if isinstance(out, torch.Tensor) and out.dim() == 2:
        cols = [out[:, v] for v in range(out.shape[1])]
    elif isinstance(out, (tuple, list)):
        cols = list(out)
    else:
        cols = [out]
    if len(cols) != nvars:
        raise ValueError(f"the {what} must return one value per variable: the field set has nvars = {nvars}, "
                         f"it returned {len(cols)}")
    res = torch.empty((npts, nvars), dtype=torch.float64, device=device)
    for v, c in enumerate(cols):
        res[:, v] = torch.as_tensor(c, dtype=torch.float64, device=device).expand(npts)
    return res


class _HookPoints:
    """Every hook ghost point of a schedule's regions, in their numbering
    (tah_boundary_region.first; i fastest within a region): block, stored
    index, nearest evolved point, outward direction, and the positions of
    the ghost and of that point as nodeposition forms them."""

    def __init__(self, fs, schedule):
        import torch
        d = schedule.desc
        D = fs.forest.D
        P = int(d.boundary_points)
        block = np.zeros(P, np.int64)
        idx = np.zeros((P, 3), np.int64)
        ridx = np.zeros((P, 3), np.int64)
        dirs = np.zeros((P, 3), np.int64)
        for r in range(d.nboundaries):
            R = d.boundaries[r]
            lo, hi = list(R.lo), list(R.hi)
            k, j, i = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]), np.arange(lo[0], hi[0]),
                                  indexing="ij")
            a, m = int(R.first), i.size
            for dd, arr in enumerate((i, j, k)):
                idx[a:a + m, dd] = arr.ravel()
                ridx[a:a + m, dd] = R.clamps[dd] if R.clamps[dd] >= 0 else arr.ravel()
                dirs[a:a + m, dd] = R.direction[dd]
            block[a:a + m] = R.block
        dev = fs.work.device
        pos = [fs.positions(dd) for dd in range(D)]
        self.npoints = P
        self.block, self.idx, self.ridx = block, idx, ridx
        self.xg = tuple(torch.tensor(pos[dd][block, idx[:, dd]], dtype=torch.float64, device=dev)
                        for dd in range(D))
        self.xr = tuple(torch.tensor(pos[dd][block, ridx[:, dd]], dtype=torch.float64, device=dev)
                        for dd in range(D))
        self.dirs = tuple(torch.tensor(dirs[:, dd], dtype=torch.int64, device=dev) for dd in range(D))
        self.patch = torch.ones(P, dtype=torch.int64, device=dev)


def _hook_points(fs, schedule):
    cache = schedule.__dict__.setdefault("_hookpoints", {})
    key = (str(fs.work.device), fs.G, fs.centering)
    if key not in cache:
        cache[key] = _HookPoints(fs, schedule)
    return cache[key]


def boundary_values(fs, schedule, hook):
    """The hook's values at the schedule's hook points: (kind, values, fr),
    kind _abi.TAH_HOOK_VALUES (values: g at every ghost, fr None) or
    _abi.TAH_HOOK_RELATIVE (values = f(x_g), fr = f(x_r)), each a contiguous
    (points, nvars) float64 tensor on the field set's device."""
    if not isinstance(hook, CellBoundary):
        raise ValueError("this boundary hook takes a whole region, which is a host form: it indexes the "
                         "working array cell by cell. Wrap a per-cell function in CellBoundary(g) "
                         "(g(x, v, delta) -> value), which the package evaluates at every hook point; a "
                         "condition relative to the nearest evolved point is CellBoundary(Relative(f)), and "
                         "a zero-gradient one an outflow face")
    pts = _hook_points(fs, schedule)
    P, nv, dev = pts.npoints, fs.nvars, fs.work.device
    g = hook.g
    if isinstance(g, Relative):
        f = g.f
        if isinstance(f, AllVariables):
            fg = _values_of(f.f(pts.xg), P, nv, "boundary hook", dev)
            fr = _values_of(f.f(pts.xr), P, nv, "boundary hook", dev)
        else:
            if nv != 1:
                raise ValueError(f"CellBoundary(Relative(f)) sets one variable from f(x), but the field set has "
                                 f"nvars = {nv}; for several, wrap the function in AllVariables, "
                                 "CellBoundary(Relative(AllVariables(lambda x: (u1, u2, ...))))")
            fg = _values_of(f(pts.xg), P, 1, "boundary hook", dev)
            fr = _values_of(f(pts.xr), P, 1, "boundary hook", dev)
        return _abi.TAH_HOOK_RELATIVE, fg, fr
    patch = isinstance(g, WithPatch)
    inner = g.f if patch else g
    if isinstance(inner, AllVariables):
        out = inner.f(pts.xg, pts.patch, pts.dirs) if patch else inner.f(pts.xg, pts.dirs)
        return _abi.TAH_HOOK_VALUES, _values_of(out, P, nv, "boundary hook", dev), None
    cols = [inner(pts.xg, pts.patch, v + 1, pts.dirs) if patch else inner(pts.xg, v + 1, pts.dirs)
            for v in range(nv)]
    return _abi.TAH_HOOK_VALUES, _values_of(cols, P, nv, "boundary hook", dev), None


def fill_by_coordinates(f, fs):
    """fill_by_coordinates!(f, fs) (src/storage.jl:1359-1418): every own point
    of every variable from f(x, v), x the positions over (blocks, k, j, i) of
    the own range as nodeposition forms them; or AllVariables(f), f(x) ->
    nvars values; or WithPatch(f), f(x, p, v) / f(x, p)."""
    import torch
    dev = fs.work.device
    own = fs.own_slices()
    D = fs.forest.D
    x = []
    for d in range(D):
        p = torch.tensor(fs.positions(d)[:, own[2 - d]], dtype=torch.float64, device=dev)
        shape = [fs.nblocks, 1, 1, 1]
        shape[3 - d] = p.shape[1]
        x.append(p.reshape(shape))
    ext = [own[2 - d].stop - own[2 - d].start if d < D else 1 for d in range(3)]
    full = (fs.nblocks, ext[2], ext[1], ext[0])
    x = tuple(t.expand(full) for t in x)
    patch = torch.ones(full, dtype=torch.int64, device=dev)
    withpatch = isinstance(f, WithPatch)
    inner = f.f if withpatch else f
    a = fs.array()
    if isinstance(inner, AllVariables):
        out = inner.f(x, patch) if withpatch else inner.f(x)
        cols = [out[..., v] for v in range(out.shape[-1])] if isinstance(out, torch.Tensor) and \
            out.dim() == 5 else list(out) if isinstance(out, (tuple, list)) else [out]
        if len(cols) != fs.nvars:
            raise ValueError(f"the fill callback must return one value per variable: the field set has nvars = "
                             f"{fs.nvars}, it returned {len(cols)}")
    else:
        cols = [inner(x, patch, v + 1) if withpatch else inner(x, v + 1) for v in range(fs.nvars)]
    for v, c in enumerate(cols):
        a[(slice(None), v) + own] = torch.as_tensor(c, dtype=torch.float64, device=dev).expand(full)
    return fs


def fill_ghosts(fs, schedule, *, boundary=None, method="staged", ctx=None):
    """fill_ghosts!(fs, schedule; boundary): every ghost of fs.work from the
    schedule, on the GPU.

    method="staged" (one launch per stage group, the default) or "plain" (one
    launch per transfer group, the reference's device form). boundary: a
    CellBoundary hook, evaluated at the schedule's hook points and run
    between the stage groups as the reference runs it (tah_fill_boundary).
    """
    _check_fill(fs, schedule)
    if fs.work.device.type != "cuda":
        raise RuntimeError("fill_ghosts runs on the GPU only: this field set lives on "
                           f"{fs.work.device}; there is no CPU fallback")
    plan = schedule.plan(fs.layout, ctx)
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    if boundary is not None:
        if method != "staged":
            raise ValueError("the boundary hook runs with the staged fill (method='staged')")
        kind, values, fr = boundary_values(fs, schedule, boundary)
        check(lib().tah_fill_boundary(plan._h, C.c_void_p(fs.work.data_ptr()), fptr, kind,
                                      C.c_void_p(values.data_ptr()),
                                      C.c_void_p(fr.data_ptr()) if fr is not None else None, _stream_ptr(fs)))
        return fs
    fn = lib().tah_fill if method == "staged" else lib().tah_fill_plain
    check(fn(plan._h, C.c_void_p(fs.work.data_ptr()), fptr, _stream_ptr(fs)))
    return fs


def tune(fs, schedule, cache_path=None, *, ctx=None):
    """tah_tune: time every combination of the plan-wide launch options on
    fs.work and keep the fastest, then each interpolation family's options
    (cached in cache_path by mesh signature). Returns (options with the
    families' under "families", median ms of the choice, or None when taken
    from the cache)."""
    _check_fill(fs, schedule)
    if fs.work.device.type != "cuda":
        raise RuntimeError("tune runs on the GPU only")
    plan = schedule.plan(fs.layout, ctx)
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    ms = C.c_double()
    path = None if cache_path is None else str(cache_path).encode()
    check(lib().tah_tune(plan._h, C.c_void_p(fs.work.data_ptr()), fptr, path, C.byref(ms)))
    opts = dict(plan.options)
    opts["families"] = plan.families
    return opts, (None if ms.value < 0 else ms.value)


def kernel_class_name(cls):
    buf = C.create_string_buffer(96)
    check(lib().tah_kernel_class_name(int(cls), buf, 96))
    return buf.value.decode()


def fill_ghosts_timed(fs, schedule, *, ctx=None):
    """The staged fill with its launches run in order, each timed by HIP
    events on fs's stream: [(kernel name, values written, ms)] in fill order."""
    _check_fill(fs, schedule)
    if fs.work.device.type != "cuda":
        raise RuntimeError("fill_ghosts_timed runs on the GPU only")
    plan = schedule.plan(fs.layout, ctx)
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    n = C.c_int32()
    work = C.c_void_p(fs.work.data_ptr())
    check(lib().tah_fill_timed(plan._h, work, fptr, _stream_ptr(fs), C.byref(n), None, None, None))
    k = n.value
    cls = (C.c_int32 * max(k, 1))()
    vals = (C.c_int64 * max(k, 1))()
    ms = (C.c_float * max(k, 1))()
    check(lib().tah_fill_timed(plan._h, work, fptr, _stream_ptr(fs), C.byref(n), cls, vals, ms))
    return [(kernel_class_name(cls[i]), int(vals[i]), float(ms[i])) for i in range(n.value)]


def fill_ghosts_adjoint(fs, schedule, adj, *, ctx=None):
    """The transpose of the fill applied in place to adj (a tensor shaped like fs.work)."""
    _check_fill(fs, schedule)
    plan = schedule.plan(fs.layout, ctx)
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    check(lib().tah_fill_adjoint(plan._h, C.c_void_p(adj.data_ptr()), fptr, _stream_ptr(fs)))
    return adj


ADJOINT_KINDS = ("adjoint_pass_kernel<2>", "adjoint_pass_kernel<1>", "adjoint_pass_kernel<0>",
                 "adjoint_gather_kernel", "zero_kernel")


def fill_ghosts_adjoint_timed(fs, schedule, adj, *, ctx=None):
    """The adjoint with its launches timed one by one (tah_fill_adjoint_timed):
    a list of (kernel name, algorithmic bytes, ms) in launch order."""
    _check_fill(fs, schedule)
    plan = schedule.plan(fs.layout, ctx)
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    n = C.c_int32()
    a = C.c_void_p(adj.data_ptr())
    check(lib().tah_fill_adjoint_timed(plan._h, a, fptr, _stream_ptr(fs), C.byref(n), None, None, None))
    kind = (C.c_int32 * max(1, n.value))()
    nbytes = (C.c_int64 * max(1, n.value))()
    ms = (C.c_float * max(1, n.value))()
    check(lib().tah_fill_adjoint_timed(plan._h, a, fptr, _stream_ptr(fs), C.byref(n), kind, nbytes, ms))
    return [(ADJOINT_KINDS[kind[i]], int(nbytes[i]), float(ms[i])) for i in range(n.value)]


# ---------------------------------------------------------------------------
# The state vector (src/state.jl): every own point of every block, no ghosts,
# laid out (x, y, z, variable, block) with x fastest.

def ownextents(fs):
    """Own points per dimension (src/storage.jl:939): N + c_d, c_d = 1 along a
    vertex-like dimension; 1 beyond the forest's dimension."""
    D = fs.forest.D
    return tuple(fs.forest.N + fs.centering[d] for d in range(D)) + (1,) * (3 - D)


def _own_first(fs):
    D = fs.forest.D
    return tuple(fs.G[d] for d in range(D)) + (0,) * (3 - D)


def _check_sets(sets):
    """check_sets (src/state.jl:283-310)."""
    if len(sets) == 0:
        raise ValueError("a state vector over several field sets needs at least one set; for one set, "
                         "the single-set functions take the field set itself")
    fs = sets[0]
    for i, s in enumerate(sets, 1):
        if s.forest is not fs.forest:
            raise ValueError(f"field set {i} is over a different forest from field set 1: the sets of one "
                             "state vector are regridded together, so they must share the forest")
        if s.work.device != fs.work.device:
            raise ValueError(f"field set {i} is on {s.work.device} and field set 1 on {fs.work.device}: the "
                             "state vector is one array, on one device")
        if s.work.dtype != fs.work.dtype:
            raise ValueError(f"field set {i} has element type {s.work.dtype} and field set 1 has "
                             f"{fs.work.dtype}: the state vector is one array, of one element type")
        if s.nblocks != s.forest.nleaves:
            raise ValueError(f"field set {i} has {s.nblocks} blocks but the forest has {s.forest.nleaves}: "
                             "the set was not resized with the forest's last regrid")


def statelength(fs):
    """statelength (src/state.jl:25, 171-174): prod(ownextents) * nvars * nblocks,
    summed over the sets of a tuple."""
    if isinstance(fs, tuple):
        _check_sets(fs)
        return sum(statelength(s) for s in fs)
    e = ownextents(fs)
    return e[0] * e[1] * e[2] * fs.nvars * fs.nblocks


def _check_statelength(u, fs):
    n = statelength(fs)
    if u.dim() != 1 or u.numel() != n:
        what = "this field set needs" if not isinstance(fs, tuple) else f"these {len(fs)} field sets need"
        raise ValueError(f"state vector has {u.numel()} entries but {what} {n}")


def statevector(fs):
    """A zeroed state vector on the field set's device (src/state.jl:37-42, 214-224)."""
    import torch
    first = fs[0] if isinstance(fs, tuple) else fs
    return torch.zeros(statelength(fs), dtype=first.work.dtype, device=first.work.device)


def staterange(sets, i):
    """The entries of the i-th set's segment (0-based here): src/state.jl:181-189."""
    _check_sets(sets)
    if not 0 <= i < len(sets):
        raise ValueError(f"set index {i} is outside 0:{len(sets) - 1}: the state vector has one segment "
                         "per field set in the tuple")
    start = sum(statelength(sets[j]) for j in range(i))
    return range(start, start + statelength(sets[i]))


def stateview(u, sets, i):
    """The i-th set's segment of u, sharing memory (src/state.jl:233-236)."""
    _check_statelength(u, sets)
    r = staterange(sets, i)
    return u[r.start:r.stop]


def statearray(u, fs, i=None):
    """u viewed as (nblocks, nvars, e2, e1, e0) -- the reference's (e0, e1, e2,
    nvars, nblocks) column-major array (src/state.jl:54-58, 244-248)."""
    if isinstance(fs, tuple):
        return statearray(stateview(u, fs, i), fs[i])
    _check_statelength(u, fs)
    e = ownextents(fs)
    return u.view(fs.nblocks, fs.nvars, e[2], e[1], e[0])


def _state_copy(fn, fs, u):
    _check_statelength(u, fs)
    if str(u.dtype) != "torch.float64" or str(fs.work.dtype) != "torch.float64":
        raise ValueError("the state vector and the field set hold Float64")
    if u.device != fs.work.device or fs.work.device.type != "cuda":
        raise ValueError("the state vector and the field set must be on the same GPU")
    if not u.is_contiguous():
        raise ValueError("the state vector must be contiguous")
    check(fn(C.byref(fs.layout), _i32(_own_first(fs)), _i32(ownextents(fs)), C.c_void_p(fs.work.data_ptr()),
             C.c_void_p(u.data_ptr()), _stream_ptr(fs)))


def scatter(fs, u):
    """scatter! (src/state.jl:135-139, 257-263): u -> the working array's own
    points, ghosts untouched; on the field set's current stream."""
    if isinstance(fs, tuple):
        _check_statelength(u, fs)
        for i, s in enumerate(fs):
            scatter(s, stateview(u, fs, i))
        return fs
    _state_copy(lib().tah_scatter, fs, u)
    return fs


def gather(u, fs):
    """gather! (src/state.jl:150-154, 270-276): the working array's own points
    -> u; on the field set's current stream."""
    if isinstance(fs, tuple):
        _check_statelength(u, fs)
        for i, s in enumerate(fs):
            gather(stateview(u, fs, i), s)
        return u
    _state_copy(lambda L, f, e, w, st, sp: lib().tah_gather(L, f, e, st, w, sp), fs, u)
    return u


# ---------------------------------------------------------------------------
# unify! (src/ghosts.jl:1836-1906, src/unify.jl): stages P and 0 alone.

class UnifySchedule:
    """UnifySchedule(fs) (src/unify.jl:76-170): stage P and stage 0 of a set
    with a vertex-like dimension, alone, with the restriction the geometry
    fixes (injection along a vertex-like dimension, the exact two-cell
    average along a cell-like one) and any ghost width."""

    def __init__(self, fs):
        self.forest = fs.forest
        self.G, self.centering, self.nvars, self.parity = fs.G, fs.centering, fs.nvars, fs.parity
        par = None
        if self.parity is not None:
            flat = [x for p in self.parity for x in p]
            par = (C.c_int8 * len(flat))(*flat)
        h = C.c_void_p()
        check(lib().tah_unify_schedule_create(fs.forest._h, _i32(self.G), _i32(self.centering), self.nvars, par,
                                              C.byref(h)))
        self._h = h
        self.generation = fs.forest.generation
        dp = C.POINTER(_abi.tah_schedule_desc)()
        check(lib().tah_schedule_desc_get(h, C.byref(dp)))
        self.desc = dp.contents
        self._descptr = dp
        self._plans = {}

    def __del__(self):
        self._plans = {}
        h = getattr(self, "_h", None)
        if h and _abi is not None and _abi._lib is not None:
            _abi._lib.tah_schedule_destroy(h)

    def isstale(self):
        return self.forest.generation != self.generation

    def stage_tags(self):
        return [self.desc.stage_tags[i] for i in range(self.desc.nstages)]

    def groups(self):
        return GhostSchedule.groups(self)


class _Desc:
    """A derived descriptor (stages P and 0 of a schedule), kept with it."""

    def __init__(self, owner, dp):
        self.owner = owner
        self._descptr = dp
        self.desc = dp.contents

    def groups(self):
        return GhostSchedule.groups(self)


def unify_desc(schedule, state=False):
    """Stages P and 0 of a GhostSchedule or UnifySchedule as a descriptor of
    their own: in the working array's indices, or with state=True a state
    vector's (tah_unify_desc_get)."""
    dp = C.POINTER(_abi.tah_schedule_desc)()
    check(lib().tah_unify_desc_get(schedule._h, 1 if state else 0, C.byref(dp)))
    return _Desc(schedule, dp)


def _check_unify(fs, schedule):
    """check_unify (src/unify.jl): the schedule is current and built for this
    set's forest and layout."""
    if isinstance(schedule, GhostSchedule):
        return _check_fill(fs, schedule)
    if fs.forest is not schedule.forest:
        raise ValueError("the field set and the unify schedule were built over different forests")
    if schedule.isstale():
        raise _abi.StaleScheduleError(_abi.TAH_ESTALE, "the unify schedule is stale: the forest has changed "
                                      "since it was built; rebuild it")
    if fs.nblocks != schedule.forest.nleaves:
        raise ValueError(f"the field set has {fs.nblocks} blocks but the forest has {schedule.forest.nleaves}")
    if fs.G != schedule.G or fs.centering != schedule.centering or fs.nvars != schedule.nvars:
        raise ValueError("the unify schedule was built for another layout (G, centering or nvars)")
    if schedule.parity is not None and fs.parity != schedule.parity:
        raise ValueError("the unify schedule was built for other parities than the field set's")


def unify(*args, ctx=None):
    """unify!(fs, schedule) and unify!(v, fs, schedule) (src/ghosts.jl:1836-1860):
    stages P and 0 of the schedule (a GhostSchedule or a UnifySchedule) on
    the field set's working array, or on the state vector v, whose every
    secondary copy of a shared point then holds its primary's value. On the
    field set's current stream."""
    import torch
    if len(args) == 2:
        (fs, schedule), v = args, None
    elif len(args) == 3:
        v, fs, schedule = args
    else:
        raise TypeError("unify(fs, schedule) or unify(v, fs, schedule)")
    _check_unify(fs, schedule)
    ctx = ctx or default_context()
    state = v is not None
    if state:
        _check_statelength(v, fs)
        if str(v.dtype) != "torch.float64" or v.device != fs.work.device or not v.is_contiguous():
            raise ValueError("the state vector must be a contiguous Float64 tensor on the field set's GPU")
    dsc = unify_desc(schedule, state)
    layout = fs.layout
    if state:
        layout = _abi.tah_layout()
        check(lib().tah_layout_dense(dsc._descptr, fs.nvars, C.byref(layout)))
    key = (state, tuple(layout.stride), layout.nvars, ctx.device)
    plans = schedule._plans.setdefault("unify", {})
    if key not in plans:
        plans[key] = (Plan(ctx, dsc, layout), dsc)
    plan = plans[key][0]
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    target = v if state else fs.work
    check(lib().tah_fill(plan._h, C.c_void_p(target.data_ptr()), fptr, _stream_ptr(fs)))
    return v if state else fs


# ---------------------------------------------------------------------------
# Regrid (src/regrid.jl): refine and coarsen the forest and move every field
# set's own points onto the new leaves.

class RegridSchedule:
    """The regrid transfer of one field-set layout from the leaves `old`
    (root, level, coords arrays as Forest.leaf_arrays gives them) onto the
    forest's current leaves (transfer_groups / regrid_stage and the
    kept-over-created pass, src/regrid.jl:755-960)."""

    def __init__(self, forest, old, operators, *, G, centering, nvars, parity=None):
        root, level, coords = (np.ascontiguousarray(a) for a in old)
        D = forest.D
        par = None
        if parity is not None:
            flat = [int(x) for p in parity for x in p]
            par = (C.c_int8 * len(flat))(*flat)
        h = C.c_void_p()
        check(lib().tah_regrid_schedule_create(
            forest._h, len(root), root.ctypes.data_as(_abi.I32P), level.ctypes.data_as(_abi.I32P),
            coords.astype(np.uint32).ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(operators._c()),
            _i32(G if hasattr(G, "__len__") else (G,) * D), _i32(_centering_codes(centering, D)), int(nvars),
            par, C.byref(h)))
        self._h = h
        dp = C.POINTER(_abi.tah_schedule_desc)()
        check(lib().tah_schedule_desc_get(h, C.byref(dp)))
        self.desc = dp.contents
        self._descptr = dp

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _abi is not None and _abi._lib is not None:
            _abi._lib.tah_schedule_destroy(h)

    def stage_tags(self):
        return [self.desc.stage_tags[i] for i in range(self.desc.nstages)]

    def groups(self):
        return GhostSchedule.groups(self)


def regrid_transfer(fs, old_layout, old_work, schedule, *, ctx=None):
    """The transfer alone: a fresh zeroed array of fs's layout (fs.layout
    already the new one) filled from old_work over `schedule` (a
    RegridSchedule). Returns the new array; ghosts are left for the next fill."""
    import torch
    ctx = ctx or default_context()
    plan = Plan(ctx, schedule, fs.layout)
    fresh = torch.zeros(fs.nelements, dtype=torch.float64, device=old_work.device)
    check(lib().tah_regrid_transfer(plan._h, C.byref(old_layout), C.c_void_p(old_work.data_ptr()),
                                    C.c_void_p(fresh.data_ptr()),
                                    C.c_void_p(torch.cuda.current_stream(old_work.device).cuda_stream)))
    return fresh, plan


def regrid(forest, pairs, *, refine=(), coarsen=(), ctx=None):
    """regrid! (src/regrid.jl:1269-1395) with refinement and coarsening given
    as leaf indices: fill every set's ghosts (a prolongation reads its
    parent's), coarsen the families whose children are all listed (where
    2:1 balance allows), refine the listed leaves and rebalance, then move
    each set's own points onto the new leaves -- kept blocks copied, new
    ones prolongated from their parent or restricted from their children --
    and, for a set with a vertex-like dimension, give every created copy of
    a shared point the kept block's value. Each set's storage is replaced
    in place (fs.work, fs.layout); schedules built before are stale.
    Returns True if the mesh changed."""
    import torch
    pairs = [pairs] if isinstance(pairs, tuple) and len(pairs) == 2 and isinstance(pairs[0], FieldSet) else \
        list(pairs)
    for fs, sched in pairs:
        if fs.forest is not forest:
            raise ValueError("every field set must be over the forest being regridded")
        _check_fill(fs, sched)
    refine, coarsen = list(refine), list(coarsen)
    if set(refine) & set(coarsen):
        raise ValueError("a leaf flagged both for refinement and for coarsening")
    old = forest.leaf_arrays()
    keys = forest.leaves
    for fs, sched in pairs:
        fill_ghosts(fs, sched, ctx=ctx)
    if coarsen:
        forest.coarsen(coarsen)
    if refine:
        now = {k: i for i, k in enumerate(forest.leaves)}
        forest.refine([now[keys[i]] for i in refine])
    forest.balance()
    if forest.leaves == keys:
        return False
    for fs, sched in pairs:
        rs = RegridSchedule(forest, old, sched.operators, G=fs.G, centering=fs.centering, nvars=fs.nvars,
                            parity=fs.parity)
        old_layout, old_work = fs.layout, fs.work
        fs._relayout()
        fs.work, _ = regrid_transfer(fs, old_layout, old_work, rs, ctx=ctx)
        torch.cuda.current_stream(old_work.device).synchronize()
    return True


def _batch_args(pairs, ctx):
    ctx = ctx or default_context()
    n = len(pairs)
    plans = (C.c_void_p * n)()
    works = (C.c_void_p * n)()
    facs = (C.c_void_p * n)()
    keep = []
    for i, (fs, sched) in enumerate(pairs):
        _check_fill(fs, sched)
        if fs.work.device.type != "cuda":
            raise RuntimeError("batched fills run on the GPU only")
        pl = sched.plan(fs.layout, ctx)
        keep.append(pl)
        plans[i] = pl._h
        works[i] = fs.work.data_ptr()
        facs[i] = fs.factors.data_ptr() if fs.factors is not None else None
    return ctx, n, plans, works, facs, keep


def fill_ghosts_batched(pairs, *, ctx=None):
    """Fill many independent field sets ((fs, schedule) pairs, distinct working
    arrays) with one launch per kernel class and stage position."""
    ctx, n, plans, works, facs, keep = _batch_args(pairs, ctx)
    check(lib().tah_fill_batched(ctx._h, n, plans, works, facs, _stream_ptr(pairs[0][0])))
    return pairs


def tune_batched(pairs, cache_path=None, *, ctx=None):
    """tah_tune_batched: time the batched fill of these (fs, schedule) pairs
    for every combination of fork, defer and copy_per_thread (the same on
    every plan), keep the fastest, then the copies' tile order at it;
    cached by the batch's signature. Returns (dict(fork=, defer=,
    copy_per_thread=, copy_order=), median ms of the choice, or None when
    taken from the cache)."""
    ctx, n, plans, works, facs, keep = _batch_args(pairs, ctx)
    ms = C.c_double()
    path = None if cache_path is None else str(cache_path).encode()
    check(lib().tah_tune_batched(ctx._h, n, plans, works, facs, path, C.byref(ms)))
    o = pairs[0][1].plan(pairs[0][0].layout, ctx).options
    return (dict(fork=o["fork"], defer=o["defer"], copy_per_thread=o["copy_per_thread"],
                 copy_order=o["copy_order"]),
            None if ms.value < 0 else ms.value)


class FillGraph:
    """A batched fill captured once in a hipGraph and replayed by launch().
    The field sets' working arrays must stay where they are."""

    def __init__(self, pairs, *, ctx=None):
        ctx, n, plans, works, facs, keep = _batch_args(pairs, ctx)
        self._keep = (pairs, keep)
        h = C.c_void_p()
        check(lib().tah_graph_create(ctx._h, n, plans, works, facs, C.byref(h)))
        self._h = h
        self._stream_of = pairs[0][0]

    def launch(self):
        check(lib().tah_graph_launch(self._h, _stream_ptr(self._stream_of)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _abi is not None and _abi._lib is not None:
            _abi._lib.tah_graph_destroy(h)
