"""Seeded synthetic meshes and fields of the benchmark configurations.

Each builder returns (forest, operators, fieldset-kwargs, description). Data
generators write own points only; ghosts start as NaN (or a given value) so
that a ghost the fill misses stays visible.

  config1: 2x2x2 periodic roots, N=8, cell, G=2, 4 vars, Conservative(3,2);
           every root refined once, then every level-1 leaf whose centre is
           within 0.3 of the domain centre: 120 blocks (56 + 64).
  config2: 4x4x4 roots, N=16, vertex, G=3, 8 vars, PointValue(6,6); nested
           refinement around the domain centre at radii 0.35/0.2/0.1 (leaf
           centre within the radius, then balance): 232 blocks (56+56+56+64);
           reflecting z-low face (variables 1-6 even, 7-8 odd in z), outflow
           elsewhere; data sin(2 pi k.x + phase) from seed 1 plus 1e-3 noise.
  config5: 8x8x8 roots, N=32, cell, G=3, 20 vars, PointValue(6,6); refined
           twice where leaves meet 4 seeded spheres, 2:1 balanced; reflecting
           z-low face (parities alternating even/odd in z), outflow elsewhere.
  config3: 64 independent forests drawn from seeds 2..65 (see random_forest).
"""
import math

import numpy as np

from .core import (Conservative, EvenParity, FieldSet, Forest, NoParity, OddParity, Operators,
                   PointValue)


def config1_forest():
    f = Forest((2, 2, 2), N=8, periodic=(True,) * 3, extents=((0.0, 1.0),) * 3)
    f.refine(range(f.nleaves))
    c = f.block_centers()
    idx = [i for i in range(f.nleaves) if np.linalg.norm(c[i] - 0.5) < 0.3]
    f.refine(idx)
    f.balance()
    return f


def config1():
    f = config1_forest()
    ops = Operators(prolongation=3, restriction=2, family=Conservative)
    return f, ops, dict(nvars=4, G=2), "config1: 2^3 periodic roots, N=8, cell, G=2, 4 vars, Conservative(3,2)"


def config2_forest(roots=4, N=16, radii=(0.35, 0.2, 0.1)):
    f = Forest((roots,) * 3, N=N, extents=((0.0, 1.0),) * 3,
               reflecting=((False, False), (False, False), (True, False)),
               outflow=((True, True), (True, True), (False, True)))
    for lvl, r in enumerate(radii):
        c = f.block_centers()
        _, _, level = f.block_geometry()
        f.refine([i for i in range(f.nleaves) if level[i] == lvl and np.linalg.norm(c[i] - 0.5) < r])
        f.balance()
    return f


def config2(nvars=8, roots=4, N=16, radii=(0.35, 0.2, 0.1)):
    f = config2_forest(roots=roots, N=N, radii=radii)
    ops = Operators(prolongation=6, restriction=6)
    parity = tuple((NoParity, NoParity, EvenParity if v < nvars - 2 else OddParity)
                   for v in range(nvars))
    return f, ops, dict(nvars=nvars, G=3, centering="vertex", parity=parity), (
        f"config2: {roots}^3 roots, N={N}, vertex, G=3, {nvars} vars, PointValue(6,6), "
        "nested refinement, reflecting z-lo, outflow elsewhere")


def fill_waves(fs, seed, noise=1e-3):
    """Own points = sin(2 pi k_v.x + phase_v) per variable (k_v, phase_v from
    `seed`), plus uniform noise of amplitude `noise` on every stored own copy."""
    import torch
    rng = np.random.default_rng(seed)
    dev = fs.work.device
    X = [torch.tensor(x, device=dev) for x in fs.coordinates()]
    a = fs.array()
    own = fs.own_slices()
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    for v in range(fs.nvars):
        k = rng.integers(1, 4, size=3).astype(float)
        ph = rng.uniform(0.0, 2 * math.pi)
        val = torch.sin(2 * math.pi * (k[0] * X[0] + k[1] * X[1] + k[2] * X[2]) + ph)
        val = val[(slice(None),) + own]
        val = val + noise * (torch.rand(val.shape, generator=g, dtype=torch.float64, device=dev) * 2 - 1)
        a[(slice(None), v) + own] = val
    return fs


# --- the reference's own benchmark meshes (H200 timings published with them) ---

def uniform_vertex(N=16, roots=8, nvars=20, G=3, p=4):
    """bench/copies.jl's default: a uniform periodic mesh of roots^3 blocks of
    N^3, vertex-centred, G=3, 20 variables, PointValue(p) (p=4)."""
    f = Forest((roots,) * 3, N=N, periodic=(True,) * 3, extents=((0.0, 1.0),) * 3)
    return f, Operators(prolongation=p, restriction=p), dict(nvars=nvars, G=G, centering="vertex")


def nested_octant(N=32, nvars=20):
    """bench/nested.jl's mesh: the reflecting octant [0,32]^3 (low faces
    reflecting, high faces outer), 4^3 roots, the cubes [0,16)^3, [0,8)^3 and
    [0,4)^3 refined to levels 1-3 (232 leaves), vertex-centred, G=3,
    PointValue(6,6); odd variables (1-based) scalars, even ones a vector
    component odd along dimension mod1(v/2, 3)."""
    f = Forest((4,) * 3, N=N, reflecting=((True, False),) * 3, extents=((0.0, 32.0),) * 3)
    for lvl in range(1, 4):
        edge = 32.0 / 2 ** lvl
        lows, h, level = f.block_geometry()
        hi = lows + (h * f.N)[:, None]
        f.refine([i for i in range(f.nleaves) if level[i] == lvl - 1 and np.all(hi[i] <= edge)])
    f.balance()
    parity = []
    for v in range(1, nvars + 1):
        if v % 2 == 1:
            parity.append((EvenParity,) * 3)
        else:
            d = ((v // 2) - 1) % 3
            parity.append(tuple(OddParity if e == d else EvenParity for e in range(3)))
    return f, Operators(prolongation=6, restriction=6), dict(nvars=nvars, G=3, centering="vertex",
                                                               parity=tuple(parity))


def two_level(N=32, roots=8, nvars=2):
    """bench/gpu.jl's mesh: roots^3 periodic blocks of N^3 with the middle half
    of the domain refined once (960 blocks at 8^3 roots), cell-centred, G=2,
    2 variables, PointValue(4,4)."""
    f = Forest((roots,) * 3, N=N, periodic=(True,) * 3, extents=((0.0, 1.0),) * 3)
    c = f.block_centers()
    f.refine([i for i in range(f.nleaves) if np.all((c[i] > 0.25) & (c[i] < 0.75))])
    f.balance()
    return f, Operators(prolongation=4, restriction=4), dict(nvars=nvars, G=2)


def _spheres(seed, n, rlo, rhi):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.15, 0.85, size=(n, 3))
    radii = rng.uniform(rlo, rhi, size=n)
    return centres, radii


def _box_meets_sphere(lo, hi, c, r):
    q = np.clip(c[None, :], lo, hi)
    return np.sum((q - c[None, :]) ** 2, axis=1) <= r * r


def config5_forest(roots=8, N=32, seed=5, levels=2, spheres=4, radius=(0.12, 0.16)):
    f = Forest((roots,) * 3, N=N, extents=((0.0, 1.0),) * 3,
               reflecting=((False, False), (False, False), (True, False)),
               outflow=((True, True), (True, True), (False, True)))
    centres, radii = _spheres(seed, spheres, *radius)
    for lvl in range(levels):
        lows, h, level = f.block_geometry()
        hi = lows + (h * f.N)[:, None]
        hit = np.zeros(f.nleaves, bool)
        for c, r in zip(centres, radii):
            hit |= _box_meets_sphere(lows, hi, c, r)
        hit &= level == lvl
        f.refine(np.nonzero(hit)[0])
        f.balance()
    return f


def config5(nvars=20, roots=8, N=32):
    f = config5_forest(roots=roots, N=N)
    ops = Operators(prolongation=6, restriction=6)
    parity = tuple((EvenParity, EvenParity, EvenParity if v % 2 == 0 else OddParity)
                   for v in range(nvars))
    return f, ops, dict(nvars=nvars, G=3, parity=parity), (
        f"config5: {roots}^3 roots, N={N}, cell, G=3, {nvars} vars, PointValue(6,6), "
        "reflecting z-lo, outflow elsewhere")


def fill_uniform(fs, seed, lo=-1.0, hi=1.0):
    """Own points uniform in [lo, hi) from a seeded generator (torch, on fs's device)."""
    import torch
    g = torch.Generator(device=fs.work.device)
    g.manual_seed(seed)
    a = fs.array()
    own = (slice(None), slice(None)) + fs.own_slices()
    shape = a[own].shape
    a[own] = torch.rand(shape, generator=g, dtype=torch.float64, device=fs.work.device) * (hi - lo) + lo
    return fs


def fill_gaussians(fs, seed, noise=1e-4):
    """Own points = a Gaussian bump per variable (seeded centre and width) plus noise."""
    import torch
    rng = np.random.default_rng(seed)
    dev = fs.work.device
    a = fs.array()
    own = fs.own_slices()
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    # own points' positions per dimension, (nb, own extent), one rounding each
    x = [torch.tensor(fs.positions(d)[:, own[2 - d]], dtype=torch.float64, device=dev) for d in range(3)]
    for v in range(fs.nvars):
        c = rng.uniform(0.2, 0.8, size=3)
        s = rng.uniform(0.05, 0.2)
        ex = [torch.exp(-((x[d] - c[d]) ** 2) / (2 * s * s)) for d in range(3)]
        val = ex[2][:, :, None, None] * ex[1][:, None, :, None] * ex[0][:, None, None, :]
        val = val + noise * (torch.rand(val.shape, generator=g, dtype=torch.float64, device=dev) * 2 - 1)
        a[(slice(None), v) + own] = val
    return fs


def config3(count=64, first_seed=2, max_bytes=2 << 30, layout="arena", device="cuda", seed_data=True):
    """The batched configuration: `count` independent forests from seeds
    first_seed.., each with its 1-3 field sets; returns (fs, schedule) pairs."""
    from .core import GhostSchedule
    pairs = []
    for i in range(count):
        sd = first_seed + i
        f, ops, G, nvars, centering = random_forest(sd, max_blocks=4096, max_bytes=max_bytes)
        for k, nv in enumerate(nvars):
            fs = FieldSet(f, nv, G=G, centering=centering, device=device, layout=layout, fill=0.0)
            if seed_data:
                fill_uniform(fs, 1000 * sd + k)
            pairs.append((fs, GhostSchedule(fs, ops)))
    return pairs


def make_fieldset(forest, device="cuda", layout="dense", fill=float("nan"), **kw):
    return FieldSet(forest, kw.pop("nvars"), device=device, layout=layout, fill=fill, **kw)


def random_forest(seed, max_blocks=4096, max_bytes=None):
    """One forest of the batched configuration, drawn from `seed`: brick of
    r^3 roots (r in 1..4), N in {8,16,32}, 1-4 refinement levels around 1-3
    spheres of radius 0.05-0.3, G in {2,3}, 1-3 field sets of 2-10 variables.
    All faces periodic; PointValue of order 2G; cell or vertex centering.
    Draws whose blocks exceed max_blocks, or whose field sets together exceed
    max_bytes, are redrawn with fewer levels. Returns (forest, operators, G,
    nvars per field set, centering)."""
    rng = np.random.default_rng(seed)
    r = int(rng.integers(1, 5))
    N = int(rng.choice([8, 16, 32]))
    levels = int(rng.integers(1, 5))
    nsph = int(rng.integers(1, 4))
    centres = rng.uniform(0.1, 0.9, size=(nsph, 3))
    radii = rng.uniform(0.05, 0.3, size=nsph)
    G = int(rng.choice([2, 3]))
    nsets = int(rng.integers(1, 4))
    nvars = [int(rng.integers(2, 11)) for _ in range(nsets)]
    centering = "vertex" if rng.integers(0, 2) == 1 else "cell"
    cv = 1 if centering == "vertex" else 0

    def too_big(f):
        if f.nleaves > max_blocks:
            return True
        if max_bytes is not None:
            per_block = (N + 2 * G + cv) ** 3 * 8 * sum(nvars)
            return f.nleaves * per_block > max_bytes
        return False

    while True:
        f = Forest((r,) * 3, N=N, periodic=(True,) * 3, extents=((0.0, 1.0),) * 3)
        for lvl in range(levels):
            lows, h, level = f.block_geometry()
            hi = lows + (h * f.N)[:, None]
            hit = np.zeros(f.nleaves, bool)
            for c, rad in zip(centres, radii):
                hit |= _box_meets_sphere(lows, hi, c, rad)
            hit &= level == lvl
            f.refine(np.nonzero(hit)[0])
            f.balance()
            if too_big(f):
                break
        if not too_big(f) or levels == 1:
            break
        levels -= 1
    ops = Operators(prolongation=2 * G, restriction=2 * G)
    return f, ops, G, nvars, centering
