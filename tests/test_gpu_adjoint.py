"""The adjoint fill on the GPU (the Laplacian-loss configuration: the
120-block two-level mesh with 4 variables, Conservative(3,2)).

Pinned by the dot-product identity <fill(x), y> = <x, fill^T(y)>, by the
oracle's CPU transpose assembled from the same schedule, and by central
differences of a loss through fill + a 7-point Laplacian; plus a checkpointed
20-step RK4 wave differentiated end to end.
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import FieldSet, GhostSchedule, fill_ghosts, fill_ghosts_adjoint
from treeamr_hip import autodiff as AD
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _setup(which="config1", layout="dense"):
    if which == "config1":
        f, ops, kw, _ = W.config1()
    elif which == "vertex":
        f, ops, kw, _ = W.config2(roots=2, N=8, radii=(0.5, 0.3))
    else:
        f, ops, kw, _ = W.config5(nvars=3, roots=2, N=16)
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout=layout, fill=0.0, **kw)
    return fs, GhostSchedule(fs, ops)


@pytest.mark.parametrize("which,layout", [("config1", "dense"), ("config1", "arena"),
                                          ("reflecting", "dense"), ("reflecting", "arena"),
                                          ("vertex", "dense"), ("vertex", "arena")])
def test_dot_product_identity(which, layout):
    fs, s = _setup(which, layout)
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    own = AD.own_mask(fs)
    x = torch.where(own, torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1,
                    torch.zeros_like(fs.work))
    y = torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    fx = x.clone()
    AD._fill_tensor(fs, s, fx)
    aty = y.clone()
    fill_ghosts_adjoint(fs, s, aty)
    torch.cuda.synchronize()
    lhs = float(torch.dot(fx, y))
    rhs = float(torch.dot(x, aty))
    assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), 1.0), (lhs, rhs)


def test_dot_product_identity_random_forests():
    """The identity on twelve forests of the batched configuration (seeds
    20..31, cell and vertex, G 2 or 3, PointValue of order 2G), each in the
    dense layout and in the arena layout with an even variable count."""
    failed = []
    for seed in range(20, 32):
        f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=400)
        for layout in ("dense", "arena"):
            nv = nvars[0] + (nvars[0] % 2 if layout == "arena" else 0)
            print(f"[adjoint] seed={seed} {layout} {centering} G={G} nvars={nv} blocks={f.nleaves}", flush=True)
            fs = FieldSet(f, nv, G=G, centering=centering, device="cuda", layout=layout, fill=0.0)
            s = GhostSchedule(fs, ops)
            g = torch.Generator(device="cuda")
            g.manual_seed(seed)
            own = AD.own_mask(fs)
            x = torch.where(own, torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1,
                            torch.zeros_like(fs.work))
            y = torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
            fx = x.clone()
            AD._fill_tensor(fs, s, fx)
            aty = y.clone()
            fill_ghosts_adjoint(fs, s, aty)
            torch.cuda.synchronize()
            lhs = float(torch.dot(fx, y))
            rhs = float(torch.dot(x, aty))
            if abs(lhs - rhs) > 1e-13 * max(abs(lhs), 1.0):
                failed.append((seed, layout, lhs, rhs))
    assert not failed, failed


@pytest.mark.parametrize("which,layout,offset", [("config1", "dense", 0), ("reflecting", "dense", 0),
                                                 ("vertex", "dense", 0), ("config1", "arena", 0),
                                                 ("config1", "arena", 1), ("vertex", "arena", 0),
                                                 ("reflecting", "arena", 0)])
def test_adjoint_matches_cpu_transpose(which, layout, offset):
    """Offset 1 starts the array 8 bytes past a 16-byte boundary: the
    working-array accesses then move single values, not pairs."""
    fs, s = _setup(which, layout)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    y = torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    if offset:
        buf = torch.zeros(fs.nelements + 1, dtype=torch.float64, device="cuda")
        buf[1:].copy_(y)
        y = buf[1:]
        assert y.data_ptr() % 16 == 8
    host = y.cpu().numpy().copy()
    fac = fs.factors.cpu().numpy().copy() if fs.factors is not None else None
    oracle.fill_adjoint(s._descptr, fs.layout, host, fac)
    fill_ghosts_adjoint(fs, s, y)
    torch.cuda.synchronize()
    got = y.cpu().numpy()
    np.testing.assert_allclose(got, host, rtol=1e-12, atol=1e-13)


def test_gradient_by_central_differences():
    """loss = 0.5 sum L(fill(u))^2: autograd gradient against central
    differences, step 1e-6, on 32 seeded directions."""
    fs, s = _setup("config1")
    h = AD.spacings(fs)
    N = fs.forest.N
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    shape = (fs.nblocks, fs.nvars, N, N, N)
    u = (torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1).requires_grad_(True)
    loss = AD.loss_laplacian(fs, s, u, h)
    grad, = torch.autograd.grad(loss, u)
    eps = 1e-6
    worst = 0.0
    with torch.no_grad():
        for _ in range(32):
            d = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
            lp = AD.loss_laplacian(fs, s, u + eps * d, h)
            lm = AD.loss_laplacian(fs, s, u - eps * d, h)
            fd = float((lp - lm) / (2 * eps))
            an = float((grad * d).sum())
            worst = max(worst, abs(fd - an) / max(abs(an), 1e-30))
    assert worst < 1e-6, worst


def test_rk4_wave_gradient_with_checkpoints():
    """20 RK4 steps of u_t = v, v_t = L(fill(u)), checkpointed every 5 steps:
    the gradient of 0.5*|u_final|^2 against central differences, and equal
    to the gradient without checkpoints."""
    fs, s = _setup("config1")
    N = fs.forest.N
    h = AD.spacings(fs)
    dt = 0.2 * float(h.min())
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    shape = (fs.nblocks, fs.nvars, N, N, N)
    u0 = (torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1) * 1e-2
    v0 = torch.zeros(shape, device="cuda", dtype=torch.float64)

    def objective(u, every):
        uf, _ = AD.wave_rk4(fs, s, u, v0, dt, 20, checkpoint_every=every)
        return 0.5 * (uf * uf).sum()

    u = u0.clone().requires_grad_(True)
    g_ck, = torch.autograd.grad(objective(u, 5), u)
    u = u0.clone().requires_grad_(True)
    g_full, = torch.autograd.grad(objective(u, 0), u)
    assert torch.allclose(g_ck, g_full, rtol=1e-12, atol=1e-15)
    eps = 1e-6
    with torch.no_grad():
        for _ in range(1):
            d = torch.rand(shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
            fd = float((objective(u0 + eps * d, 0) - objective(u0 - eps * d, 0)) / (2 * eps))
            an = float((g_ck * d).sum())
            assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)


def test_adjoint_timed_bytes_closed_form():
    """tah_fill_adjoint_timed on a uniform periodic cell mesh, where every
    transfer is a copy: the launches are gathers only, their results are
    tah_fill_adjoint's bit for bit, and their algorithmic bytes are the
    closed form 16 x nvars x blocks x (the own points within G of a face,
    each a source read and written once, plus the ghost points, each read
    and zeroed once)."""
    from treeamr_hip import Forest, Operators, fill_ghosts_adjoint_timed
    N, G, nv = 8, 2, 3
    f = Forest((2, 2, 2), N=N, periodic=(True,) * 3)
    fs = FieldSet(f, nv, G=G, device="cuda", fill=0.0)
    s = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    y = torch.rand_like(fs.work)
    a, b = y.clone(), y.clone()
    fill_ghosts_adjoint(fs, s, a)
    recs = fill_ghosts_adjoint_timed(fs, s, b)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int64), b.view(torch.int64))
    assert recs and all(name == "adjoint_gather_kernel" for name, _, _ in recs)
    own_near_face = N ** 3 - (N - 2 * G) ** 3
    ghosts = (N + 2 * G) ** 3 - N ** 3
    assert sum(nb for _, nb, _ in recs) == 16 * nv * f.nleaves * (own_near_face + ghosts)
    assert all(ms > 0 for _, _, ms in recs)


@pytest.mark.parametrize("which,layout", [("vertex", "arena"), ("reflecting", "arena"), ("reflecting", "dense"),
                                          ("config1", "arena")])
def test_forked_zeroing_equals_launch_by_launch(which, layout):
    """tah_fill_adjoint runs a stage's zeroing of its interpolated targets on
    a second stream beside the stage's y and x passes and gather;
    tah_fill_adjoint_timed runs every launch in order on one. On meshes with
    restriction and prolongation both give the same bits, twice in a row."""
    from treeamr_hip import fill_ghosts_adjoint_timed
    fs, s = _setup(which, layout)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    y = torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    for _ in range(2):
        a, b = y.clone(), y.clone()
        fill_ghosts_adjoint(fs, s, a)
        recs = fill_ghosts_adjoint_timed(fs, s, b)
        torch.cuda.synchronize()
        assert any(name == "zero_kernel" for name, _, _ in recs)
        assert torch.equal(a.view(torch.int64), b.view(torch.int64))
        y = a
