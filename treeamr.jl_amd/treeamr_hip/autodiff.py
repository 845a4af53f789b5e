"""Differentiating through the fill.

The fill is linear: filled = F(work), ghosts overwritten from own points
stage by stage. Its transpose runs on the device (tah_fill_adjoint), and
FillGhosts wires the pair into torch's autograd -- the counterpart of a
ChainRulesCore rrule on fill_ghosts! for a Julia caller. On top of it:

  laplacian(fs, filled)      the 7-point Laplacian (sum_d u(i+e_d) + u(i-e_d) - 6u) / h^2
                             over own points, h the block's spacing
  wave_rk4(fs, sched, u, v, dt, steps, checkpoint_every)
                             RK4 of u_t = v, v_t = L(fill(u)), differentiable, with
                             recomputation from checkpoints every few steps
"""
import torch
import torch.utils.checkpoint as cp

from .core import fill_ghosts, fill_ghosts_adjoint


class FillGhosts(torch.autograd.Function):
    """y = fill(x) on a copy of x; dL/dx = fill^T(dL/dy)."""

    @staticmethod
    def forward(ctx, work, fs, schedule):
        ctx.fs = fs
        ctx.schedule = schedule
        out = work.detach().clone()
        _fill_tensor(fs, schedule, out)
        return out

    @staticmethod
    def backward(ctx, grad):
        g = grad.detach().clone().contiguous()
        _adjoint_tensor(ctx.fs, ctx.schedule, g)
        return g, None, None


def _swap(fs, t):
    saved = fs.work
    fs.work = t
    return saved


def _fill_tensor(fs, schedule, t):
    saved = _swap(fs, t)
    try:
        fill_ghosts(fs, schedule)
    finally:
        fs.work = saved


def _adjoint_tensor(fs, schedule, t):
    fill_ghosts_adjoint(fs, schedule, t)


def fill(fs, schedule, work):
    return FillGhosts.apply(work, fs, schedule)


def view(fs, work):
    """work as [block, var, k, j, i], through reshape and permute where the
    layout allows (cheap under autograd), else through as_strided."""
    s = fs.layout.stride
    S = fs.stored
    nb, nv = fs.nblocks, fs.nvars
    dense = (s[0] == 1 and s[1] == S[0] and s[2] == S[0] * S[1] and s[3] == S[0] * S[1] * S[2]
             and s[4] == s[3] * nv)
    vin = (s[3] == 1 and s[0] == nv and s[1] == nv * S[0] and s[2] == nv * S[0] * S[1]
           and s[4] == s[2] * S[2])
    if dense:
        return work.reshape(nb, nv, S[2], S[1], S[0])
    if vin:
        return work.reshape(nb, S[2], S[1], S[0], nv).permute(0, 4, 1, 2, 3)
    return work.as_strided((nb, nv, S[2], S[1], S[0]), (s[4], s[3], s[2], s[1], s[0]))


def spacings(fs):
    _, h, _ = fs.forest.block_geometry()
    return torch.tensor(h, dtype=torch.float64, device=fs.work.device)


def laplacian(fs, filled, h=None):
    """(sum over dims of the two neighbours - 6u) / h^2 at own points: [b, v, N, N, N]."""
    a = view(fs, filled)
    G, N = fs.G, fs.forest.N
    if h is None:
        h = spacings(fs)
    k0, j0, i0 = G[2], G[1], G[0]
    c = a[:, :, k0:k0 + N, j0:j0 + N, i0:i0 + N]
    acc = (a[:, :, k0:k0 + N, j0:j0 + N, i0 - 1:i0 - 1 + N] + a[:, :, k0:k0 + N, j0:j0 + N, i0 + 1:i0 + 1 + N]
           + a[:, :, k0:k0 + N, j0 - 1:j0 - 1 + N, i0:i0 + N] + a[:, :, k0:k0 + N, j0 + 1:j0 + 1 + N, i0:i0 + N]
           + a[:, :, k0 - 1:k0 - 1 + N, j0:j0 + N, i0:i0 + N] + a[:, :, k0 + 1:k0 + 1 + N, j0:j0 + N, i0:i0 + N]
           - 6.0 * c)
    return acc / (h * h)[:, None, None, None, None]


def own_mask(fs):
    m = torch.zeros_like(fs.work, dtype=torch.bool)
    view(fs, m)[(slice(None), slice(None)) + fs.own_slices()] = True
    return m


def scatter_own(fs, own_values):
    """A working array holding own_values [b, v, N, N, N] at own points, 0 elsewhere
    (scatter!, src/state.jl:136-140)."""
    G, N, S = fs.G, fs.forest.N, fs.stored
    pad = []
    for d in range(3):  # F.pad takes the last dimension first: i, j, k
        pad += [G[d], S[d] - G[d] - N - fs.centering[d]]
    full = torch.nn.functional.pad(own_values, pad)  # [b, v, k, j, i]
    s = fs.layout.stride
    dense = s[0] == 1 and s[4] == s[3] * fs.nvars and s[3] == S[0] * S[1] * S[2]
    if dense:
        return full.reshape(-1)
    w = torch.zeros_like(fs.work)
    view(fs, w)[...] = full
    return w


def loss_laplacian(fs, schedule, own_u, h=None):
    """0.5 * sum over own points of L(fill(u))^2."""
    w = scatter_own(fs, own_u)
    L = laplacian(fs, fill(fs, schedule, w), h)
    return 0.5 * (L * L).sum()


def wave_rk4(fs, schedule, u, v, dt, steps, checkpoint_every=5):
    """RK4 for u_t = v, v_t = L(fill(u)) on own points ([b, v, N, N, N] tensors)."""
    h = spacings(fs)

    def rhs(uu, vv):
        return vv, laplacian(fs, fill(fs, schedule, scatter_own(fs, uu)), h)

    def step(uu, vv):
        k1u, k1v = rhs(uu, vv)
        k2u, k2v = rhs(uu + 0.5 * dt * k1u, vv + 0.5 * dt * k1v)
        k3u, k3v = rhs(uu + 0.5 * dt * k2u, vv + 0.5 * dt * k2v)
        k4u, k4v = rhs(uu + dt * k3u, vv + dt * k3v)
        return (uu + dt / 6 * (k1u + 2 * k2u + 2 * k3u + k4u),
                vv + dt / 6 * (k1v + 2 * k2v + 2 * k3v + k4v))

    def segment(n):
        def run(uu, vv):
            for _ in range(n):
                uu, vv = step(uu, vv)
            return uu, vv
        return run

    if not checkpoint_every or checkpoint_every <= 0:
        return segment(steps)(u, v)
    done = 0
    while done < steps:
        n = min(checkpoint_every, steps - done)
        if torch.is_grad_enabled():
            u, v = cp.checkpoint(segment(n), u, v, use_reentrant=False)
        else:
            u, v = segment(n)(u, v)
        done += n
    return u, v
