"""scatter!/gather! on the GPU (tah_scatter / tah_gather) against the
oracle's loops over the state vector (src/state.jl:101-115), bit for bit,
and the reference's round-trip properties at full size
(test/closed_tests.jl:798-912): gather after scatter gives the state back,
and a scatter leaves every ghost as it was.
"""
import numpy as np
import pytest
import torch

import oracle
from treeamr_hip import FieldSet, Forest, gather, ownextents, scatter, statelength, statevector
from treeamr_hip.core import _own_first
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _forest(D=3, N=8, roots=2):
    f = Forest((roots,) * D, N=N, periodic=(True,) * D)
    if roots > 1:
        f.refine([0])
        f.balance()
    return f


def _bits(t):
    return t.detach().cpu().contiguous().view(torch.int64)


# (D, N, roots, centering, layout, nvars); 200 variables of 33-point rows
# cut the rows into runs of x (arena) or the variables into tiles (dense).
CASES = [(3, 8, 2, "cell", "dense", 4), (3, 8, 2, "vertex", "arena", 4), (3, 8, 2, "cell", "arena", 3),
         (3, 16, 2, "vertex", "arena", 20), (3, 32, 1, "vertex", "arena", 200),
         (3, 32, 1, "vertex", "dense", 200), (3, 128, 1, "cell", "arena", 20),
         (2, 8, 2, "vertex", "dense", 2), (2, 16, 2, "cell", "arena", 5), (1, 16, 2, "cell", "dense", 3)]


@pytest.mark.parametrize("D,N,roots,centering,layout,nvars", CASES)
def test_scatter_and_gather_match_the_oracle(D, N, roots, centering, layout, nvars):
    print(f"[state] D={D} N={N} roots={roots} {centering} {layout} nvars={nvars}", flush=True)
    fs = FieldSet(_forest(D, N, roots), nvars, G=2, centering=centering, device="cuda", layout=layout,
                  fill=float("nan"))
    g = torch.Generator(device="cuda")
    g.manual_seed(D * 100 + nvars)
    u = torch.rand(statelength(fs), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    host = fs.work.cpu().numpy().copy()
    oracle.scatter(fs.layout, _own_first(fs), ownextents(fs), host, u.cpu().numpy().copy())
    scatter(fs, u)
    torch.cuda.synchronize()
    assert (_bits(fs.work) == torch.from_numpy(host).view(torch.int64)).all()
    # gather from an array with data in the ghosts too
    fs.work.copy_(torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64))
    want = np.zeros(statelength(fs))
    oracle.gather(fs.layout, _own_first(fs), ownextents(fs), want, fs.work.cpu().numpy().copy())
    got = torch.full_like(u, float("nan"))
    gather(got, fs)
    torch.cuda.synchronize()
    assert (_bits(got) == torch.from_numpy(want).view(torch.int64)).all()


def test_several_sets_in_one_state_vector():
    f = _forest()
    a = FieldSet(f, 3, G=2, device="cuda", layout="arena", fill=0.0)
    b = FieldSet(f, 2, G=3, centering="vertex", device="cuda", fill=0.0)
    u = statevector((a, b))
    u.copy_(torch.arange(u.numel(), device="cuda", dtype=torch.float64))
    scatter((a, b), u)
    back = torch.zeros_like(u)
    gather(back, (a, b))
    torch.cuda.synchronize()
    assert torch.equal(back, u)
    # each set holds its own segment
    ua = torch.zeros(statelength(a), device="cuda", dtype=torch.float64)
    gather(ua, a)
    torch.cuda.synchronize()
    assert torch.equal(ua, u[:statelength(a)])


def test_unaligned_state_vector():
    fs = FieldSet(_forest(), 4, G=2, device="cuda", layout="arena", fill=0.0)
    n = statelength(fs)
    buf = torch.zeros(n + 1, device="cuda", dtype=torch.float64)
    u = buf[1:]
    u.copy_(torch.linspace(-1, 1, n, device="cuda", dtype=torch.float64))
    scatter(fs, u)
    back = torch.zeros(n + 1, device="cuda", dtype=torch.float64)[1:]
    gather(back, fs)
    torch.cuda.synchronize()
    assert torch.equal(back, u)


@pytest.mark.parametrize("layout", ["arena", "dense"])
def test_full_size_round_trip_leaves_ghosts(layout):
    """The large forest (about 5k blocks of 32^3, 20 variables)."""
    f, ops, kw, _ = W.config5(nvars=20)
    kw = dict(kw)
    fs = FieldSet(f, kw.pop("nvars"), device="cuda", layout=layout, fill=0.0, **kw)
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    fs.work.copy_(torch.rand(fs.work.shape, generator=g, device="cuda", dtype=torch.float64))
    k, j, i = fs.own_slices()
    ghost = torch.ones(fs.array().shape[2:], dtype=torch.bool, device="cuda")
    ghost[k, j, i] = False
    before = fs.array()[:, :, ghost].clone()
    u = torch.rand(statelength(fs), generator=g, device="cuda", dtype=torch.float64)
    scatter(fs, u)
    back = torch.zeros_like(u)
    gather(back, fs)
    torch.cuda.synchronize()
    assert torch.equal(back, u)
    assert torch.equal(fs.array()[:, :, ghost], before)


def test_refusals():
    fs = FieldSet(_forest(), 2, G=2, device="cuda", fill=0.0)
    with pytest.raises(ValueError, match="entries but this field set needs"):
        scatter(fs, torch.zeros(5, device="cuda", dtype=torch.float64))
    with pytest.raises(ValueError, match="same GPU"):
        scatter(fs, torch.zeros(statelength(fs), dtype=torch.float64))
    with pytest.raises(ValueError, match="Float64"):
        gather(torch.zeros(statelength(fs), device="cuda", dtype=torch.float32), fs)
