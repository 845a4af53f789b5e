"""The LDS-staged interpolation kernels (plan option lds: bit 0 restriction
from chunks of the fine source box, bit 1 prolongation of a face's inner
units from chunks of the coarse source box) against the oracle, bit for bit:
on a mesh whose faces are 32^3 blocks' (inner unit classes of 225 units),
on its 16-byte-misaligned twin (single loads), with a reflecting face
(factors), and on forty random forests of the batched configuration."""
import pytest
import torch

import oracle
from treeamr_hip import FieldSet, GhostSchedule, fill_ghosts
from treeamr_hip.core import fill_ghosts_timed
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _oracle(fs, sched):
    host = fs.work.detach().cpu().numpy().copy()
    fac = fs.factors.cpu().numpy().copy() if fs.factors is not None else None
    oracle.fill(sched._descptr, fs.layout, host, fac)
    return torch.from_numpy(host)


def _fill_and_compare(fs, sched, lds):
    sched.plan(fs.layout).set_options(lds=lds)
    expected = _oracle(fs, sched)
    init = fs.work.clone()
    fill_ghosts(fs, sched)
    torch.cuda.synchronize()
    bad = int((fs.work.cpu().view(torch.int64) != expected.view(torch.int64)).sum())
    fs.work.copy_(init)
    return bad


@pytest.mark.parametrize("offset", [0, 1])
def test_lds_kernels_on_32_cubed_faces(offset):
    f, ops, kw, _ = W.config5(nvars=4, roots=4, N=32)
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout="arena", fill=float("nan"), **kw)
    W.fill_gaussians(fs, 3)
    if offset:
        buf = torch.empty(fs.nelements + 1, dtype=torch.float64, device="cuda")
        view = buf[1:]
        view.copy_(fs.work)
        fs.work = view
    sched = GhostSchedule(fs, ops)
    for lds in (1, 2, 3):
        print(f"[lds] 32^3 faces offset={offset} lds={lds}", flush=True)
        sched.plan(fs.layout).set_options(lds=lds)
        names = {name for name, _, _ in fill_ghosts_timed(fs, sched)}
        if lds & 1:
            assert "interp_lds_kernel<6>" in names
        if lds & 2:
            assert "interp_lds_prolong_kernel<6>" in names
        assert _fill_and_compare(fs, sched, lds) == 0, lds


def test_lds_kernels_on_forty_random_forests():
    failed = []
    for seed in range(20, 60):
        f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=500)
        nv = nvars[0] + nvars[0] % 2
        fs = FieldSet(f, nv, G=G, centering=centering, device="cuda", layout="arena", fill=float("nan"))
        W.fill_uniform(fs, seed)
        sched = GhostSchedule(fs, ops)
        print(f"[lds] seed={seed} {centering} G={G} nvars={nv} blocks={f.nleaves}", flush=True)
        bad = _fill_and_compare(fs, sched, 3)
        if bad:
            failed.append((seed, bad))
    assert not failed, failed
