"""Interpolation streamed along z through LDS (plan option lds bit 2 for
restriction, bit 3 for prolongation; family option lds 2: interp_stream_kernel)
against the oracle, bit for bit: on a mesh of 32^3 blocks (faces, edges and
corners of the large forest's shape, with a reflecting face: parity factors),
in the arena layout 16-byte aligned (paired loads and stores) and 8 bytes off
(single values), with an odd variable count and in the dense layout; on config
1 (Conservative(3,2)), on PointValue(4,4), on vertex-centred meshes, and on
forty random forests of the batched configuration in both layouts."""
import pytest
import torch

import oracle
from treeamr_hip import FieldSet, GhostSchedule, Operators, fill_ghosts
from treeamr_hip.core import fill_ghosts_timed
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu

STREAM = 12  # plan option lds: bits 2 and 3, restriction and prolongation streamed along z


def _oracle(fs, sched):
    host = fs.work.detach().cpu().numpy().copy()
    fac = fs.factors.cpu().numpy().copy() if fs.factors is not None else None
    oracle.fill(sched._descptr, fs.layout, host, fac)
    return torch.from_numpy(host)


def _mismatches(fs, sched, lds=STREAM):
    if lds is not None:
        sched.plan(fs.layout).set_options(lds=lds)
    expected = _oracle(fs, sched)
    init = fs.work.clone()
    fill_ghosts(fs, sched)
    torch.cuda.synchronize()
    bad = int((fs.work.cpu().view(torch.int64) != expected.view(torch.int64)).sum())
    fs.work.copy_(init)
    return bad


def _names(fs, sched):
    return {name for name, _, _ in fill_ghosts_timed(fs, sched)}


def _unaligned(fs):
    buf = torch.empty(fs.nelements + 1, dtype=torch.float64, device="cuda")
    view = buf[1:]
    view.copy_(fs.work)
    fs.work = view


@pytest.mark.parametrize("layout,nvars,offset", [("arena", 4, 0), ("arena", 4, 1), ("arena", 3, 0),
                                                 ("arena", 20, 0), ("dense", 4, 0)])
def test_stream_on_32_cubed_blocks(layout, nvars, offset):
    f, ops, kw, _ = W.config5(nvars=nvars, roots=4, N=32)
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout=layout, fill=float("nan"), **kw)
    W.fill_gaussians(fs, 3)
    if offset:
        _unaligned(fs)
    sched = GhostSchedule(fs, ops)
    print(f"[stream] 32^3 blocks layout={layout} nvars={nv} offset={offset} blocks={f.nleaves}", flush=True)
    sched.plan(fs.layout).set_options(lds=STREAM)
    names = _names(fs, sched)
    assert "interp_stream_kernel<6, 4>" in names and "interp_stream_kernel<6, 12>" in names
    assert not any(n.startswith("interp_all_kernel") or n.startswith("interp_pair_kernel") for n in names)
    assert _mismatches(fs, sched) == 0
    for lds in (4, 8):
        assert _mismatches(fs, sched, lds=lds) == 0, lds


@pytest.mark.parametrize("layout", ["arena", "dense"])
def test_stream_conservative_order_2(layout):
    f, ops, kw, _ = W.config1()
    fs = FieldSet(f, kw["nvars"], G=kw["G"], device="cuda", layout=layout, fill=float("nan"))
    W.fill_uniform(fs, 0)
    sched = GhostSchedule(fs, ops)
    print(f"[stream] config 1 layout={layout}", flush=True)
    sched.plan(fs.layout).set_options(lds=STREAM)
    names = _names(fs, sched)
    assert "interp_stream_kernel<2, 2>" in names and "interp_stream_kernel<3, 6>" in names
    assert _mismatches(fs, sched) == 0


def test_stream_order_4():
    f, _, kw, _ = W.config5(nvars=6, roots=2, N=16)
    kw = dict(kw)
    nv = kw.pop("nvars")
    kw["G"] = 2
    fs = FieldSet(f, nv, device="cuda", layout="arena", fill=float("nan"), **kw)
    W.fill_gaussians(fs, 4)
    sched = GhostSchedule(fs, Operators(prolongation=4, restriction=4))
    print(f"[stream] PointValue(4,4) blocks={f.nleaves}", flush=True)
    sched.plan(fs.layout).set_options(lds=STREAM)
    names = _names(fs, sched)
    assert "interp_stream_kernel<4, 3>" in names and "interp_stream_kernel<4, 8>" in names
    assert _mismatches(fs, sched) == 0


def test_stream_family_option_and_refusal():
    f, ops, kw, _ = W.config5(nvars=2, roots=2, N=16)
    kw = dict(kw)
    nv = kw.pop("nvars")
    fs = FieldSet(f, nv, device="cuda", layout="arena", fill=float("nan"), **kw)
    W.fill_gaussians(fs, 5)
    sched = GhostSchedule(fs, ops)
    plan = sched.plan(fs.layout)
    plan.set_options(lds=0)
    plan.set_family(6, True, lds=2)
    names = _names(fs, sched)
    assert "interp_stream_kernel<6, 4>" in names and "interp_stream_kernel<6, 12>" not in names
    assert _mismatches(fs, sched, lds=None) == 0
    plan.set_family(6, False, lds=2)
    assert "interp_stream_kernel<6, 12>" in _names(fs, sched)
    assert _mismatches(fs, sched, lds=None) == 0
    before = plan.families
    with pytest.raises(Exception):
        plan.set_family(6, False, lds=3)
    assert plan.families == before
    with pytest.raises(Exception):
        plan.set_options(lds=16)


@pytest.mark.parametrize("layout,offset", [("arena", 0), ("arena", 1), ("dense", 0)])
def test_stream_vertex_centred(layout, offset):
    f, ops, kw, _ = W.config2(nvars=8, roots=4, N=16, radii=(0.35, 0.2))
    fs = FieldSet(f, kw["nvars"], G=kw["G"], centering=kw["centering"], parity=kw["parity"], device="cuda",
                  layout=layout, fill=float("nan"))
    W.fill_waves(fs, 1)
    if offset:
        _unaligned(fs)
    sched = GhostSchedule(fs, ops)
    print(f"[stream] vertex layout={layout} offset={offset} blocks={f.nleaves}", flush=True)
    sched.plan(fs.layout).set_options(lds=STREAM)
    assert "interp_stream_kernel<6, 12>" in _names(fs, sched)
    assert _mismatches(fs, sched) == 0


@pytest.mark.parametrize("layout", ["arena", "dense"])
def test_stream_on_forty_random_forests(layout):
    failed = []
    for seed in range(20, 60):
        f, ops, G, nvars, centering = W.random_forest(seed, max_blocks=500)
        nv = nvars[0]
        fs = FieldSet(f, nv, G=G, centering=centering, device="cuda", layout=layout, fill=float("nan"))
        W.fill_uniform(fs, seed)
        sched = GhostSchedule(fs, ops)
        print(f"[stream] seed={seed} {layout} {centering} G={G} nvars={nv} blocks={f.nleaves}", flush=True)
        bad = _mismatches(fs, sched)
        if bad:
            failed.append((seed, bad))
    assert not failed, failed
