"""Plan options and the autotuner: every combination of the launch options
fills bit for bit as the defaults do, and a tuned choice is cached by mesh
signature and taken back from the cache without timing."""
import itertools

import pytest
import torch

from treeamr_hip import FieldSet, GhostSchedule, fill_ghosts, tune
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _case(which):
    if which in ("cell", "cell-even"):
        f, ops, kw, _ = W.config5(nvars=3 if which == "cell" else 4, roots=4, N=16)
    else:
        f, ops, kw, _ = W.config2(roots=2, N=8, radii=(0.5, 0.3))
    kw = dict(kw)
    fs = FieldSet(f, kw.pop("nvars"), device="cuda", layout="arena", fill=float("nan"), **kw)
    W.fill_uniform(fs, 9)
    return fs, GhostSchedule(fs, ops)


@pytest.mark.parametrize("which", ["cell", "cell-even", "vertex"])
def test_every_option_combination_fills_alike(which):
    """Odd (3) and even (4, 8) variable counts: the even ones take the
    variable-pair interpolations when interp_pairs is on."""
    fs, s = _case(which)
    init = fs.work.clone()
    fill_ghosts(fs, s)
    torch.cuda.synchronize()
    ref = fs.work.clone()
    plan = s.plan(fs.layout)
    for fork, merge, defer, cpt, wgs, ip, lds in itertools.product((0, 1), (0, 1), (0, 1), (4, 8, 16),
                                                                   (0, 64, 512), (0, 1), (0, 1, 2, 3, 4, 6, 8, 12)):
        print(f"[tune] {which} fork={fork} merge={merge} defer={defer} copy_per_thread={cpt} copy_wgs={wgs} "
              f"interp_pairs={ip} lds={lds}", flush=True)
        plan.set_options(fork=fork, merge=merge, fuse=0, defer=defer, copy_per_thread=cpt, copy_wgs=wgs,
                         interp_pairs=ip, lds=lds)
        fs.work.copy_(init)
        fill_ghosts(fs, s)
        torch.cuda.synchronize()
        assert torch.equal(fs.work.view(torch.int64), ref.view(torch.int64)), (fork, merge, defer, cpt, wgs, ip,
                                                                                lds)
    # copy tiles by target block
    for fork, defer, cpt, wgs in itertools.product((0, 1), (0, 1), (4, 8, 16), (0, 64)):
        print(f"[tune] {which} copy_order=1 fork={fork} defer={defer} copy_per_thread={cpt} copy_wgs={wgs}",
              flush=True)
        plan.set_options(fork=fork, merge=1, fuse=0, defer=defer, copy_per_thread=cpt, copy_wgs=wgs,
                         interp_pairs=1, lds=4, copy_order=1)
        fs.work.copy_(init)
        fill_ghosts(fs, s)
        torch.cuda.synchronize()
        assert torch.equal(fs.work.view(torch.int64), ref.view(torch.int64)), (fork, defer, cpt, wgs)
    plan.set_options(fork=0, merge=1, fuse=0, defer=1, copy_per_thread=8, copy_wgs=0, interp_pairs=1, lds=4,
                     copy_order=0)


@pytest.mark.parametrize("which", ["cell", "cell-even", "vertex"])
def test_every_family_option_fills_alike(which):
    """Every interpolation family's workgroup size, merged or per-class
    launches, variable pairs and LDS staging, one family at a time, with
    the copies forked beside them or not: bit for bit the defaults."""
    fs, s = _case(which)
    init = fs.work.clone()
    fill_ghosts(fs, s)
    torch.cuda.synchronize()
    ref = fs.work.clone()
    plan = s.plan(fs.layout)
    fams = plan.families
    assert fams and all(f["threads"] == 256 for f in fams)
    for fork in (0, 1):
        plan.set_options(fork=fork)
        for fam in fams:
            for threads, merge, pairs, lds in itertools.product((64, 128, 256), (0, 1), (0, 1), (0, 1, 2)):
                if fam["restriction"] and pairs or pairs and not merge or lds == 2 and not fam["restriction"]:
                    continue
                print(f"[family] {which} fork={fork} order={fam['order']} restriction={fam['restriction']} "
                      f"threads={threads} merge={merge} pairs={pairs} lds={lds}", flush=True)
                plan.set_family(fam["order"], fam["restriction"], threads=threads, merge=merge, pairs=pairs, lds=lds,
                                vars=0)
                fs.work.copy_(init)
                fill_ghosts(fs, s)
                torch.cuda.synchronize()
                assert torch.equal(fs.work.view(torch.int64), ref.view(torch.int64)), (fam, threads, merge, pairs,
                                                                                        lds)
            for threads, nv in itertools.product((64, 256), (1, 2, 3, 4)):
                print(f"[family] {which} fork={fork} order={fam['order']} restriction={fam['restriction']} "
                      f"threads={threads} vars={nv}", flush=True)
                plan.set_family(fam["order"], fam["restriction"], threads=threads, merge=1, lds=0, vars=nv,
                                pairs=0 if fam["restriction"] else 1)
                fs.work.copy_(init)
                fill_ghosts(fs, s)
                torch.cuda.synchronize()
                assert torch.equal(fs.work.view(torch.int64), ref.view(torch.int64)), (fam, threads, nv)
            plan.set_family(fam["order"], fam["restriction"], threads=256, merge=1,
                            pairs=0 if fam["restriction"] else 1, lds=2 if fam["restriction"] else 0, vars=0)
    # the plan-wide options set every family back
    plan.set_family(fams[0]["order"], fams[0]["restriction"], threads=64)
    plan.set_options(fork=0, merge=1, defer=1, copy_per_thread=8, copy_wgs=0, interp_pairs=1, lds=4)
    assert all(f["threads"] == 256 for f in plan.families)


def test_family_options_refused():
    from treeamr_hip import _abi
    fs, s = _case("cell")
    plan = s.plan(fs.layout)
    fam = plan.families[0]
    before = plan.families
    for bad in (dict(threads=96), dict(threads=512), dict(merge=2), dict(lds=-1), dict(lds=3), dict(vars=-1),
                dict(vars=300)):
        with pytest.raises(ValueError):
            plan.set_family(fam["order"], fam["restriction"], **bad)
        assert plan.families == before, bad
    with pytest.raises(ValueError):
        plan.set_family(7, 0)
    with pytest.raises(ValueError):
        plan.set_family(fam["order"], 1, pairs=1)
    assert _abi.tah_family_options().__class__.__name__ == "tah_family_options"


def test_tune_caches_by_signature(tmp_path):
    fs, s = _case("cell")
    path = tmp_path / "tune.txt"
    opts, ms = tune(fs, s, path)
    assert ms is not None and ms > 0
    lines = path.read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("tah-tune-7 ")
    assert opts["families"] and all(f["threads"] in (64, 128, 256) for f in opts["families"])
    s.plan(fs.layout).set_options(fork=1 - opts["fork"])
    s.plan(fs.layout).set_family(opts["families"][0]["order"], opts["families"][0]["restriction"],
                                 threads=64 if opts["families"][0]["threads"] != 64 else 128)
    opts2, ms2 = tune(fs, s, path)
    assert ms2 is None and opts2 == opts
    assert len(path.read_text().splitlines()) == 1
    # the tuned plan still fills correctly
    init = fs.work.clone()
    fill_ghosts(fs, s)
    torch.cuda.synchronize()
    once = fs.work.clone()
    fill_ghosts(fs, s)
    torch.cuda.synchronize()
    assert torch.equal(once.view(torch.int64), fs.work.view(torch.int64))
    del init


def test_fused_launch_refused():
    """The fused copy-and-interpolation launch is switched off: asking for it
    is an error raised on the host, before any launch."""
    from treeamr_hip import UnsupportedError
    fs, s = _case("cell")
    plan = s.plan(fs.layout)
    with pytest.raises(UnsupportedError):
        plan.set_options(fuse=1)
    assert plan.options["fuse"] == 0


def test_option_values_refused():
    """Out-of-range options are refused on the host, before any launch, and
    leave the plan's options as they were."""
    fs, s = _case("cell-even")
    plan = s.plan(fs.layout)
    before = plan.options
    for bad in (dict(interp_pairs=2), dict(interp_pairs=-1), dict(copy_per_thread=5), dict(copy_wgs=12),
                dict(fork=2), dict(lds=16), dict(lds=-1), dict(queue=2), dict(queue=-1), dict(copy_order=2),
                dict(copy_order=-1)):
        with pytest.raises(ValueError):
            plan.set_options(**bad)
        assert plan.options == before, bad


@pytest.mark.skip(reason="the fused copy-and-interpolation launch faulted the GPU (illegal address) in a timing "
                         "run on the large forest; its cause was not found, so the launch is switched off and "
                         "this case is not run")
@pytest.mark.parametrize("which", ["cell", "vertex"])
def test_fused_launch_fills_alike(which):
    fs, s = _case(which)
    init = fs.work.clone()
    fill_ghosts(fs, s)
    torch.cuda.synchronize()
    ref = fs.work.clone()
    plan = s.plan(fs.layout)
    for defer, cpt in itertools.product((0, 1), (4, 8, 16)):
        plan.set_options(merge=1, fuse=1, defer=defer, copy_per_thread=cpt)
        fs.work.copy_(init)
        fill_ghosts(fs, s)
        torch.cuda.synchronize()
        assert torch.equal(fs.work.view(torch.int64), ref.view(torch.int64))
