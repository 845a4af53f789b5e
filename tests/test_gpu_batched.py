"""Batched fills of many independent forests (the 64-forest configuration,
smaller draws here, cell and vertex centering) and their hipGraph form: bit
for bit the per-forest fills, graph replays checked one by one with eager work
in between."""
import pytest
import torch

from treeamr_hip import FieldSet, FillGraph, GhostSchedule, fill_ghosts, fill_ghosts_batched
from treeamr_hip import workloads as W

pytestmark = pytest.mark.gpu


def _forests(seeds, max_blocks=300):
    out = []
    for sd in seeds:
        f, ops, G, nvars, centering = W.random_forest(sd, max_blocks=max_blocks)
        for k, nv in enumerate(nvars):
            layout = "arena" if (sd + k) % 2 else "dense"
            fs = FieldSet(f, nv, G=G, centering=centering, device="cuda", layout=layout, fill=float("nan"))
            W.fill_uniform(fs, 100 * sd + k)
            out.append((fs, GhostSchedule(fs, ops)))
    return out


def _reference(pairs):
    ref = []
    for fs, s in pairs:
        saved = fs.work.clone()
        fill_ghosts(fs, s)
        ref.append(fs.work.clone())
        fs.work.copy_(saved)
    torch.cuda.synchronize()
    return ref


def test_batched_equals_single_fills():
    pairs = _forests(range(2, 10))
    print(f"[batched] {len(pairs)} field sets", flush=True)
    ref = _reference(pairs)
    fill_ghosts_batched(pairs)
    torch.cuda.synchronize()
    for (fs, _), r in zip(pairs, ref):
        assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64))


@pytest.mark.parametrize("queue", ["all", "none", "mixed"])
def test_batched_queue_modes(queue):
    """The work-queue launch (every plan's option queue set), the per-class
    launches (none set) and a batch with one plan opted out (no queue for the
    batch): each bit for bit the single fills, two calls in a row (the queue's
    counters start again from zero), then again after the per-plan copy tile
    changes (a batch built anew)."""
    pairs = _forests(range(20, 26), max_blocks=250)
    ref = _reference(pairs)
    init = [fs.work.clone() for fs, _ in pairs]
    for i, (fs, s) in enumerate(pairs):
        on = queue == "all" or (queue == "mixed" and i != 2)
        s.plan(fs.layout).set_options(queue=int(on))
    print(f"[batched queue={queue}] {len(pairs)} field sets", flush=True)
    for rep in range(3):
        if rep == 2:
            for fs, s in pairs:
                s.plan(fs.layout).set_options(copy_per_thread=16 if s.plan(fs.layout).options["merge"] else 4)
        for (fs, _), i0 in zip(pairs, init):
            fs.work.copy_(i0)
        fill_ghosts_batched(pairs)
        torch.cuda.synchronize()
        for k, ((fs, _), r) in enumerate(zip(pairs, ref)):
            assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64)), f"call {rep}, field set {k}"
    for fs, s in pairs:
        s.plan(fs.layout).set_options(queue=0, copy_per_thread=8)


def test_graph_replays_equal_single_fills():
    pairs = _forests(range(10, 16))
    ref = _reference(pairs)
    init = [fs.work.clone() for fs, _ in pairs]
    graph = FillGraph(pairs)
    for rep in range(3):
        for (fs, _), i0 in zip(pairs, init):
            fs.work.copy_(i0)          # eager work between replays
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        for (fs, _), r in zip(pairs, ref):
            assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64)), f"replay {rep}"


def test_batched_after_schedules_rebuilt():
    """A batched fill, then its schedules dropped and others built on the same
    field sets (other operators; new plans may take the old plans' addresses),
    then a batched fill again: the cached batch of the old plans must not run.
    Also a change of launch options between two batched fills."""
    import gc

    from treeamr_hip import Operators
    pairs = _forests(range(2, 6), max_blocks=200)
    init = [fs.work.clone() for fs, _ in pairs]  # NaN ghosts
    fill_ghosts_batched(pairs)
    torch.cuda.synchronize()
    sets = [fs for fs, _ in pairs]
    del pairs
    gc.collect()
    rebuilt = [(fs, GhostSchedule(fs, Operators(prolongation=2, restriction=2))) for fs in sets]
    for fs, i0 in zip(sets, init):
        fs.work.copy_(i0)
    ref = _reference(rebuilt)
    fill_ghosts_batched(rebuilt)
    torch.cuda.synchronize()
    for (fs, _), r in zip(rebuilt, ref):
        assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64)), "stale batch after a rebuild"
    # options changed on one plan: the batch is built again from its new tiles
    fs0, s0 = rebuilt[0]
    s0.plan(fs0.layout).set_options(merge=0, copy_per_thread=4)
    for fs, i0 in zip(sets, init):
        fs.work.copy_(i0)
    fill_ghosts_batched(rebuilt)
    torch.cuda.synchronize()
    for (fs, _), r in zip(rebuilt, ref):
        assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64)), "stale batch after set_options"


def test_one_plan_two_streams():
    """Two field sets over one forest and layout share one plan (and its
    device Job). Filled back to back on two streams, each must see its own
    array: the second call waits for the first before it rewrites the Job."""
    from treeamr_hip import workloads as W2
    f, ops, kw, _ = W2.config5(nvars=4, roots=4, N=16)
    kw = dict(kw)
    nv = kw.pop("nvars")
    a = FieldSet(f, nv, device="cuda", layout="arena", fill=float("nan"), **kw)
    b = FieldSet(f, nv, device="cuda", layout="arena", fill=float("nan"), **kw)
    W2.fill_uniform(a, 1)
    W2.fill_uniform(b, 2)
    s = GhostSchedule(a, ops)
    ra, rb = _reference([(a, s), (b, s)])
    for rep in range(3):
        sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
        sa.wait_stream(torch.cuda.current_stream())
        sb.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(sa):
            fill_ghosts(a, s)
        with torch.cuda.stream(sb):
            fill_ghosts(b, s)
        torch.cuda.synchronize()
        assert torch.equal(a.work.view(torch.int64), ra.view(torch.int64)), f"rep {rep}: first array"
        assert torch.equal(b.work.view(torch.int64), rb.view(torch.int64)), f"rep {rep}: second array"


def test_tune_batched_picks_options_and_fills_alike(tmp_path):
    """tah_tune_batched sets one choice of fork, defer and copy_per_thread on
    every plan of the batch, the batched fill and its graph stay bit for bit
    the single fills, and a second call reads the choice from the cache."""
    from treeamr_hip import tune_batched
    pairs = _forests(range(2, 8))
    print(f"[tune batched] {len(pairs)} field sets", flush=True)
    ref = _reference(pairs)
    cache = tmp_path / "tune.txt"
    choice, ms = tune_batched(pairs, cache)
    assert ms is not None and ms > 0
    assert choice["copy_per_thread"] in (4, 8, 16) and choice["copy_order"] in (0, 1)
    for fs, s in pairs:
        o = s.plan(fs.layout).options
        assert (o["fork"], o["defer"], o["copy_per_thread"], o["copy_order"]) == (
            choice["fork"], choice["defer"], choice["copy_per_thread"], choice["copy_order"])
    inits = [fs.work.clone() for fs, _ in pairs]
    fill_ghosts_batched(pairs)
    torch.cuda.synchronize()
    for (fs, _), r in zip(pairs, ref):
        assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64))
    for (fs, _), i in zip(pairs, inits):
        fs.work.copy_(i)
    g = FillGraph(pairs)
    g.launch()
    torch.cuda.synchronize()
    for (fs, _), r in zip(pairs, ref):
        assert torch.equal(fs.work.view(torch.int64), r.view(torch.int64))
    # from the cache, after the plans were set otherwise
    for fs, s in pairs:
        s.plan(fs.layout).set_options(fork=1 - choice["fork"], copy_per_thread=4 if choice["copy_per_thread"] != 4 else 8,
                                      copy_order=1 - choice["copy_order"])
    again, ms2 = tune_batched(pairs, cache)
    assert ms2 is None and again == choice
    with pytest.raises(ValueError):
        tune_batched([])
