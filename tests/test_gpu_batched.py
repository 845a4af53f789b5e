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
