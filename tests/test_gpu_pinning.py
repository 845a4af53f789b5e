"""The pinning cases of tests/test_pinning_host.py through the HIP path: the
tap-order known answers, stage 0's primary copy, periodic = tiled bit for
bit, reflecting = doubled; in the dense layout and in the arena layout with
an even variable count (prolongation over variable pairs, 16-byte accesses).
"""
import numpy as np
import pytest

from test_pinning_host import (_host, check_primary, check_tap_kats, periodic_vs_tiled, reflecting_vs_doubled,
                               tap_case, vertex_case, _fill)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_tap_order_known_answers_gpu(layout):
    fs, s = tap_case(device="cuda", layout=layout)
    print(f"[pinning] tap order, {layout}", flush=True)
    _fill(fs, s)
    check_tap_kats(_host(fs), s)


@pytest.mark.parametrize("which", ["config2", "config2_small", "random3", "random12"])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_stage0_value_is_the_primary_copys_gpu(which, layout):
    print(f"[pinning] stage-0 primary, {which} {layout}", flush=True)
    check_primary(*vertex_case(which, "cuda", layout))


@pytest.mark.parametrize("M,N,G,ops,centering", [(1, 4, 1, (2, 2), "cell"), (2, 4, 1, (2, 2), "cell"),
                                                   (1, 8, 3, (6, 6), "cell"), (1, 8, 2, (4, 4), "vertex")])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_periodic_equals_tiled_gpu(M, N, G, ops, centering, layout):
    print(f"[pinning] periodic vs tiled M={M} N={N} G={G} {ops} {centering} {layout}", flush=True)
    p, t, pairs = periodic_vs_tiled(M, N=N, G=G, ops=ops, centering=centering, device="cuda", layout=layout)
    ap, at = _host(p), _host(t)
    assert not np.isnan(ap).any()
    for b, tb in pairs:
        assert np.array_equal(ap[b].view(np.int64), at[tb].view(np.int64)), b


@pytest.mark.parametrize("side,centering", [("lo", "cell"), ("hi", "cell"), ("lo", "vertex"), ("hi", "vertex")])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_reflecting_equals_doubled_gpu(side, centering, layout):
    print(f"[pinning] reflecting vs doubled {side} {centering} {layout}", flush=True)
    h, f, pairs = reflecting_vs_doubled(side, centering, device="cuda", layout=layout)
    ah, af = _host(h), _host(f)
    assert not np.isnan(ah).any()
    worst = max(np.abs(ah[b] - af[fb]).max() for b, fb in pairs)
    assert worst < 1e-13, worst


@pytest.mark.parametrize("which", ["config2", "config5_small", "random10"])
@pytest.mark.parametrize("layout", ["dense", "arena"])
def test_grouped_fill_in_any_order_within_its_groups_gpu(which, layout):
    """Plans of the schedule with every stage group's stages reversed and
    with every stage alone fill as the schedule's own plan does, from NaN
    ghosts, bit for bit and with no NaN left: the levels a plan launches
    together are independent."""
    import ctypes as C

    import torch

    from test_pinning_host import grouped_case, restaged
    from treeamr_hip._abi import check, lib
    from treeamr_hip.core import _stream_ptr, default_context
    print(f"[pinning] stage groups, {which} {layout}", flush=True)
    fs, s = grouped_case(which, device="cuda", layout=layout)
    init = fs.work.clone()
    _fill(fs, s)
    ref = fs.work.clone()
    ctx = default_context()
    fptr = C.c_void_p(fs.factors.data_ptr()) if fs.factors is not None else None
    for mode in ("reversed", "alone"):
        ptr, keep = restaged(s, mode)
        plan = C.c_void_p()
        check(lib().tah_plan_create(ctx._h, ptr, C.byref(fs.layout), C.byref(plan)))
        try:
            fs.work.copy_(init)
            check(lib().tah_fill(plan, C.c_void_p(fs.work.data_ptr()), fptr, _stream_ptr(fs)))
            torch.cuda.synchronize()
        finally:
            lib().tah_plan_destroy(plan)
        assert not torch.isnan(fs.array()).any(), mode  # (the arena's padding stays NaN)
        assert torch.equal(fs.work.view(torch.int64), ref.view(torch.int64)), mode
