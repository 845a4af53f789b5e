"""Ghost-fill benchmark: effective HBM GB/s of one full ghost fill on one GPU.

Workload (default): the large single forest -- 8^3 roots of 32^3 cells,
cell-centred, two refinement levels over 4 seeded spheres (about 12% of the
volume refined, 2:1 balanced, about 5.2k blocks), G=3, 20 Float64 variables
(Gaussian bumps plus 1e-4 noise), PointValue(6,6), reflecting z-low face with
alternating parities, outflow elsewhere.

A step is one full ghost fill of that field set, inputs resident in HBM.
Effective bytes per fill = 16 x (values written), the reference's accounting
(bench/copies.jl:115-125): one 8-byte read and one 8-byte write per value.

Reported: the graph-captured staged fill (headline), the staged fill launched
eagerly, the plain per-group fill, the adjoint fill, the 64-forest batch
(sequential plain fills against one graph-captured batched fill), the
roofline figure of the fill, and the restated reference CPU fill (the oracle)
on a bounded sample, on this host's cores.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
# The CPU baseline's threads stay on their cores (the reference pins its
# CPU fill's threads the same way); set before any OpenMP runtime starts.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# the CPUs this process may use, read before an OpenMP runtime binds the
# main thread to one of them
USABLE_CPUS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
sys.path.insert(0, os.path.join(ROOT, "treeamr.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "ghost-fill effective HBM GB/s (% of peak); batched vs plain fill time, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_fill(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    times = []
    t_all0 = time.perf_counter()
    for _ in range(steps):
        start.record(stream)
        fn()
        stop.record(stream)
        stop.synchronize()
        times.append(start.elapsed_time(stop))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_all0
    times.sort()
    return dict(min_ms=times[0], median_ms=times[len(times) // 2],
                mean_ms=sum(times) / len(times), wall_s=wall)


def cpu_baseline(sched_builder, seconds=12.0):
    """The oracle (the restated reference CPU fill, threaded by block owner as
    run_phase!(::CPU) is) on the same mesh with one of the 20 variables."""
    import numpy as np
    import oracle
    from treeamr_hip import FieldSet, GhostSchedule
    from treeamr_hip import workloads as W
    f, ops, kw = sched_builder(1)
    kw = dict(kw)
    nvars = kw.pop("nvars")
    fs = FieldSet(f, nvars, device="cpu", fill=0.0, **kw)
    W.fill_gaussians(fs, 5)
    sched = GhostSchedule(fs, ops)
    work = fs.work.numpy()
    fac = fs.factors.numpy().copy() if fs.factors is not None else None
    nthreads = oracle.max_threads()
    oracle.fill(sched._descptr, fs.layout, work, fac, nthreads=nthreads)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.fill(sched._descptr, fs.layout, work, fac, nthreads=nthreads)
        reps += 1
        if time.perf_counter() - t0 > seconds or reps >= 50:
            break
    dt = (time.perf_counter() - t0) / reps
    values = sched.points() * nvars
    return dict(value=16.0 * values / dt / 1e9, unit="GB/s", cores=nthreads, kind="port",
                sample=f"the same mesh with 1 of the 20 variables ({values} values per fill), "
                       f"{reps} fills, {dt * 1e3:.1f} ms per fill, OpenMP threads owning contiguous "
                       f"block ranges", host=host_cpu())


def host_cpu():
    """The host the CPU baseline runs on: CPU model, sockets and logical CPUs
    (from /proc/cpuinfo), and the CPUs this process may use."""
    model, sockets, logical = None, set(), 0
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                key, _, val = line.partition(":")
                key = key.strip()
                if key == "model name" and model is None:
                    model = val.strip()
                elif key == "physical id":
                    sockets.add(val.strip())
                elif key == "processor":
                    logical += 1
    except OSError:
        pass
    return dict(model=model, sockets=len(sockets) or None, logical_cpus=logical or os.cpu_count(),
                usable_cpus=USABLE_CPUS, omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                omp_places=os.environ.get("OMP_PLACES"))


def footprint(sched, nvars):
    """The reference's footprint figure (bench/nested.jl:67-73): the distinct
    points a fill reads -- the union, per source block, of every transfer's
    source windows -- and the points it writes (each ghost once), as
    (points read, points written, bytes = 8 * nvars * their sum)."""
    import numpy as np
    groups = sched.groups()
    nblocks = 1 + max((int(max(g["sourceblocks"].max(), g["targetblocks"].max()))
                       for g in groups if g["ntransfers"]), default=-1)
    ext = [0, 0, 0]
    for g in groups:
        for d, st in enumerate(g["stencils"]):
            if st["ntarget"]:
                ext[d] = max(ext[d], int(st["srcstart"].max()) + st["order"])
    read = np.zeros((max(nblocks, 1), ext[2], ext[1], ext[0]), dtype=bool)
    written = 0
    for g in groups:
        st = g["stencils"]
        n = [x["ntarget"] for x in st]
        if g["ntransfers"] == 0 or min(n) == 0:
            continue
        lo = [int(x["srcstart"].min()) for x in st]
        hi = [int(x["srcstart"].max()) + x["order"] for x in st]
        written += n[0] * n[1] * n[2] * g["ntransfers"]
        for b in np.unique(g["sourceblocks"]):
            read[b, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = True
    nread = int(read.sum())
    return nread, written, 8 * nvars * (nread + written)


def device_copy_rate(gib=2.0, reps=10):
    """GB/s (bytes read plus bytes written) of torch's device-to-device copy of a
    gib-GiB buffer on this GPU, median of reps: what a plain streaming copy
    reaches here, beside the fill's copy kernel."""
    import torch
    n = int(gib * (1 << 30)) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        y.copy_(x)
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    ms = sorted(ts)[len(ts) // 2]
    del x, y
    return round(2 * n * 8 / (ms * 1e-3) / 1e9, 1)


def traffic_of(kernel, rec, layout, nvars):
    """HBM bytes per launch of `kernel` from the committed PMC summary of this
    workload (profiles/hbm_traffic_large_forest_arena.json: FETCH_SIZE and WRITE_SIZE per dispatch,
    scaled against a copy of known size), when it covers this kernel at this
    layout and launch size; else None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_large_forest_arena.json")
    # (the adjoint's launches differ per stage: the record's mean per launch
    # is compared, as the summary averages its dispatches)
    try:
        with open(path) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    e = table.get("kernels", {}).get(kernel)
    if not e or table.get("layout") != layout or table.get("nvars") != nvars:
        return None
    if e.get("bytes_per_launch") != rec["bytes_per_launch"]:
        return None
    return e["hbm_bytes_per_launch"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nvars", type=int, default=20)
    ap.add_argument("--layout", default="arena", choices=["dense", "arena"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-config2", action="store_true")
    ap.add_argument("--no-reference-meshes", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-state", action="store_true")
    ap.add_argument("--no-regrid", action="store_true")
    ap.add_argument("--no-unify", action="store_true")
    ap.add_argument("--config2-repeats", type=int, default=100)
    ap.add_argument("--batched-steps", type=int, default=3)
    ap.add_argument("--only-batched", action="store_true",
                    help="run the 64-forest batch alone (for a kernel trace of it) and print its record")
    ap.add_argument("--batch-queue", type=int, default=0, choices=[0, 1],
                    help="the batch's plans' option queue: 1 one persistent work-queue launch per stage position")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="at the end, copy a buffer of this many GiB once (a kernel of known "
                         "traffic, to scale HBM counters by)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import torch
    from treeamr_hip import _abi
    from treeamr_hip import (FieldSet, FillGraph, GhostSchedule, fill_ghosts, fill_ghosts_adjoint,
                             fill_ghosts_adjoint_timed, fill_ghosts_timed, gather, ownextents, regrid_transfer,
                             scatter, statelength, tune, tune_batched, Operators, RegridSchedule, unify)
    from treeamr_hip import workloads as W

    def run_batched(stream):
        log("[bench] building the 64-forest batch")
        tb = time.perf_counter()
        pairs = W.config3(count=64, layout=args.layout)
        for f2, s2 in pairs:
            s2.plan(f2.layout).set_options(queue=args.batch_queue)
        bvalues = sum(s.points() * f2.nvars for f2, s in pairs)
        bblocks = sum({id(f2.forest): f2.nblocks for f2, _ in pairs}.values())
        log(f"[bench] batch: {len(pairs)} field sets, {bblocks} blocks, {bvalues} values; "
            f"{time.perf_counter() - tb:.1f}s")

        def seq_plain():
            for f2, s2 in pairs:
                fill_ghosts(f2, s2, method="plain")

        bplain = time_fill(seq_plain, args.batched_steps, 1, stream)
        log(f"[bench] batch, 64 sequential plain fills: {bplain}")
        btune = None
        if not args.no_tune:
            # the batch's launch options, picked by the library's batch tuner
            # before the timed region (as a warm-up would)
            choice, tms = tune_batched(pairs, None)
            btune = {"options": choice, "tuner_median_ms": round(tms, 4) if tms else None}
            log(f"[bench] batch tuner: {btune}")
        bgraph_obj = FillGraph(pairs)
        bgraph = time_fill(bgraph_obj.launch, args.batched_steps * 3, 1, stream)
        log(f"[bench] batch, graph-captured batched fill: {bgraph}")
        rec = {
            "workload": "64 independent forests, seeds 2..65 (1^3-4^3 roots, N in {8,16,32}, 1-4 levels, "
                        "G in {2,3}, cell or vertex, 1-3 field sets of 2-10 vars, PointValue(2G), periodic, "
                        "<=2 GiB each)",
            "field_sets": len(pairs), "blocks": bblocks, "values_per_fill": bvalues,
            "queue": args.batch_queue,
            "tuned": btune,
            "plain_sequential_ms": round(bplain["mean_ms"], 3),
            "graph_batched_ms": round(bgraph["mean_ms"], 3),
            "speedup": round(bplain["mean_ms"] / bgraph["mean_ms"], 2),
            "graph_effective_GBs": round(16.0 * bvalues / (bgraph["mean_ms"] * 1e-3) / 1e9, 1),
            "graph_pct_of_peak": round(100.0 * 16.0 * bvalues / (bgraph["mean_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
            "plain_effective_GBs": round(16.0 * bvalues / (bplain["mean_ms"] * 1e-3) / 1e9, 1),
        }
        del bgraph_obj, pairs
        return rec

    if args.only_batched:
        rec = run_batched(torch.cuda.current_stream())
        print(json.dumps({"batched": rec}), flush=True)
        return

    def builder(nvars):
        f, ops, kw, _ = W.config5(nvars=nvars)
        return f, ops, kw

    log("[bench] building mesh and schedule")
    t0 = time.perf_counter()
    f, ops, kw = builder(args.nvars)
    kw = dict(kw)
    nvars = kw.pop("nvars")
    fs = FieldSet(f, nvars, device="cuda", layout=args.layout, fill=0.0, **kw)
    W.fill_gaussians(fs, 5)
    torch.cuda.synchronize()
    # the host schedule alone, then its device plan (classes, tiles, upload)
    sched_t0 = time.perf_counter()
    sched = GhostSchedule(fs, ops)
    build_s = time.perf_counter() - sched_t0
    plan_t0 = time.perf_counter()
    sched.plan(fs.layout)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - plan_t0
    values = sched.points() * nvars
    bytes_eff = 16.0 * values
    log(f"[bench] {f.nleaves} blocks, {values} values/fill, {bytes_eff / 1e9:.2f} GB effective; "
        f"setup {time.perf_counter() - t0:.1f}s")
    stream = torch.cuda.current_stream()

    def gbs_of(ms):
        return bytes_eff / (ms * 1e-3) / 1e9

    plain = None
    plain_dense = None
    if not args.no_plain:
        log("[bench] timing the plain per-group fill")
        plain = time_fill(lambda: fill_ghosts(fs, sched, method="plain"), args.steps, args.warmup, stream)
        log(f"[bench] plain: {plain}")
        if args.layout != "dense":
            # the plain fill on the reference's own layout (work[i, j, k, var,
            # block], i fastest), where its i-fastest threads read contiguous rows
            log("[bench] timing the plain per-group fill on the dense layout")
            fsd = FieldSet(f, nvars, device="cuda", layout="dense", fill=0.0, **kw)
            W.fill_gaussians(fsd, 5)
            plain_dense = time_fill(lambda: fill_ghosts(fsd, sched, method="plain"), args.steps, args.warmup,
                                    stream)
            log(f"[bench] plain (dense layout): {plain_dense}")
            sched._plans.pop((tuple(fsd.layout.stride), fsd.layout.nvars, 0), None)
            del fsd
    log("[bench] timing the staged fill")
    staged = time_fill(lambda: fill_ghosts(fs, sched, method="staged"), args.steps, args.warmup, stream)
    log(f"[bench] staged: {staged}")
    log("[bench] timing the graph-captured fill")
    graph = FillGraph([(fs, sched)])
    graphed = time_fill(graph.launch, args.steps, args.warmup, stream)
    log(f"[bench] graph: {graphed}")
    log("[bench] timing every launch of the fill (in order, HIP events on the fill's stream)")
    kernels = {}
    for it in range(args.warmup + args.steps):
        recs = fill_ghosts_timed(fs, sched)
        if it < args.warmup:
            continue
        for name, vals, ms in recs:
            k = kernels.setdefault(name, dict(launches=0, values=0, ms=0.0))
            k["launches"] += 1
            k["values"] += vals
            k["ms"] += ms
    for k in kernels.values():
        k["avg_launch_ms"] = round(k["ms"] / k["launches"], 4)
        k["bytes_per_launch"] = 16 * k["values"] // k["launches"]
        k["GBs"] = round(16.0 * k["values"] / (k["ms"] * 1e-3) / 1e9, 1)
        k["ms_per_fill"] = round(k["ms"] / args.steps, 4)
        k["ms"] = round(k["ms"], 3)
    dominant = max(kernels, key=lambda n: kernels[n]["ms"])
    dk = kernels[dominant]
    log(f"[bench] dominant kernel {dominant}: {dk}")
    tuned = None
    tg = None
    if not args.no_tune:
        # The autotuner picks the plan's launch options for this mesh (before
        # the timed region, as a warm-up would); the headline is the fill
        # with its choice.
        log("[bench] autotuning the plan's launch options (tah_tune, no cache)")
        opts, tune_ms = tune(fs, sched, None)
        tgraph = FillGraph([(fs, sched)])
        tg = time_fill(tgraph.launch, args.steps, args.warmup, stream)
        tuned = {"options": opts, "tuner_median_ms": round(tune_ms, 4),
                 "graph_ms": {k: round(v, 4) for k, v in tg.items()},
                 "GBs": round(gbs_of(tg["mean_ms"]), 1)}
        log(f"[bench] tuned: {tuned}")
        del tgraph
        sched.plan(fs.layout).set_options(fork=0, merge=1, defer=1, copy_per_thread=8, copy_wgs=0)
    adjoint = None
    adj_kernels = {}
    if not args.no_adjoint:
        log("[bench] timing the adjoint fill")
        # The adjoint runs in place on the working array's own memory (the
        # forward timings are done; random values replace the field), so
        # that forward and adjoint share one physical placement: on one box
        # the same kernels' HBM rate moved by 10-15% between fresh
        # allocations of the same size, the copies between 5.9 and 6.7 ms.
        adj = fs.work
        adj.uniform_()
        adjoint = time_fill(lambda: fill_ghosts_adjoint(fs, sched, adj), args.steps, args.warmup, stream)
        log(f"[bench] adjoint: {adjoint}")
        log("[bench] timing every launch of the adjoint (in order, HIP events on its stream)")
        for it in range(args.warmup + args.steps):
            recs = fill_ghosts_adjoint_timed(fs, sched, adj)
            if it < args.warmup:
                continue
            for name, nbytes, ms in recs:
                k = adj_kernels.setdefault(name, dict(launches=0, bytes=0, ms=0.0))
                k["launches"] += 1
                k["bytes"] += nbytes
                k["ms"] += ms
        for k in adj_kernels.values():
            k["avg_launch_ms"] = round(k["ms"] / k["launches"], 4)
            k["bytes_per_launch"] = k["bytes"] // k["launches"]
            k["GBs"] = round(k["bytes"] / (k["ms"] * 1e-3) / 1e9, 1)
            k["ms_per_fill"] = round(k["ms"] / args.steps, 4)
            k["ms"] = round(k["ms"], 3)
        log(f"[bench] adjoint kernels: {adj_kernels}")
        del adj
    head = tg if tg is not None else graphed
    t_ms = head["mean_ms"]
    gbs = gbs_of(t_ms)

    batched = None
    if not args.no_batched:
        batched = run_batched(stream)

    nested = None
    if not args.no_config2:
        log("[bench] the nested vertex-centred mesh (config 2)")
        f2, ops2, kw2, _ = W.config2()
        kw2 = dict(kw2)
        nv2 = kw2.pop("nvars")
        fs2 = FieldSet(f2, nv2, device="cuda", layout="dense", fill=0.0, **kw2)
        W.fill_waves(fs2, 1)
        s2 = GhostSchedule(fs2, ops2)
        v2 = s2.points() * nv2
        reps = args.config2_repeats
        p2 = time_fill(lambda: fill_ghosts(fs2, s2, method="plain"), reps, args.warmup, stream)
        st2 = time_fill(lambda: fill_ghosts(fs2, s2, method="staged"), reps, args.warmup, stream)
        g2obj = FillGraph([(fs2, s2)])
        g2 = time_fill(g2obj.launch, reps, args.warmup, stream)
        c2 = None
        if not args.no_cpu:
            import oracle
            host = fs2.work.cpu().numpy().copy()
            fac = fs2.factors.cpu().numpy().copy()
            nth = oracle.max_threads()
            oracle.fill(s2._descptr, fs2.layout, host, fac, nthreads=nth)
            times = []
            for _ in range(20):
                t0 = time.perf_counter()
                oracle.fill(s2._descptr, fs2.layout, host, fac, nthreads=nth)
                times.append(time.perf_counter() - t0)
            times.sort()
            c2 = dict(min_ms=round(times[0] * 1e3, 3), median_ms=round(times[10] * 1e3, 3), threads=nth)
        nested = {
            "workload": "config 2: 4^3 roots, N=16, vertex, G=3, 8 vars, PointValue(6,6), nested "
                        "refinement at radii 0.35/0.2/0.1 (232 blocks), reflecting z-lo (vars 7-8 odd), "
                        "outflow elsewhere; dense layout",
            "blocks": f2.nleaves, "values_per_fill": v2, "repeats": reps,
            "plain_ms": {"min": round(p2["min_ms"], 4), "median": round(p2["median_ms"], 4)},
            "staged_ms": {"min": round(st2["min_ms"], 4), "median": round(st2["median_ms"], 4)},
            "graph_ms": {"min": round(g2["min_ms"], 4), "median": round(g2["median_ms"], 4)},
            "graph_GBs_at_min": round(16.0 * v2 / (g2["min_ms"] * 1e-3) / 1e9, 1),
            "cpu_oracle_ms": c2,
        }
        log(f"[bench] config 2: {nested}")
        del g2obj, fs2, s2

    refmeshes = None
    if not args.no_reference_meshes:
        refmeshes = []
        cases = [
            ("uniform 8^3 roots of 16^3, periodic, vertex, G=3, 20 vars, PointValue(4) (bench/copies.jl)",
             lambda: W.uniform_vertex(N=16), 1.716, "CODE.md:2992-2997"),
            ("uniform 8^3 roots of 32^3, periodic, vertex, G=3, 20 vars, PointValue(4) (bench/copies.jl)",
             lambda: W.uniform_vertex(N=32), 4.969, "CODE.md:2992-2997"),
            ("nested reflecting octant, 232 blocks of 32^3, vertex, G=3, 20 vars, PointValue(6) "
             "(bench/nested.jl; outer-face hook regions not written here)",
             lambda: W.nested_octant(N=32), 7.90, "CODE.md:2752-2760"),
            ("two-level, 960 blocks of 32^3, periodic, cell, G=2, 2 vars, PointValue(4) (bench/gpu.jl)",
             lambda: W.two_level(N=32), 2.8, "CODE.md:5532-5548"),
        ]
        for name, make, h200_ms, src in cases:
            log(f"[bench] reference mesh: {name}")
            fR, opsR, kwR = make()
            kwR = dict(kwR)
            nvR = kwR.pop("nvars")
            rec = {"workload": name, "blocks": fR.nleaves, "h200_ms": h200_ms, "h200_source": src}
            for layout in ("arena", "dense"):
                fsR = FieldSet(fR, nvR, device="cuda", layout=layout, fill=0.0, **kwR)
                W.fill_uniform(fsR, 3)
                sR = GhostSchedule(fsR, opsR)
                vR = sR.points() * nvR
                gR = FillGraph([(fsR, sR)])
                tR = time_fill(gR.launch, 20, args.warmup, stream)
                rec["values_per_fill"] = vR
                rec[f"{layout}_graph_ms"] = {"min": round(tR["min_ms"], 4), "median": round(tR["median_ms"], 4)}
                rec[f"{layout}_GBs_at_min"] = round(16.0 * vR / (tR["min_ms"] * 1e-3) / 1e9, 1)
                del gR, fsR, sR
            rec["speedup_vs_h200_best_of"] = round(h200_ms / min(rec["arena_graph_ms"]["min"],
                                                                  rec["dense_graph_ms"]["min"]), 2)
            log(f"[bench] {rec}")
            refmeshes.append(rec)

    state = None
    if not args.no_state:
        # scatter!/gather! per owned point on the reference's copy meshes
        # (bench/copies.jl), beside a linear copy of the state, the floor.
        state = []
        h200 = {16: (0.182, 0.147, 0.100), 32: (0.140, 0.136, 0.093), 128: (0.116, 0.134, 0.093)}
        for N, roots in ((16, 8), (32, 8), (128, 2)):
            fS, _, kwS = W.uniform_vertex(N=N, roots=roots)
            kwS = dict(kwS)
            nvS = kwS.pop("nvars")
            rec = {"workload": f"uniform {roots}^3 roots of {N}^3, vertex, G=3, {nvS} vars (bench/copies.jl)",
                   "blocks": fS.nleaves,
                   "h200_ns_per_point": {"scatter": h200[N][0], "gather": h200[N][1], "floor": h200[N][2]},
                   "h200_source": "CODE.md:2942-2953"}
            for layout in ("arena", "dense"):
                fsS = FieldSet(fS, nvS, device="cuda", layout=layout, fill=0.0, **kwS)
                e = ownextents(fsS)
                points = fS.nleaves * e[0] * e[1] * e[2]
                u = torch.rand(statelength(fsS), device="cuda", dtype=torch.float64)
                tsc = time_fill(lambda: scatter(fsS, u), 20, args.warmup, stream)
                tga = time_fill(lambda: gather(u, fsS), 20, args.warmup, stream)
                ns = lambda t: round(t["min_ms"] * 1e6 / points, 4)
                gb = lambda t: round(16.0 * u.numel() / (t["min_ms"] * 1e-3) / 1e9, 1)
                rec[layout] = {"scatter_ns_per_point": ns(tsc), "gather_ns_per_point": ns(tga),
                               "scatter_GBs": gb(tsc), "gather_GBs": gb(tga)}
                if layout == "arena":
                    v = torch.empty_like(u)
                    tfl = time_fill(lambda: v.copy_(u), 20, args.warmup, stream)
                    rec["floor_ns_per_point"] = ns(tfl)
                    rec["floor_GBs"] = gb(tfl)
                    del v
                del fsS, u
            rec["points"] = points
            log(f"[bench] state copies: {rec}")
            state.append(rec)

    regrid_rec = None
    if not args.no_regrid:
        # The reference's regrid-transfer benchmark (bench/gpu.jl:159-174):
        # old leaves = new leaves (every block copied into a fresh array),
        # on its two-level mesh; scatter! on the same mesh beside it. H200
        # times from CODE.md:5538-5546.
        fT, opsT, kwT = W.two_level(N=32)
        kwT = dict(kwT)
        nvT = kwT.pop("nvars")
        regrid_rec = {"workload": "two-level, 960 blocks of 32^3, cell, G=2, 2 vars (bench/gpu.jl), "
                                  "old leaves = new leaves",
                      "blocks": fT.nleaves, "h200_ms": {"regrid_transfer": 1.3, "scatter": 1.0},
                      "h200_source": "CODE.md:5538-5546"}
        for layout in ("arena", "dense"):
            fsT = FieldSet(fT, nvT, device="cuda", layout=layout, fill=0.0, **kwT)
            W.fill_uniform(fsT, 6)
            rsT = RegridSchedule(fT, fT.leaf_arrays(), opsT, G=fsT.G, centering=fsT.centering, nvars=nvT)
            fresh, planT = regrid_transfer(fsT, fsT.layout, fsT.work, rsT)
            lay = fsT.layout

            def move():
                _abi.check(_abi.lib().tah_regrid_transfer(planT._h, ctypes.byref(lay),
                                                          ctypes.c_void_p(fsT.work.data_ptr()),
                                                          ctypes.c_void_p(fresh.data_ptr()),
                                                          ctypes.c_void_p(stream.cuda_stream)))
            tT = time_fill(move, 20, args.warmup, stream)
            vals = statelength(fsT)
            u = torch.rand(vals, device="cuda", dtype=torch.float64)
            tS = time_fill(lambda: scatter(fsT, u), 20, args.warmup, stream)
            regrid_rec[layout] = {
                "regrid_transfer_ms": round(tT["min_ms"], 4),
                "regrid_transfer_GBs": round(16.0 * vals / (tT["min_ms"] * 1e-3) / 1e9, 1),
                "scatter_ms": round(tS["min_ms"], 4),
                "scatter_GBs": round(16.0 * vals / (tS["min_ms"] * 1e-3) / 1e9, 1)}
            del fsT, rsT, fresh, planT, u
        regrid_rec["speedup_vs_h200"] = {
            "regrid_transfer": round(1.3 / min(regrid_rec[l]["regrid_transfer_ms"] for l in ("arena", "dense")), 2),
            "scatter": round(1.0 / min(regrid_rec[l]["scatter_ms"] for l in ("arena", "dense")), 2)}
        log(f"[bench] regrid transfer: {regrid_rec}")

    unify_rec = None
    if not args.no_unify:
        # unify! alone, as bench/rhs.jl times it (unify!(du, fs, schedule) on a
        # state-shaped du; the reference's uniform mesh, N=16, vertex, G=3,
        # 20 variables, order-6 operators), and on the working array. The
        # H200 figure is derived: unify! alone is 10-23% of the 0.540 ns per
        # own point of the every-point right-hand side (CODE.md:1296-1308).
        fU, _, kwU = W.uniform_vertex(N=16, roots=8)
        kwU = dict(kwU)
        nvU = kwU.pop("nvars")
        opsU = Operators(prolongation=6, restriction=6)
        unify_rec = {"workload": "uniform 8^3 roots of 16^3, vertex, G=3, 20 vars, PointValue(6,6) (bench/rhs.jl)",
                     "h200_ns_per_point_derived": [round(0.10 * 0.540, 3), round(0.23 * 0.540, 3)],
                     "h200_source": "CODE.md:1296-1308 (10-23% of 0.540 ns per point)"}
        for layout in ("arena", "dense"):
            fsU = FieldSet(fU, nvU, device="cuda", layout=layout, fill=0.0, **kwU)
            W.fill_uniform(fsU, 9)
            sU = GhostSchedule(fsU, opsU)
            e = ownextents(fsU)
            pts = fU.nleaves * e[0] * e[1] * e[2]
            du = torch.rand(statelength(fsU), device="cuda", dtype=torch.float64)
            tv = time_fill(lambda: unify(du, fsU, sU), 20, args.warmup, stream)
            tw = time_fill(lambda: unify(fsU, sU), 20, args.warmup, stream)
            unify_rec[layout] = {"state_ns_per_point": round(tv["min_ms"] * 1e6 / pts, 4),
                                 "work_ns_per_point": round(tw["min_ms"] * 1e6 / pts, 4),
                                 "state_ms": round(tv["min_ms"], 4), "work_ms": round(tw["min_ms"], 4)}
            del fsU, sU, du
        log(f"[bench] unify: {unify_rec}")

    cpu = None
    if not args.no_cpu:
        log("[bench] timing the CPU baseline (oracle)")
        cpu = cpu_baseline(builder)
        log(f"[bench] cpu: {cpu}")

    def r(d):
        return {k: round(v, 4) for k, v in d.items()} if d else None

    def adjoint_roofline(ks):
        # the adjoint's dominant kernel against HBM, as the fill's below
        if not ks:
            return None
        name = max(ks, key=lambda n: ks[n]["ms"])
        k = ks[name]
        return {"kernel": name, "bound": "hbm", "achieved": k["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(k["GBs"] / HBM_PEAK_GBS, 4),
                "traffic": traffic_of(name, k, args.layout, nvars),
                "algorithmic_bytes_per_launch": k["bytes_per_launch"], "avg_launch_ms": k["avg_launch_ms"]}

    copy_ref = device_copy_rate()
    log(f"[bench] device copy of 2 GiB: {copy_ref} GB/s")
    fp_read, fp_written, fp_bytes = footprint(sched, nvars)
    fp = {"points_read": fp_read, "points_written": fp_written, "bytes": fp_bytes,
          "GBs": round(fp_bytes / (t_ms * 1e-3) / 1e9, 1)}
    log(f"[bench] footprint: {fp}")

    out = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian bumps + 1e-4 noise, seed 5)",
        "config": {
            "workload": "large single forest: 8^3 roots, N=32, cell, G=3, 20 vars, PointValue(6,6), "
                        "2 levels over 4 spheres (~12% refined), reflecting z-lo + outflow",
            "blocks": f.nleaves,
            "values_per_fill": values,
            "layout": args.layout,
            "fill": "graph-captured staged fill",
            "plan_options": tuned["options"] if tuned else "defaults (fork 0, merge 1, defer 1, copy tiles of 8)",
        },
        "pct_of_peak": round(100.0 * gbs / HBM_PEAK_GBS, 2),
        "footprint": fp,
        "graph_ms": r(head),
        "graph_ms_default_options": r(graphed),
        "staged_ms": r(staged),
        "plain_ms": r(plain),
        "plain_dense_layout_ms": r(plain_dense),
        "plain_GBs": round(gbs_of(plain["mean_ms"]), 1) if plain else None,
        "plain_pct_of_peak": round(100.0 * gbs_of(plain["mean_ms"]) / HBM_PEAK_GBS, 2) if plain else None,
        "speedup_vs_plain": round(plain["mean_ms"] / t_ms, 3) if plain else None,
        "speedup_vs_plain_dense_layout": round(plain_dense["mean_ms"] / t_ms, 3) if plain_dense else None,
        "tuned": tuned,
        "adjoint_ms": r(adjoint),
        "adjoint_over_forward": round(adjoint["mean_ms"] / t_ms, 3) if adjoint else None,
        "adjoint_effective_GBs": round(gbs_of(adjoint["mean_ms"]), 1) if adjoint else None,
        "adjoint_pct_of_peak": round(100.0 * gbs_of(adjoint["mean_ms"]) / HBM_PEAK_GBS, 2) if adjoint else None,
        "adjoint_roofline": adjoint_roofline(adj_kernels),
        "adjoint_kernels": adj_kernels or None,
        "batched": batched,
        "config2": nested,
        "reference_meshes": refmeshes,
        "state_copies": state,
        "regrid_transfer": regrid_rec,
        "unify": unify_rec,
        "schedule_build_s": round(build_s, 3),
        "plan_build_s": round(plan_s, 3),
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": dk["GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dk["GBs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic_of(dominant, dk, args.layout, nvars),
            "algorithmic_bytes_per_launch": dk["bytes_per_launch"],
            "avg_launch_ms": dk["avg_launch_ms"],
            "device_copy_GBs": copy_ref,
            "frac_of_device_copy": round(dk["GBs"] / copy_ref, 4),
        },
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.copy_probe_gib > 0:
        del fs, graph
        torch.cuda.synchronize()
        n = int(args.copy_probe_gib * (1 << 30)) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {n * 8} bytes read, {n * 8} bytes written")


if __name__ == "__main__":
    main()
