"""bench.py's command line on the CPU: asked for more than one GPU it stops
with a message that says this build measures one GPU, before touching a
device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout.strip() == ""          # no JSON line


def _brute_footprint(sched):
    """Every target's own window marked point by point."""
    import numpy as np
    groups = sched.groups()
    marks = set()
    written = 0
    for g in groups:
        st = g["stencils"]
        n = [x["ntarget"] for x in st]
        if g["ntransfers"] == 0 or min(n) == 0:
            continue
        written += n[0] * n[1] * n[2] * g["ntransfers"]
        for b in g["sourceblocks"]:
            for k in range(n[2]):
                for j in range(n[1]):
                    for i in range(n[0]):
                        s = [int(st[0]["srcstart"][i]), int(st[1]["srcstart"][j]), int(st[2]["srcstart"][k])]
                        for z in range(s[2], s[2] + st[2]["order"]):
                            for y in range(s[1], s[1] + st[1]["order"]):
                                for x in range(s[0], s[0] + st[0]["order"]):
                                    marks.add((int(b), z, y, x))
    return len(marks), written


def test_footprint_counts_every_window_point_once():
    """bench.footprint's union of source boxes equals the union of every
    target's window, marked point by point, on the periodic two-level cell
    mesh (Conservative(3,2)) and the nested vertex mesh (PointValue(6,6));
    points written = the schedule's points."""
    sys.path.insert(0, os.path.join(ROOT, "treeamr.jl_amd"))
    sys.path.insert(0, ROOT)
    import bench
    from treeamr_hip import FieldSet, GhostSchedule
    from treeamr_hip import workloads as W
    for f, ops, kw, _ in (W.config1(), W.config2(nvars=2, roots=2, N=8, radii=(0.35,))):
        kw = dict(kw)
        nvars = kw.pop("nvars")
        fs = FieldSet(f, nvars, device="cpu", fill=0.0, **kw)
        s = GhostSchedule(fs, ops)
        nread, written, nbytes = bench.footprint(s, nvars)
        bread, bwritten = _brute_footprint(s)
        assert (nread, written) == (bread, bwritten)
        assert written == s.points()
        assert nbytes == 8 * nvars * (nread + written)
