"""TEST INFRASTRUCTURE ONLY: ctypes access to liboracle.so, the CPU
restatement of the reference's ghost fill (oracle.c). Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
product package.

Functions:
  fill(desc, layout, work, factors, nthreads=0, skip=True)   run_phase!(::CPU) per stage
  fill_adjoint(desc, layout, adj, factors)                   CPU transpose (scatter form)
  write_counts(desc, layout)                                 writes per stored point
  regrid(desc, layout, old, new)                             the regrid transfer (src/regrid.jl)
  scatter(layout, first, extent, work, state)                scatter_kernel! (src/state.jl)
  gather(layout, first, extent, state, work)                 gather_kernel! (src/state.jl)
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        for name in ("oracle_fill", "oracle_fill_adjoint", "oracle_write_counts"):
            getattr(L, name).restype = C.c_int
        L.oracle_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.oracle_fill_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_write_counts.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_fill_boundary.restype = C.c_int
        L.oracle_fill_boundary.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_int]
        L.oracle_max_threads.restype = C.c_int
        L.oracle_regrid.restype = C.c_int
        L.oracle_regrid.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        for name in ("oracle_scatter", "oracle_gather"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return C.c_void_p(a.ctypes.data)
    return C.c_void_p(a.data_ptr())  # a CPU torch tensor


def fill(desc_ptr, layout, work, factors=None, nthreads=0, skip=True):
    rc = lib().oracle_fill(C.cast(desc_ptr, C.c_void_p), C.cast(C.pointer(layout), C.c_void_p),
                           _ptr(work), _ptr(factors), int(nthreads), 1 if skip else 0)
    if rc:
        raise RuntimeError(f"oracle_fill failed ({rc})")


def fill_boundary(desc_ptr, layout, work, factors, kind, values, fr=None, nthreads=0):
    """The fill with the outer-face hook between its stage groups (kind 0:
    values after phase 1; kind 1: relative, values = f(x_g), fr = f(x_r)),
    the values as contiguous (points, nvars) arrays."""
    rc = lib().oracle_fill_boundary(C.cast(desc_ptr, C.c_void_p), C.cast(C.pointer(layout), C.c_void_p),
                                    _ptr(work), _ptr(factors), int(kind), _ptr(values), _ptr(fr), int(nthreads))
    if rc:
        raise RuntimeError(f"oracle_fill_boundary failed ({rc})")


def fill_adjoint(desc_ptr, layout, adj, factors=None):
    rc = lib().oracle_fill_adjoint(C.cast(desc_ptr, C.c_void_p), C.cast(C.pointer(layout), C.c_void_p),
                                   _ptr(adj), _ptr(factors))
    if rc:
        raise RuntimeError(f"oracle_fill_adjoint failed ({rc})")


def write_counts(desc_ptr, layout):
    n = layout.nblocks * layout.stored[0] * layout.stored[1] * layout.stored[2]
    counts = np.zeros(n, np.int32)
    lib().oracle_write_counts(C.cast(desc_ptr, C.c_void_p), C.cast(C.pointer(layout), C.c_void_p),
                              _ptr(counts))
    return counts.reshape(layout.nblocks, layout.stored[2], layout.stored[1], layout.stored[0])


def max_threads():
    return lib().oracle_max_threads()


def _i32(seq):
    return (C.c_int32 * 3)(*seq)


def scatter(layout, first, extent, work, state):
    """work[own point] = state[n] for every state entry n, in place on host arrays."""
    lib().oracle_scatter(C.cast(C.pointer(layout), C.c_void_p), _i32(first), _i32(extent), _ptr(work),
                         _ptr(state))


def gather(layout, first, extent, state, work):
    """state[n] = work[own point] for every state entry n, in place on host arrays."""
    lib().oracle_gather(C.cast(C.pointer(layout), C.c_void_p), _i32(first), _i32(extent), _ptr(state),
                        _ptr(work))


def regrid(desc_ptr, layout, old, new, nthreads=0):
    """The regrid transfer of schedule desc_ptr from host array old into new
    (layout: the new array's; old has desc.nsources blocks of the same layout)."""
    rc = lib().oracle_regrid(C.cast(desc_ptr, C.c_void_p), C.cast(C.pointer(layout), C.c_void_p), _ptr(old),
                             _ptr(new), int(nthreads))
    if rc:
        raise RuntimeError(f"oracle_regrid failed ({rc})")
