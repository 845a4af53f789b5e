"""ctypes declarations of libtah (include/tah.h).

The library is built in-tree by ``make -C treeamr.jl_amd`` into
``treeamr.jl_amd/lib/libtah.so``; loading fails loudly when it is missing.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(_HERE), "lib", "libtah.so")

TAH_OK, TAH_EINVAL, TAH_ESTALE, TAH_ELAYOUT, TAH_EHIP, TAH_ENOMEM, TAH_EUNSUPPORTED = range(7)
TAH_HOOK_VALUES, TAH_HOOK_RELATIVE = 0, 1


class tah_operators(C.Structure):
    _fields_ = [("family", C.c_int32), ("prolongation", C.c_int32), ("restriction", C.c_int32),
                ("bounded", C.c_int32), ("vertex_prolongation", C.c_int32)]


class tah_stencil1d(C.Structure):
    _fields_ = [("targetfirst", C.c_int32), ("ntarget", C.c_int32), ("order", C.c_int32),
                ("unit", C.c_int32), ("srcstart", C.POINTER(C.c_int32)),
                ("weights", C.POINTER(C.c_double))]


class tah_group_desc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("stage", C.c_int32), ("factorcol", C.c_int32),
                ("ntransfers", C.c_int32), ("targetblocks", C.POINTER(C.c_int32)),
                ("sourceblocks", C.POINTER(C.c_int32)), ("stencils", tah_stencil1d * 3)]


class tah_boundary_region(C.Structure):
    _fields_ = [("block", C.c_int32), ("direction", C.c_int32 * 3), ("lo", C.c_int32 * 3),
                ("hi", C.c_int32 * 3), ("clamps", C.c_int32 * 3), ("slot", C.c_int32), ("first", C.c_int64)]


class tah_schedule_desc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("nblocks", C.c_int32), ("stored", C.c_int32 * 3),
                ("nstages", C.c_int32), ("stage_tags", C.POINTER(C.c_int32)),
                ("nstagegroups", C.c_int32), ("stagegroup_first", C.POINTER(C.c_int32)),
                ("stagegroup_last", C.POINTER(C.c_int32)), ("ngroups", C.c_int32),
                ("groups", C.POINTER(tah_group_desc)), ("generation", C.c_int64),
                ("nsources", C.c_int32), ("pad_", C.c_int32), ("nboundaries", C.c_int32),
                ("pad2_", C.c_int32), ("boundaries", C.POINTER(tah_boundary_region)),
                ("boundary_points", C.c_int64)]


class tah_layout(C.Structure):
    _fields_ = [("stored", C.c_int32 * 3), ("nvars", C.c_int32), ("nblocks", C.c_int32),
                ("pad_", C.c_int32), ("stride", C.c_int64 * 5)]


P = C.c_void_p
I32P = C.POINTER(C.c_int32)
DP = C.POINTER(C.c_double)

class tah_family_options(C.Structure):
    _fields_ = [("order", C.c_int32), ("restriction", C.c_int32), ("threads", C.c_int32), ("merge", C.c_int32),
                ("pairs", C.c_int32), ("lds", C.c_int32), ("vars", C.c_int32)]


class tah_plan_options(C.Structure):
    _fields_ = [("fork", C.c_int32), ("merge", C.c_int32), ("defer", C.c_int32),
                ("copy_per_thread", C.c_int32), ("fuse", C.c_int32), ("copy_wgs", C.c_int32),
                ("interp_pairs", C.c_int32), ("lds", C.c_int32), ("queue", C.c_int32),
                ("copy_order", C.c_int32), ("reserved_", C.c_int32 * 2)]


# (name, argtypes) for every entry point of include/tah.h.
SIGNATURES = {
    "tah_lagrange_weights": [C.c_int32, C.c_int64, C.c_int64, DP],
    "tah_conservative_weights": [C.c_int32, DP, DP],
    "tah_forest_create": [C.c_int32, I32P, C.c_int32, I32P, I32P, I32P, DP, C.POINTER(P)],
    "tah_forest_destroy": [P],
    "tah_forest_info": [P, I32P, C.POINTER(C.c_int64)],
    "tah_forest_leaves": [P, I32P, I32P, C.POINTER(C.c_uint32)],
    "tah_forest_refine": [P, C.c_int32, I32P],
    "tah_forest_balance": [P],
    "tah_forest_isbalanced": [P, I32P],
    "tah_operators_check": [C.POINTER(tah_operators)],
    "tah_schedule_create": [P, C.POINTER(tah_operators), I32P, I32P, C.c_int32,
                            C.POINTER(C.c_int8), C.POINTER(P)],
    "tah_schedule_destroy": [P],
    "tah_schedule_check": [P],
    "tah_parity_factors": [C.c_int32, C.c_int32, C.POINTER(C.c_int8), DP],
    "tah_schedule_desc_get": [P, C.POINTER(C.POINTER(tah_schedule_desc))],
    "tah_schedule_boundaries": [P, I32P, I32P, I32P, I32P, I32P],
    "tah_schedule_points": [P, C.POINTER(C.c_int64)],
    "tah_layout_dense": [C.POINTER(tah_schedule_desc), C.c_int32, C.POINTER(tah_layout)],
    "tah_layout_arena": [C.POINTER(tah_schedule_desc), C.c_int32, C.POINTER(tah_layout)],
    "tah_layout_size": [C.POINTER(tah_layout), C.POINTER(C.c_int64)],
    "tah_ctx_create": [C.c_int32, C.POINTER(P)],
    "tah_ctx_destroy": [P],
    "tah_plan_create": [P, C.POINTER(tah_schedule_desc), C.POINTER(tah_layout), C.POINTER(P)],
    "tah_plan_destroy": [P],
    "tah_plan_values": [P, C.POINTER(C.c_int64)],
    "tah_fill_plain": [P, P, P, P],
    "tah_fill": [P, P, P, P],
    "tah_fill_boundary": [P, P, P, C.c_int32, P, P, P],
    "tah_fill_batched": [P, C.c_int32, C.POINTER(P), C.POINTER(P), C.POINTER(P), P],
    "tah_graph_create": [P, C.c_int32, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)],
    "tah_graph_launch": [P, P],
    "tah_graph_destroy": [P],
    "tah_fill_timed": [P, P, P, P, I32P, I32P, C.POINTER(C.c_int64), C.POINTER(C.c_float)],
    "tah_kernel_class_name": [C.c_int32, C.c_char_p, C.c_int32],
    "tah_fill_adjoint": [P, P, P, P],
    "tah_fill_adjoint_noalloc": [P, P, P, P],
    "tah_fill_adjoint_timed": [P, P, P, P, I32P, I32P, C.POINTER(C.c_int64), C.POINTER(C.c_float)],
    "tah_plan_get_options": [P, C.POINTER(tah_plan_options)],
    "tah_plan_set_options": [P, C.POINTER(tah_plan_options)],
    "tah_plan_get_family_options": [P, I32P, C.POINTER(tah_family_options)],
    "tah_plan_set_family_options": [P, C.POINTER(tah_family_options)],
    "tah_tune": [P, P, P, C.c_char_p, C.POINTER(C.c_double)],
    "tah_tune_batched": [P, C.c_int32, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.c_char_p,
                         C.POINTER(C.c_double)],
    "tah_state_length": [C.POINTER(tah_layout), I32P, C.POINTER(C.c_int64)],
    "tah_forest_coarsen": [P, C.c_int32, I32P, I32P],
    "tah_forest_positions": [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, DP],
    "tah_unify_schedule_create": [P, I32P, I32P, C.c_int32, C.POINTER(C.c_int8), C.POINTER(P)],
    "tah_unify_desc_get": [P, C.c_int32, C.POINTER(C.POINTER(tah_schedule_desc))],
    "tah_regrid_schedule_create": [P, C.c_int32, I32P, I32P, C.POINTER(C.c_uint32), C.POINTER(tah_operators),
                                   I32P, I32P, C.c_int32, C.POINTER(C.c_int8), C.POINTER(P)],
    "tah_regrid_transfer": [P, C.POINTER(tah_layout), P, P, P],
    "tah_scatter": [C.POINTER(tah_layout), I32P, I32P, P, P, P],
    "tah_gather": [C.POINTER(tah_layout), I32P, I32P, P, P, P],
}

_lib = None


def lib():
    """The loaded libtah, every entry point typed. Raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"libtah is not built ({LIBPATH} missing): run `make -C treeamr.jl_amd` "
                "or __graft_entry__.build()")
        # torch first: its HIP runtime (soname libamdhip64.so.7, which
        # libtah needs) is then the one the process binds libtah to. Loaded
        # the other way round, torch brings in a second runtime beside
        # libtah's, and libtah's finds no device once torch's has the GPU.
        import torch  # noqa: F401
        L = C.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.tah_last_error.argtypes = []
        L.tah_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class TahError(Exception):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class StaleScheduleError(TahError, ValueError):
    pass


class LayoutError(TahError, ValueError):
    pass


class UnsupportedError(TahError, NotImplementedError):
    pass


def check(status):
    if status == TAH_OK:
        return
    msg = lib().tah_last_error().decode()
    if status == TAH_ESTALE:
        raise StaleScheduleError(status, msg)
    if status == TAH_ELAYOUT:
        raise LayoutError(status, msg)
    if status == TAH_EUNSUPPORTED:
        raise UnsupportedError(status, msg)
    if status == TAH_EINVAL:
        raise ValueError(msg)
    raise TahError(status, msg)
