"""treeamr_hip -- the MI355X ghost fill behind TreeAMR.jl's fill interface.

The host mirror of the reference API for this path, over libtah's C ABI:

    forest = Forest((2, 2, 2), N=8, periodic=(True,) * 3)      # src/forest.jl:314
    forest.refine([...]); forest.balance()                   # src/forest.jl:1761,1837
    ops = Operators(prolongation=3, restriction=2, family=Conservative)  # src/operators.jl:314
    fs = FieldSet(forest, 4, G=2)                            # src/storage.jl:417
    sched = GhostSchedule(fs, ops)                           # src/schedule.jl:3069
    fill_ghosts(fs, sched)                                   # src/ghosts.jl:1783

Field data lives in HBM as a torch float64 tensor; every fill runs the
library's HIP kernels. There is no CPU fill in this package: a field set on
the CPU is refused by fill_ghosts.
"""
from .core import (  # noqa: F401
    Forest, MortonKey, Operators, FieldSet, GhostSchedule, fill_ghosts, fill_ghosts_adjoint,
    PointValue, Conservative, NoParity, EvenParity, OddParity, cellcentered, vertexcentered,
    Context, default_context, Plan, fill_ghosts_batched, FillGraph, fill_ghosts_timed, kernel_class_name, tune,
    tune_batched,
    ownextents, statelength, statevector, staterange, stateview, statearray, scatter, gather,
    RegridSchedule, regrid, regrid_transfer, UnifySchedule, unify, unify_desc,
    AllVariables, WithPatch, Relative, CellBoundary, boundary_by_coordinates, boundary_values,
    fill_by_coordinates, fill_ghosts_adjoint_timed,
)
from ._abi import TahError, StaleScheduleError, LayoutError, UnsupportedError  # noqa: F401
