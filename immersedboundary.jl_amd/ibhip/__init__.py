"""ibhip — the CNAB immersed-boundary projection step of ImmersedBoundary.jl on MI355X.

The API mirrors the reference package (src/ImmersedBoundary.jl): build a Problem from a
StreamFcnFluid, a BodyGroup and a SchemeCNAB, then `solve` / `solve_` (solve!) or drive
`initstate` / `initsolver` / `advance_` (advance!) yourself.  Every step runs in
libibhip.so (hand-written HIP for gfx950) through its C ABI (include/ibh.h).
"""
from .problem import (  # noqa: F401
    AbstractFrame, BodyGroup, GridFrame, LabFrame, MultiLevel, Problem, RelativeFrame, RigidBody,
    SchemeCNAB, StreamFcnFluid, UniformGrid, circle, timesteps,
)
from .cnab import (  # noqa: F401
    Batch, BodyStress, DeviceCase, Quantities, Solver, State, StateCallback, StateResult,
    StopIteration, advance_, at_indices, at_times, each_timestep, initsolver, initstate,
    load_state, quantities, save_state, solve, solve_,
)
from ._lib import IbhError, load as load_library  # noqa: F401

__all__ = [
    "AbstractFrame", "BodyGroup", "GridFrame", "LabFrame", "MultiLevel", "Problem",
    "RelativeFrame", "RigidBody", "SchemeCNAB", "StreamFcnFluid", "UniformGrid", "circle",
    "timesteps", "Batch", "BodyStress", "DeviceCase", "Quantities", "Solver", "State", "StateCallback",
    "StateResult", "StopIteration", "advance_", "at_indices", "at_times", "each_timestep",
    "initsolver", "initstate", "load_state", "quantities", "save_state", "solve", "solve_",
    "IbhError", "load_library",
]
