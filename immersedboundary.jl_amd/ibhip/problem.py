"""Problem description types: a Python mirror of ImmersedBoundary.jl's setup API.

Names, argument order and meaning follow the reference:
  * frames            src/dynamics.jl:12-29        (GridFrame, LabFrame, RelativeFrame)
  * grids             src/fluids/fluids.jl:53-102  (UniformGrid, MultiLevel, gridstep)
  * fluid             src/streamfcn/fluids.jl:49-58 (StreamFcnFluid(domain, freestream, Re))
  * bodies            src/bodies/*.jl              (RigidBody, BodyGroup, circle)
  * scheme            src/schemes.jl:19-28          (SchemeCNAB(dt), β = [1.5, -0.5])
  * problem           src/ImmersedBoundary.jl:34-38 (Problem(fluid, bodies, scheme))

Only what the CNAB step reads is kept; the geometry derived here (nx, ny, nlevels, h, x0,
y0, stacked body points) is what the C ABI receives.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np


# --------------------------------------------------------------------------- frames
class AbstractFrame:
    pass


class GridFrame(AbstractFrame):
    """A body fixed in the grid frame (src/dynamics.jl:7)."""

    def __repr__(self):
        return "GridFrame()"


class LabFrame(AbstractFrame):
    """The laboratory frame; the only fluid frame with a base flux (fluid-ops.jl:403-417)."""

    def __repr__(self):
        return "LabFrame()"


@dataclass
class RelativeFrame(AbstractFrame):
    """A frame moving relative to `base` (src/dynamics.jl:23-29).

    r(t), v(t) -> (x, y) origin position / velocity; theta(t), omega(t) -> angle and angular
    velocity.  Points of a body in this frame are r + R(θ)·xb and their velocity is
    v + Ω·R'(θ)·xb (src/bodies/body-points.jl:47-64, with the field Ω that the frame
    actually defines).
    """

    r: Callable[[float], Sequence[float]]
    v: Callable[[float], Sequence[float]]
    theta: Callable[[float], float]
    omega: Callable[[float], float]
    base: AbstractFrame = field(default_factory=GridFrame)


# --------------------------------------------------------------------------- grids
def _range_length(x0: float, h: float, x1: float) -> int:
    """Number of points of the Julia float range x0:h:x1.

    Julia lifts start/step/stop to nearby rationals, so 4/0.02 gives exactly 201 points even
    when the float quotient is 199.999...; the same effect is obtained by rounding quotients
    that are within a relative 1e-10 of an integer.
    """
    q = (x1 - x0) / h
    k = round(q)
    if abs(q - k) <= 1e-10 * max(1.0, abs(q)):
        return int(k) + 1
    return int(math.floor(q)) + 1


@dataclass
class UniformGrid:
    """UniformGrid(h, (x0, x1), (y0, y1)) — src/fluids/fluids.jl:53-62."""

    h: float
    xlims: tuple
    ylims: tuple

    def __post_init__(self):
        self.h = float(self.h)
        self.nxp = _range_length(self.xlims[0], self.h, self.xlims[1])
        self.nyp = _range_length(self.ylims[0], self.h, self.ylims[1])

    @property
    def x0(self) -> float:
        return float(self.xlims[0])

    @property
    def y0(self) -> float:
        return float(self.ylims[0])

    def xycoords(self):
        xs = self.x0 + self.h * np.arange(self.nxp)
        ys = self.y0 + self.h * np.arange(self.nyp)
        return xs, ys

    def gridstep(self) -> float:
        return self.h


@dataclass
class MultiLevel:
    """MultiLevel(base, nlevels) — src/fluids/fluids.jl:86-102 (each level doubles the extent)."""

    base: UniformGrid
    nlevels: int

    def gridstep(self, lev: int = 1) -> float:
        return self.base.h * 2.0 ** (lev - 1)


def _freestream(f):
    if callable(f):
        return f
    ux, uy = (float(v) for v in f)
    return lambda t: (ux, uy)


@dataclass
class StreamFcnFluid:
    """StreamFcnFluid(domain, freestream, Re) — src/streamfcn/fluids.jl:49-58.

    `domain` is a UniformGrid (one level) or a MultiLevel; `freestream(t) -> (ux, uy)` or a
    constant pair.  The fluid frame is the LabFrame (the only one the reference steps).
    """

    domain: object
    freestream: object
    Re: float

    def __post_init__(self):
        if isinstance(self.domain, UniformGrid):
            self.domain = MultiLevel(self.domain, 1)
        self.freestream = _freestream(self.freestream)
        self.Re = float(self.Re)
        base = self.domain.base
        self.nx = base.nxp - 1
        self.ny = base.nyp - 1
        if self.nx % 4 or self.ny % 4:
            raise ValueError(
                f"grid of {self.nx}x{self.ny} cells: the multilevel transfers need nx, ny "
                "divisible by 4 (fluid-ops.jl:189-192, 215-232)"
            )
        self.nlevels = self.domain.nlevels
        self.frame = LabFrame()


# --------------------------------------------------------------------------- bodies
@dataclass
class RigidBody:
    """RigidBody(frame, xb, ds) — src/bodies/rigid-bodies.jl:4-8.

    xb: n×2 array of (x, y) points (in the body frame for a RelativeFrame); ds: n segment
    lengths.
    """

    frame: AbstractFrame
    xb: np.ndarray
    ds: np.ndarray

    def __post_init__(self):
        self.xb = np.ascontiguousarray(np.asarray(self.xb, dtype=np.float64).reshape(-1, 2))
        self.ds = np.asarray(self.ds, dtype=np.float64).reshape(-1)
        if self.ds.shape[0] != self.xb.shape[0]:
            raise ValueError("ds must have one entry per body point")

    @property
    def npoints(self) -> int:
        return self.xb.shape[0]

    @property
    def moving(self) -> bool:
        return not isinstance(self.frame, GridFrame)


class BodyGroup(list):
    """BodyGroup(bodies...) — src/bodies/bodies.jl:11-21."""

    def __init__(self, *bodies):
        if len(bodies) == 1 and not isinstance(bodies[0], RigidBody):
            bodies = tuple(bodies[0])
        super().__init__(bodies)

    @property
    def npoints(self) -> int:
        return sum(b.npoints for b in self)


def circle(center, r, h, frame=None) -> RigidBody:
    """circle((x, y), r, h; frame=GridFrame()) — src/bodies/body-tools.jl:1-10.

    n = floor(π r / h) points, so the spacing is about 2h; pass h/2 for spacing ≈ h.
    """
    frame = GridFrame() if frame is None else frame
    x, y = center
    n = int(math.floor(math.pi * r / h))
    t = 2 * math.pi / n * np.arange(n)
    xb = np.stack([x + r * np.cos(t), y + r * np.sin(t)], axis=1)
    ds = math.hypot(xb[1, 0] - xb[0, 0], xb[1, 1] - xb[0, 1])
    return RigidBody(frame, xb, np.full(n, ds))


# --------------------------------------------------------------------------- scheme / problem
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "mixed": 3}
FORCE_SOLVERS = {"auto": 0, "dense": 1, "cg": 2, "pcg": 3}


@dataclass
class SchemeCNAB:
    """SchemeCNAB(dt) — src/schemes.jl:19-28, 2-step Adams–Bashforth weights [1.5, -0.5].

    The extra keywords select the device implementation and have no reference counterpart:
    `precision` of the bulk sine transforms ("fp64", "fp32", "bf16x3"; "mixed": FP64 for the
    transforms whose error reaches the forces, FP32 for the coarse levels of the step's final
    Vort2Flux pass, whose ψ and q feed only the next nonlinear term) and the surface-force
    solver ("auto": the precomputed dense inverse for bodies fixed in the grid, PCG
    otherwise; "dense"; "cg": the reference's unpreconditioned cg!; "pcg": cg! with the dense
    inverse of B as preconditioner, rebuilt for moving bodies every `pc_refresh` steps, or
    when its drift has cost as much as a rebuild for pc_refresh = 0).
    cg_maxiter / cg_reltol are the reference's cg! settings (cnab/ops.jl:110); every CG
    variant stops on the same rule ‖b − B·x‖ ≤ reltol·‖b − B·x₀‖.  `refine_sweeps` (FP32 /
    bf16x3 only): FP64-residual refinement sweeps of every low-precision solve (trial Ainv_ℓ,
    both Vort2Flux Poisson solves) on the finest `refine_levels` levels (0: every level; the
    final Vort2Flux pass refines level 1 only in FP32); refine_sweeps = 0 picks 1 for FP32 and
    2 for bf16x3, a negative value turns refinement off.
    """

    dt: float
    precision: str = "fp64"
    force_solver: str = "auto"
    cg_maxiter: int = 5000
    cg_reltol: float = 1e-12
    pc_refresh: int = 0
    beta: tuple = (1.5, -0.5)
    refine_sweeps: int = 0
    refine_levels: int = 0

    def __post_init__(self):
        if self.precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        if self.force_solver not in FORCE_SOLVERS:
            raise ValueError(f"force_solver must be one of {sorted(FORCE_SOLVERS)}")

    def timestep(self) -> float:
        return self.dt


@dataclass
class Problem:
    """Problem(fluid, bodies, scheme) — src/ImmersedBoundary.jl:34-38."""

    fluid: StreamFcnFluid
    bodies: BodyGroup
    scheme: SchemeCNAB

    def __post_init__(self):
        if isinstance(self.bodies, RigidBody):
            self.bodies = BodyGroup(self.bodies)
        elif not isinstance(self.bodies, BodyGroup):
            self.bodies = BodyGroup(*self.bodies)

    # geometry handed to the native layer -------------------------------------------
    @property
    def stacked_points(self) -> np.ndarray:
        """nb×2 stacked initial points (BodyGroupPoints, body-points.jl:13-26)."""
        if not self.bodies:
            return np.zeros((0, 2))
        return np.concatenate([b.xb for b in self.bodies], axis=0)

    @property
    def stacked_ds(self) -> np.ndarray:
        if not self.bodies:
            return np.zeros(0)
        return np.concatenate([b.ds for b in self.bodies])

    @property
    def any_moving(self) -> bool:
        return any(b.moving for b in self.bodies)

    def body_points(self, t: float):
        """update_bodies! (body-points.jl:47-76): stacked positions and velocities at t.

        Bodies in the GridFrame keep their points and have zero velocity.  Bodies in a
        RelativeFrame get r0 + R(θ)·xb and v0 + Ω·R'(θ)·xb.
        """
        rs, vs = [], []
        for b in self.bodies:
            if not b.moving:
                rs.append(b.xb)
                vs.append(np.zeros_like(b.xb))
                continue
            fr = b.frame
            r0 = np.asarray(fr.r(t), dtype=np.float64)
            v0 = np.asarray(fr.v(t), dtype=np.float64)
            th = float(fr.theta(t))
            om = float(fr.omega(t))
            c, s = math.cos(th), math.sin(th)
            Rx = np.array([[c, -s], [s, c]])
            Rv = om * np.array([[-s, -c], [c, -s]])
            rs.append(r0[None, :] + b.xb @ Rx.T)
            vs.append(v0[None, :] + b.xb @ Rv.T)
        r = np.concatenate(rs, axis=0)
        v = np.concatenate(vs, axis=0)
        # column-major nb×2 (x block, then y block), as the reference stores points
        return np.ascontiguousarray(r.T).reshape(-1), np.ascontiguousarray(v.T).reshape(-1)


def timesteps(t0: float, dt: float, tf: float) -> np.ndarray:
    """The inclusive range t0:dt:tf of solve! (src/solve.jl:188-199)."""
    n = _range_length(t0, dt, tf)
    return t0 + dt * np.arange(n)
