"""Synthetic cases of the benchmark shapes (bodies and kinematics drawn from fixed seeds).

  cylinder_re100   200² cells on [-1,3]×[-2,2], 3 levels, cylinder d=1 with 157 points,
                   Re=100, dt=0.004 (the CPU-runnable case)
  cylinder_plate   1024² cells, h=0.005, 5 levels; fixed cylinder d=1 (628 points) and a flat
                   plate of chord 1 (200 points) at (2.5, 0) pitching as 30° + 10°·sin(t + φ)
                   (reduced frequency 0.5, ω = 1), φ from seed 2; Re=200, dt=0.001
  cylinder_array   2048² cells, h=0.0025, 4 levels; a 4×4 array of d=0.25 cylinders (314
                   points each) at spacing 0.6 jittered by N(0, 0.02) from seed 5; Re=150,
                   dt=5e-4
"""
from __future__ import annotations

import math

import numpy as np

from .problem import (BodyGroup, GridFrame, MultiLevel, Problem, RelativeFrame, RigidBody,
                      SchemeCNAB, StreamFcnFluid, UniformGrid, circle)


def flat_plate(chord: float, n: int, frame) -> RigidBody:
    """A flat plate of n points along its body-frame x axis, centred on the frame origin."""
    s = (np.arange(n) + 0.5) / n - 0.5
    xb = np.stack([chord * s, np.zeros(n)], axis=1)
    return RigidBody(frame, xb, np.full(n, chord / n))


def pitching_frame(center, mean_deg, amp_deg, omega, phase) -> RelativeFrame:
    """Rotation about `center` by −(mean + amp·sin(ωt + φ)): a positive angle of attack
    raises the leading edge (body-frame −x) into a +x freestream."""
    m, a = math.radians(mean_deg), math.radians(amp_deg)
    cx, cy = center
    return RelativeFrame(
        r=lambda t: (cx, cy),
        v=lambda t: (0.0, 0.0),
        theta=lambda t: -(m + a * math.sin(omega * t + phase)),
        omega=lambda t: -(a * omega * math.cos(omega * t + phase)),
        base=GridFrame(),
    )


def cylinder_re100(precision="fp64", force_solver="auto", **kw) -> Problem:
    h = 0.02
    grid = UniformGrid(h, (-1.0, 3.0), (-2.0, 2.0))
    fluid = StreamFcnFluid(MultiLevel(grid, 3), (1.0, 0.0), 100.0)
    body = circle((0.0, 0.0), 0.5, h / 2)
    return Problem(fluid, BodyGroup(body),
                   SchemeCNAB(0.004, precision=precision, force_solver=force_solver, **kw))


def cylinder_plate(n=1024, nlev=5, precision="fp64", force_solver="auto", seed=2,
                   plate_points=200, **kw) -> Problem:
    x0, x1, y0, y1 = -1.28, 3.84, -2.56, 2.56
    h = (x1 - x0) / n
    grid = UniformGrid(h, (x0, x1), (y0, y1))
    fluid = StreamFcnFluid(MultiLevel(grid, nlev), (1.0, 0.0), 200.0)
    cyl = circle((0.0, 0.0), 0.5, h / 2)
    phase = float(np.random.default_rng(seed).uniform(0.0, 2.0 * math.pi))
    plate = flat_plate(1.0, plate_points, pitching_frame((2.5, 0.0), 30.0, 10.0, 1.0, phase))
    dt = 0.001 * (1024 / n)
    return Problem(fluid, BodyGroup(cyl, plate),
                   SchemeCNAB(dt, precision=precision, force_solver=force_solver, **kw))


def cylinder_array(n=2048, nlev=4, precision="fp64", force_solver="auto", seed=5, **kw) -> Problem:
    x0, x1, y0, y1 = -2.0, 3.12, -2.56, 2.56
    h = (x1 - x0) / n
    grid = UniformGrid(h, (x0, x1), (y0, y1))
    fluid = StreamFcnFluid(MultiLevel(grid, nlev), (1.0, 0.0), 150.0)
    rng = np.random.default_rng(seed)
    bodies = []
    for a in range(4):
        for b in range(4):
            cx = -0.9 + 0.6 * a + rng.normal(0.0, 0.02)
            cy = -0.9 + 0.6 * b + rng.normal(0.0, 0.02)
            bodies.append(circle((cx, cy), 0.125, h / 2))
    dt = 5e-4 * (2048 / n)
    return Problem(fluid, BodyGroup(*bodies),
                   SchemeCNAB(dt, precision=precision, force_solver=force_solver, **kw))


def ellipse(center, a, b, rot, h) -> RigidBody:
    """An ellipse with semi-axes a, b rotated by `rot`, its points spaced ≈h in arc length."""
    s = np.linspace(0.0, 2.0 * math.pi, 4097)
    x, y = a * np.cos(s), b * np.sin(s)
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(x), np.diff(y)))])
    n = max(8, int(round(arc[-1] / h)))
    sk = np.interp(np.arange(n) * (arc[-1] / n), arc, s)
    xe, ye = a * np.cos(sk), b * np.sin(sk)
    c, sn = math.cos(rot), math.sin(rot)
    xb = np.stack([center[0] + c * xe - sn * ye, center[1] + sn * xe + c * ye], axis=1)
    return RigidBody(GridFrame(), xb, np.full(n, arc[-1] / n))


def ellipse_batch(ncases=64, seed=4, precision="fp32", sizes=(128, 192, 256, 384, 512),
                  levels=(2, 3, 4, 5), **kw) -> list:
    """Independent cases of a parameter sweep: grid n uniform over `sizes` on the 4×4 domain
    centred at (1, 0) (h = 4/n), 2–5 levels, one fixed ellipse (a = 0.5, b ~ U[0.1, 0.5],
    rotation ~ U[0, π)) with points spaced ≈h, Re ~ U[50, 400], dt = 0.2·h (CFL 0.2)."""
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(ncases):
        n = int(rng.choice(sizes))
        nlev = int(rng.choice(levels))
        b = float(rng.uniform(0.1, 0.5))
        rot = float(rng.uniform(0.0, math.pi))
        Re = float(rng.uniform(50.0, 400.0))
        h = 4.0 / n
        grid = UniformGrid(h, (-1.0, 3.0), (-2.0, 2.0))
        fluid = StreamFcnFluid(MultiLevel(grid, nlev), (1.0, 0.0), Re)
        body = ellipse((0.0, 0.0), 0.5, b, rot, h)
        probs.append(Problem(fluid, BodyGroup(body),
                             SchemeCNAB(0.2 * h, precision=precision, force_solver="dense", **kw)))
    return probs


WORKLOADS = {
    "cylinder_re100": cylinder_re100,
    "cylinder_plate": cylinder_plate,
    "cylinder_array": cylinder_array,
}
