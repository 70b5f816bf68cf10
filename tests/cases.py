"""Shared problem builders for the tests (shrunk versions of the benchmark configurations)."""
import math

import numpy as np

from ibhip.problem import (BodyGroup, GridFrame, MultiLevel, Problem, RelativeFrame, RigidBody,
                           SchemeCNAB, StreamFcnFluid, UniformGrid, circle)


def cylinder_case(n=64, nlev=3, Re=100.0, dt=None, force_solver="auto", precision="fp64",
                  spacing=0.5, cg_reltol=1e-12):
    """Stationary circular cylinder of diameter 1 on [-1, 3]x[-2, 2] with n×n finest cells
    (config 1 is n=200, L=3, dt=0.004, Re=100)."""
    h = 4.0 / n
    grid = UniformGrid(h, (-1.0, 3.0), (-2.0, 2.0))
    fluid = StreamFcnFluid(MultiLevel(grid, nlev), (1.0, 0.0), Re)
    body = circle((0.0, 0.0), 0.5, h * spacing)
    dt = dt if dt is not None else 0.2 * h
    sch = SchemeCNAB(dt, precision=precision, force_solver=force_solver, cg_reltol=cg_reltol)
    return Problem(fluid, BodyGroup(body), sch)


def plate_points(chord, n):
    """Flat plate of `chord` with n points along its body-frame x axis, centred at 0."""
    s = (np.arange(n) + 0.5) / n - 0.5
    xb = np.stack([chord * s, np.zeros(n)], axis=1)
    return xb, np.full(n, chord / n)


def pitching_plate_frame(center, theta0_deg, amp_deg, omega, phase):
    th0, amp = math.radians(theta0_deg), math.radians(amp_deg)
    cx, cy = center
    # angle of attack α: the plate's body frame is rotated by -α (nose up into +x flow)
    return RelativeFrame(
        r=lambda t: (cx, cy),
        v=lambda t: (0.0, 0.0),
        theta=lambda t: -(th0 + amp * math.sin(omega * t + phase)),
        omega=lambda t: -(amp * omega * math.cos(omega * t + phase)),
        base=GridFrame(),
    )


def cylinder_plate_case(n=64, nlev=3, Re=200.0, dt=None, precision="fp64", cg_reltol=1e-12,
                        seed=2, domain=None, plate_pts=None, cyl_spacing=0.5,
                        force_solver="auto", pc_refresh=50):
    """Config 2 geometry (cylinder d=1 at the origin + pitching flat plate of chord 1 at
    (2.5, 0), 30° ± 10°, ω = 1) on a smaller grid."""
    x0, x1, y0, y1 = domain or (-1.28, 3.84, -2.56, 2.56)
    h = (x1 - x0) / n
    grid = UniformGrid(h, (x0, x1), (y0, y1))
    fluid = StreamFcnFluid(MultiLevel(grid, nlev), (1.0, 0.0), Re)
    cyl = circle((0.0, 0.0), 0.5, h * cyl_spacing)
    npl = plate_pts or max(8, int(round(1.0 / h)))
    xb, ds = plate_points(1.0, npl)
    phase = float(np.random.default_rng(seed).uniform(0, 2 * math.pi))
    plate = RigidBody(pitching_plate_frame((2.5, 0.0), 30.0, 10.0, 1.0, phase), xb, ds)
    dt = dt if dt is not None else 0.2 * h
    sch = SchemeCNAB(dt, precision=precision, cg_reltol=cg_reltol, force_solver=force_solver,
                     pc_refresh=pc_refresh)
    return Problem(fluid, BodyGroup(cyl, plate), sch)
