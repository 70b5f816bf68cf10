"""Failure detection through the C ABI: the reference's BoundsError (structure-ops.jl:70-72,
97-99), error("infinite flux") (structure-ops.jl:80-82) and cg! stopping at maxiter
(cnab/ops.jl:110) surface as IBH_EBOUNDS, IBH_ENONFINITE and IBH_ENOCONV.

Every case here is rejected on the host before a launch (bounds), or runs kernels that stay
in bounds and only carry non-finite values (NaN state) or stop early (cg_maxiter = 1).
"""
import ctypes as C
import math
import warnings

import numpy as np
import pytest

from cases import cylinder_case, cylinder_plate_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ibhip():
    import ibhip as m

    m.load_library()
    return m


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_points_outside_the_grid_are_rejected_before_upload(ibhip):
    from ibhip import _lib

    prob = cylinder_plate_case(n=64, nlev=2, cg_reltol=1e-12)
    st = ibhip.initstate(prob)
    sv = ibhip.initsolver(prob, st)
    case = st.case
    ibhip.advance_(st, sv, 0.0)
    before = case.field(_lib.F_POINTS_R, 2 * case.nb).copy()
    r, v = prob.body_points(prob.scheme.dt)
    lib = case.lib
    h = prob.fluid.domain.base.h
    nb = case.nb
    # one point of the plate 3.5 cells from the right edge: its ±6 support leaves the grid
    for bad_x in (prob.fluid.domain.base.xlims[1] - 3.5 * h, float("nan"), 1e300):
        rb = r.copy()
        rb[nb - 1] = bad_x
        rc = lib.ibh_set_body_points(case.h, dptr(rb), dptr(v))
        print(f"x = {bad_x}: rc {rc}, {lib.ibh_last_error(case.h).decode()}")
        assert rc == _lib.EBOUNDS
        rc = lib.ibh_set_field(case.h, _lib.F_POINTS_R, dptr(rb))
        assert rc == _lib.EBOUNDS
        # nothing was uploaded
        assert np.array_equal(case.field(_lib.F_POINTS_R, 2 * nb), before)
    # the case still steps with valid points
    ibhip.advance_(st, sv, prob.scheme.dt)
    assert np.isfinite(case.Fb()).all()


def test_fixed_body_near_the_edge_fails_setup(ibhip):
    from ibhip import _lib
    from ibhip.problem import BodyGroup, MultiLevel, Problem, SchemeCNAB, StreamFcnFluid, UniformGrid, circle

    n = 64
    h = 4.0 / n
    grid = UniformGrid(h, (-1.0, 3.0), (-2.0, 2.0))
    # a circle centred 0.55 from the left edge: points reach within 0.05 = 0.8 h of it
    prob = Problem(StreamFcnFluid(MultiLevel(grid, 2), (1.0, 0.0), 100.0),
                   BodyGroup(circle((-0.45, 0.0), 0.5, h / 2)), SchemeCNAB(0.2 * h))
    case = ibhip.DeviceCase(prob)
    with pytest.raises(_lib.IbhError) as e:
        case.setup()
    print(e.value)
    assert e.value.code == _lib.EBOUNDS


def test_cg_maxiter_gives_enoconv(ibhip):
    from ibhip import _lib

    prob = cylinder_case(n=64, nlev=2, force_solver="cg")
    prob.scheme.cg_maxiter = 1
    st = ibhip.initstate(prob)
    sv = ibhip.initsolver(prob, st)
    case = st.case
    ux, uy = prob.fluid.freestream(0.0)
    rc = case.lib.ibh_advance(case.h, 0.0, ux, uy)
    print(rc, case.lib.ibh_last_error(case.h).decode())
    assert rc == _lib.ENOCONV
    assert case.cg_iters() == 1
    # the step completed: the state is finite and the next step runs
    assert np.isfinite(case.gamma()).all()
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        ibhip.advance_(st, sv, prob.scheme.dt)
    assert any(issubclass(x.category, _lib.CGNotConverged) for x in w)


@pytest.mark.parametrize("solver", ["dense", "cg"])
def test_nonfinite_state_gives_enonfinite(ibhip, solver):
    from ibhip import _lib

    prob = cylinder_case(n=64, nlev=2, force_solver=solver)
    st = ibhip.initstate(prob)
    sv = ibhip.initsolver(prob, st)
    case = st.case
    ibhip.advance_(st, sv, 0.0)
    g = case.gamma().T.reshape(-1).copy()
    g[case.ng // 2 + 7] = math.nan
    case.set_field(_lib.F_GAMMA, g)
    with pytest.raises(_lib.IbhError) as e:
        ibhip.advance_(st, sv, prob.scheme.dt)
    print(e.value)
    assert e.value.code == _lib.ENONFINITE


def test_advance_n_refuses_moving_bodies(ibhip):
    from ibhip import _lib

    prob = cylinder_plate_case(n=64, nlev=2)
    st = ibhip.initstate(prob)
    ibhip.initsolver(prob, st)
    with pytest.raises(_lib.IbhError) as e:
        st.case.advance_n(0.0, 2)
    assert e.value.code == _lib.EINVAL
