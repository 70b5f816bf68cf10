"""The benchmark configurations on the GPU, checked against the FP64 oracle at sizes it runs in
seconds, and at full size through properties that do not need it.

  mixed precision (cylinder + pitching plate in FP32 and bf16x3 with FP64 forces): lift and
      drag within 1e-6 of the oracle, every field's error printed
  batch of ellipses in FP32 (grouped launches, HIP graph): lift and drag within 1e-6
  array of 16 cylinders: FP64 within 1e-10, FP32 within 1e-6
  cylinder + pitching plate at 1024², 5 levels (1656 unknowns: the Cholesky-inverted
      preconditioner, column-batched setup): finite fields and the no-slip constraint
      E(q₁ + q₀) = v·h after every step (timestepping.jl:122-139, 141-171)

Error norms: fields ‖x − x_ref‖∞/‖x_ref‖∞ per level; CD max over steps 2..N of |ΔCD| over
max |CD_ref|, CL the same but relative to max |CD_ref| (CL ≈ 0 for symmetric bodies); step 1
(the impulsive start) is printed separately.
"""
import numpy as np
import pytest

import ibref
from cases import cylinder_plate_case

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


@pytest.fixture(scope="module")
def ibhip():
    import ibhip as m

    m.load_library()
    return m


def field_errors(case, orc):
    out = {}
    for name, a, b in [("Γ", case.gamma(), orc.gamma()), ("ψ", case.psi(), orc.psi()),
                       ("q", case.q(), orc.q())]:
        out[name] = [rel(a[:, lev], b[:, lev]) for lev in range(case.nlev)]
    if case.nb:
        out["Fb"] = [rel(case.Fb(), orc.Fb())]
    return out


def force_errors(cd, cl, cd_ref, cl_ref):
    cd, cl, cd_ref, cl_ref = map(np.asarray, (cd, cl, cd_ref, cl_ref))
    den = np.abs(cd_ref[1:]).max(axis=0)
    e_cd = (np.abs(cd[1:] - cd_ref[1:]).max(axis=0) / den).max()
    e_cl = (np.abs(cl[1:] - cl_ref[1:]).max(axis=0) / den).max()
    e1 = (np.abs(cd[0] - cd_ref[0]) / np.abs(cd_ref[0])).max()
    return e_cd, e_cl, e1


def run_with_oracle(ibhip, prob, nsteps):
    st = ibhip.initstate(prob)
    sv = ibhip.initsolver(prob, st)
    orc = ibref.OracleRun(prob)
    cd, cl, cdr, clr = [], [], [], []
    for k in range(nsteps):
        t = k * prob.scheme.dt
        ibhip.advance_(st, sv, t)
        orc.advance(t)
        cd.append(st.case.CD())
        cl.append(st.case.CL())
        cdr.append(orc.CD())
        clr.append(orc.CL())
    return st, orc, force_errors(cd, cl, cdr, clr)


def report(label, st, orc, forces):
    e_cd, e_cl, e1 = forces
    fe = field_errors(st.case, orc)
    parts = [f"{k} " + "/".join(f"{v:.1e}" for v in vs) for k, vs in fe.items()]
    print(f"{label}: CD {e_cd:.2e}, CL {e_cl:.2e} (step 1 CD {e1:.2e}); " + ", ".join(parts))
    return fe


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "mixed"])
def test_mixed_precision_cylinder_pitching_plate(ibhip, precision):
    # the cylinder + pitching plate geometry on 128² cells, 3 levels; the plate moves, so the
    # forces come from PCG (FP64 B-products) and the bulk transforms are low precision with
    # FP64-residual refinement (mixed: FP32 on the final pass's coarse levels, unrefined).  Both CG iterations stop at 1e-14 so that their histories
    # differ by less than the precision effect being measured.
    prob = cylinder_plate_case(n=128, nlev=3, precision=precision, cg_reltol=1e-14)
    st, orc, forces = run_with_oracle(ibhip, prob, 5)
    fe = report(f"cylinder+plate 128² L3 {precision}", st, orc, forces)
    e_cd, e_cl, e1 = forces
    assert e_cd < 1e-6 and e_cl < 1e-6 and e1 < 1e-6
    assert max(max(v) for v in fe.values()) < 1e-5


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_low_precision_ellipse_batch_matches_oracle(ibhip, precision):
    import ibhip as m
    from ibhip.workloads import ellipse_batch

    probs = ellipse_batch(ncases=3, seed=4, precision=precision, sizes=(64, 96, 128),
                          levels=(2, 3, 4))
    nsteps = 5
    states = [m.initstate(p) for p in probs]
    for p, st in zip(probs, states):
        m.initsolver(p, st)
    batch = m.Batch(states)
    cd = [[] for _ in probs]
    cl = [[] for _ in probs]
    for k in range(nsteps):
        batch.advance()
        batch.sync()
        for i, st in enumerate(states):
            cd[i].append(st.case.CD())
            cl[i].append(st.case.CL())
    for i, p in enumerate(probs):
        orc = ibref.OracleRun(p)
        cdr, clr = [], []
        for k in range(nsteps):
            orc.advance(k * p.scheme.dt)
            cdr.append(orc.CD())
            clr.append(orc.CL())
        forces = force_errors(cd[i], cl[i], cdr, clr)
        report(f"batch case {i} ({p.fluid.nx}², L{p.fluid.nlevels}, {precision})", states[i], orc,
               forces)
        assert max(forces) < 1e-6


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_cylinder_array(ibhip, precision, tol):
    from ibhip.workloads import cylinder_array

    # the 4×4 array of d = 0.25 cylinders (jittered from seed 5) on 256² cells, 3 levels:
    # 16 bodies, 624 points, dense Binv of 1248 unknowns (rocSOLVER above 768)
    prob = cylinder_array(n=256, nlev=3, precision=precision)
    assert len(prob.bodies) == 16
    st, orc, forces = run_with_oracle(ibhip, prob, 5)
    fe = report(f"cylinder array 256² L3 {precision}", st, orc, forces)
    assert max(forces) < tol
    if precision == "fp64":
        assert max(max(v) for v in fe.values()) < tol


def test_full_size_cylinder_plate_no_slip(ibhip):
    from ibhip.workloads import cylinder_plate

    prob = cylinder_plate()  # 1024², h = 0.005, 5 levels, 628 + 200 points, Re 200, dt 1e-3
    assert (prob.fluid.nx, prob.fluid.nlevels, prob.bodies.npoints) == (1024, 5, 828)
    st = ibhip.initstate(prob)
    sv = ibhip.initsolver(prob, st)
    case = st.case
    h = prob.fluid.domain.base.h
    for k in range(3):
        t = k * prob.scheme.dt
        ibhip.advance_(st, sv, t)
        q, q0 = case.q(), case.q0()
        for a in (case.gamma(), case.psi(), q, case.Fb()):
            assert np.isfinite(a).all()
        _, v = prob.body_points(t)
        slip = case.E(q[:, 0] + q0[:, 0]) - v * h
        free = case.E(q0[:, 0])
        err = np.abs(slip).max() / np.abs(free).max()
        print(f"1024² L5 step {k + 1}: no-slip residual {err:.2e}, CG iterations "
              f"{case.cg_iters()}, CD {case.CD()}")
        assert err < 1e-9
    assert np.all(case.CD() > 0)
