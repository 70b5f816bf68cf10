"""solve(prob, (t0, tf); save) on the device (src/solve.jl:188-231): the saved drag history
equals a hand-written advance! loop bit for bit and the oracle's within 1e-10."""
import numpy as np
import pytest

import ibref
from cases import cylinder_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ibhip():
    import ibhip as m

    m.load_library()
    return m


def test_solve_saves_match_manual_loop_and_oracle(ibhip):
    prob = cylinder_case(n=64, nlev=3)
    dt = prob.scheme.dt
    nsteps = 6
    cd = ibhip.each_timestep(lambda st: st.qty.stress[0].CD, float)
    cl = ibhip.at_indices(lambda st: st.qty.stress[0].CL, [2, 4, 6], float)
    res_cd, res_cl = ibhip.solve(prob, (0.0, (nsteps - 1) * dt), save=[cd, cl])
    assert len(res_cd) == nsteps and len(res_cl) == 3
    assert np.allclose(res_cd.t, dt * np.arange(nsteps), rtol=0, atol=1e-15)

    st = ibhip.initstate(prob)
    sv = ibhip.initsolver(prob, st)
    manual = []
    for k in range(nsteps):
        ibhip.advance_(st, sv, k * dt)
        manual.append(st.qty.stress[0].CD)
    assert list(res_cd) == manual  # bitwise: same device path

    orc = ibref.OracleRun(prob)
    ref_cd, ref_cl = [], []
    for k in range(nsteps):
        orc.advance(k * dt)
        ref_cd.append(float(orc.CD()[0]))
        ref_cl.append(float(orc.CL()[0]))
    ref_cd = np.array(ref_cd)
    err = np.abs(np.array(res_cd) - ref_cd).max() / np.abs(ref_cd).max()
    err_cl = np.abs(np.array(res_cl) - np.array(ref_cl)[[1, 3, 5]]).max() / np.abs(ref_cd).max()
    print(f"solve: CD history rel err {err:.2e}, CL (rel. max CD) {err_cl:.2e}")
    assert err < 1e-10 and err_cl < 1e-10
