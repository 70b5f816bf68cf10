"""The block schedules of E and Eᵀ that the tuner picks between give the same bits: E forms
each output in one wave in a fixed order whatever the waves per workgroup, and Eᵀ's fixed-point
adds do not depend on the order in which its threads walk the stencil entries.  Each schedule is
forced in turn (IBH_REG_SCHED="e,et") on a moving-body CG case and on a batch, and every result
must equal the tuned run's exactly."""
import numpy as np
import pytest

from cases import cylinder_plate_case
from test_gpu_batch import mixed_problems

pytestmark = pytest.mark.gpu

E_SCHEDS, ET_SCHEDS = 5, 10
FORCED = [f"{e},0" for e in range(E_SCHEDS)] + [f"0,{et}" for et in range(1, ET_SCHEDS)]


@pytest.fixture(scope="module")
def ibhip():
    import ibhip as m

    m.load_library()
    return m


def single_run(m, nsteps=3):
    prob = cylinder_plate_case(n=64, nlev=3)
    st = m.initstate(prob)
    sv = m.initsolver(prob, st)
    for k in range(nsteps):
        m.advance_(st, sv, k * prob.scheme.dt)
    c = st.case
    out = [c.gamma(), c.q(), c.Fb(), c.CD(), c.CL()]
    report = c.tune_report()
    c.close()
    return out, report


def batch_run(m, nsteps=2):
    probs = mixed_problems("fp64")[:4]
    states = [m.initstate(p) for p in probs]
    for p, st in zip(probs, states):
        m.initsolver(p, st)
    batch = m.Batch(states)
    for _ in range(nsteps):
        batch.advance()
    batch.sync()
    out = [st.case.gamma() for st in states] + [st.case.Fb() for st in states if st.case.nb]
    report = batch.tune_report()
    batch.close()
    for st in states:
        st.case.close()
    return out, report


def test_single_case_schedules_identical(ibhip, monkeypatch):
    monkeypatch.delenv("IBH_REG_SCHED", raising=False)
    ref, report = single_run(ibhip)
    line = [l for l in report if l.startswith("E schedule")]
    assert line, report
    print(line[0])
    for f in FORCED:
        monkeypatch.setenv("IBH_REG_SCHED", f)
        got, _ = single_run(ibhip)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), f"IBH_REG_SCHED={f}"


def test_batch_schedules_identical(ibhip, monkeypatch):
    monkeypatch.delenv("IBH_REG_SCHED", raising=False)
    ref, report = batch_run(ibhip)
    lines = [l for l in report if " schedule " in l]
    assert any(l.startswith("E schedule") for l in lines), report
    assert any(l.startswith("Eᵀ schedule") for l in lines), report
    print("\n".join(lines))
    for f in FORCED:
        monkeypatch.setenv("IBH_REG_SCHED", f)
        got, _ = batch_run(ibhip)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), f"IBH_REG_SCHED={f}"
