"""solve! / solve and their callbacks (src/solve.jl:104-231) on the host: the inclusive time
range, when each kind of StateCallback runs, StopIteration and the saved StateResults.

The stepping itself is replaced by a stub that records the times it was called with, so these
run without a GPU; the same loop drives the device case in tests/test_gpu_solve.py.
"""
import numpy as np
import pytest

from ibhip import cnab
from ibhip.problem import timesteps


class StubState:
    def __init__(self, t0):
        self.t = t0
        self.steps = []


class StubScheme:
    def __init__(self, dt):
        self.dt = dt


class StubProblem:
    def __init__(self, dt):
        self.scheme = StubScheme(dt)


@pytest.fixture
def stub(monkeypatch):
    def advance(state, solver, t):
        state.steps.append(t)
        state.t = t

    monkeypatch.setattr(cnab, "initsolver", lambda prob, state: None)
    monkeypatch.setattr(cnab, "advance_", advance)


@pytest.mark.parametrize("t0,dt,tf,n", [
    (0.0, 0.02, 4.0, 201),         # 4/0.02 = 199.99999999999997 in floats; Julia's range: 201
    (0.0, 0.1, 0.3, 4),            # 0.3/0.1 = 2.9999999999999996
    (0.0, 0.004, 499 * 0.004, 500),  # config 1: 500 advance! calls
    (0.5, 0.25, 1.6, 5),           # tf not on the grid: stops at 1.5
    (1.0, 0.1, 1.0, 1),            # t0 = tf: one step
])
def test_timesteps_inclusive_range(t0, dt, tf, n):
    ts = timesteps(t0, dt, tf)
    assert len(ts) == n
    assert ts[0] == t0 and ts[-1] <= tf + 1e-12 * max(1.0, abs(tf))
    assert np.allclose(np.diff(ts), dt, rtol=0, atol=1e-15 * max(1.0, abs(tf)))


def test_each_timestep_and_indices_and_times(stub):
    dt = 0.02
    prob = StubProblem(dt)
    st = StubState(0.0)
    seen = {"all": [], "idx": [], "times": []}
    call = [
        cnab.each_timestep(lambda s: seen["all"].append(s.t)),
        cnab.at_indices(lambda s: seen["idx"].append(len(s.steps)), [2, 5, 9]),
        # 0.05 runs at the first step with t + dt/2 > 0.05 (t = 0.06); -1 (before t0) at the
        # first step; 0.0799 and 0.08 are both due at t = 0.08: one call
        cnab.at_times(lambda s: seen["times"].append(s.t), [-1.0, 0.05, 0.0799, 0.08, 0.5]),
    ]
    cnab.solve_(st, prob, (0.0, 0.1), call=call)
    assert st.steps == pytest.approx([0.0, 0.02, 0.04, 0.06, 0.08, 0.1])
    assert seen["all"] == st.steps
    assert seen["idx"] == [2, 5]  # index 9 is never reached
    assert seen["times"] == pytest.approx([0.0, 0.06, 0.08])


def test_stop_iteration_ends_after_running_every_callback(stub):
    prob = StubProblem(0.1)
    st = StubState(0.0)
    later = []
    call = [
        cnab.each_timestep(lambda s: cnab.StopIteration() if len(s.steps) == 3 else None),
        cnab.each_timestep(lambda s: later.append(len(s.steps))),  # still runs at step 3
    ]
    cnab.solve_(st, prob, (0.0, 1.0), call=call)
    assert len(st.steps) == 3
    assert later == [1, 2, 3]


def test_saved_results_carry_their_times(stub):
    prob = StubProblem(0.25)
    st = StubState(0.0)
    save = [
        cnab.each_timestep(lambda s: 10 * s.t),
        cnab.at_indices(lambda s: len(s.steps), [1, 3]),
        cnab.each_timestep(lambda s: None),  # nothing returned: nothing saved
    ]
    res = cnab.solve_(st, prob, (0.0, 1.0), save=save)
    assert list(res[0]) == pytest.approx([0.0, 2.5, 5.0, 7.5, 10.0])
    assert res[0].t == pytest.approx([0.0, 0.25, 0.5, 0.75, 1.0])
    assert list(res[1]) == [1, 3] and res[1].t == pytest.approx([0.0, 0.5])
    assert list(res[2]) == [] and res[2].t == []
