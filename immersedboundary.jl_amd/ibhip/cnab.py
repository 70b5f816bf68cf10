"""CNAB state, solver and driver on the GPU: the Python mirror of the reference's
initstate / initsolver / advance! / solve! / solve (src/ImmersedBoundary.jl:51-52,
src/solve.jl:52-231, src/streamfcn/cnab/*.jl) over the libibhip C ABI.

The device context holds the whole State (q, Γ, ψ, Fb, nonlin) and the Solver's work
buffers; `State.qty` is a lazily refreshed host mirror so that callbacks written for the
reference (`state.qty.stress[k].CD`, `state.qty.Γ`) keep working.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from dataclasses import dataclass
from typing import Any, Callable

import numpy as np

from . import _lib
from ._lib import dptr, f64
from .problem import FORCE_SOLVERS, PRECISIONS, Problem, timesteps


@dataclass
class BodyStress:
    """BodyStress(fb, CD, CL) — src/streamfcn/states.jl:3-7."""

    fb: np.ndarray
    CD: float
    CL: float


class DeviceCase:
    """One ibh_ctx: device-resident state and solver of a Problem."""

    def __init__(self, prob: Problem, t0: float = 0.0):
        L = _lib.load()
        self.lib = L
        self.prob = prob
        fl, sch = prob.fluid, prob.scheme
        base = fl.domain.base
        self._npts = np.array([b.npoints for b in prob.bodies] or [0], dtype=np.int32)
        self._mov = np.array([int(b.moving) for b in prob.bodies] or [0], dtype=np.int32)
        xb = prob.stacked_points
        self._xb = f64(xb.T.reshape(-1)) if xb.size else np.zeros(1)
        self._ds = f64(prob.stacked_ds) if xb.size else np.zeros(1)
        d = _lib.CaseDesc()
        d.nx, d.ny, d.nlev = fl.nx, fl.ny, fl.nlevels
        d.h, d.x0, d.y0 = base.h, base.x0, base.y0
        d.Re, d.dt = fl.Re, sch.dt
        d.beta[0], d.beta[1] = sch.beta
        d.nbodies = len(prob.bodies)
        d.npts = self._npts.ctypes.data_as(C.POINTER(C.c_int32))
        d.moving = self._mov.ctypes.data_as(C.POINTER(C.c_int32))
        d.xb = dptr(self._xb)
        d.ds = dptr(self._ds)
        d.precision = PRECISIONS[sch.precision]
        d.force_solver = FORCE_SOLVERS[sch.force_solver]
        d.cg_maxiter = sch.cg_maxiter
        d.cg_reltol = sch.cg_reltol
        d.pc_refresh = sch.pc_refresh
        d.refine_sweeps = sch.refine_sweeps
        d.refine_levels = sch.refine_levels
        self._desc = d
        h = C.c_void_p()
        rc = L.ibh_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise _lib.IbhError(rc, "ibh_create rejected the case (sizes, precision or no GPU)")
        self.h = h
        dims = (C.c_int32 * 8)()
        L.ibh_dims(h, dims)
        self.nx, self.ny, self.nlev, self.nq, self.ng, self.nb, self.nbc, self.nbodies = list(dims)
        ux, uy = fl.freestream(t0)
        self._check(L.ibh_init_state(h, float(t0), float(ux), float(uy)))

    # -------------------------------------------------------------------------- plumbing
    def _check(self, rc):
        if rc != 0:
            raise _lib.IbhError(rc, self.lib.ibh_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ibh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self):
        self._check(self.lib.ibh_setup(self.h))

    def advance(self, t: float):
        prob = self.prob
        if prob.any_moving:
            r, v = prob.body_points(t)
            r, v = f64(r), f64(v)
            self._check(self.lib.ibh_set_body_points(self.h, dptr(r), dptr(v)))
        ux, uy = prob.fluid.freestream(t)
        rc = self.lib.ibh_advance(self.h, float(t), float(ux), float(uy))
        if rc == _lib.ENOCONV:  # the step completed; cg! returns its last iterate silently
            warnings.warn(self.lib.ibh_last_error(self.h).decode(), _lib.CGNotConverged,
                          stacklevel=3)
            return
        self._check(rc)

    def advance_n(self, t0: float, n: int):
        """n steps at t0 + k·dt with the freestream of t0 (fixed bodies only)."""
        ux, uy = self.prob.fluid.freestream(t0)
        self._check(self.lib.ibh_advance_n(self.h, float(t0), int(n), float(ux), float(uy)))

    def sync(self):
        self._check(self.lib.ibh_sync(self.h))

    def field(self, which: int, n: int) -> np.ndarray:
        out = np.empty(max(n, 1))
        self._check(self.lib.ibh_get_field(self.h, which, dptr(out)))
        return out[:n]

    def set_field(self, which: int, a):
        a = f64(a).reshape(-1)
        self._check(self.lib.ibh_set_field(self.h, which, dptr(a)))

    # quantities as (n, nlev) arrays (column = level), like the reference's matrices
    def q(self):
        return self.field(_lib.F_Q, self.nq * self.nlev).reshape(self.nlev, self.nq).T

    def q0(self):
        return self.field(_lib.F_Q0, self.nq * self.nlev).reshape(self.nlev, self.nq).T

    def gamma(self):
        return self.field(_lib.F_GAMMA, self.ng * self.nlev).reshape(self.nlev, self.ng).T

    def psi(self):
        return self.field(_lib.F_PSI, self.ng * self.nlev).reshape(self.nlev, self.ng).T

    def nonlin(self, k: int):
        w = _lib.F_NONLIN1 if k == 0 else _lib.F_NONLIN2
        return self.field(w, self.ng * self.nlev).reshape(self.nlev, self.ng).T

    def Fb(self):
        return self.field(_lib.F_FB, 2 * self.nb)

    def CD(self):
        return self.field(_lib.F_CD, self.nbodies)

    def CL(self):
        return self.field(_lib.F_CL, self.nbodies)

    def fb_stress(self):
        return self.field(_lib.F_FBSTRESS, 2 * self.nb)

    def cg_iters(self) -> int:
        return self.lib.ibh_cg_iterations(self.h)

    def solver_stats(self) -> dict:
        """PCG preconditioner rebuilds: count, seconds in total / column assembly / inversion,
        and the per-iteration and per-rebuild costs the adaptive rebuild rule weighs."""
        out = np.zeros(6)
        self._check(self.lib.ibh_solver_stats(self.h, dptr(out)))
        return {"pc_rebuilds": int(out[0]), "pc_rebuild_s": out[1], "pc_columns_s": out[2],
                "pc_inverse_s": out[3], "iteration_s": out[4], "rebuild_s_weighed": out[5]}

    def lowrank_stats(self) -> dict:
        """The low-rank coarse-level part of the CG B-products: rank in use (0 when the full
        product runs), its B-product error at the last build, builds and build seconds."""
        out = np.zeros(4)
        self._check(self.lib.ibh_lowrank_stats(self.h, dptr(out)))
        return {"rank": int(out[0]), "rel_err": out[1], "builds": int(out[2]),
                "build_s": out[3]}

    def time(self) -> float:
        return self.lib.ibh_time(self.h)

    # -------------------------------------------------------------------------- operators
    def _op(self, name, *args):
        self._check(getattr(self.lib, name)(self.h, *args))

    def lapinv(self, mode: int, lev: int, b):
        out = np.zeros(self.ng)
        self._op("ibh_op_lapinv", int(mode), int(lev), dptr(f64(b)), dptr(out))
        return out

    def dst2(self, x):
        out = np.zeros(self.ng)
        self._op("ibh_op_dst2", dptr(f64(x)), dptr(out))
        return out

    def rot(self, q):
        out = np.zeros(self.ng)
        self._op("ibh_op_rot", dptr(f64(q)), dptr(out))
        return out

    def curl(self, psi, psibc):
        out = np.zeros(self.nq)
        self._op("ibh_op_curl", dptr(f64(psi)), dptr(f64(psibc)), dptr(out))
        return out

    def coarsify(self, gc, g):
        gc = f64(gc).copy()
        self._op("ibh_op_coarsify", dptr(gc), dptr(f64(g)))
        return gc

    def get_bc(self, r):
        out = np.zeros(self.nbc)
        self._op("ibh_op_get_bc", dptr(f64(r)), dptr(out))
        return out

    def apply_bc(self, r, rbc, fac):
        r = f64(r).copy()
        self._op("ibh_op_apply_bc", dptr(r), dptr(f64(rbc)), float(fac))
        return r

    def nonlinear(self, lev, q, q0, g, gbc):
        out = np.zeros(self.ng)
        self._op("ibh_op_nonlinear", int(lev), dptr(f64(q)), dptr(f64(q0)), dptr(f64(g)),
                 dptr(f64(gbc)), dptr(out))
        return out

    def vort2flux(self, g):
        g = f64(np.asarray(g).T.reshape(-1)).copy()
        psi = np.zeros(self.ng * self.nlev)
        q = np.zeros(self.nq * self.nlev)
        self._op("ibh_op_vort2flux", dptr(g), dptr(psi), dptr(q))
        return (psi.reshape(self.nlev, self.ng).T, q.reshape(self.nlev, self.nq).T,
                g.reshape(self.nlev, self.ng).T)

    def reg_update(self, r):
        self._op("ibh_op_reg_update", dptr(f64(r)))

    def E(self, q):
        out = np.zeros(2 * self.nb)
        self._op("ibh_op_E", dptr(f64(q)), dptr(out))
        return out

    def ET(self, fb):
        out = np.zeros(self.nq)
        self._op("ibh_op_ET", dptr(f64(fb)), dptr(out))
        return out

    def B_times(self, z):
        out = np.zeros(2 * self.nb)
        self._op("ibh_op_B", dptr(f64(z)), dptr(out))
        return out

    def tune_report(self) -> list:
        """The autotuner's choices so far: one line per contraction shape, and the block
        schedules of E and Eᵀ."""
        buf = C.create_string_buffer(1 << 16)
        self._check(self.lib.ibh_tune_report(self.h, buf, len(buf)))
        return [ln for ln in buf.value.decode().splitlines() if ln]

    def profile(self, on: bool):
        self._check(self.lib.ibh_profile_enable(self.h, int(on)))

    def profile_read(self, kind: int = 0, events: bool = False):
        """(ms, launches, flops) of the profiled contraction launches of one precision: device
        kernel time of the launches that ran, or (events=True) HIP-event time of all."""
        ms = C.c_double()
        n = C.c_int64()
        fl = C.c_double()
        fn = self.lib.ibh_profile_read_events if events else self.lib.ibh_profile_read
        self._check(fn(self.h, kind, C.byref(ms), C.byref(n), C.byref(fl)))
        return ms.value, n.value, fl.value

    def profile_launches(self, kind: int = 0):
        """Per-launch device time (us), algorithmic flops and bytes of the profiled contraction
        launches of one precision, in launch order (skipped launches left out)."""
        return _launch_records(self.lib.ibh_profile_launches, self.h, kind, self._check)


def _launch_records(fn, h, kind, check, max_launches=8192):
    """Per-launch (us, flops, bytes) arrays of the profiled contraction launches."""
    us, fl, by = (np.zeros(max_launches) for _ in range(3))
    n = fn(h, kind, max_launches, dptr(us), dptr(fl), dptr(by))
    if n < 0:
        check(-n)
    return us[:n], fl[:n], by[:n]


# ------------------------------------------------------------------------------ state / solver
class Quantities:
    """Host mirror of StreamFcnQuantities (src/streamfcn/states.jl:9-17), read on access."""

    def __init__(self, case: DeviceCase):
        self._case = case

    @property
    def q(self):
        return self._case.q()

    @property
    def q0(self):
        return self._case.q0()

    @property
    def Γ(self):
        return self._case.gamma()

    Gamma = Γ

    @property
    def ψ(self):
        return self._case.psi()

    psi = ψ

    @property
    def Fb(self):
        return self._case.Fb()

    @property
    def stress(self):
        c = self._case
        fbs = c.fb_stress().reshape(2, c.nb).T
        cd, cl = c.CD(), c.CL()
        out, o = [], 0
        for k, b in enumerate(c.prob.bodies):
            n = b.npoints
            out.append(BodyStress(fbs[o:o + n].copy(), float(cd[k]), float(cl[k])))
            o += n
        return out


class State:
    """CNAB.State (src/streamfcn/cnab/states.jl:1-16) backed by a device case."""

    def __init__(self, prob: Problem, t: float = 0.0):
        self.case = DeviceCase(prob, t)
        self.qty = Quantities(self.case)
        self.cfl = 0.0
        self.t = float(t)

    @property
    def nonlin(self):
        return [self.case.nonlin(0), self.case.nonlin(1)]


class Solver:
    """CNAB.Solver (src/streamfcn/cnab/timestepping.jl:173-249): regularization and the force
    solver (dense Binv precomputed on the GPU, or CG)."""

    def __init__(self, prob: Problem, state: State):
        self.prob = prob
        self.case = state.case
        self.case.setup()


def quantities(state: State) -> Quantities:
    return state.qty


def initstate(prob: Problem, t0: float = 0.0) -> State:
    """initstate(prob, t0) — src/ImmersedBoundary.jl:51."""
    return State(prob, t0)


def initsolver(prob: Problem, state: State) -> Solver:
    """initsolver(prob, state) — src/ImmersedBoundary.jl:52."""
    return Solver(prob, state)


def advance_(state: State, solver: Solver, t: float) -> None:
    """advance!(state, solver, t) — src/streamfcn/cnab/timestepping.jl:253-290."""
    solver.case.advance(t)
    state.t = float(t)


SNAPSHOT_FIELDS = (("gamma", _lib.F_GAMMA), ("psi", _lib.F_PSI), ("q", _lib.F_Q),
                   ("nonlin1", _lib.F_NONLIN1), ("nonlin2", _lib.F_NONLIN2), ("Fb", _lib.F_FB))


def save_state(state: State, path) -> None:
    """Write a snapshot of a CNAB state to an .npz file: Γ, ψ, q (one column per level, the
    reference's matrices, streamfcn/states.jl:9-17), nonlin[1:2] in cycle order
    (cnab/states.jl:1-6), Fb, q0 and t, plus the grid shape for checking on load."""
    c = state.case
    np.savez(path, gamma=c.gamma(), psi=c.psi(), q=c.q(), q0=c.q0(), nonlin1=c.nonlin(0),
             nonlin2=c.nonlin(1), Fb=c.Fb(), t=np.float64(state.t),
             shape=np.array([c.nx, c.ny, c.nlev, c.nb], dtype=np.int64))


def load_state(state: State, path) -> None:
    """Restore a snapshot written by save_state into a state of the same problem shape (the
    device fields are overwritten; the solver's regularization and force solver are kept)."""
    c = state.case
    with np.load(path) as d:
        shape = tuple(int(v) for v in d["shape"])
        if shape != (c.nx, c.ny, c.nlev, c.nb):
            raise ValueError(f"snapshot of shape {shape} does not fit a case of shape "
                             f"{(c.nx, c.ny, c.nlev, c.nb)}")
        for name, which in SNAPSHOT_FIELDS:
            a = np.asarray(d[name], dtype=np.float64)
            if name == "Fb":
                if c.nb:
                    c.set_field(which, a)
                continue
            c.set_field(which, a.T.reshape(-1))  # levels stacked, as on the device
        state.t = float(d["t"])


class Batch:
    """Independent cases advanced together: every substep of advance! is one grouped launch
    over all cases (each with its own grid, levels, Re, dt and bodies) and the whole step is
    one HIP graph.  A parameter sweep that the reference runs as one solve per case
    (src/solve.jl:154-186) steps here as one batch.

    The states must come from initstate + initsolver of problems with fixed bodies (dense
    force solver) and one precision, and share the freestream: the first problem's
    freestream at its current time is used for every case.  `advance()` is one advance! of
    every case at its own t = t0 + k·dt; the states' fields read the batch's results after
    `sync()`.
    """

    def __init__(self, states):
        self.lib = _lib.load()
        self.states = list(states)
        if not self.states:
            raise ValueError("a batch needs at least one case")
        arr = (C.c_void_p * len(self.states))(*[st.case.h.value for st in self.states])
        h = C.c_void_p()
        rc = self.lib.ibh_batch_create(len(self.states), arr, C.byref(h))
        if rc != 0:
            raise _lib.IbhError(rc, self.lib.ibh_last_error(self.states[0].case.h).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise _lib.IbhError(rc, self.lib.ibh_batch_last_error(self.h).decode())

    def _freestream(self):
        st = self.states[0]
        return st.case.prob.fluid.freestream(st.t)

    def advance(self):
        ux, uy = self._freestream()
        self._check(self.lib.ibh_batch_advance(self.h, float(ux), float(uy)))
        for st in self.states:
            st.t = st.case.time()

    def advance_n(self, n: int):
        ux, uy = self._freestream()
        self._check(self.lib.ibh_batch_advance_n(self.h, int(n), float(ux), float(uy)))
        for st in self.states:
            st.t = st.case.time()

    def sync(self):
        self._check(self.lib.ibh_batch_sync(self.h))

    def launches_per_step(self) -> int:
        return self.lib.ibh_batch_launches_per_step(self.h)

    def profile(self, on: bool):
        self._check(self.lib.ibh_batch_profile_enable(self.h, int(on)))

    def profile_read(self, kind: int = 0, events: bool = False):
        ms, n, fl = C.c_double(), C.c_int64(), C.c_double()
        fn = self.lib.ibh_batch_profile_read_events if events else self.lib.ibh_batch_profile_read
        self._check(fn(self.h, kind, C.byref(ms), C.byref(n), C.byref(fl)))
        return ms.value, n.value, fl.value

    def profile_launches(self, kind: int = 0):
        """Per-launch (us, flops, bytes) of the profiled grouped contraction launches."""
        return _launch_records(self.lib.ibh_batch_profile_launches, self.h, kind, self._check)

    def tune_report(self) -> list:
        buf = C.create_string_buffer(1 << 16)
        self._check(self.lib.ibh_batch_tune_report(self.h, buf, len(buf)))
        return [ln for ln in buf.value.decode().splitlines() if ln]

    def close(self):
        if getattr(self, "h", None):
            self.lib.ibh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------ callbacks
class StopIteration_:
    """StopIteration() — returned by a callback to stop solve (src/solve.jl:1-6)."""

    def __eq__(self, other):
        return isinstance(other, StopIteration_)

    def __hash__(self):
        return 0


StopIteration = StopIteration_  # noqa: A001  (the reference's name)


@dataclass
class StateCallback:
    """StateCallback(f, timesteps, ret) — src/solve.jl:52-58."""

    f: Callable
    kind: str  # "all" | "indices" | "times"
    sel: Any = None
    ret: type = object


def each_timestep(f, ret=object):
    return StateCallback(f, "all", None, ret)


def at_times(f, t, ret=object):
    return StateCallback(f, "times", list(t), ret)


def at_indices(f, i, ret=object):
    return StateCallback(f, "indices", list(i), ret)


class StateResult(list):
    """StateResult{T} — src/solve.jl:145-157: data entries with their times in `.t`."""

    def __init__(self):
        super().__init__()
        self.t = []


def _caller(dt: float, cb: StateCallback):
    """_callback_caller, src/solve.jl:106-134."""
    if cb.kind == "all":
        return lambda i, state: cb.f(state)
    if cb.kind == "indices":
        pending = list(cb.sel)

        def call_idx(i, state):
            if pending and pending[0] == i:
                pending.pop(0)
                return cb.f(state)
            return None

        return call_idx
    pending_t = list(cb.sel)
    half = dt / 2

    def call_t(i, state):
        call = False
        while pending_t and pending_t[0] < state.t + half:
            pending_t.pop(0)
            call = True
        return cb.f(state) if call else None

    return call_t


def solve_(state: State, prob: Problem, tlims, save=(), call=()):
    """solve!(state, problem, (t0, tf); save, call) — src/solve.jl:188-215.

    Advances at every t of the inclusive range t0:dt:tf, then runs the save pushers and the
    callbacks; stops after a step where any callback returns StopIteration().
    """
    solver = initsolver(prob, state)
    dt = prob.scheme.dt
    ts = timesteps(float(tlims[0]), dt, float(tlims[1]))
    callers = [_caller(dt, cb) for cb in call]
    results = [StateResult() for _ in save]

    def pusher(cb, res):
        def push(st):
            u = cb.f(st)
            if u is not None:
                res.append(u)
                res.t.append(st.t)
            return None

        return _caller(dt, StateCallback(push, cb.kind, cb.sel, cb.ret))

    pushers = [pusher(cb, r) for cb, r in zip(save, results)]
    for i, t in enumerate(ts, start=1):
        advance_(state, solver, float(t))
        for p in pushers:
            p(i, state)
        stop = False
        for f in callers:
            stop |= f(i, state) == StopIteration_()
        if stop:
            break
    return results


def solve(prob: Problem, tlims, save=(), call=()):
    """solve(problem, (t0, tf); save, call) — src/solve.jl:231."""
    return solve_(initstate(prob, float(tlims[0])), prob, tlims, save=save, call=call)
