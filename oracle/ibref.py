"""ctypes driver for the CPU FP64 oracle (oracle/libibref.so).

TEST INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's CPU-baseline leg use it
as the checker; the product path never imports it.

`OracleRun(problem)` mirrors the reference's initstate / initsolver / advance! sequence
(src/ImmersedBoundary.jl:51-52, src/solve.jl:188-215) on the C restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libibref.so")

F_Q, F_Q0, F_GAMMA, F_PSI, F_FB, F_NL1, F_NL2, F_CD, F_CL, F_FBS, F_R, F_V, F_BINV, F_QS, F_GS = range(15)


class Desc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nlev", C.c_int32),
        ("h", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
        ("Re", C.c_double), ("dt", C.c_double), ("beta", C.c_double * 2),
        ("nbodies", C.c_int32),
        ("npts", C.POINTER(C.c_int32)), ("moving", C.POINTER(C.c_int32)),
        ("xb", C.POINTER(C.c_double)), ("ds", C.POINTER(C.c_double)),
        ("force_mode", C.c_int32), ("cg_maxiter", C.c_int32), ("cg_reltol", C.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.ibr_create.restype = vp
        L.ibr_create.argtypes = [C.POINTER(Desc)]
        L.ibr_destroy.argtypes = [vp]
        L.ibr_last_error.restype = C.c_char_p
        L.ibr_last_error.argtypes = [vp]
        L.ibr_init_state.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.ibr_setup_solver.argtypes = [vp]
        L.ibr_set_body_points.argtypes = [vp, dp, dp]
        L.ibr_advance.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.ibr_field.restype = dp
        L.ibr_field.argtypes = [vp, C.c_int]
        L.ibr_cg_iters.argtypes = [vp]
        L.ibr_dims.argtypes = [vp, C.POINTER(C.c_int32)]
        L.ibr_dst2.argtypes = [C.c_int, C.c_int, dp, dp]
        L.ibr_lap_eigs.argtypes = [C.c_int, C.c_int, dp]
        L.ibr_lapinv.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        for nm in ("ibr_curl_plain", "ibr_rot", "ibr_coarsify", "ibr_get_bc"):
            getattr(L, nm).argtypes = [C.c_int, C.c_int, dp, dp]
        L.ibr_curl_bc.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.ibr_apply_bc.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double]
        L.ibr_avg_flux.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.ibr_direct_product.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp]
        L.ibr_nonlinear.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp]
        L.ibr_vort2flux.argtypes = [vp, dp, dp, dp]
        L.ibr_deltah.restype = C.c_double
        L.ibr_deltah.argtypes = [C.c_double, C.c_double, C.c_double]
        L.ibr_reg_update.argtypes = [vp, dp]
        L.ibr_E.argtypes = [vp, dp, dp]
        L.ibr_ET.argtypes = [vp, dp, dp]
        L.ibr_B_times.argtypes = [vp, dp, dp]
        L.ibr_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# ----------------------------------------------------------------------- stateless operators
def dst2(x: np.ndarray) -> np.ndarray:
    """RODFT00 along both axes of an (m, p) array given in the reference layout (x fastest)."""
    m, p = x.shape
    a = f64(x.T)  # (p, m) C-order == (m, p) Fortran order
    out = np.empty_like(a)
    lib().ibr_dst2(m, p, ptr(a), ptr(out))
    return out.T.copy()


def lap_eigs(nx, ny):
    out = np.empty((ny - 1) * (nx - 1))
    lib().ibr_lap_eigs(nx, ny, ptr(out))
    return out


def _sizes(nx, ny):
    nq = (nx + 1) * ny + nx * (ny + 1)
    ng = (nx - 1) * (ny - 1)
    nbc = 2 * (nx + 1) + 2 * (ny + 1)
    return nq, ng, nbc


def curl_plain(nx, ny, psi):
    nq, ng, _ = _sizes(nx, ny)
    q = np.zeros(nq)
    lib().ibr_curl_plain(nx, ny, ptr(f64(psi)), ptr(q))
    return q


def rot(nx, ny, q):
    nq, ng, _ = _sizes(nx, ny)
    g = np.zeros(ng)
    lib().ibr_rot(nx, ny, ptr(f64(q)), ptr(g))
    return g


def curl_bc(nx, ny, psi, psibc):
    nq, ng, _ = _sizes(nx, ny)
    q = np.zeros(nq)
    lib().ibr_curl_bc(nx, ny, ptr(f64(psi)), ptr(f64(psibc)), ptr(q))
    return q


def coarsify(nx, ny, gc, g):
    gc = f64(gc).copy()
    lib().ibr_coarsify(nx, ny, ptr(gc), ptr(f64(g)))
    return gc


def get_bc(nx, ny, r):
    _, _, nbc = _sizes(nx, ny)
    rbc = np.zeros(nbc)
    lib().ibr_get_bc(nx, ny, ptr(rbc), ptr(f64(r)))
    return rbc


def apply_bc(nx, ny, r, rbc, fac):
    r = f64(r).copy()
    lib().ibr_apply_bc(nx, ny, ptr(r), ptr(f64(rbc)), float(fac))
    return r


def avg_flux(nx, ny, q, q0):
    nq, _, _ = _sizes(nx, ny)
    Q = np.zeros(nq)
    lib().ibr_avg_flux(nx, ny, ptr(Q), ptr(f64(q)), ptr(f64(q0)))
    return Q


def direct_product(nx, ny, Q, g, gbc):
    nq, _, _ = _sizes(nx, ny)
    fq = np.zeros(nq)
    lib().ibr_direct_product(nx, ny, ptr(fq), ptr(f64(Q)), ptr(f64(g)), ptr(f64(gbc)))
    return fq


def deltah(rf, rb, dr):
    return lib().ibr_deltah(float(rf), float(rb), float(dr))


# ----------------------------------------------------------------------- stateful case
FORCE = {"auto": 0, "dense": 1, "cg": 2, "pcg": 2}  # the oracle solves every CG variant with plain cg!


class OracleRun:
    """The reference's State + Solver on the CPU restatement for one Problem."""

    def __init__(self, prob, t0: float = 0.0, setup: bool = True):
        L = lib()
        self.prob = prob
        fl = prob.fluid
        base = fl.domain.base
        sch = prob.scheme
        self._npts = np.array([b.npoints for b in prob.bodies], dtype=np.int32)
        self._mov = np.array([int(b.moving) for b in prob.bodies], dtype=np.int32)
        xb = prob.stacked_points
        self._xb = f64(xb.T.reshape(-1)) if xb.size else np.zeros(1)
        self._ds = f64(prob.stacked_ds) if xb.size else np.zeros(1)
        d = Desc()
        d.nx, d.ny, d.nlev = fl.nx, fl.ny, fl.nlevels
        d.h, d.x0, d.y0 = base.h, base.x0, base.y0
        d.Re, d.dt = fl.Re, sch.dt
        d.beta[0], d.beta[1] = sch.beta
        d.nbodies = len(prob.bodies)
        d.npts = self._npts.ctypes.data_as(C.POINTER(C.c_int32))
        d.moving = self._mov.ctypes.data_as(C.POINTER(C.c_int32))
        d.xb = ptr(self._xb)
        d.ds = ptr(self._ds)
        d.force_mode = FORCE[sch.force_solver]
        d.cg_maxiter = sch.cg_maxiter
        d.cg_reltol = sch.cg_reltol
        self._desc = d
        h = L.ibr_create(C.byref(d))
        if not h:
            raise ValueError("oracle rejected the case description")
        self.h = h
        dims = (C.c_int32 * 8)()
        L.ibr_dims(h, dims)
        self.nx, self.ny, self.nlev, self.nq, self.ng, self.nb, self.nbc, self.nbodies = list(dims)
        ux, uy = fl.freestream(t0)
        L.ibr_init_state(h, t0, ux, uy)
        self.t = t0
        if setup:
            self.setup()

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().ibr_last_error(self.h).decode())

    def setup(self):
        self._check(lib().ibr_setup_solver(self.h))

    def advance(self, t: float):
        prob = self.prob
        if prob.any_moving:
            r, v = prob.body_points(t)
            r, v = f64(r), f64(v)
            lib().ibr_set_body_points(self.h, ptr(r), ptr(v))
        ux, uy = prob.fluid.freestream(t)
        self._check(lib().ibr_advance(self.h, float(t), float(ux), float(uy)))
        self.t = t

    def field(self, which: int, n: int) -> np.ndarray:
        p = lib().ibr_field(self.h, which)
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    # quantities (streamfcn/states.jl:9-17) as (n, nlev) arrays, column = level
    def q(self):
        return self.field(F_Q, self.nq * self.nlev).reshape(self.nlev, self.nq).T

    def q0(self):
        return self.field(F_Q0, self.nq * self.nlev).reshape(self.nlev, self.nq).T

    def gamma(self):
        return self.field(F_GAMMA, self.ng * self.nlev).reshape(self.nlev, self.ng).T

    def psi(self):
        return self.field(F_PSI, self.ng * self.nlev).reshape(self.nlev, self.ng).T

    def nonlin(self, k):
        return self.field(F_NL1 if k == 0 else F_NL2, self.ng * self.nlev).reshape(self.nlev, self.ng).T

    def Fb(self):
        return self.field(F_FB, 2 * self.nb)

    def CD(self):
        return self.field(F_CD, self.nbodies)

    def CL(self):
        return self.field(F_CL, self.nbodies)

    def fb_stress(self):
        return self.field(F_FBS, 2 * self.nb)

    def binv(self):
        n = 2 * self.nb
        return self.field(F_BINV, n * n).reshape(n, n).T

    def cg_iters(self):
        return lib().ibr_cg_iters(self.h)

    # operators bound to this case
    def lapinv(self, mode: int, lev: int, b):
        out = np.zeros(self.ng)
        self._check(lib().ibr_lapinv(self.h, mode, lev, ptr(f64(b)), ptr(out)))
        return out

    def reg_update(self, r):
        self._check(lib().ibr_reg_update(self.h, ptr(f64(r))))

    def E(self, q):
        out = np.zeros(2 * self.nb)
        self._check(lib().ibr_E(self.h, ptr(f64(q)), ptr(out)))
        return out

    def ET(self, fb):
        out = np.zeros(self.nq)
        self._check(lib().ibr_ET(self.h, ptr(f64(fb)), ptr(out)))
        return out

    def B_times(self, z):
        out = np.zeros(2 * self.nb)
        self._check(lib().ibr_B_times(self.h, ptr(f64(z)), ptr(out)))
        return out

    def nonlinear(self, lev, q, q0, g, gbc):
        out = np.zeros(self.ng)
        self._check(lib().ibr_nonlinear(self.h, lev, ptr(f64(q)), ptr(f64(q0)), ptr(f64(g)),
                                        ptr(f64(gbc)), ptr(out)))
        return out

    def vort2flux(self, g):
        g = f64(np.asarray(g).T.reshape(-1)).copy()  # (ng, nlev) -> level-major flat
        psi = np.zeros(self.ng * self.nlev)
        q = np.zeros(self.nq * self.nlev)
        self._check(lib().ibr_vort2flux(self.h, ptr(psi), ptr(q), ptr(g)))
        return (psi.reshape(self.nlev, self.ng).T, q.reshape(self.nlev, self.nq).T,
                g.reshape(self.nlev, self.ng).T)

    def close(self):
        if self.h:
            lib().ibr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
