"""ibh_op_lapinv_noalloc (the LaplacianInv apply that allocates nothing, for callers that cannot
allocate device memory) against ibh_op_lapinv on the same inputs: two calls in a row on two
contexts, the second with a smaller grid, bitwise equal (same kernels, and every tile variant
gives the same bits); grids beyond the scratch limit are rejected."""
import ctypes as C

import numpy as np
import pytest

from cases import cylinder_case

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyKind


class DeviceArray:
    """Device memory for the test's inputs and outputs, from the HIP runtime that libibhip
    itself links (loaded with it), so that no second runtime enters the process."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), 8 * self.n, H2D) == 0

    def get(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, 8 * self.n, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def hip_runtime():
    from ibhip import _lib

    _lib.load()  # brings in the runtime libibhip links
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def test_lapinv_noalloc_matches_lapinv():
    import ibhip
    from ibhip import _lib

    hip = hip_runtime()
    L = _lib.load()
    cases = []
    for n, nlev in ((256, 3), (64, 2)):  # the second call has the smaller input
        prob = cylinder_case(n=n, nlev=nlev)
        cases.append(ibhip.DeviceCase(prob))
    rng = np.random.default_rng(21)
    for case in cases:
        for mode in (0, 2):
            lev = case.nlev - 1
            x = rng.standard_normal(case.ng)
            want = case.lapinv(mode, lev, x)
            xin, out = DeviceArray(hip, case.ng), DeviceArray(hip, case.ng)
            xin.put(x)
            out.put(np.full(case.ng, np.nan))
            print(f"noalloc: {case.nx}x{case.ny} mode {mode} level {lev}", flush=True)
            rc = L.ibh_op_lapinv_noalloc(case.h, mode, lev, xin.p, out.p)
            assert rc == 0, L.ibh_last_error(case.h).decode()
            assert L.ibh_sync(case.h) == 0
            got = out.get()
            xin.free()
            out.free()
            assert np.array_equal(got, want), np.abs(got - want).max()


def test_lapinv_noalloc_rejects_grids_beyond_the_limit():
    import ibhip
    from ibhip import _lib
    from ibhip.problem import BodyGroup, MultiLevel, Problem, SchemeCNAB, StreamFcnFluid, UniformGrid

    hip = hip_runtime()
    L = _lib.load()
    n = 1028  # 1027² interior points > 1024²
    grid = UniformGrid(1.0 / n, (0.0, 1.0), (0.0, 1.0))
    prob = Problem(StreamFcnFluid(MultiLevel(grid, 1), (1.0, 0.0), 100.0), BodyGroup(),
                   SchemeCNAB(0.001))
    case = ibhip.DeviceCase(prob)
    assert case.ng > _lib.NOALLOC_MAX_NG
    xin, out = DeviceArray(hip, case.ng), DeviceArray(hip, case.ng)
    rc = L.ibh_op_lapinv_noalloc(case.h, 0, 0, xin.p, out.p)
    print(L.ibh_last_error(case.h).decode())
    xin.free()
    out.free()
    assert rc == _lib.EINVAL
