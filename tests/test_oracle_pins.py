"""Pins of the CPU oracle against closed forms, invariants and library routines.

The reference ships no tests or golden vectors, so each restated function is checked here
against what the mathematics fixes, chosen so that a dropped term, a wrong sign or index or a
transposed operand fails:

  DST-I (RODFT00)          scipy.fft.dstn(type=1) and the dense 4·S·X·S closed form
  LaplacianInv Δinv/A/Ainv an independent 5-point Laplacian: (−L₅)x = b, A = I − α(−L₅),
                           (I + α(−L₅))·Ainv(b) = b
  curl!/rot!               ⟨Cψ, q⟩ = ⟨ψ, C′q⟩ and C′C = −L₅;  curl! with boundary data is
                           divergence-free cell by cell and vanishes on a constant ψ
  coarsify!/get_bc!        linear fields: full weighting gives 4×f at the coincident point,
                           boundary interpolation gives 0.25×f at the fine boundary vertex
  nonlinear term           uniform flow over a single vorticity spike gives the centred
                           −u·∂ω/∂x (and −v·∂ω/∂y) pattern; Γ = 0 gives 0
  δh (Yang3)               zeroth/first moments, symmetry, continuity, support 3h (to the
                           1e-9 that the reference's 10-digit literals allow)
  E / Eᵀ                   adjointness and partition of unity
  B, Binv                  B·Binv = I; B·z equals the composition of the pinned operators
  full step                no-slip after projection: E(q₁ + q₀) = 0 for fixed bodies
                           (= v·h for moving ones); uniform flow without bodies stays exact
"""
import math

import numpy as np
import pytest
import scipy.fft as sf

import ibref
from cases import cylinder_case, cylinder_plate_case
from ibhip.problem import BodyGroup, MultiLevel, Problem, SchemeCNAB, StreamFcnFluid, UniformGrid


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(b).max()
    return np.abs(a - b).max() / (d if d > 0 else 1.0)


# ------------------------------------------------------------------ independent helpers
def neg_lap5(x, m, p):
    """−L₅ x on the m×p interior with zero Dirichlet values outside (x fastest)."""
    X = x.reshape(p, m).T
    P = np.zeros((m + 2, p + 2))
    P[1:-1, 1:-1] = X
    Y = 4 * P[1:-1, 1:-1] - P[2:, 1:-1] - P[:-2, 1:-1] - P[1:-1, 2:] - P[1:-1, :-2]
    return Y.T.reshape(-1)


def sine_matrix(n):
    k = np.arange(1, n)
    return np.sin(np.pi * np.outer(k, k) / n)


def split(q, nx, ny):
    nu = (nx + 1) * ny
    return q[:nu].reshape(ny, nx + 1).T, q[nu:].reshape(ny + 1, nx).T  # [i, j], 0-based


def gam2d(g, nx, ny):
    return g.reshape(ny - 1, nx - 1).T


# ------------------------------------------------------------------ DST
@pytest.mark.parametrize("m,p", [(7, 7), (15, 11), (199, 199), (63, 47), (255, 383)])
def test_dst2_scipy_and_closed_form(m, p):
    x = np.random.default_rng(m * p).standard_normal((m, p))
    got = ibref.dst2(x)
    assert rel(got, sf.dstn(x, type=1)) < 1e-13
    if m < 300:
        want = 4 * sine_matrix(m + 1) @ x @ sine_matrix(p + 1)
        assert rel(got, want) < 1e-13


def oracle_case(n=32, nlev=3, **kw):
    prob = cylinder_case(n=n, nlev=nlev, **kw) if n >= 64 else _nobody(n, nlev)
    return prob, ibref.OracleRun(prob, setup=False)


def _nobody(n, nlev, Re=100.0, dt=0.01):
    h = 1.0 / n
    grid = UniformGrid(h, (0.0, 1.0), (0.0, 1.0))
    return Problem(StreamFcnFluid(MultiLevel(grid, nlev), (1.0, 0.0), Re), BodyGroup(),
                   SchemeCNAB(dt))


# ------------------------------------------------------------------ LaplacianInv
def test_lapinv_inverts_five_point_laplacian():
    prob, o = oracle_case(32, 3)
    m = p = 31
    b = np.random.default_rng(0).standard_normal(m * p)
    x = o.lapinv(0, 0, b)
    assert rel(neg_lap5(x, m, p), b) < 1e-12


@pytest.mark.parametrize("lev", [0, 1, 2])
def test_A_and_Ainv_are_crank_nicolson_halves(lev):
    prob, o = oracle_case(32, 3)
    m = p = 31
    hc = prob.fluid.domain.gridstep(lev + 1)
    alpha = prob.scheme.dt / (2 * prob.fluid.Re * hc * hc)
    b = np.random.default_rng(lev).standard_normal(m * p)
    # A_ℓ = I − α(−L₅) (explicit half), cnab/ops.jl:26
    assert rel(o.lapinv(1, lev, b), b - alpha * neg_lap5(b, m, p)) < 1e-12
    # Ainv_ℓ solves (I + α(−L₅)) x = b (implicit half), cnab/ops.jl:29
    x = o.lapinv(2, lev, b)
    assert rel(x + alpha * neg_lap5(x, m, p), b) < 1e-12


# ------------------------------------------------------------------ curl / rot
def test_curl_rot_adjoint_and_laplacian():
    nx, ny = 24, 20
    nq = (nx + 1) * ny + nx * (ny + 1)
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((nx - 1) * (ny - 1))
    q = rng.standard_normal(nq)
    Cpsi = ibref.curl_plain(nx, ny, psi)
    assert abs(Cpsi @ q - psi @ ibref.rot(nx, ny, q)) < 1e-12 * np.abs(q).sum()
    assert rel(ibref.rot(nx, ny, Cpsi), neg_lap5(psi, nx - 1, ny - 1)) < 1e-13


def test_curl_bc_divergence_free_and_constant():
    nx, ny = 24, 20
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((nx - 1) * (ny - 1))
    # boundary data taken from one vertex field, so each corner vertex has one value
    # (L/R run over y, B/T over x; get_bc! produces such data)
    V = rng.standard_normal((nx + 1, ny + 1))
    bc = np.concatenate([V[0, :], V[nx, :], V[:, 0], V[:, ny]])
    qx, qy = split(ibref.curl_bc(nx, ny, psi, bc), nx, ny)
    div = qx[1:, :] - qx[:-1, :] + qy[:, 1:] - qy[:, :-1]
    assert np.abs(div).max() < 1e-13
    # constant streamfunction with matching boundary values has zero flux
    q = ibref.curl_bc(nx, ny, np.full_like(psi, 2.5), np.full_like(bc, 2.5))
    assert np.abs(q).max() == 0.0
    # zero boundary data reproduces the plain curl except on the outermost flux columns
    q0 = ibref.curl_bc(nx, ny, psi, np.zeros_like(bc))
    q1 = ibref.curl_plain(nx, ny, psi)
    a0, b0 = split(q0, nx, ny)
    a1, b1 = split(q1, nx, ny)
    assert rel(a0[1:-1, :], a1[1:-1, :]) == 0 and rel(b0[:, 1:-1], b1[:, 1:-1]) == 0


# ------------------------------------------------------------------ multilevel transfers
def _linear_gamma(nx, ny, a, bx, cy, spacing):
    i = np.arange(1, nx)[:, None]
    j = np.arange(1, ny)[None, :]
    return (a + bx * i * spacing + cy * j * spacing).T.reshape(-1)


def test_coarsify_full_weighting_of_linear_field():
    nx, ny = 32, 24
    f = lambda x, y: 0.3 + 1.7 * x - 0.6 * y  # noqa: E731  (fine vertex coordinates, h = 1)
    i = np.arange(1, nx)[:, None]
    j = np.arange(1, ny)[None, :]
    g = f(i, j).T.reshape(-1).astype(float)
    gc0 = np.full((nx - 1) * (ny - 1), -9.0)
    gc = gam2d(ibref.coarsify(nx, ny, gc0, g), nx, ny)
    for ic in range(1, nx):
        for jc in range(1, ny):
            inside = nx // 4 + 1 <= ic <= 3 * nx // 4 - 1 and ny // 4 + 1 <= jc <= 3 * ny // 4 - 1
            if inside:  # coarse vertex ic sits on fine vertex 2ic − nx/2 (levels share the centre)
                want = 4 * f(2 * ic - nx // 2, 2 * jc - ny // 2)
                assert abs(gc[ic - 1, jc - 1] - want) < 1e-12
            else:
                assert gc[ic - 1, jc - 1] == -9.0


def test_get_bc_interpolates_linear_field_to_fine_boundary():
    nx, ny = 32, 24
    # coarse vertex i ↔ fine coordinate 2i − nx/2 (fine vertex index, h = 1)
    f = lambda x, y: 1.1 - 0.4 * x + 2.3 * y  # noqa: E731
    i = np.arange(1, nx)[:, None]
    j = np.arange(1, ny)[None, :]
    r = f(2 * i - nx // 2, 2 * j - ny // 2).T.reshape(-1).astype(float)
    rbc = ibref.get_bc(nx, ny, r)
    Lo, Ro, Bo, To = 0, ny + 1, 2 * (ny + 1), 2 * (ny + 1) + nx + 1
    k = np.arange(nx + 1)
    assert rel(rbc[Bo:Bo + nx + 1], 0.25 * f(k, 0)) < 1e-14
    assert rel(rbc[To:To + nx + 1], 0.25 * f(k, ny)) < 1e-14
    k = np.arange(ny + 1)
    assert rel(rbc[Lo:Lo + ny + 1], 0.25 * f(0, k)) < 1e-14
    assert rel(rbc[Ro:Ro + ny + 1], 0.25 * f(nx, k)) < 1e-14


def test_apply_bc_edges():
    nx, ny = 16, 12
    rbc = np.arange(2 * (nx + 1) + 2 * (ny + 1), dtype=float) + 1
    r = gam2d(ibref.apply_bc(nx, ny, np.zeros((nx - 1) * (ny - 1)), rbc, 2.0), nx, ny)
    Lo, Ro, Bo, To = 0, ny + 1, 2 * (ny + 1), 2 * (ny + 1) + nx + 1
    # the reference reads rbc[(B+1)+i] for i = 1..nx-1, i.e. the interior edge vertices
    assert r[3, 0] == 2.0 * rbc[Bo + 4] and r[3, -1] == 2.0 * rbc[To + 4]
    assert r[0, 5] == 2.0 * rbc[Lo + 6] and r[-1, 5] == 2.0 * rbc[Ro + 6]
    assert r[0, 0] == 2.0 * (rbc[Bo + 1] + rbc[Lo + 1])
    assert np.count_nonzero(r[1:-1, 1:-1]) == 0


# ------------------------------------------------------------------ nonlinear term
def test_nonlinear_uniform_flow_advects_spike():
    prob = _nobody(32, 2)
    o = ibref.OracleRun(prob, setup=False)
    nx = ny = 32
    nq = (nx + 1) * ny + nx * (ny + 1)
    nu = (nx + 1) * ny
    for lev in (0, 1):
        hc = prob.fluid.domain.gridstep(lev + 1)
        for U, V in ((1.3, 0.0), (0.0, -0.7)):
            q0 = np.concatenate([np.full(nu, U * hc), np.full(nq - nu, V * hc)])
            g = np.zeros((nx - 1) * (ny - 1))
            a0, b0 = 12, 9  # spike at Γ[a0, b0] (0-based)
            g[a0 + (nx - 1) * b0] = 1.0
            N = gam2d(o.nonlinear(lev, np.zeros(nq), q0, g, np.zeros(2 * (nx + 1) + 2 * (ny + 1))), nx, ny)
            want = np.zeros_like(N)
            # −(u ∂ω/∂x + v ∂ω/∂y) by centred differences, with Γ = ω·hc² ⇒ N = −u(Γ[a+1]−Γ[a−1])/(2hc)
            want[a0 - 1, b0] -= U / (2 * hc)
            want[a0 + 1, b0] += U / (2 * hc)
            want[a0, b0 - 1] -= V / (2 * hc)
            want[a0, b0 + 1] += V / (2 * hc)
            assert np.abs(N - want).max() < 1e-12 / hc
    g = np.zeros((nx - 1) * (ny - 1))
    q = np.random.default_rng(1).standard_normal(nq)
    assert np.abs(o.nonlinear(0, q, q, g, np.zeros(2 * (nx + 1) + 2 * (ny + 1)))).max() == 0.0


# ------------------------------------------------------------------ regularized delta
def test_deltah_yang3_properties():
    h = 0.37
    d = np.vectorize(lambda r: ibref.deltah(r, 0.0, h))
    for X in np.linspace(0.0, 1.0, 17):
        xs = (np.arange(-5, 6) - X) * h
        w = d(xs)
        assert abs(w.sum() - 1.0) < 1e-8
        assert abs((xs * w).sum()) < 1e-8 * h
    r = np.linspace(0, 3.5, 71) * h
    assert np.allclose(d(r), d(-r), rtol=0, atol=0)
    assert np.all(d(np.array([3.01, 3.5, 5.0]) * h) == 0.0)
    for r1 in (1.0, 2.0, 3.0):
        lo, hi = d((r1 - 1e-12) * h), d((r1 + 1e-12) * h)
        assert abs(lo - hi) < 1e-8


def test_E_ET_adjoint_and_partition_of_unity():
    prob = cylinder_case(n=64, nlev=2)
    o = ibref.OracleRun(prob, setup=False)
    r = prob.stacked_points.T.reshape(-1).copy()
    o.reg_update(r)
    rng = np.random.default_rng(4)
    q = rng.standard_normal(o.nq)
    f = rng.standard_normal(2 * o.nb)
    assert abs(o.E(q) @ f - q @ o.ET(f)) < 1e-11 * np.abs(q).sum()
    ones = np.ones(o.nq)
    assert np.abs(o.E(ones) - 1.0).max() < 1e-8


# ------------------------------------------------------------------ Schur complement
def test_B_composition_and_dense_inverse():
    prob = cylinder_case(n=64, nlev=2, spacing=1.0)
    o = ibref.OracleRun(prob)  # builds the dense Binv
    nf = 2 * o.nb
    z = np.random.default_rng(5).standard_normal(nf)
    # B·z = E·[vort2flux(Ainv₁·C′·Eᵀ·z)]₁ from the separately pinned operators
    g1 = o.lapinv(2, 0, ibref.rot(o.nx, o.ny, o.ET(z)))
    G = np.zeros((o.ng, o.nlev))
    G[:, 0] = g1
    psi, q, _ = o.vort2flux(G)
    assert rel(o.B_times(z), o.E(q[:, 0])) < 1e-13
    Binv = o.binv()
    Bm = np.stack([o.B_times(e) for e in np.eye(nf)], axis=1)
    assert np.abs(Bm @ Binv - np.eye(nf)).max() < 1e-8


# ------------------------------------------------------------------ full step
def test_uniform_flow_without_bodies_stays_uniform():
    prob = _nobody(32, 3)
    o = ibref.OracleRun(prob)
    for k in range(3):
        o.advance(k * prob.scheme.dt)
    assert np.abs(o.gamma()).max() == 0.0 and np.abs(o.q()).max() == 0.0


def test_step_enforces_no_slip_fixed_body():
    prob = cylinder_case(n=64, nlev=3)
    o = ibref.OracleRun(prob)
    for k in range(4):
        o.advance(k * prob.scheme.dt)
        slip = o.E(o.q()[:, 0] + o.q0()[:, 0])
        free = o.E(o.q0()[:, 0])
        assert np.abs(slip).max() < 1e-9 * np.abs(free).max()
    cd, cl = o.CD()[0], o.CL()[0]
    assert cd > 0 and abs(cl) < 1e-9 * cd  # symmetric body in a uniform stream


def test_step_enforces_body_velocity_moving_plate():
    prob = cylinder_plate_case(n=64, nlev=2, cg_reltol=1e-15, cyl_spacing=1.0)
    o = ibref.OracleRun(prob)
    for k in range(2):
        t = k * prob.scheme.dt
        o.advance(t)
        r, v = prob.body_points(t)
        slip = o.E(o.q()[:, 0] + o.q0()[:, 0]) - v * prob.fluid.domain.base.h
        assert np.abs(slip).max() < 1e-8 * np.abs(o.E(o.q0()[:, 0])).max()
