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
  E / Eᵀ                   adjointness and partition of unity; E of a linear velocity
                           sampled at the staggered faces returns it at every body point
                           (pins the support position and each component's half-cell offset)
  avg_flux!                linear fluxes average to their midpoint values
  direct_product!          on a sub-grid whose Γbc is the larger grid's Γ, the edge branches
                           reproduce the larger grid's interior products (L/R/B/T blocks)
  Vort2Flux                per level: coarsify! chain, −L₅ψ_ℓ = Γ_ℓ + edges(get_bc!(ψ_{ℓ+1})),
                           q_ℓ = curl!(ψ_ℓ, bc) divergence-free
  update_bodies!           r = r0 + R(θ)xb (= r0 + xb at θ = 0), v = dr/dt by central
                           differences, fixed bodies unchanged
  update_stress!           two mirror-image ellipses (y ↦ −y) in the symmetric domain get
                           equal CD and opposite CL, each away from zero (per-body slices of
                           Fb and the x/y component order)
  B, Binv                  B·Binv = I; B·z equals the composition of the pinned operators
  full step                no-slip after projection: E(q₁ + q₀) = 0 for fixed bodies
                           (= v·h for moving ones); uniform flow without bodies stays exact

Each pin was checked to fail under an injected mistake in a scratch copy of the oracle: a
one-cell shift of the delta support, a dropped half-cell offset, B↔T and L↔R swaps of Γbc, a
j+1 averaging index, a skipped Vort2Flux level, v := r and a flipped angular velocity.
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


# ------------------------------------------------------------------ pins of index patterns
def _face_coords(prob):
    """Coordinates of the flux unknowns (structure-ops.jl:48-57): qx[i, j] (1-based) at
    (x0 + (i−1)h, y0 + (j−½)h), qy[i, j] at (x0 + (i−½)h, y0 + (j−1)h)."""
    fl = prob.fluid
    base = fl.domain.base
    nx, ny, h, x0, y0 = fl.nx, fl.ny, base.h, base.x0, base.y0
    ix, jx = np.meshgrid(np.arange(nx + 1), np.arange(ny), indexing="xy")  # x fastest
    iy, jy = np.meshgrid(np.arange(nx), np.arange(ny + 1), indexing="xy")
    ux = (x0 + ix * h, y0 + (jx + 0.5) * h)
    vy = (x0 + (iy + 0.5) * h, y0 + jy * h)
    return ux, vy


def test_E_reproduces_linear_velocity_at_the_body_points():
    """E of q sampled from a linear field at the staggered face positions returns the field
    at each body point (zeroth and first moments of δh).  A one-cell shift of the support or
    of a component's half-cell offset moves the result by b·h or c·h/2."""
    prob = cylinder_case(n=64, nlev=2)
    o = ibref.OracleRun(prob, setup=False)
    rng = np.random.default_rng(11)
    nb = o.nb
    # points anywhere in the interior (not just the circle), all cell offsets
    r = np.concatenate([rng.uniform(-0.4, 2.4, nb), rng.uniform(-1.4, 1.4, nb)])
    o.reg_update(r)
    (xu, yu), (xv, yv) = _face_coords(prob)
    fu = lambda x, y: 0.7 + 1.9 * x - 2.3 * y  # noqa: E731
    fv = lambda x, y: -0.4 - 1.1 * x + 0.8 * y  # noqa: E731
    q = np.concatenate([fu(xu, yu).reshape(-1), fv(xv, yv).reshape(-1)])
    got = o.E(q)
    X, Y = r[:nb], r[nb:]
    scale = np.abs(q).max()
    assert np.abs(got[:nb] - fu(X, Y)).max() < 2e-8 * scale
    assert np.abs(got[nb:] - fv(X, Y)).max() < 2e-8 * scale


def test_avg_flux_of_linear_flux_is_the_midpoint_value():
    """avg_flux! (fluid-ops.jl:277-297): Qx[i, j] = (qx + q0x)(i, j − ½), Qy[i, j] =
    (qy + q0y)(i − ½, j) for linear fields of the face index (the /2 of four terms)."""
    nx, ny = 20, 16
    nu = (nx + 1) * ny
    i, j = np.meshgrid(np.arange(1, nx + 2), np.arange(1, ny + 1), indexing="xy")
    a, b = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 2), indexing="xy")
    fx = lambda i, j: 0.3 + 1.2 * i - 0.7 * j  # noqa: E731
    gx = lambda i, j: -1.1 + 0.4 * i + 2.1 * j  # noqa: E731
    fy = lambda i, j: 2.2 - 0.9 * i + 0.5 * j  # noqa: E731
    gy = lambda i, j: 0.6 + 1.7 * i + 0.2 * j  # noqa: E731
    q = np.concatenate([fx(i, j).reshape(-1), fy(a, b).reshape(-1)])
    q0 = np.concatenate([gx(i, j).reshape(-1), gy(a, b).reshape(-1)])
    Q = ibref.avg_flux(nx, ny, q, q0)
    Qx, Qy = split(Q, nx, ny)  # 0-based [i−1, j−1]
    I, J = np.meshgrid(np.arange(1, nx + 2), np.arange(2, ny + 1), indexing="ij")
    assert np.abs(Qx[:, 1:] - (fx(I, J - 0.5) + gx(I, J - 0.5))).max() < 1e-12
    I, J = np.meshgrid(np.arange(2, nx + 1), np.arange(1, ny + 2), indexing="ij")
    assert np.abs(Qy[1:, :] - (fy(I - 0.5, J) + gy(I - 0.5, J))).max() < 1e-12


def _sub_fields(big, nx2, ny2, ox, oy, nx, ny):
    """Restrict a flux vector, a vertex field and boundary values of an nx2×ny2 grid to the
    nx×ny grid whose lower-left vertex is big vertex (ox, oy)."""
    qx2, qy2 = split(big["q"], nx2, ny2)
    q0x2, q0y2 = split(big["q0"], nx2, ny2)
    sq = lambda ax, ay: np.concatenate([ax[ox:ox + nx + 1, oy:oy + ny].T.reshape(-1),  # noqa: E731
                                        ay[ox:ox + nx, oy:oy + ny + 1].T.reshape(-1)])
    V = big["V"]  # all vertices (nx2+1)×(ny2+1), including the big grid's boundary
    Vs = V[ox:ox + nx + 1, oy:oy + ny + 1]
    g = Vs[1:-1, 1:-1].T.reshape(-1)
    bc = np.concatenate([Vs[0, :], Vs[nx, :], Vs[:, 0], Vs[:, ny]])  # L, R, B, T
    return sq(qx2, qy2), sq(q0x2, q0y2), g, bc


def test_direct_product_edges_extend_the_interior_formula():
    """direct_product! (fluid-ops.jl:307-401): on a sub-grid whose Γbc holds the larger grid's
    Γ on the sub-grid's boundary vertices, the product equals the larger grid's on every
    face it computes, so each edge branch reads the right Γbc block (L, R, B, T) and index."""
    nx2, ny2 = 32, 28
    rng = np.random.default_rng(12)
    nq2 = (nx2 + 1) * ny2 + nx2 * (ny2 + 1)
    V = rng.standard_normal((nx2 + 1, ny2 + 1))
    big = {"q": rng.standard_normal(nq2), "q0": rng.standard_normal(nq2), "V": V}
    g2 = V[1:-1, 1:-1].T.reshape(-1)
    bc2 = np.concatenate([V[0, :], V[nx2, :], V[:, 0], V[:, ny2]])
    fq2 = ibref.direct_product(nx2, ny2, ibref.avg_flux(nx2, ny2, big["q"], big["q0"]), g2, bc2)
    fqx2, fqy2 = split(fq2, nx2, ny2)
    for ox, oy, nx, ny in ((6, 5, 16, 12), (1, 1, 20, 24), (12, 3, 20, 24)):
        q, q0, g, bc = _sub_fields(big, nx2, ny2, ox, oy, nx, ny)
        fq = ibref.direct_product(nx, ny, ibref.avg_flux(nx, ny, q, q0), g, bc)
        fqx, fqy = split(fq, nx, ny)
        # x fluxes i = 2..nx (all j), y fluxes j = 2..ny (all i), 1-based; the edge branches
        # list the two products in the other order, which changes the rounding of a fused
        # multiply-add, so equality holds to roundoff
        assert rel(fqx[1:nx, :], fqx2[ox + 1:ox + nx, oy:oy + ny]) < 1e-15
        assert rel(fqy[:, 1:ny], fqy2[ox:ox + nx, oy + 1:oy + ny]) < 1e-15
        assert np.count_nonzero(fqx[[0, nx], :]) == 0 and np.count_nonzero(fqy[:, [0, ny]]) == 0


def test_vort2flux_levels_telescope_by_poisson_residuals():
    """Vort2Flux (fluid-ops.jl:440-472), level by level with independent operators: the coarse
    centres hold the coarsify! chain of the finer Γ; the coarsest ψ solves −L₅ψ = Γ with zero
    boundary values; every finer ψ_ℓ solves −L₅ψ_ℓ = Γ_ℓ + edges(get_bc!(ψ_{ℓ+1})); q_ℓ is
    curl!(ψ_ℓ) with those boundary values and is divergence-free."""
    nx = ny = 32
    L = 4
    prob = _nobody(nx, L)
    o = ibref.OracleRun(prob, setup=False)
    m = p = nx - 1
    rng = np.random.default_rng(13)
    G0 = rng.standard_normal((o.ng, L))
    psi, q, G = o.vort2flux(G0)
    # coarse centres: coarsify! of the (already modified) finer level, pinned separately
    assert np.array_equal(G[:, 0], G0[:, 0])
    for lev in range(1, L):
        assert np.array_equal(G[:, lev], ibref.coarsify(nx, ny, G0[:, lev], G[:, lev - 1]))
    zero = np.zeros(o.nbc)
    for lev in range(L):
        bc = zero if lev == L - 1 else ibref.get_bc(nx, ny, psi[:, lev + 1])
        rhs = ibref.apply_bc(nx, ny, G[:, lev], bc, 1.0)
        assert rel(neg_lap5(psi[:, lev], m, p), rhs) < 1e-12, lev
        assert rel(q[:, lev], ibref.curl_bc(nx, ny, psi[:, lev], bc)) == 0, lev
        qx, qy = split(q[:, lev], nx, ny)
        div = qx[1:, :] - qx[:-1, :] + qy[:, 1:] - qy[:, :-1]
        assert np.abs(div).max() < 1e-12 * np.abs(q[:, lev]).max()


# ------------------------------------------------------------------ body kinematics
def test_body_points_kinematics():
    """update_bodies! (body-points.jl:47-76, with the frame's Ω and v from the kinematics):
    a RelativeFrame body's points are r0 + R(θ)·xb, so at θ = 0 they are r0 + xb, and their
    velocity is the time derivative of those positions (central difference, O(ε²)); bodies
    fixed in the grid keep xb with zero velocity; bodies are stacked in order, x block first."""
    prob = cylinder_plate_case(n=64, nlev=2)
    cyl, plate = prob.bodies
    nb, nc = prob.bodies.npoints, cyl.npoints
    fr = plate.frame
    # the plate's angle crosses zero nowhere (30° ± 10°): use a frame with θ(t0) = 0
    from ibhip.problem import RelativeFrame, RigidBody, BodyGroup, Problem, GridFrame

    om = 0.9
    spin = RelativeFrame(r=lambda t: (0.3 * t + 1.0, -0.2 * t * t), v=lambda t: (0.3, -0.4 * t),
                         theta=lambda t: math.sin(om * t), omega=lambda t: om * math.cos(om * t),
                         base=GridFrame())
    body = RigidBody(spin, plate.xb, plate.ds)
    prob2 = Problem(prob.fluid, BodyGroup(cyl, body), prob.scheme)
    r, v = prob2.body_points(0.0)
    npl = body.npoints
    assert np.array_equal(r[:nc], cyl.xb[:, 0]) and np.array_equal(r[nb:nb + nc], cyl.xb[:, 1])
    assert np.count_nonzero(v[:nc]) == 0 and np.count_nonzero(v[nb:nb + nc]) == 0
    assert np.abs(r[nc:nb] - (1.0 + plate.xb[:, 0])).max() < 1e-15
    assert np.abs(r[nb + nc:] - plate.xb[:, 1]).max() < 1e-15
    for pb in (prob, prob2):
        for t in (0.0, 0.37, 2.1):
            eps = 1e-5
            rp, _ = pb.body_points(t + eps)
            rm, _ = pb.body_points(t - eps)
            _, vt = pb.body_points(t)
            assert np.abs((rp - rm) / (2 * eps) - vt).max() < 1e-8
    assert fr is plate.frame and npl == plate.npoints


def test_update_stress_mirror_bodies():
    # update_stress! (streamfcn/states.jl:53-84) sums each body's slice of Fb: a mirror pair
    # in the y-symmetric domain must get equal drag and opposite lift.  A wrong body offset or
    # swapped force components breaks the symmetry.
    from ibhip.problem import GridFrame, RigidBody
    from ibhip.workloads import ellipse

    n = 64
    h = 4.0 / n
    upper = ellipse((0.2, 0.7), 0.4, 0.15, 0.5, h)
    lower = RigidBody(GridFrame(), upper.xb * np.array([1.0, -1.0]), upper.ds)
    grid = UniformGrid(h, (-1.0, 3.0), (-2.0, 2.0))
    prob = Problem(StreamFcnFluid(MultiLevel(grid, 2), (1.0, 0.0), 100.0),
                   BodyGroup(upper, lower), SchemeCNAB(0.2 * h, force_solver="dense"))
    o = ibref.OracleRun(prob)
    for k in range(3):
        o.advance(k * prob.scheme.dt)
    cd, cl = np.asarray(o.CD()), np.asarray(o.CL())
    assert cd.shape == (2,) and cd[0] > 0 and abs(cl[0]) > 1e-3 * cd[0]
    assert abs(cd[0] - cd[1]) < 1e-9 * cd[0]
    assert abs(cl[0] + cl[1]) < 1e-9 * cd[0]
