/*
 * ibref.c — CPU FP64 restatement of ImmersedBoundary.jl's CNAB immersed-boundary projection
 * step, written from the reference's Julia source.  TEST INFRASTRUCTURE ONLY (the parity
 * oracle and the CPU baseline); the product library never links it.
 *
 * Each function names the reference file:line it restates.  Index macros below take the
 * reference's 1-based (i, j) so that every loop reads like the Julia it follows.
 *
 * Third-party arithmetic restated here (not part of the reference tree):
 *   - FFTW 3.3.10 RODFT00 (DST-I), FFTW.jl 1.5.0: computed through a mixed-radix complex FFT
 *     of length 2(m+1) on the odd extension (FFTW's documented RODFT00 definition).
 *   - IterativeSolvers.jl 0.9.2 `cg!` (unpreconditioned, warm start, tolerance
 *     max(reltol·‖b−A·x0‖, abstol), stop test before each iteration).
 *   - LinearAlgebra `inv` (LAPACK getrf/getri): LU with partial pivoting, then the inverse.
 */
#include "ibref.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------------------------
 * Mixed-radix complex FFT (decimation in time), used only for DST-I.
 * ---------------------------------------------------------------------------------------- */
typedef struct {
    int n;
    int nfac;
    int fac[64]; /* pairs (radix, remaining length) */
    cplx *tw;    /* exp(-2πik/n) */
} fftplan;

static void fft_plan_init(fftplan *p, int n) {
    p->n = n;
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->tw[k] = cos(a) + I * sin(a);
    }
    int rem = n, nf = 0;
    int radices[] = {4, 2, 3, 5};
    while (rem > 1) {
        int r = 0;
        for (int t = 0; t < 4; ++t)
            if (rem % radices[t] == 0) { r = radices[t]; break; }
        if (!r) {
            for (int d = 7; d <= rem; d += 2)
                if (rem % d == 0) { r = d; break; }
        }
        rem /= r;
        p->fac[2 * nf] = r;
        p->fac[2 * nf + 1] = rem;
        ++nf;
    }
    p->nfac = nf;
}

static void fft_plan_free(fftplan *p) { free(p->tw); p->tw = NULL; }

static void fft_rec(cplx *out, const cplx *in, size_t fstride, const int *fac, const fftplan *p,
                    cplx *scratch) {
    const int r = fac[0], m = fac[1];
    if (m == 1) {
        for (int k = 0; k < r; ++k) out[k] = in[(size_t)k * fstride];
    } else {
        for (int k = 0; k < r; ++k)
            fft_rec(out + (size_t)k * m, in + (size_t)k * fstride, fstride * (size_t)r, fac + 2, p,
                    scratch);
    }
    const cplx *tw = p->tw;
    const int n = p->n;
    if (r == 2) {
        for (int u = 0; u < m; ++u) {
            cplx t = out[u + m] * tw[(size_t)u * fstride];
            out[u + m] = out[u] - t;
            out[u] = out[u] + t;
        }
    } else if (r == 4) {
        for (int u = 0; u < m; ++u) {
            cplx a0 = out[u];
            cplx a1 = out[u + m] * tw[(size_t)u * fstride];
            cplx a2 = out[u + 2 * m] * tw[(size_t)2 * u * fstride];
            cplx a3 = out[u + 3 * m] * tw[(size_t)3 * u * fstride];
            cplx s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = a1 - a3;
            cplx md13 = -I * d13; /* forward transform: multiply by -i */
            out[u] = s02 + s13;
            out[u + m] = d02 + md13;
            out[u + 2 * m] = s02 - s13;
            out[u + 3 * m] = d02 - md13;
        }
    } else {
        for (int u = 0; u < m; ++u) {
            for (int q1 = 0; q1 < r; ++q1) scratch[q1] = out[u + q1 * m];
            for (int q1 = 0; q1 < r; ++q1) {
                int k = u + q1 * m;
                cplx acc = scratch[0];
                size_t twidx = 0;
                for (int q = 1; q < r; ++q) {
                    twidx += fstride * (size_t)k;
                    twidx %= (size_t)n;
                    acc += scratch[q] * tw[twidx];
                }
                out[k] = acc;
            }
        }
    }
}

static void fft_exec(const fftplan *p, const cplx *in, cplx *out, cplx *scratch) {
    fft_rec(out, in, 1, p->fac, p, scratch);
}

/* RODFT00 of two real length-m sequences x, y (strides sx, sy) at once, in place:
 * y_k = 2 Σ_j x_j sin(π(j+1)(k+1)/(m+1)).  z = xodd + i·yodd, Z = FFT(z):
 * DST(x)_{k-1} = -Im Z_k,  DST(y)_{k-1} = Re Z_k. */
static void dst1_pair(const fftplan *p, int m, double *x, size_t sx, double *y, size_t sy,
                      cplx *zin, cplx *zout, cplx *scratch) {
    const int n = m + 1, N = 2 * n;
    zin[0] = 0;
    zin[n] = 0;
    for (int j = 1; j <= m; ++j) {
        double xr = x[(size_t)(j - 1) * sx];
        double yi = y ? y[(size_t)(j - 1) * sy] : 0.0;
        zin[j] = xr + I * yi;
        zin[N - j] = -(xr + I * yi);
    }
    fft_exec(p, zin, zout, scratch);
    for (int k = 1; k <= m; ++k) {
        x[(size_t)(k - 1) * sx] = -cimag(zout[k]);
        if (y) y[(size_t)(k - 1) * sy] = creal(zout[k]);
    }
}

/* 2-D RODFT00 along both dimensions (FFTW.plan_r2r(b, RODFT00, [1,2]),
 * fluid-ops.jl:25,49,51).  in and out may alias. */
static void dst2_planned(const fftplan *pm, const fftplan *pp, int m, int p, const double *in,
                         double *out) {
    if (in != out) memcpy(out, in, sizeof(double) * (size_t)m * (size_t)p);
    /* along x (contiguous columns of length m) */
#pragma omp parallel
    {
        cplx *zin = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)(m + 1));
        cplx *zout = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)(m + 1));
        cplx *scr = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)(m + 1));
#pragma omp for schedule(static)
        for (int jj = 0; jj < (p + 1) / 2; ++jj) {
            int j0 = 2 * jj, j1 = 2 * jj + 1;
            dst1_pair(pm, m, out + (size_t)m * j0, 1, j1 < p ? out + (size_t)m * j1 : NULL, 1, zin,
                      zout, scr);
        }
        free(zin);
        free(zout);
        free(scr);
    }
    /* along y (stride m) */
#pragma omp parallel
    {
        cplx *zin = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)(p + 1));
        cplx *zout = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)(p + 1));
        cplx *scr = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)(p + 1));
#pragma omp for schedule(static)
        for (int ii = 0; ii < (m + 1) / 2; ++ii) {
            int i0 = 2 * ii, i1 = 2 * ii + 1;
            dst1_pair(pp, p, out + i0, (size_t)m, i1 < m ? out + i1 : NULL, (size_t)m, zin, zout,
                      scr);
        }
        free(zin);
        free(zout);
        free(scr);
    }
}

void ibr_dst2(int m, int p, const double *in, double *out) {
    fftplan pm, pp;
    fft_plan_init(&pm, 2 * (m + 1));
    fft_plan_init(&pp, 2 * (p + 1));
    dst2_planned(&pm, &pp, m, p, in, out);
    fft_plan_free(&pm);
    fft_plan_free(&pp);
}

void ibr_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------
 * Index helpers, 1-based like the reference (streamfcn/fluids.jl:58-98).
 * ---------------------------------------------------------------------------------------- */
#define GI(a, i, j) (a)[((i)-1) + (size_t)m * ((j)-1)]           /* Γ/ψ (nx-1)×(ny-1) */
#define QX(a, i, j) (a)[((i)-1) + (size_t)(nx + 1) * ((j)-1)]    /* qx (nx+1)×ny      */
#define QY(a, i, j) (a)[nu + ((i)-1) + (size_t)nx * ((j)-1)]     /* qy nx×(ny+1)      */
#define BCV(a, k) (a)[(k)-1]

#define DOMAIN_DECL                                                                  \
    const int m = nx - 1, p = ny - 1;                                                \
    const size_t nu = (size_t)(nx + 1) * ny;                                         \
    const int Lo = 0, Ro = ny + 1, Bo = 2 * (ny + 1), To = 2 * (ny + 1) + nx + 1;    \
    (void)m; (void)p; (void)nu; (void)Lo; (void)Ro; (void)Bo; (void)To;

/* lap_eigs, fluid-ops.jl:58-61 */
void ibr_lap_eigs(int nx, int ny, double *lam) {
    const int m = nx - 1, p = ny - 1;
    for (int j = 1; j <= p; ++j)
        for (int i = 1; i <= m; ++i)
            lam[(i - 1) + (size_t)m * (j - 1)] =
                -2.0 * (cos(M_PI * i / nx) + cos(M_PI * j / ny) - 2.0);
}

/* curl!(q, ψ, domain) without boundary data, fluid-ops.jl:79-110 (the adjoint of rot!). */
void ibr_curl_plain(int nx, int ny, const double *psi, double *q) {
    DOMAIN_DECL
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx; ++i) QX(q, i, j) = GI(psi, i - 1, j) - GI(psi, i - 1, j - 1);
    for (int i = 2; i <= nx; ++i) QX(q, i, 1) = GI(psi, i - 1, 1);
    for (int i = 2; i <= nx; ++i) QX(q, i, ny) = -GI(psi, i - 1, ny - 1);
    for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx - 1; ++i) QY(q, i, j) = GI(psi, i - 1, j - 1) - GI(psi, i, j - 1);
    for (int j = 2; j <= ny; ++j) QY(q, 1, j) = -GI(psi, 1, j - 1);
    for (int j = 2; j <= ny; ++j) QY(q, nx, j) = GI(psi, nx - 1, j - 1);
}

/* rot!(Γ, q, domain), fluid-ops.jl:117-129 */
void ibr_rot(int nx, int ny, const double *q, double *gam) {
    DOMAIN_DECL
    for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx; ++i)
            GI(gam, i - 1, j - 1) = (QX(q, i, j - 1) - QX(q, i, j)) + (QY(q, i, j) - QY(q, i - 1, j));
}

/* curl!(q, ψ, ψbc, domain), fluid-ops.jl:138-182 */
void ibr_curl_bc(int nx, int ny, const double *psi, const double *pbc, double *q) {
    DOMAIN_DECL
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx; ++i) QX(q, i, j) = GI(psi, i - 1, j) - GI(psi, i - 1, j - 1);
    for (int i = 2; i <= nx; ++i) QX(q, i, 1) = GI(psi, i - 1, 1) - BCV(pbc, Bo + i);
    for (int i = 2; i <= nx; ++i) QX(q, i, ny) = BCV(pbc, i + To) - GI(psi, i - 1, ny - 1);
    for (int j = 1; j <= ny; ++j) QX(q, 1, j) = BCV(pbc, (Lo + 1) + j) - BCV(pbc, Lo + j);
    for (int j = 1; j <= ny; ++j) QX(q, nx + 1, j) = BCV(pbc, (Ro + 1) + j) - BCV(pbc, Ro + j);
    for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx - 1; ++i) QY(q, i, j) = GI(psi, i - 1, j - 1) - GI(psi, i, j - 1);
    for (int j = 2; j <= ny; ++j) QY(q, 1, j) = BCV(pbc, Lo + j) - GI(psi, 1, j - 1);
    for (int j = 2; j <= ny; ++j) QY(q, nx, j) = GI(psi, nx - 1, j - 1) - BCV(pbc, Ro + j);
    for (int i = 1; i <= nx; ++i) QY(q, i, 1) = BCV(pbc, Bo + i) - BCV(pbc, (Bo + 1) + i);
    for (int i = 1; i <= nx; ++i) QY(q, i, ny + 1) = BCV(pbc, To + i) - BCV(pbc, (To + 1) + i);
}

/* coarsify!(Γc, Γ, domain), fluid-ops.jl:184-202.  Only the centre of Γc is written. */
void ibr_coarsify(int nx, int ny, double *gc, const double *g) {
    DOMAIN_DECL
    /* js = ny÷2 .+ ((-ny÷2+2):2:(ny÷2-2)), jcs = ny÷2 .+ ((-ny÷4+1):(ny÷4-1)), likewise i */
    int nj = 0;
    for (int t = -ny / 2 + 2; t <= ny / 2 - 2; t += 2) ++nj;
    int ni = 0;
    for (int t = -nx / 2 + 2; t <= nx / 2 - 2; t += 2) ++ni;
    for (int b = 0; b < nj; ++b) {
        int j = ny / 2 + (-ny / 2 + 2) + 2 * b;
        int jc = ny / 2 + (-ny / 4 + 1) + b;
        for (int a = 0; a < ni; ++a) {
            int i = nx / 2 + (-nx / 2 + 2) + 2 * a;
            int ic = nx / 2 + (-nx / 4 + 1) + a;
            GI(gc, ic, jc) = GI(g, i, j) +
                             0.5 * (GI(g, i + 1, j) + GI(g, i, j + 1) + GI(g, i - 1, j) +
                                    GI(g, i, j - 1)) +
                             0.25 * (GI(g, i + 1, j + 1) + GI(g, i + 1, j - 1) +
                                     GI(g, i - 1, j - 1) + GI(g, i - 1, j + 1));
        }
    }
}

/* get_bc!(rbc, r, domain), fluid-ops.jl:210-239 (including the final ×0.25). */
void ibr_get_bc(int nx, int ny, double *rbc, const double *r) {
    DOMAIN_DECL
    for (int t = 0; t <= nx / 2; ++t) {
        int i = nx / 4 + t, ibc = 1 + 2 * t;
        BCV(rbc, Bo + ibc) = GI(r, i, ny / 4);
        BCV(rbc, To + ibc) = GI(r, i, 3 * ny / 4);
    }
    for (int t = 1; t <= nx / 2; ++t) {
        int i = nx / 4 + t, ibc = 2 * t;
        BCV(rbc, Bo + ibc) = 0.5 * (GI(r, i, ny / 4) + GI(r, i - 1, ny / 4));
        BCV(rbc, To + ibc) = 0.5 * (GI(r, i, 3 * ny / 4) + GI(r, i - 1, 3 * ny / 4));
    }
    for (int t = 0; t <= ny / 2; ++t) {
        int j = ny / 4 + t, jbc = 1 + 2 * t;
        BCV(rbc, Lo + jbc) = GI(r, nx / 4, j);
        BCV(rbc, Ro + jbc) = GI(r, 3 * nx / 4, j);
    }
    for (int t = 1; t <= ny / 2; ++t) {
        int j = ny / 4 + t, jbc = 2 * t;
        BCV(rbc, Lo + jbc) = 0.5 * (GI(r, nx / 4, j) + GI(r, nx / 4, j - 1));
        BCV(rbc, Ro + jbc) = 0.5 * (GI(r, 3 * nx / 4, j) + GI(r, 3 * nx / 4, j - 1));
    }
    const int nbc = 2 * (nx + 1) + 2 * (ny + 1);
    for (int k = 0; k < nbc; ++k) rbc[k] *= 0.25;
}

/* apply_bc!(r, rbc, fac, domain), fluid-ops.jl:249-275 */
void ibr_apply_bc(int nx, int ny, double *r, const double *rbc, double fac) {
    DOMAIN_DECL
    for (int i = 1; i <= nx - 1; ++i) GI(r, i, 1) += fac * BCV(rbc, (Bo + 1) + i);
    for (int i = 1; i <= nx - 1; ++i) GI(r, i, ny - 1) += fac * BCV(rbc, (To + 1) + i);
    for (int j = 1; j <= ny - 1; ++j) GI(r, 1, j) += fac * BCV(rbc, (Lo + 1) + j);
    for (int j = 1; j <= ny - 1; ++j) GI(r, nx - 1, j) += fac * BCV(rbc, (Ro + 1) + j);
}

/* avg_flux!(Q, qty, domain, lev), fluid-ops.jl:277-297 (q, q0 are that level's slices) */
void ibr_avg_flux(int nx, int ny, double *Q, const double *q, const double *q0) {
    DOMAIN_DECL
    for (int j = 2; j <= ny; ++j)
        for (int i = 1; i <= nx + 1; ++i)
            QX(Q, i, j) = (QX(q, i, j) + QX(q, i, j - 1) + QX(q0, i, j) + QX(q0, i, j - 1)) / 2;
    for (int j = 1; j <= ny + 1; ++j)
        for (int i = 2; i <= nx; ++i)
            QY(Q, i, j) = (QY(q, i, j) + QY(q, i - 1, j) + QY(q0, i, j) + QY(q0, i - 1, j)) / 2;
}

/* direct_product!(fq, Q, Γ, Γbc, domain), fluid-ops.jl:307-401 */
void ibr_direct_product(int nx, int ny, double *fq, const double *Q, const double *g,
                        const double *gbc) {
    DOMAIN_DECL
    const size_t nq = nu + (size_t)nx * (ny + 1);
    memset(fq, 0, sizeof(double) * nq);
    for (int j = 2; j <= ny - 1; ++j)
        for (int i = 2; i <= nx; ++i)
            QX(fq, i, j) = (QY(Q, i, j + 1) * GI(g, i - 1, j) + QY(Q, i, j) * GI(g, i - 1, j - 1)) / 2;
    for (int i = 2; i <= nx; ++i) {
        int j = 1;
        QX(fq, i, j) = (QY(Q, i, j + 1) * GI(g, i - 1, j) + QY(Q, i, j) * BCV(gbc, Bo + i)) / 2;
    }
    for (int i = 2; i <= nx; ++i) {
        int j = ny;
        QX(fq, i, j) = (QY(Q, i, j) * GI(g, i - 1, j - 1) + QY(Q, i, j + 1) * BCV(gbc, To + i)) / 2;
    }
    for (int j = 2; j <= ny; ++j)
        for (int i = 2; i <= nx - 1; ++i)
            QY(fq, i, j) =
                -(QX(Q, i + 1, j) * GI(g, i, j - 1) + QX(Q, i, j) * GI(g, i - 1, j - 1)) / 2;
    for (int j = 2; j <= ny; ++j) {
        int i = 1;
        QY(fq, i, j) = -(QX(Q, i + 1, j) * GI(g, i, j - 1) + QX(Q, i, j) * BCV(gbc, Lo + j)) / 2;
    }
    for (int j = 2; j <= ny; ++j) {
        int i = nx;
        QY(fq, i, j) = -(QX(Q, i, j) * GI(g, i - 1, j - 1) + QX(Q, i + 1, j) * BCV(gbc, Ro + j)) / 2;
    }
}

/* δh, structure-ops.jl:114-171: Yang3 smoothed delta (Yang et al., JCP 2009) times h.
 * The coefficients are the reference's literals. */
double ibr_deltah(double rf, double rb, double dr) {
    double r = fabs(rf - rb);
    double r1 = r / dr;
    double r2 = r1 * r1;
    double r3 = r2 * r1;
    double r4 = r3 * r1;
    if (r1 <= 1.0) {
        double a5 = asin((1.0 / 2.0) * sqrt(3.0) * (2.0 * r1 - 1.0));
        double a8 = sqrt(1.0 - 12.0 * r2 + 12.0 * r1);
        return 4.166666667e-2 * r4 + (-0.1388888889 + 3.472222222e-2 * a8) * r3 +
               (-7.121664902e-2 - 5.208333333e-2 * a8 + 0.2405626122 * a5) * r2 +
               (-0.2405626122 * a5 - 0.3792313933 + 0.1012731481 * a8) * r1 +
               8.0187537413e-2 * a5 - 4.195601852e-2 * a8 + 0.6485698427;
    } else if (r1 <= 2.0) {
        double a6 = asin((1.0 / 2.0) * sqrt(3.0) * (-3.0 + 2.0 * r1));
        double a9 = sqrt(-23.0 + 36.0 * r1 - 12.0 * r2);
        return -6.250000000e-2 * r4 + (0.4861111111 - 1.736111111e-2 * a9) * r3 +
               (-1.143175026 + 7.812500000e-2 * a9 - 0.1202813061 * a6) * r2 +
               (0.8751991178 + 0.3608439183 * a6 - 0.1548032407 * a9) * r1 - 0.2806563809 * a6 +
               8.22848104e-3 + 0.1150173611 * a9;
    } else if (r1 <= 3.0) {
        double a1 = asin((1.0 / 2.0 * (2.0 * r1 - 5.0)) * sqrt(3.0));
        double a7 = sqrt(-71.0 - 12.0 * r2 + 60.0 * r1);
        return 2.083333333e-2 * r4 + (3.472222222e-3 * a7 - 0.2638888889) * r3 +
               (1.214391675 - 2.604166667e-2 * a7 + 2.405626122e-2 * a1) * r2 +
               (-0.1202813061 * a1 - 2.449273192 + 7.262731481e-2 * a7) * r1 +
               0.1523563211 * a1 + 1.843201677 - 7.306134259e-2 * a7;
    }
    return 0.0;
}

/* ------------------------------------------------------------------------------------------
 * Case: the Solver + State of cnab/timestepping.jl:173-249, cnab/states.jl:1-16,
 * streamfcn/states.jl:9-41.
 * ---------------------------------------------------------------------------------------- */
#define SUPP 6
#define NS (2 * SUPP + 1)

struct ibr_case {
    int nx, ny, nlev, m, p;
    size_t nq, ng, nbc;
    double h, x0, y0, Re, dt, beta[2];
    int nbodies, nb;
    int *npts, *moving, *boff;
    double *ds;
    int force_mode, use_cg, cg_maxiter, cg_last_iters;
    double cg_reltol;
    fftplan pm, pp;
    double *lam;              /* Λ */
    double *sc_A, *sc_Ainv;   /* per level: scale/Λ' for A_ℓ and Ainv_ℓ, nlev·nΓ */
    double *sc_D;             /* scale/Λ for Δinv */
    /* state */
    double t;
    double *q, *q0, *gam, *psi, *Fb, *nl[2];
    double *r, *v;            /* points nb×2 */
    double *xb0;              /* initial stacked points */
    double *CD, *CL, *fbs;    /* per-body coefficients, stress fb nb×2 */
    /* Reg, structure-ops.jl:4-9 */
    int *bidx;                /* nb×2 */
    double *wt;               /* nb×2×13×13 (only nb rows, see structure-ops.jl:31) */
    /* solver work */
    double *qs, *gs, *gbc, *gtmpL, *psitmp, *qtmp, *Ftmp, *g1, *g2, *g3, *q1, *q2, *work;
    double *Binv;
    double *cg_r, *cg_c, *cg_u;
    char err[256];
};

const char *ibr_last_error(ibr_case *c) { return c ? c->err : "null case"; }

static void *zalloc(size_t n) { return calloc(n ? n : 1, sizeof(double)); }

ibr_case *ibr_create(const ibr_desc *d) {
    if (d->nx % 4 || d->ny % 4 || d->nx < 8 || d->ny < 8 || d->nlev < 1) return NULL;
    ibr_case *c = (ibr_case *)calloc(1, sizeof(ibr_case));
    c->nx = d->nx;
    c->ny = d->ny;
    c->nlev = d->nlev;
    c->m = d->nx - 1;
    c->p = d->ny - 1;
    c->nq = (size_t)(d->nx + 1) * d->ny + (size_t)d->nx * (d->ny + 1);
    c->ng = (size_t)c->m * c->p;
    c->nbc = 2 * (size_t)(d->nx + 1) + 2 * (size_t)(d->ny + 1);
    c->h = d->h;
    c->x0 = d->x0;
    c->y0 = d->y0;
    c->Re = d->Re;
    c->dt = d->dt;
    c->beta[0] = d->beta[0];
    c->beta[1] = d->beta[1];
    c->nbodies = d->nbodies;
    c->npts = (int *)calloc(d->nbodies + 1, sizeof(int));
    c->moving = (int *)calloc(d->nbodies + 1, sizeof(int));
    c->boff = (int *)calloc(d->nbodies + 1, sizeof(int));
    int nb = 0, anymove = 0;
    for (int b = 0; b < d->nbodies; ++b) {
        c->npts[b] = d->npts[b];
        c->moving[b] = d->moving ? d->moving[b] : 0;
        anymove |= c->moving[b];
        c->boff[b] = nb;
        nb += d->npts[b];
    }
    c->boff[d->nbodies] = nb;
    c->nb = nb;
    c->ds = (double *)zalloc(nb);
    memcpy(c->ds, d->ds, sizeof(double) * nb);
    c->force_mode = d->force_mode;
    c->use_cg = d->force_mode == 2 || (d->force_mode == 0 && anymove);
    c->cg_maxiter = d->cg_maxiter > 0 ? d->cg_maxiter : 5000;
    c->cg_reltol = d->cg_reltol > 0 ? d->cg_reltol : 1e-12;
    fft_plan_init(&c->pm, 2 * (c->m + 1));
    fft_plan_init(&c->pp, 2 * (c->p + 1));
    const size_t ng = c->ng, nq = c->nq, L = c->nlev;
    c->lam = (double *)zalloc(ng);
    ibr_lap_eigs(c->nx, c->ny, c->lam);
    c->sc_A = (double *)zalloc(ng * L);
    c->sc_Ainv = (double *)zalloc(ng * L);
    c->sc_D = (double *)zalloc(ng);
    const double scale = 1.0 / (4.0 * c->nx * c->ny); /* fluid-ops.jl:37 */
    for (size_t k = 0; k < ng; ++k) c->sc_D[k] = scale / c->lam[k];
    for (int lev = 1; lev <= c->nlev; ++lev) {
        /* A_Ainv_linearmaps, cnab/ops.jl:14-33 */
        double hc = c->h * pow(2.0, lev - 1); /* gridstep, fluids/fluids.jl:102 */
        for (size_t k = 0; k < ng; ++k) {
            double lexpl = 1.0 / (1.0 - c->lam[k] * c->dt / (2.0 * c->Re * hc * hc));
            double limpl = 1.0 + c->lam[k] * c->dt / (2.0 * c->Re * hc * hc);
            c->sc_A[(lev - 1) * ng + k] = scale / lexpl;
            c->sc_Ainv[(lev - 1) * ng + k] = scale / limpl;
        }
    }
    c->q = (double *)zalloc(nq * L);
    c->q0 = (double *)zalloc(nq * L);
    c->gam = (double *)zalloc(ng * L);
    c->psi = (double *)zalloc(ng * L);
    c->Fb = (double *)zalloc(2 * (size_t)nb);
    c->nl[0] = (double *)zalloc(ng * L);
    c->nl[1] = (double *)zalloc(ng * L);
    c->r = (double *)zalloc(2 * (size_t)nb);
    c->v = (double *)zalloc(2 * (size_t)nb);
    c->xb0 = (double *)zalloc(2 * (size_t)nb);
    memcpy(c->xb0, d->xb, sizeof(double) * 2 * nb);
    memcpy(c->r, d->xb, sizeof(double) * 2 * nb); /* BodyPoints(body), body-points.jl:13-26 */
    c->CD = (double *)zalloc(d->nbodies);
    c->CL = (double *)zalloc(d->nbodies);
    c->fbs = (double *)zalloc(2 * (size_t)nb);
    c->bidx = (int *)calloc(2 * (size_t)nb + 1, sizeof(int));
    c->wt = (double *)zalloc((size_t)nb * 2 * NS * NS);
    c->qs = (double *)zalloc(nq * L);
    c->gs = (double *)zalloc(ng * L);
    c->gbc = (double *)zalloc(c->nbc);
    c->gtmpL = (double *)zalloc(ng * L);
    c->psitmp = (double *)zalloc(ng * L);
    c->qtmp = (double *)zalloc(nq * L);
    c->Ftmp = (double *)zalloc(2 * (size_t)nb);
    c->g1 = (double *)zalloc(ng);
    c->g2 = (double *)zalloc(ng);
    c->g3 = (double *)zalloc(ng);
    c->q1 = (double *)zalloc(nq);
    c->q2 = (double *)zalloc(nq);
    c->work = (double *)zalloc(ng);
    c->cg_r = (double *)zalloc(2 * (size_t)nb);
    c->cg_c = (double *)zalloc(2 * (size_t)nb);
    c->cg_u = (double *)zalloc(2 * (size_t)nb);
    return c;
}

void ibr_destroy(ibr_case *c) {
    if (!c) return;
    fft_plan_free(&c->pm);
    fft_plan_free(&c->pp);
    void *ptrs[] = {c->npts, c->moving, c->boff, c->ds, c->lam, c->sc_A, c->sc_Ainv, c->sc_D,
                    c->q, c->q0, c->gam, c->psi, c->Fb, c->nl[0], c->nl[1], c->r, c->v, c->xb0,
                    c->CD, c->CL, c->fbs, c->bidx, c->wt, c->qs, c->gs, c->gbc, c->gtmpL,
                    c->psitmp, c->qtmp, c->Ftmp, c->g1, c->g2, c->g3, c->q1, c->q2, c->work,
                    c->Binv, c->cg_r, c->cg_c, c->cg_u};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
    free(c);
}

int ibr_dims(ibr_case *c, int32_t *o) {
    o[0] = c->nx; o[1] = c->ny; o[2] = c->nlev; o[3] = (int32_t)c->nq; o[4] = (int32_t)c->ng;
    o[5] = c->nb; o[6] = (int32_t)c->nbc; o[7] = c->nbodies;
    return 0;
}

int ibr_cg_iters(ibr_case *c) { return c->cg_last_iters; }

double *ibr_field(ibr_case *c, int which) {
    switch (which) {
    case IBR_F_Q: return c->q;
    case IBR_F_Q0: return c->q0;
    case IBR_F_GAMMA: return c->gam;
    case IBR_F_PSI: return c->psi;
    case IBR_F_FB: return c->Fb;
    case IBR_F_NONLIN1: return c->nl[0];
    case IBR_F_NONLIN2: return c->nl[1];
    case IBR_F_CD: return c->CD;
    case IBR_F_CL: return c->CL;
    case IBR_F_FBSTRESS: return c->fbs;
    case IBR_F_POINTS_R: return c->r;
    case IBR_F_POINTS_V: return c->v;
    case IBR_F_BINV: return c->Binv;
    case IBR_F_QS: return c->qs;
    case IBR_F_GAMMAS: return c->gs;
    }
    return NULL;
}

/* LaplacianInv apply, fluid-ops.jl:42-56: x = DST(DST(b) .* (scale/Λ')) */
static void lapinv_apply(ibr_case *c, const double *scl, const double *b, double *x) {
    double *w = c->work;
    dst2_planned(&c->pm, &c->pp, c->m, c->p, b, w);
    for (size_t k = 0; k < c->ng; ++k) w[k] *= scl[k];
    dst2_planned(&c->pm, &c->pp, c->m, c->p, w, x);
}

int ibr_lapinv(ibr_case *c, int mode, int lev, const double *b, double *x) {
    if (lev < 0 || lev >= c->nlev) return -1;
    const double *scl = mode == 0 ? c->sc_D : mode == 1 ? c->sc_A + lev * c->ng
                                                        : c->sc_Ainv + lev * c->ng;
    lapinv_apply(c, scl, b, x);
    return 0;
}

/* base_flux!, fluid-ops.jl:403-417 (LabFrame) */
static void base_flux(ibr_case *c, double ux, double uy) {
    const int nx = c->nx, ny = c->ny;
    const size_t nu = (size_t)(nx + 1) * ny;
    for (int lev = 1; lev <= c->nlev; ++lev) {
        double hc = c->h * pow(2.0, lev - 1);
        double *q0 = c->q0 + (lev - 1) * c->nq;
        for (size_t k = 0; k < nu; ++k) q0[k] = ux * hc;
        for (size_t k = nu; k < c->nq; ++k) q0[k] = uy * hc;
    }
}

int ibr_init_state(ibr_case *c, double t0, double ux, double uy) {
    /* StreamFcnQuantities(prob, t), streamfcn/states.jl:19-41; State, cnab/states.jl:8-16 */
    const size_t L = c->nlev;
    memset(c->q, 0, sizeof(double) * c->nq * L);
    memset(c->gam, 0, sizeof(double) * c->ng * L);
    memset(c->psi, 0, sizeof(double) * c->ng * L);
    memset(c->Fb, 0, sizeof(double) * 2 * c->nb);
    memset(c->nl[0], 0, sizeof(double) * c->ng * L);
    memset(c->nl[1], 0, sizeof(double) * c->ng * L);
    memcpy(c->r, c->xb0, sizeof(double) * 2 * c->nb);
    memset(c->v, 0, sizeof(double) * 2 * c->nb);
    base_flux(c, ux, uy);
    c->t = t0;
    return 0;
}

/* Reg update!, structure-ops.jl:37-61 */
int ibr_reg_update(ibr_case *c, const double *r) {
    const int nb = c->nb;
    const double h = c->h, x0 = c->x0, y0 = c->y0;
    for (int k = 0; k < nb; ++k) {
        c->bidx[k] = (int)floor((r[k] - x0) / h);
        c->bidx[nb + k] = (int)floor((r[nb + k] - y0) / h);
    }
    for (int k = 0; k < nb; ++k) {
        double xs[NS], ys[NS];
        for (int s = 0; s < NS; ++s) {
            xs[s] = h * (c->bidx[k] - 1 + (s - SUPP)) + x0;
            ys[s] = h * (c->bidx[nb + k] - 1 + (s - SUPP)) + y0;
        }
        double *w1 = c->wt + ((size_t)k * 2 + 0) * NS * NS;
        double *w2 = c->wt + ((size_t)k * 2 + 1) * NS * NS;
        for (int t = 0; t < NS; ++t)
            for (int s = 0; s < NS; ++s) {
                w1[s + NS * t] = ibr_deltah(xs[s], r[k], h) * ibr_deltah(ys[t] + h / 2, r[nb + k], h);
                w2[s + NS * t] = ibr_deltah(xs[s] + h / 2, r[k], h) * ibr_deltah(ys[t], r[nb + k], h);
            }
    }
    /* bounds: the 13×13 support must stay inside the finest flux arrays (reference: BoundsError) */
    for (int k = 0; k < nb; ++k) {
        int bx = c->bidx[k], by = c->bidx[nb + k];
        if (bx - SUPP < 1 || bx + SUPP > c->nx || by - SUPP < 1 || by + SUPP > c->ny) {
            snprintf(c->err, sizeof c->err, "body point %d support leaves the finest grid", k + 1);
            return -2;
        }
    }
    return 0;
}

/* Eᵀ: (reg::Reg)(q, fb), structure-ops.jl:63-86 */
int ibr_ET(ibr_case *c, const double *fb, double *q) {
    const int nx = c->nx, ny = c->ny, nb = c->nb;
    const size_t nu = (size_t)(nx + 1) * ny;
    (void)ny;
    memset(q, 0, sizeof(double) * c->nq);
    for (int k = 0; k < nb; ++k) {
        const double *w1 = c->wt + ((size_t)k * 2 + 0) * NS * NS;
        const double *w2 = c->wt + ((size_t)k * 2 + 1) * NS * NS;
        int bx = c->bidx[k], by = c->bidx[nb + k];
        double ss = 0;
        for (int t = 0; t < NS; ++t)
            for (int s = 0; s < NS; ++s) {
                int i = bx + s - SUPP, j = by + t - SUPP;
                QX(q, i, j) += w1[s + NS * t] * fb[k];
                QY(q, i, j) += w2[s + NS * t] * fb[nb + k];
            }
        for (int t = 0; t < NS; ++t)
            for (int s = 0; s < NS; ++s) {
                double vv = QX(q, bx + s - SUPP, by + t - SUPP);
                ss += vv * vv;
            }
        if (!isfinite(ss)) {
            snprintf(c->err, sizeof c->err, "infinite flux");
            return -3;
        }
    }
    return 0;
}

/* E: (regT::RegT)(fb, q), structure-ops.jl:88-105 */
int ibr_E(ibr_case *c, const double *q, double *fb) {
    const int nx = c->nx, nb = c->nb;
    const size_t nu = (size_t)(nx + 1) * c->ny;
    for (int k = 0; k < nb; ++k) {
        const double *w1 = c->wt + ((size_t)k * 2 + 0) * NS * NS;
        const double *w2 = c->wt + ((size_t)k * 2 + 1) * NS * NS;
        int bx = c->bidx[k], by = c->bidx[nb + k];
        double sx = 0, sy = 0;
        for (int t = 0; t < NS; ++t)
            for (int s = 0; s < NS; ++s) {
                int i = bx + s - SUPP, j = by + t - SUPP;
                sx += QX(q, i, j) * w1[s + NS * t];
                sy += QY(q, i, j) * w2[s + NS * t];
            }
        fb[k] = sx;
        fb[nb + k] = sy;
    }
    return 0;
}

/* Vort2Flux, fluid-ops.jl:440-472.  Modifies the coarse-level centres of g. */
int ibr_vort2flux(ibr_case *c, double *psi, double *q, double *g) {
    const int nx = c->nx, ny = c->ny, L = c->nlev;
    const size_t ng = c->ng, nq = c->nq;
    double *psibc = c->gbc, *gtmp = c->g1;
    for (int lev = 2; lev <= L; ++lev) ibr_coarsify(nx, ny, g + (lev - 1) * ng, g + (lev - 2) * ng);
    memset(psi, 0, sizeof(double) * ng * L);
    memset(psibc, 0, sizeof(double) * c->nbc);
    lapinv_apply(c, c->sc_D, g + (L - 1) * ng, psi + (L - 1) * ng);
    ibr_curl_bc(nx, ny, psi + (L - 1) * ng, psibc, q + (L - 1) * nq);
    for (int lev = L - 1; lev >= 1; --lev) {
        memcpy(gtmp, g + (lev - 1) * ng, sizeof(double) * ng);
        ibr_get_bc(nx, ny, psibc, psi + lev * ng);
        ibr_apply_bc(nx, ny, gtmp, psibc, 1.0);
        lapinv_apply(c, c->sc_D, gtmp, psi + (lev - 1) * ng);
        ibr_curl_bc(nx, ny, psi + (lev - 1) * ng, psibc, q + (lev - 1) * nq);
    }
    return 0;
}

/* Nonlinear, cnab/ops.jl:151-168 with RhsForce, fluid-ops.jl:479-492.
 * Q is qtmp1 and fq is qtmp2 (timestepping.jl:222-223). */
int ibr_nonlinear(ibr_case *c, int lev, const double *q, const double *q0, const double *g,
                  const double *gbc, double *nonlin) {
    const int nx = c->nx, ny = c->ny;
    ibr_avg_flux(nx, ny, c->q1, q, q0);
    ibr_direct_product(nx, ny, c->q2, c->q1, g, gbc);
    ibr_rot(nx, ny, c->q2, nonlin);
    double hc = c->h * pow(2.0, lev);
    double s = 1 / (hc * hc);
    for (size_t k = 0; k < c->ng; ++k) nonlin[k] *= s;
    return 0;
}

/* B_Times, cnab/ops.jl:125-143 */
int ibr_B_times(ibr_case *c, const double *z, double *x) {
    const size_t ng = c->ng;
    int rc;
    memset(c->gtmpL, 0, sizeof(double) * ng * c->nlev);
    if ((rc = ibr_ET(c, z, c->q2))) return rc;        /* qtmp = qtmp2 */
    ibr_rot(c->nx, c->ny, c->q2, c->g2);                /* Γtmp = Γtmp2 */
    lapinv_apply(c, c->sc_Ainv, c->g2, c->gtmpL);       /* Γ[:,1] = Ainv₁ Γtmp */
    ibr_vort2flux(c, c->psitmp, c->qtmp, c->gtmpL);
    return ibr_E(c, c->qtmp, x);
}

/* LU with partial pivoting and explicit inverse (LinearAlgebra.inv, cnab/ops.jl:96). */
static int dense_inverse(int n, double *A, double *Ainv) {
    int *piv = (int *)malloc(sizeof(int) * n);
    for (int k = 0; k < n; ++k) {
        int pr = k;
        double best = fabs(A[k + (size_t)n * k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i + (size_t)n * k]) > best) { best = fabs(A[i + (size_t)n * k]); pr = i; }
        piv[k] = pr;
        if (best == 0.0) { free(piv); return -1; }
        if (pr != k)
            for (int j = 0; j < n; ++j) {
                double t = A[k + (size_t)n * j];
                A[k + (size_t)n * j] = A[pr + (size_t)n * j];
                A[pr + (size_t)n * j] = t;
            }
        double d = A[k + (size_t)n * k];
        for (int i = k + 1; i < n; ++i) A[i + (size_t)n * k] /= d;
#pragma omp parallel for schedule(static)
        for (int j = k + 1; j < n; ++j) {
            double akj = A[k + (size_t)n * j];
            if (akj != 0.0)
                for (int i = k + 1; i < n; ++i) A[i + (size_t)n * j] -= A[i + (size_t)n * k] * akj;
        }
    }
    /* solve A X = P I column by column */
#pragma omp parallel for schedule(static)
    for (int col = 0; col < n; ++col) {
        double *x = Ainv + (size_t)n * col;
        for (int i = 0; i < n; ++i) x[i] = (i == col) ? 1.0 : 0.0;
        for (int k = 0; k < n; ++k)
            if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
        for (int k = 0; k < n; ++k) {
            double xk = x[k];
            if (xk != 0.0)
                for (int i = k + 1; i < n; ++i) x[i] -= A[i + (size_t)n * k] * xk;
        }
        for (int k = n - 1; k >= 0; --k) {
            x[k] /= A[k + (size_t)n * k];
            double xk = x[k];
            for (int i = 0; i < k; ++i) x[i] -= A[i + (size_t)n * k] * xk;
        }
    }
    free(piv);
    return 0;
}

/* Solver construction, timestepping.jl:193-249: Reg from the state's points, then the dense
 * Binv of cnab/ops.jl:79-102 (unit-vector B-products, then inv) unless CG is used. */
int ibr_setup_solver(ibr_case *c) {
    int anymove = 0;
    for (int b = 0; b < c->nbodies; ++b) anymove |= c->moving[b];
    /* moving bodies under CG: the regularization is rebuilt every step from the points that
     * update_bodies! produces, so nothing is precomputed here */
    if (c->use_cg && anymove) return 0;
    int rc = ibr_reg_update(c, c->r);
    if (rc) return rc;
    if (c->use_cg) return 0;
    const int nf = 2 * c->nb;
    double *Bmat = (double *)zalloc((size_t)nf * nf);
    double *e = (double *)zalloc(nf);
    for (int j = 0; j < nf; ++j) {
        e[j > 0 ? j - 1 : 0] = 0;
        e[j] = 1;
        if ((rc = ibr_B_times(c, e, Bmat + (size_t)nf * j))) { free(Bmat); free(e); return rc; }
    }
    free(c->Binv);
    c->Binv = (double *)zalloc((size_t)nf * nf);
    rc = dense_inverse(nf, Bmat, c->Binv);
    free(Bmat);
    free(e);
    if (rc) snprintf(c->err, sizeof c->err, "singular B matrix");
    return rc;
}

static double dot(int n, const double *a, const double *b) {
    double s = 0;
    for (int k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

/* cg!(f, B, g; maxiter, reltol), cnab/ops.jl:104-114 via IterativeSolvers 0.9.2:
 * x is the warm start and is overwritten. */
static int cg_solve(ibr_case *c, double *x, const double *b) {
    const int n = 2 * c->nb;
    double *r = c->cg_r, *cc = c->cg_c, *u = c->cg_u;
    int rc;
    memset(u, 0, sizeof(double) * n);
    memcpy(r, b, sizeof(double) * n);
    if ((rc = ibr_B_times(c, x, cc))) return rc;
    for (int k = 0; k < n; ++k) r[k] -= cc[k];
    double residual = sqrt(dot(n, r, r));
    double tol = c->cg_reltol * residual; /* max(reltol·‖r0‖, abstol = 0) */
    double prev = 1.0;
    int it = 0;
    while (!(it >= c->cg_maxiter || residual <= tol)) {
        double beta = residual * residual / (prev * prev);
        for (int k = 0; k < n; ++k) u[k] = r[k] + beta * u[k];
        if ((rc = ibr_B_times(c, u, cc))) return rc;
        double alpha = residual * residual / dot(n, u, cc);
        for (int k = 0; k < n; ++k) x[k] += alpha * u[k];
        for (int k = 0; k < n; ++k) r[k] -= alpha * cc[k];
        prev = residual;
        residual = sqrt(dot(n, r, r));
        ++it;
    }
    c->cg_last_iters = it;
    return 0;
}

int ibr_set_body_points(ibr_case *c, const double *r, const double *v) {
    /* update_bodies! result (body-points.jl:47-76), computed by the caller */
    if (r) memcpy(c->r, r, sizeof(double) * 2 * c->nb);
    if (v) memcpy(c->v, v, sizeof(double) * 2 * c->nb);
    return 0;
}

/* advance!(state, solver, t), timestepping.jl:253-290 */
int ibr_advance(ibr_case *c, double t, double ux, double uy) {
    const int nx = c->nx, ny = c->ny, L = c->nlev, nb = c->nb;
    const size_t ng = c->ng, nq = c->nq;
    const double dt = c->dt;
    int rc, anymove = 0;
    for (int b = 0; b < c->nbodies; ++b) anymove |= c->moving[b];
    /* update!(reg) unless every body is RigidBody{GridFrame} (timestepping.jl:262) */
    if (anymove && (rc = ibr_reg_update(c, c->r))) return rc;
    base_flux(c, ux, uy);

    /* GetTrialState, timestepping.jl:28-79 */
    double *bc = c->gbc, *rhsbc = c->g2, *rhs = c->g3;
    memset(bc, 0, sizeof(double) * c->nbc);
    memset(rhsbc, 0, sizeof(double) * ng);
    for (int lev = L; lev >= 1; --lev) {
        double hc = c->h * pow(2.0, lev - 1);
        if (lev < L) {
            ibr_get_bc(nx, ny, bc, c->gam + (size_t)lev * ng);
            double fac = 0.25 * dt / (c->Re * hc * hc);
            ibr_apply_bc(nx, ny, rhsbc, bc, fac);
        }
        for (size_t k = 0; k < c->nbc; ++k) bc[k] *= 0.25;
        ibr_nonlinear(c, lev - 1, c->q + (lev - 1) * nq, c->q0 + (lev - 1) * nq,
                      c->gam + (lev - 1) * ng, bc, c->nl[0] + (lev - 1) * ng);
        lapinv_apply(c, c->sc_A + (lev - 1) * ng, c->gam + (lev - 1) * ng, rhs);
        for (int n = 0; n < 2; ++n) {
            const double *nn = c->nl[n] + (lev - 1) * ng;
            double f = dt * c->beta[n];
            for (size_t k = 0; k < ng; ++k) rhs[k] += f * nn[k];
        }
        for (size_t k = 0; k < ng; ++k) rhs[k] += rhsbc[k];
        lapinv_apply(c, c->sc_Ainv + (lev - 1) * ng, rhs, c->gs + (lev - 1) * ng);
    }
    { /* cycle!(nonlin), timestepping.jl:81 */
        double *tmp = c->nl[0];
        c->nl[0] = c->nl[1];
        c->nl[1] = tmp;
    }
    ibr_vort2flux(c, c->psi, c->qs, c->gs);

    /* BoundaryForcer, timestepping.jl:106-139 (Q = qtmp1) */
    double *Q = c->q1;
    for (size_t k = 0; k < nq; ++k) Q[k] = c->qs[k] + c->q0[k];
    ibr_E(c, Q, c->Ftmp);
    if (anymove)
        for (int k = 0; k < 2 * nb; ++k) c->Ftmp[k] -= c->v[k] * c->h;
    if (c->use_cg) {
        if ((rc = cg_solve(c, c->Fb, c->Ftmp))) return rc;
    } else {
        const int nf = 2 * nb;
        for (int i = 0; i < nf; ++i) {
            double s = 0;
            for (int j = 0; j < nf; ++j) s += c->Binv[i + (size_t)nf * j] * c->Ftmp[j];
            c->Fb[i] = s;
        }
    }

    /* update_stress!, streamfcn/states.jl:53-84 */
    for (int b = 0; b < c->nbodies; ++b) {
        int o = c->boff[b], n = c->npts[b];
        double sx = 0, sy = 0;
        for (int k = 0; k < n; ++k) {
            double fx = c->Fb[o + k] * c->h / dt, fy = c->Fb[nb + o + k] * c->h / dt;
            c->fbs[o + k] = fx;
            c->fbs[nb + o + k] = fy;
            sx += fx;
            sy += fy;
        }
        c->CD[b] = 2 * sx;
        c->CL[b] = 2 * sy;
        for (int k = 0; k < n; ++k) {
            c->fbs[o + k] /= c->ds[o + k];
            c->fbs[nb + o + k] /= c->ds[o + k];
        }
    }

    /* CircProjecter, timestepping.jl:156-171 (Γtmp = Γtmp2, qtmp = qtmp1) */
    memcpy(c->gam, c->gs, sizeof(double) * ng * L);
    if ((rc = ibr_ET(c, c->Fb, c->q1))) return rc;
    ibr_rot(nx, ny, c->q1, c->g2);
    lapinv_apply(c, c->sc_Ainv, c->g2, c->gs);
    for (size_t k = 0; k < ng; ++k) c->gam[k] -= c->gs[k];

    /* vort2flux!(qty), fluid-ops.jl:438 */
    ibr_vort2flux(c, c->psi, c->q, c->gam);
    c->t = t;
    return 0;
}
