/*
 * ibref — CPU FP64 restatement of ImmersedBoundary.jl's CNAB immersed-boundary step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: tests, the smoke check and
 * the benchmark's CPU-baseline leg may load it, the product library never does.  Every
 * function follows a reference file:line (paths relative to the reference repository root).
 *
 * Layout (identical to the reference, 0-based here):
 *   Γ, ψ  : per level an (nx-1)×(ny-1) array, x fastest;  level ℓ at offset ℓ·nΓ.
 *   q     : per level [qx (nx+1)×ny | qy nx×(ny+1)], x fastest; level ℓ at offset ℓ·nq.
 *   bc    : 2(nx+1)+2(ny+1) edge values, offsets L=0, R=ny+1, B=2(ny+1), T=2(ny+1)+nx+1.
 *   Fb    : [fx_1..fx_nb, fy_1..fy_nb].
 */
#ifndef IBREF_H
#define IBREF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ibr_desc {
    int32_t nx, ny, nlev;       /* finest cells; nx%4 == ny%4 == 0 */
    double h, x0, y0;           /* finest spacing and lower-left corner of level 1 */
    double Re, dt;
    double beta[2];             /* Adams–Bashforth weights (1.5, -0.5) */
    int32_t nbodies;
    const int32_t *npts;        /* points per body */
    const int32_t *moving;      /* 0: fixed in the grid frame; 1: points set per step */
    const double *xb;           /* nb×2 stacked initial points, column-major */
    const double *ds;           /* nb segment lengths */
    int32_t force_mode;         /* 0 auto (dense when all fixed, else CG), 1 dense, 2 CG */
    int32_t cg_maxiter;
    double cg_reltol;
} ibr_desc;

typedef struct ibr_case ibr_case;

/* lifecycle */
ibr_case *ibr_create(const ibr_desc *d);
void ibr_destroy(ibr_case *c);
const char *ibr_last_error(ibr_case *c);
int ibr_init_state(ibr_case *c, double t0, double ux, double uy);
int ibr_setup_solver(ibr_case *c);              /* Reg + (dense) Binv precompute */
int ibr_set_body_points(ibr_case *c, const double *r, const double *v);
int ibr_advance(ibr_case *c, double t, double ux, double uy);
double *ibr_field(ibr_case *c, int which);      /* see IBR_F_* */
int ibr_cg_iters(ibr_case *c);                  /* iterations of the last CG solve */
int ibr_dims(ibr_case *c, int32_t *out8);       /* nx ny nlev nq nΓ nb nbc nbodies */

enum {
    IBR_F_Q = 0, IBR_F_Q0, IBR_F_GAMMA, IBR_F_PSI, IBR_F_FB, IBR_F_NONLIN1, IBR_F_NONLIN2,
    IBR_F_CD, IBR_F_CL, IBR_F_FBSTRESS, IBR_F_POINTS_R, IBR_F_POINTS_V, IBR_F_BINV,
    IBR_F_QS, IBR_F_GAMMAS
};

/* operator-level entry points (stateless helpers take raw sizes) */
void ibr_dst2(int m, int p, const double *in, double *out);
void ibr_lap_eigs(int nx, int ny, double *lam);
/* mode 0: Δinv (Λ'=Λ), 1: A_ℓ, 2: Ainv_ℓ ; lev is 0-based */
int ibr_lapinv(ibr_case *c, int mode, int lev, const double *b, double *x);
void ibr_curl_plain(int nx, int ny, const double *psi, double *q);
void ibr_rot(int nx, int ny, const double *q, double *gam);
void ibr_curl_bc(int nx, int ny, const double *psi, const double *psibc, double *q);
void ibr_coarsify(int nx, int ny, double *gc, const double *g);
void ibr_get_bc(int nx, int ny, double *rbc, const double *r);
void ibr_apply_bc(int nx, int ny, double *r, const double *rbc, double fac);
void ibr_avg_flux(int nx, int ny, double *Q, const double *q, const double *q0);
void ibr_direct_product(int nx, int ny, double *fq, const double *Q, const double *g,
                        const double *gbc);
int ibr_nonlinear(ibr_case *c, int lev, const double *q, const double *q0, const double *g,
                  const double *gbc, double *nonlin);
int ibr_vort2flux(ibr_case *c, double *psi, double *q, double *g);
double ibr_deltah(double rf, double rb, double dr);
int ibr_reg_update(ibr_case *c, const double *r);
int ibr_E(ibr_case *c, const double *q, double *fb);
int ibr_ET(ibr_case *c, const double *fb, double *q);
int ibr_B_times(ibr_case *c, const double *z, double *x);
void ibr_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
