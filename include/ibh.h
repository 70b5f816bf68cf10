/*
 * ibh.h — C ABI of libibhip.so: the CNAB immersed-boundary projection step of
 * ImmersedBoundary.jl on one MI355X (gfx950), hand-written HIP.
 *
 * Plain C types only.  Arrays cross the boundary as host pointers in the reference's layout
 * (Float64, column-major, x fastest):
 *   Γ, ψ      (nx-1)(ny-1) per level, levels stacked:   [lev1 | lev2 | ...]
 *   q, q0     nq = (nx+1)ny + nx(ny+1) per level, [qx | qy] per level
 *   Fb        [fx_1..fx_nb, fy_1..fy_nb]
 *   points    nb×2 column-major: [x_1..x_nb, y_1..y_nb]
 * The library copies host data synchronously and never keeps host pointers.  A context owns
 * its device memory and one HIP stream; it is not thread-safe (one host thread per context).
 * Every function returns IBH_OK (0) or an ibh_status code; ibh_last_error() has the message.
 *
 * Reference citations are relative to the ImmersedBoundary.jl repository root.
 */
#ifndef IBH_H
#define IBH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IBH_ABI_VERSION 1

enum ibh_status {
    IBH_OK = 0,
    IBH_EINVAL = 1,     /* bad sizes / arguments (e.g. nx % 4 != 0)                        */
    IBH_EBOUNDS = 2,    /* a body point's 13×13 delta support leaves the finest grid       */
    IBH_ENONFINITE = 3, /* non-finite spread flux ("infinite flux", structure-ops.jl:80-82) */
    IBH_ENOCONV = 4,    /* CG reached cg_maxiter without meeting the tolerance; the step   */
                        /* completed with the last iterate (cg! returns it, cnab/ops.jl:110) */
    IBH_EDEVICE = 5,    /* HIP runtime error or missing GPU                                 */
    IBH_ESINGULAR = 6   /* zero pivot in the dense Schur-complement inverse                 */
};

enum ibh_precision {
    IBH_FP64 = 0,   /* every sine transform in FP64 (the drop-in path)                     */
    IBH_FP32 = 1,   /* bulk sine transforms FP32 on MFMA, forces/CG/residuals FP64          */
    IBH_BF16X3 = 2, /* bulk sine transforms as 3 bf16 MFMA products, FP32 accumulation     */
    IBH_MIXED = 3   /* FP64 for the transforms whose error reaches the forces (trial Ainv,  */
                    /* the trial Vort2Flux pass, level 1 of the final pass), FP32 for the   */
                    /* coarse levels of the final Vort2Flux pass (their ψ and q feed only   */
                    /* the next step's nonlinear term); no refinement                       */
};

enum ibh_force_solver {
    IBH_FORCE_AUTO = 0,  /* dense precomputed inverse when every body is fixed, else PCG   */
    IBH_FORCE_DENSE = 1, /* Binv_linearmap dense branch, cnab/ops.jl:79-102                  */
    IBH_FORCE_CG = 2,    /* Binv_linearmap CG branch, cnab/ops.jl:104-114                    */
    IBH_FORCE_PCG = 3    /* the CG branch with a preconditioner: the dense inverse of B,     */
                         /* rebuilt for the moving bodies' columns (see pc_refresh);        */
                         /* same stopping rule (‖b−Bx‖ ≤ reltol·‖b−Bx₀‖), far fewer products */
};

/* Field selectors for ibh_get_field / ibh_set_field (StreamFcnQuantities, streamfcn/states.jl:9-17;
 * CNAB.State.nonlin, cnab/states.jl:1-6). */
enum ibh_field {
    IBH_F_Q = 0,        /* q   nq×L                                        */
    IBH_F_Q0 = 1,       /* q0  nq×L (base flux, fluid-ops.jl:403-417)      */
    IBH_F_GAMMA = 2,    /* Γ   nΓ×L                                        */
    IBH_F_PSI = 3,      /* ψ   nΓ×L                                        */
    IBH_F_FB = 4,       /* Fb  2nb                                         */
    IBH_F_NONLIN1 = 5,  /* nonlin[1] nΓ×L                                  */
    IBH_F_NONLIN2 = 6,  /* nonlin[2] nΓ×L                                  */
    IBH_F_CD = 7,       /* per-body drag coefficient (nbodies)             */
    IBH_F_CL = 8,       /* per-body lift coefficient (nbodies)             */
    IBH_F_FBSTRESS = 9, /* BodyStress.fb stacked nb×2                       */
    IBH_F_POINTS_R = 10,/* body point positions nb×2                       */
    IBH_F_POINTS_V = 11 /* body point velocities nb×2                      */
};

/* Problem description: Problem(fluid, bodies, scheme), src/ImmersedBoundary.jl:34-38, with
 * UniformGrid/MultiLevel (src/fluids/fluids.jl:53-102), StreamFcnFluid Re
 * (src/streamfcn/fluids.jl:49-58), RigidBody xb/ds (src/bodies/rigid-bodies.jl:4-8) and
 * SchemeCNAB dt/β (src/schemes.jl:19-26). */
typedef struct ibh_case_desc {
    int32_t nx, ny, nlev;   /* finest cells (multiples of 4, ≥ 16) and level count        */
    double h, x0, y0;       /* finest spacing and lower-left corner of the finest grid    */
    double Re, dt;
    double beta[2];         /* Adams–Bashforth weights, (1.5, -0.5)                       */
    int32_t nbodies;
    const int32_t *npts;    /* points per body                                            */
    const int32_t *moving;  /* per body: 0 fixed in the grid frame, 1 moving (may be NULL)*/
    const double *xb;       /* nb×2 initial points, column-major                          */
    const double *ds;       /* nb segment lengths                                         */
    int32_t precision;      /* enum ibh_precision                                         */
    int32_t force_solver;   /* enum ibh_force_solver                                      */
    int32_t cg_maxiter;     /* ≤0 → 5000 (cnab/ops.jl:110)                                */
    double cg_reltol;       /* ≤0 → 1e-12 (cnab/ops.jl:110)                               */
    int32_t pc_refresh;     /* PCG: steps between preconditioner rebuilds; ≤0: rebuild when
                               the iterations spent above the post-rebuild minimum have cost
                               as much as the last rebuild (timed on the device)          */
    int32_t refine_sweeps;  /* FP32/BF16X3: iterative-refinement sweeps of every solve in
                               the transform precision on the levels refine_levels selects
                               (trial Ainv_ℓ and both Vort2Flux Poisson solves), each an FP64
                               5-point residual plus a low-precision correction;
                               0 → 1 for FP32 and 2 for BF16X3, < 0 → none.  FP64, MIXED:
                               unused. */
    int32_t refine_levels;  /* FP32/BF16X3: refine the solves of the finest refine_levels
                               levels; ≤ 0 → every level.  The final Vort2Flux pass refines
                               level 1 only (FP32) or every level (BF16X3).  FP64, MIXED:
                               unused.                                                      */
} ibh_case_desc;

typedef struct ibh_ctx ibh_ctx;

/* ---- lifecycle ----------------------------------------------------------------------- */
int ibh_version(void);
/* Allocates device state for one case (replaces the array allocations of
 * StreamFcnQuantities, streamfcn/states.jl:19-41, and Solver, cnab/timestepping.jl:193-249). */
int ibh_create(const ibh_case_desc *desc, ibh_ctx **out);
void ibh_destroy(ibh_ctx *ctx);
const char *ibh_last_error(const ibh_ctx *ctx);
/* initstate(prob, t0): zero q, Γ, ψ, Fb, nonlin; base flux at t0
 * (src/ImmersedBoundary.jl:51 → cnab/states.jl:8-16, streamfcn/states.jl:19-41). */
int ibh_init_state(ibh_ctx *ctx, double t0, double ux, double uy);
/* initsolver(prob, state): Reg(domain, points) (structure-ops.jl:24-35) and, for the dense
 * force solver, Bmat by unit-vector B-products then inv(Bmat) (cnab/ops.jl:79-97), on the GPU
 * (src/ImmersedBoundary.jl:52 → cnab/timestepping.jl:193-249). */
int ibh_setup(ibh_ctx *ctx);
/* update_bodies! output for moving bodies (src/bodies/body-points.jl:47-76): stacked points
 * and velocities at the coming step's t.  NULL keeps the current array.  Points that are not
 * finite or whose 13×13 delta support leaves the finest grid are rejected with IBH_EBOUNDS
 * before anything is uploaded (the reference's BoundsError, structure-ops.jl:70-72, 97-99). */
int ibh_set_body_points(ibh_ctx *ctx, const double *r, const double *v);
/* advance!(state, solver, t) (cnab/timestepping.jl:253-290), with the freestream (ux, uy)
 * = fluid.freestream(t) evaluated by the caller (fluid-ops.jl:403-417).  Returns once the
 * step's device error flags are read: IBH_ENONFINITE ("infinite flux",
 * structure-ops.jl:80-82), IBH_EBOUNDS, or IBH_ENOCONV from the CG force solve. */
int ibh_advance(ibh_ctx *ctx, double t, double ux, double uy);
/* advance! for nsteps steps at t0 + k·dt (k = 0..nsteps-1) with a constant freestream and
 * fixed bodies (IBH_EINVAL if any body moves): the inner loop of solve!
 * (src/solve.jl:198-212) without callbacks.  Device error flags raised by any of the steps
 * are returned at the end. */
int ibh_advance_n(ibh_ctx *ctx, double t0, int32_t nsteps, double ux, double uy);
int ibh_sync(ibh_ctx *ctx);

/* ---- quantities / checkpoint ------------------------------------------------------------ */
/* Copy a field to / from host memory (copy!/similar of streamfcn/states.jl:90-112 and
 * cnab/states.jl:23-27).  Sizes follow enum ibh_field.  ibh_get_field copies the field and
 * then returns any device error flag still set (and clears it).  ibh_set_field of
 * IBH_F_POINTS_R checks the points as ibh_set_body_points does. */
int ibh_get_field(ibh_ctx *ctx, int32_t which, double *host_out);
int ibh_set_field(ibh_ctx *ctx, int32_t which, const double *host_in);
/* out8 = {nx, ny, nlev, nq, nΓ, nb, nbc, nbodies} */
int ibh_dims(ibh_ctx *ctx, int32_t *out8);
double ibh_time(ibh_ctx *ctx);
/* iterations of the last CG force solve (IterativeSolvers cg!, cnab/ops.jl:110) */
int ibh_cg_iterations(ibh_ctx *ctx);
/* PCG preconditioner statistics (no reference counterpart; the reference has no
 * preconditioner): out6 = {rebuilds, total rebuild s, column assembly s, inversion s,
 * last measured seconds per iteration, last rebuild s as the adaptive rule weighs it} */
int ibh_solver_stats(ibh_ctx *ctx, double *out6);

/* Low-rank coarse-level part of the CG B-products (no reference counterpart; a faster way to
 * apply the same B of cnab/ops.jl:125-143): out4 = {rank in use (0: the full product runs),
 * relative max-norm difference of B-products through it from full ones at its last build,
 * builds so far, total build seconds}.  IBH_LOWRANK_B=0 in the environment turns it off. */
int ibh_lowrank_stats(ibh_ctx *ctx, double *out4);

/* ---- operator-level entry points (host arrays in/out; each runs the device kernels) ---- */
/* LaplacianInv apply (fluid-ops.jl:42-56): mode 0 Δinv (Λ), 1 A_ℓ (Λexpl), 2 Ainv_ℓ (Λimpl),
 * cnab/ops.jl:14-33; lev is 0-based.  in/out: nΓ. */
int ibh_op_lapinv(ibh_ctx *ctx, int32_t mode, int32_t lev, const double *in, double *out);
/* FFTW RODFT00 along both axes of the (nx-1)×(ny-1) array (fluid-ops.jl:25,49,51). */
int ibh_op_dst2(ibh_ctx *ctx, const double *in, double *out);
/* rot! (C'), fluid-ops.jl:117-129: q (nq) → Γ (nΓ). */
int ibh_op_rot(ibh_ctx *ctx, const double *q, double *gam);
/* curl!(q, ψ, ψbc), fluid-ops.jl:138-182: ψ (nΓ), ψbc (nbc) → q (nq). */
int ibh_op_curl(ibh_ctx *ctx, const double *psi, const double *psibc, double *q);
/* coarsify!(Γc, Γ), fluid-ops.jl:184-202: writes the centre of gc (in/out, nΓ). */
int ibh_op_coarsify(ibh_ctx *ctx, double *gc, const double *g);
/* get_bc!(rbc, r), fluid-ops.jl:210-239: r (nΓ) → rbc (nbc). */
int ibh_op_get_bc(ibh_ctx *ctx, const double *r, double *rbc);
/* apply_bc!(r, rbc, fac), fluid-ops.jl:249-275: r (nΓ, in/out). */
int ibh_op_apply_bc(ibh_ctx *ctx, double *r, const double *rbc, double fac);
/* Nonlinear (cnab/ops.jl:151-168; avg_flux!/direct_product!/rot!, fluid-ops.jl:277-401,474-492)
 * at level lev (0-based): q, q0 (nq), Γ (nΓ), Γbc (nbc) → nonlin (nΓ). */
int ibh_op_nonlinear(ibh_ctx *ctx, int32_t lev, const double *q, const double *q0,
                     const double *gam, const double *gbc, double *nonlin);
/* Vort2Flux (fluid-ops.jl:440-472): Γ (nΓ×L, in/out: coarse centres are overwritten) →
 * ψ (nΓ×L), q (nq×L). */
int ibh_op_vort2flux(ibh_ctx *ctx, double *gam, double *psi, double *q);
/* Reg update! (structure-ops.jl:37-61) from stacked points r (nb×2). */
int ibh_op_reg_update(ibh_ctx *ctx, const double *r);
/* E (RegT, structure-ops.jl:88-105): q (nq, level 1) → fb (2nb). */
int ibh_op_E(ibh_ctx *ctx, const double *q, double *fb);
/* Eᵀ (Reg, structure-ops.jl:63-86): fb (2nb) → q (nq). */
int ibh_op_ET(ibh_ctx *ctx, const double *fb, double *q);
/* B·z (B_Times, cnab/ops.jl:125-143): z (2nb) → x (2nb). */
int ibh_op_B(ibh_ctx *ctx, const double *z, double *x);

/* LaplacianInv apply as ibh_op_lapinv, on device arrays, allocating nothing.  No reference
 * counterpart: for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  in and out are distinct device arrays of
 * nΓ doubles; the apply is enqueued on the context's stream and the call returns without
 * synchronizing (ibh_sync waits).  The two intermediates live in one library-wide device array
 * per GPU of 2·IBH_NOALLOC_MAX_NG doubles (16 MiB), held until the process ends: grids with
 * more than IBH_NOALLOC_MAX_NG interior points (1024², e.g. 1025×1025 cells) are rejected with
 * IBH_EINVAL, and calls that overlap on one GPU (contexts on different streams, or threads)
 * race on that array.  Contraction shapes not tuned before by the context run an untuned tile
 * variant (same results). */
#define IBH_NOALLOC_MAX_NG 1048576
int ibh_op_lapinv_noalloc(ibh_ctx *ctx, int32_t mode, int32_t lev, const double *in_dev,
                          double *out_dev);

/* ---- batches of independent cases ------------------------------------------------------- */
/* Many independent simulations (the dozens of cases a parameter sweep runs through
 * solve(prob, tspan) one after another, src/solve.jl:154-186) advanced together: every
 * substep of advance! is one launch over all cases, each with its own grid size, level count,
 * Re, dt and bodies.  The cases are contexts after ibh_init_state + ibh_setup, with fixed
 * bodies and the dense force solver (the reference's reachable path, cnab/ops.jl:79-102), all
 * of one precision.  The batch borrows them (ibh_destroy them after ibh_batch_destroy) and
 * keeps their state current: after ibh_batch_sync, ibh_get_field on a case reads its fields.
 * A case stepped on its own after ibh_batch_create makes the next batch step fail. */
typedef struct ibh_batch ibh_batch;
int ibh_batch_create(int32_t ncases, ibh_ctx *const *cases, ibh_batch **out);
void ibh_batch_destroy(ibh_batch *batch);
const char *ibh_batch_last_error(const ibh_batch *batch);
/* one advance! of every case (its own dt; t = t0 + k·dt as in solve!'s loop) with the same
 * freestream (ux, uy) for all cases */
int ibh_batch_advance(ibh_batch *batch, double ux, double uy);
int ibh_batch_advance_n(ibh_batch *batch, int32_t nsteps, double ux, double uy);
int ibh_batch_sync(ibh_batch *batch);
/* grouped launches per batch step (one HIP graph replays them) */
int ibh_batch_launches_per_step(ibh_batch *batch);
int ibh_batch_profile_enable(ibh_batch *batch, int32_t on);
/* kind as in ibh_profile_read: the grouped contraction launches of that precision */
int ibh_batch_profile_read(ibh_batch *batch, int32_t kind, double *ms, int64_t *launches,
                           double *flops);
int ibh_batch_profile_read_events(ibh_batch *batch, int32_t kind, double *ms, int64_t *launches,
                                  double *flops);
int ibh_batch_tune_report(ibh_batch *batch, char *buf, int32_t len);
/* per-launch records of the grouped contraction launches, as ibh_profile_launches */
int ibh_batch_profile_launches(ibh_batch *batch, int32_t kind, int32_t max, double *us,
                               double *flops, double *bytes);

/* ---- profiling of the sine-transform contraction launches (no reference counterpart) --- */
int ibh_profile_enable(ibh_ctx *ctx, int32_t on);
/* The GEMM autotuner's choices so far (one line per contraction shape: tile variant, sizes,
 * measured time), as text.  No reference counterpart. */
int ibh_tune_report(ibh_ctx *ctx, char *buf, int32_t len);
/* kind 0: FP64 transform GEMMs, 1: FP32, 2: bf16x3.  Returns the summed kernel time (ms),
 * launch count and algorithmic flops since the last enable.  Kernel time is the device wall
 * clock from the first workgroup's start to the last one's end, as a profiler's kernel trace
 * measures it; launches that a device flag skipped (B-products after the CG converged) are
 * left out of all three figures. */
int ibh_profile_read(ibh_ctx *ctx, int32_t kind, double *ms, int64_t *launches, double *flops);
/* The same from HIP events recorded around every launch: includes the dispatch gap and the
 * skipped launches. */
int ibh_profile_read_events(ibh_ctx *ctx, int32_t kind, double *ms, int64_t *launches,
                            double *flops);
/* Per-launch records since the last enable, launches of precision `kind` in order (skipped
 * ones left out): device time (us, as ibh_profile_read), algorithmic flops and algorithmic
 * bytes (the data operand over the summed pairs, both sine parts, the output and an epilogue's
 * operand, each once).  Returns the number written (at most max), or -IBH_E... on error. */
int ibh_profile_launches(ibh_ctx *ctx, int32_t kind, int32_t max, double *us, double *flops,
                         double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* IBH_H */
