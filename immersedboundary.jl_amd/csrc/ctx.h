// The context of one case (device state, work buffers, step graphs, tuner), shared by the
// translation units that orchestrate single cases and batches.
#pragma once
#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/ibh.h"
#include "ibh_internal.h"

namespace ibh {

// kind: 0 full contraction, 1 split left (C = S·X), 2 split right (C = X·S)
struct TuneKey {
    int prec, kind, M, N, K, batch;
    bool tout;
    // split contractions: launched output extents (grid x, y) and summed pairs, which the
    // body-box narrowing makes differ between launches of one M, N, K
    int ex, ey, kp;
    bool operator<(const TuneKey &o) const {
        return std::tie(prec, kind, M, N, K, batch, tout, ex, ey, kp) <
               std::tie(o.prec, o.kind, o.M, o.N, o.K, o.batch, o.tout, o.ex, o.ey, o.kp);
    }
};

struct Prof {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Rec {
        int kind;       // precision of the launch
        double flops;   // algorithmic flops
        double bytes;   // algorithmic bytes: each operand read and each output written once
    };
    std::vector<Rec> rec;  // per recorded launch
    unsigned long long *ts = nullptr;  // device wall clock: PROF_SLOTS records of PROF_REC
};

// Profiling of the contraction launches.  prof_arm (at profile enable) allocates and presets
// every timestamp slot at once, so that nothing runs between the profiled launches;
// prof_slot returns the slot of the record about to be pushed (null past the capacity);
// prof_device_sum adds up the device-timed kernel time, launches and flops of one precision,
// leaving out launches that their device flag skipped; prof_event_sum does the same with the
// HIP events around each launch (which also count the dispatch gap and the skipped launches).
int prof_arm(Prof &P, hipStream_t s);
unsigned long long *prof_slot(Prof &P);
int prof_device_sum(const Prof &P, int kind, hipStream_t s, double *ms, int64_t *n, double *fl);
int prof_event_sum(const Prof &P, int kind, double *ms, int64_t *n, double *fl);
// per-launch records of one precision (device time in us, flops, bytes), skipped launches left
// out; returns the number written (at most max)
int prof_launches(const Prof &P, int kind, hipStream_t s, int max, double *us, double *fl,
                  double *by);

}  // namespace ibh

using ibh::Dom;
using ibh::RegDev;
using ibh::CGDev;

struct ibh_ctx {
    Dom d;
    int L = 1;
    double h = 0, x0 = 0, y0 = 0, Re = 0, dt = 0, beta[2] = {1.5, -0.5};
    int nbodies = 0, nb = 0;
    std::vector<int> npts, moving, boff;
    bool anymove = false, use_cg = false;
    int precision = 0, force_solver = 0, cg_maxiter = 5000;
    double cg_reltol = 1e-12;
    hipStream_t s = nullptr;

    // transforms
    double *Sm = nullptr, *Sp = nullptr;
    // odd/even parts of the sine matrices for the split contraction (x and y dimensions)
    double *Som = nullptr, *Sem = nullptr, *Sop = nullptr, *Sep = nullptr;
    double *Smoe = nullptr, *Spoe = nullptr;  // the FP64 parts interleaved (GemmArgs::Soe)
    float *Smoef = nullptr, *Spoef = nullptr;  // the FP32 parts interleaved (GemmArgs::Soef)
    // the split parts rounded for the FP32 and bf16x3 contractions
    float *Somf = nullptr, *Semf = nullptr, *Sopf = nullptr, *Sepf = nullptr;
    unsigned *Somb = nullptr, *Semb = nullptr, *Sopb = nullptr, *Sepb = nullptr;
    bool use_eo = true;
    double *sD = nullptr, *sA = nullptr, *sAinv = nullptr;
    // state (StreamFcnQuantities + CNAB.State.nonlin)
    double *gam = nullptr, *psi = nullptr, *q = nullptr, *nl[2] = {nullptr, nullptr};
    double ux = 0, uy = 0, t = 0;
    // solver work
    double *qs = nullptr, *gs = nullptr, *bcraw = nullptr, *accbc = nullptr, *psibc = nullptr;
    double *gtmp = nullptr, *fac = nullptr, *qtmp = nullptr, *g1 = nullptr;
    double *T1 = nullptr, *T2 = nullptr;
    long Tcap = 0;  // elements in T1/T2
    double *rhsb = nullptr;  // trial right-hand side, L levels
    int refine = 0;          // FP64-residual refinement sweeps of the low-precision solves
    int refine_levels = 1;   // ... on the finest refine_levels levels
    // the same for the step's final Vort2Flux pass, whose ψ and q feed only the next step's
    // nonlinear term (the forces see the trial pass)
    int refine_final = 0, refine_final_levels = 1;
    // IBH_MIXED: precision is FP64 and the final pass's coarse levels run in FP32
    bool mixed = false;
    // The state's flux q is curl!(ψ, ψbc) of the final Vort2Flux pass.  The step does not store
    // it: the next nonlinear term forms it from ψ and ψbc while staging, and ibh_get_field
    // forms it on demand (q_stale: q's buffer does not hold it yet).  q_given: q (or ψ) was set
    // through ibh_set_field, so the next nonlinear term reads the stored q, as the reference's
    // step reads its state's q.
    bool q_stale = false, q_given = false;
    double *rres = nullptr;  // refinement residuals, L levels
    // scaled spectra kept for the level-1 Poisson solves (trial Ainv over L levels, projection)
    // and the step's work for ψ on get_bc!'s lines and for boundary terms in the sine basis
    double *specT = nullptr, *specP = nullptr, *lineS = nullptr, *edgeS = nullptr;
    // B-product work (batched over nc columns)
    int ncB = 0;
    double *zB = nullptr, *qB = nullptr, *gB = nullptr, *psiB = nullptr, *qLB = nullptr,
           *bcB = nullptr, *g1B = nullptr, *edgeW = nullptr, *lineW = nullptr,
           *specB = nullptr;  // scaled spectrum of the B-product's Ainv₁ (nc × nΓ)
    // In a B-product, level ℓ ≥ 1 of Γ is zero outside the region that the coarsify! chain
    // writes: 0-based index ranges [lo, hi) per level, in x (i) and y (j)
    std::vector<int> bx_lo, bx_hi, by_lo, by_hi;
    bool sparse_B = true;   // skip those zero tiles in the B-products' forward transforms
    bool lines_B = true;    // ... and form coarse-level ψ only on the lines get_bc! reads
    // Level 1 of a B-product needs Γ and ψ only within reach of the body points' deltas: an
    // inclusive box of interior indices [bxa, bxb] × [bya, byb] with slack for the motion.
    // Its forward transforms sum only the box's pairs and its Poisson solve forms ψ only on it.
    bool box_B = true, box_on = false;
    int bxa = 0, bxb = 0, bya = 0, byb = 0;
    // Coarse-level part of the B-products, kept at low rank for the CG iterations (see
    // ibh_api.hip): T ≈ Q·Zᵀ on the body box, stored as the flux fields Q and V = curl Z
    // (lr_k columns of nq each), with EQ = E·Q and EV = E·V (2nb × lr_k) at the current body
    // positions; sAD = κ·sAinv₁ ⊙ sD, the level-1 Poisson spectrum of Ainv₁'s input.
    bool lr_on = true, lr_valid = false, lr_failed = false, lr_bypass = false;
    int lr_k = 0, lr_builds = 0;
    double lr_err = 0, lr_build_s = 0;
    double *lrQ = nullptr, *lrV = nullptr, *lrEQ = nullptr, *lrEV = nullptr, *lrW = nullptr;
    double *sAD = nullptr, *zerobc = nullptr;
    // bodies
    double *pts = nullptr, *vel = nullptr, *ds = nullptr;
    double *pin_pts = nullptr;  // pinned host staging of uploaded points and velocities (2·2nb)
    int *pin_stat = nullptr;    // pinned host words: the status read [0, 3) and the flag read [3]
    // host copy of the points last given (nb×2, x then y) and whether every point's delta
    // support lies inside the finest grid (checked on the host before anything reads them)
    std::vector<double> hpts;
    bool pts_ok = true;
    int *d_boff = nullptr, *d_npts = nullptr;
    RegDev reg{};
    int *derr = nullptr;       // device error flags (bit 1 bounds, 2 non-finite, 4 singular)
    double *etscale = nullptr; // fixed-point scale of the spreading, one per column
    int etcap = 0;
    double *Fb = nullptr, *Ftmp = nullptr, *CD = nullptr, *CL = nullptr, *fbs = nullptr;
    double *Binv = nullptr;
    CGDev cg{};
    int cg_last = 0;
    // PCG preconditioner: dense inverse of the assembled B (moving columns refreshed)
    bool use_pcg = false, pc_valid = false;
    int pc_refresh = 50, steps_since_pc = 0;
    bool pc_adaptive = false;            // rebuild by cost (pc_refresh ≤ 0 in the description)
    int pc_prec = 1;                     // transform precision of the preconditioner's columns
    int pc_levels = 1;                   // Vort2Flux levels of the moving columns (≤ 0: all)
    double pc_t_refresh = 0, pc_t_iter = 0, pc_extra = 0;  // seconds
    int pc_base = 1 << 30;               // fewest iterations since the last rebuild
    int pc_rebuilds = 0;
    double pc_time_total = 0, pc_time_cols = 0, pc_time_inv = 0;  // seconds, all rebuilds
    double *Bfull = nullptr, *Bwork = nullptr, *Pc = nullptr;
    void *rb = nullptr;  // rocBLAS handle of the dense inverses (created on first use)
    int *dJ = nullptr, *dinJ = nullptr, *gjws = nullptr;
    std::vector<int> J;                      // unknowns of moving bodies
    std::vector<std::pair<int, int>> Jrange; // contiguous [start, count) runs of J
    bool setup_done = false;
    // step graphs (per pointer parity) and the device parameter/status words
    double *uv = nullptr;      // freestream (ux, uy) of the current step
    int *status = nullptr;     // [CG done, CG iterations, error flags, pad]
    bool use_graphs = true, capturing = false;
    bool no_alloc = false;  // inside ibh_op_lapinv_noalloc: no tuning runs, no profiling
    int dev = 0;            // the device current at ibh_create
    int parity = 0;
    long steps_done = 0;
    // step segments as HIP graphs: the trial state (two pointer parities), chunks of 1, 2, 4,
    // 8 and 16 CG iterations, the projection and Vort2Flux (two parities)
    hipGraphExec_t gTrial[2] = {nullptr, nullptr}, gIter[5] = {nullptr, nullptr, nullptr, nullptr, nullptr},
                   gPost[2] = {nullptr, nullptr};
    ibh::Prof prof;
    std::string err;
    std::vector<void *> allocs;
    // GEMM autotuner: chosen variant per shape, scratch output, log of the choices
    std::map<ibh::TuneKey, int> tuned;
    int tune_fixed = -1;
    bool reg_tuned = false;  // E / Eᵀ block schedules picked (reg.esched, reg.etsched)
    bool tune_verbose = false;  // log every candidate's time (IBH_TUNE_VERBOSE)
    double *tune_buf = nullptr;
    size_t tune_cap = 0;
    std::vector<std::string> tune_log;
};


namespace ibh {

// EPI_RHS extras of the last contraction of a CN solve
struct RhsExtras {
    const double *N1, *N2, *bcacc;
    double f1, f2;
};

// The four split contractions of one LaplacianInv apply (fluid-ops.jl:42-56) on nbatch fields:
// kind 1 = left (S·X, transposed store), kind 2 = right (X·S, epilogue epi[k]).
struct LapInvPlan {
    GemmArgs g[4];
    int epi[4], kind[4];
};

void lapinv_plan(ibh_ctx *c, const double *in, long sin, double *out, long sout,
                 const double *scale, long sscale, int nbatch, int epi, const RhsExtras *rx,
                 const int *skip, LapInvPlan *plan);

// Algorithmic flops of one split contraction (kind 1 left, 2 right), narrowing included
double gemm_flops(const GemmArgs &g, int kind);
// Algorithmic bytes of the same: the data operand over the summed pairs, the two sine parts
// (in the operand precision prec), the output (both butterfly rows of each pair) and, for the
// scaling and subtracting epilogues, the operand D; each once
double gemm_bytes(const GemmArgs &g, int kind, int epi, int prec);

// A forward pair whose second contraction sums only the pairs [klo, khi): the first forms only
// the (128-aligned) rows those K tiles read
void rows_for_pairs(GemmArgs &g0, const GemmArgs &g1);

}  // namespace ibh
