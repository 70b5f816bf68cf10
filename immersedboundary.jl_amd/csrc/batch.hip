// Batched CNAB steps: many independent cases of uneven grid size, level count and body-point
// count advanced together.  Each substep of advance! (cnab/timestepping.jl:253-290) becomes
// one grouped launch over every case that takes part in it: the trial state of all levels of
// all cases (timestepping.jl:28-79), each level of the two Vort2Flux sweeps
// (fluid-ops.jl:440-472, coarsest level first; cases with fewer levels join later), the
// dense-inverse force solve (cnab/ops.jl:79-102, timestepping.jl:89-120), update_stress!
// (streamfcn/states.jl:53-84) and the circulation projection (timestepping.jl:156-171).  The
// grouped kernels run the single-case per-element bodies, so a case's arithmetic is the same
// as when it steps alone.  The whole step is one HIP graph per pointer parity.
//
// Cases are contexts made by ibh_create / ibh_init_state / ibh_setup; the batch borrows them
// (it never frees them) and keeps their host-side state (pointer swaps, t, step count) in
// step with what it runs, so ibh_get_field on a case reads the batch's results.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ibh.h"
#include "ctx.h"
#include "ibh_internal.h"

using namespace ibh;

namespace {

enum Op {
    OP_GEMM, OP_CURL, OP_ROT, OP_COARS, OP_GETBC, OP_COPYBC, OP_COPYGETBC, OP_NONLIN, OP_BCPREF, OP_FILL,
    OP_E, OP_ETZ, OP_ETS, OP_ET, OP_ETF, OP_GEMV, OP_STRESS, OP_RHS, OP_RESID,
    OP_ETZS, OP_ROTI, OP_EDGE, OP_SPEC, OP_LNP, OP_LNC, OP_LNO, OP_AXPY
};

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

struct Stage {
    int op = 0, n = 0;
    std::vector<char> args;    // n argument structs
    std::vector<int> blocks;   // per problem (non-GEMM ops)
    // E and Eᵀ: workgroups per problem of each block schedule (the variants the tuner times);
    // Eᵀ's tuning re-runs the zeroing stage tune_pre before each timed scatter
    std::vector<std::vector<int>> vblocks;
    int tune_pre = -1;
    // contractions
    std::vector<GemmArgs> gargs;
    int side = 0, epi = 0, kind = 0, prec = 0;
    int tune_ref = -1;         // index of the stage whose variant this one reuses
    bool dma_ok = true;        // every problem fits the LDS-direct variants' 32-bit offsets
    double flops = 0, bytes = 0;
    int variant = -1;
    // device copies
    const void *d_args = nullptr;
    const int *d_off = nullptr;   // GEMM: one (n+1) offset row per variant
    std::vector<int> total;       // workgroups (per variant for GEMM)
};

}  // namespace

struct ibh_batch {
    std::vector<ibh_ctx *> cs;
    int prec = 0, Lmax = 1;
    // IBH_MIXED cases: the coarse levels of the final Vort2Flux pass run in FP32
    bool mix = false;
    std::vector<NonlinA> nl_patch;  // nonlinear-term arguments reading a case's stored q
    hipStream_t s = nullptr;
    std::vector<Stage> st[2];
    void *arena = nullptr;
    double *uv = nullptr;
    int parity = 0;
    long steps = 0;
    std::vector<double *> gam0, gs0, nl00, nl10;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    bool use_graphs = true, tuned = false;
    int tune_fixed = -1;  // grouped contraction variant forced for every stage (IBH_GRP_VARIANT)
    int e_fixed = -1, et_fixed = -1;  // E / Eᵀ block schedules forced (IBH_REG_SCHED="e,et")
    std::vector<std::string> tune_log;
    ibh::Prof prof;
    std::string err;
};

namespace {

int bfail(ibh_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    return code;
}

#define BCK(expr)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return bfail(b, IBH_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_));  \
    } while (0)

template <class A>
void add(std::vector<Stage> &v, int op, const std::vector<A> &a, const std::vector<int> &bl) {
    if (a.empty()) return;
    Stage S;
    S.op = op;
    S.n = (int)a.size();
    S.args.resize(a.size() * sizeof(A));
    std::memcpy(S.args.data(), a.data(), S.args.size());
    S.blocks = bl;
    v.push_back(std::move(S));
}

// One grouped contraction stage over the listed split contractions (all of one kind and
// epilogue).  `tune_ref` names an earlier stage of the same shapes whose tile variant it reuses.
void add_gemm(std::vector<Stage> &v, const std::vector<GemmArgs> &gs, int kind, int epi, int prec,
              int tune_ref = -1) {
    if (gs.empty()) return;
    Stage S;
    S.op = OP_GEMM;
    S.prec = prec;
    S.kind = kind;
    S.side = kind - 1;
    S.epi = epi;
    S.n = (int)gs.size();
    for (const GemmArgs &g : gs) {
        if (!sine_ready(g, prec)) {
            std::fprintf(stderr, "libibhip: grouped contraction without its sine parts\n");
            std::abort();
        }
        S.gargs.push_back(g);
        S.dma_ok = S.dma_ok && dma_fits(g, kind - 1, prec);
        S.flops += gemm_flops(g, kind);
        S.bytes += gemm_bytes(g, kind, epi, prec);
    }
    S.tune_ref = tune_ref;
    v.push_back(std::move(S));
}

// the four contractions of one LaplacianInv apply per case, as four grouped stages; with
// `share` the second left/right pair reuses the first pair's tuned tile variant (same shapes).
// The last contraction always takes the second one's variant: its epilogue may read its own
// output (EPI_SUBFROM), which the tuner cannot re-run.
void add_lapinv(std::vector<Stage> &v, const std::vector<LapInvPlan> &plans, int prec,
                bool share = true) {
    if (plans.empty()) return;
    const int base = (int)v.size();
    for (int k = 0; k < 4; ++k) {
        std::vector<GemmArgs> gs;
        for (const LapInvPlan &P : plans) gs.push_back(P.g[k]);
        add_gemm(v, gs, plans[0].kind[k], plans[0].epi[k], prec,
                 k == 3 || (share && k == 2) ? base + k - 2 : -1);
    }
}

// One Poisson solve ψ_ℓ = Δinv(rhs) per listed case (each on its own level), in one of three
// forms that agree up to roundoff:
// * SV_FULL: the four contractions of the LaplacianInv apply;
// * SV_LINES: the forward pair, then ψ only on the rows and columns get_bc! reads (a coarse
//   level whose ψ feeds nothing but the next finer level's boundary values);
// * SV_SPEC: level 1 from the kept scaled spectrum F of its Γ = S F S (S·S = (n/2)·I per
//   dimension): the inverse pair of (nx·ny/4)·F ⊙ sD plus the boundary terms in the sine basis.
enum { SV_FULL, SV_LINES, SV_SPEC };
struct SolveItem {
    int i, mode;
    const double *in;    // FULL, LINES: right-hand side with its boundary values applied
    double *psi;         // the level's ψ
    const double *bc;    // SPEC: the level's boundary values (null: zero)
    const double *spec;  // SPEC: the kept spectrum
    const double *sub;   // SPEC: subtracted from it (the projection's spectrum), or null
};

void add_solves(ibh_batch *b, std::vector<Stage> &v, const std::vector<SolveItem> &items,
                int prec) {
    std::vector<EdgeA> ed;
    std::vector<SpecA> sp;
    std::vector<LinesA> ln;
    std::vector<int> edb, spb, lb[3];
    std::vector<GemmArgs> left1, right1, left2, right2;
    for (const SolveItem &it : items) {
        ibh_ctx *c = b->cs[it.i];
        const Dom &d = c->d;
        LapInvPlan P;
        lapinv_plan(c, it.in, d.ng, it.psi, d.ng, c->sD, 0, 1, EPI_STORE, nullptr, nullptr, &P);
        if (it.mode == SV_SPEC) {
            if (it.bc) {
                ed.push_back({d, c->Sm, c->Sp, it.bc, c->edgeS});
                edb.push_back((int)nblk(d.nx / 2 + d.ny / 2, GRP_BS / 64));  // a wave per pair
            }
            sp.push_back({d, it.spec, c->sD, 0.25 * d.nx * d.ny, c->Sm, c->Sp,
                          it.bc ? c->edgeS : nullptr, c->T2, it.sub});
            spb.push_back(nblk(d.ng, GRP_BS));
            left1.push_back(P.g[2]);  // reads T2
            right2.push_back(P.g[3]);
            continue;
        }
        left1.push_back(P.g[0]);
        right1.push_back(P.g[1]);
        if (it.mode == SV_LINES) {
            ln.push_back({d, c->T2, c->Sm, c->Sp, nullptr, c->sD, c->lineS, it.psi});
            for (int k = 0; k < 3; ++k) lb[k].push_back((int)grp_lines_blocks(d, k));
        } else {
            left2.push_back(P.g[2]);
            right2.push_back(P.g[3]);
        }
    }
    add(v, OP_EDGE, ed, edb);
    add(v, OP_SPEC, sp, spb);
    add_gemm(v, left1, 1, EPI_STORE, prec);
    add_gemm(v, right1, 2, EPI_SCALE, prec);
    add(v, OP_LNP, ln, lb[0]);
    add(v, OP_LNC, ln, lb[1]);
    add(v, OP_LNO, ln, lb[2]);
    add_gemm(v, left2, 1, EPI_STORE, prec);
    add_gemm(v, right2, 2, EPI_STORE, prec);
}

// cases whose trial transforms are unrefined keep the trial spectrum of every level (specT)
// and the projection's (specP), so that level 1's solves in both Vort2Flux passes start from it
bool keeps_spectrum(const ibh_ctx *c) { return c->use_eo && c->refine == 0; }

// FP64-residual refinement of the solves on nlev consecutive levels per listed case
// (x -= Solve(M x − rhs), the correction in the batch precision), `sweeps` sweeps per item.
struct RefineItem {
    int i;
    double *x;
    const double *rhs;
    double a;             // < 0: Poisson, else the implicit CN operator I + (a/4^ℓ)(−L)
    const double *scale;  // eigenvalue tables of the solves (level stride sscale)
    long sscale;
    int nlev;
    int sweeps;
};

void add_refine(ibh_batch *b, std::vector<Stage> &v, const std::vector<RefineItem> &items) {
    int maxref = 0;
    for (const RefineItem &it : items) maxref = std::max(maxref, it.sweeps);
    for (int k = 0; k < maxref; ++k) {
        std::vector<ResidA> ra;
        std::vector<int> bl;
        std::vector<LapInvPlan> plans;
        for (const RefineItem &it : items) {
            ibh_ctx *c = b->cs[it.i];
            if (it.sweeps <= k) continue;
            const Dom &d = c->d;
            ra.push_back({d, it.x, it.rhs, it.a, it.a < 0 ? 1.0 : 0.25, c->rres, it.nlev});
            bl.push_back((int)grp_blocks(d.ng) * it.nlev);
            LapInvPlan P;
            lapinv_plan(c, c->rres, d.ng, it.x, d.ng, it.scale, it.sscale, it.nlev, EPI_SUBFROM,
                        nullptr, nullptr, &P);
            plans.push_back(P);
        }
        add(v, OP_RESID, ra, bl);
        add_lapinv(v, plans, b->prec);
    }
}

// Vort2Flux (fluid-ops.jl:440-472) of every case: coarsify up, then the Poisson solves from
// each case's coarsest level down.  Stage j of the descent solves level Lmax-1-j of every case
// that has it: cases with fewer levels join later, and every case solves level 1 in the last
// stage.  With `refine`, the solves on each case's finest refine_levels levels are refined
// before their curl.  Cases that keep their spectrum solve level 1 from it (SV_SPEC), and
// without `curl_all` their coarser levels on get_bc!'s lines only (SV_LINES).
void add_vort2flux(ibh_batch *b, std::vector<Stage> &v, const std::vector<double *> &G,
                   const std::vector<double *> &Psi, const std::vector<double *> &Q,
                   bool curl_all, bool final_pass) {
    const int n = (int)b->cs.size();
    // refinement of the trial pass, or of the final pass (its own sweeps and levels)
    auto sweeps = [&](const ibh_ctx *c) { return final_pass ? c->refine_final : c->refine; };
    auto rlevels = [&](const ibh_ctx *c) {
        return final_pass ? c->refine_final_levels : c->refine_levels;
    };
    // the final pass's level 1 from the spectrum: Γ₁ = S(F_trial − F_proj)S for cases whose
    // projection kept its spectrum
    auto sub = [&](const ibh_ctx *c, int lev) -> const double * {
        return final_pass && lev == 0 && keeps_spectrum(c) && c->nb > 0 ? c->specP : nullptr;
    };
    auto mode = [&](const ibh_ctx *c, int lev) {
        if (!keeps_spectrum(c)) return (int)SV_FULL;
        if (lev == 0) return (int)SV_SPEC;
        return !curl_all && c->lines_B ? (int)SV_LINES : (int)SV_FULL;
    };
    for (int j = 1; j < b->Lmax; ++j) {
        std::vector<CoarsA> a;
        std::vector<int> bl;
        for (int i = 0; i < n; ++i) {
            ibh_ctx *c = b->cs[i];
            if (c->L <= j) continue;
            const Dom &d = c->d;
            a.push_back({d, G[i] + j * d.ng, G[i] + (j - 1) * d.ng, 0, 0, 1});
            bl.push_back(nblk(grp_elems_coarsify(d), GRP_BS));
        }
        add(v, OP_COARS, a, bl);
    }
    {   // ψbc = 0 on every case's coarsest level
        std::vector<FillA> f;
        std::vector<int> fb;
        for (int i = 0; i < n; ++i) {
            ibh_ctx *c = b->cs[i];
            f.push_back({c->psibc + (long)(c->L - 1) * c->d.nbc, c->d.nbc, 0.0});
            fb.push_back(nblk(c->d.nbc, GRP_BS));
        }
        add(v, OP_FILL, f, fb);
    }
    for (int lev = b->Lmax - 1; lev >= 0; --lev) {
        std::vector<GetBcA> gb;
        std::vector<CopyGetBcA> cg;
        std::vector<int> b1, b2, cb;
        std::vector<SolveItem> sv;
        std::vector<CurlA> cu;
        std::vector<RefineItem> rf;
        for (int i = 0; i < n; ++i) {
            ibh_ctx *c = b->cs[i];
            if (c->L <= lev) continue;
            const Dom &d = c->d;
            double *bcl = c->psibc + (long)lev * d.nbc;
            const int md = mode(c, lev);
            const double *rhs = G[i] + lev * d.ng;
            if (lev == c->L - 1) {  // the coarsest level: ψ_L = Δinv Γ_L
                sv.push_back({i, md, rhs, Psi[i] + lev * d.ng, nullptr, c->specT + lev * d.ng,
                               sub(c, lev)});
            } else {
                if (md != SV_SPEC) {  // get_bc! fused with the copy into Γtmp + edges
                    cg.push_back({d, rhs, c->gtmp, Psi[i] + (lev + 1) * d.ng, bcl, 1});
                    b2.push_back((int)grp_blocks(d.ng + d.nbc));
                } else {
                    gb.push_back({d, Psi[i] + (lev + 1) * d.ng, 0, bcl, 0, 1});
                    b1.push_back(nblk(d.nbc, GRP_BS));
                }
                rhs = c->gtmp;
                sv.push_back({i, md, rhs, Psi[i] + lev * d.ng, bcl, c->specT + lev * d.ng,
                              sub(c, lev)});
            }
            if (sweeps(c) && lev < rlevels(c))
                rf.push_back({i, Psi[i] + lev * d.ng, rhs, -1.0, c->sD, 0, 1, sweeps(c)});
            // the final pass stores no flux: the next nonlinear term forms it from ψ and ψbc
            if (!final_pass && (curl_all || lev == 0)) {
                cu.push_back({d, Psi[i] + lev * d.ng, 0, bcl, 0, Q[i] + (curl_all ? lev * d.nq : 0),
                              0, 1});
                cb.push_back((int)grp_blocks(d.nq));
            }
        }
        add(v, OP_GETBC, gb, b1);
        add(v, OP_COPYGETBC, cg, b2);
        add_solves(b, v, sv, lev >= 1 && final_pass && b->mix ? IBH_FP32 : b->prec);
        add_refine(b, v, rf);
        add(v, OP_CURL, cu, cb);
    }
}

// The stage list of one step at batch parity P (pointer set of the creation snapshot, swapped
// when P = 1), in the order of advance!.
void build_step(ibh_batch *b, int P, std::vector<Stage> &v) {
    const int n = (int)b->cs.size();
    std::vector<double *> gam(n), gs(n), n0(n), n1(n), psi(n), qs(n), q(n);
    for (int i = 0; i < n; ++i) {
        gam[i] = P ? b->gs0[i] : b->gam0[i];
        gs[i] = P ? b->gam0[i] : b->gs0[i];
        n0[i] = P ? b->nl10[i] : b->nl00[i];
        n1[i] = P ? b->nl00[i] : b->nl10[i];
        psi[i] = b->cs[i]->psi;
        qs[i] = b->cs[i]->qs;
        q[i] = b->cs[i]->q;
    }
    // ---- GetTrialState (timestepping.jl:28-79), all levels of all cases per launch
    {
        std::vector<GetBcA> gb;
        std::vector<BcPrefA> pf;
        std::vector<int> b1, b2;
        for (int i = 0; i < n; ++i) {
            ibh_ctx *c = b->cs[i];
            if (c->L < 2) continue;
            const Dom &d = c->d;
            gb.push_back({d, gam[i] + d.ng, d.ng, c->bcraw, d.nbc, c->L - 1});
            b1.push_back(nblk(d.nbc, GRP_BS) * (c->L - 1));
            pf.push_back({d.nbc, c->bcraw, c->accbc, c->fac, c->L});
            b2.push_back(nblk(d.nbc, GRP_BS));
        }
        add(v, OP_GETBC, gb, b1);
        add(v, OP_BCPREF, pf, b2);
    }
    {
        std::vector<NonlinA> a;
        std::vector<RhsA> ra;
        std::vector<int> bl, bln;
        std::vector<LapInvPlan> pI;
        std::vector<RefineItem> rf;
        for (int i = 0; i < n; ++i) {
            ibh_ctx *c = b->cs[i];
            const Dom &d = c->d;
            // the previous step's q formed from its ψ and ψbc (batch_step patches in the stored
            // q for a case whose q was set)
            a.push_back({d, c->q, d.nq, gam[i], d.ng, c->bcraw, d.nbc, 0.25, n0[i], d.ng, c->h, 0,
                         b->uv, c->L, c->psi, d.ng, c->psibc, d.nbc});
            bl.push_back((int)grp_blocks(d.ng) * c->L);
            bln.push_back((int)(nl_tiles(d) * c->L));
            // A_ℓΓ as its exact stencil fused with the right-hand side, then Ainv_ℓ
            ra.push_back({d, gam[i], n0[i], n1[i], c->accbc, c->rhsb, c->h, c->dt, c->Re,
                          c->dt * c->beta[0], c->dt * c->beta[1], c->L});
            LapInvPlan P2;
            lapinv_plan(c, c->rhsb, d.ng, gs[i], d.ng, c->sAinv, d.ng, c->L, EPI_STORE, nullptr,
                        nullptr, &P2);
            if (keeps_spectrum(c)) {  // Γs = S F S: keep F of every level
                P2.g[1].C = c->specT;
                P2.g[2].B = c->specT;
            }
            pI.push_back(P2);
            rf.push_back({i, gs[i], c->rhsb, c->dt / (2 * c->Re * c->h * c->h), c->sAinv, d.ng,
                          std::min(c->refine_levels, c->L), c->refine});
        }
        add(v, OP_NONLIN, a, bln);
        add(v, OP_RHS, ra, bl);
        add_lapinv(v, pI, b->prec);
        add_refine(b, v, rf);
    }
    add_vort2flux(b, v, gs, psi, qs, false, false);
    // ---- BoundaryForcer (timestepping.jl:106-120): Fb = Binv·E(qs + q0)
    {
        std::vector<EA> e;
        std::vector<GemvA> g;
        std::vector<int> be, bp;
        for (int i = 0; i < n; ++i) {
            ibh_ctx *c = b->cs[i];
            if (c->nb == 0) continue;
            const int nf = 2 * c->nb;
            e.push_back({c->d, c->reg, c->qs, c->Ftmp, b->uv, c->h});
            be.push_back(nblk(nf, GRP_BS / 64));  // one wave per output
            g.push_back({nf, c->Binv, c->Ftmp, c->Fb});
            bp.push_back(nf);  // one workgroup per row of Binv (stored transposed)
        }
        add(v, OP_E, e, be);
        if (!e.empty()) {
            Stage &S = v.back();
            for (int k = 0; k < e_sched_count(); ++k) {
                std::vector<int> vb;
                for (const EA &a : e) vb.push_back((int)grp_E_blocks(a.r.nb, k));
                S.vblocks.push_back(vb);
            }
        }
        add(v, OP_GEMV, g, bp);
    }
    // ---- update_stress! and CircProjecter (timestepping.jl:156-171): Γ = Γs;
    //      Γ₁ -= Ainv₁ C' Eᵀ Fb
    {
        std::vector<StressA> sa;
        std::vector<ETA> et;
        std::vector<RotIA> ro;
        std::vector<int> b1, b2, b3, b4;
        std::vector<LapInvPlan> plans;
        for (int i = 0; i < n; ++i) {
            ibh_ctx *c = b->cs[i];
            if (c->nb == 0) continue;
            const Dom &d = c->d;
            sa.push_back({c->nbodies, c->d_boff, c->d_npts, c->nb, c->Fb, c->ds, c->h, c->dt, c->CD,
                          c->CL, c->fbs});
            b1.push_back(c->nbodies);
            // Eᵀ into rot: zeroing and scale in one stage, rot reads the fixed-point image
            // the image cleared and rot formed on the body box (the case keeps qtmp and g1
            // zero outside it, as its single-case projection does)
            GBox bx;
            if (c->box_on) {
                bx.x0 = c->bxa;
                bx.x1 = c->bxb;
                bx.y0 = c->bya;
                bx.y1 = c->byb;
            }
            et.push_back({d, c->reg, c->Fb, c->qtmp, c->etscale, bx});
            b2.push_back(nblk(gbox_faces(d, bx), GRP_BS) + 1);
            b3.push_back(nblk(2L * c->nb * GRP_NS * GRP_NS, GRP_BS));
            ro.push_back({d, c->qtmp, c->etscale, c->g1, 1, bx});
            b4.push_back(nblk(gbox_vertices(d, bx), GRP_BS));
            LapInvPlan P;
            lapinv_plan(c, c->g1, d.ng, gs[i], d.ng, c->sAinv, 0, 1, EPI_SUBFROM, nullptr, nullptr, &P);
            if (keeps_spectrum(c)) {  // Γ₁ = S(F_trial − F_proj)S: keep F_proj (the final
                P.g[1].C = c->specP;   // pass's level-1 solve reads F_trial − F_proj)
                P.g[2].B = c->specP;
            }
            if (c->box_on) {  // rot Eᵀ Fb vanishes outside the body box
                P.g[0].klo = c->bxa / 2;
                P.g[0].khi = c->bxb / 2 + 1;
                P.g[1].klo = c->bya / 2;
                P.g[1].khi = c->byb / 2 + 1;
                rows_for_pairs(P.g[0], P.g[1]);
            }
            plans.push_back(P);
        }
        add(v, OP_STRESS, sa, b1);
        add(v, OP_ETZS, et, b2);
        add(v, OP_ET, et, b3);
        if (!et.empty()) {
            Stage &S = v.back();
            S.tune_pre = (int)v.size() - 2;  // the zeroing + scale stage
            for (int k = 0; k < et_sched_count(); ++k) {
                std::vector<int> vb;
                for (const ETA &a : et) vb.push_back((int)grp_ET_blocks(a.r.nb, k));
                S.vblocks.push_back(vb);
            }
        }
        add(v, OP_ROTI, ro, b4);
        add_lapinv(v, plans, IBH_FP64, false);  // the projection stays FP64

    }
    add_vort2flux(b, v, gs, psi, q, true, true);
}

// copy every stage's tables into one device arena
int upload(ibh_batch *b) {
    size_t bytes = 0;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    for (int P = 0; P < 2; ++P)
        for (Stage &S : b->st[P]) {
            const int nv = gemm_eo_grp_variant_count(S.prec);
            if (S.op == OP_GEMM) bytes += al(S.gargs.size() * sizeof(GemmArgs)) + al(sizeof(int) * (S.n + 1) * nv);
            else bytes += al(S.args.size()) + al(sizeof(int) * (S.n + 1) * std::max<size_t>(1, S.vblocks.size()));
        }
    BCK(hipMalloc(&b->arena, bytes + 256));
    std::vector<char> host(bytes + 256, 0);
    size_t o = 0;
    for (int P = 0; P < 2; ++P)
        for (Stage &S : b->st[P]) {
            char *base = static_cast<char *>(b->arena);
            const int nv = gemm_eo_grp_variant_count(S.prec);
            if (S.op == OP_GEMM) {
                const size_t na = S.gargs.size() * sizeof(GemmArgs);
                std::memcpy(host.data() + o, S.gargs.data(), na);
                S.d_args = base + o;
                o += al(na);
                std::vector<int> off((size_t)(S.n + 1) * nv);
                S.total.assign(nv, 0);
                for (int v = 0; v < nv; ++v) {
                    int acc = 0;
                    for (int i = 0; i < S.n; ++i) {
                        off[(size_t)v * (S.n + 1) + i] = acc;
                        acc += (int)grouped_eo_tiles(S.gargs[i], S.prec, v, S.side);
                    }
                    off[(size_t)v * (S.n + 1) + S.n] = acc;
                    S.total[v] = acc;
                }
                std::memcpy(host.data() + o, off.data(), off.size() * sizeof(int));
                S.d_off = reinterpret_cast<const int *>(base + o);
                o += al(off.size() * sizeof(int));
            } else {
                std::memcpy(host.data() + o, S.args.data(), S.args.size());
                S.d_args = base + o;
                o += al(S.args.size());
                // one offset row per block schedule (a single row for the other ops)
                const std::vector<std::vector<int>> rows =
                    S.vblocks.empty() ? std::vector<std::vector<int>>{S.blocks} : S.vblocks;
                std::vector<int> off((size_t)(S.n + 1) * rows.size());
                S.total.assign(rows.size(), 0);
                for (size_t r = 0; r < rows.size(); ++r) {
                    int acc = 0;
                    for (int i = 0; i < S.n; ++i) {
                        off[r * (S.n + 1) + i] = acc;
                        acc += rows[r][i];
                    }
                    off[r * (S.n + 1) + S.n] = acc;
                    S.total[r] = acc;
                }
                std::memcpy(host.data() + o, off.data(), off.size() * sizeof(int));
                S.d_off = reinterpret_cast<const int *>(base + o);
                o += al(off.size() * sizeof(int));
            }
        }
    BCK(hipMemcpy(b->arena, host.data(), o, hipMemcpyHostToDevice));
    return IBH_OK;
}

void launch_stage(ibh_batch *b, const Stage &S, int variant, unsigned long long *ts = nullptr) {
    Grp g{S.d_args, S.d_off, S.n, S.total.empty() ? 0 : S.total[0]};
    hipStream_t s = b->s;
    switch (S.op) {
    case OP_GEMM:
        g.off = S.d_off + (size_t)variant * (S.n + 1);
        g.blocks = S.total[variant];
        launch_gemm_eo_grp(g, S.prec, variant, S.side, S.epi, s, ts);
        break;
    case OP_CURL: launch_grp_curl(g, s); break;
    case OP_ROT: launch_grp_rot(g, s); break;
    case OP_COARS: launch_grp_coarsify(g, s); break;
    case OP_GETBC: launch_grp_get_bc(g, s); break;
    case OP_COPYBC: launch_grp_copy_apply_bc(g, s); break;
    case OP_COPYGETBC: launch_grp_copy_getbc(g, s); break;
    case OP_NONLIN: launch_grp_nonlinear(g, s); break;
    case OP_BCPREF: launch_grp_bc_prefix(g, s); break;
    case OP_FILL: launch_grp_fill(g, s); break;
    case OP_E:
        g.off = S.d_off + (size_t)variant * (S.n + 1);
        g.blocks = S.total[variant];
        launch_grp_E(g, variant, s);
        break;
    case OP_ETZ: launch_grp_ET_zero(g, s); break;
    case OP_ETS: launch_grp_ET_scale(g, s); break;
    case OP_ET:
        g.off = S.d_off + (size_t)variant * (S.n + 1);
        g.blocks = S.total[variant];
        launch_grp_ET(g, variant, s);
        break;
    case OP_ETF: launch_grp_ET_fix(g, s); break;
    case OP_ETZS: launch_grp_ET_zero_scale(g, s); break;
    case OP_ROTI: launch_grp_rot_image(g, s); break;
    case OP_EDGE: launch_grp_edge_gemv(g, s); break;
    case OP_SPEC: launch_grp_spectral_scale(g, s); break;
    case OP_LNP: launch_grp_lines_part(g, s); break;
    case OP_LNC: launch_grp_lines_colsum(g, s); break;
    case OP_LNO: launch_grp_lines_out(g, s); break;
    case OP_AXPY: launch_grp_axpy_sub(g, s); break;
    case OP_GEMV: launch_grp_gemv_t(g, s); break;
    case OP_STRESS: launch_grp_stress(g, s); break;
    case OP_RHS: launch_grp_rhs(g, s); break;
    case OP_RESID: launch_grp_resid(g, s); break;
    }
}

// Time every tile variant of a contraction stage on its real operands (each grouped
// contraction reads operands it does not write, so re-running it leaves the same result) and
// keep the fastest.  Variants of one precision sum k in the same order: identical results.
int tune_stage(ibh_batch *b, std::vector<Stage> &st, size_t k) {
    Stage &S = st[k];
    const bool gemm = S.op == OP_GEMM;
    const int nv = gemm ? gemm_eo_grp_variant_count(S.prec) : (int)S.vblocks.size();
    // a stage that accumulates into its output (Eᵀ) re-runs the stage that clears it first
    auto run = [&](int v) {
        if (S.tune_pre >= 0) launch_stage(b, st[S.tune_pre], st[S.tune_pre].variant);
        launch_stage(b, S, v);
    };
    hipEvent_t e0, e1;
    BCK(hipEventCreate(&e0));
    BCK(hipEventCreate(&e1));
    int best = 0;
    float best_ms = 1e30f;
    for (int v = 0; v < nv; ++v) {
        if (gemm && !S.dma_ok && gemm_eo_grp_variant_dma(S.prec, v)) continue;
        run(v);
        BCK(hipEventRecord(e0, b->s));
        for (int r = 0; r < 3; ++r) run(v);
        BCK(hipEventRecord(e1, b->s));
        BCK(hipEventSynchronize(e1));
        float ms = 0;
        BCK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    // every variant leaves the same output (the same bits); a stage that accumulates has its
    // output cleared once more, for the launch that follows the tuning in run_stages
    if (S.tune_pre >= 0) launch_stage(b, st[S.tune_pre], st[S.tune_pre].variant);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    S.variant = best;
    char line[256];
    if (gemm)
        std::snprintf(line, sizeof line, "%s %s problems=%d %.1fus",
                      gemm_eo_grp_variant_name(S.prec, best), S.kind == 1 ? "left" : "right", S.n,
                      best_ms / 3 * 1e3);
    else
        std::snprintf(line, sizeof line, "%s schedule %s problems=%d %.1fus%s",
                      S.op == OP_E ? "E" : "Eᵀ", S.op == OP_E ? e_sched_name(best) : et_sched_name(best),
                      S.n, best_ms / 3 * 1e3, S.tune_pre >= 0 ? " (with its zeroing)" : "");
    b->tune_log.push_back(line);
    return IBH_OK;
}

int run_stages(ibh_batch *b, std::vector<Stage> &v, bool may_tune) {
    for (size_t k = 0; k < v.size(); ++k) {
        Stage &S = v[k];
        if (S.op == OP_GEMM && S.variant < 0) {
            if (b->tune_fixed >= 0 && b->tune_fixed < gemm_eo_grp_variant_count(S.prec) &&
                (S.dma_ok || !gemm_eo_grp_variant_dma(S.prec, b->tune_fixed)))
                S.variant = b->tune_fixed;
            else if (S.tune_ref >= 0) S.variant = v[S.tune_ref].variant;
            else if (may_tune && S.epi != EPI_SUBFROM) {  // re-running it would subtract again
                int rc = tune_stage(b, v, k);
                if (rc) return rc;
            }
            if (S.variant < 0) S.variant = 0;
        }
        if (!S.vblocks.empty() && S.variant < 0) {  // E and Eᵀ block schedules
            const int fixed = S.op == OP_E ? b->e_fixed : b->et_fixed;
            if (fixed >= 0 && fixed < (int)S.vblocks.size()) S.variant = fixed;
            else if (may_tune) {
                int rc = tune_stage(b, v, k);
                if (rc) return rc;
            }
            if (S.variant < 0) S.variant = 0;
        }
        if (S.op == OP_GEMM && b->prof.on) {
            Prof &P = b->prof;
            if (P.used + 2 > P.pool.size())
                for (int k2 = 0; k2 < 64; ++k2) {
                    hipEvent_t e;
                    hipEventCreate(&e);
                    P.pool.push_back(e);
                }
            unsigned long long *ts = prof_slot(P);
            hipEventRecord(P.pool[P.used], b->s);
            launch_stage(b, S, S.variant, ts);
            hipEventRecord(P.pool[P.used + 1], b->s);
            P.used += 2;
            P.rec.push_back({S.prec, S.flops, S.bytes});
        } else {
            launch_stage(b, S, S.variant);
        }
    }
    return IBH_OK;
}

int batch_step(ibh_batch *b, double ux, double uy) {
    const int P = b->parity;
    const int n = (int)b->cs.size();
    for (int i = 0; i < n; ++i) {
        ibh_ctx *c = b->cs[i];
        if (c->gam != (P ? b->gs0[i] : b->gam0[i]) || c->nl[0] != (P ? b->nl10[i] : b->nl00[i]))
            return bfail(b, IBH_EINVAL, "case " + std::to_string(i) +
                                            " was stepped outside the batch since the batch was created");
    }
    launch_fill(b->uv, 1, ux, b->s);
    launch_fill(b->uv + 1, 1, uy, b->s);
    std::vector<Stage> &v = b->st[P];
    // a case whose q was set through ibh_set_field has its nonlinear term read the stored q
    // for this step: its argument entry loses ψ (restored after the step)
    Stage *nl = nullptr;
    for (Stage &S : v)
        if (S.op == OP_NONLIN) nl = &S;
    std::vector<int> patched;
    char *nl_dev = nl ? const_cast<char *>(static_cast<const char *>(nl->d_args)) : nullptr;
    b->nl_patch.resize(n);
    for (int i = 0; i < n && nl; ++i)
        if (b->cs[i]->q_given) {
            NonlinA &a = b->nl_patch[i];
            std::memcpy(&a, nl->args.data() + i * sizeof(NonlinA), sizeof a);
            a.psi = nullptr;
            BCK(hipMemcpyAsync(nl_dev + i * sizeof(NonlinA), &a, sizeof a, hipMemcpyHostToDevice,
                               b->s));
            patched.push_back(i);
        }
    const bool eager = !b->use_graphs || b->prof.on || !b->tuned;
    int rc;
    if (eager) {
        if ((rc = run_stages(b, v, !b->tuned))) return rc;
        if (!b->tuned) {
            // both parities have the same shapes in the same order
            for (size_t k = 0; k < v.size(); ++k) b->st[1 - P][k].variant = v[k].variant;
            b->tuned = true;
        }
    } else {
        if (!b->graph[P]) {
            BCK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
            rc = run_stages(b, v, false);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->s, &g);
            if (rc) return rc;
            if (e != hipSuccess) return bfail(b, IBH_EDEVICE, std::string("capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph[P], g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->graph[P] = nullptr;
                return bfail(b, IBH_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
        }
        BCK(hipGraphLaunch(b->graph[P], b->s));
    }
    for (int i : patched)
        BCK(hipMemcpyAsync(nl_dev + i * sizeof(NonlinA), nl->args.data() + i * sizeof(NonlinA),
                           sizeof(NonlinA), hipMemcpyHostToDevice, b->s));
    // the cases' host-side state follows what ran: cycle!(nonlin), Γ = Γs, t, step count
    for (ibh_ctx *c : b->cs) {
        c->q_given = false;
        c->q_stale = true;
        std::swap(c->nl[0], c->nl[1]);
        std::swap(c->gam, c->gs);
        c->parity ^= 1;
        c->t = c->steps_done == 0 ? c->t : c->t + c->dt;
        ++c->steps_done;
        c->ux = ux;
        c->uy = uy;
    }
    b->parity ^= 1;
    ++b->steps;
    return IBH_OK;
}

}  // namespace

extern "C" {

int ibh_batch_create(int32_t n, ibh_ctx *const *cases, ibh_batch **out) {
    if (!out || !cases || n <= 0) return IBH_EINVAL;
    *out = nullptr;
    ibh_batch *b = new ibh_batch();
    auto bad = [&](const std::string &m) {
        std::string msg = m;
        delete b;
        // no batch to hold the message: report through the first case
        if (cases[0]) cases[0]->err = msg;
        return IBH_EINVAL;
    };
    for (int i = 0; i < n; ++i) {
        ibh_ctx *c = cases[i];
        if (!c) return bad("null case");
        if (!c->setup_done) return bad("case " + std::to_string(i) + ": ibh_setup has not run");
        if (c->use_cg || c->anymove)
            return bad("case " + std::to_string(i) + ": batches take fixed bodies with the dense force solver");
        if (!c->use_eo) return bad("case " + std::to_string(i) + ": batches need the split transforms");
        if (c->precision != cases[0]->precision || c->mixed != cases[0]->mixed)
            return bad("cases differ in precision");
        for (int j = 0; j < i; ++j)
            if (cases[j] == c) return bad("a case appears twice");
        b->cs.push_back(c);
        b->Lmax = std::max(b->Lmax, c->L);
        b->gam0.push_back(c->gam);
        b->gs0.push_back(c->gs);
        b->nl00.push_back(c->nl[0]);
        b->nl10.push_back(c->nl[1]);
    }
    b->prec = cases[0]->precision;
    b->mix = cases[0]->mixed;
    if (const char *ge = std::getenv("IBH_GRAPH")) b->use_graphs = std::atoi(ge) != 0;
    if (const char *gv = std::getenv("IBH_GRP_VARIANT")) b->tune_fixed = std::atoi(gv);
    if (const char *rs = std::getenv("IBH_REG_SCHED")) parse_reg_sched(rs, &b->e_fixed, &b->et_fixed);
    if (hipDeviceSynchronize() != hipSuccess || hipStreamCreate(&b->s) != hipSuccess ||
        hipMalloc(&b->uv, 2 * sizeof(double)) != hipSuccess) {
        ibh_batch_destroy(b);
        return IBH_EDEVICE;
    }
    build_step(b, 0, b->st[0]);
    build_step(b, 1, b->st[1]);
    int rc = upload(b);
    if (rc) {
        ibh_batch_destroy(b);
        return rc;
    }
    *out = b;
    return IBH_OK;
}

void ibh_batch_destroy(ibh_batch *b) {
    if (!b) return;
    if (b->s) hipStreamSynchronize(b->s);
    for (hipGraphExec_t &g : b->graph)
        if (g) hipGraphExecDestroy(g);
    if (b->arena) hipFree(b->arena);
    if (b->uv) hipFree(b->uv);
    for (hipEvent_t e : b->prof.pool) hipEventDestroy(e);
    if (b->prof.ts) hipFree(b->prof.ts);
    if (b->s) hipStreamDestroy(b->s);
    delete b;
}

const char *ibh_batch_last_error(const ibh_batch *b) { return b ? b->err.c_str() : "null batch"; }

int ibh_batch_advance(ibh_batch *b, double ux, double uy) {
    if (!b) return IBH_EINVAL;
    return batch_step(b, ux, uy);
}

int ibh_batch_advance_n(ibh_batch *b, int32_t nsteps, double ux, double uy) {
    if (!b) return IBH_EINVAL;
    for (int k = 0; k < nsteps; ++k) {
        int rc = batch_step(b, ux, uy);
        if (rc) return rc;
    }
    return IBH_OK;
}

int ibh_batch_sync(ibh_batch *b) {
    if (!b) return IBH_EINVAL;
    BCK(hipStreamSynchronize(b->s));
    BCK(hipGetLastError());
    for (size_t i = 0; i < b->cs.size(); ++i) {
        int f = 0;
        BCK(hipMemcpy(&f, b->cs[i]->derr, sizeof(int), hipMemcpyDeviceToHost));
        if (f) {
            BCK(hipMemset(b->cs[i]->derr, 0, sizeof(int)));
            return bfail(b, (f & 2) ? IBH_ENONFINITE : IBH_EBOUNDS,
                         "case " + std::to_string(i) + ((f & 2) ? ": infinite flux" : ": device error flag"));
        }
    }
    return IBH_OK;
}

int ibh_batch_profile_enable(ibh_batch *b, int32_t on) {
    if (!b) return IBH_EINVAL;
    BCK(hipStreamSynchronize(b->s));
    b->prof.on = on != 0;
    b->prof.used = 0;
    b->prof.rec.clear();
    if (b->prof.on && prof_arm(b->prof, b->s))
        return bfail(b, IBH_EDEVICE, "allocating the profiling timestamps failed");
    return IBH_OK;
}

int ibh_batch_profile_read(ibh_batch *b, int32_t kind, double *ms, int64_t *launches,
                           double *flops) {
    if (!b) return IBH_EINVAL;
    BCK(hipStreamSynchronize(b->s));
    const int rc = prof_device_sum(b->prof, kind, b->s, ms, launches, flops);
    return rc ? bfail(b, rc, "reading the profiling timestamps failed") : IBH_OK;
}

int ibh_batch_profile_read_events(ibh_batch *b, int32_t kind, double *ms, int64_t *launches,
                                  double *flops) {
    if (!b) return IBH_EINVAL;
    BCK(hipStreamSynchronize(b->s));
    const int rc = prof_event_sum(b->prof, kind, ms, launches, flops);
    return rc ? bfail(b, rc, "reading the profiling events failed") : IBH_OK;
}

int ibh_batch_profile_launches(ibh_batch *b, int32_t kind, int32_t max, double *us,
                               double *flops, double *bytes) {
    if (!b || max < 0 || (max > 0 && (!us || !flops || !bytes))) return -IBH_EINVAL;
    return prof_launches(b->prof, kind, b->s, max, us, flops, bytes);
}

int ibh_batch_tune_report(ibh_batch *b, char *buf, int32_t len) {
    if (!b || !buf || len <= 0) return IBH_EINVAL;
    std::string r;
    for (auto &l : b->tune_log) r += l + "\n";
    std::snprintf(buf, (size_t)len, "%s", r.c_str());
    return IBH_OK;
}

int ibh_batch_launches_per_step(ibh_batch *b) {
    return b ? (int)b->st[0].size() : -1;
}

}  // extern "C"
