// libibhip C ABI: context lifecycle, the CNAB step orchestration and operator entry points.
//
// One context = one case, all state resident in HBM, one HIP stream.  The step follows
// advance! (cnab/timestepping.jl:253-290); the independent per-level work of GetTrialState
// (timestepping.jl:28-79) is batched over levels in single launches, and the Schur-complement
// products B·z (cnab/ops.jl:125-143) are batched over columns for the dense-inverse setup.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "../../include/ibh.h"
#include "ctx.h"
#include "ibh_internal.h"

using namespace ibh;

namespace {

int fail(ibh_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define HIPCK(expr)                                                                   \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(c, IBH_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(ibh_ctx *c, T **p, size_t n) {
    void *v = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&v, n * sizeof(T));
    if (e != hipSuccess) return fail(c, IBH_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    hipMemsetAsync(v, 0, n * sizeof(T), c->s);
    c->allocs.push_back(v);
    *p = static_cast<T *>(v);
    return IBH_OK;
}

#define ALLOC(ptr, n)                                  \
    do {                                               \
        int rc_ = dalloc(c, &(ptr), (size_t)(n));      \
        if (rc_) return rc_;                           \
    } while (0)

// --------------------------------------------------------------------------- profiling
}  // namespace

namespace ibh {

int prof_arm(Prof &P, hipStream_t s) {
    const size_t bytes = (size_t)PROF_SLOTS * PROF_REC * sizeof(unsigned long long);
    if (!P.ts && hipMalloc(&P.ts, bytes) != hipSuccess) {
        P.ts = nullptr;
        return IBH_EDEVICE;
    }
    // enabling profiling runs outside any graph capture, so a plain memset is fine here
    if (hipMemsetAsync(P.ts, 0, bytes, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return IBH_EDEVICE;
    return IBH_OK;
}

unsigned long long *prof_slot(Prof &P) {
    const size_t k = P.rec.size();
    return (P.ts && k < (size_t)PROF_SLOTS) ? P.ts + k * PROF_REC : nullptr;
}

int prof_device_sum(const Prof &P, int kind, hipStream_t s, double *ms, int64_t *n, double *fl) {
    const size_t nr = P.ts ? std::min(P.rec.size(), (size_t)PROF_SLOTS) : 0;
    std::vector<unsigned long long> h(nr * PROF_REC + 1);
    if (nr) {
        if (hipMemcpyAsync(h.data(), P.ts, nr * PROF_REC * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return IBH_EDEVICE;
    }
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        return IBH_EDEVICE;
    double tot = 0, f = 0;
    int64_t cnt = 0;
    for (size_t k = 0; k < nr; ++k) {
        if (P.rec[k].kind != kind) continue;
        const unsigned long long *r = h.data() + k * PROF_REC;
        unsigned long long b = 0;
        for (int j = 0; j < 16; ++j) b = std::max(b, r[16 + 16 * j]);
        const unsigned long long a = r[0];
        if (a == 0 || b == 0 || b < a) continue;  // skipped on the device
        tot += (double)(b - a) / khz;                  // ticks / kHz = ms
        f += P.rec[k].flops;
        ++cnt;
    }
    if (ms) *ms = tot;
    if (n) *n = cnt;
    if (fl) *fl = f;
    return IBH_OK;
}

int prof_launches(const Prof &P, int kind, hipStream_t s, int max, double *us, double *fl,
                  double *by) {
    const size_t nr = P.ts ? std::min(P.rec.size(), (size_t)PROF_SLOTS) : 0;
    std::vector<unsigned long long> h(nr * PROF_REC + 1);
    if (nr) {
        if (hipMemcpyAsync(h.data(), P.ts, nr * PROF_REC * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return -IBH_EDEVICE;
    }
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        return -IBH_EDEVICE;
    int cnt = 0;
    for (size_t k = 0; k < nr && cnt < max; ++k) {
        if (P.rec[k].kind != kind) continue;
        const unsigned long long *r = h.data() + k * PROF_REC;
        unsigned long long b = 0;
        for (int j = 0; j < 16; ++j) b = std::max(b, r[16 + 16 * j]);
        const unsigned long long a = r[0];
        if (a == 0 || b == 0 || b < a) continue;  // skipped on the device
        us[cnt] = (double)(b - a) * 1e3 / khz;
        fl[cnt] = P.rec[k].flops;
        by[cnt] = P.rec[k].bytes;
        ++cnt;
    }
    return cnt;
}

int prof_event_sum(const Prof &P, int kind, double *ms, int64_t *n, double *fl) {
    double tot = 0, f = 0;
    int64_t cnt = 0;
    for (size_t k = 0; k < P.rec.size(); ++k) {
        if (P.rec[k].kind != kind) continue;
        float e = 0;
        if (hipEventElapsedTime(&e, P.pool[2 * k], P.pool[2 * k + 1]) != hipSuccess)
            return IBH_EDEVICE;
        tot += e;
        f += P.rec[k].flops;
        ++cnt;
    }
    if (ms) *ms = tot;
    if (n) *n = cnt;
    if (fl) *fl = f;
    return IBH_OK;
}

}  // namespace ibh

namespace {

void prof_begin(ibh_ctx *c, hipEvent_t *e0) {
    Prof &P = c->prof;
    if (P.used + 2 > P.pool.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            hipEventCreate(&e);
            P.pool.push_back(e);
        }
    }
    *e0 = P.pool[P.used];
    hipEventRecord(*e0, c->s);
}

// --------------------------------------------------------------------------- autotuner
// Every tile variant of a precision sums k in the same order, so they give identical
// results; the first use of a GEMM shape times each variant on the real operands (output to
// a scratch buffer) and keeps the fastest.
void launch_kind(const GemmArgs &g, int prec, int kind, int v, int epi, bool tout,
                 hipStream_t s) {
    if (kind == 0) launch_gemm(g, prec, v, epi, tout, s);
    else launch_gemm_eo(g, prec, v, kind - 1, epi, s);
}

int tune(ibh_ctx *c, const GemmArgs &g0, int prec, int kind, bool tout) {
    // batch counts share a choice within power-of-two buckets (3-4, 5-8, ..., 33+): the column
    // batches of the preconditioner's rebuilds (e.g. 79, 79, 42) then reuse what the first
    // build tuned instead of timing every variant inside a step; all variants give the same bits
    const int bb = g0.batch <= 2 ? g0.batch : g0.batch <= 4 ? 4 : g0.batch <= 8 ? 8
                 : g0.batch <= 16 ? 16 : g0.batch <= 32 ? 32 : 64;
    int ex = 0, ey = 0, kp = 0;
    if (kind != 0) {
        const int gm = kind == 1 ? g0.h : g0.M, gn = kind == 1 ? g0.N : g0.h;
        ex = g0.ohi[0] > 0 ? std::min(g0.ohi[0], gm) - g0.olo[0] : gm;
        ey = g0.ohi[1] > 0 ? std::min(g0.ohi[1], gn) - g0.olo[1] : gn;
        kp = g0.khi > 0 ? g0.khi - g0.klo : g0.h;
    }
    const TuneKey key{prec, kind, g0.M, g0.N, g0.K, bb, tout, ex, ey, kp};
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) return it->second;
    if (c->tune_fixed >= 0) return c->tuned[key] = c->tune_fixed;
    // never time inside a graph capture (shapes are tuned before) or in the entry point that
    // allocates nothing (the timing needs a scratch output): untuned shapes take variant 0,
    // which gives the same bits as every other
    if (c->capturing || c->no_alloc) return 0;
    const long ext = std::max((long)g0.M, (long)g0.N) * g0.ldc;
    const size_t need = (size_t)std::max(0L, (long)(g0.batch - 1)) * std::max(g0.sC, 0L) + ext;
    if (need > c->tune_cap) {
        if (c->tune_buf) hipFree(c->tune_buf);
        c->tune_buf = nullptr;
        if (hipMalloc(&c->tune_buf, need * sizeof(double)) != hipSuccess) return 0;
        c->tune_cap = need;
    }
    GemmArgs g = g0;
    g.C = c->tune_buf;
    g.skip = nullptr;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int best = 0;
    float best_ms = 1e30f;
    const int nv = kind == 0 ? gemm_variant_count(prec) : gemm_eo_variant_count(prec);
    for (int v = 0; v < nv; ++v) {
        launch_kind(g, prec, kind, v, EPI_STORE, tout, c->s);
        hipEventRecord(e0, c->s);
        for (int r = 0; r < 3; ++r) launch_kind(g, prec, kind, v, EPI_STORE, tout, c->s);
        hipEventRecord(e1, c->s);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (hipGetLastError() == hipSuccess && ms < best_ms) {
            best_ms = ms;
            best = v;
        }
        if (c->tune_verbose)
            c->tune_log.push_back(std::string("  candidate ") +
                                  (kind == 0 ? gemm_variant_name(prec, v) : gemm_eo_variant_name(prec, v)) +
                                  " " + std::to_string(ms / 3 * 1e3) + "us");
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    const char *vname = kind == 0 ? gemm_variant_name(prec, best) : gemm_eo_variant_name(prec, best);
    const bool narrowed = kind != 0 && (g0.ohi[0] > 0 || g0.ohi[1] > 0 || g0.khi > 0);
    c->tune_log.push_back(std::string(vname) + (kind == 2 ? " right" : kind == 1 ? " left" : "") +
                          " M=" + std::to_string(g0.M) + " N=" + std::to_string(g0.N) + " K=" +
                          std::to_string(g0.K) + " batch=" + std::to_string(g0.batch) +
                          (tout ? " T" : "") +
                          (narrowed ? " out " + std::to_string(ex) + "x" + std::to_string(ey) +
                                          " pairs " + std::to_string(kp)
                                    : std::string()) +
                          " " + std::to_string(best_ms / 3 * 1e3) + "us");
    return c->tuned[key] = best;
}

}  // namespace

// Block schedules of E and Eᵀ for this case's points, picked once with valid points (outside a
// graph capture): each schedule runs on the real operands (E's output and Eᵀ's image into the
// tuning scratch, Eᵀ spreading the current forces) and the fastest is kept.  Every schedule
// gives the same bits.  IBH_REG_SCHED="e,et" forces them.
void tune_reg(ibh_ctx *c) {
    if (c->reg_tuned || c->nb == 0 || c->capturing || c->no_alloc) return;
    c->reg_tuned = true;
    int fe = -1, fet = -1;
    if (const char *rs = std::getenv("IBH_REG_SCHED")) parse_reg_sched(rs, &fe, &fet);
    const Dom &d = c->d;
    const int nf = 2 * c->nb;
    const size_t need = (size_t)std::max((long)d.nq, (long)nf);
    if (need > c->tune_cap) {
        if (c->tune_buf) hipFree(c->tune_buf);
        c->tune_buf = nullptr;
        c->tune_cap = 0;
        if (hipMalloc(&c->tune_buf, need * sizeof(double)) != hipSuccess) return;
        c->tune_cap = need;
    }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    auto time = [&](auto &&run) {
        run();
        hipEventRecord(e0, c->s);
        for (int r = 0; r < 5; ++r) run();
        hipEventRecord(e1, c->s);
        hipEventSynchronize(e1);
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        return ms / 5;
    };
    RegDev r = c->reg;
    float best_e = 1e30f, best_et = 1e30f;
    int be = 0, bet = 0;
    if (fe >= 0 && fe < e_sched_count()) {
        be = fe;
    } else {
        for (int v = 0; v < e_sched_count(); ++v) {
            r.esched = v;
            const float ms = time([&] {
                launch_E(d, r, c->qs, d.nq, 0.0, 0.0, c->tune_buf, nf, 1, nullptr, c->s);
            });
            if (ms < best_e) { best_e = ms; be = v; }
        }
    }
    if (fet >= 0 && fet < et_sched_count()) {
        bet = fet;
    } else {
        for (int v = 0; v < et_sched_count(); ++v) {
            r.etsched = v;
            const float ms = time([&] {
                launch_ET_image(d, r, c->Fb, nf, c->tune_buf, d.nq, 1, nullptr, c->etscale, c->s);
            });
            if (ms < best_et) { best_et = ms; bet = v; }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    c->reg.esched = be;
    c->reg.etsched = bet;
    char line[200];
    std::snprintf(line, sizeof line, "E schedule %s %.1fus; Eᵀ schedule %s %.1fus (with its zeroing), %d points",
                  e_sched_name(be), best_e < 1e29f ? best_e * 1e3 : 0.0, et_sched_name(bet),
                  best_et < 1e29f ? best_et * 1e3 : 0.0, c->nb);
    c->tune_log.push_back(line);
}

// Algorithmic flops of a launch: 2MNK for the full contraction, 2·2·h·k·(data rows) for the
// split one (two h×k parts; k = h pairs, or the khi - klo pairs a narrowed launch sums)
double ibh::gemm_bytes(const GemmArgs &g, int kind, int epi, int prec) {
    if (kind == 0) return 8.0 * ((double)g.M * g.K + (double)g.K * g.N + 2.0 * g.M * g.N) * g.batch;
    const int ext[2] = {kind == 1 ? g.h : g.M, kind == 1 ? g.N : g.h};
    double e[2];
    for (int a = 0; a < 2; ++a)
        e[a] = g.ohi[a] > 0 ? (double)(std::min(g.ohi[a], ext[a]) - g.olo[a]) : (double)ext[a];
    const double pairs = kind == 1 ? e[0] : e[1], rows = kind == 1 ? e[1] : e[0];
    const double kp = g.khi > 0 ? (double)(g.khi - g.klo) : (double)g.h;
    const double sine = prec == 0 ? 8.0 : 4.0;  // FP64, or FP32 / packed bf16 pairs
    const double out = 2.0 * pairs * rows * 8.0;
    const double d = (kind == 2 && (epi == EPI_SCALE || epi == EPI_SUBFROM)) ? out : 0.0;
    return (2.0 * kp * rows * 8.0 + out + d) * g.batch + 2.0 * pairs * kp * sine;
}

double ibh::gemm_flops(const GemmArgs &g, int kind) {
    if (kind == 0) return 2.0 * g.M * (double)g.N * g.K * g.batch;
    const int ext[2] = {kind == 1 ? g.h : g.M, kind == 1 ? g.N : g.h};
    double e[2];
    for (int a = 0; a < 2; ++a)
        e[a] = g.ohi[a] > 0 ? (double)(std::min(g.ohi[a], ext[a]) - g.olo[a]) : (double)ext[a];
    const double pairs = kind == 1 ? e[0] : e[1], rows = kind == 1 ? e[1] : e[0];
    const double kp = g.khi > 0 ? (double)(g.khi - g.klo) : (double)g.h;
    return 4.0 * pairs * kp * rows * g.batch;
}

namespace {

void gemm(ibh_ctx *c, GemmArgs g, int epi, int prec, bool tout = false, int kind = 0) {
    const int v = tune(c, g, prec, kind, tout);
    if (c->prof.on && !c->no_alloc) {
        g.ts = prof_slot(c->prof);
        hipEvent_t e0;
        prof_begin(c, &e0);
        launch_kind(g, prec, kind, v, epi, tout, c->s);
        hipEvent_t e1 = c->prof.pool[c->prof.used + 1];
        hipEventRecord(e1, c->s);
        c->prof.used += 2;
        c->prof.rec.push_back({prec, gemm_flops(g, kind), gemm_bytes(g, kind, epi, prec)});
    } else {
        launch_kind(g, prec, kind, v, epi, tout, c->s);
    }
}

GemmArgs gargs() {
    GemmArgs g;
    std::memset(&g, 0, sizeof g);
    g.batch = 1;
    return g;
}

}  // namespace

// --------------------------------------------------------------------------- transforms
// LaplacianInv apply (fluid-ops.jl:42-56) on `nbatch` fields: out = S((S·in·S)⊙scale)S where
// the 16·s factor of two RODFT00 passes lives in the scale table.  The last contraction's
// epilogue is `epi` (EPI_STORE, EPI_SUBFROM with D = out, or EPI_RHS with the extras in rhs).
// sine parts of the x (m) or y (p) dimension in every operand precision
static void set_sine(GemmArgs &g, const ibh_ctx *c, bool y) {
    g.So = y ? c->Sop : c->Som;
    g.Se = y ? c->Sep : c->Sem;
    g.Soe = y ? c->Spoe : c->Smoe;
    g.Soef = y ? c->Spoef : c->Smoef;
    g.Sof = y ? c->Sopf : c->Somf;
    g.Sef = y ? c->Sepf : c->Semf;
    g.Sob = y ? c->Sopb : c->Somb;
    g.Seb = y ? c->Sepb : c->Semb;
}

// The four contractions are TN GEMMs (both operands k-fastest): the left ones read S_m
// through its symmetry and write their result transposed, so the right ones read it
// k-fastest too.  Through the odd/even split each costs half the MFMA work of the full one.
void ibh::lapinv_plan(ibh_ctx *c, const double *in, long sin, double *out, long sout,
                      const double *scale, long sscale, int nbatch, int epi, const RhsExtras *rx,
                      const int *skip, LapInvPlan *P) {
    const Dom &d = c->d;
    const int m = d.m, p = d.p;
    const long ng = d.ng;
    GemmArgs g = gargs();
    g.batch = nbatch;
    g.skip = skip;
    g.M = m; g.N = p; g.K = m; g.h = d.nx / 2;
    set_sine(g, c, false);
    g.B = in; g.ldb = m; g.sB = sin;
    g.C = c->T1; g.ldc = p; g.sC = ng;
    P->g[0] = g; P->epi[0] = EPI_STORE; P->kind[0] = 1;
    g.K = p; g.h = d.ny / 2;
    set_sine(g, c, true);
    g.B = nullptr;
    g.A = c->T1; g.lda = p; g.sA = ng;
    g.C = c->T2; g.ldc = m; g.sC = ng;
    g.D = scale; g.ldd = m; g.sD = sscale;
    P->g[1] = g; P->epi[1] = EPI_SCALE; P->kind[1] = 2;
    g.D = nullptr; g.A = nullptr;
    g.K = m; g.h = d.nx / 2;
    set_sine(g, c, false);
    g.B = c->T2; g.ldb = m; g.sB = ng;
    g.C = c->T1; g.ldc = p; g.sC = ng;
    P->g[2] = g; P->epi[2] = EPI_STORE; P->kind[2] = 1;
    g.K = p; g.h = d.ny / 2;
    set_sine(g, c, true);
    g.B = nullptr;
    g.A = c->T1; g.lda = p; g.sA = ng;
    g.C = out; g.ldc = m; g.sC = sout;
    if (epi == EPI_SUBFROM) {
        g.D = out; g.ldd = m; g.sD = sout;
    } else if (epi == EPI_RHS) {
        g.N1 = rx->N1; g.N2 = rx->N2; g.bcacc = rx->bcacc;
        g.f1 = rx->f1; g.f2 = rx->f2;
        g.nbc = d.nbc; g.Lo = d.Lo; g.Ro = d.Ro; g.Bo = d.Bo; g.To = d.To;
    }
    P->g[3] = g; P->epi[3] = epi; P->kind[3] = 2;
}

namespace {

void lapinv(ibh_ctx *c, const double *in, long sin, double *out, long sout, const double *scale,
            long sscale, int nbatch, int epi, const RhsExtras *rx, const int *skip,
            int prec = -1) {
    const Dom &d = c->d;
    const int m = d.m, p = d.p;
    const long ng = d.ng;
    if (prec < 0) prec = c->precision;
    if (c->use_eo) {
        LapInvPlan P;
        lapinv_plan(c, in, sin, out, sout, scale, sscale, nbatch, epi, rx, skip, &P);
        for (int k = 0; k < 4; ++k) gemm(c, P.g[k], P.epi[k], prec, P.kind[k] == 1, P.kind[k]);
        return;
    }
    GemmArgs g = gargs();
    g.batch = nbatch;
    g.skip = skip;
    // T1ᵀ = (S_m · in)ᵀ
    g.M = m; g.N = p; g.K = m;
    g.A = c->Sm; g.lda = m; g.sA = 0;
    g.B = in; g.ldb = m; g.sB = sin;
    g.C = c->T1; g.ldc = p; g.sC = ng;
    gemm(c, g, EPI_STORE, prec, true);
    // T2 = (T1 · S_p) ⊙ scale
    g.M = m; g.N = p; g.K = p;
    g.A = c->T1; g.lda = p; g.sA = ng;
    g.B = c->Sp; g.ldb = p; g.sB = 0;
    g.C = c->T2; g.ldc = m; g.sC = ng;
    g.D = scale; g.ldd = m; g.sD = sscale;
    gemm(c, g, EPI_SCALE, prec);
    // T1ᵀ = (S_m · T2)ᵀ
    g.D = nullptr;
    g.M = m; g.N = p; g.K = m;
    g.A = c->Sm; g.lda = m; g.sA = 0;
    g.B = c->T2; g.ldb = m; g.sB = ng;
    g.C = c->T1; g.ldc = p; g.sC = ng;
    gemm(c, g, EPI_STORE, prec, true);
    // out = T1 · S_p (+ epilogue)
    g.M = m; g.N = p; g.K = p;
    g.A = c->T1; g.lda = p; g.sA = ng;
    g.B = c->Sp; g.ldb = p; g.sB = 0;
    g.C = out; g.ldc = m; g.sC = sout;
    if (epi == EPI_SUBFROM) {
        g.D = out; g.ldd = m; g.sD = sout;
    } else if (epi == EPI_RHS) {
        g.N1 = rx->N1; g.N2 = rx->N2; g.bcacc = rx->bcacc;
        g.f1 = rx->f1; g.f2 = rx->f2;
        g.nbc = d.nbc; g.Lo = d.Lo; g.Ro = d.Ro; g.Bo = d.Bo; g.To = d.To;
    }
    gemm(c, g, epi, prec);
}

// Vort2Flux (fluid-ops.jl:440-472) on nc independent multilevel fields.  G, Psi: per column
// L levels of ng (column stride L·ng); Q: per column nq (level 0 only, curl_all = false) or
// L·nq; bc: per column L·nbc.  Coarse centres of G are overwritten as in the reference.
// `sweeps` sweeps of FP64-residual refinement of the solves M_ℓ x_ℓ = rhs_ℓ on nlev consecutive
// levels (single column): x -= Solve(M x − rhs), the correction in the transform precision.
// M = −L (a < 0, scale table sD) or I + a_ℓ(−L) with a_ℓ = a/4^ℓ (per-level tables, stride sscale).
void refine_sweeps(ibh_ctx *c, double *x, const double *rhs, double a, const double *scale,
                   long sscale, int nlev, int sweeps, int prec) {
    const long ng = c->d.ng;
    for (int k = 0; k < sweeps; ++k) {
        launch_resid(c->d, x, rhs, a, 0.25, c->rres, nlev, c->s);
        lapinv(c, c->rres, ng, x, ng, scale, sscale, nlev, EPI_SUBFROM, nullptr, nullptr, prec);
    }
}

}  // namespace

// A forward pair whose second contraction sums only the pairs [klo, khi) of the first one's
// output rows: the first contraction forms only the rows those K tiles read (128-aligned, so
// every tile variant, BK ≤ 64 pairs, reads rows inside the range); other rows of its output keep
// stale values that nothing reads.
void ibh::rows_for_pairs(GemmArgs &g0, const GemmArgs &g1) {
    if (g1.khi <= 0) return;
    g0.olo[1] = std::max(0, (2 * g1.klo / 128) * 128);
    g0.ohi[1] = std::min(g0.N, ((2 * g1.khi + 127) / 128) * 128);
}

namespace {

// Poisson solve ψ = Δinv(G + edges(bc)) on level `lev` of a B-product, where G is zero
// outside the rows/columns [bx_lo, bx_hi) × [by_lo, by_hi) of that level: the two forward
// contractions sum only those K tiles, and the boundary terms enter in the sine basis as a
// rank-4 update of the scaled spectrum (bc null: zero boundary values).  Up to roundoff this
// is the LaplacianInv apply of G + edges (fluid-ops.jl:42-56, 249-275).  With `lines_only`
// the inverse transform is evaluated only on the rows and columns get_bc! reads.
void lapinv_sparse(ibh_ctx *c, const double *G, long sG, double *Psi, long sPsi, int nc, int lev,
                   const double *bc, long sbc, const int *skip, int prec, bool lines_only) {
    LapInvPlan P;
    lapinv_plan(c, G, sG, Psi, sPsi, c->sD, 0, nc, EPI_STORE, nullptr, skip, &P);
    P.g[0].klo = c->bx_lo[lev] / 2;
    P.g[0].khi = (c->bx_hi[lev] + 1) / 2;
    P.g[1].klo = c->by_lo[lev] / 2;
    P.g[1].khi = (c->by_hi[lev] + 1) / 2;
    rows_for_pairs(P.g[0], P.g[1]);
    gemm(c, P.g[0], P.epi[0], prec, true, P.kind[0]);
    gemm(c, P.g[1], P.epi[1], prec, false, P.kind[1]);
    if (lines_only) {  // the boundary term is added where the line passes read the spectrum
        if (bc) launch_edge_gemv(c->d, c->Sm, c->Sp, bc, sbc, c->edgeW, nc, skip, c->s);
        launch_lines(c->d, c->Sm, c->Sp, c->T2, c->d.ng, Psi, sPsi, c->lineW, nc, skip, c->s,
                     bc ? c->edgeW : nullptr, c->sD);
        return;
    }
    if (bc)
        launch_edge_spectral(c->d, c->Sm, c->Sp, bc, sbc, c->edgeW, c->sD, c->T2, c->d.ng, nc,
                             skip, c->s);
    gemm(c, P.g[2], P.epi[2], prec, true, P.kind[2]);
    gemm(c, P.g[3], P.epi[3], prec, false, P.kind[3]);
}

// Pairs a'' of a split transform of size n whose output rows (a'' and n − 2 − a'') meet the
// rows [lo, hi]: [*plo, *phi).  min(r, n − 2 − r) rises to the middle row h − 1 and falls after.
void pair_range(int lo, int hi, int n, int *plo, int *phi) {
    const int h = n / 2;
    auto f = [n](int r) { return std::min(r, n - 2 - r); };
    *plo = std::min(f(lo), f(hi));
    *phi = (lo <= h - 1 && hi >= h - 1) ? h : std::max(f(lo), f(hi)) + 1;
}

// `skip_q`: ψ on every level as with curl_all, but no flux is stored (the step's final pass,
// whose q the next nonlinear term forms from ψ and ψbc).
// `nlev` > 0 solves only the finest nlev levels (the coarsest of them with zero boundary
// values), as for a domain of nlev levels; the arrays keep the case's L-level layout.
// Options that change only the order of roundoff:
// * `sparse` (B-products): levels ≥ 1 hold only the coarsified image of level 1, so their
//   solves go through lapinv_sparse;
// * `lines_ok` (callers that curl level 1 only): a coarser level's ψ only feeds get_bc! for the
//   next finer level, so it is formed on the lines get_bc! reads (levels that are refined
//   excepted);
// * `spec0`: the scaled spectrum F of level 1's Γ = S F S (the Ainv that produced it, stride
//   nΓ per column); S·S = (n/2)·I per dimension, so level 1's Poisson solve needs only the
//   inverse transform of (nx·ny/4)·F ⊙ sD plus the boundary terms in the sine basis.
void vort2flux(ibh_ctx *c, double *G, double *Psi, double *Q, double *bc, int nc, bool curl_all,
               const int *skip, int prec = -1, int refine = 0, int nlev = -1,
               bool sparse = false, const double *spec0 = nullptr, bool lines_ok = false,
               int rlev = -1, int prec_coarse = -1, bool skip_q = false,
               const double *spec0_sub = nullptr) {
    const Dom &d = c->d;
    if (rlev < 0) rlev = c->refine_levels;  // refined levels (the finest rlev)
    const int Ls = c->L, L = nlev > 0 ? std::min(nlev, Ls) : Ls;
    const long ng = d.ng, nq = d.nq;
    const long sG = Ls * ng, sQ = curl_all ? Ls * nq : nq, sB = (long)Ls * d.nbc;
    if (prec < 0) prec = c->precision;
    if (prec_coarse < 0) prec_coarse = prec;  // the transforms of levels ≥ 2
    const int prec0 = prec;
    const bool bsp = sparse && c->use_eo && c->sparse_B && refine == 0 && !curl_all;
    const bool lines = lines_ok && c->use_eo && c->lines_B && !curl_all;
    const bool fuse0 = spec0 && c->use_eo && !(refine && 0 < rlev);
    // B-products work on their column batch's buffers, the step on its own
    double *lineW = sparse ? c->lineW : c->lineS, *edgeW = sparse ? c->edgeW : c->edgeS;
    // pc: the next coarser level's ψ when the copy path forms bcl = get_bc!(pc) itself
    auto solve = [&](int lev, const double *bcl, const double *pc) {
        double *Gl = G + lev * ng, *Pl = Psi + lev * ng;
        const int prec = lev >= 1 ? prec_coarse : prec0;
        const bool ln = lines && lev >= 1 && !(refine && lev < rlev);
        if (bsp && lev >= 1) {
            lapinv_sparse(c, Gl, sG, Pl, sG, nc, lev, bcl, sB, skip, prec, ln);
            return;
        }
        if (lev == 0 && fuse0) {
            LapInvPlan P;
            lapinv_plan(c, nullptr, 0, Pl, sG, c->sD, 0, nc, EPI_STORE, nullptr, skip, &P);
            if (bcl) launch_edge_gemv(d, c->Sm, c->Sp, bcl, sB, edgeW, nc, skip, c->s);
            launch_spectral_scale(d, spec0, ng, c->sD, 0.25 * d.nx * d.ny, c->T2, ng, nc, skip,
                                  c->s, c->Sm, c->Sp, bcl ? edgeW : nullptr, spec0_sub);
            if (sparse && c->box_on) {
                // E reads level 1 only inside the body box: the x transform forms the pairs
                // whose rows meet it, the y transform its rows and the pairs of its columns
                pair_range(c->bxa, c->bxb, d.nx, &P.g[2].olo[0], &P.g[2].ohi[0]);
                P.g[3].olo[0] = c->bxa;
                P.g[3].ohi[0] = c->bxb + 1;
                pair_range(c->bya, c->byb, d.ny, &P.g[3].olo[1], &P.g[3].ohi[1]);
            }
            gemm(c, P.g[2], P.epi[2], prec, true, P.kind[2]);
            gemm(c, P.g[3], P.epi[3], prec, false, P.kind[3]);
            return;
        }
        const double *in = Gl;
        long sin = sG;
        if (bcl) {
            if (pc)
                launch_copy_getbc(d, Gl, sG, c->gtmp, ng, pc, sG, const_cast<double *>(bcl), sB,
                                  nc, skip, c->s);
            else
                launch_copy_apply_bc(d, Gl, sG, c->gtmp, ng, bcl, sB, 1.0, nc, skip, c->s);
            in = c->gtmp;
            sin = ng;
        }
        if (ln) {
            LapInvPlan P;
            lapinv_plan(c, in, sin, Pl, sG, c->sD, 0, nc, EPI_STORE, nullptr, skip, &P);
            gemm(c, P.g[0], P.epi[0], prec, true, P.kind[0]);
            gemm(c, P.g[1], P.epi[1], prec, false, P.kind[1]);
            launch_lines(d, c->Sm, c->Sp, c->T2, ng, Pl, sG, lineW, nc, skip, c->s);
            return;
        }
        lapinv(c, in, sin, Pl, sG, c->sD, 0, nc, EPI_STORE, nullptr, skip, prec);
    };
    for (int lev = 1; lev < L; ++lev)
        launch_coarsify(d, G + lev * ng, G + (lev - 1) * ng, sG, sG, nc, skip, c->s);
    // ψ_L = Δinv Γ_L with ψbc = 0
    double *bcL = bc + (long)(L - 1) * d.nbc;
    // a B-product over every level finds its coarsest boundary values zero already: their
    // buffer is zeroed when allocated and get_bc! writes only the finer levels' values
    if (!(sparse && L == Ls)) launch_fill_strided(bcL, d.nbc, sB, nc, 0.0, c->s);
    solve(L - 1, nullptr, nullptr);
    if (refine && L - 1 < rlev)
        refine_sweeps(c, Psi + (L - 1) * ng, G + (L - 1) * ng, -1.0, c->sD, 0, 1, refine,
                      L > 1 ? prec_coarse : prec0);
    // E reads level 1 only inside the body box in a B-product (whose Poisson solve formed ψ there)
    const bool qbox = !curl_all && sparse && c->box_on && c->use_eo && refine == 0;
    if ((curl_all || L == 1) && !skip_q) {
        if (qbox && L == 1)
            launch_curl_box(d, Psi, sG, bcL, sB, Q, sQ, nc, skip, c->s, c->bxa + 1, c->bxb + 2,
                            c->bya + 1, c->byb + 2);
        else
            launch_curl_bc(d, Psi + (L - 1) * ng, sG, bcL, sB, Q + (curl_all ? (L - 1) * nq : 0),
                           sQ, nc, skip, c->s);
    }
    for (int lev = L - 2; lev >= 0; --lev) {
        double *bcl = bc + (long)lev * d.nbc;
        // the copy path forms the boundary values in its copy of Γ (launch_copy_getbc)
        const bool copy_path = !(bsp && lev >= 1) && !(lev == 0 && fuse0);
        if (!copy_path) launch_get_bc(d, Psi + (lev + 1) * ng, sG, bcl, sB, nc, skip, c->s);
        solve(lev, bcl, copy_path ? Psi + (lev + 1) * ng : nullptr);
        if (refine && lev < rlev)
            refine_sweeps(c, Psi + lev * ng, c->gtmp, -1.0, c->sD, 0, 1, refine,
                          lev >= 1 ? prec_coarse : prec0);
        if ((curl_all || lev == 0) && !skip_q) {
            if (qbox && lev == 0)
                launch_curl_box(d, Psi, sG, bcl, sB, Q, sQ, nc, skip, c->s, c->bxa + 1,
                                c->bxb + 2, c->bya + 1, c->byb + 2);
            else
                launch_curl_bc(d, Psi + lev * ng, sG, bcl, sB, Q + (curl_all ? lev * nq : 0), sQ,
                               nc, skip, c->s);
        }
    }
}

// x = B·z for nc columns (B_Times, cnab/ops.jl:125-143): Γ = 0; Γ₁ = Ainv₁ C' Eᵀ z;
// vort2flux; x = E q₁.  z, x: column stride sz, sx.  The transforms run in FP64 (the
// Schur-complement side of the step) unless `prec` asks otherwise, and Vort2Flux runs on every
// level unless `nlev` asks for fewer (both for the preconditioner's columns).
// B·z with the coarse levels' part at low rank (valid factorization, FP64, every level):
// x = E·curl ψ₁ + EQ·(EVᵀ z), with ψ₁ = Δinv Ainv₁ C' Eᵀ z on level 1 alone (zero boundary
// values).  ψ₁ needs only the forward transform of C'Eᵀz, scaled by κ·sAinv₁ ⊙ sD (S·S =
// (n/2)·I per dimension), and its inverse transform on the body box: four narrowed
// contractions where the full product makes about fourteen and forty small launches.
constexpr int LR_MAXNC = 128;

// The body box as a GBox: Eᵀ's image is cleared and rot formed only there (the whole grid
// when the box is off).  The buffers they write (qB, g1B of the B-products, qtmp, g1 of the
// projection) then hold zeros outside it: clear_box_work restores that when the box changes
// and after the low-rank build, which uses qB and g1B as full-grid scratch.
GBox body_gbox(const ibh_ctx *c) {
    GBox b;
    if (c->box_on) {
        b.x0 = c->bxa;
        b.x1 = c->bxb;
        b.y0 = c->bya;
        b.y1 = c->byb;
    }
    return b;
}

void clear_box_work(ibh_ctx *c) {
    const Dom &d = c->d;
    if (c->qB) launch_fill(c->qB, (long)c->ncB * d.nq, 0.0, c->s);
    if (c->g1B) launch_fill(c->g1B, (long)c->ncB * d.ng, 0.0, c->s);
    if (c->qtmp) launch_fill(c->qtmp, d.nq, 0.0, c->s);
    if (c->g1) launch_fill(c->g1, d.ng, 0.0, c->s);
}

// `chained` (PCG iterations): the direction kernel has set z's Eᵀ scale and the previous
// product's E cleared the image, so Eᵀ is the scatter alone and rot's extra workgroups form
// w = EVᵀz.  Every product leaves the image zero on the box (E's extra workgroups).
// Without `lowrank` this is the one-level product B₁ (zero boundary values on level 1: the
// preconditioner's columns), in the transform precision `prec`.
void B_times_lr(ibh_ctx *c, const double *z, long sz, double *x, long sx, int nc,
                const int *skip, bool chained = false, int prec = IBH_FP64, bool lowrank = true) {
    const Dom &d = c->d;
    const int L = c->L;
    const long ng = d.ng;
    const GBox box = body_gbox(c);
    const int lrk = lowrank ? c->lr_k : 0;
    const double *EV = lowrank ? c->lrEV : nullptr, *EQ = lowrank ? c->lrEQ : nullptr;
    double *W = lowrank ? c->lrW : nullptr;
    if (chained) {
        launch_ET_scatter_only(d, c->reg, z, sz, c->qB, d.nq, nc, skip, c->etscale, c->s);
        launch_rot_image(d, c->qB, d.nq, c->etscale, c->g1B, ng, nc, skip, c->s, box, EV, z, sz,
                         2 * c->nb, W, lrk);
    } else {
        launch_ET_image(d, c->reg, z, sz, c->qB, d.nq, nc, skip, c->etscale, c->s, EV, W, lrk,
                        box);  // w = EVᵀ z in the same launch
        launch_rot_image(d, c->qB, d.nq, c->etscale, c->g1B, ng, nc, skip, c->s, box);
    }
    LapInvPlan P;
    lapinv_plan(c, c->g1B, ng, c->psiB, L * ng, c->sAD, 0, nc, EPI_STORE, nullptr, skip, &P);
    if (c->box_on) {
        P.g[0].klo = c->bxa / 2;  // C'Eᵀz vanishes outside the body box
        P.g[0].khi = c->bxb / 2 + 1;
        P.g[1].klo = c->bya / 2;
        P.g[1].khi = c->byb / 2 + 1;
        rows_for_pairs(P.g[0], P.g[1]);
        pair_range(c->bxa, c->bxb, d.nx, &P.g[2].olo[0], &P.g[2].ohi[0]);  // ψ₁ on the box only
        P.g[3].olo[0] = c->bxa;
        P.g[3].ohi[0] = c->bxb + 1;
        pair_range(c->bya, c->byb, d.ny, &P.g[3].olo[1], &P.g[3].ohi[1]);
    }
    for (int k = 0; k < 4; ++k) gemm(c, P.g[k], P.epi[k], prec, P.kind[k] == 1, P.kind[k]);
    // E of curl ψ₁ (zero boundary values), the flux formed inside E's wave sums, + EQ·w; its
    // extra workgroups clear Eᵀ's image on the box for the next product
    launch_E(d, c->reg, nullptr, 0, 0.0, 0.0, x, sx, nc, skip, c->s, nullptr, 1.0, EQ, W, lrk, 0,
             -1, c->psiB, L * ng, c->zerobc, 0, c->qB, d.nq, box);
}

// the product of a column batch of nc goes through B_times_lr
bool lr_path(const ibh_ctx *c, int prec, int nlev, int nc) {
    return c->lr_valid && !c->lr_bypass && prec == IBH_FP64 && nlev <= 0 && nc <= LR_MAXNC;
}

void B_times(ibh_ctx *c, const double *z, long sz, double *x, long sx, int nc, const int *skip,
             int prec = IBH_FP64, int nlev = -1, bool chained = false) {
    const Dom &d = c->d;
    const int L = c->L, Lv = nlev > 0 ? std::min(nlev, L) : L;
    const long ng = d.ng;
    if (lr_path(c, prec, nlev, nc)) {
        B_times_lr(c, z, sz, x, sx, nc, skip, chained);
        return;
    }
    // one level (the preconditioner's columns): Δinv·Ainv₁ is one diagonal in the sine basis
    // (S·S = (n/2)·I per dimension), so ψ₁ is one forward pair, the scale and one inverse pair
    if (nlev == 1 && c->use_eo && c->sAD && c->zerobc) {
        B_times_lr(c, z, sz, x, sx, nc, skip, false, prec, false);
        return;
    }
    const GBox box = body_gbox(c);
    launch_ET_image(d, c->reg, z, sz, c->qB, d.nq, nc, skip, c->etscale, c->s, nullptr, nullptr,
                    0, box);
    launch_rot_image(d, c->qB, d.nq, c->etscale, c->g1B, ng, nc, skip, c->s, box);
    // Γ's coarse levels (zero here) need no clearing: their buffer is zeroed when allocated and
    // only ever receives coarsify!'s fixed centre blocks, which every product overwrites
    // Ainv₁ with its scaled spectrum kept (specB) for the level-1 Poisson solve
    const bool fuse = c->use_eo && c->sparse_B;
    if (fuse) {
        LapInvPlan P;
        lapinv_plan(c, c->g1B, ng, c->gB, L * ng, c->sAinv, 0, nc, EPI_STORE, nullptr, skip, &P);
        P.g[1].C = c->specB;
        P.g[2].B = c->specB;
        if (c->box_on) {  // rot Eᵀz vanishes outside the body box
            P.g[0].klo = c->bxa / 2;
            P.g[0].khi = c->bxb / 2 + 1;
            P.g[1].klo = c->bya / 2;
            P.g[1].khi = c->byb / 2 + 1;
            rows_for_pairs(P.g[0], P.g[1]);
        }
        if (prec < 0) prec = c->precision;
        for (int k = 0; k < 4; ++k) gemm(c, P.g[k], P.epi[k], prec, P.kind[k] == 1, P.kind[k]);
    } else {
        lapinv(c, c->g1B, ng, c->gB, L * ng, c->sAinv, 0, nc, EPI_STORE, nullptr, skip, prec);
    }
    vort2flux(c, c->gB, c->psiB, c->qLB, c->bcB, nc, false, skip, prec, 0, Lv, true,
              fuse ? c->specB : nullptr, true);
    launch_E(d, c->reg, c->qLB, d.nq, 0.0, 0.0, x, sx, nc, skip, c->s);
}

int check_flags(ibh_ctx *c) {
    int f = 0;
    int *h = c->pin_stat ? c->pin_stat + 3 : &f;
    HIPCK(hipMemcpyAsync(h, c->derr, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    f = *h;
    if (f) {
        HIPCK(hipMemsetAsync(c->derr, 0, sizeof(int), c->s));
        if (f & 1) return fail(c, IBH_EBOUNDS, "a body point's delta support leaves the finest grid");
        if (f & 2) return fail(c, IBH_ENONFINITE, "infinite flux");
        if (f & 4) return fail(c, IBH_ESINGULAR, "singular B matrix");
    }
    return IBH_OK;
}

void drop_graphs(ibh_ctx *c);

void dfree(ibh_ctx *c, void *p);

int ensure_B_work(ibh_ctx *c, int nc) {
    if (c->ncB >= nc) return IBH_OK;
    drop_graphs(c);  // captured graphs hold the old work pointers
    HIPCK(hipStreamSynchronize(c->s));
    const Dom &d = c->d;
    const long L = c->L;
    for (double **p : {&c->qB, &c->g1B, &c->gB, &c->psiB, &c->qLB, &c->bcB, &c->zB, &c->edgeW,
                       &c->lineW, &c->specB}) {
        dfree(c, *p);
        *p = nullptr;
    }
    ALLOC(c->qB, (size_t)nc * d.nq);
    ALLOC(c->g1B, (size_t)nc * d.ng);
    ALLOC(c->gB, (size_t)nc * L * d.ng);
    ALLOC(c->psiB, (size_t)nc * L * d.ng);
    ALLOC(c->qLB, (size_t)nc * d.nq);
    ALLOC(c->bcB, (size_t)nc * L * d.nbc);
    ALLOC(c->zB, (size_t)nc * 2 * c->nb + 1);
    ALLOC(c->edgeW, (size_t)nc * 2 * (d.m + d.p));
    ALLOC(c->lineW, (size_t)nc * lines_work_size(d));
    ALLOC(c->specB, (size_t)nc * d.ng);
    if (nc > c->etcap) {
        dfree(c, c->etscale);
        ALLOC(c->etscale, nc);
        c->etcap = nc;
    }
    if ((long)nc * d.ng > c->Tcap) {
        dfree(c, c->T1);
        dfree(c, c->T2);
        ALLOC(c->T1, (size_t)nc * d.ng);
        ALLOC(c->T2, (size_t)nc * d.ng);
        c->Tcap = (long)nc * d.ng;
    }
    if ((long)nc * d.ng > (long)L * d.ng) {
        dfree(c, c->gtmp);
        ALLOC(c->gtmp, (size_t)nc * d.ng);
    }
    c->ncB = nc;
    return IBH_OK;
}

// the dense Binv of cnab/ops.jl:79-97: Bmat[:, j] = B·e_j in column batches, then the inverse
__global__ void k_unit_cols(double *Z, int nf, int j0, int nc) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)nf * nc) return;
    const int i = (int)(t % nf), col = (int)(t / nf);
    Z[t] = (i == j0 + col) ? 1.0 : 0.0;
}

// Bm[:, j0:j0+cnt] = B·e_j (unit-vector products in column batches, cnab/ops.jl:88-94)
int ensure_level1_tables(ibh_ctx *c);

int B_columns(ibh_ctx *c, double *Bm, int j0, int cnt, int prec = IBH_FP64, int nlev = -1) {
    const int nf = 2 * c->nb;
    // column batch sized to keep the B-product work under ~12 GB
    const Dom &d = c->d;
    const double per_col = 8.0 * (2.0 * d.nq + d.ng * (1.0 + 2.0 * c->L + 3.0));
    int nc = (int)std::max(1.0, std::min(512.0, 12.0e9 / per_col));
    nc = std::min(nc, cnt);
    // the columns [j0, j0 + cnt) of the nf×nf matrix, and the unit-column grid of a batch
    // (nf·nc entries) within 32-bit workgroup counts
    if (j0 < 0 || cnt < 0 || j0 + cnt > nf || (long)nf * nc >= (1L << 31))
        return fail(c, IBH_EINVAL, "B_columns: column range outside the force unknowns");
    int rc = ensure_B_work(c, nc);
    if (!rc && nlev == 1) rc = ensure_level1_tables(c);
    if (rc) return rc;
    for (int j = j0; j < j0 + cnt; j += nc) {
        const int n = std::min(nc, j0 + cnt - j);
        hipLaunchKernelGGL(k_unit_cols, dim3((unsigned)(((long)nf * n + 255) / 256)), dim3(256), 0,
                           c->s, c->zB, nf, j, n);
        B_times(c, c->zB, nf, Bm + (long)nf * j, nf, n, nullptr, prec, nlev);
    }
    return IBH_OK;
}

void dfree(ibh_ctx *c, void *p);

// Inv = inverse(A) through LAPACK's getrf + getri as rocSOLVER implements them (the routines
// behind inv(Bmat), cnab/ops.jl:96): blocked, MFMA-backed, for systems where Gauss–Jordan's
// one HBM pass per pivot is too slow.  A is destroyed.
rocblas_handle rb_handle(ibh_ctx *c) {
    if (!c->rb) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
        c->rb = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(c->rb);
    rocblas_set_stream(h, c->s);
    return h;
}

int invert_lapack(ibh_ctx *c, double *A, double *Inv) {
    const int nf = 2 * c->nb;
    rocblas_handle h = rb_handle(c);
    if (!h) return fail(c, IBH_EDEVICE, "rocblas_create_handle failed");
    rocblas_int *piv = nullptr, *info = nullptr;
    int rc = dalloc(c, &piv, (size_t)nf);
    if (!rc) rc = dalloc(c, &info, 2);
    rocblas_status st = rocblas_status_success;
    if (!rc) st = rocsolver_dgetrf(h, nf, nf, A, nf, piv, info);
    if (!rc && st == rocblas_status_success) st = rocsolver_dgetri(h, nf, A, nf, piv, info + 1);
    int hinfo[2] = {0, 0};
    if (!rc && st == rocblas_status_success) {
        launch_copy(Inv, A, (long)nf * nf, c->s);
        if (hipMemcpyAsync(hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, c->s) != hipSuccess ||
            hipStreamSynchronize(c->s) != hipSuccess)
            rc = fail(c, IBH_EDEVICE, "rocSOLVER inverse: device error");
    }
    dfree(c, piv);
    dfree(c, info);
    if (rc) return rc;
    if (st != rocblas_status_success) return fail(c, IBH_EDEVICE, "rocSOLVER getrf/getri failed");
    if (hinfo[0] || hinfo[1]) return fail(c, IBH_ESINGULAR, "singular B matrix");
    return IBH_OK;
}

// Inv = lower(A)⁻¹ of a symmetric positive definite A by Cholesky (rocSOLVER potrf + potri on
// the lower triangle, then mirrored), for the PCG preconditioner; falls back to LU when the
// factorization finds A not positive definite.  A is destroyed.
__global__ void k_mirror_lower(int n, const double *__restrict__ A, double *__restrict__ Inv) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)n * n) return;
    const int i = (int)(t % n), j = (int)(t / n);
    Inv[t] = i >= j ? A[t] : A[j + (long)n * i];
}

int invert_spd(ibh_ctx *c, double *A, double *Inv) {
    const int nf = 2 * c->nb;
    rocblas_handle h = rb_handle(c);
    if (!h) return fail(c, IBH_EDEVICE, "rocblas_create_handle failed");
    // keep A for the LU fallback
    launch_copy(Inv, A, (long)nf * nf, c->s);
    rocblas_int *info = nullptr;
    int rc = dalloc(c, &info, 2);
    if (rc) return rc;
    rocblas_status st = rocsolver_dpotrf(h, rocblas_fill_lower, nf, Inv, nf, info);
    if (st == rocblas_status_success) st = rocsolver_dpotri(h, rocblas_fill_lower, nf, Inv, nf, info + 1);
    int hinfo[2] = {0, 0};
    if (hipMemcpyAsync(hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, c->s) != hipSuccess ||
        hipStreamSynchronize(c->s) != hipSuccess)
        rc = fail(c, IBH_EDEVICE, "rocSOLVER Cholesky inverse: device error");
    dfree(c, info);
    if (rc) return rc;
    if (st != rocblas_status_success || hinfo[0] || hinfo[1]) return invert_lapack(c, A, Inv);
    // Inv holds the inverse in its lower triangle: mirror it into A, then back
    hipLaunchKernelGGL(k_mirror_lower, dim3((unsigned)(((long)nf * nf + 255) / 256)), dim3(256), 0,
                       c->s, nf, Inv, A);
    launch_copy(Inv, A, (long)nf * nf, c->s);
    return IBH_OK;
}

// Inv = inverse(A) by Gauss–Jordan with partial pivoting (up to 768 unknowns) or rocSOLVER's
// LU (larger systems, or IBH_DENSE_INV=lapack|gj to force one); A is destroyed
int invert(ibh_ctx *c, double *A, double *Inv) {
    const int nf = 2 * c->nb;
    // measured on gfx950: Gauss–Jordan inverts 1656² in 39 ms, rocSOLVER getrf+getri in 16 ms
    bool lapack = nf > 768;
    if (const char *e = std::getenv("IBH_DENSE_INV")) lapack = std::strcmp(e, "lapack") == 0;
    if (lapack) return invert_lapack(c, A, Inv);
    hipLaunchKernelGGL(k_unit_cols, dim3((unsigned)(((long)nf * nf + 255) / 256)), dim3(256), 0,
                       c->s, Inv, nf, 0, nf);
    if (!c->gjws) ALLOC(c->gjws, (size_t)2 * (nf + 4) + 8);
    for (int k = 0; k < nf; ++k) launch_gj_step(nf, A, Inv, k, c->gjws, c->derr, c->s);
    HIPCK(hipStreamSynchronize(c->s));
    return check_flags(c);
}

void dfree(ibh_ctx *c, void *p) {
    if (!p) return;
    for (size_t k = 0; k < c->allocs.size(); ++k)
        if (c->allocs[k] == p) {
            c->allocs.erase(c->allocs.begin() + k);
            hipFree(p);
            return;
        }
}

// Give back the column-batch work of the dense-inverse setup (a few GB on large grids); the
// step's own T1/T2/gtmp return to L·nΓ.
int release_B_work(ibh_ctx *c) {
    HIPCK(hipStreamSynchronize(c->s));
    drop_graphs(c);
    for (double **p : {&c->qB, &c->g1B, &c->gB, &c->psiB, &c->qLB, &c->bcB, &c->zB, &c->edgeW,
                       &c->lineW, &c->specB}) {
        dfree(c, *p);
        *p = nullptr;
    }
    c->ncB = 0;
    const long base = (long)c->L * c->d.ng;
    if (c->Tcap > base) {
        dfree(c, c->T1);
        dfree(c, c->T2);
        dfree(c, c->gtmp);
        ALLOC(c->T1, base);
        ALLOC(c->T2, base);
        ALLOC(c->gtmp, base);
        c->Tcap = base;
    }
    HIPCK(hipStreamSynchronize(c->s));
    return IBH_OK;
}

// ---------------------------------------------------------------- low-rank coarse part
// In a B-product (B_Times, cnab/ops.jl:125-143) the coarse levels of Vort2Flux reach the body
// points only through level 1's boundary values: q₁ = curl(Δinv(Γ₁ + edges(bc₁)), bc₁) with
// bc₁ = G(Γ₁) from the coarsify!/Poisson/get_bc! chain (fluid-ops.jl:448-469).  So B = B₁ + B_c
// with B₁ the one-level product and B_c z = E·T·(C'Eᵀz), where T maps a vertex field on the
// body box to curl Δinv edges G Ainv₁ of it on the box.  T passes through the 2(nx+ny)+4
// boundary values and is smooth there, so its numerical rank is small (~60 at 1e-12 for the
// cylinder + plate on 256², 5 levels, against its 30000 box unknowns).  A randomized range
// finder (Y = T·Ω, Q = orth(Y), Z = Tᵀ·Q) gives T ≈ Q·Zᵀ; B_c z = (E Q)·((E V)ᵀ z) with
// V = curl Z, since Zᵀ C' = (C Z)ᵀ.  The factorization depends on the grid and the body box
// only; EQ and EV follow the body points.  It is accepted only if B-products through it match
// the full ones to LR_TOL (relative, max norm) on random forces, else the rank doubles.
constexpr double LR_TOL = 1e-12;
constexpr int LR_K0 = 64, LR_KMAX = 256, LR_NC = 16;

__global__ void k_scale_product(double *out, const double *a, const double *b, double kappa,
                                long n) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < n) out[t] = kappa * a[t] * b[t];
}

// Y[:, j] = T Γin[:, j] on the body box (zero elsewhere), nc columns (≤ ncB)
void lr_T(ibh_ctx *c, const double *Gin, double *Y, int nc) {
    const Dom &d = c->d;
    const int L = c->L;
    const long ng = d.ng;
    lapinv(c, Gin, ng, c->gB, L * ng, c->sAinv, 0, nc, EPI_STORE, nullptr, nullptr, IBH_FP64);
    // q₁ of the full multilevel solve and of level 1 alone, both formed on the box
    vort2flux(c, c->gB, c->psiB, c->qLB, c->bcB, nc, false, nullptr, IBH_FP64, 0, -1, true,
              nullptr, true);
    vort2flux(c, c->gB, c->psiB, c->qB, c->bcB, nc, false, nullptr, IBH_FP64, 0, 1, true,
              nullptr, true);
    launch_box_diff(d, c->qLB, c->qB, d.nq, Y, nc, c->bxa + 1, c->bxb + 2, c->bya + 1,
                    c->byb + 2, c->s);
}

// Zout[:, j] = Tᵀ W[:, j] (W: flux fields zero off the box): Ainv₁ Gᵀ edgesᵀ Δinv C'.  Gᵀ runs
// the coarse chain backwards: aℓ = get_bc!ᵀ βℓ₋₁, yℓ = Δinv aℓ, βℓ = edgesᵀ yℓ, and
// Gᵀβ₁ = coarsifyᵀ(y₂ + coarsifyᵀ(y₃ + …)).  A holds nc·L·nΓ doubles of work.
void lr_Tt(ibh_ctx *c, const double *W, double *Zout, int nc, double *A) {
    const Dom &d = c->d;
    const int L = c->L;
    const long ng = d.ng, sL = L * ng, sB = (long)L * d.nbc;
    launch_rot(d, W, d.nq, c->g1B, ng, nc, nullptr, c->s);
    lapinv(c, c->g1B, ng, c->psiB, sL, c->sD, 0, nc, EPI_STORE, nullptr, nullptr, IBH_FP64);
    launch_edges_adj(d, c->psiB, sL, c->bcB, sB, nc, c->s);
    for (int lev = 1; lev < L; ++lev) {
        launch_get_bc_adj(d, c->bcB + (lev - 1) * d.nbc, sB, A + lev * ng, sL, nc, c->s);
        lapinv(c, A + lev * ng, sL, c->psiB + lev * ng, sL, c->sD, 0, nc, EPI_STORE, nullptr,
               nullptr, IBH_FP64);
        if (lev < L - 1) launch_edges_adj(d, c->psiB + lev * ng, sL, c->bcB + lev * d.nbc, sB, nc, c->s);
    }
    const double *acc = c->psiB + (L - 1) * ng;
    for (int lev = L - 2; lev >= 1; --lev) {
        launch_coarsify_adj(d, acc, sL, c->psiB + lev * ng, sL, A + lev * ng, sL, nc, c->s);
        acc = A + lev * ng;
    }
    launch_coarsify_adj(d, acc, sL, nullptr, 0, c->g1B, ng, nc, c->s);
    lapinv(c, c->g1B, ng, Zout, ng, c->sAinv, 0, nc, EPI_STORE, nullptr, nullptr, IBH_FP64);
}

// EQ = E·Q, EV = E·V at the regularization's current points; with moving_only the rows of
// the moving bodies' unknowns (those of fixed bodies keep the values of the last full pass)
void lr_points(ibh_ctx *c, bool moving_only = false) {
    const Dom &d = c->d;
    const int nf = 2 * c->nb;
    auto rows = [&](int t0, int tn) {
        launch_E(d, c->reg, c->lrQ, d.nq, 0.0, 0.0, c->lrEQ, nf, c->lr_k, nullptr, c->s, nullptr,
                 1.0, nullptr, nullptr, 0, t0, tn);
        launch_E(d, c->reg, c->lrV, d.nq, 0.0, 0.0, c->lrEV, nf, c->lr_k, nullptr, c->s, nullptr,
                 1.0, nullptr, nullptr, 0, t0, tn);
    };
    if (!moving_only) {
        rows(0, nf);
        return;
    }
    for (const auto &run : c->Jrange) rows(run.first, run.second);
}

void lr_release(ibh_ctx *c) {
    for (double **p : {&c->lrQ, &c->lrV, &c->lrEQ, &c->lrEV, &c->lrW}) {
        dfree(c, *p);
        *p = nullptr;
    }
    c->lr_valid = false;
    c->lr_k = 0;
}

// level 1's combined scale κ·sAinv₁ ⊙ sD (Δinv·Ainv₁ in the sine basis, κ = nx·ny/4 from
// S·S) and zero boundary values, for the one-level products
int ensure_level1_tables(ibh_ctx *c) {
    const Dom &d = c->d;
    const long ng = d.ng;
    if (!c->sAD) {
        ALLOC(c->sAD, ng);
        hipLaunchKernelGGL(k_scale_product, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0,
                           c->s, c->sAD, c->sAinv, c->sD, 0.25 * d.nx * d.ny, ng);
    }
    if (!c->zerobc) ALLOC(c->zerobc, d.nbc);
    return IBH_OK;
}

// the factorization for the current body box, with the regularization of the current points
int lr_build(ibh_ctx *c) {
    const Dom &d = c->d;
    const int L = c->L, nf = 2 * c->nb;
    const long ng = d.ng, nq = d.nq;
    const double t0 = now_s();
    HIPCK(hipStreamSynchronize(c->s));
    drop_graphs(c);
    lr_release(c);
    if (int rc0 = ensure_level1_tables(c)) return rc0;
    int rc = ensure_B_work(c, LR_NC);
    if (rc) return rc;
    // validation forces (host random, fixed seed) and their full products
    std::vector<double> zh(3 * (size_t)nf), xf(3 * (size_t)nf), xl(3 * (size_t)nf);
    unsigned long long st = 0x2545F4914F6CDD1Dull;
    for (auto &v : zh) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        v = (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    double *zd = nullptr, *xd = nullptr, *tau = nullptr, *Om = nullptr, *Z = nullptr, *A = nullptr;
    ALLOC(zd, 3 * (size_t)nf);
    ALLOC(xd, 3 * (size_t)nf);
    HIPCK(hipMemcpyAsync(zd, zh.data(), sizeof(double) * 3 * nf, hipMemcpyHostToDevice, c->s));
    c->lr_bypass = true;
    for (int j = 0; j < 3; ++j) B_times(c, zd + (long)j * nf, 0, xd + (long)j * nf, 0, 1, nullptr);
    c->lr_bypass = false;
    HIPCK(hipMemcpyAsync(xf.data(), xd, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    ALLOC(A, (size_t)LR_NC * L * ng);
    bool ok = false;
    for (int k = LR_K0; k <= LR_KMAX && !ok; k *= 2) {
        lr_release(c);
        dfree(c, tau); dfree(c, Om); dfree(c, Z);
        ALLOC(c->lrQ, (size_t)k * nq);
        ALLOC(c->lrV, (size_t)k * nq);
        ALLOC(c->lrEQ, (size_t)k * nf);
        ALLOC(c->lrEV, (size_t)k * nf);
        ALLOC(c->lrW, (size_t)k * LR_MAXNC);
        ALLOC(tau, k);
        ALLOC(Om, (size_t)k * ng);
        ALLOC(Z, (size_t)k * ng);
        c->lr_k = k;
        launch_fill_random_box(d, Om, ng, k, 0x1234567ull + (unsigned long long)k, c->bxa, c->bxb,
                               c->bya, c->byb, c->s);
        for (int j = 0; j < k; j += LR_NC)
            lr_T(c, Om + (long)j * ng, c->lrQ + (long)j * nq, std::min(LR_NC, k - j));
        rocblas_handle rb = rb_handle(c);
        if (!rb) return fail(c, IBH_EDEVICE, "rocBLAS handle");
        if (rocsolver_dgeqrf(rb, (rocblas_int)nq, k, c->lrQ, (rocblas_int)nq, tau) !=
                rocblas_status_success ||
            rocsolver_dorgqr(rb, (rocblas_int)nq, k, k, c->lrQ, (rocblas_int)nq, tau) !=
                rocblas_status_success)
            return fail(c, IBH_EDEVICE, "rocSOLVER QR of the coarse-level range");
        // the reflectors may leave Q nonzero off the box, where T's range is zero
        launch_box_diff(d, c->lrQ, nullptr, nq, c->lrQ, k, c->bxa + 1, c->bxb + 2, c->bya + 1,
                        c->byb + 2, c->s);
        for (int j = 0; j < k; j += LR_NC)
            lr_Tt(c, c->lrQ + (long)j * nq, Z + (long)j * ng, std::min(LR_NC, k - j), A);
        launch_curl_bc(d, Z, ng, c->zerobc, 0, c->lrV, nq, k, nullptr, c->s);
        clear_box_work(c);  // lr_T and lr_Tt used qB and g1B over the whole grid
        lr_points(c);
        c->lr_valid = true;
        for (int j = 0; j < 3; ++j)
            B_times(c, zd + (long)j * nf, 0, xd + (long)j * nf, 0, 1, nullptr);
        HIPCK(hipMemcpyAsync(xl.data(), xd, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost, c->s));
        HIPCK(hipStreamSynchronize(c->s));
        double em = 0.0, xm = 0.0;
        for (size_t i = 0; i < xf.size(); ++i) {
            em = std::max(em, std::fabs(xl[i] - xf[i]));
            xm = std::max(xm, std::fabs(xf[i]));
        }
        c->lr_err = em / std::max(xm, 1e-300);
        ok = c->lr_err <= LR_TOL && std::isfinite(c->lr_err);
        char line[160];
        std::snprintf(line, sizeof line,
                      "low-rank coarse levels: rank %d, B-product rel. err %.2e (%s), %.3f s", k,
                      c->lr_err, ok ? "accepted" : "rejected", now_s() - t0);
        c->tune_log.push_back(line);
        if (!ok) c->lr_valid = false;
    }
    dfree(c, tau); dfree(c, Om); dfree(c, Z); dfree(c, A); dfree(c, zd); dfree(c, xd);
    if (!ok) {
        lr_release(c);
        c->lr_failed = true;
    }
    ++c->lr_builds;
    HIPCK(hipStreamSynchronize(c->s));
    c->lr_build_s += now_s() - t0;
    return IBH_OK;
}

// A low-rank coarse part is used by CG force solves with the body box on, over every level,
// when the box's flux faces all take curl!'s interior formula (fluid-ops.jl:138-182): the
// faces next to the domain edge read the boundary values directly, a path T leaves out.
bool lr_wanted(const ibh_ctx *c) {
    const bool inner = c->bxa >= 1 && c->bxb <= c->d.m - 2 && c->bya >= 1 &&
                       c->byb <= c->d.p - 2;
    return c->lr_on && !c->lr_failed && c->use_cg && c->box_on && inner && c->use_eo &&
           c->sparse_B && c->lines_B && c->L >= 2 && c->nb > 0;
}

int precompute_binv(ibh_ctx *c) {
    const int nf = 2 * c->nb;
    if (nf == 0) return IBH_OK;
    double *Bmat = nullptr;
    ALLOC(Bmat, (size_t)nf * nf);
    if (!c->Binv) ALLOC(c->Binv, (size_t)nf * nf);
    int rc = B_columns(c, Bmat, 0, nf);
    if (rc) return rc;
    rc = invert(c, Bmat, c->Binv);  // Binv = inv(Bmat), cnab/ops.jl:96
    if (rc) return rc;
    launch_transpose_inplace(nf, c->Binv, c->s);  // the GEMV reads Binv by rows
    dfree(c, Bmat);
    return release_B_work(c);
}

// PCG preconditioner: the first build assembles every column of B; later builds refresh the
// moving bodies' columns and take their rows from the fresh columns (B is symmetric up to
// the multilevel boundary coupling), then invert.
int refresh_pc(ibh_ctx *c) {
    const int nf = 2 * c->nb;
    int rc;
    if (!c->Bfull) {
        ALLOC(c->Bfull, (size_t)nf * nf);
        ALLOC(c->Bwork, (size_t)nf * nf);
        ALLOC(c->Pc, (size_t)nf * nf);
        std::vector<int> inJ(nf, 0);
        for (int j : c->J) inJ[j] = 1;
        ALLOC(c->dJ, c->J.size() + 1);
        ALLOC(c->dinJ, nf);
        if (!c->J.empty())
            HIPCK(hipMemcpyAsync(c->dJ, c->J.data(), sizeof(int) * c->J.size(),
                                 hipMemcpyHostToDevice, c->s));
        HIPCK(hipMemcpyAsync(c->dinJ, inJ.data(), sizeof(int) * nf, hipMemcpyHostToDevice, c->s));
        HIPCK(hipStreamSynchronize(c->s));
    }
    const double t0 = now_s();
    // fewer levels only where rebuilds recur (moving bodies); a fixed set is built once, exact
    const int lv = c->J.empty() ? -1 : c->pc_levels;
    if (!c->pc_valid) {
        rc = B_columns(c, c->Bfull, 0, nf, c->pc_prec, lv);
        if (rc) return rc;
    } else {
        for (auto &r : c->Jrange) {
            rc = B_columns(c, c->Bfull, r.first, r.second, c->pc_prec, lv);
            if (rc) return rc;
        }
        launch_mirror_rows(nf, c->Bfull, c->dJ, (int)c->J.size(), c->dinJ, c->s);
    }
    launch_copy(c->Bwork, c->Bfull, (long)nf * nf, c->s);
    HIPCK(hipStreamSynchronize(c->s));
    const double t1 = now_s();
    // P only steers the iteration: the Cholesky inverse of the (symmetric) assembled B
    // (IBH_PC_INV=spd|lu forces a route)
    bool spd = nf > 768;
    if (const char *pi = std::getenv("IBH_PC_INV")) spd = std::strcmp(pi, "spd") == 0;
    rc = spd ? invert_spd(c, c->Bwork, c->Pc) : invert(c, c->Bwork, c->Pc);
    if (rc) return rc;
    launch_transpose_inplace(nf, c->Pc, c->s);  // the PCG step reads P by rows
    HIPCK(hipStreamSynchronize(c->s));
    const double t2 = now_s();
    ++c->pc_rebuilds;
    c->pc_time_cols += t1 - t0;
    c->pc_time_inv += t2 - t1;
    c->pc_time_total += t2 - t0;
    c->pc_valid = true;
    c->steps_since_pc = 0;
    return IBH_OK;
}

void drop_graphs(ibh_ctx *c);

// Body box of the B-products (host positions r, x then y): the interior indices of the points'
// cells ± BOX_REACH (the 13-point delta stencil of E and Eᵀ, 6 cells, plus the neighbour of
// curl and rot and the node offset) and BOX_SLACK cells more for the motion.  It is recomputed
// when a point comes within BOX_REACH of its edge; the captured graphs carry it in their
// launches, so they are dropped then (after the stream drains).
constexpr int BOX_REACH = 10, BOX_SLACK = 16;
int update_box(ibh_ctx *c, const double *r) {
    const int nb = c->nb;
    bool ok = c->box_B && c->use_eo && c->sparse_B && nb > 0;
    int xa = 0, xb = 0, ya = 0, yb = 0;
    for (int k = 0; ok && k < nb; ++k) {
        const double fx = std::floor((r[k] - c->x0) / c->h), fy = std::floor((r[nb + k] - c->y0) / c->h);
        if (!(std::fabs(fx) < 1e9 && std::fabs(fy) < 1e9)) {
            ok = false;  // non-finite or far outside: the regularization reports it
            break;
        }
        const int ix = (int)fx - 1, iy = (int)fy - 1;
        if (k == 0) { xa = xb = ix; ya = yb = iy; }
        xa = std::min(xa, ix); xb = std::max(xb, ix);
        ya = std::min(ya, iy); yb = std::max(yb, iy);
    }
    const int m = c->d.m, p = c->d.p;
    if (ok && c->box_on && xa - BOX_REACH >= c->bxa && xb + BOX_REACH <= c->bxb &&
        ya - BOX_REACH >= c->bya && yb + BOX_REACH <= c->byb)
        return IBH_OK;
    if (!ok && !c->box_on) return IBH_OK;
    HIPCK(hipStreamSynchronize(c->s));
    drop_graphs(c);
    c->lr_valid = false;  // the low-rank coarse part lives on the box: rebuilt at the next step
    c->box_on = ok;
    if (ok) {
        c->bxa = std::max(0, xa - BOX_REACH - BOX_SLACK);
        c->bxb = std::min(m - 1, xb + BOX_REACH + BOX_SLACK);
        c->bya = std::max(0, ya - BOX_REACH - BOX_SLACK);
        c->byb = std::min(p - 1, yb + BOX_REACH + BOX_SLACK);
        if (c->bxa > c->bxb || c->bya > c->byb) c->box_on = false;
    }
    if (c->box_on) clear_box_work(c);
    return IBH_OK;
}

// Host check of body points before anything on the device reads them (the reference raises
// BoundsError at the first E or Eᵀ, structure-ops.jl:70-72, 97-99, before writing anything):
// every point must be finite with its 13×13 delta support inside the finest grid, computed as
// k_reg_update computes body_idx.  A failing set of points is not uploaded.
//
// It also bounds the fixed-point spreading of Eᵀ: a contribution is at most
// max δh(x)·δh(y) < 0.274 times 2^51 (the scale puts max|f| below 2^51), so an int64 cell
// cannot overflow while fewer than 2^63 / (0.274·2^51) ≈ 14900 points reach one cell; the
// check asks for half of that.  With at most SPREAD_MAX_COVER points in all it holds
// trivially; beyond, the points in each occupied cell's ±2·6 window are counted.
constexpr int SPREAD_MAX_COVER = 7000;
int check_points(ibh_ctx *c, const double *r) {
    const int nb = c->nb;
    std::vector<std::pair<int, int>> cells(nb);
    for (int k = 0; k < nb; ++k) {
        const double fx = std::floor((r[k] - c->x0) / c->h);
        const double fy = std::floor((r[nb + k] - c->y0) / c->h);
        const bool fin = std::fabs(fx) < 1e9 && std::fabs(fy) < 1e9;
        if (!fin || !support_inside(c->d, (int)fx, (int)fy)) {
            char msg[200];
            std::snprintf(msg, sizeof msg,
                          "body point %d at (%.17g, %.17g): its delta support leaves the finest "
                          "grid (BoundsError)", k + 1, r[k], r[nb + k]);
            return fail(c, IBH_EBOUNDS, msg);
        }
        cells[k] = {(int)fx, (int)fy};
    }
    if (nb > SPREAD_MAX_COVER) {
        std::map<std::pair<int, int>, int> cnt;
        for (auto &e : cells) ++cnt[e];
        const int w = 2 * DELTA_SUPP;
        for (auto &e : cnt) {
            long n = 0;
            for (auto it = cnt.lower_bound({e.first.first - w, e.first.second - w});
                 it != cnt.end() && it->first.first <= e.first.first + w; ++it)
                if (std::abs(it->first.second - e.first.second) <= w) n += it->second;
            if (n > SPREAD_MAX_COVER)
                return fail(c, IBH_EINVAL,
                            "more than " + std::to_string(SPREAD_MAX_COVER) +
                                " body points within one delta support: the fixed-point "
                                "spreading would overflow");
        }
    }
    return IBH_OK;
}

int upload_points(ibh_ctx *c, const double *r, const double *v) {
    if (c->nb == 0) return IBH_OK;
    if (r) {
        int rc = check_points(c, r);
        if (rc) return rc;
        c->hpts.assign(r, r + 2 * c->nb);
        c->pts_ok = true;
        rc = update_box(c, r);
        if (rc) return rc;
    }
    // through pinned staging, so the copies are queued behind the running step without waiting
    // for it (the staging is reused only after the next step's status read has passed them)
    const size_t nf = 2 * (size_t)c->nb;
    if (r) {
        std::memcpy(c->pin_pts, r, sizeof(double) * nf);
        HIPCK(hipMemcpyAsync(c->pts, c->pin_pts, sizeof(double) * nf, hipMemcpyHostToDevice, c->s));
    }
    if (v) {
        std::memcpy(c->pin_pts + nf, v, sizeof(double) * nf);
        HIPCK(hipMemcpyAsync(c->vel, c->pin_pts + nf, sizeof(double) * nf, hipMemcpyHostToDevice,
                             c->s));
    }
    return IBH_OK;
}

// ---------------------------------------------------------------------- the step
// advance!(state, solver, t), cnab/timestepping.jl:253-290, as three segments that only
// enqueue device work (no host synchronisation inside), so each can be captured in a HIP
// graph: trial (+ start of the force solve), CG iterations, and post (stress, projection,
// vort2flux).  The freestream is read from the device parameter block uv.

// cycle!(nonlin) (timestepping.jl:81) and Γ = Γs (timestepping.jl:161) are pointer swaps
inline void trial_swap(ibh_ctx *c) { std::swap(c->nl[0], c->nl[1]); }
inline void post_swap(ibh_ctx *c) { std::swap(c->gam, c->gs); }

void seg_trial(ibh_ctx *c) {
    const Dom &d = c->d;
    const int L = c->L;
    const long ng = d.ng;
    // update!(reg) for moving bodies (timestepping.jl:262); points came from set_body_points
    if (c->anymove) {
        launch_reg_update(d, c->reg, c->pts, c->h, c->x0, c->y0, c->s);
        if (c->lr_valid) lr_points(c, true);  // the low-rank coarse part at the new points
    }
    // ---- GetTrialState (timestepping.jl:28-79), all levels batched
    if (L > 1) {
        launch_get_bc(d, c->gam + ng, ng, c->bcraw, d.nbc, L - 1, nullptr, c->s);
        launch_bc_prefix(d, c->bcraw, c->accbc, c->fac, L, c->s);
    }
    // bc of the coarsest level is zero; bcraw[L-1] stays zero
    // q of the previous step formed from its ψ and ψbc (or the stored q the caller set)
    launch_nonlinear(d, c->q, d.nq, c->gam, ng, c->bcraw, d.nbc, 0.25, c->nl[0], ng, 0.0, 0.0, c->h,
                     0, L, c->s, 0, c->uv, c->q_given ? nullptr : c->psi, ng, c->psibc, d.nbc);
    RhsExtras rx{c->nl[0], c->nl[1], c->accbc, c->dt * c->beta[0], c->dt * c->beta[1]};
    // A_ℓΓ by its exact stencil, fused with the right-hand side terms; then Ainv_ℓ
    launch_rhs(d, c->gam, rx.N1, rx.N2, rx.bcacc, c->rhsb, L, c->h, c->dt, c->Re, rx.f1, rx.f2,
               c->s);
    // without refinement Γs₁ = S F S exactly for the trial spectrum F, which is kept (specT)
    // for the Poisson solves of level 1 in this step's two Vort2Flux passes
    const bool fuse = c->use_eo && c->refine == 0;
    if (fuse) {
        LapInvPlan P;
        lapinv_plan(c, c->rhsb, ng, c->gs, ng, c->sAinv, ng, L, EPI_STORE, nullptr, nullptr, &P);
        P.g[1].C = c->specT;
        P.g[2].B = c->specT;
        for (int k = 0; k < 4; ++k)
            gemm(c, P.g[k], P.epi[k], c->precision, P.kind[k] == 1, P.kind[k]);
    } else {
        lapinv(c, c->rhsb, ng, c->gs, ng, c->sAinv, ng, L, EPI_STORE, nullptr, nullptr);
    }
    // low-precision transforms: the finest level's two solves are refined against their FP64
    // residuals (the forces see the trial flux at the body through Binv, which amplifies
    // high-frequency noise there; coarser levels reach it only through smooth boundary values)
    if (c->refine)
        refine_sweeps(c, c->gs, c->rhsb, c->dt / (2 * c->Re * c->h * c->h), c->sAinv, ng,
                      std::min(c->refine_levels, L), c->refine, c->precision);
    trial_swap(c);
    vort2flux(c, c->gs, c->psi, c->qs, c->psibc, 1, false, nullptr, -1, c->refine, -1, false,
              fuse ? c->specT : nullptr, true);
    // ---- BoundaryForcer right-hand side (timestepping.jl:106-139): E(qs + q0) [- v·h]
    const int nf = 2 * c->nb;
    if (nf) {
        launch_E(d, c->reg, c->qs, d.nq, 0.0, 0.0, c->Ftmp, nf, 1, nullptr, c->s, c->uv, c->h);
        if (c->anymove) launch_axpy_sub(nf, c->Ftmp, c->vel, c->h, c->s);
    }
}

// dense: Fb = Binv·Ftmp; CG: r = b − B·x0 with the previous Fb as warm start, then init
void seg_solve_begin(ibh_ctx *c) {
    const int nf = 2 * c->nb;
    if (nf == 0) return;
    if (!c->use_cg) {
        launch_gemv_t(nf, c->Binv, c->Ftmp, c->Fb, nullptr, c->s);
        return;
    }
    CGDev &cg = c->cg;
    cg.P = c->use_pcg ? c->Pc : nullptr;
    B_times(c, cg.x, 0, cg.c, 0, 1, nullptr);
    launch_cg_init(cg, c->s);
}

void seg_cg_iters(ibh_ctx *c, int k) {
    CGDev &cg = c->cg;
    // PCG over low-rank products: the direction kernel sets u's Eᵀ scale, and each product
    // follows one (the initial residual's or the previous iteration's) that cleared the image
    const bool chain = c->use_pcg && lr_path(c, IBH_FP64, -1, 1);
    cg.etscale = chain ? c->etscale : nullptr;
    for (int i = 0; i < k; ++i) {
        if (c->use_pcg) launch_pcg_dir(cg, c->s);
        else launch_cg_dir(cg, c->s);
        B_times(c, cg.u, 0, cg.c, 0, 1, cg.done, IBH_FP64, -1, chain);
        launch_cg_update(cg, c->s);
    }
    cg.etscale = nullptr;
}

void seg_post(ibh_ctx *c) {
    const Dom &d = c->d;
    const long ng = d.ng;
    const int nf = 2 * c->nb;
    // update_stress! (streamfcn/states.jl:53-84)
    if (nf)
        launch_stress(c->nbodies, c->d_boff, c->d_npts, c->nb, c->Fb, c->ds, c->h, c->dt, c->CD,
                      c->CL, c->fbs, c->s);
    // ---- CircProjecter (timestepping.jl:156-171): Γ = Γs; Γ₁ -= Ainv₁ C' Eᵀ Fb
    post_swap(c);
    const bool fuse = c->use_eo && c->refine == 0;  // Γ₁'s spectrum known (seg_trial)
    if (nf) {
        const GBox box = body_gbox(c);
        launch_ET_image(d, c->reg, c->Fb, nf, c->qtmp, d.nq, 1, nullptr, c->etscale, c->s,
                        nullptr, nullptr, 0, box);
        launch_rot_image(d, c->qtmp, d.nq, c->etscale, c->g1, ng, 1, nullptr, c->s, box);
        if (fuse) {
            // Γ₁ = S(F_trial − F_proj)S: keep the projection's spectrum and subtract it
            LapInvPlan P;
            lapinv_plan(c, c->g1, ng, c->gam, ng, c->sAinv, 0, 1, EPI_SUBFROM, nullptr, nullptr,
                        &P);
            P.g[1].C = c->specP;
            P.g[2].B = c->specP;
            if (c->box_on) {  // rot Eᵀ Fb vanishes outside the body box
                P.g[0].klo = c->bxa / 2;
                P.g[0].khi = c->bxb / 2 + 1;
                P.g[1].klo = c->bya / 2;
                P.g[1].khi = c->byb / 2 + 1;
                rows_for_pairs(P.g[0], P.g[1]);
            }
            for (int k = 0; k < 4; ++k)
                gemm(c, P.g[k], P.epi[k], IBH_FP64, P.kind[k] == 1, P.kind[k]);
        } else {
            lapinv(c, c->g1, ng, c->gam, ng, c->sAinv, 0, 1, EPI_SUBFROM, nullptr, nullptr,
                   IBH_FP64);
        }
    }
    // ---- vort2flux!(qty) (fluid-ops.jl:438); the finest level is refined like the trial's (its
    // flux feeds the next step's nonlinear term next to the bodies)
    // level 1's spectrum is F_trial − F_proj (formed where the spectral scale reads it)
    vort2flux(c, c->gam, c->psi, c->q, c->psibc, 1, true, nullptr, -1, c->refine_final, -1,
              false, fuse ? c->specT : nullptr, false, c->refine_final_levels,
              c->mixed ? IBH_FP32 : -1, true, fuse && nf ? c->specP : nullptr);
}

// status word: [CG done, CG iterations, error flags]
// The status words (and the error flag among them) are read through pinned host memory: one
// copy into it instead of a staged copy into pageable memory, on the path of every CG step.
int read_status(ibh_ctx *c, int *st) {
    int *h = c->pin_stat ? c->pin_stat : st;
    HIPCK(hipMemcpyAsync(h, c->status, 3 * sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    if (h != st) std::memcpy(st, h, 3 * sizeof(int));
    return IBH_OK;
}

int flag_error(ibh_ctx *c, int f) {
    if (!f) return IBH_OK;
    HIPCK(hipMemsetAsync(c->derr, 0, sizeof(int), c->s));
    if (f & 1) return fail(c, IBH_EBOUNDS, "a body point's delta support leaves the finest grid");
    if (f & 2) return fail(c, IBH_ENONFINITE, "infinite flux");
    return fail(c, IBH_ESINGULAR, "singular B matrix");
}

// Run one segment: eagerly, or from its graph (captured on first use, then replayed).  The
// host-side pointer swaps inside a segment are redone by `after` on replay.
template <class F>
int run_seg(ibh_ctx *c, hipGraphExec_t *slot, bool use_graph, F body, void (*after)(ibh_ctx *)) {
    if (!use_graph) {
        body();
        return IBH_OK;
    }
    if (*slot) {
        HIPCK(hipGraphLaunch(*slot, c->s));
        if (after) after(c);
        return IBH_OK;
    }
    HIPCK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    body();
    c->capturing = false;
    hipGraph_t g = nullptr;
    HIPCK(hipStreamEndCapture(c->s, &g));
    hipError_t e = hipGraphInstantiate(slot, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        *slot = nullptr;
        return fail(c, IBH_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    HIPCK(hipGraphLaunch(*slot, c->s));
    return IBH_OK;
}

void drop_graphs(ibh_ctx *c) {
    for (hipGraphExec_t *g : {&c->gTrial[0], &c->gTrial[1], &c->gIter[0], &c->gIter[1],
                              &c->gIter[2], &c->gIter[3], &c->gIter[4], &c->gPost[0], &c->gPost[1]})
        if (*g) {
            hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

int step(ibh_ctx *c, double t, double ux, double uy) {
    // moving bodies: points that were never checked (the description's body-frame points,
    // when no ibh_set_body_points came first) are rejected before anything is enqueued
    if (c->anymove && c->nb && !c->pts_ok) return check_points(c, c->hpts.data());
    if (!c->reg_tuned && c->nb) {  // first step: the points are valid now
        if (c->anymove) launch_reg_update(c->d, c->reg, c->pts, c->h, c->x0, c->y0, c->s);
        tune_reg(c);
    }
    if (lr_wanted(c) && !c->lr_valid) {  // first CG step, or the body box moved
        if (c->anymove) launch_reg_update(c->d, c->reg, c->pts, c->h, c->x0, c->y0, c->s);
        int rc0 = lr_build(c);
        if (rc0) return rc0;
    }
    c->ux = ux;
    c->uy = uy;
    launch_fill(c->uv, 1, ux, c->s);
    launch_fill(c->uv + 1, 1, uy, c->s);
    const bool refresh = c->use_pcg && (!c->pc_valid || c->steps_since_pc >= c->pc_refresh);
    // graphs from the second step on (the first one runs the autotuner), not while profiling
    // and not on steps that rebuild the preconditioner
    const bool g = c->use_graphs && !c->prof.on && !refresh && c->steps_done >= 1 && !c->q_given;
    const int par = c->parity;
    // CG iterations: k1 in the trial segment; for PCG then as many as the last step needed
    // (the count drifts slowly), queued as chunks of 8, 4, 2 and 1 before the first status
    // read, and chunks of k2 until converged (each status read is a host round trip)
    const int k1 = c->use_pcg ? 2 : 16, k2 = c->use_pcg ? 2 : 16;
    auto chunk = [&](int k) {  // k ∈ {1, 2, 4, 8, 16}
        const int slot = k == 1 ? 0 : k == 2 ? 1 : k == 4 ? 2 : k == 8 ? 3 : 4;
        return run_seg(c, &c->gIter[slot], g, [&] { seg_cg_iters(c, k); }, nullptr);
    };
    int rc;
    if (refresh) {  // eager: the preconditioner needs this step's regularization
        seg_trial(c);
        HIPCK(hipStreamSynchronize(c->s));
        const bool first = !c->pc_valid;
        const double t0 = now_s();
        if ((rc = refresh_pc(c))) return rc;
        // the adaptive rule weighs later refreshes (moving bodies' columns only)
        const double nf = 2.0 * c->nb;
        c->pc_t_refresh = (now_s() - t0) * (first ? c->J.size() / std::max(nf, 1.0) : 1.0);
        c->pc_extra = 0.0;
        c->pc_base = 1 << 30;
        if ((rc = ensure_B_work(c, 1))) return rc;
        const double t1 = now_s();
        seg_solve_begin(c);
        if (c->use_cg) seg_cg_iters(c, k1);
        int s0[3];
        if ((rc = read_status(c, s0))) return rc;
        // one B-product per iteration plus the initial residual
        c->pc_t_iter = (now_s() - t1) / (std::min(s0[1], k1) + 1);
    } else {
        if (c->use_cg && (rc = ensure_B_work(c, 1))) return rc;
        rc = run_seg(c, &c->gTrial[par], g, [&] {
            seg_trial(c);
            seg_solve_begin(c);
            if (c->use_cg) seg_cg_iters(c, k1);
        }, trial_swap);
        if (rc) return rc;
    }
    int st[3] = {0, 0, 0};
    if (c->use_cg) {
        if (c->use_pcg && !refresh) {
            int ahead = std::max(0, (c->cg_last > 0 ? c->cg_last : 8) - k1);
            for (int k = 8; k >= 1; k /= 2)
                while (ahead >= k) {
                    if ((rc = chunk(k))) return rc;
                    ahead -= k;
                }
        }
        if ((rc = read_status(c, st))) return rc;
        while (!st[0]) {
            if ((rc = chunk(k2))) return rc;
            if ((rc = read_status(c, st))) return rc;
        }
        c->cg_last = st[1];
        if (c->use_pcg) {
            ++c->steps_since_pc;
            if (c->pc_adaptive) {
                // rent or buy: rebuild once the iterations spent above the post-rebuild
                // minimum have cost as much as the last rebuild (at most twice the best
                // fixed interval's cost for a steadily drifting B)
                c->pc_base = std::min(c->pc_base, st[1]);
                c->pc_extra += (st[1] - c->pc_base) * c->pc_t_iter;
                if (c->pc_extra >= c->pc_t_refresh) c->steps_since_pc = c->pc_refresh;
            }
            // a preconditioner that has drifted far from B is rebuilt at the next step
            if (st[1] > 24) c->steps_since_pc = c->pc_refresh;
            // a fresh one that is already slow lacks coarse-level coupling: add a level
            if (refresh && !c->J.empty() && st[1] > 8 && c->pc_levels > 0 &&
                c->pc_levels < c->L) {
                ++c->pc_levels;
                c->steps_since_pc = c->pc_refresh;
            }
        }
    }
    if ((rc = run_seg(c, &c->gPost[par], g, [&] { seg_post(c); }, post_swap))) return rc;
    c->q_given = false;
    c->q_stale = true;
    c->parity ^= 1;
    ++c->steps_done;
    c->t = t;
    if (c->use_cg) {
        // the step is complete, as after the reference's cg!, which returns its last iterate
        // without raising; the status says why the iteration stopped
        int rc2 = flag_error(c, st[2]);
        if (rc2) return rc2;
        if (st[0] == CG_NONFINITE)
            return fail(c, IBH_ENONFINITE, "non-finite residual in the force solve");
        if (st[0] == CG_MAXITER) {
            char msg[160];
            std::snprintf(msg, sizeof msg,
                          "CG stopped at maxiter = %d without meeting reltol = %g (the step "
                          "completed with the last iterate)", c->cg_maxiter, c->cg_reltol);
            return fail(c, IBH_ENOCONV, msg);
        }
    }
    return IBH_OK;
}

long field_len(ibh_ctx *c, int which) {
    const long L = c->L;
    switch (which) {
    case IBH_F_Q: case IBH_F_Q0: return L * c->d.nq;
    case IBH_F_GAMMA: case IBH_F_PSI: case IBH_F_NONLIN1: case IBH_F_NONLIN2: return L * c->d.ng;
    case IBH_F_FB: case IBH_F_FBSTRESS: case IBH_F_POINTS_R: case IBH_F_POINTS_V: return 2L * c->nb;
    case IBH_F_CD: case IBH_F_CL: return c->nbodies;
    }
    return -1;
}

double *field_ptr(ibh_ctx *c, int which) {
    switch (which) {
    case IBH_F_Q: return c->q;
    case IBH_F_GAMMA: return c->gam;
    case IBH_F_PSI: return c->psi;
    case IBH_F_NONLIN1: return c->nl[0];
    case IBH_F_NONLIN2: return c->nl[1];
    case IBH_F_FB: return c->Fb;
    case IBH_F_FBSTRESS: return c->fbs;
    case IBH_F_POINTS_R: return c->pts;
    case IBH_F_POINTS_V: return c->vel;
    case IBH_F_CD: return c->CD;
    case IBH_F_CL: return c->CL;
    }
    return nullptr;
}

}  // namespace

// =========================================================================== C ABI
extern "C" {

int ibh_version(void) { return IBH_ABI_VERSION; }

const char *ibh_last_error(const ibh_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ibh_create(const ibh_case_desc *desc, ibh_ctx **out) {
    if (!out || !desc) return IBH_EINVAL;
    *out = nullptr;
    const ibh_case_desc &D = *desc;
    if (D.nx % 4 || D.ny % 4 || D.nx < 16 || D.ny < 16 || D.nlev < 1 || D.nlev > 30 || D.h <= 0 ||
        D.nbodies < 0)
        return IBH_EINVAL;
    if (D.precision < IBH_FP64 || D.precision > IBH_MIXED) return IBH_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return IBH_EDEVICE;
    ibh_ctx *c = new ibh_ctx();
    c->d = make_dom(D.nx, D.ny);
    c->L = D.nlev;
    c->h = D.h; c->x0 = D.x0; c->y0 = D.y0; c->Re = D.Re; c->dt = D.dt;
    c->beta[0] = D.beta[0]; c->beta[1] = D.beta[1];
    c->nbodies = D.nbodies;
    c->npts.resize(D.nbodies);
    c->moving.resize(D.nbodies);
    c->boff.resize(D.nbodies + 1);
    int nb = 0;
    for (int b = 0; b < D.nbodies; ++b) {
        c->npts[b] = D.npts[b];
        c->moving[b] = D.moving ? D.moving[b] : 0;
        c->anymove |= c->moving[b] != 0;
        c->boff[b] = nb;
        nb += D.npts[b];
    }
    c->boff[D.nbodies] = nb;
    c->nb = nb;
    c->mixed = D.precision == IBH_MIXED;
    c->precision = c->mixed ? IBH_FP64 : D.precision;
    if (const char *fv = std::getenv("IBH_GEMM_VARIANT")) c->tune_fixed = std::atoi(fv);
    if (const char *tv = std::getenv("IBH_TUNE_VERBOSE")) c->tune_verbose = std::atoi(tv) != 0;
    c->force_solver = D.force_solver;
    c->use_pcg = D.force_solver == IBH_FORCE_PCG || (D.force_solver == IBH_FORCE_AUTO && c->anymove);
    c->use_cg = D.force_solver == IBH_FORCE_CG || c->use_pcg;
    c->pc_adaptive = D.pc_refresh <= 0;
    // the PCG preconditioner only approximates B⁻¹, so its columns use FP32 transforms (PCG
    // still applies the FP64 B); IBH_PC_PREC=0/1/2 picks another precision
    c->pc_prec = IBH_FP32;
    if (const char *pp = std::getenv("IBH_PC_PREC")) c->pc_prec = std::atoi(pp);
    // ... and, for moving bodies, a Vort2Flux over the finest level only: the coarser levels
    // change B by ~1e-5 relative (measured on the 1024², 5-level plate case: the inverse of
    // the one-level B takes 4-5 PCG iterations on the five-level B where its exact inverse
    // takes 2, the two-level one 3), at a third of the transforms per column.  A fresh
    // preconditioner that needs more than 8 iterations adds a level (see step).
    // IBH_PC_LEVELS sets the starting count (≤ 0: every level).
    c->pc_levels = 1;
    if (const char *pl = std::getenv("IBH_PC_LEVELS")) c->pc_levels = std::atoi(pl);
    c->pc_refresh = D.pc_refresh > 0 ? D.pc_refresh : (1 << 30);
    if (c->precision != IBH_FP64)
        c->refine = D.refine_sweeps > 0 ? D.refine_sweeps
                    : D.refine_sweeps < 0 ? 0 : (D.precision == IBH_FP32 ? 1 : 2);
    if (const char *rf = std::getenv("IBH_REFINE")) c->refine = std::max(0, std::atoi(rf));
    c->refine_levels = D.refine_levels > 0 ? std::min(D.refine_levels, c->L) : c->L;
    {
        // the region of each level that a B-product's Vort2Flux can make nonzero: level 0 in
        // full, level ℓ the coarsify! image (fluid-ops.jl:184-202) of level ℓ-1 — coarse ic
        // (1-based) reads fine i-1..i+1 with i = 2(ic - n/4) and is written for
        // ic in [n/4+1, 3n/4-1] — widened by one cell; 0-based [lo, hi)
        auto coarse = [](int n, int lo, int hi, int *clo, int *chi) {
            const int a = std::max(n / 4 + 1, n / 4 + (lo + 1) / 2) - 1;
            const int b = std::min(3 * n / 4 - 1, n / 4 + (hi + 1) / 2) + 1;
            *clo = std::max(1, a) - 1;
            *chi = std::min(n - 1, b);
        };
        c->bx_lo.assign(c->L, 0);
        c->bx_hi.assign(c->L, D.nx - 1);
        c->by_lo.assign(c->L, 0);
        c->by_hi.assign(c->L, D.ny - 1);
        for (int l = 1; l < c->L; ++l) {
            coarse(D.nx, c->bx_lo[l - 1], c->bx_hi[l - 1], &c->bx_lo[l], &c->bx_hi[l]);
            coarse(D.ny, c->by_lo[l - 1], c->by_hi[l - 1], &c->by_lo[l], &c->by_hi[l]);
        }
        if (const char *sb = std::getenv("IBH_SPARSE_B")) c->sparse_B = std::atoi(sb) != 0;
        if (const char *bb = std::getenv("IBH_BODY_BOX")) c->box_B = std::atoi(bb) != 0;
        if (const char *sl = std::getenv("IBH_SPARSE_LINES")) c->lines_B = std::atoi(sl) != 0;
        if (const char *lr = std::getenv("IBH_LOWRANK_B")) c->lr_on = std::atoi(lr) != 0;
    }
    if (const char *rl = std::getenv("IBH_REFINE_LEVELS"))
        c->refine_levels = std::min(std::max(1, std::atoi(rl)), c->L);
    // the final Vort2Flux pass: FP32 refines level 1 only (measured on the 64-case ellipse
    // batch over 300 steps: CD within 1.1e-8 of FP64 as with every level refined, 3.3e-6
    // with none); bf16x3 keeps every level
    c->refine_final = c->refine;
    c->refine_final_levels = c->precision == IBH_FP32 ? 1 : c->refine_levels;
    if (const char *rf = std::getenv("IBH_REFINE_FINAL")) c->refine_final = std::max(0, std::atoi(rf));
    if (const char *rl = std::getenv("IBH_REFINE_FINAL_LEVELS"))
        c->refine_final_levels = std::min(std::max(0, std::atoi(rl)), c->L);
    {
        // unknowns of moving bodies, as contiguous runs in [fx | fy]
        for (int comp = 0; comp < 2; ++comp)
            for (int b = 0; b < D.nbodies; ++b)
                if (D.moving && D.moving[b] && D.npts[b] > 0) {
                    const int j0 = comp * nb + c->boff[b];
                    c->Jrange.push_back({j0, D.npts[b]});
                    for (int k = 0; k < D.npts[b]; ++k) c->J.push_back(j0 + k);
                }
    }
    c->cg_maxiter = D.cg_maxiter > 0 ? D.cg_maxiter : 5000;
    c->cg_reltol = D.cg_reltol > 0 ? D.cg_reltol : 1e-12;
    if (hipGetDevice(&c->dev) != hipSuccess || hipStreamCreate(&c->s) != hipSuccess) {
        delete c;
        return IBH_EDEVICE;
    }
    const Dom &d = c->d;
    const long L = c->L, ng = d.ng, nq = d.nq;
    int rc = 0;
#define CK(x) do { if ((rc = (x)) != 0) { ibh_destroy(c); return rc; } } while (0)
    CK(dalloc(c, &c->Sm, (size_t)d.m * d.m));
    c->Sp = c->Sm;
    if (d.p != d.m) CK(dalloc(c, &c->Sp, (size_t)d.p * d.p));
    {
        const size_t hx = D.nx / 2, hy = D.ny / 2;
        CK(dalloc(c, &c->Som, hx * hx));
        CK(dalloc(c, &c->Sem, hx * hx));
        CK(dalloc(c, &c->Somf, hx * hx));
        CK(dalloc(c, &c->Semf, hx * hx));
        CK(dalloc(c, &c->Somb, hx * hx));
        CK(dalloc(c, &c->Semb, hx * hx));
        CK(dalloc(c, &c->Smoe, 2 * hx * hx));
        CK(dalloc(c, &c->Smoef, 4 * hx * hx));  // and the slot-swapped copy
        c->Spoe = c->Smoe;
        c->Spoef = c->Smoef;
        c->Sop = c->Som;
        c->Sep = c->Sem;
        c->Sopf = c->Somf;
        c->Sepf = c->Semf;
        c->Sopb = c->Somb;
        c->Sepb = c->Semb;
        if (D.ny != D.nx) {
            CK(dalloc(c, &c->Sop, hy * hy));
            CK(dalloc(c, &c->Sep, hy * hy));
            CK(dalloc(c, &c->Sopf, hy * hy));
            CK(dalloc(c, &c->Sepf, hy * hy));
            CK(dalloc(c, &c->Sopb, hy * hy));
            CK(dalloc(c, &c->Sepb, hy * hy));
            CK(dalloc(c, &c->Spoe, 2 * hy * hy));
            CK(dalloc(c, &c->Spoef, 4 * hy * hy));
        }
        if (const char *e = std::getenv("IBH_DST_SPLIT")) c->use_eo = std::atoi(e) != 0;
    }
    CK(dalloc(c, &c->sD, ng));
    CK(dalloc(c, &c->sA, L * ng));
    CK(dalloc(c, &c->sAinv, L * ng));
    CK(dalloc(c, &c->gam, L * ng));
    CK(dalloc(c, &c->psi, L * ng));
    CK(dalloc(c, &c->q, L * nq));
    CK(dalloc(c, &c->nl[0], L * ng));
    CK(dalloc(c, &c->nl[1], L * ng));
    CK(dalloc(c, &c->qs, nq));
    CK(dalloc(c, &c->gs, L * ng));
    CK(dalloc(c, &c->bcraw, L * d.nbc));
    CK(dalloc(c, &c->accbc, L * d.nbc));
    CK(dalloc(c, &c->psibc, L * d.nbc));
    CK(dalloc(c, &c->gtmp, L * ng));
    CK(dalloc(c, &c->fac, L + 1));
    CK(dalloc(c, &c->qtmp, nq));
    CK(dalloc(c, &c->g1, ng));
    CK(dalloc(c, &c->etscale, 1));
    c->etcap = 1;
    CK(dalloc(c, &c->rhsb, L * ng));
    CK(dalloc(c, &c->specT, L * ng));
    CK(dalloc(c, &c->specP, ng));
    CK(dalloc(c, &c->lineS, lines_work_size(d)));
    CK(dalloc(c, &c->edgeS, 2 * (d.m + d.p)));
    if (c->refine) CK(dalloc(c, &c->rres, L * ng));
    CK(dalloc(c, &c->T1, L * ng));
    CK(dalloc(c, &c->T2, L * ng));
    c->Tcap = L * ng;
    const int nf = 2 * nb;
    CK(dalloc(c, &c->pts, nf));
    CK(dalloc(c, &c->vel, nf));
    if (nf && hipHostMalloc((void **)&c->pin_pts, sizeof(double) * 2 * nf, hipHostMallocDefault) !=
                  hipSuccess) {
        c->pin_pts = nullptr;
        return fail(c, IBH_EDEVICE, "hipHostMalloc of the points' staging buffer failed");
    }
    CK(dalloc(c, &c->ds, nb));
    CK(dalloc(c, &c->d_boff, D.nbodies + 1));
    CK(dalloc(c, &c->d_npts, D.nbodies + 1));
    CK(dalloc(c, &c->status, 4));
    c->derr = c->status + 2;
    if (hipHostMalloc((void **)&c->pin_stat, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        c->pin_stat = nullptr;
        return fail(c, IBH_EDEVICE, "hipHostMalloc of the status words failed");
    }
    CK(dalloc(c, &c->uv, 2));
    if (const char *ge = std::getenv("IBH_GRAPH")) c->use_graphs = std::atoi(ge) != 0;
    CK(dalloc(c, &c->reg.bidx, nf));
    CK(dalloc(c, &c->reg.w, (size_t)nb * 4 * 13));
    c->reg.nb = nb;
    c->reg.err = c->derr;
    c->reg.esched = c->reg.etsched = 0;
    CK(dalloc(c, &c->Fb, nf));
    CK(dalloc(c, &c->Ftmp, nf));
    CK(dalloc(c, &c->CD, D.nbodies + 1));
    CK(dalloc(c, &c->CL, D.nbodies + 1));
    CK(dalloc(c, &c->fbs, nf));
    // CG state
    CK(dalloc(c, &c->cg.r, nf));
    CK(dalloc(c, &c->cg.c, nf));
    CK(dalloc(c, &c->cg.u, nf));
    CK(dalloc(c, &c->cg.z, nf));
    CK(dalloc(c, &c->cg.part, std::max(1, pcg_parts(nf))));
    CK(dalloc(c, &c->cg.sc, 4));
    c->cg.etscale = nullptr;
    c->cg.done = c->status;
    c->cg.it = c->status + 1;
    c->cg.x = c->Fb;
    c->cg.b = c->Ftmp;
    c->cg.n = nf;
    c->cg.maxiter = c->cg_maxiter;
    c->cg.reltol = c->cg_reltol;
    // constants
    launch_sine_matrix(c->Sm, D.nx, c->s);
    if (c->Sp != c->Sm) launch_sine_matrix(c->Sp, D.ny, c->s);
    launch_sine_split(c->Som, c->Sem, D.nx, c->s);
    if (c->Sop != c->Som) launch_sine_split(c->Sop, c->Sep, D.ny, c->s);
    {
        const long hx = D.nx / 2, hy = D.ny / 2;
        launch_sine_low(c->Som, c->Somf, c->Somb, hx * hx, c->s);
        launch_sine_low(c->Sem, c->Semf, c->Semb, hx * hx, c->s);
        launch_sine_interleave(c->Som, c->Sem, c->Smoe, hx * hx, c->s);
        if (c->Spoe != c->Smoe) launch_sine_interleave(c->Sop, c->Sep, c->Spoe, hy * hy, c->s);
        if (c->Sop != c->Som) {
            launch_sine_low(c->Sop, c->Sopf, c->Sopb, hy * hy, c->s);
            launch_sine_low(c->Sep, c->Sepf, c->Sepb, hy * hy, c->s);
        }
        launch_sine_interleave(c->Somf, c->Semf, c->Smoef, hx * hx, c->s);
        if (c->Spoef != c->Smoef) launch_sine_interleave(c->Sopf, c->Sepf, c->Spoef, hy * hy, c->s);
    }
    launch_scale_tables(d, c->sD, c->sA, c->sAinv, c->L, c->h, c->dt, c->Re, c->s);
    std::vector<double> fac(L + 1, 0.0);
    for (int l = 0; l < L; ++l) {
        const double hc = c->h * std::ldexp(1.0, l);
        fac[l] = 0.25 * c->dt / (c->Re * (hc * hc));  // timestepping.jl:50
    }
    hipMemcpyAsync(c->fac, fac.data(), sizeof(double) * (L + 1), hipMemcpyHostToDevice, c->s);
    if (nb) {
        // the initial points of moving bodies may be body-frame coordinates that the first
        // ibh_set_body_points replaces: a failed check here is reported where they are used
        c->hpts.assign(D.xb, D.xb + nf);
        c->pts_ok = check_points(c, D.xb) == IBH_OK;
        c->err.clear();
        CK(update_box(c, D.xb));
        hipMemcpyAsync(c->pts, D.xb, sizeof(double) * nf, hipMemcpyHostToDevice, c->s);
        hipMemcpyAsync(c->ds, D.ds, sizeof(double) * nb, hipMemcpyHostToDevice, c->s);
    }
    std::vector<int> bo(c->boff.begin(), c->boff.end()), np(c->npts);
    np.push_back(0);
    hipMemcpyAsync(c->d_boff, bo.data(), sizeof(int) * (D.nbodies + 1), hipMemcpyHostToDevice, c->s);
    hipMemcpyAsync(c->d_npts, np.data(), sizeof(int) * (D.nbodies + 1), hipMemcpyHostToDevice, c->s);
    if (hipStreamSynchronize(c->s) != hipSuccess) {
        ibh_destroy(c);
        return IBH_EDEVICE;
    }
#undef CK
    *out = c;
    return IBH_OK;
}

void ibh_destroy(ibh_ctx *c) {
    if (!c) return;
    if (c->s) hipStreamSynchronize(c->s);
    drop_graphs(c);
    for (void *p : c->allocs) hipFree(p);
    if (c->tune_buf) hipFree(c->tune_buf);
    for (hipEvent_t e : c->prof.pool) hipEventDestroy(e);
    if (c->prof.ts) hipFree(c->prof.ts);
    if (c->rb) rocblas_destroy_handle(static_cast<rocblas_handle>(c->rb));
    if (c->pin_pts) hipHostFree(c->pin_pts);
    if (c->pin_stat) hipHostFree(c->pin_stat);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
}

int ibh_init_state(ibh_ctx *c, double t0, double ux, double uy) {
    if (!c) return IBH_EINVAL;
    const long L = c->L;
    HIPCK(hipMemsetAsync(c->q, 0, sizeof(double) * L * c->d.nq, c->s));
    HIPCK(hipMemsetAsync(c->gam, 0, sizeof(double) * L * c->d.ng, c->s));
    HIPCK(hipMemsetAsync(c->psi, 0, sizeof(double) * L * c->d.ng, c->s));
    HIPCK(hipMemsetAsync(c->psibc, 0, sizeof(double) * L * c->d.nbc, c->s));
    c->q_stale = c->q_given = false;
    HIPCK(hipMemsetAsync(c->nl[0], 0, sizeof(double) * L * c->d.ng, c->s));
    HIPCK(hipMemsetAsync(c->nl[1], 0, sizeof(double) * L * c->d.ng, c->s));
    if (c->nb) {
        HIPCK(hipMemsetAsync(c->Fb, 0, sizeof(double) * 2 * c->nb, c->s));
        HIPCK(hipMemsetAsync(c->vel, 0, sizeof(double) * 2 * c->nb, c->s));
    }
    c->ux = ux;
    c->uy = uy;
    c->t = t0;
    HIPCK(hipStreamSynchronize(c->s));
    return IBH_OK;
}

int ibh_setup(ibh_ctx *c) {
    if (!c) return IBH_EINVAL;
    // moving bodies under CG rebuild the regularization every step
    if (!(c->use_cg && c->anymove)) {
        if (c->nb && !c->pts_ok) return check_points(c, c->hpts.data());
        launch_reg_update(c->d, c->reg, c->pts, c->h, c->x0, c->y0, c->s);
        int rc = check_flags(c);
        if (rc) return rc;
        tune_reg(c);
        if (!c->use_cg) {
            rc = precompute_binv(c);
            if (rc) return rc;
        }
    }
    c->setup_done = true;
    HIPCK(hipStreamSynchronize(c->s));
    return IBH_OK;
}

// q = curl!(ψ, ψbc) on every level into the state's flux buffer, if the last step left it to be
// formed on demand
static hipError_t materialize_q(ibh_ctx *c) {
    if (!c->q_stale) return hipSuccess;
    const Dom &d = c->d;
    launch_curl_bc(d, c->psi, d.ng, c->psibc, d.nbc, c->q, d.nq, c->L, nullptr, c->s);
    c->q_stale = false;
    return hipGetLastError();
}

int ibh_set_body_points(ibh_ctx *c, const double *r, const double *v) {
    if (!c) return IBH_EINVAL;
    return upload_points(c, r, v);
}

int ibh_advance(ibh_ctx *c, double t, double ux, double uy) {
    if (!c) return IBH_EINVAL;
    if (!c->setup_done) return fail(c, IBH_EINVAL, "ibh_setup must run before ibh_advance");
    int rc = step(c, t, ux, uy);
    if (rc) return rc;
    // the CG path read the step's flags mid-step; the dense path reads them here (one 4-byte
    // copy after the step), so "infinite flux" and bounds errors surface from this call
    if (!c->use_cg) return check_flags(c);
    return IBH_OK;
}

int ibh_advance_n(ibh_ctx *c, double t0, int32_t nsteps, double ux, double uy) {
    if (!c) return IBH_EINVAL;
    if (!c->setup_done) return fail(c, IBH_EINVAL, "ibh_setup must run before ibh_advance_n");
    if (c->anymove)
        return fail(c, IBH_EINVAL,
                    "ibh_advance_n steps fixed bodies only: moving bodies need their points "
                    "for every t (ibh_set_body_points + ibh_advance)");
    for (int k = 0; k < nsteps; ++k) {
        int rc = step(c, t0 + k * c->dt, ux, uy);
        if (rc) return rc;
    }
    // flags raised in any of the steps (they stay set until read) surface here
    return check_flags(c);
}

int ibh_sync(ibh_ctx *c) {
    if (!c) return IBH_EINVAL;
    HIPCK(hipStreamSynchronize(c->s));
    return check_flags(c);
}

int ibh_get_field(ibh_ctx *c, int32_t which, double *out) {
    if (!c || !out) return IBH_EINVAL;
    const long n = field_len(c, which);
    if (n < 0) return fail(c, IBH_EINVAL, "unknown field");
    if (which == IBH_F_Q) HIPCK(materialize_q(c));
    HIPCK(hipStreamSynchronize(c->s));
    // a device error flag still set (e.g. from ibh_advance_n or an operator call) is returned
    // after the copy, so the caller gets both the values and the error
    int flags = 0;
    HIPCK(hipMemcpy(&flags, c->derr, sizeof(int), hipMemcpyDeviceToHost));
    if (which == IBH_F_Q0) {  // base_flux!, fluid-ops.jl:403-417 (never materialized)
        const Dom &d = c->d;
        for (int l = 0; l < c->L; ++l) {
            const double hc = c->h * std::ldexp(1.0, l);
            for (long k = 0; k < d.nq; ++k) out[l * d.nq + k] = k < d.nu ? c->ux * hc : c->uy * hc;
        }
        return flags ? check_flags(c) : IBH_OK;
    }
    if (n > 0)
        HIPCK(hipMemcpy(out, field_ptr(c, which), sizeof(double) * n, hipMemcpyDeviceToHost));
    return flags ? check_flags(c) : IBH_OK;
}

int ibh_set_field(ibh_ctx *c, int32_t which, const double *in) {
    if (!c || !in) return IBH_EINVAL;
    const long n = field_len(c, which);
    if (n < 0 || which == IBH_F_Q0 || which == IBH_F_CD || which == IBH_F_CL)
        return fail(c, IBH_EINVAL, "field cannot be set");
    if (n == 0) return IBH_OK;
    // the next step reads q as stored: a new ψ must not change the q it sees
    if (which == IBH_F_PSI) HIPCK(materialize_q(c));
    if (which == IBH_F_Q || which == IBH_F_PSI) {
        c->q_given = true;
        c->q_stale = false;
    }
    if (which == IBH_F_POINTS_R) {
        int rc = check_points(c, in);
        if (rc) return rc;
        c->hpts.assign(in, in + n);
        c->pts_ok = true;
        rc = update_box(c, in);
        if (rc) return rc;
    }
    HIPCK(hipMemcpyAsync(field_ptr(c, which), in, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    return IBH_OK;
}

int ibh_dims(ibh_ctx *c, int32_t *o) {
    if (!c || !o) return IBH_EINVAL;
    o[0] = c->d.nx; o[1] = c->d.ny; o[2] = c->L; o[3] = (int32_t)c->d.nq; o[4] = (int32_t)c->d.ng;
    o[5] = c->nb; o[6] = c->d.nbc; o[7] = c->nbodies;
    return IBH_OK;
}

double ibh_time(ibh_ctx *c) { return c ? c->t : 0.0; }

int ibh_cg_iterations(ibh_ctx *c) { return c ? c->cg_last : -1; }

int ibh_lowrank_stats(ibh_ctx *c, double *out4) {
    if (!c || !out4) return IBH_EINVAL;
    out4[0] = c->lr_valid ? (double)c->lr_k : 0.0;
    out4[1] = c->lr_err;
    out4[2] = (double)c->lr_builds;
    out4[3] = c->lr_build_s;
    return IBH_OK;
}

int ibh_solver_stats(ibh_ctx *c, double *out6) {
    if (!c || !out6) return IBH_EINVAL;
    const double v[6] = {(double)c->pc_rebuilds, c->pc_time_total, c->pc_time_cols,
                         c->pc_time_inv, c->pc_t_iter, c->pc_t_refresh};
    for (int k = 0; k < 6; ++k) out6[k] = v[k];
    return IBH_OK;
}

// --------------------------------------------------------------------------- operators
namespace {
struct DevBuf {
    ibh_ctx *c;
    std::vector<void *> v;
    double *get(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, (n ? n : 1) * sizeof(double)) != hipSuccess) return nullptr;
        hipMemset(p, 0, (n ? n : 1) * sizeof(double));
        v.push_back(p);
        return (double *)p;
    }
    double *up(const double *h, size_t n) {
        double *p = get(n);
        if (p && h && n) hipMemcpy(p, h, n * sizeof(double), hipMemcpyHostToDevice);
        return p;
    }
    ~DevBuf() {
        hipDeviceSynchronize();
        for (void *p : v) hipFree(p);
    }
};
int down(ibh_ctx *c, double *h, const double *d, size_t n) {
    HIPCK(hipStreamSynchronize(c->s));
    HIPCK(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return IBH_OK;
}
}  // namespace

int ibh_op_lapinv(ibh_ctx *c, int32_t mode, int32_t lev, const double *in, double *out) {
    if (!c || lev < 0 || lev >= c->L || mode < 0 || mode > 2) return IBH_EINVAL;
    DevBuf B{c};
    const long ng = c->d.ng;
    double *din = B.up(in, ng), *dout = B.get(ng);
    const double *sc = mode == 0 ? c->sD : (mode == 1 ? c->sA : c->sAinv) + lev * ng;
    lapinv(c, din, ng, dout, ng, sc, 0, 1, EPI_STORE, nullptr, nullptr);
    return down(c, out, dout, ng);
}

int ibh_op_dst2(ibh_ctx *c, const double *in, double *out) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    const Dom &d = c->d;
    double *din = B.up(in, d.ng), *dout = B.get(d.ng);
    // RODFT00 in both dimensions = 4·S_m·X·S_p: two contractions, then ×4
    GemmArgs g = gargs();
    g.M = d.m; g.N = d.p; g.K = d.m;
    g.A = c->Sm; g.lda = d.m; g.B = din; g.ldb = d.m; g.C = c->T1; g.ldc = d.p;
    gemm(c, g, EPI_STORE, c->precision, true);
    double *four = B.get(d.ng);
    launch_fill(four, d.ng, 4.0, c->s);
    g.M = d.m; g.N = d.p; g.K = d.p;
    g.A = c->T1; g.lda = d.p; g.B = c->Sp; g.ldb = d.p; g.C = dout; g.ldc = d.m;
    g.D = four; g.ldd = d.m;
    gemm(c, g, EPI_SCALE, c->precision);
    return down(c, out, dout, d.ng);
}

int ibh_op_rot(ibh_ctx *c, const double *q, double *gam) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    double *dq = B.up(q, c->d.nq), *dg = B.get(c->d.ng);
    launch_rot(c->d, dq, 0, dg, 0, 1, nullptr, c->s);
    return down(c, gam, dg, c->d.ng);
}

int ibh_op_curl(ibh_ctx *c, const double *psi, const double *psibc, double *q) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    double *dp = B.up(psi, c->d.ng), *db = B.up(psibc, c->d.nbc), *dq = B.get(c->d.nq);
    launch_curl_bc(c->d, dp, 0, db, 0, dq, 0, 1, nullptr, c->s);
    return down(c, q, dq, c->d.nq);
}

int ibh_op_coarsify(ibh_ctx *c, double *gc, const double *g) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    double *dgc = B.up(gc, c->d.ng), *dg = B.up(g, c->d.ng);
    launch_coarsify(c->d, dgc, dg, 0, 0, 1, nullptr, c->s);
    return down(c, gc, dgc, c->d.ng);
}

int ibh_op_get_bc(ibh_ctx *c, const double *r, double *rbc) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    double *dr = B.up(r, c->d.ng), *db = B.get(c->d.nbc);
    launch_get_bc(c->d, dr, 0, db, 0, 1, nullptr, c->s);
    return down(c, rbc, db, c->d.nbc);
}

int ibh_op_apply_bc(ibh_ctx *c, double *r, const double *rbc, double fac) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    double *dr = B.up(r, c->d.ng), *db = B.up(rbc, c->d.nbc);
    launch_apply_bc(c->d, dr, 0, db, 0, fac, 1, nullptr, c->s);
    return down(c, r, dr, c->d.ng);
}

int ibh_op_nonlinear(ibh_ctx *c, int32_t lev, const double *q, const double *q0,
                     const double *gam, const double *gbc, double *nonlin) {
    if (!c || lev < 0 || lev >= c->L) return IBH_EINVAL;
    // q0 must be the constant base flux of the level (fluid-ops.jl:403-417)
    const Dom &d = c->d;
    const double U = q0[0], V = q0[d.nu];
    DevBuf B{c};
    double *dq = B.up(q, d.nq), *dg = B.up(gam, d.ng), *db = B.up(gbc, d.nbc), *dn = B.get(d.ng);
    launch_nonlinear(d, dq, 0, dg, 0, db, 0, 1.0, dn, 0, U, V, c->h, lev, 1, c->s, 1);
    return down(c, nonlin, dn, d.ng);
}

int ibh_op_vort2flux(ibh_ctx *c, double *gam, double *psi, double *q) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    const Dom &d = c->d;
    const long L = c->L;
    double *dg = B.up(gam, L * d.ng), *dp = B.get(L * d.ng), *dq = B.get(L * d.nq),
           *db = B.get(L * d.nbc);
    vort2flux(c, dg, dp, dq, db, 1, true, nullptr);
    int rc = down(c, gam, dg, L * d.ng);
    if (!rc) rc = down(c, psi, dp, L * d.ng);
    if (!rc) rc = down(c, q, dq, L * d.nq);
    return rc;
}

int ibh_op_reg_update(ibh_ctx *c, const double *r) {
    if (!c) return IBH_EINVAL;
    int rc = upload_points(c, r, nullptr);
    if (rc) return rc;
    launch_reg_update(c->d, c->reg, c->pts, c->h, c->x0, c->y0, c->s);
    if (c->lr_valid) lr_points(c);
    return check_flags(c);
}

int ibh_op_E(ibh_ctx *c, const double *q, double *fb) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    double *dq = B.up(q, c->d.nq), *df = B.get(2 * c->nb);
    launch_E(c->d, c->reg, dq, 0, 0.0, 0.0, df, 0, 1, nullptr, c->s);
    return down(c, fb, df, 2 * c->nb);
}

int ibh_op_ET(ibh_ctx *c, const double *fb, double *q) {
    if (!c) return IBH_EINVAL;
    DevBuf B{c};
    double *df = B.up(fb, 2 * c->nb), *dq = B.get(c->d.nq);
    launch_ET(c->d, c->reg, df, 0, dq, 0, 1, nullptr, c->etscale, c->s);
    int rc = down(c, q, dq, c->d.nq);
    if (rc) return rc;
    return check_flags(c);
}

int ibh_op_B(ibh_ctx *c, const double *z, double *x) {
    if (!c) return IBH_EINVAL;
    int rc = lr_wanted(c) && !c->lr_valid ? lr_build(c) : IBH_OK;
    if (rc) return rc;
    rc = ensure_B_work(c, 1);
    if (rc) return rc;
    DevBuf B{c};
    const int nf = 2 * c->nb;
    double *dz = B.up(z, nf), *dx = B.get(nf);
    B_times(c, dz, 0, dx, 0, 1, nullptr);
    return down(c, x, dx, nf);
}

// ---------------------------------------------------------------- allocation-free operator
// The two intermediates of a LaplacianInv apply (T1, T2 of lapinv) for callers that cannot
// allocate: one library-wide array per GPU, held until the process ends.  Its host-side name
// is not a device address: kernels get the address hipGetSymbolAddress returns for the device
// (looked up once per device on the first call there and kept in noalloc_ptr).
constexpr long NOALLOC_MAX_NG = IBH_NOALLOC_MAX_NG;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_NG];
static constexpr int NOALLOC_MAX_DEV = 64;
static double *noalloc_ptr[NOALLOC_MAX_DEV];
static std::mutex noalloc_mu;

static int noalloc_scratch(ibh_ctx *c, double **out) {
    if (c->dev < 0 || c->dev >= NOALLOC_MAX_DEV)
        return fail(c, IBH_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[c->dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(c, IBH_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
        noalloc_ptr[c->dev] = static_cast<double *>(p);
    }
    *out = noalloc_ptr[c->dev];
    return IBH_OK;
}

int ibh_op_lapinv_noalloc(ibh_ctx *c, int32_t mode, int32_t lev, const double *in,
                          double *out) {
    if (!c || lev < 0 || lev >= c->L || mode < 0 || mode > 2 || !in || !out)
        return IBH_EINVAL;
    const long ng = c->d.ng;
    if (ng > NOALLOC_MAX_NG)
        return fail(c, IBH_EINVAL,
                    "ibh_op_lapinv_noalloc: " + std::to_string(ng) + " interior points exceed "
                    "the scratch limit of " + std::to_string(NOALLOC_MAX_NG) +
                    " (use ibh_op_lapinv)");
    if (in == out) return fail(c, IBH_EINVAL, "ibh_op_lapinv_noalloc: in and out must differ");
    int cur = 0;
    HIPCK(hipGetDevice(&cur));
    if (cur != c->dev) HIPCK(hipSetDevice(c->dev));
    double *S = nullptr;
    int rc = noalloc_scratch(c, &S);
    if (!rc) {
        const double *sc = mode == 0 ? c->sD : (mode == 1 ? c->sA : c->sAinv) + lev * ng;
        // lapinv's intermediates, pointed at the scratch array for this call only
        double *T1 = c->T1, *T2 = c->T2;
        c->T1 = S;
        c->T2 = S + NOALLOC_MAX_NG;
        c->no_alloc = true;
        lapinv(c, in, ng, out, ng, sc, 0, 1, EPI_STORE, nullptr, nullptr);
        c->no_alloc = false;
        c->T1 = T1;
        c->T2 = T2;
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(c, IBH_EDEVICE, hipGetErrorString(e));
    }
    if (cur != c->dev) hipSetDevice(cur);
    return rc;
}

int ibh_tune_report(ibh_ctx *c, char *buf, int32_t len) {
    if (!c || !buf || len <= 0) return IBH_EINVAL;
    std::string r;
    for (auto &l : c->tune_log) r += l + "\n";
    std::snprintf(buf, (size_t)len, "%s", r.c_str());
    return IBH_OK;
}

int ibh_profile_enable(ibh_ctx *c, int32_t on) {
    if (!c) return IBH_EINVAL;
    HIPCK(hipStreamSynchronize(c->s));
    c->prof.on = on != 0;
    c->prof.used = 0;
    c->prof.rec.clear();
    if (c->prof.on && ibh::prof_arm(c->prof, c->s))
        return fail(c, IBH_EDEVICE, "allocating the profiling timestamps failed");
    return IBH_OK;
}

int ibh_profile_read(ibh_ctx *c, int32_t kind, double *ms, int64_t *launches, double *flops) {
    if (!c) return IBH_EINVAL;
    HIPCK(hipStreamSynchronize(c->s));
    const int rc = ibh::prof_device_sum(c->prof, kind, c->s, ms, launches, flops);
    return rc ? fail(c, rc, "reading the profiling timestamps failed") : IBH_OK;
}

int ibh_profile_launches(ibh_ctx *c, int32_t kind, int32_t max, double *us, double *flops,
                         double *bytes) {
    if (!c || max < 0 || (max > 0 && (!us || !flops || !bytes))) return -IBH_EINVAL;
    return ibh::prof_launches(c->prof, kind, c->s, max, us, flops, bytes);
}

int ibh_profile_read_events(ibh_ctx *c, int32_t kind, double *ms, int64_t *launches,
                            double *flops) {
    if (!c) return IBH_EINVAL;
    HIPCK(hipStreamSynchronize(c->s));
    const int rc = ibh::prof_event_sum(c->prof, kind, ms, launches, flops);
    return rc ? fail(c, rc, "reading the profiling events failed") : IBH_OK;
}

}  // extern "C"
