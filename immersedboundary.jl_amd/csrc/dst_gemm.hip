// Sine-matrix contractions on MFMA (gfx950): FP64, FP32 and bf16x3.
//
// A 2-D DST-I of an m×p field X is S_m·X·S_p with S[i,k] = sin(π i k / n) (FFTW RODFT00 is
// 2·S per dimension, fluid-ops.jl:25,49,51).  Every contraction is computed as a "TN" GEMM
// C = A·B whose operands are both read k-fastest:
//   left  contraction  T = S·X : A = S (symmetric, so column i of S is row i), B = X,
//   right contraction  W = T·S : A = Tᵀ (written transposed by the previous contraction),
//                                B = S.
// Tiles are staged in LDS as [row][k].  The output is written column-major (x fastest) or
// transposed, by choosing which staged tile feeds the MFMA's A operand, so that lane%16
// always walks the contiguous index of C.  Block tile BM×BN, K tile BK, WM×WN waves; K tiles
// are staged in one or two LDS buffers with the next one or two tiles prefetched into
// registers.
//
// Precision policies:
//   FP64   v_mfma_f64_16x16x4f64,  D row = lane/16 + 4r
//   FP32   v_mfma_f32_16x16x4f32,  D row = 4(lane/16) + r, operands rounded from FP64
//   BF16X3 v_mfma_f32_16x16x32_bf16, a·b ≈ ah·bh + ah·bl + al·bh with a = ah + al split on
//          the LDS store, FP32 accumulation, D row = 4(lane/16) + r
// (the D layouts were measured on gfx950 with an in-bounds probe).
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "ibh_internal.h"

namespace ibh {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------------------ precision policies
struct PF64 {
    typedef double E;    // LDS element
    typedef double SL;   // sine part as stored in global memory
    __device__ static const SL *so(const GemmArgs &g) { return g.So; }
    __device__ static const SL *se(const GemmArgs &g) { return g.Se; }
    typedef d4 Acc;
    static constexpr int KS = 4;      // K per MFMA
    // LDS row padding (elements): an odd row stride in doubles; measured with rocprofv3 on the
    // 1023² contraction, padding 1 has 13× fewer bank-conflict cycles than 4 (8 is worse still)
    static constexpr int PAD = 1;
    __device__ static Acc zero() { return d4{0.0, 0.0, 0.0, 0.0}; }
    __device__ static int drow(int lq, int r) { return lq + 4 * r; }
};
struct PF32 {
    typedef float E;
    typedef float SL;
    __device__ static const SL *so(const GemmArgs &g) { return g.Sof; }
    __device__ static const SL *se(const GemmArgs &g) { return g.Sef; }
    typedef f4 Acc;
    static constexpr int KS = 4;
    static constexpr int PAD = 4;
    __device__ static Acc zero() { return f4{0.f, 0.f, 0.f, 0.f}; }
    __device__ static int drow(int lq, int r) { return 4 * lq + r; }
};
struct PBF3 {
    typedef __bf16 E;    // hi and lo planes
    typedef unsigned SL; // hi | lo << 16
    __device__ static const SL *so(const GemmArgs &g) { return g.Sob; }
    __device__ static const SL *se(const GemmArgs &g) { return g.Seb; }
    typedef f4 Acc;
    static constexpr int KS = 32;
    static constexpr int PAD = 8;
    __device__ static Acc zero() { return f4{0.f, 0.f, 0.f, 0.f}; }
    __device__ static int drow(int lq, int r) { return 4 * lq + r; }
};

template <class P> struct Frag;
template <> struct Frag<PF64> {
    double v;
    __device__ void load(const double *row, int kk, int lq) { v = row[kk * 4 + lq]; }
};
template <> struct Frag<PF32> {
    float v;
    __device__ void load(const float *row, int kk, int lq) { v = row[kk * 4 + lq]; }
};
template <> struct Frag<PBF3> {
    bf8 hi, lo;
};

__device__ inline d4 mma(const Frag<PF64> &a, const Frag<PF64> &b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a.v, b.v, c, 0, 0, 0);
}
__device__ inline f4 mma(const Frag<PF32> &a, const Frag<PF32> &b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a.v, b.v, c, 0, 0, 0);
}
__device__ inline f4 mma(const Frag<PBF3> &a, const Frag<PBF3> &b, f4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
}

// ------------------------------------------------------------------------ LDS tiles
template <class P, int ROWS, int BK> struct Tile {
    typename P::E s[ROWS][BK + P::PAD];
};
template <int ROWS, int BK> struct Tile<PBF3, ROWS, BK> {
    __bf16 hi[ROWS][BK + PBF3::PAD];
    __bf16 lo[ROWS][BK + PBF3::PAD];
};

template <class P, int ROWS, int BK>
__device__ inline void tput(Tile<P, ROWS, BK> &t, int r, int k, double v) {
    t.s[r][k] = (typename P::E)v;
}
// bf16x3 operand split: f = fl32(v), hi = bf16(f), lo = bf16(f − hi), both round-to-nearest-even
__device__ inline void bf3_split(double v, __bf16 &hi, __bf16 &lo) {
    const float f = (float)v;
    hi = __float2bfloat16(f);
    lo = __float2bfloat16(f - __bfloat162float(hi));
}
template <int ROWS, int BK>
__device__ inline void tput(Tile<PBF3, ROWS, BK> &t, int r, int k, double v) {
    bf3_split(v, t.hi[r][k], t.lo[r][k]);
}
// sine elements arrive already rounded (launch_sine_low)
template <class P, int ROWS, int BK>
__device__ inline void tputs(Tile<P, ROWS, BK> &t, int r, int k, typename P::SL v) {
    t.s[r][k] = v;
}
template <int ROWS, int BK>
__device__ inline void tputs(Tile<PBF3, ROWS, BK> &t, int r, int k, unsigned v) {
    t.hi[r][k] = __builtin_bit_cast(__bf16, (unsigned short)(v & 0xffffu));
    t.lo[r][k] = __builtin_bit_cast(__bf16, (unsigned short)(v >> 16));
}

__global__ void k_sine_low(const double *S, float *f, unsigned *b, long n) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    __bf16 hi, lo;
    bf3_split(S[t], hi, lo);
    f[t] = (float)S[t];
    b[t] = (unsigned)__builtin_bit_cast(unsigned short, hi) |
           ((unsigned)__builtin_bit_cast(unsigned short, lo) << 16);
}

// FP32: [Soe | the same with the two pairs of every 16-byte slot swapped] (n pairs each, n
// even, rows of an even number of pairs): the LDS-direct tiles take rows r with r & 8 set from
// the swapped copy, so that the 16 rows of a fragment read meet in different LDS banks
__global__ void k_sine_interleave_f(const float *So, const float *Se, float *Soe, long n) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    Soe[2 * t] = So[t];
    Soe[2 * t + 1] = Se[t];
    Soe[2 * n + 2 * t] = So[t ^ 1];
    Soe[2 * n + 2 * t + 1] = Se[t ^ 1];
}
__global__ void k_sine_interleave(const double *So, const double *Se, double *Soe, long n) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    Soe[2 * t] = So[t];
    Soe[2 * t + 1] = Se[t];
}

template <class P, int ROWS, int BK>
__device__ inline void fload(Frag<P> &f, const Tile<P, ROWS, BK> &t, int row, int kk, int lq) {
    f.load(t.s[row], kk, lq);
}
template <int ROWS, int BK>
__device__ inline void fload(Frag<PBF3> &f, const Tile<PBF3, ROWS, BK> &t, int row, int kk,
                             int lq) {
    f.hi = *reinterpret_cast<const bf8 *>(&t.hi[row][kk * 32 + lq * 8]);
    f.lo = *reinterpret_cast<const bf8 *>(&t.lo[row][kk * 32 + lq * 8]);
}

// ------------------------------------------------------------------------ XCD-aware order
// Workgroups are dispatched round-robin over the 8 XCDs (linear id mod 8).  Remap so that each
// XCD walks a contiguous range of tiles (x fastest): tiles sharing a data or sine panel then
// run on the same XCD and hit its L2.  Bijective for any grid size.
__device__ inline void xcd_tile(int &x, int &y, int &z) {
    const int gx = gridDim.x, gy = gridDim.y;
    const int total = gx * gy * gridDim.z;
    const int L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int q = total / 8, r = total % 8;
    const int xcd = L % 8, slot = L / 8;
    const int T = xcd * q + min(xcd, r) + slot;
    x = T % gx;
    y = (T / gx) % gy;
    z = T / (gx * gy);
}

// ------------------------------------------------------------------------ profiling
// Kernel duration as the profiler defines it (first workgroup start to last workgroup end), on
// the constant-rate device wall clock, only when a profiling record is passed: workgroup 0
// (linear id `lin`) stores the start; every workgroup folds its end into a sub-slot with a
// non-returning vector atomic.
__device__ inline void ts_mark(unsigned long long *ts, bool end, unsigned lin) {
    if (ts && threadIdx.x == 0) {
        const unsigned long long t = (unsigned long long)wall_clock64();
        if (end) atomicMax(ts + 16 + 16 * (lin & 15), t);
        else if (lin == 0) ts[0] = t;
    }
}
__device__ inline unsigned lin_block() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// ------------------------------------------------------------------------ kernel
template <class P, int BM, int BN, int BK, int WM, int WN, int EPI, bool TOUT>
__global__ __launch_bounds__(WM *WN * 64) void k_gemm(GemmArgs g) {
    constexpr int NT = WM * WN * 64;
    constexpr int TM = BM / WM, TN = BN / WN;  // wave tile
    constexpr int FM = TM / 16, FN = TN / 16;  // MFMA blocks per wave
    constexpr int KS = P::KS;
    static_assert(BK % KS == 0 && TM % 16 == 0 && TN % 16 == 0, "tile shape");
    constexpr int LA = BM * BK / NT, LB = BN * BK / NT;  // elements per thread per tile
    static_assert(LA * NT == BM * BK && LB * NT == BN * BK, "load split");

    if (g.skip && *g.skip) return;
    ts_mark(g.ts, false, lin_block());
    int bxt, byt, bz;
    xcd_tile(bxt, byt, bz);
    const double *__restrict__ A = g.A + bz * g.sA;  // A(m, k) at A[k + lda·m]
    const double *__restrict__ B = g.B + bz * g.sB;  // B(k, n) at B[k + ldb·n]
    double *__restrict__ C = g.C + bz * g.sC;
    const int M = g.M, N = g.N, K = g.K;
    const int m0 = bxt * BM, n0 = byt * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int l16 = lane & 15, lq = lane >> 4;

    __shared__ Tile<P, BM, BK> As[2];
    __shared__ Tile<P, BN, BK> Bs[2];

    typename P::Acc acc[FN][FM];
#pragma unroll
    for (int a = 0; a < FN; ++a)
#pragma unroll
        for (int b = 0; b < FM; ++b) acc[a][b] = P::zero();

    // global → register prefetch: thread t covers element (row = e / BK, k = e % BK) for
    // e = t + NT·u, so a wave reads whole k-runs of consecutive rows
    double ra[LA], rb[LB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int e = tid + NT * u, r = e / BK, k = k0 + e % BK;
            ra[u] = (m0 + r < M && k < K) ? A[k + g.lda * (long)(m0 + r)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int e = tid + NT * u, r = e / BK, k = k0 + e % BK;
            rb[u] = (n0 + r < N && k < K) ? B[k + g.ldb * (long)(n0 + r)] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < LA; ++u) {
            const int e = tid + NT * u;
            tput(As[buf], e / BK, e % BK, ra[u]);
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int e = tid + NT * u;
            tput(Bs[buf], e / BK, e % BK, rb[u]);
        }
    };

    const int nk = (K + BK - 1) / BK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK / KS; ++kk) {
            Frag<P> fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) fload(fa[i], As[cur], wm * TM + i * 16 + l16, kk, lq);
#pragma unroll
            for (int j = 0; j < FN; ++j) fload(fb[j], Bs[cur], wn * TN + j * 16 + l16, kk, lq);
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    // column-major C: MFMA rows ↔ n (B tile), lanes ↔ m (A tile);
                    // transposed C: MFMA rows ↔ m, lanes ↔ n
                    if (TOUT) acc[j][i] = mma(fa[i], fb[j], acc[j][i]);
                    else acc[j][i] = mma(fb[j], fa[i], acc[j][i]);
                }
        }
        if (kt + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }

    const double *D = g.D ? g.D + bz * g.sD : nullptr;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int m, n;
                if (TOUT) {
                    m = m0 + wm * TM + i * 16 + P::drow(lq, r);
                    n = n0 + wn * TN + j * 16 + l16;
                } else {
                    m = m0 + wm * TM + i * 16 + l16;
                    n = n0 + wn * TN + j * 16 + P::drow(lq, r);
                }
                if (m >= M || n >= N) continue;
                double v = (double)acc[j][i][r];
                if (TOUT) {  // C(m, n) at C[n + ldc·m]
                    C[n + g.ldc * (long)m] = v;
                    continue;
                }
                const long ci = m + g.ldc * (long)n;
                if (EPI == EPI_STORE) {
                    C[ci] = v;
                } else if (EPI == EPI_SCALE) {
                    C[ci] = v * D[m + g.ldd * (long)n];
                } else if (EPI == EPI_SUBFROM) {
                    C[ci] = D[m + g.ldd * (long)n] - v;
                } else {  // EPI_RHS: rhs = AΓ; rhs += dtβ1 N1; rhs += dtβ2 N2; rhs += rhsbc
                    const double *N1 = g.N1 + bz * g.sC, *N2 = g.N2 + bz * g.sC;
                    v = v + g.f1 * N1[ci];
                    v = v + g.f2 * N2[ci];
                    const double *bc = g.bcacc + (long)bz * g.nbc;
                    // apply_bc! edge terms (fluid-ops.jl:256-272), fac already folded in
                    double e = 0.0;
                    if (n == 0) e += bc[g.Bo + 1 + m];
                    if (n == N - 1) e += bc[g.To + 1 + m];
                    if (m == 0) e += bc[g.Lo + 1 + n];
                    if (m == M - 1) e += bc[g.Ro + 1 + n];
                    C[ci] = v + e;
                }
            }
    ts_mark(g.ts, true, lin_block());
}

// ------------------------------------------------------------------------ even/odd split
// DST-I symmetry: with n even, h = n/2 and a'' = 0..h-1, rows a = a''+1 and n−a of S·x are
//   O ± E,  O = Σ_k'' So[a''][k'']·x[2k''],  E = Σ_k'' Se[a''][k'']·x[2k''+1]
// (0-based x), So[a''][k''] = sin(π(a''+1)(2k''+1)/n), Se[a''][k''] = sin(2π(a''+1)(k''+1)/n)
// with Se's last row and column zero.  Each part is an h×h dense contraction, so a transform
// costs half the MFMA work of the full m×m one.  SIDE 0: C = S·X (pairs on the M index, data
// X as the B operand); SIDE 1: C = X·S (pairs on the N index, data as the A operand).  The
// data operand is read as 2·BK contiguous k values and split by parity on the LDS store; the
// butterfly runs in the epilogue.
// WS = 2 splits the parts over the waves: WM×WN waves compute O and as many compute E from the
// same LDS tiles (twice the resident waves for latency hiding when the grid is small); the E
// waves hand their accumulators to the O waves through LDS for the butterfly.  Each part's
// sums run in the same order either way.
template <class P, int BM, int BN, int BK, int WM, int WN, int EPI, bool TOUT, int SIDE, int NB,
          int WS = 1>
__device__ __forceinline__ void eo_tile(const GemmArgs &g, int bxt, int byt, int bz) {
    constexpr int NW = WM * WN;              // waves per part
    constexpr int NT = NW * 64 * WS;
    constexpr int TM = BM / WM, TN = BN / WN;
    constexpr int FM = TM / 16, FN = TN / 16;
    constexpr int KS = P::KS;
    static_assert(BK % KS == 0 && TM % 16 == 0 && TN % 16 == 0, "tile shape");
    constexpr int RS = SIDE == 0 ? BM : BN;  // rows of the sine tiles (pair index)
    constexpr int RD = SIDE == 0 ? BN : BM;  // rows of the data tile
    constexpr int LS = RS * BK / NT, LD = RD * 2 * BK / NT;
    static_assert(LS * NT == RS * BK && LD * NT == RD * 2 * BK, "load split");
    // callers that form only the rows a narrowed second contraction reads align them to 128
    static_assert(BK <= 64 && (BK & (BK - 1)) == 0, "K tile of at most 64 pairs, a power of 2");

    const int h = g.h;                       // pairs
    const int Kd = g.K;                      // data length along k (= n - 1)
    const double *__restrict__ X = SIDE == 0 ? g.B + bz * g.sB : g.A + bz * g.sA;
    const long ldx = SIDE == 0 ? g.ldb : g.lda;
    double *__restrict__ C = g.C + bz * g.sC;
    const int M = g.M, N = g.N;
    const int m0 = bxt * BM, n0 = byt * BN;
    const int r0s = SIDE == 0 ? m0 : n0, r0d = SIDE == 0 ? n0 : m0;
    const int nrd = SIDE == 0 ? N : M;       // valid data rows
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int part = WS == 2 ? wave / NW : 0;  // 0: O (or both parts), 1: E
    const int wl = wave % NW, wm = wl % WM, wn = wl / WM;
    const int l16 = lane & 15, lq = lane >> 4;

    // NB = 1: one LDS buffer; 2: two buffers, the next K tile prefetched into registers;
    // NB = R + 1 ≥ 3: two buffers and R register stages, so global loads are issued R K tiles
    // ahead
    constexpr int NBUF = NB == 1 ? 1 : 2;
    constexpr int NREG = NB >= 3 ? NB - 1 : 1;
    __shared__ Tile<P, BM, BK> AsO[NBUF], AsE[NBUF];
    __shared__ Tile<P, BN, BK> BsO[NBUF], BsE[NBUF];

    typename P::Acc accO[FN][FM], accE[FN][FM];
#pragma unroll
    for (int a = 0; a < FN; ++a)
#pragma unroll
        for (int b = 0; b < FM; ++b) accO[a][b] = accE[a][b] = P::zero();

    typename P::SL rso[NREG][LS], rse[NREG][LS];
    double rd[NREG][LD];
    // predicated loads: out-of-range elements are zero.  (Loading them unconditionally from a
    // clamped index was measured slower on every variant, 4-39%.)
    auto gload = [&](int k0, int st) {
#pragma unroll
        for (int u = 0; u < LS; ++u) {
            const int e = tid + NT * u, r = r0s + e / BK, k = k0 + e % BK;
            const bool in = r < h && k < h;
            rso[st][u] = in ? P::so(g)[k + (long)h * r] : typename P::SL(0);
            rse[st][u] = in ? P::se(g)[k + (long)h * r] : typename P::SL(0);
        }
#pragma unroll
        for (int u = 0; u < LD; ++u) {
            const int e = tid + NT * u, r = r0d + e / (2 * BK), q = 2 * k0 + e % (2 * BK);
            rd[st][u] = (r < nrd && q < Kd) ? X[q + ldx * (long)r] : 0.0;
        }
    };
    auto sstore = [&](int buf, int st) {
#pragma unroll
        for (int u = 0; u < LS; ++u) {
            const int e = tid + NT * u;
            if (SIDE == 0) {
                tputs(AsO[buf], e / BK, e % BK, rso[st][u]);
                tputs(AsE[buf], e / BK, e % BK, rse[st][u]);
            } else {
                tputs(BsO[buf], e / BK, e % BK, rso[st][u]);
                tputs(BsE[buf], e / BK, e % BK, rse[st][u]);
            }
        }
#pragma unroll
        for (int u = 0; u < LD; ++u) {
            const int e = tid + NT * u, r = e / (2 * BK), q = e % (2 * BK);
            if (SIDE == 0) {
                if (q & 1) tput(BsE[buf], r, q >> 1, rd[st][u]);
                else tput(BsO[buf], r, q >> 1, rd[st][u]);
            } else {
                if (q & 1) tput(AsE[buf], r, q >> 1, rd[st][u]);
                else tput(AsO[buf], r, q >> 1, rd[st][u]);
            }
        }
    };
    auto compute = [&](int cur) {
        if (WS == 2) {  // one part per wave: its accumulators live in accO
            const Tile<P, BM, BK> &Ta = part ? AsE[cur] : AsO[cur];
            const Tile<P, BN, BK> &Tb = part ? BsE[cur] : BsO[cur];
#pragma unroll
            for (int kk = 0; kk < BK / KS; ++kk) {
                Frag<P> fa[FM], fb[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i) fload(fa[i], Ta, wm * TM + i * 16 + l16, kk, lq);
#pragma unroll
                for (int j = 0; j < FN; ++j) fload(fb[j], Tb, wn * TN + j * 16 + l16, kk, lq);
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int i = 0; i < FM; ++i) {
                        if (TOUT) accO[j][i] = mma(fa[i], fb[j], accO[j][i]);
                        else accO[j][i] = mma(fb[j], fa[i], accO[j][i]);
                    }
            }
            return;
        }
#pragma unroll
        for (int kk = 0; kk < BK / KS; ++kk) {
            Frag<P> fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) fload(fa[i], AsO[cur], wm * TM + i * 16 + l16, kk, lq);
#pragma unroll
            for (int j = 0; j < FN; ++j) fload(fb[j], BsO[cur], wn * TN + j * 16 + l16, kk, lq);
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    if (TOUT) accO[j][i] = mma(fa[i], fb[j], accO[j][i]);
                    else accO[j][i] = mma(fb[j], fa[i], accO[j][i]);
                }
#pragma unroll
            for (int i = 0; i < FM; ++i) fload(fa[i], AsE[cur], wm * TM + i * 16 + l16, kk, lq);
#pragma unroll
            for (int j = 0; j < FN; ++j) fload(fb[j], BsE[cur], wn * TN + j * 16 + l16, kk, lq);
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int i = 0; i < FM; ++i) {
                    if (TOUT) accE[j][i] = mma(fa[i], fb[j], accE[j][i]);
                    else accE[j][i] = mma(fb[j], fa[i], accE[j][i]);
                }
        }
    };

    // K tiles [kt0, kt0 + nk) hold the pairs to sum (all of them unless klo/khi narrow it)
    const int nkall = (h + BK - 1) / BK;
    const int kt0 = g.khi > 0 ? g.klo / BK : 0;
    const int nk = g.khi > 0 ? min(nkall, (g.khi + BK - 1) / BK) - kt0 : nkall;
    if (nk <= 0) {
        // nothing to sum: the accumulators stay zero
    } else if (NB >= 3) {
        // tile t computes from LDS buffer t&1 while tiles t+1 .. t+R-1 wait in register
        // stages and the loads of tile t+R land in the stage tile t left; the loop is
        // unrolled by R so that the register stage of every tile is a compile-time index
        constexpr int R = NREG;
        gload(kt0 * BK, 0);
        sstore(0, 0);
#pragma unroll
        for (int u = 1; u < R; ++u)
            if (u < nk) gload((kt0 + u) * BK, u);
        __syncthreads();
        for (int kt = 0; kt < nk; kt += R) {
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int t = kt + u;
                if (t >= nk) break;
                if (t + R < nk) gload((kt0 + t + R) * BK, u);
                compute(t & 1);
                if (t + 1 < nk) sstore((t + 1) & 1, (u + 1) % R);
                __syncthreads();
            }
        }
    } else {
        gload(kt0 * BK, 0);
        sstore(0, 0);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const int cur = NB == 2 ? (kt & 1) : 0;
            if (kt + 1 < nk) gload((kt0 + kt + 1) * BK, 0);
            compute(cur);
            if (NB == 2) {
                if (kt + 1 < nk) sstore(cur ^ 1, 0);
                __syncthreads();
            } else {  // one LDS buffer: wait for every wave to finish reading before refilling
                __syncthreads();
                if (kt + 1 < nk) {
                    sstore(0, 0);
                    __syncthreads();
                }
            }
        }
    }

    // WS = 2: the E waves pass their accumulators to the O waves (same wave-local index and
    // lane), then leave.  Every wave of the group reaches this barrier exactly once.
    constexpr int NX = WS == 2 ? NW * FN * FM * 4 * 64 : 1;
    __shared__ double xch[NX];
    if (WS == 2) {
        if (part == 1) {
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        xch[(((wl * FN + j) * FM + i) * 4 + r) * 64 + lane] = (double)accO[j][i][r];
        }
        __syncthreads();
        if (part == 1) return;
    }

    const double *D = g.D ? g.D + bz * g.sD : nullptr;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int mi, ni;  // tile-local row (M side) and column (N side) of this value
                if (TOUT) {
                    mi = wm * TM + i * 16 + P::drow(lq, r);
                    ni = wn * TN + j * 16 + l16;
                } else {
                    mi = wm * TM + i * 16 + l16;
                    ni = wn * TN + j * 16 + P::drow(lq, r);
                }
                const int pi = SIDE == 0 ? m0 + mi : n0 + ni;   // pair index a''
                const int di = SIDE == 0 ? n0 + ni : m0 + mi;   // data row index
                if (pi >= h || di >= nrd) continue;
                const double o = (double)accO[j][i][r];
                const double e = WS == 2 ? xch[(((wl * FN + j) * FM + i) * 4 + r) * 64 + lane]
                                         : (double)accE[j][i][r];
                for (int side2 = 0; side2 < 2; ++side2) {
                    const int pr = side2 == 0 ? pi : Kd - 1 - pi;
                    if (side2 == 1 && pr == pi) break;
                    double v = side2 == 0 ? o + e : o - e;
                    const int m = SIDE == 0 ? pr : di, n = SIDE == 0 ? di : pr;
                    if (TOUT) {
                        C[n + g.ldc * (long)m] = v;
                        continue;
                    }
                    const long ci = m + g.ldc * (long)n;
                    if (EPI == EPI_STORE) {
                        C[ci] = v;
                    } else if (EPI == EPI_SCALE) {
                        C[ci] = v * D[m + g.ldd * (long)n];
                    } else if (EPI == EPI_SUBFROM) {
                        C[ci] = D[m + g.ldd * (long)n] - v;
                    } else {
                        const double *N1 = g.N1 + bz * g.sC, *N2 = g.N2 + bz * g.sC;
                        v = v + g.f1 * N1[ci];
                        v = v + g.f2 * N2[ci];
                        const double *bc = g.bcacc + (long)bz * g.nbc;
                        double eb = 0.0;
                        if (n == 0) eb += bc[g.Bo + 1 + m];
                        if (n == N - 1) eb += bc[g.To + 1 + m];
                        if (m == 0) eb += bc[g.Lo + 1 + n];
                        if (m == M - 1) eb += bc[g.Ro + 1 + n];
                        C[ci] = v + eb;
                    }
                }
            }
}

template <class P, int BM, int BN, int BK, int WM, int WN, int EPI, bool TOUT, int SIDE, int NB = 2,
          int WS = 1>
__global__ __launch_bounds__(WM *WN * 64 * WS) void k_gemm_eo(GemmArgs g) {
    if (g.skip && *g.skip) return;
    ts_mark(g.ts, false, lin_block());
    int bxt, byt, bz;
    xcd_tile(bxt, byt, bz);
    bxt += g.olo[0] / BM;
    byt += g.olo[1] / BN;
    eo_tile<P, BM, BN, BK, WM, WN, EPI, TOUT, SIDE, NB, WS>(g, bxt, byt, bz);
    ts_mark(g.ts, true, lin_block());
}

// Grouped form: one launch over the contractions of many cases (each its own sizes, sine
// matrices and buffers).  The linear workgroup id is remapped XCD-contiguously over the whole
// grid, then located in the tile-offset table; the tile walks x fastest inside its problem.
template <class P, int BM, int BN, int BK, int WM, int WN, int EPI, bool TOUT, int SIDE, int NB,
          int WS = 1>
__global__ __launch_bounds__(WM *WN * 64 * WS) void kg_gemm_eo(const GemmArgs *gs, const int *off,
                                                               int n, unsigned long long *ts) {
    ts_mark(ts, false, blockIdx.x);
    const int total = gridDim.x, q = total / 8, r = total % 8;
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const int T = xcd * q + min(xcd, r) + slot;
    const int p = grp_find(off, n, T);
    const GemmArgs g = gs[p];
    const int local = T - off[p];
    const int rows = SIDE == 0 ? g.h : g.M, cols = SIDE == 0 ? g.N : g.h;
    // output tiles [olo, ohi) per dimension where ohi > 0 (as k_gemm_eo's grid)
    const int x0 = g.ohi[0] > 0 ? g.olo[0] / BM : 0, y0 = g.ohi[1] > 0 ? g.olo[1] / BN : 0;
    const int gx = ((g.ohi[0] > 0 ? min(g.ohi[0], rows) : rows) + BM - 1) / BM - x0;
    const int gy = ((g.ohi[1] > 0 ? min(g.ohi[1], cols) : cols) + BN - 1) / BN - y0;
    eo_tile<P, BM, BN, BK, WM, WN, EPI, TOUT, SIDE, NB, WS>(g, x0 + local % gx,
                                                            y0 + (local / gx) % gy,
                                                            local / (gx * gy));
    ts_mark(ts, true, blockIdx.x);
}

// ------------------------------------------------------------------------ LDS-direct split tiles
// The split contraction of eo_tile (FP64) with its operand tiles staged by buffer loads that
// write LDS directly (buffer_load ... lds): nothing passes through registers, two LDS stages
// alternate, and the loads of K tile t+1 are in flight while tile t is multiplied.
// * The sine parts come interleaved (GemmArgs::Soe, 16 bytes per pair) with 16-byte loads into
//   XOR-swizzled slots: pair k of tile row r sits in slot k ^ (r & 15) of its row, so the 16 rows
//   one fragment read touches fall in distinct LDS banks.
// * Data rows have odd lengths (not 16-byte aligned), so they are staged by 4-byte loads, one
//   256-byte run per wave instruction, into rows padded by 16 bytes; one 16-byte read gives
//   (x[2k], x[2k+1]) for the O and the E part.
// * Out-of-range elements (rows past the operand, pairs past h, and x[K], which the last pair
//   would take from the next row) are loaded from an offset past the buffer's size, which the
//   buffer load returns as zero: the zeros eo_tile stores.
// * Fragments are read from LDS two K steps ahead of the products that use them.
// Each part's products run over the same k groups in the same order as eo_tile, so the results
// are bitwise those of every other split variant.  WS = 2 runs the O and E parts on separate
// waves (twice the resident waves); the E waves hand their sums over through LDS.

typedef __attribute__((address_space(3))) void lds_void;
typedef double d2 __attribute__((ext_vector_type(2)));

// buffer resource of `bytes` bytes at p (raw, stride 0; the third dword selects 32-bit data,
// which raw loads need to return data at all): offsets at or past `bytes` read as zero.
// Arguments are made wave-uniform, as the resource lives in scalar registers.
__device__ inline __amdgpu_buffer_rsrc_t wave_rsrc(const void *p, long bytes) {
    const unsigned long long a = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    const int nb = __builtin_amdgcn_readfirstlane((int)bytes);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)hi << 32) | lo), 0, nb,
                                             0x00020000);
}
constexpr unsigned DMA_OOB = 0x80000000u;  // an offset past every operand's size

// the split contraction's store of o ± e for pair index pi and data row di (eo_tile's epilogue)
template <int EPI, bool TOUT, int SIDE>
__device__ __forceinline__ void eo_store(const GemmArgs &g, double *__restrict__ C, int bz, int pi,
                                         int di, double o, double e) {
    const int Kd = g.K;
    for (int side2 = 0; side2 < 2; ++side2) {
        const int pr = side2 == 0 ? pi : Kd - 1 - pi;
        if (side2 == 1 && pr == pi) break;
        double v = side2 == 0 ? o + e : o - e;
        const int m = SIDE == 0 ? pr : di, n = SIDE == 0 ? di : pr;
        if (TOUT) {
            C[n + g.ldc * (long)m] = v;
            continue;
        }
        const long ci = m + g.ldc * (long)n;
        if (EPI == EPI_STORE) {
            C[ci] = v;
        } else if (EPI == EPI_SCALE) {
            C[ci] = v * g.D[bz * g.sD + m + g.ldd * (long)n];
        } else if (EPI == EPI_SUBFROM) {
            C[ci] = g.D[bz * g.sD + m + g.ldd * (long)n] - v;
        } else {
            const double *N1 = g.N1 + bz * g.sC, *N2 = g.N2 + bz * g.sC;
            v = v + g.f1 * N1[ci];
            v = v + g.f2 * N2[ci];
            const double *bc = g.bcacc + (long)bz * g.nbc;
            double eb = 0.0;
            if (n == 0) eb += bc[g.Bo + 1 + m];
            if (n == g.N - 1) eb += bc[g.To + 1 + m];
            if (m == 0) eb += bc[g.Lo + 1 + n];
            if (m == g.M - 1) eb += bc[g.Ro + 1 + n];
            C[ci] = v + eb;
        }
    }
}

__device__ inline void dma_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)
// workgroup barrier ordering LDS only (the staged loads are waited for by dma_wait_all)
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// the interleaved sine parts of a precision: SB bytes per pair, (So, Se) of one element type
template <class P> struct DmaSine;
// SWAP: the array holds a second copy with the two pairs of each 16-byte slot swapped, which
// the tile rows with bit 3 set are loaded from (their pair k then sits in the other half of the
// slot, so rows r and r + 8 of a fragment read hit different banks)
template <> struct DmaSine<PF64> {
    static constexpr int SB = 16;
    static constexpr bool SWAP = false;
    typedef d2 T;
    __device__ static const void *ptr(const GemmArgs &g) { return g.Soe; }
};
template <> struct DmaSine<PF32> {
    static constexpr int SB = 8;
    static constexpr bool SWAP = true;
    typedef float2 T;
    __device__ static const void *ptr(const GemmArgs &g) { return g.Soef; }
};
__device__ inline d4 mfma16(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma16(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class P, int BM, int BN, int BK, int WM, int WN, int SIDE, int WS>
struct DmaCfg {
    static constexpr int NW = WM * WN, NWT = NW * WS;            // waves per part, in all
    static constexpr int RS = SIDE == 0 ? BM : BN;              // sine rows (pairs)
    static constexpr int RD = SIDE == 0 ? BN : BM;              // data rows
    static constexpr int TP = SIDE == 0 ? BM / WM : BN / WN;    // a wave's pairs
    static constexpr int TD = SIDE == 0 ? BN / WN : BM / WM;    // a wave's data rows
    static constexpr int FP = TP / 16, FD = TD / 16;
    static constexpr int SB = DmaSine<P>::SB;                   // sine bytes per pair
    static constexpr int PPS = 16 / SB;                         // pairs per 16-byte slot
    static constexpr int SPR = BK / PPS;                        // slots per sine row
    static constexpr int SWM = (SPR < 16 ? SPR : 16) - 1;       // slot swizzle mask
    static constexpr int PITCH = 2 * BK * 8 + 16;               // data row pitch (bytes)
    static constexpr int SBYTES = RS * SPR * 16, DBYTES = RD * PITCH, STAGE = SBYTES + DBYTES;
    static constexpr int SI = RS * SPR / 64;                    // 16-byte wave loads (sine)
    static constexpr int SEGS = 2 * BK / 32;                    // 256-byte runs per data row
    static constexpr int DI = RD * SEGS;                        // 4-byte wave loads (data)
    static constexpr int NSI = (SI + NWT - 1) / NWT, NDI = (DI + NWT - 1) / NWT;
    static_assert(BK >= 16 && BK <= 64 && (BK & (BK - 1)) == 0, "K tile of 16..64 pairs");
    static_assert(SPR >= 8, "eight slots per sine row for the swizzle");
    static_assert(TP % 16 == 0 && TD % 16 == 0, "wave tile");
    static_assert(SI * 64 == RS * SPR && NWT % SEGS == 0, "load split");
    static_assert(WS == 1 || NW * FP * FD * 4 * 64 * 8 <= STAGE, "hand-over fits one stage");
};

template <class P, int BM, int BN, int BK, int WM, int WN, int EPI, bool TOUT, int SIDE, int WS>
__device__ __forceinline__ void eo_tile_dma(const GemmArgs &g, int bxt, int byt, int bz) {
    using Q = DmaCfg<P, BM, BN, BK, WM, WN, SIDE, WS>;
    using ST = typename DmaSine<P>::T;
    typedef typename P::E E;
    constexpr int FP = Q::FP, FD = Q::FD, SB = Q::SB, PPS = Q::PPS;
    const int h = g.h, Kd = g.K;
    const double *X = SIDE == 0 ? g.B + bz * g.sB : g.A + bz * g.sA;
    const long ldx = SIDE == 0 ? g.ldb : g.lda;
    double *__restrict__ C = g.C + bz * g.sC;
    const int m0 = bxt * BM, n0 = byt * BN;
    const int r0s = SIDE == 0 ? m0 : n0, r0d = SIDE == 0 ? n0 : m0;
    const int nrd = SIDE == 0 ? g.N : g.M;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int part = WS == 2 ? wave / Q::NW : 0, wl = wave % Q::NW;
    const int wm = wl % WM, wn = wl / WM;
    const int wp = SIDE == 0 ? wm : wn, wd = SIDE == 0 ? wn : wm;
    const int l16 = lane & 15, lq = lane >> 4;

    __shared__ __attribute__((aligned(16))) char st0[Q::STAGE];
    __shared__ __attribute__((aligned(16))) char st1[Q::STAGE];

    constexpr bool SWAP = DmaSine<P>::SWAP;
    const __amdgpu_buffer_rsrc_t rs = wave_rsrc(DmaSine<P>::ptr(g), (long)h * h * SB * (SWAP ? 2 : 1));
    const __amdgpu_buffer_rsrc_t rd = wave_rsrc(X, (ldx * (long)(nrd - 1) + Kd) * 8);
    // byte offsets of this lane's elements in K tile 0 (DMA_OOB for rows past the operand):
    // sine slot L of the tile holds pairs from k = PPS·((L % SPR) ^ (row & SWM)) of its row
    int sbase[Q::NSI], sk[Q::NSI];
#pragma unroll
    for (int u = 0; u < Q::NSI; ++u) {
        const int L = (u * Q::NWT + wave) * 64 + lane;
        const int row = L / Q::SPR, k = PPS * ((L % Q::SPR) ^ (row & Q::SWM));
        sk[u] = k;
        sbase[u] = r0s + row < h
                       ? (k + h * (r0s + row)) * SB + (SWAP && (row & 8) ? h * h * SB : 0)
                       : (int)DMA_OOB;
    }
    int dbase[Q::NDI];
    const int de0 = (wave % Q::SEGS) * 32 + (lane >> 1);  // element of the tile row
#pragma unroll
    for (int v = 0; v < Q::NDI; ++v) {
        const int row = (v * Q::NWT + wave) / Q::SEGS;
        dbase[v] = r0d + row < nrd ? (int)((ldx * (r0d + row) + de0) * 8 + (lane & 1) * 4)
                                   : (int)DMA_OOB;
    }
    auto issue = [&](char *stage, int kt) {
        const int kp = kt * BK;  // first pair of the tile
        const bool full = kp + BK <= h && 2 * (kp + BK) <= Kd;
        const bool eok = 2 * kp + de0 < Kd;
#pragma unroll
        for (int u = 0; u < Q::NSI; ++u) {
            const int inst = u * Q::NWT + wave;
            if (Q::SI % Q::NWT != 0 && inst >= Q::SI) break;
            // h is even, so a slot's pairs are in range together
            const unsigned off = full || (sbase[u] != (int)DMA_OOB && kp + sk[u] < h)
                                     ? (unsigned)sbase[u] + kp * SB : DMA_OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void *)(stage + inst * 1024), 16, off,
                                                     0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < Q::NDI; ++v) {
            const int inst = v * Q::NWT + wave;
            if (Q::DI % Q::NWT != 0 && inst >= Q::DI) break;
            const int row = inst / Q::SEGS, seg = inst % Q::SEGS;
            const unsigned off = full || (dbase[v] != (int)DMA_OOB && eok)
                                     ? (unsigned)dbase[v] + kp * 16 : DMA_OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                rd, (lds_void *)(stage + Q::SBYTES + row * Q::PITCH + seg * 256), 4, off, 0, 0, 0);
        }
    };

    // WS = 1: the O and E sums; WS = 2: this wave's part in accO (accE unused)
    typename P::Acc accO[FP][FD], accE[FP][FD];
#pragma unroll
    for (int i = 0; i < FP; ++i)
#pragma unroll
        for (int j = 0; j < FD; ++j) accO[i][j] = accE[i][j] = P::zero();

    // fragment of K step kk: the sine pair (So, Se) of row i, the data pair (x even, x odd) in
    // the operand type (FP32 rounds the data here, as eo_tile's tile store does)
    // (WS = 2: only this wave's part, slot half `part`)
    auto ldf = [&](const char *stage, int kk, ST *sf, E (*df)[2]) {
        const int k = 4 * kk + lq;
#pragma unroll
        for (int i = 0; i < FP; ++i) {
            const int row = wp * Q::TP + i * 16 + l16;
            const int half = (k % PPS) ^ (SWAP ? (row >> 3) & 1 : 0);
            const char *a = stage + (row * Q::SPR + ((k / PPS) ^ (row & Q::SWM))) * 16 + half * SB;
            if (WS == 2) sf[i].x = *(const E *)(a + part * (SB / 2));
            else sf[i] = *(const ST *)a;
        }
#pragma unroll
        for (int j = 0; j < FD; ++j) {
            const int row = wd * Q::TD + j * 16 + l16;
            const char *a = stage + Q::SBYTES + row * Q::PITCH + k * 16;
            if (WS == 2) {
                df[j][0] = (E) * (const double *)(a + part * 8);
            } else {
                const d2 x = *(const d2 *)a;
                df[j][0] = (E)x.x;
                df[j][1] = (E)x.y;
            }
        }
    };
    auto mm = [&](const ST *sf, const E (*df)[2]) {
#pragma unroll
        for (int i = 0; i < FP; ++i)
#pragma unroll
            for (int j = 0; j < FD; ++j) {
                if (WS == 2) {
                    accO[i][j] = mfma16(sf[i].x, df[j][0], accO[i][j]);
                } else {
                    accO[i][j] = mfma16(sf[i].x, df[j][0], accO[i][j]);
                    accE[i][j] = mfma16(sf[i].y, df[j][1], accE[i][j]);
                }
            }
    };
    auto compute = [&](const char *stage) {
        ST sf[3][FP];
        E df[3][FD][2];
        ldf(stage, 0, sf[0], df[0]);
        ldf(stage, 1, sf[1], df[1]);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            if (kk + 2 < BK / 4) ldf(stage, kk + 2, sf[(kk + 2) % 3], df[(kk + 2) % 3]);
            mm(sf[kk % 3], df[kk % 3]);
        }
        // the reads of step kk+2 are issued ahead of the products of step kk
        __builtin_amdgcn_sched_group_barrier(0x100, 2 * (FP + FD), 0);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            if (kk + 2 < BK / 4) __builtin_amdgcn_sched_group_barrier(0x100, FP + FD, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, (WS == 1 ? 2 : 1) * FP * FD, 0);
        }
    };

    const int nkall = (h + BK - 1) / BK;
    const int kt0 = g.khi > 0 ? g.klo / BK : 0;
    const int nk = g.khi > 0 ? min(nkall, (g.khi + BK - 1) / BK) - kt0 : nkall;
    if (nk > 0) {
        issue(st0, kt0);
        for (int t = 0; t < nk; t += 2) {
            dma_wait_all();  // tile t has landed (this wave's loads) ...
            lds_barrier();   // ... for every wave, and every wave is done with tile t-1
            if (t + 1 < nk) issue(st1, kt0 + t + 1);
            compute(st0);
            if (t + 1 >= nk) break;
            dma_wait_all();
            lds_barrier();
            if (t + 2 < nk) issue(st0, kt0 + t + 2);
            compute(st1);
        }
    }

    double *xch = reinterpret_cast<double *>(st0);
    if (WS == 2) {
        // no load is in flight (the last tile's were waited for); every wave is past its products
        lds_barrier();
        if (part == 1) {
#pragma unroll
            for (int i = 0; i < FP; ++i)
#pragma unroll
                for (int j = 0; j < FD; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        xch[(((wl * FP + i) * FD + j) * 4 + r) * 64 + lane] = (double)accO[i][j][r];
        }
        lds_barrier();
        if (part == 1) return;
    }
#pragma unroll
    for (int i = 0; i < FP; ++i)
#pragma unroll
        for (int j = 0; j < FD; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int pi = r0s + wp * Q::TP + i * 16 + P::drow(lq, r);  // pair index a''
                const int di = r0d + wd * Q::TD + j * 16 + l16;             // data row
                if (pi >= h || di >= nrd) continue;
                const double e = WS == 2 ? xch[(((wl * FP + i) * FD + j) * 4 + r) * 64 + lane]
                                         : (double)accE[i][j][r];
                eo_store<EPI, TOUT, SIDE>(g, C, bz, pi, di, (double)accO[i][j][r], e);
            }
}

template <class P, int BM, int BN, int BK, int WM, int WN, int EPI, bool TOUT, int SIDE, int WS>
__global__ __launch_bounds__(WM *WN * 64 * WS) void k_gemm_eo_dma(GemmArgs g) {
    if (g.skip && *g.skip) return;
    ts_mark(g.ts, false, lin_block());
    int bxt, byt, bz;
    xcd_tile(bxt, byt, bz);
    bxt += g.olo[0] / BM;
    byt += g.olo[1] / BN;
    eo_tile_dma<P, BM, BN, BK, WM, WN, EPI, TOUT, SIDE, WS>(g, bxt, byt, bz);
    ts_mark(g.ts, true, lin_block());
}

template <class P, int BM, int BN, int BK, int WM, int WN, int EPI, bool TOUT, int SIDE, int WS>
__global__ __launch_bounds__(WM *WN * 64 * WS) void kg_gemm_eo_dma(const GemmArgs *gs, const int *off,
                                                                   int n, unsigned long long *ts) {
    ts_mark(ts, false, blockIdx.x);
    const int total = gridDim.x, q = total / 8, r = total % 8;
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const int T = xcd * q + min(xcd, r) + slot;
    const int p = grp_find(off, n, T);
    const GemmArgs g = gs[p];
    const int local = T - off[p];
    const int rows = SIDE == 0 ? g.h : g.M, cols = SIDE == 0 ? g.N : g.h;
    const int x0 = g.ohi[0] > 0 ? g.olo[0] / BM : 0, y0 = g.ohi[1] > 0 ? g.olo[1] / BN : 0;
    const int gx = ((g.ohi[0] > 0 ? min(g.ohi[0], rows) : rows) + BM - 1) / BM - x0;
    const int gy = ((g.ohi[1] > 0 ? min(g.ohi[1], cols) : cols) + BN - 1) / BN - y0;
    eo_tile_dma<P, BM, BN, BK, WM, WN, EPI, TOUT, SIDE, WS>(g, x0 + local % gx, y0 + (local / gx) % gy,
                                                        local / (gx * gy));
    ts_mark(ts, true, blockIdx.x);
}

// ------------------------------------------------------------------------ variant tables
typedef void (*KFn)(GemmArgs);
struct Variant {
    int bm, bn, nthreads;
    KFn fn[4][2];  // [epi][tout]
    const char *name;
};

#define EPIS(P, BM, BN, BK, WM, WN)                                                       \
    {{k_gemm<P, BM, BN, BK, WM, WN, EPI_STORE, false>,                                    \
      k_gemm<P, BM, BN, BK, WM, WN, EPI_STORE, true>},                                    \
     {k_gemm<P, BM, BN, BK, WM, WN, EPI_SCALE, false>, nullptr},                          \
     {k_gemm<P, BM, BN, BK, WM, WN, EPI_SUBFROM, false>, nullptr},                        \
     {k_gemm<P, BM, BN, BK, WM, WN, EPI_RHS, false>, nullptr}}
#define VAR(P, BM, BN, BK, WM, WN) \
    {BM, BN, WM * WN * 64, EPIS(P, BM, BN, BK, WM, WN), #P " " #BM "x" #BN "x" #BK " " #WM "x" #WN}

static const Variant VF64[] = {
    VAR(PF64, 64, 64, 16, 2, 2), VAR(PF64, 64, 64, 32, 2, 2), VAR(PF64, 128, 64, 16, 2, 2),
    VAR(PF64, 64, 128, 16, 2, 2), VAR(PF64, 128, 128, 16, 2, 2), VAR(PF64, 32, 64, 32, 1, 2),
    VAR(PF64, 64, 32, 32, 2, 1), VAR(PF64, 128, 64, 32, 2, 2),
};
static const Variant VF32[] = {
    VAR(PF32, 64, 64, 32, 2, 2), VAR(PF32, 128, 64, 32, 2, 2), VAR(PF32, 64, 128, 32, 2, 2),
    VAR(PF32, 128, 128, 16, 2, 2), VAR(PF32, 128, 128, 32, 2, 2), VAR(PF32, 64, 64, 64, 2, 2),
};
static const Variant VBF3[] = {
    VAR(PBF3, 64, 64, 32, 2, 2), VAR(PBF3, 128, 64, 32, 2, 2), VAR(PBF3, 64, 128, 32, 2, 2),
    VAR(PBF3, 128, 128, 32, 2, 2), VAR(PBF3, 128, 128, 64, 2, 2), VAR(PBF3, 64, 64, 64, 2, 2),
};

static const Variant *table(int prec, int *n) {
    if (prec == 1) { *n = sizeof(VF32) / sizeof(VF32[0]); return VF32; }
    if (prec == 2) { *n = sizeof(VBF3) / sizeof(VBF3[0]); return VBF3; }
    *n = sizeof(VF64) / sizeof(VF64[0]);
    return VF64;
}

int gemm_variant_count(int prec) {
    int n;
    table(prec, &n);
    return n;
}

const char *gemm_variant_name(int prec, int v) {
    int n;
    const Variant *t = table(prec, &n);
    return (v >= 0 && v < n) ? t[v].name : "?";
}

void launch_gemm(const GemmArgs &g, int prec, int variant, int epi, bool tout, hipStream_t s) {
    int n;
    const Variant *t = table(prec, &n);
    const Variant &V = t[(variant >= 0 && variant < n) ? variant : 0];
    dim3 grid((g.M + V.bm - 1) / V.bm, (g.N + V.bn - 1) / V.bn, g.batch);
    KFn fn = V.fn[epi][tout ? 1 : 0];
    hipLaunchKernelGGL(fn, grid, dim3(V.nthreads), 0, s, g);
}

// even/odd split variants: [side][epi] (side 0 is only used with the transposed store)
struct VariantEO {
    int bm, bn, nthreads;
    KFn fn[2][4];
    const char *name;
    int dma;  // operand tiles staged by LDS-direct loads (32-bit buffer offsets)
};
#define EOVW(P, BM, BN, BK, WM, WN, NB, WS, SUF)                                             \
    {BM, BN, WM * WN * 64 * WS,                                                                  \
     {{k_gemm_eo<P, BM, BN, BK, WM, WN, EPI_STORE, true, 0, NB, WS>, nullptr, nullptr, nullptr}, \
      {k_gemm_eo<P, BM, BN, BK, WM, WN, EPI_STORE, false, 1, NB, WS>,                            \
       k_gemm_eo<P, BM, BN, BK, WM, WN, EPI_SCALE, false, 1, NB, WS>,                            \
       k_gemm_eo<P, BM, BN, BK, WM, WN, EPI_SUBFROM, false, 1, NB, WS>,                          \
       k_gemm_eo<P, BM, BN, BK, WM, WN, EPI_RHS, false, 1, NB, WS>}},                            \
     "eo " #P " " #BM "x" #BN "x" #BK " " #WM "x" #WN " buf" #NB SUF, 0}
#define EOV(P, BM, BN, BK, WM, WN, NB) EOVW(P, BM, BN, BK, WM, WN, NB, 1, "")
// odd/even parts on separate waves (WS = 2)
#define EOVS(P, BM, BN, BK, WM, WN, NB) EOVW(P, BM, BN, BK, WM, WN, NB, 2, " oe-waves")

// LDS-direct variants (eo_tile_dma); WS = 2 puts the two parts on separate waves
#define EODMA(P, BM, BN, BK, WM, WN, WS, SUF)                                                      \
    {BM, BN, WM * WN * 64 * WS,                                                                     \
     {{k_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_STORE, true, 0, WS>, nullptr, nullptr, nullptr},    \
      {k_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_STORE, false, 1, WS>,                               \
       k_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_SCALE, false, 1, WS>,                               \
       k_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_SUBFROM, false, 1, WS>,                             \
       k_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_RHS, false, 1, WS>}},                               \
     "eo-dma " #P " " #BM "x" #BN "x" #BK " " #WM "x" #WN SUF, 1}

static const VariantEO EF64[] = {
    EOV(PF64, 64, 64, 16, 2, 2, 2), EOV(PF64, 32, 64, 16, 1, 2, 2), EOV(PF64, 64, 32, 16, 2, 1, 2),
    EOV(PF64, 32, 64, 32, 2, 2, 2), EOV(PF64, 64, 32, 32, 2, 2, 2), EOV(PF64, 64, 64, 32, 2, 2, 2),
    EOV(PF64, 64, 64, 16, 2, 2, 1), EOV(PF64, 32, 64, 16, 2, 2, 1), EOV(PF64, 64, 32, 16, 2, 2, 1),
    EOV(PF64, 32, 32, 16, 2, 2, 1), EOV(PF64, 64, 64, 8, 2, 2, 1),  EOV(PF64, 32, 64, 8, 2, 2, 1),
    EOV(PF64, 64, 64, 16, 2, 4, 1), EOV(PF64, 64, 64, 16, 4, 2, 1), EOV(PF64, 128, 64, 16, 4, 2, 1),
    EOV(PF64, 64, 128, 16, 2, 4, 1),
    // two-stage register prefetch (loads issued two K tiles ahead)
    EOV(PF64, 32, 64, 32, 2, 2, 3), EOV(PF64, 64, 32, 32, 2, 2, 3), EOV(PF64, 32, 64, 16, 2, 2, 3),
    EOV(PF64, 64, 64, 16, 2, 2, 3), EOV(PF64, 32, 32, 32, 2, 2, 3),
    // 64-deep K tiles: four times the MFMA work per tile's loads, stores and barriers
    EOV(PF64, 32, 32, 64, 2, 2, 1), EOV(PF64, 32, 64, 64, 2, 2, 1), EOV(PF64, 64, 32, 64, 2, 2, 1),
    EOV(PF64, 32, 32, 64, 2, 2, 2), EOV(PF64, 64, 64, 32, 2, 2, 1),
    // odd and even parts on separate waves: twice the resident waves for small grids
    EOVS(PF64, 32, 32, 16, 2, 2, 1), EOVS(PF64, 32, 32, 16, 2, 2, 2), EOVS(PF64, 32, 32, 32, 2, 2, 3),
    EOVS(PF64, 32, 32, 32, 2, 2, 1), EOVS(PF64, 32, 64, 16, 2, 2, 1), EOVS(PF64, 64, 32, 16, 2, 2, 1),
    EOVS(PF64, 16, 32, 16, 1, 2, 1), EOVS(PF64, 32, 16, 16, 2, 1, 1),
    // three and four register stages (loads issued three / four K tiles ahead)
    EOVS(PF64, 32, 32, 32, 2, 2, 4), EOVS(PF64, 32, 32, 32, 2, 2, 5), EOVS(PF64, 32, 32, 16, 2, 2, 4),
    EOVS(PF64, 32, 32, 16, 2, 2, 5), EOVS(PF64, 32, 64, 16, 2, 2, 4), EOVS(PF64, 64, 32, 16, 2, 2, 4),
    // operand tiles staged by LDS-direct loads
    EODMA(PF64, 32, 32, 16, 2, 2, 1, ""), EODMA(PF64, 64, 32, 16, 2, 2, 1, ""), EODMA(PF64, 32, 64, 16, 2, 2, 1, ""),
    EODMA(PF64, 32, 32, 32, 2, 2, 1, ""), EODMA(PF64, 32, 32, 16, 2, 2, 2, " oe-waves"),
    EODMA(PF64, 32, 32, 32, 2, 2, 2, " oe-waves"), EODMA(PF64, 32, 16, 16, 2, 1, 2, " oe-waves"),
    EODMA(PF64, 16, 32, 16, 1, 2, 2, " oe-waves"), EODMA(PF64, 32, 16, 32, 2, 1, 2, " oe-waves"),
    EODMA(PF64, 16, 32, 32, 1, 2, 2, " oe-waves"), EODMA(PF64, 64, 32, 16, 4, 2, 2, " oe-waves"),
    EODMA(PF64, 32, 64, 16, 2, 4, 2, " oe-waves"),
};
static const VariantEO EF32[] = {
    EOV(PF32, 64, 64, 16, 2, 2, 2), EOV(PF32, 32, 64, 32, 1, 2, 2), EOV(PF32, 64, 32, 32, 2, 1, 2),
    EOV(PF32, 64, 64, 32, 2, 2, 2), EOV(PF32, 64, 64, 16, 2, 2, 1), EOV(PF32, 32, 64, 16, 2, 2, 1),
    EOV(PF32, 64, 32, 16, 2, 2, 1), EOV(PF32, 64, 64, 32, 2, 2, 1),
    EOV(PF32, 32, 64, 32, 2, 2, 3), EOV(PF32, 64, 64, 32, 2, 2, 3),
    EOVS(PF32, 32, 64, 16, 2, 2, 1), EOVS(PF32, 64, 32, 16, 2, 2, 1),
    EODMA(PF32, 32, 32, 16, 2, 2, 1, ""), EODMA(PF32, 64, 32, 16, 2, 2, 1, ""),
    EODMA(PF32, 32, 64, 16, 2, 2, 1, ""), EODMA(PF32, 64, 64, 16, 2, 2, 1, ""),
    EODMA(PF32, 32, 32, 16, 2, 2, 2, " oe-waves"), EODMA(PF32, 32, 16, 32, 2, 1, 2, " oe-waves"),
    EODMA(PF32, 16, 32, 32, 1, 2, 2, " oe-waves"), EODMA(PF32, 64, 32, 16, 4, 2, 2, " oe-waves"),
    EODMA(PF32, 32, 64, 16, 2, 4, 2, " oe-waves"),
};
static const VariantEO EBF3[] = {
    EOV(PBF3, 64, 64, 32, 2, 2, 2), EOV(PBF3, 128, 64, 32, 2, 2, 2), EOV(PBF3, 64, 128, 32, 2, 2, 2),
    EOV(PBF3, 32, 64, 32, 1, 2, 2), EOV(PBF3, 64, 64, 32, 2, 2, 1), EOV(PBF3, 128, 64, 32, 2, 2, 1),
    EOV(PBF3, 64, 128, 32, 2, 2, 1), EOV(PBF3, 128, 128, 32, 2, 2, 1),
    EOV(PBF3, 64, 64, 32, 2, 2, 3),
};

static const VariantEO *table_eo(int prec, int *n) {
    if (prec == 1) { *n = sizeof(EF32) / sizeof(EF32[0]); return EF32; }
    if (prec == 2) { *n = sizeof(EBF3) / sizeof(EBF3[0]); return EBF3; }
    *n = sizeof(EF64) / sizeof(EF64[0]);
    return EF64;
}

int gemm_eo_variant_count(int prec) {
    int n;
    table_eo(prec, &n);
    return n;
}

const char *gemm_eo_variant_name(int prec, int v) {
    int n;
    const VariantEO *t = table_eo(prec, &n);
    return (v >= 0 && v < n) ? t[v].name : "?";
}

void launch_sine_low(const double *S, float *f, unsigned *b, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_sine_low, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, f, b, n);
}

bool dma_fits(const GemmArgs &g, int side, int prec) {
    const long ldx = side == 0 ? g.ldb : g.lda, rows = side == 0 ? g.N : g.M;
    const long data = (ldx * (rows - 1) + g.K) * 8, sine = (long)g.h * g.h * (prec == 0 ? 16 : 8);
    return (prec == 0 ? g.Soe != nullptr : prec == 1 ? g.Soef != nullptr : false) &&
           data < (1L << 31) && sine < (1L << 31);
}

void launch_sine_interleave(const float *So, const float *Se, float *Soe, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_sine_interleave_f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, So,
                       Se, Soe, n);
}

void launch_sine_interleave(const double *So, const double *Se, double *Soe, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_sine_interleave, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, So, Se,
                       Soe, n);
}

void launch_gemm_eo(const GemmArgs &g, int prec, int variant, int side, int epi, hipStream_t s) {
    if (!sine_ready(g, prec)) {
        std::fprintf(stderr, "libibhip: split contraction without its precision's sine parts\n");
        std::abort();
    }
    int n;
    const VariantEO *t = table_eo(prec, &n);
    if (variant < 0 || variant >= n || (t[variant].dma && !dma_fits(g, side, prec))) variant = 0;
    const VariantEO &V = t[variant];
    const int gm = side == 0 ? g.h : g.M, gn = side == 0 ? g.N : g.h;
    const int ex = g.ohi[0] > 0 ? std::min(g.ohi[0], gm) : gm;
    const int ey = g.ohi[1] > 0 ? std::min(g.ohi[1], gn) : gn;
    const int tx = (ex + V.bm - 1) / V.bm - (g.ohi[0] > 0 ? g.olo[0] / V.bm : 0);
    const int ty = (ey + V.bn - 1) / V.bn - (g.ohi[1] > 0 ? g.olo[1] / V.bn : 0);
    if (tx <= 0 || ty <= 0) return;
    GemmArgs ga = g;
    if (g.ohi[0] <= 0) ga.olo[0] = 0;
    if (g.ohi[1] <= 0) ga.olo[1] = 0;
    dim3 grid(tx, ty, g.batch);
    hipLaunchKernelGGL(V.fn[side][side == 0 ? 0 : epi], grid, dim3(V.nthreads), 0, s, ga);
}

// grouped split variants: [side][epi]
typedef void (*KGFn)(const GemmArgs *, const int *, int, unsigned long long *);
struct VariantGEO {
    int bm, bn, nthreads;
    KGFn fn[2][4];
    const char *name;
    int dma;
};
#define GEOVW(P, BM, BN, BK, WM, WN, NB, WS, SUF)                                             \
    {BM, BN, WM * WN * 64 * WS,                                                                   \
     {{kg_gemm_eo<P, BM, BN, BK, WM, WN, EPI_STORE, true, 0, NB, WS>, nullptr, nullptr, nullptr}, \
      {kg_gemm_eo<P, BM, BN, BK, WM, WN, EPI_STORE, false, 1, NB, WS>,                            \
       kg_gemm_eo<P, BM, BN, BK, WM, WN, EPI_SCALE, false, 1, NB, WS>,                            \
       kg_gemm_eo<P, BM, BN, BK, WM, WN, EPI_SUBFROM, false, 1, NB, WS>,                          \
       kg_gemm_eo<P, BM, BN, BK, WM, WN, EPI_RHS, false, 1, NB, WS>}},                            \
     "grouped eo " #P " " #BM "x" #BN "x" #BK " " #WM "x" #WN " buf" #NB SUF, 0}
#define GEOV(P, BM, BN, BK, WM, WN, NB) GEOVW(P, BM, BN, BK, WM, WN, NB, 1, "")
#define GEOVS(P, BM, BN, BK, WM, WN, NB) GEOVW(P, BM, BN, BK, WM, WN, NB, 2, " oe-waves")

#define GEODMA(P, BM, BN, BK, WM, WN, WS, SUF)                                                     \
    {BM, BN, WM * WN * 64 * WS,                                                                     \
     {{kg_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_STORE, true, 0, WS>, nullptr, nullptr, nullptr},   \
      {kg_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_STORE, false, 1, WS>,                              \
       kg_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_SCALE, false, 1, WS>,                              \
       kg_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_SUBFROM, false, 1, WS>,                            \
       kg_gemm_eo_dma<P, BM, BN, BK, WM, WN, EPI_RHS, false, 1, WS>}},                              \
     "grouped eo-dma " #P " " #BM "x" #BN "x" #BK " " #WM "x" #WN SUF, 1}

static const VariantGEO GF64[] = {
    GEOV(PF64, 64, 64, 16, 2, 2, 1), GEOV(PF64, 32, 64, 16, 2, 2, 1), GEOV(PF64, 64, 32, 16, 2, 2, 1),
    GEOV(PF64, 32, 64, 32, 2, 2, 2), GEOV(PF64, 64, 32, 32, 2, 2, 2), GEOV(PF64, 64, 64, 8, 2, 2, 1),
    GEOV(PF64, 32, 32, 16, 2, 2, 1), GEOV(PF64, 64, 64, 16, 2, 2, 3), GEOV(PF64, 32, 64, 32, 2, 2, 3),
    GEOV(PF64, 32, 32, 64, 2, 2, 1), GEOV(PF64, 32, 64, 64, 2, 2, 1),
    GEOVS(PF64, 32, 32, 16, 2, 2, 1), GEOVS(PF64, 32, 64, 16, 2, 2, 1),
    GEOVS(PF64, 32, 32, 32, 2, 2, 4), GEOVS(PF64, 32, 64, 16, 2, 2, 4),
    GEODMA(PF64, 32, 32, 16, 2, 2, 1, ""), GEODMA(PF64, 64, 32, 16, 2, 2, 1, ""), GEODMA(PF64, 32, 64, 16, 2, 2, 1, ""),
    GEODMA(PF64, 32, 32, 16, 2, 2, 2, " oe-waves"), GEODMA(PF64, 64, 32, 16, 4, 2, 2, " oe-waves"),
    GEODMA(PF64, 32, 64, 16, 2, 4, 2, " oe-waves"), GEODMA(PF64, 32, 16, 32, 2, 1, 2, " oe-waves"),
    GEODMA(PF64, 16, 32, 32, 1, 2, 2, " oe-waves"),
};
static const VariantGEO GF32[] = {
    GEOV(PF32, 64, 64, 16, 2, 2, 1), GEOV(PF32, 64, 64, 32, 2, 2, 1), GEOV(PF32, 32, 64, 16, 2, 2, 1),
    GEOV(PF32, 64, 32, 16, 2, 2, 1), GEOV(PF32, 64, 64, 32, 2, 2, 3),
    GEOVS(PF32, 32, 32, 16, 2, 2, 1), GEOVS(PF32, 32, 64, 16, 2, 2, 1), GEOVS(PF32, 64, 64, 16, 2, 2, 1),
    GEOVS(PF32, 32, 32, 32, 2, 2, 4), GEOVS(PF32, 32, 64, 32, 2, 2, 3),
    GEODMA(PF32, 32, 32, 16, 2, 2, 1, ""), GEODMA(PF32, 64, 32, 16, 2, 2, 1, ""),
    GEODMA(PF32, 32, 64, 16, 2, 2, 1, ""), GEODMA(PF32, 64, 64, 16, 2, 2, 1, ""),
    GEODMA(PF32, 32, 32, 16, 2, 2, 2, " oe-waves"), GEODMA(PF32, 64, 32, 16, 4, 2, 2, " oe-waves"),
    GEODMA(PF32, 32, 64, 16, 2, 4, 2, " oe-waves"),
};
static const VariantGEO GBF3[] = {
    GEOV(PBF3, 64, 64, 32, 2, 2, 1), GEOV(PBF3, 128, 64, 32, 2, 2, 1), GEOV(PBF3, 64, 128, 32, 2, 2, 1),
    GEOV(PBF3, 64, 64, 32, 2, 2, 2), GEOV(PBF3, 64, 64, 32, 2, 2, 3),
    GEOVS(PBF3, 32, 32, 32, 2, 2, 1), GEOVS(PBF3, 32, 64, 32, 2, 2, 1), GEOVS(PBF3, 64, 64, 32, 2, 2, 1),
    GEOVS(PBF3, 32, 32, 32, 2, 2, 3), GEOVS(PBF3, 32, 64, 64, 2, 2, 1),
};

static const VariantGEO *table_geo(int prec, int *n) {
    if (prec == 1) { *n = sizeof(GF32) / sizeof(GF32[0]); return GF32; }
    if (prec == 2) { *n = sizeof(GBF3) / sizeof(GBF3[0]); return GBF3; }
    *n = sizeof(GF64) / sizeof(GF64[0]);
    return GF64;
}

bool gemm_eo_grp_variant_dma(int prec, int v) {
    int n;
    const VariantGEO *t = table_geo(prec, &n);
    return v >= 0 && v < n && t[v].dma;
}

int gemm_eo_grp_variant_count(int prec) {
    int n;
    table_geo(prec, &n);
    return n;
}

const char *gemm_eo_grp_variant_name(int prec, int v) {
    int n;
    const VariantGEO *t = table_geo(prec, &n);
    return (v >= 0 && v < n) ? t[v].name : "?";
}

long grouped_eo_tiles(const GemmArgs &g, int prec, int variant, int side) {
    int n;
    const VariantGEO *t = table_geo(prec, &n);
    const VariantGEO &V = t[(variant >= 0 && variant < n) ? variant : 0];
    const int gm = side == 0 ? g.h : g.M, gn = side == 0 ? g.N : g.h;
    const int ex = g.ohi[0] > 0 ? std::min(g.ohi[0], gm) : gm;
    const int ey = g.ohi[1] > 0 ? std::min(g.ohi[1], gn) : gn;
    const long tx = (ex + V.bm - 1) / V.bm - (g.ohi[0] > 0 ? g.olo[0] / V.bm : 0);
    const long ty = (ey + V.bn - 1) / V.bn - (g.ohi[1] > 0 ? g.olo[1] / V.bn : 0);
    return std::max(0L, tx) * std::max(0L, ty) * g.batch;
}

void launch_gemm_eo_grp(const Grp &gr, int prec, int variant, int side, int epi, hipStream_t s,
                        unsigned long long *ts) {
    if (gr.n == 0 || gr.blocks == 0) return;
    int n;
    const VariantGEO *t = table_geo(prec, &n);
    const VariantGEO &V = t[(variant >= 0 && variant < n) ? variant : 0];
    hipLaunchKernelGGL(V.fn[side][side == 0 ? 0 : epi], dim3(gr.blocks), dim3(V.nthreads), 0, s,
                       static_cast<const GemmArgs *>(gr.args), gr.off, gr.n, ts);
}

}  // namespace ibh
