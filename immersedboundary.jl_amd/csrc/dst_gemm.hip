// Sine-matrix contractions on MFMA (gfx950), FP64.
//
// A 2-D DST-I of an m×p field X is S_m·X·S_p with S[i,k] = sin(π i k / n) (FFTW RODFT00 is
// 2·S per dimension, fluid-ops.jl:25,49,51).  Each contraction is a column-major batched GEMM
// C = A·B computed here with v_mfma_f64_16x16x4f64.  The MFMA is fed the transposed problem
// (Cᵀ = Bᵀ·Aᵀ) so that lane%16 walks the contiguous (x) index of C and stores coalesce.
//
// Block tile 64×64, K tile 16, 256 threads = 4 waves in 2×2; each wave owns a 32×32 sub-tile
// (2×2 MFMA blocks).  K tiles are double-buffered through LDS; global loads of the next tile
// overlap the MFMAs of the current one.  Epilogues fuse the eigenvalue scaling between the
// two transforms, the Crank–Nicolson right-hand side, and the circulation projection.
#include "ibh_internal.h"

namespace ibh {

typedef double d4 __attribute__((ext_vector_type(4)));

static constexpr int BM = 64, BN = 64, BK = 16, LDP = 80;  // LDS row pitch (doubles)

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f64(GemmArgs g) {
    if (g.skip && *g.skip) return;
    const int bz = blockIdx.z;
    const double *__restrict__ A = g.A + bz * g.sA;
    const double *__restrict__ B = g.B + bz * g.sB;
    double *__restrict__ C = g.C + bz * g.sC;
    const int M = g.M, N = g.N, K = g.K;
    const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    __shared__ double As[2][BK][LDP];  // [k][m]
    __shared__ double Bs[2][BK][LDP];  // [k][n]

    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    double ra[4], rb[4];
    const int am = tid & 63, ak = (tid >> 6) * 4;   // A: 64 m × 4 k-quads
    const int bn = tid >> 2, bk = (tid & 3) * 4;    // B: 64 n × 4 k-quads
    const bool a_in = (m0 + am) < M;
    const bool b_in = (n0 + bn) < N;
    const double *Ap = A + (m0 + am);
    const double *Bp = B + ldb * (long)(n0 + bn);

    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int k = k0 + ak + u;
            ra[u] = (a_in && k < K) ? Ap[lda * (long)k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int k = k0 + bk + u;
            rb[u] = (b_in && k < K) ? Bp[k] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) As[buf][ak + u][am] = ra[u];
#pragma unroll
        for (int u = 0; u < 4; ++u) Bs[buf][bk + u][bn] = rb[u];
    };

    const int nk = (K + BK - 1) / BK;
    gload(0);
    sstore(0);
    __syncthreads();
    const int l16 = lane & 15, lq = lane >> 4;
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            const int kr = kk * 4 + lq;
            double a0 = Bs[cur][kr][wn * 32 + l16];
            double a1 = Bs[cur][kr][wn * 32 + 16 + l16];
            double b0 = As[cur][kr][wm * 32 + l16];
            double b1 = As[cur][kr][wm * 32 + 16 + l16];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (kt + 1 < nk) sstore(cur ^ 1);
        __syncthreads();
    }

    // epilogue.  v_mfma_f64_16x16x4f64 leaves D[row][col] in lane col + 16·(row % 4),
    // register row / 4 (row = lane/16 + 4·r); here row ↔ n and col ↔ m, so acc[ni][mi][r]
    // holds C[m][n] with m = l16 (+16 mi), n = lq + 4·r (+16 ni).
    const double *D = g.D ? g.D + bz * g.sD : nullptr;
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm * 32 + mi * 16 + l16;
                const int n = n0 + wn * 32 + ni * 16 + lq + 4 * r;
                if (m < M && n < N) {
                    double v = acc[ni][mi][r];
                    const long ci = m + ldc * (long)n;
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
            }
}

void launch_gemm_f64(const GemmArgs &g, int epi, hipStream_t s) {
    dim3 grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, g.batch);
    dim3 block(256);
    switch (epi) {
    case EPI_STORE: hipLaunchKernelGGL(k_gemm_f64<EPI_STORE>, grid, block, 0, s, g); break;
    case EPI_SCALE: hipLaunchKernelGGL(k_gemm_f64<EPI_SCALE>, grid, block, 0, s, g); break;
    case EPI_SUBFROM: hipLaunchKernelGGL(k_gemm_f64<EPI_SUBFROM>, grid, block, 0, s, g); break;
    default: hipLaunchKernelGGL(k_gemm_f64<EPI_RHS>, grid, block, 0, s, g); break;
    }
}

}  // namespace ibh
