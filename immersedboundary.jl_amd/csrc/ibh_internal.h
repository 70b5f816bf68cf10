// Internal declarations shared by the libibhip translation units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ibh {

// ---------------------------------------------------------------------------- grid geometry
// Sizes of one level of the reference layout (src/streamfcn/fluids.jl:28-45).
struct Dom {
    int nx, ny;     // cells
    int m, p;       // Γ is m×p = (nx-1)×(ny-1)
    long nu, nq;    // qx block (nx+1)·ny, total flux nq
    long ng;        // m·p
    int nbc;        // 2(nx+1)+2(ny+1)
    int Lo, Ro, Bo, To;  // bc offsets
};

inline Dom make_dom(int nx, int ny) {
    Dom d;
    d.nx = nx; d.ny = ny; d.m = nx - 1; d.p = ny - 1;
    d.nu = (long)(nx + 1) * ny;
    d.nq = d.nu + (long)nx * (ny + 1);
    d.ng = (long)d.m * d.p;
    d.nbc = 2 * (nx + 1) + 2 * (ny + 1);
    d.Lo = 0; d.Ro = ny + 1; d.Bo = 2 * (ny + 1); d.To = 2 * (ny + 1) + nx + 1;
    return d;
}

// Index arithmetic of the grid kernels: an element index within one level (always < 2^31)
// split by a run-time extent in 32-bit unsigned arithmetic, which the compiler expands to a
// few instructions (a 64-bit division is a long sequence).
#ifdef __HIPCC__
__device__ __forceinline__ int idiv(long t, long n) { return (int)((unsigned)t / (unsigned)n); }
__device__ __forceinline__ int imod(long t, long n) { return (int)((unsigned)t % (unsigned)n); }
// The linear workgroup id (over a 1-D or 2-D grid) renumbered so that each XCD runs one
// contiguous range of ids: the hardware deals workgroup i to XCD i mod 8, so without this the
// neighbouring tiles of a stencil land on different XCDs and fetch their shared halo rows into
// different L2s.  A bijection of [0, total).
__device__ __forceinline__ int xcd_block() {
    const int total = gridDim.x * gridDim.y, q = total >> 3, r = total & 7;
    const int lin = blockIdx.x + gridDim.x * blockIdx.y;
    const int x = lin & 7, slot = lin >> 3;
    return x * q + min(x, r) + slot;
}
#endif

// curl!(q, ψ, ψbc, domain), fluid-ops.jl:138-182: the u face (i, j) and the v face (i, j)
// (1-based within each block), and the value of flux entry t (u faces first, then v faces);
// each q entry is written once by its caller.  (Inside a HIP source only: the curl kernels, E
// and the nonlinear term evaluate it.)
#ifdef __HIPCC__
// G(a, b): ψ at interior vertex (a, b) (1-based), from device memory or a staged copy
template <class GA>
__device__ inline double curl_u_g(const Dom &d, const GA &G, const double *__restrict__ pbc, int i,
                                  int j) {
    const int nx = d.nx, ny = d.ny;
    auto B = [&](int k) { return pbc[k - 1]; };
    if (i == 1) return B((d.Lo + 1) + j) - B(d.Lo + j);
    if (i == nx + 1) return B((d.Ro + 1) + j) - B(d.Ro + j);
    if (j == 1) return G(i - 1, j) - B(d.Bo + i);
    if (j == ny) return B(i + d.To) - G(i - 1, j - 1);
    return G(i - 1, j) - G(i - 1, j - 1);
}
template <class GA>
__device__ inline double curl_v_g(const Dom &d, const GA &G, const double *__restrict__ pbc, int i,
                                  int j) {
    const int nx = d.nx, ny = d.ny;
    auto B = [&](int k) { return pbc[k - 1]; };
    if (j == 1) return B(d.Bo + i) - B((d.Bo + 1) + i);
    if (j == ny + 1) return B(d.To + i) - B((d.To + 1) + i);
    if (i == 1) return B(d.Lo + j) - G(i, j - 1);
    if (i == nx) return G(i - 1, j - 1) - B(d.Ro + j);
    return G(i - 1, j - 1) - G(i, j - 1);
}
__device__ inline double curl_u(const Dom &d, const double *__restrict__ psi,
                                const double *__restrict__ pbc, int i, int j) {
    return curl_u_g(d, [&](int a, int b) { return psi[(a - 1) + (long)d.m * (b - 1)]; }, pbc, i, j);
}
__device__ inline double curl_v(const Dom &d, const double *__restrict__ psi,
                                const double *__restrict__ pbc, int i, int j) {
    return curl_v_g(d, [&](int a, int b) { return psi[(a - 1) + (long)d.m * (b - 1)]; }, pbc, i, j);
}
__device__ inline double curl_at(const Dom &d, const double *__restrict__ psi,
                                 const double *__restrict__ pbc, long t) {
    if (t < d.nu) return curl_u(d, psi, pbc, imod(t, d.nx + 1) + 1, idiv(t, d.nx + 1) + 1);
    const long u = t - d.nu;
    return curl_v(d, psi, pbc, imod(u, d.nx) + 1, idiv(u, d.nx) + 1);
}
#endif

// A rectangle of the interior vertices (0-based [x0, x1] × [y0, y1]) and the flux faces a
// kernel touches around it (1-based [x0, x1 + 3] × [y0, y1 + 3], clipped to each block): the
// body box of the B-products.  x0 < 0 means the whole grid.
struct GBox { int x0 = -1, x1 = 0, y0 = 0, y1 = 0; };
__host__ __device__ inline long gbox_vertices(const Dom &d, const GBox &b) {
    return b.x0 < 0 ? d.ng : (long)(b.x1 - b.x0 + 1) * (b.y1 - b.y0 + 1);
}
__host__ __device__ inline long gbox_vertex(const Dom &d, const GBox &b, long k) {
    if (b.x0 < 0) return k;
    const int w = b.x1 - b.x0 + 1;
    return (b.x0 + k % w) + (long)d.m * (b.y0 + k / w);
}
// faces of the box: the u block's [max(1, x0), min(nx+1, x1+3)] × [max(1, y0), min(ny, y1+3)]
// first, then the v block's [max(1, x0), min(nx, x1+3)] × [max(1, y0), min(ny+1, y1+3)]
__host__ __device__ inline void gbox_face_ranges(const Dom &d, const GBox &b, int *r) {
    const int i0 = b.x0 > 1 ? b.x0 : 1, j0 = b.y0 > 1 ? b.y0 : 1;
    r[0] = i0; r[1] = b.x1 + 3 < d.nx + 1 ? b.x1 + 3 : d.nx + 1;   // u: i
    r[2] = j0; r[3] = b.y1 + 3 < d.ny ? b.y1 + 3 : d.ny;           // u: j
    r[4] = i0; r[5] = b.x1 + 3 < d.nx ? b.x1 + 3 : d.nx;           // v: i
    r[6] = j0; r[7] = b.y1 + 3 < d.ny + 1 ? b.y1 + 3 : d.ny + 1;   // v: j
}
__host__ __device__ inline long gbox_faces(const Dom &d, const GBox &b) {
    if (b.x0 < 0) return d.nq;
    int r[8];
    gbox_face_ranges(d, b, r);
    return (long)(r[1] - r[0] + 1) * (r[3] - r[2] + 1) + (long)(r[5] - r[4] + 1) * (r[7] - r[6] + 1);
}
__host__ __device__ inline long gbox_face(const Dom &d, const GBox &b, long k) {
    if (b.x0 < 0) return k;
    int r[8];
    gbox_face_ranges(d, b, r);
    const int wu = r[1] - r[0] + 1;
    const long nu = (long)wu * (r[3] - r[2] + 1);
    if (k < nu) return (r[0] - 1 + k % wu) + (long)(d.nx + 1) * (r[2] - 1 + k / wu);
    k -= nu;
    const int wv = r[5] - r[4] + 1;
    return d.nu + (r[4] - 1 + k % wv) + (long)d.nx * (r[6] - 1 + k / wv);
}

// ---------------------------------------------------------------------------- sine GEMM
// C[b] = op(A[b]·B[b]) with A(m, k) at A[k + lda·m] and B(k, n) at B[k + ldb·n] (both
// k-fastest); batch b = blockIdx.z.
enum Epi {
    EPI_STORE = 0,   // C = acc
    EPI_SCALE = 1,   // C = acc ⊙ D          (eigenvalue scaling between the two transforms)
    EPI_SUBFROM = 2, // C = D - acc           (circulation projection Γ₁ = Γs₁ - Ainv·…)
    EPI_RHS = 3      // C = acc + f1·N1 + f2·N2 + edges(bcacc)  (CN right-hand side)
};

struct GemmArgs {
    int M, N, K;
    const double *A; long lda, sA;
    const double *B; long ldb, sB;
    double *C; long ldc, sC;
    const double *D; long ldd, sD;
    // EPI_RHS extras: N1, N2 (ld = ldc, batch stride sC), edge accumulators (nbc per batch)
    const double *N1, *N2, *bcacc;
    double f1, f2;
    int nbc, Lo, Ro, Bo, To;
    const int *skip;  // if non-null and *skip != 0 the launch does nothing
    // even/odd split: h = n/2 pairs, So/Se h×h sine parts stored k-fastest ([k + h·a])
    const double *So, *Se;
    // the same parts rounded for the low-precision contractions: FP32 (Sof, Sef) and bf16
    // pairs (Sob, Seb: the rounded value in the low 16 bits, the rounded remainder in the high)
    const float *Sof, *Sef;
    const unsigned *Sob, *Seb;
    // the FP64 parts interleaved, Soe[2(k + h·a)] = So[a][k], Soe[2(k + h·a) + 1] = Se[a][k]
    // (one 16-byte element per pair: the tiles staged by LDS-direct loads read both at once)
    const double *Soe;
    const float *Soef;  // the FP32 parts interleaved likewise
    int h;
    int batch;
    // even/odd split: only pair indices [klo, khi) of the data's contiguous dimension are
    // summed (the data is zero elsewhere); khi = 0 sums all of them
    int klo, khi;
    // split launches of one case: only the tiles that cover elements [olo[0], ohi[0]) of the
    // grid's x extent (pairs for SIDE 0, data rows for SIDE 1) and [olo[1], ohi[1]) of its y
    // extent (data rows, pairs) run; ohi = 0 covers the whole extent.  Outputs outside the
    // launched tiles keep their previous contents.
    int olo[2], ohi[2];
    // profiling only: a PROF_REC record of device wall-clock ticks, zeroed beforehand: the
    // start of workgroup 0 (the first dispatched) at ts[0], each workgroup's end folded by
    // atomicMax into one of 16 sub-slots (ts[16 + 16·j], 128 B apart, so that the end
    // atomics do not pile onto one address); a launch skipped by its flag writes nothing.
    // Null when not profiling.
    unsigned long long *ts;
};
constexpr int PROF_SLOTS = 8192;  // profiled launches per profiling session
constexpr int PROF_REC = 272;     // u64 per profiled launch

// prec: 0 FP64, 1 FP32, 2 bf16x3; variant indexes the tile table of that precision;
// tout writes C transposed (C(m, n) at C[n + ldc·m], EPI_STORE only).
void launch_gemm(const GemmArgs &g, int prec, int variant, int epi, bool tout, hipStream_t s);
int gemm_variant_count(int prec);
const char *gemm_variant_name(int prec, int v);
// even/odd split contraction: side 0 C = S·X (transposed store), side 1 C = X·S (any epi)
void launch_gemm_eo(const GemmArgs &g, int prec, int variant, int side, int epi, hipStream_t s);
int gemm_eo_variant_count(int prec);
const char *gemm_eo_variant_name(int prec, int v);

// ---------------------------------------------------------------------------- fluid kernels
// All take a batch count `nb` over levels (or columns) with element strides.
void launch_curl_bc(const Dom &d, const double *psi, long spsi, const double *pbc, long sbc,
                    double *q, long sq, int nbatch, const int *skip, hipStream_t s);
// curl on the faces with 1-based indices [i0, i1] × [j0, j1] only (clipped to the grid)
void launch_curl_box(const Dom &d, const double *psi, long spsi, const double *pbc, long sbc,
                     double *q, long sq, int nbatch, const int *skip, hipStream_t s, int i0,
                     int i1, int j0, int j1);
void launch_rot(const Dom &d, const double *q, long sq, double *g, long sg, int nbatch,
                const int *skip, hipStream_t s);
// rot of the q whose fixed-point image (scale per batch entry) launch_ET_image left in img
// (with a box: only Γ inside it; the caller keeps Γ zero outside)
// With lrk > 0 its extra workgroups also form w[b, j] = V[:, j]·fb[b] (V n2 × lrk
// column-major), as launch_ET_image would.
void launch_rot_image(const Dom &d, const double *img, long sq, const double *scale, double *g,
                      long sg, int nbatch, const int *skip, hipStream_t s, GBox box = GBox(),
                      const double *V = nullptr, const double *fb = nullptr, long sfb = 0,
                      int n2 = 0, double *w = nullptr, int lrk = 0);
void launch_coarsify(const Dom &d, double *gc, const double *g, long sgc, long sg, int nbatch,
                     const int *skip, hipStream_t s);
void launch_get_bc(const Dom &d, const double *r, long sr, double *rbc, long sbc, int nbatch,
                   const int *skip, hipStream_t s);
// r[b] += fac[b]·edges(rbc[b])   (fac per batch entry on the host side, passed by value array)
void launch_apply_bc(const Dom &d, double *r, long sr, const double *rbc, long sbc, double fac,
                     int nbatch, const int *skip, hipStream_t s);
// out[b] = in[b] + edges(rbc[b])·fac
// rbc = get_bc!(r) and out = in + edges(rbc) in one pass (the Poisson right-hand side of a level
// from the next coarser level's ψ r)
void launch_copy_getbc(const Dom &d, const double *in, long sin, double *out, long sout,
                       const double *r, long sr, double *rbc, long sbc, int nbatch,
                       const int *skip, hipStream_t s);
void launch_copy_apply_bc(const Dom &d, const double *in, long sin, double *out, long sout,
                          const double *rbc, long sbc, double fac, int nbatch, const int *skip,
                          hipStream_t s);
// nonlinear term for levels lev0..lev0+nbatch-1 (level stride sq / sg); q0 is the constant
// (ux·hc, uy·hc) per level (or (ux, uy) as given when q0_given); gbc per level with stride
// sbc, multiplied by gbc_scale.
// With psi set the flux of level b is curl!(ψ_b, ψbc_b) (ψ, ψbc level strides spsi, spbc),
// formed while the tile is staged, and q is not read.
void launch_nonlinear(const Dom &d, const double *q, long sq, const double *g, long sg,
                      const double *gbc, long sbc, double gbc_scale, double *out, long sout,
                      double ux, double uy, double h, int lev0, int nbatch, hipStream_t s,
                      int q0_given = 0, const double *uv = nullptr, const double *psi = nullptr,
                      long spsi = 0, const double *pbc = nullptr, long spbc = 0);
// accbc[ℓ] = Σ_{l=ℓ}^{L-2} fac_l·bcraw[l]  for ℓ = 0..L-2, accbc[L-1] = 0
void launch_bc_prefix(const Dom &d, const double *bcraw, double *accbc, const double *fac, int L,
                      hipStream_t s);
void launch_fill(double *p, long n, double v, hipStream_t s);
// p[b·stride + i] = v for i < n, b < count
void launch_fill_strided(double *p, long n, long stride, int count, double v, hipStream_t s);
// F[b] += D ⊙ (S_m E_b S_p) for the edge field E_b of the bc array bc[b] (apply_bc! with fac 1):
// the rank-4 sine transform of a Poisson right-hand side's boundary terms.  W: work of
// 2(m + p) doubles per batch entry; D: the m×p scale table (shared by the batch).
void launch_edge_spectral(const Dom &d, const double *Sm, const double *Sp, const double *bc,
                          long sbc, double *W, const double *D, double *F, long sF, int nbatch,
                          const int *skip, hipStream_t s);
// the GEMV half of it only: W = [S_m cB | S_m cT | S_p cL | S_p cR], for kernels that add the
// rank-4 term where they read the spectrum
void launch_edge_gemv(const Dom &d, const double *Sm, const double *Sp, const double *bc,
                      long sbc, double *W, int nbatch, const int *skip, hipStream_t s);
// ψ[b] = S_m F[b] S_p on the rows i = nx/4, 3nx/4 and the columns j = ny/4, 3ny/4 only (the
// values get_bc! reads); other entries of ψ[b] are left as they were.  W: lines_work_size(d)
// doubles per batch entry.
long lines_work_size(const Dom &d);
// F[b] = kappa·A[b] ⊙ D (D shared by the batch), plus D ⊙ (rank-4 boundary term) with EW
void launch_spectral_scale(const Dom &d, const double *A, long sA, const double *D, double kappa,
                           double *F, long sF, int nbatch, const int *skip, hipStream_t s,
                           const double *Sm = nullptr, const double *Sp = nullptr,
                           const double *EW = nullptr, const double *sub = nullptr);
// EW (optional, from launch_edge_gemv): read F + D ⊙ (rank-4 boundary term) instead of F.
void launch_lines(const Dom &d, const double *Sm, const double *Sp, const double *F, long sF,
                  double *psi, long spsi, double *W, int nbatch, const int *skip, hipStream_t s,
                  const double *EW = nullptr, const double *D = nullptr);
// trial right-hand side of L levels: out = A_ℓ·g + f1·N1 + f2·N2 + edges(bcacc), A_ℓ as the
// exact 5-point stencil I − a_ℓ(−L), a_ℓ = dt/(2 Re (h·2^ℓ)²)
void launch_rhs(const Dom &d, const double *g, const double *N1, const double *N2,
                const double *bcacc, double *out, int L, double h, double dt, double Re,
                double f1, double f2, hipStream_t s);
// r = M·x − rhs (FP64) with M = −L (a < 0) or I + a(−L), zero values beyond the edges, on
// nlev consecutive levels (stride ng); level b uses a·ar^b (ar = 1/4 for the CN operator)
void launch_resid(const Dom &d, const double *x, const double *rhs, double a, double ar,
                  double *r, int nlev, hipStream_t s);
void launch_sine_matrix(double *S, int n, hipStream_t s);  // S[i,k] = sin(π(i+1)(k+1)/n), (n-1)²
// So[a][k] = sin(π(a+1)(2k+1)/n), Se[a][k] = sin(2π(a+1)(k+1)/n) (0 on the last row/column),
// both h×h (h = n/2) stored [k + h·a]
void launch_sine_split(double *So, double *Se, int n, hipStream_t s);
// FP32 and bf16-pair copies of n sine elements, rounded as the contraction rounds FP64 operands
void launch_sine_low(const double *S, float *f, unsigned *b, long n, hipStream_t s);
// Soe[2t] = So[t], Soe[2t + 1] = Se[t] for t < n (the interleaved parts of GemmArgs::Soe)
void launch_sine_interleave(const double *So, const double *Se, double *Soe, long n, hipStream_t s);
void launch_sine_interleave(const float *So, const float *Se, float *Soe, long n, hipStream_t s);
// a split contraction in precision prec (0 FP64, 1 FP32, 2 bf16x3) has its sine parts
inline bool sine_ready(const GemmArgs &g, int prec) {
    return prec == 0 ? g.So && g.Se && g.Soe : prec == 1 ? g.Sof && g.Sef && g.Soef : g.Sob && g.Seb;
}
void launch_scale_tables(const Dom &d, double *sD, double *sA, double *sAinv, int L, double h,
                         double dt, double Re, hipStream_t s);

// adjoints of the multilevel transfers (the coarse-level part of the B-product, transposed)
// bc[b] = edgesᵀ(f[b]): the transpose of apply_bc!'s edge addition
void launch_edges_adj(const Dom &d, const double *f, long sf, double *bc, long sbc, int nbatch,
                      hipStream_t s);
// a[b] = get_bcᵀ(bc[b]) over the whole field (zero off get_bc!'s four lines)
void launch_get_bc_adj(const Dom &d, const double *bc, long sbc, double *a, long sa, int nbatch,
                       hipStream_t s);
// out[b] = base[b] + coarsifyᵀ(y[b]) (base may be null)
void launch_coarsify_adj(const Dom &d, const double *y, long sy, const double *base, long sbase,
                         double *out, long sout, int nbatch, hipStream_t s);
// out[b] = a[b] − b[b] on the flux faces [i0, i1] × [j0, j1] (1-based), zero elsewhere
void launch_box_diff(const Dom &d, const double *a, const double *b, long sq, double *out,
                     int nbatch, int i0, int i1, int j0, int j1, hipStream_t s);
// deterministic uniform values in [−1, 1) on the vertex box [xa, xb] × [ya, yb], zero elsewhere
void launch_fill_random_box(const Dom &d, double *p, long stride, int count,
                            unsigned long long seed, int xa, int xb, int ya, int yb,
                            hipStream_t s);

// ---------------------------------------------------------------------------- body kernels
// The delta support of a point in cell (bx, by) (1-based, body_idx of structure-ops.jl:48-49)
// covers qx[i, j], qy[i, j] for i ∈ bx-6..bx+6, j ∈ by-6..by+6; it stays inside both flux
// arrays (qx (nx+1)×ny, qy nx×(ny+1)) when 1 ≤ i ≤ nx and 1 ≤ j ≤ ny.  The reference raises
// BoundsError otherwise (structure-ops.jl:70-72, 97-99).
constexpr int DELTA_SUPP = 6;
__host__ __device__ inline bool support_inside(const Dom &d, int bx, int by) {
    return bx >= 1 + DELTA_SUPP && bx <= d.nx - DELTA_SUPP && by >= 1 + DELTA_SUPP &&
           by <= d.ny - DELTA_SUPP;
}

struct RegDev {
    int nb;
    int *bidx;     // 2·nb  (x index block, then y)
    double *w;     // nb × 4 × 13 : (wx_u, wy_u, wx_v, wy_v) 1-D delta factors
    int *err;      // device error flag (1: out of bounds, 2: non-finite flux)
    int esched, etsched;  // block schedules of E and Eᵀ (host side; see e_sched_name)
};
// Block schedules of E and Eᵀ.  E: waves (outputs) per workgroup; Eᵀ: workgroup size and the
// order in which a workgroup's threads walk the 2nb·169 stencil entries (entry-major: the 169
// entries of one point together; point-major: one entry of consecutive points together).  Every
// schedule gives the same bits: E forms each output in one wave in a fixed order, and Eᵀ's
// fixed-point adds do not depend on their order.  Schedule 0 is the default.
int e_sched_count();
int et_sched_count();
const char *e_sched_name(int v);
const char *et_sched_name(int v);
// "e,et" (IBH_REG_SCHED): schedules forced instead of timed; a missing or negative entry is left
// to the tuner
void parse_reg_sched(const char *s, int *e, int *et);
void launch_reg_update(const Dom &d, const RegDev &r, const double *pts, double h, double x0,
                       double y0, hipStream_t s);
// fb = E·(q + q0) where q0 = (u0, v0) constant (pass 0, 0 for plain E)
// With lrk > 0 also fb[b] += U·w[b] (U 2nb × lrk, w lrk per batch entry): the low-rank
// coarse-level part of a B-product, summed inside E's wave reduction.  Only the entries
// [t0, t0 + tn) of each fb[b] are formed (tn < 0: through the end).
// With psi set, q is not read: each flux value E needs is curl_at(ψ, ψbc) evaluated on the fly
// (ψ, ψbc: level-1 column stride spsi, spbc), the same values curl! would have stored.
void launch_E(const Dom &d, const RegDev &r, const double *q, long sq, double u0, double v0,
              double *fb, long sfb, int nbatch, const int *skip, hipStream_t s,
              const double *uv = nullptr, double uvscale = 1.0, const double *U = nullptr,
              const double *w = nullptr, int lrk = 0, int t0 = 0, int tn = -1,
              const double *psi = nullptr, long spsi = 0, const double *pbc = nullptr,
              long spbc = 0, double *zimg = nullptr, long szimg = 0, GBox zbox = GBox());
// q = Eᵀ fb (zero, scale, fixed-point scatter, conversion); scale: nbatch doubles of work
void launch_ET(const Dom &d, const RegDev &r, const double *fb, long sfb, double *q, long sq,
               int nbatch, const int *skip, double *scale, hipStream_t s);
// Eᵀ up to its fixed-point image (zeroing and the scales in one launch, then the scatter);
// launch_rot_image reads that image directly, converting as launch_ET's last pass would.
// With lrk > 0 the first launch also forms w[b] = Vᵀ fb[b] (V 2nb × lrk), the low-rank
// coarse-level part's input
// With a box, only its faces are cleared: every face Eᵀ writes lies in it (the box holds the
// points' supports), and the caller keeps the image zero outside it.
void launch_ET_image(const Dom &d, const RegDev &r, const double *fb, long sfb, double *img,
                     long sq, int nbatch, const int *skip, double *scale, hipStream_t s,
                     const double *V = nullptr, double *w = nullptr, int lrk = 0,
                     GBox box = GBox());
// the scatter of launch_ET_image alone: the image must be zero on the box and the scales set
// (the PCG direction kernel sets them; E of the previous product cleared the image)
void launch_ET_scatter_only(const Dom &d, const RegDev &r, const double *fb, long sfb,
                            double *img, long sq, int nbatch, const int *skip,
                            const double *scale, hipStream_t s);
// y = A x with A stored transposed (AT column-major: A's rows contiguous), one workgroup per row
void launch_gemv_t(int n, const double *AT, const double *x, double *y, const int *skip,
                   hipStream_t s);
void launch_axpy_sub(int n, double *y, const double *v, double a, hipStream_t s);  // y -= a v

void launch_stress(int nbodies, const int *boff, const int *npts, int nb, const double *Fb,
                   const double *ds, double h, double dt, double *CD, double *CL, double *fbs,
                   hipStream_t s);

// CG state lives on the device; see body.hip.  The `done` word: 0 running, otherwise why it
// stopped (every nonzero value makes the B-products of later iterations skip).
constexpr int CG_CONVERGED = 1, CG_MAXITER = 2, CG_NONFINITE = 3;
struct CGDev {
    double *x, *b, *r, *c, *u;  // 2nb each
    double *z;                  // preconditioned residual (PCG)
    const double *P;            // dense preconditioner n×n (PCG), stored transposed, or null
    double *sc;                 // [residual, prev_residual, tol, ρ]
    double *part;               // z_i·r_i of the PCG step (pcg_parts(n))
    double *etscale;            // PCG: the Eᵀ fixed-point scale of u is set here (or null)
    int *it, *done;
    int n, maxiter;
    double reltol;
};
void launch_cg_init(const CGDev &cg, hipStream_t s);        // after c = B·x: r = b - c, ...
void launch_cg_dir(const CGDev &cg, hipStream_t s);         // u = r + β u   (CG)
void launch_pcg_dir(const CGDev &cg, hipStream_t s);        // z = P r; u = z + β u (PCG)
int pcg_parts(int n);                                        // doubles of CGDev::part
// in-place transpose of an n×n matrix (the PCG preconditioner, read by rows)
void launch_transpose_inplace(int n, double *A, hipStream_t s);
void launch_cg_update(const CGDev &cg, hipStream_t s);      // after c = B·u
// mirror fresh columns' transposes into the stale rows of an assembled n×n matrix
void launch_mirror_rows(int n, double *Bm, const int *cols, int ncols, const int *is_col,
                        hipStream_t s);
void launch_copy(double *dst, const double *src, long n, hipStream_t s);

// dense inverse (Gauss–Jordan with partial pivoting) of an n×n column-major matrix
void launch_gj_step(int n, double *A, double *Inv, int k, int *piv_row, int *err,
                    hipStream_t s);

// ---------------------------------------------------------------------------- grouped launches
// One launch runs a list of independent problems (the cases of a batch, each with its own
// grid size, level count, body points and buffers).  Problem i owns the workgroups
// [off[i], off[i+1]) of a 1-D grid of off[n] workgroups; args and off live in device memory.
struct Grp {
    const void *args;
    const int *off;
    int n, blocks;
};

// The problem p with off[p] <= b < off[p + 1] (off[0] = 0, non-decreasing; a problem without
// blocks repeats its offset): the number of k in [1, n) with off[k] <= b.  Each wave reads 64
// offsets per load and counts them with a ballot, one memory round trip per 64 problems
// instead of a chain of dependent loads.  Called by whole waves (kernel entry).
__device__ inline int grp_find(const int *off, int n, int b) {
    const int lane = threadIdx.x & 63;
    int cnt = 0;
    for (int k0 = 1; k0 < n; k0 += 64) {
        const int k = k0 + lane;
        const bool le = k < n && off[k] <= b;
        cnt += __popcll(__ballot(le));
    }
    return cnt;
}

// Per-problem arguments; nbatch entries per problem with the given element strides (the same
// meaning as the batched single-case launchers above).  Every grouped stencil kernel uses
// 256-thread workgroups; blocks(problem) = ceil(elements / 256) · nbatch.
struct CurlA { Dom d; const double *psi; long spsi; const double *pbc; long sbc; double *q; long sq; int nbatch; };
struct RotA { Dom d; const double *q; long sq; double *g; long sg; int nbatch; };
// rot of Eᵀ's fixed-point image (img, as written by the ET stages, scale one double)
struct RotIA { Dom d; const double *img; const double *scale; double *g; int nbatch; GBox box; };
struct CoarsA { Dom d; double *gc; const double *g; long sgc, sg; int nbatch; };
struct GetBcA { Dom d; const double *r; long sr; double *rbc; long sbc; int nbatch; };
struct CopyBcA { Dom d; const double *in; long sin; double *out; long sout; const double *rbc; long sbc; double fac; int nbatch; };
// get_bc! of r into rbc and out = in + edges(rbc) (one batch entry); blocks grp_blocks(nΓ + nbc)
struct CopyGetBcA { Dom d; const double *in; double *out; const double *r; double *rbc; int nbatch; };
// the nonlinear term's output tiles per level (NL_TX × NL_TY outputs each)
constexpr int NL_TX = 64, NL_TY = 4;
__host__ __device__ inline long nl_tiles(const Dom &d) {
    return (long)((d.m + NL_TX - 1) / NL_TX) * ((d.p + NL_TY - 1) / NL_TY);
}
// psi set: the flux is curl!(ψ, ψbc) formed while staging (q unread), else q is read
struct NonlinA {
    Dom d; const double *q; long sq; const double *g; long sg; const double *gbc; long sbc;
    double gbc_scale; double *out; long sout; double h; int lev0; const double *uv; int nbatch;
    const double *psi; long spsi; const double *pbc; long spbc;
};
struct BcPrefA { int nbc; const double *bcraw; double *acc; const double *fac; int L; };
struct RhsA {
    Dom d; const double *g, *N1, *N2, *bc; double *out; double h, dt, Re, f1, f2; int nbatch;
};
struct ResidA { Dom d; const double *x, *rhs; double a, ar; double *r; int nbatch; };
struct FillA { double *p; long n; double v; };
struct EA { Dom d; RegDev r; const double *q; double *fb; const double *uv; double uvscale; };
struct ETA { Dom d; RegDev r; const double *fb; double *q; double *scale; GBox box; };
struct GemvA { int n; const double *A; const double *x; double *y; };  // A transposed
struct StressA {
    int nbodies; const int *boff, *npts; int nb; const double *Fb, *ds; double h, dt;
    double *CD, *CL, *fbs;
};
// The Poisson solve of the finest level from a kept spectrum and the coarse levels' ψ on
// get_bc!'s lines (grouped forms of launch_edge_gemv, launch_spectral_scale and launch_lines,
// one batch entry per problem; W of EdgeA holds 2(m + p), of LinesA lines_work_size(d)
// doubles), and y -= a·x.
struct EdgeA { Dom d; const double *Sm, *Sp, *bc; double *W; };
// sub (or null): the spectrum read is A − sub (the trial's minus the projection's)
struct SpecA { Dom d; const double *A, *D; double kappa; const double *Sm, *Sp, *EW; double *F; const double *sub; };
struct LinesA { Dom d; const double *F, *Sm, *Sp, *EW, *D; double *W, *psi; };
struct AxpyA { long n; double *y; const double *x; double a; };
// workgroups per problem of the three lines passes
long grp_lines_blocks(const Dom &d, int pass);

// element counts per batch entry of each grouped op (host side, for the block offsets)
__host__ __device__ inline long grp_elems_coarsify(const Dom &d) {
    return (long)(d.nx / 2 - 1) * (d.ny / 2 - 1);
}
static constexpr int GRP_BS = 256;
// heavier grouped elementwise kernels take GRP_EPT elements per thread (at stride GRP_BS, so
// each pass is coalesced): one problem lookup per GRP_EPT·GRP_BS elements
static constexpr int GRP_EPT = 4;
__host__ __device__ inline long grp_blocks(long elems) {
    return (elems + (long)GRP_BS * GRP_EPT - 1) / ((long)GRP_BS * GRP_EPT);
}

// sum over the 64 lanes of a wave, in a fixed tree order; the total lands in lane 0
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// one 256-thread workgroup: *out = V_j·f over n2 entries in a fixed order (the low-rank dots
// w = (E·V)ᵀ z of a B-product; the Eᵀ preparation and rot's extra workgroups both use it)
__device__ inline void lr_dot256(const double *__restrict__ Vj, const double *__restrict__ f,
                                 int n2, double *out) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n2; i += 256) s += Vj[i] * f[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}
static constexpr int GRP_NS = 13;  // delta support per dimension

void launch_grp_curl(const Grp &g, hipStream_t s);
void launch_grp_rot(const Grp &g, hipStream_t s);
void launch_grp_rot_image(const Grp &g, hipStream_t s);    // RotIA; blocks: ceil(nΓ/256)
void launch_grp_coarsify(const Grp &g, hipStream_t s);
void launch_grp_get_bc(const Grp &g, hipStream_t s);
void launch_grp_copy_apply_bc(const Grp &g, hipStream_t s);
void launch_grp_copy_getbc(const Grp &g, hipStream_t s);
void launch_grp_nonlinear(const Grp &g, hipStream_t s);
void launch_grp_bc_prefix(const Grp &g, hipStream_t s);   // blocks: ceil(nbc/256)
void launch_grp_rhs(const Grp &g, hipStream_t s);         // blocks: ceil(nΓ/256)·levels
void launch_grp_resid(const Grp &g, hipStream_t s);       // blocks: ceil(nΓ/256)
void launch_grp_fill(const Grp &g, hipStream_t s);
void launch_grp_edge_gemv(const Grp &g, hipStream_t s);
void launch_grp_spectral_scale(const Grp &g, hipStream_t s);
void launch_grp_lines_part(const Grp &g, hipStream_t s);
void launch_grp_lines_colsum(const Grp &g, hipStream_t s);
void launch_grp_lines_out(const Grp &g, hipStream_t s);
void launch_grp_axpy_sub(const Grp &g, hipStream_t s);        // blocks: ceil(n/256)
void launch_grp_E(const Grp &g, int sched, hipStream_t s);  // blocks: grp_E_blocks
void launch_grp_ET_zero(const Grp &g, hipStream_t s);     // ETA; blocks: ceil(nq/256)
void launch_grp_ET_scale(const Grp &g, hipStream_t s);    // ETA; blocks: 1
void launch_grp_ET_zero_scale(const Grp &g, hipStream_t s); // ETA; blocks: ceil(nq/256) + 1
void launch_grp_ET(const Grp &g, int sched, hipStream_t s);  // ETA; blocks: grp_ET_blocks
long grp_E_blocks(int nb, int sched);
long grp_ET_blocks(int nb, int sched);
void launch_grp_ET_fix(const Grp &g, hipStream_t s);      // ETA; blocks: ceil(nq/256)
void launch_grp_gemv_t(const Grp &g, hipStream_t s);      // blocks: n (one per row)
void launch_grp_stress(const Grp &g, hipStream_t s);      // blocks: nbodies

// grouped even/odd split contraction: every problem's GemmArgs in device memory; problem i
// owns tiles [off[i], off[i+1]) with tiles = ceil(rows/BM)·ceil(cols/BN)·batch for the
// variant's tile shape (grouped_eo_tiles).
void launch_gemm_eo_grp(const Grp &g, int prec, int variant, int side, int epi, hipStream_t s,
                        unsigned long long *ts = nullptr);
int gemm_eo_grp_variant_count(int prec);
// whether grouped variant v stages its tiles by LDS-direct loads (32-bit offsets: dma_fits)
bool gemm_eo_grp_variant_dma(int prec, int v);
// the LDS-direct split tiles address the sine parts and one batch entry's data operand by
// 32-bit byte offsets; false if either exceeds them (or the interleaved parts are missing)
bool dma_fits(const GemmArgs &g, int side, int prec);
const char *gemm_eo_grp_variant_name(int prec, int v);
long grouped_eo_tiles(const GemmArgs &g, int prec, int variant, int side);

}  // namespace ibh
