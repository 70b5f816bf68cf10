// Body coupling: Yang3 regularized delta (structure-ops.jl), interpolation E, spreading Eᵀ,
// surface-force solvers (dense inverse GEMV and CG, cnab/ops.jl:79-114) and the per-body
// force coefficients (streamfcn/states.jl:53-84).  Reductions are FP64.
#include <cstdio>

#include "ibh_internal.h"

namespace ibh {

static constexpr int SUPP = 6, NS = 13;

__host__ __device__ static inline unsigned nblk(long n, int bs) {
    return (unsigned)((n + bs - 1) / bs);
}

// δh, structure-ops.jl:114-171 — Yang et al. (JCP 2009) smoothed delta times h; the
// coefficients are the reference's literals.
__device__ double deltah(double rf, double rb, double dr) {
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

// Reg update!, structure-ops.jl:37-61.  The 13×13 weight of point k is the product of two
// 1-D factors, stored separately: w[k][0][s] = δh(x_s), w[k][1][t] = δh(y_t + h/2) for the
// x-flux, w[k][2][s] = δh(x_s + h/2), w[k][3][t] = δh(y_t) for the y-flux.
__global__ void k_reg_update(Dom d, RegDev r, const double *__restrict__ pts, double h, double x0,
                             double y0) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= r.nb * 4 * NS) return;
    const int k = t / (4 * NS), c = (t / NS) % 4, s = t % NS;
    const double xb = pts[k], yb = pts[r.nb + k];
    // non-finite or far-off points get an index that support_inside rejects
    const double fx = floor((xb - x0) / h), fy = floor((yb - y0) / h);
    const int bx = fabs(fx) < 1e9 ? (int)fx : -(1 << 30);
    const int by = fabs(fy) < 1e9 ? (int)fy : -(1 << 30);
    double v;
    if (c == 0) v = deltah(h * (double)(bx - 1 + (s - SUPP)) + x0, xb, h);
    else if (c == 1) v = deltah(h * (double)(by - 1 + (s - SUPP)) + y0 + h / 2, yb, h);
    else if (c == 2) v = deltah(h * (double)(bx - 1 + (s - SUPP)) + x0 + h / 2, xb, h);
    else v = deltah(h * (double)(by - 1 + (s - SUPP)) + y0, yb, h);
    r.w[t] = v;
    if (c == 0 && s == 0) {
        r.bidx[k] = bx;
        r.bidx[r.nb + k] = by;
        // E and Eᵀ skip such a point (see support_inside), so nothing outside q is touched
        if (!support_inside(d, bx, by)) atomicOr(r.err, 1);
    }
}

void launch_reg_update(const Dom &d, const RegDev &r, const double *pts, double h, double x0,
                       double y0, hipStream_t s) {
    if (r.nb == 0) return;
    hipLaunchKernelGGL(k_reg_update, dim3(nblk((long)r.nb * 4 * NS, 256)), dim3(256), 0, s, d, r,
                       pts, h, x0, y0);
}

// ---- block schedules (see e_sched_name): E's waves per workgroup; Eᵀ's workgroup size and
// entry order
constexpr int E_WAVES[] = {4, 1, 2, 8, 16};
constexpr int E_NSCHED = sizeof(E_WAVES) / sizeof(E_WAVES[0]);
struct ETSched {
    int bs;
    bool pm;  // point-major
};
constexpr ETSched ET_SCHED[] = {{256, false}, {64, false},  {128, false}, {512, false},
                                {1024, false}, {64, true},  {128, true},  {256, true},
                                {512, true},  {1024, true}};
constexpr int ET_NSCHED = sizeof(ET_SCHED) / sizeof(ET_SCHED[0]);

int e_sched_count() { return E_NSCHED; }
int et_sched_count() { return ET_NSCHED; }
const char *e_sched_name(int v) {
    static const char *names[] = {"4 waves/workgroup", "1 wave/workgroup", "2 waves/workgroup",
                                  "8 waves/workgroup", "16 waves/workgroup"};
    return v >= 0 && v < E_NSCHED ? names[v] : "?";
}
const char *et_sched_name(int v) {
    static const char *names[] = {"256 threads entry-major", "64 threads entry-major",
                                  "128 threads entry-major", "512 threads entry-major",
                                  "1024 threads entry-major", "64 threads point-major",
                                  "128 threads point-major", "256 threads point-major",
                                  "512 threads point-major", "1024 threads point-major"};
    return v >= 0 && v < ET_NSCHED ? names[v] : "?";
}
void parse_reg_sched(const char *s, int *e, int *et) {
    int a = -1, b = -1;
    if (s && std::sscanf(s, "%d,%d", &a, &b) >= 1) {
        *e = a;
        *et = b;
    }
}

static inline int et_sched_index(int v) { return v >= 0 && v < ET_NSCHED ? v : 0; }
static inline int e_sched_index(int v) { return v >= 0 && v < E_NSCHED ? v : 0; }

// stencil entry (point-component kc, then tt, s) of thread t: entry-major walks the 169 entries
// of one point-component together, point-major one entry of the n2 point-components together
__device__ inline int et_entry(int t, int n2, bool pm) {
    return pm ? (t % n2) * (NS * NS) + t / n2 : t;
}

// E (RegT), structure-ops.jl:88-105: fb[k] = Σ_t Σ_s q[i, j]·w over the 13×13 stencil of point
// k, with q0 (constant per component) added to q first, as in Q = qs + q0.  One wave per output:
// lane l sums the entries l, l + 64, l + 128 (column-major, s fastest), then a fixed-order wave
// sum; the total is returned in lane 0.  (The reference's sumprod adds the 169 terms in one
// column-major sweep; the two orders differ by roundoff.)
// The flux values E reads: stored (QStored) or curl!'s values formed from ψ on the fly (QCurl).
struct QStored {
    const double *q;
    __device__ double operator()(long t) const { return q[t]; }
};
struct QCurl {
    Dom d;
    const double *psi, *pbc;
    __device__ double operator()(long t) const { return curl_at(d, psi, pbc, t); }
};

template <class QA>
__device__ inline double E_wave(const Dom &d, const RegDev &r, const QA &qa, double u0, double v0,
                                int t, int lane) {
    const int comp = t / r.nb, k = t % r.nb;
    const int bx = r.bidx[k], by = r.bidx[r.nb + k];
    if (!support_inside(d, bx, by)) {  // uniform over the wave; the host reports the flag
        if (lane == 0) atomicOr(r.err, 1);
        return 0.0;
    }
    const double *wx = r.w + ((long)k * 4 + 2 * comp) * NS;
    const double *wy = wx + NS;
    const long off = comp == 0 ? -1 : d.nu - 1;  // flux index of (i, j) = off + ld·(j−1) + i
    const long ld = comp == 0 ? d.nx + 1 : d.nx;
    const double q0 = comp == 0 ? u0 : v0;
    double acc = 0.0;
    for (int e = lane; e < NS * NS; e += 64) {
        const int tt = e / NS, s = e % NS;
        const int j = by + tt - SUPP;
        acc += (qa(off + ld * (j - 1) + bx + s - SUPP) + q0) * (wx[s] * wy[tt]);
    }
    return wave_sum(acc);
}

// With lrk > 0 the wave also adds the low-rank term Σ_j U[t, j]·w[b, j] (U 2nb × lrk column-
// major, w lrk per batch entry), summed over the lanes in a fixed order.
// W waves (outputs) per workgroup.
template <int W>
__global__ __launch_bounds__(64 * W) void k_E(Dom d, RegDev r, const double *__restrict__ q,
                                              long sq, double u0, double v0,
                                              double *__restrict__ fb, long sfb, const int *skip,
                                              const double *uv, double uvscale,
                                              const double *__restrict__ U,
                                              const double *__restrict__ w, int lrk, int t0,
                                              int tn, const double *__restrict__ psi, long spsi,
                                              const double *__restrict__ pbc, long spbc,
                                              double *__restrict__ zimg, long szimg, GBox zbox) {
    if (skip && *skip) return;
    const int ne = (tn + W - 1) / W;  // E's workgroups; those after them clear zimg on zbox
    if ((int)blockIdx.x >= ne) {
        const long n = gbox_faces(d, zbox);
        for (long t = (blockIdx.x - ne) * (long)blockDim.x + threadIdx.x; t < n;
             t += (gridDim.x - ne) * (long)blockDim.x)
            zimg[blockIdx.y * szimg + gbox_face(d, zbox, t)] = 0.0;
        return;
    }
    if (uv) {  // q0 = freestream·h from the device parameter block
        u0 = uv[0] * uvscale;
        v0 = uv[1] * uvscale;
    }
    const int tl = blockIdx.x * W + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (tl >= tn) return;
    const int t = t0 + tl;
    const int b = blockIdx.y;
    double v = psi ? E_wave(d, r, QCurl{d, psi + b * spsi, pbc + b * spbc}, u0, v0, t, lane)
                   : E_wave(d, r, QStored{q + b * sq}, u0, v0, t, lane);
    if (lrk > 0) {
        const long n2 = 2L * r.nb;
        double s = 0.0;
        for (int j = lane; j < lrk; j += 64) s += U[t + n2 * j] * w[(long)b * lrk + j];
        s = wave_sum(s);
        v += s;
    }
    if (lane == 0) fb[b * sfb + t] = v;
}

void launch_E(const Dom &d, const RegDev &r, const double *q, long sq, double u0, double v0,
              double *fb, long sfb, int nbatch, const int *skip, hipStream_t s, const double *uv,
              double uvscale, const double *U, const double *w, int lrk, int t0, int tn,
              const double *psi, long spsi, const double *pbc, long spbc, double *zimg,
              long szimg, GBox zbox) {
    if (r.nb == 0) return;
    if (tn < 0) tn = 2 * r.nb - t0;
    if (tn <= 0) return;
    // zimg: extra workgroups clear its faces on zbox (512 of them at most, striding)
    const long zf = zimg ? gbox_faces(d, zbox) : 0;
#define LAUNCH_E(W)                                                                           \
    hipLaunchKernelGGL(k_E<W>, dim3(nblk(tn, W) + (unsigned)std::min<long>(512, nblk(zf, 64 * W)), \
                       nbatch), dim3(64 * W), 0, s, d, r, q, sq, u0, v0, fb, sfb, skip, uv,      \
                       uvscale, U, w, lrk, t0, tn, psi, spsi, pbc, spbc, zimg, szimg, zbox)
    switch (E_WAVES[(r.esched >= 0 && r.esched < E_NSCHED) ? r.esched : 0]) {
    case 1: LAUNCH_E(1); break;
    case 2: LAUNCH_E(2); break;
    case 8: LAUNCH_E(8); break;
    case 16: LAUNCH_E(16); break;
    default: LAUNCH_E(4); break;
    }
#undef LAUNCH_E
}

// Eᵀ (Reg), structure-ops.jl:63-86: q .= 0 then scatter-add the 13×13 stencils of every point.
// One thread per stencil entry.  The adds are 64-bit integer atomics on a fixed-point image of
// q, so the sum does not depend on the order in which the adds land: runs are bitwise
// reproducible (the order of floating-point atomics is not, and low-precision transforms
// downstream amplify such last-bit differences).  Per column, the scale 2^(50-e) with
// 2^e ≤ max|f| < 2^(e+1) keeps 2048 full-weight contributions of the largest force below
// 2^62; each contribution is rounded to 2^-51·max|f|, the size of an FP64 rounding at the top
// of the range.  A last pass converts q back to FP64 in place.
__global__ void k_zero(double *__restrict__ p, long n, long sp, const int *skip) {
    if (skip && *skip) return;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < n) p[blockIdx.y * sp + t] = 0.0;
}

// one 256-thread workgroup per column: scale[b] = 2^(50-e) for the column's max|f|
__device__ inline void et_scale_body(const double *__restrict__ fb, int n2, double *scale) {
    __shared__ double sh[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) m = fmax(m, fabs(fb[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
        *scale = (m > 0.0 && isfinite(m)) ? ldexp(1.0, 50 - ilogb(m)) : 1.0;
    }
}

__global__ __launch_bounds__(256) void k_et_scale(const double *__restrict__ fb, long sfb,
                                                  int n2, double *scale, const int *skip) {
    if (skip && *skip) return;
    et_scale_body(fb + blockIdx.x * sfb, n2, scale + blockIdx.x);
}

__device__ inline void ET_at(const Dom &d, const RegDev &r, const double *__restrict__ fb,
                             double *__restrict__ q, double sc, int t) {
    const int s = t % NS, tt = (t / NS) % NS, kc = t / (NS * NS);
    const int comp = kc / r.nb, k = kc % r.nb;
    const int bx = r.bidx[k], by = r.bidx[r.nb + k];
    if (!support_inside(d, bx, by)) {
        atomicOr(r.err, 1);
        return;
    }
    const double *wx = r.w + ((long)k * 4 + 2 * comp) * NS;
    const double f = fb[comp * r.nb + k];
    const double v = (wx[s] * wx[NS + tt]) * f;
    const int i = bx + s - SUPP, j = by + tt - SUPP;
    double *dst = comp == 0 ? q + (i - 1) + (long)(d.nx + 1) * (j - 1)
                            : q + d.nu + (i - 1) + (long)d.nx * (j - 1);
    if (!isfinite(v)) {
        atomicOr(r.err, 2);
        return;
    }
    atomicAdd(reinterpret_cast<unsigned long long *>(dst), (unsigned long long)llrint(v * sc));
}

template <int BS, bool PM>
__global__ __launch_bounds__(BS) void k_ET(Dom d, RegDev r, const double *__restrict__ fb,
                                           long sfb, double *__restrict__ q, long sq,
                                           const double *__restrict__ scale, const int *skip) {
    if (skip && *skip) return;
    const int t = blockIdx.x * BS + threadIdx.x;
    if (t >= 2 * r.nb * NS * NS) return;
    const int b = blockIdx.y;
    ET_at(d, r, fb + b * sfb, q + b * sq, scale[b], et_entry(t, 2 * r.nb, PM));
}

// the scatter-add of Eᵀ (the image zeroed and the scales set before) in r's schedule
static void launch_ET_scatter(const Dom &d, const RegDev &r, const double *fb, long sfb,
                              double *q, long sq, int nbatch, const int *skip,
                              const double *scale, hipStream_t s) {
    const long n = 2L * r.nb * NS * NS;
    const ETSched E = ET_SCHED[et_sched_index(r.etsched)];
#define LAUNCH_ET(BS, PM)                                                                     \
    hipLaunchKernelGGL((k_ET<BS, PM>), dim3(nblk(n, BS), nbatch), dim3(BS), 0, s, d, r, fb,   \
                       sfb, q, sq, scale, skip)
    if (E.pm) {
        switch (E.bs) {
        case 64: LAUNCH_ET(64, true); break;
        case 128: LAUNCH_ET(128, true); break;
        case 512: LAUNCH_ET(512, true); break;
        case 1024: LAUNCH_ET(1024, true); break;
        default: LAUNCH_ET(256, true); break;
        }
    } else {
        switch (E.bs) {
        case 64: LAUNCH_ET(64, false); break;
        case 128: LAUNCH_ET(128, false); break;
        case 512: LAUNCH_ET(512, false); break;
        case 1024: LAUNCH_ET(1024, false); break;
        default: LAUNCH_ET(256, false); break;
        }
    }
#undef LAUNCH_ET
}

// fixed point → FP64 in place (exact power-of-two rescale after the int64 → double rounding)
__device__ inline void et_fix_at(double *q, double sc, long t) {
    const long long iv = reinterpret_cast<const long long *>(q)[t];
    q[t] = (double)iv * (1.0 / sc);
}

__global__ void k_et_fix(double *__restrict__ q, long n, long sq, const double *__restrict__ scale,
                         const int *skip) {
    if (skip && *skip) return;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= n) return;
    et_fix_at(q + blockIdx.y * sq, scale[blockIdx.y], t);
}

// zeroing of the fixed-point image (the first ceil(n/256) blocks), the column's scale (the
// next one) and, with lrk > 0, the low-rank dots w[b, j] = V[:, j]·fb[b] (one block per j;
// V n2 × lrk column-major)
__global__ __launch_bounds__(256) void k_et_zero_scale(Dom d, GBox box, double *__restrict__ q,
                                                       long sq, const double *__restrict__ fb,
                                                       long sfb, int n2, double *scale,
                                                       const int *skip,
                                                       const double *__restrict__ V,
                                                       double *__restrict__ w, int lrk) {
    if (skip && *skip) return;
    const int b = blockIdx.y;
    const long n = gbox_faces(d, box);
    const long nz = nblk(n, 256);
    if (blockIdx.x == nz) {
        et_scale_body(fb + b * sfb, n2, scale + b);
        return;
    }
    if (blockIdx.x > nz) {
        const int j = blockIdx.x - (int)nz - 1;
        lr_dot256(V + (long)n2 * j, fb + b * sfb, n2, w + (long)b * lrk + j);
        return;
    }
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < n) q[b * sq + gbox_face(d, box, t)] = 0.0;
}

void launch_ET_image(const Dom &d, const RegDev &r, const double *fb, long sfb, double *img,
                     long sq, int nbatch, const int *skip, double *scale, hipStream_t s,
                     const double *V, double *w, int lrk, GBox box) {
    hipLaunchKernelGGL(k_et_zero_scale, dim3(nblk(gbox_faces(d, box), 256) + 1 + lrk, nbatch),
                       dim3(256), 0, s, d, box, img, sq, fb, sfb, 2 * r.nb, scale, skip, V, w, lrk);
    if (r.nb == 0) return;
    launch_ET_scatter(d, r, fb, sfb, img, sq, nbatch, skip, scale, s);
}

void launch_ET_scatter_only(const Dom &d, const RegDev &r, const double *fb, long sfb,
                            double *img, long sq, int nbatch, const int *skip,
                            const double *scale, hipStream_t s) {
    if (r.nb == 0) return;
    launch_ET_scatter(d, r, fb, sfb, img, sq, nbatch, skip, scale, s);
}

void launch_ET(const Dom &d, const RegDev &r, const double *fb, long sfb, double *q, long sq,
               int nbatch, const int *skip, double *scale, hipStream_t s) {
    hipLaunchKernelGGL(k_zero, dim3(nblk(d.nq, 256), nbatch), dim3(256), 0, s, q, d.nq, sq, skip);
    if (r.nb == 0) return;
    hipLaunchKernelGGL(k_et_scale, dim3(nbatch), dim3(256), 0, s, fb, sfb, 2 * r.nb, scale, skip);
    launch_ET_scatter(d, r, fb, sfb, q, sq, nbatch, skip, scale, s);
    hipLaunchKernelGGL(k_et_fix, dim3(nblk(d.nq, 256), nbatch), dim3(256), 0, s, q, d.nq, sq,
                       (const double *)scale, skip);
}

__global__ void k_axpy_sub(int n, double *y, const double *v, double a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] -= v[i] * a;
}

void launch_axpy_sub(int n, double *y, const double *v, double a, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_axpy_sub, dim3(nblk(n, 256)), dim3(256), 0, s, n, y, v, a);
}

// ---------------------------------------------------------------- block reductions (FP64)
template <int BS>
__device__ double block_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int k = 0; k < BS / 64; ++k) t += sh[k];
        sh[BS / 64] = t;
    }
    __syncthreads();
    t = sh[BS / 64];
    __syncthreads();
    return t;
}

// y = A·x for a dense n×n matrix stored transposed (AT column-major, i.e. A by rows): the Binv
// GEMV of cnab/ops.jl:97 and the PCG preconditioner's.  One workgroup per output reads its row
// contiguously; four independent partial sums per thread keep the loads in flight, and a
// fixed-order block sum finishes the row (deterministic).
constexpr int GEMV_BS = 256;

__device__ inline double gemv_row(int n, const double *__restrict__ AT,
                                  const double *__restrict__ x, int i, double *sh) {
    const double *__restrict__ row = AT + (long)n * i;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = threadIdx.x;
    for (; k + 3 * GEMV_BS < n; k += 4 * GEMV_BS) {
        a0 += row[k] * x[k];
        a1 += row[k + GEMV_BS] * x[k + GEMV_BS];
        a2 += row[k + 2 * GEMV_BS] * x[k + 2 * GEMV_BS];
        a3 += row[k + 3 * GEMV_BS] * x[k + 3 * GEMV_BS];
    }
    for (; k < n; k += GEMV_BS) a0 += row[k] * x[k];
    return block_sum<GEMV_BS>((a0 + a1) + (a2 + a3), sh);
}

__global__ __launch_bounds__(GEMV_BS) void k_gemv_t(int n, const double *__restrict__ AT,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ y, const int *skip) {
    __shared__ double sh[GEMV_BS / 64 + 1];
    if (skip && *skip) return;
    const double v = gemv_row(n, AT, x, blockIdx.x, sh);
    if (threadIdx.x == 0) y[blockIdx.x] = v;
}

void launch_gemv_t(int n, const double *AT, const double *x, double *y, const int *skip,
                   hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_gemv_t, dim3(n), dim3(GEMV_BS), 0, s, n, AT, x, y, skip);
}

// update_stress!, streamfcn/states.jl:53-84: per body fb = F·h/dt, CD = 2Σfb_x, CL = 2Σfb_y,
// then fb ./= ds.  One block per body.
__device__ inline void stress_body(const int *boff, const int *npts, int nb,
                                   const double *__restrict__ Fb, const double *__restrict__ ds,
                                   double h, double dt, double *CD, double *CL, double *fbs,
                                   int b) {
    __shared__ double sh[8];
    const int o = boff[b], n = npts[b];
    double sx = 0.0, sy = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double fx = Fb[o + k] * h / dt, fy = Fb[nb + o + k] * h / dt;
        sx += fx;
        sy += fy;
        fbs[o + k] = fx / ds[o + k];
        fbs[nb + o + k] = fy / ds[o + k];
    }
    sx = block_sum<256>(sx, sh);
    sy = block_sum<256>(sy, sh);
    if (threadIdx.x == 0) {
        CD[b] = 2 * sx;
        CL[b] = 2 * sy;
    }
}

__global__ void k_stress(const int *boff, const int *npts, int nb, const double *__restrict__ Fb,
                         const double *__restrict__ ds, double h, double dt, double *CD,
                         double *CL, double *fbs) {
    stress_body(boff, npts, nb, Fb, ds, h, dt, CD, CL, fbs, blockIdx.x);
}

void launch_stress(int nbodies, const int *boff, const int *npts, int nb, const double *Fb,
                   const double *ds, double h, double dt, double *CD, double *CL, double *fbs,
                   hipStream_t s) {
    if (nbodies == 0) return;
    hipLaunchKernelGGL(k_stress, dim3(nbodies), dim3(256), 0, s, boff, npts, nb, Fb, ds, h, dt,
                       CD, CL, fbs);
}

// ---------------------------------------------------------------- CG (IterativeSolvers 0.9.2)
// cg!(x, A, b; reltol, maxiter) with the warm start x: r = b - A x, tol = reltol·‖r‖; each
// iteration u = r + (‖r‖²/‖r_prev‖²) u, c = A u, α = ‖r‖²/(u·c), x += α u, r -= α c.  The stop
// test (it ≥ maxiter or ‖r‖ ≤ tol) runs before every iteration.  One block does the vector
// work; the device flag `done` makes the surrounding B-products skip once converged.
// PCG direction (IterativeSolvers 0.9.2 PCGIterable): z = P r; ρ_prev = ρ; ρ = z·r;
// u = z + (ρ/ρ_prev) u, skipped once converged.  P is stored transposed: z_i comes from
// gemv_row, whose workgroup also leaves z_i·r_i; one workgroup then sums those, forms β and
// updates u.  Every sum runs in a fixed order.

int pcg_parts(int n) { return n; }

__global__ __launch_bounds__(GEMV_BS) void k_pcg_gemv(CGDev cg) {
    __shared__ double sh[GEMV_BS / 64 + 1];
    if (*cg.done) return;
    const int i = blockIdx.x;
    const double z = gemv_row(cg.n, cg.P, cg.r, i, sh);
    if (threadIdx.x == 0) {
        cg.z[i] = z;
        cg.part[i] = z * cg.r[i];
    }
}

__global__ void k_pcg_dir(CGDev cg) {
    __shared__ double sh[17];
    if (*cg.done) return;
    const int n = cg.n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += cg.part[i];
    s = block_sum<1024>(s, sh);
    const double beta = s / cg.sc[3];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double u = cg.z[i] + beta * cg.u[i];
        cg.u[i] = u;
        m = fmax(m, fabs(u));
    }
    __syncthreads();
    if (threadIdx.x == 0) cg.sc[3] = s;
    if (cg.etscale) {  // the Eᵀ scale of u for the B-product that follows (as et_scale_body)
        __shared__ double mx[16];
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
        if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            m = 0.0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, mx[k]);
            *cg.etscale = (m > 0.0 && isfinite(m)) ? ldexp(1.0, 50 - ilogb(m)) : 1.0;
        }
    }
}

void launch_pcg_dir(const CGDev &cg, hipStream_t s) {
    if (cg.n == 0) return;
    hipLaunchKernelGGL(k_pcg_gemv, dim3(cg.n), dim3(GEMV_BS), 0, s, cg);
    hipLaunchKernelGGL(k_pcg_dir, dim3(1), dim3(1024), 0, s, cg);
}

__global__ void k_transpose_inplace(int n, double *A) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long i = t / n, j = t % n;
    if (i >= n || j <= i) return;
    const double a = A[i + n * j], b = A[j + n * i];
    A[i + n * j] = b;
    A[j + n * i] = a;
}

void launch_transpose_inplace(int n, double *A, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_transpose_inplace, dim3(nblk((long)n * n, 256)), dim3(256), 0, s, n, A);
}

__global__ void k_cg_init(CGDev cg) {
    __shared__ double sh[17];
    const int n = cg.n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double ri = cg.b[i] - cg.c[i];
        cg.r[i] = ri;
        cg.u[i] = 0.0;
        s += ri * ri;
    }
    s = block_sum<1024>(s, sh);
    if (threadIdx.x == 0) {
        const double res = sqrt(s);
        cg.sc[0] = res;
        cg.sc[1] = 1.0;
        cg.sc[2] = cg.reltol * res;
        cg.sc[3] = 1.0;
        *cg.it = 0;
        *cg.done = !isfinite(res)                             ? CG_NONFINITE
                   : res <= cg.reltol * res                   ? CG_CONVERGED
                   : cg.maxiter <= 0                          ? CG_MAXITER
                                                              : 0;
    }
}

__global__ void k_cg_dir(CGDev cg) {
    if (*cg.done) return;
    const double res = cg.sc[0], prev = cg.sc[1];
    const double beta = (res * res) / (prev * prev);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cg.n; i += gridDim.x * blockDim.x)
        cg.u[i] = cg.r[i] + beta * cg.u[i];
}

__global__ void k_cg_update(CGDev cg) {
    __shared__ double sh[17];
    if (*cg.done) return;
    const int n = cg.n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += cg.u[i] * cg.c[i];
    s = block_sum<1024>(s, sh);
    const double res = cg.sc[0];
    const double alpha = (cg.P ? cg.sc[3] : res * res) / s;
    double s2 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        cg.x[i] += alpha * cg.u[i];
        const double ri = cg.r[i] - alpha * cg.c[i];
        cg.r[i] = ri;
        s2 += ri * ri;
    }
    s2 = block_sum<1024>(s2, sh);
    if (threadIdx.x == 0) {
        const double nres = sqrt(s2);
        cg.sc[1] = res;
        cg.sc[0] = nres;
        const int it = *cg.it + 1;
        *cg.it = it;
        if (!isfinite(nres)) *cg.done = CG_NONFINITE;
        else if (nres <= cg.sc[2]) *cg.done = CG_CONVERGED;
        else if (it >= cg.maxiter) *cg.done = CG_MAXITER;
    }
}

void launch_cg_init(const CGDev &cg, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1024), 0, s, cg);
}
void launch_cg_dir(const CGDev &cg, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_dir, dim3(nblk(cg.n, 256)), dim3(256), 0, s, cg);
}
void launch_cg_update(const CGDev &cg, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_update, dim3(1), dim3(1024), 0, s, cg);
}

// ---------------------------------------------------------------- preconditioner assembly
// Bm[j, c] = Bm[c, j] for rows j in J (fresh columns) and columns c outside J: the rows of
// the moving unknowns are taken from the fresh columns by the near-symmetry of B.
__global__ void k_mirror_rows(int n, double *Bm, const int *J, int nJ, const int *inJ) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)n * nJ) return;
    const int c = (int)(t % n), j = J[t / n];
    if (inJ[c]) return;
    Bm[j + (long)n * c] = Bm[c + (long)n * j];
}

void launch_mirror_rows(int n, double *Bm, const int *cols, int ncols, const int *is_col,
                        hipStream_t s) {
    if (ncols == 0) return;
    hipLaunchKernelGGL(k_mirror_rows, dim3(nblk((long)n * ncols, 256)), dim3(256), 0, s, n, Bm,
                       cols, ncols, is_col);
}

__global__ void k_copy(double *dst, const double *src, long n) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

void launch_copy(double *dst, const double *src, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy, dim3(nblk(n, 256)), dim3(256), 0, s, dst, src, n);
}

// ---------------------------------------------------------------- dense inverse
// Gauss–Jordan elimination with partial pivoting on [A | Inv] (the role of LAPACK getrf/getri
// behind inv(Bmat), cnab/ops.jl:96).  Step k: pick the pivot row, swap and scale row k, then
// eliminate column k from every other row.  ws = [pivot value, multipliers (n)].
__global__ void k_gj_pivot(int n, const double *__restrict__ A, int k, int *piv, double *ws,
                           int *err) {
    __shared__ double bv[1024];
    __shared__ int bi[1024];
    double best = -1.0;
    int bidx = k;
    for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
        const double v = fabs(A[i + (long)n * k]);
        if (v > best) { best = v; bidx = i; }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = bidx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const double v2 = bv[threadIdx.x + o];
            const int i2 = bi[threadIdx.x + o];
            if (v2 > bv[threadIdx.x] || (v2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
                bv[threadIdx.x] = v2;
                bi[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *piv = bi[0];
        ws[0] = A[bi[0] + (long)n * k];
        if (bv[0] == 0.0) atomicOr(err, 4);
    }
}

// swap rows k and piv of both matrices, then scale row k by 1/pivot
__global__ void k_gj_swap_scale(int n, double *A, double *Inv, int k, const int *piv,
                                const double *ws) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 2 * n) return;
    double *M = j < n ? A : Inv;
    const int c = j < n ? j : j - n;
    const int p = *piv;
    const double pv = ws[0];
    double a = M[k + (long)n * c], b = M[p + (long)n * c];
    if (p != k) {
        M[p + (long)n * c] = a;
        a = b;
    }
    M[k + (long)n * c] = a / pv;
}

__global__ void k_gj_col(int n, const double *A, int k, double *ws) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ws[1 + i] = i == k ? 0.0 : A[i + (long)n * k];
}

__global__ void k_gj_elim(int n, double *A, double *Inv, int k, const double *ws) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int i = (int)(t % n);
    const long jj = t / n;
    if (jj >= 2L * n) return;
    const int j = (int)jj;
    double *M = j < n ? A : Inv;
    const int c = j < n ? j : j - n;
    if (j < n && c < k) return;  // already eliminated columns of A stay e_c
    const double f = ws[1 + i];
    if (f != 0.0) M[i + (long)n * c] -= f * M[k + (long)n * c];
}

void launch_gj_step(int n, double *A, double *Inv, int k, int *piv_row, int *err,
                    hipStream_t s) {
    double *ws = reinterpret_cast<double *>(piv_row + 2);  // caller sizes the workspace
    hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(1024), 0, s, n, A, k, piv_row, ws, err);
    hipLaunchKernelGGL(k_gj_swap_scale, dim3(nblk(2L * n, 256)), dim3(256), 0, s, n, A, Inv, k,
                       piv_row, ws);
    hipLaunchKernelGGL(k_gj_col, dim3(nblk(n, 256)), dim3(256), 0, s, n, A, k, ws);
    hipLaunchKernelGGL(k_gj_elim, dim3(nblk(2L * n * n, 256)), dim3(256), 0, s, n, A, Inv, k, ws);
}

// ------------------------------------------------------------------------ grouped launches
// E, Eᵀ, the dense-inverse GEMV and the stress sums over a list of cases (one problem per
// case; see Grp).  Same per-element bodies as the single-case kernels.
static_assert(GRP_NS == NS, "grouped constants");

template <int W>
__global__ __launch_bounds__(64 * W) void kg_E(const EA *args, const int *off, int n) {
    const int p = grp_find(off, n, blockIdx.x);
    const EA A = args[p];
    const int t = (blockIdx.x - off[p]) * W + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= 2 * A.r.nb) return;
    const double v = E_wave(A.d, A.r, QStored{A.q}, A.uv[0] * A.uvscale, A.uv[1] * A.uvscale, t,
                            lane);
    if (lane == 0) A.fb[t] = v;
}

__global__ __launch_bounds__(GRP_BS) void kg_ET_zero(const ETA *args, const int *off, int n) {
    const int p = grp_find(off, n, blockIdx.x);
    const ETA A = args[p];
    const long t = (long)(blockIdx.x - off[p]) * GRP_BS + threadIdx.x;
    if (t < A.d.nq) A.q[t] = 0.0;
}

// zeroing of the fixed-point image (the first ceil(nq/256) workgroups of a problem) and its
// scale (the last one)
__global__ __launch_bounds__(GRP_BS) void kg_ET_zero_scale(const ETA *args, const int *off,
                                                           int n) {
    const int p = grp_find(off, n, blockIdx.x);
    const ETA A = args[p];
    const long lb = blockIdx.x - off[p];
    const long nfaces = gbox_faces(A.d, A.box);  // the image is kept zero outside the box
    if (lb == (long)nblk(nfaces, GRP_BS)) {
        et_scale_body(A.fb, 2 * A.r.nb, A.scale);
        return;
    }
    const long t = lb * GRP_BS + threadIdx.x;
    if (t < nfaces) A.q[gbox_face(A.d, A.box, t)] = 0.0;
}

// one workgroup per problem
__global__ __launch_bounds__(GRP_BS) void kg_ET_scale(const ETA *args, const int *off, int n) {
    const int p = grp_find(off, n, blockIdx.x);
    const ETA A = args[p];
    et_scale_body(A.fb, 2 * A.r.nb, A.scale);
}

template <int BS, bool PM>
__global__ __launch_bounds__(BS) void kg_ET(const ETA *args, const int *off, int n) {
    const int p = grp_find(off, n, blockIdx.x);
    const ETA A = args[p];
    const int t = (blockIdx.x - off[p]) * BS + threadIdx.x;
    if (t >= 2 * A.r.nb * NS * NS) return;
    ET_at(A.d, A.r, A.fb, A.q, *A.scale, et_entry(t, 2 * A.r.nb, PM));
}

long grp_E_blocks(int nb, int sched) { return nblk(2L * nb, E_WAVES[e_sched_index(sched)]); }
long grp_ET_blocks(int nb, int sched) {
    return nblk(2L * nb * NS * NS, ET_SCHED[et_sched_index(sched)].bs);
}

void launch_grp_E(const Grp &g, int sched, hipStream_t s) {
    if (g.n == 0 || g.blocks == 0) return;
    const EA *a = static_cast<const EA *>(g.args);
#define LAUNCH_GE(W) hipLaunchKernelGGL(kg_E<W>, dim3(g.blocks), dim3(64 * W), 0, s, a, g.off, g.n)
    switch (E_WAVES[e_sched_index(sched)]) {
    case 1: LAUNCH_GE(1); break;
    case 2: LAUNCH_GE(2); break;
    case 8: LAUNCH_GE(8); break;
    case 16: LAUNCH_GE(16); break;
    default: LAUNCH_GE(4); break;
    }
#undef LAUNCH_GE
}

void launch_grp_ET(const Grp &g, int sched, hipStream_t s) {
    if (g.n == 0 || g.blocks == 0) return;
    const ETA *a = static_cast<const ETA *>(g.args);
    const ETSched E = ET_SCHED[et_sched_index(sched)];
#define LAUNCH_GET(BS, PM) \
    hipLaunchKernelGGL((kg_ET<BS, PM>), dim3(g.blocks), dim3(BS), 0, s, a, g.off, g.n)
    if (E.pm) {
        switch (E.bs) {
        case 64: LAUNCH_GET(64, true); break;
        case 128: LAUNCH_GET(128, true); break;
        case 512: LAUNCH_GET(512, true); break;
        case 1024: LAUNCH_GET(1024, true); break;
        default: LAUNCH_GET(256, true); break;
        }
    } else {
        switch (E.bs) {
        case 64: LAUNCH_GET(64, false); break;
        case 128: LAUNCH_GET(128, false); break;
        case 512: LAUNCH_GET(512, false); break;
        case 1024: LAUNCH_GET(1024, false); break;
        default: LAUNCH_GET(256, false); break;
        }
    }
#undef LAUNCH_GET
}

__global__ __launch_bounds__(GRP_BS) void kg_ET_fix(const ETA *args, const int *off, int n) {
    const int p = grp_find(off, n, blockIdx.x);
    const ETA A = args[p];
    const long t = (long)(blockIdx.x - off[p]) * GRP_BS + threadIdx.x;
    if (t < A.d.nq) et_fix_at(A.q, *A.scale, t);
}

__global__ __launch_bounds__(GRP_BS) void kg_gemv_t(const GemvA *args, const int *off, int n) {
    static_assert(GRP_BS == GEMV_BS, "one row per workgroup");
    __shared__ double sh[GEMV_BS / 64 + 1];
    const int p = grp_find(off, n, blockIdx.x);
    const GemvA A = args[p];
    const int i = blockIdx.x - off[p];
    const double v = gemv_row(A.n, A.A, A.x, i, sh);
    if (threadIdx.x == 0) A.y[i] = v;
}

__global__ __launch_bounds__(GRP_BS) void kg_stress(const StressA *args, const int *off, int n) {
    const int p = grp_find(off, n, blockIdx.x);
    const StressA A = args[p];
    stress_body(A.boff, A.npts, A.nb, A.Fb, A.ds, A.h, A.dt, A.CD, A.CL, A.fbs,
                blockIdx.x - off[p]);
}

#define GRP_LAUNCH(NAME, KERNEL, TYPE)                                                       \
    void NAME(const Grp &g, hipStream_t s) {                                                 \
        if (g.n == 0 || g.blocks == 0) return;                                               \
        hipLaunchKernelGGL(KERNEL, dim3(g.blocks), dim3(GRP_BS), 0, s,                       \
                           static_cast<const TYPE *>(g.args), g.off, g.n);                   \
    }
GRP_LAUNCH(launch_grp_ET_zero, kg_ET_zero, ETA)
GRP_LAUNCH(launch_grp_ET_scale, kg_ET_scale, ETA)
GRP_LAUNCH(launch_grp_ET_zero_scale, kg_ET_zero_scale, ETA)
GRP_LAUNCH(launch_grp_ET_fix, kg_ET_fix, ETA)
GRP_LAUNCH(launch_grp_gemv_t, kg_gemv_t, GemvA)
GRP_LAUNCH(launch_grp_stress, kg_stress, StressA)

}  // namespace ibh
