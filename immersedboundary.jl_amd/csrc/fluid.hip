// Grid stencils of the streamfunction–vorticity discretization (fluid-ops.jl), one thread per
// output value, x-fastest so consecutive lanes touch consecutive addresses.  Every kernel is
// batched over levels (or over independent right-hand sides) with element strides.
#include <algorithm>

#include "ibh_internal.h"

namespace ibh {

#define SKIP_GUARD \
    if (skip && *skip) return;

// 1-based accessors matching the reference's indexing
#define GI(a, i, j) (a)[((i)-1) + (long)d.m * ((j)-1)]
#define QX(a, i, j) (a)[((i)-1) + (long)(d.nx + 1) * ((j)-1)]
#define QY(a, i, j) (a)[d.nu + ((i)-1) + (long)d.nx * ((j)-1)]
#define BC(a, k) (a)[(k)-1]

__host__ __device__ static inline unsigned nblk(long n, int bs) {
    return (unsigned)((n + bs - 1) / bs);
}

// curl_at (curl!, fluid-ops.jl:138-182) is in ibh_internal.h: E evaluates it too.

__global__ void k_curl_bc(Dom d, const double *__restrict__ psi, long spsi,
                          const double *__restrict__ pbc, long sbc, double *__restrict__ q,
                          long sq, const int *skip) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.nq) return;
    const int b = blockIdx.y;
    q[b * sq + t] = curl_at(d, psi + b * spsi, pbc + b * sbc, t);
}

void launch_curl_bc(const Dom &d, const double *psi, long spsi, const double *pbc, long sbc,
                    double *q, long sq, int nbatch, const int *skip, hipStream_t s) {
    hipLaunchKernelGGL(k_curl_bc, dim3(nblk(d.nq, 256), nbatch), dim3(256), 0, s, d, psi, spsi,
                       pbc, sbc, q, sq, skip);
}

// The same q entries on the faces with 1-based indices i ∈ [i0, i1], j ∈ [j0, j1] only (u and
// v each clipped to their own extent); the rest of q is left as it was.  For products whose
// reader (E) stays inside that box.
__global__ void k_curl_box(Dom d, const double *__restrict__ psi, long spsi,
                           const double *__restrict__ pbc, long sbc, double *__restrict__ q,
                           long sq, const int *skip, int i0, int i1, int j0, int j1) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    const int wi = i1 - i0 + 1, wj = j1 - j0 + 1;
    const long nb = (long)wi * wj;
    if (t >= 2 * nb) return;
    const bool uq = t < nb;
    const long e = uq ? t : t - nb;
    const int i = i0 + imod(e, wi), j = j0 + idiv(e, wi);
    if (uq ? (i > d.nx + 1 || j > d.ny) : (i > d.nx || j > d.ny + 1)) return;
    const long idx = uq ? (i - 1) + (long)(d.nx + 1) * (j - 1) : d.nu + (i - 1) + (long)d.nx * (j - 1);
    q[b * sq + idx] = curl_at(d, psi + b * spsi, pbc + b * sbc, idx);
}

void launch_curl_box(const Dom &d, const double *psi, long spsi, const double *pbc, long sbc,
                     double *q, long sq, int nbatch, const int *skip, hipStream_t s, int i0,
                     int i1, int j0, int j1) {
    i0 = std::max(i0, 1);
    j0 = std::max(j0, 1);
    i1 = std::min(i1, d.nx + 1);
    j1 = std::min(j1, d.ny + 1);
    if (i1 < i0 || j1 < j0) return;
    const long n = 2L * (i1 - i0 + 1) * (j1 - j0 + 1);
    hipLaunchKernelGGL(k_curl_box, dim3(nblk(n, 256), nbatch), dim3(256), 0, s, d, psi, spsi, pbc,
                       sbc, q, sq, skip, i0, i1, j0, j1);
}

// rot! (C'), fluid-ops.jl:117-129
__device__ inline double rot_at(const Dom &d, const double *__restrict__ q, long t) {
    const int i = imod(t, d.m) + 2, j = idiv(t, d.m) + 2;  // Γ[i-1, j-1]
    return (QX(q, i, j - 1) - QX(q, i, j)) + (QY(q, i, j) - QY(q, i - 1, j));
}

__global__ void k_rot(Dom d, const double *__restrict__ q, long sq, double *__restrict__ g,
                      long sg, const int *skip) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const int b = blockIdx.y;
    g[b * sg + t] = rot_at(d, q + b * sq, t);
}

void launch_rot(const Dom &d, const double *q, long sq, double *g, long sg, int nbatch,
                const int *skip, hipStream_t s) {
    hipLaunchKernelGGL(k_rot, dim3(nblk(d.ng, 256), nbatch), dim3(256), 0, s, d, q, sq, g, sg,
                       skip);
}

// the same from the fixed-point image of q that Eᵀ leaves: each entry converted exactly as the
// conversion pass of Eᵀ does ((double)v · (1/scale), inv = 1/scale), then the same differences
__device__ inline double rot_image_at(const Dom &d, const long long *__restrict__ q, double inv,
                                      long t) {
    const int i = imod(t, d.m) + 2, j = idiv(t, d.m) + 2;
    const double ux0 = (double)QX(q, i, j - 1) * inv, ux1 = (double)QX(q, i, j) * inv;
    const double vy1 = (double)QY(q, i, j) * inv, vy0 = (double)QY(q, i - 1, j) * inv;
    return (ux0 - ux1) + (vy1 - vy0);
}

__global__ void k_rot_image(Dom d, GBox box, const long long *__restrict__ img, long sq,
                            const double *__restrict__ scale, double *__restrict__ g, long sg,
                            const int *skip, const double *__restrict__ V,
                            const double *__restrict__ fb, long sfb, int n2,
                            double *__restrict__ w, int lrk) {
    SKIP_GUARD
    const long nv = gbox_vertices(d, box);
    const int nrb = (int)nblk(nv, 256);
    if ((int)blockIdx.x >= nrb) {  // the low-rank dots (lrk workgroups)
        const int j = blockIdx.x - nrb, b = blockIdx.y;
        lr_dot256(V + (long)n2 * j, fb + b * sfb, n2, w + (long)b * lrk + j);
        return;
    }
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= nv) return;
    const long t = gbox_vertex(d, box, k);
    const int b = blockIdx.y;
    g[b * sg + t] = rot_image_at(d, img + b * sq, 1.0 / scale[b], t);
}

void launch_rot_image(const Dom &d, const double *img, long sq, const double *scale, double *g,
                      long sg, int nbatch, const int *skip, hipStream_t s, GBox box,
                      const double *V, const double *fb, long sfb, int n2, double *w, int lrk) {
    hipLaunchKernelGGL(k_rot_image, dim3(nblk(gbox_vertices(d, box), 256) + lrk, nbatch),
                       dim3(256), 0, s, d, box, reinterpret_cast<const long long *>(img), sq, scale,
                       g, sg, skip, V, fb, sfb, n2, w, lrk);
}

// coarsify!(Γc, Γ), fluid-ops.jl:184-202: fine (i, j) = (2a + 2, 2b + 2) → coarse
// (nx/4 + 1 + a, ny/4 + 1 + b), a in 0..nx/2-3, b in 0..ny/2-3.
__device__ inline void coarsify_at(const Dom &d, double *__restrict__ gc,
                                   const double *__restrict__ g, long t) {
    const int na = d.nx / 2 - 1;
    const int a = imod(t, na), b = idiv(t, na);
    const int i = 2 + 2 * a, j = 2 + 2 * b;
    const int ic = d.nx / 4 + 1 + a, jc = d.ny / 4 + 1 + b;
    GI(gc, ic, jc) = GI(g, i, j) +
                     0.5 * (GI(g, i + 1, j) + GI(g, i, j + 1) + GI(g, i - 1, j) + GI(g, i, j - 1)) +
                     0.25 * (GI(g, i + 1, j + 1) + GI(g, i + 1, j - 1) + GI(g, i - 1, j - 1) +
                             GI(g, i - 1, j + 1));
}

__global__ void k_coarsify(Dom d, double *__restrict__ gc, const double *__restrict__ g,
                           long sgc, long sg, const int *skip) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)(d.nx / 2 - 1) * (d.ny / 2 - 1)) return;
    const int bz = blockIdx.y;
    coarsify_at(d, gc + bz * sgc, g + bz * sg, t);
}

void launch_coarsify(const Dom &d, double *gc, const double *g, long sgc, long sg, int nbatch,
                     const int *skip, hipStream_t s) {
    long n = (long)(d.nx / 2 - 1) * (d.ny / 2 - 1);
    hipLaunchKernelGGL(k_coarsify, dim3(nblk(n, 256), nbatch), dim3(256), 0, s, d, gc, g, sgc,
                       sg, skip);
}

// get_bc!(rbc, r), fluid-ops.jl:210-239, including the trailing ×0.25
__device__ inline double get_bc_at(const Dom &d, const double *__restrict__ r, int t) {
    const int nx = d.nx, ny = d.ny;
    const int k = t + 1;  // 1-based bc index
    double v;
    if (k > d.To) {  // top: ibc = k - T in 1..nx+1
        const int ibc = k - d.To;
        if (ibc & 1) v = GI(r, nx / 4 + (ibc - 1) / 2, 3 * ny / 4);
        else {
            const int i = nx / 4 + ibc / 2;
            v = 0.5 * (GI(r, i, 3 * ny / 4) + GI(r, i - 1, 3 * ny / 4));
        }
    } else if (k > d.Bo) {  // bottom
        const int ibc = k - d.Bo;
        if (ibc & 1) v = GI(r, nx / 4 + (ibc - 1) / 2, ny / 4);
        else {
            const int i = nx / 4 + ibc / 2;
            v = 0.5 * (GI(r, i, ny / 4) + GI(r, i - 1, ny / 4));
        }
    } else if (k > d.Ro) {  // right: jbc in 1..ny+1
        const int jbc = k - d.Ro;
        if (jbc & 1) v = GI(r, 3 * nx / 4, ny / 4 + (jbc - 1) / 2);
        else {
            const int j = ny / 4 + jbc / 2;
            v = 0.5 * (GI(r, 3 * nx / 4, j) + GI(r, 3 * nx / 4, j - 1));
        }
    } else {  // left
        const int jbc = k - d.Lo;
        if (jbc & 1) v = GI(r, nx / 4, ny / 4 + (jbc - 1) / 2);
        else {
            const int j = ny / 4 + jbc / 2;
            v = 0.5 * (GI(r, nx / 4, j) + GI(r, nx / 4, j - 1));
        }
    }
    return v * 0.25;
}

__global__ void k_get_bc(Dom d, const double *__restrict__ r, long sr, double *__restrict__ rbc,
                         long sbc, const int *skip) {
    SKIP_GUARD
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.nbc) return;
    const int b = blockIdx.y;
    rbc[b * sbc + t] = get_bc_at(d, r + b * sr, t);
}

void launch_get_bc(const Dom &d, const double *r, long sr, double *rbc, long sbc, int nbatch,
                   const int *skip, hipStream_t s) {
    hipLaunchKernelGGL(k_get_bc, dim3(nblk(d.nbc, 256), nbatch), dim3(256), 0, s, d, r, sr, rbc,
                       sbc, skip);
}

// out = in + fac·edges(rbc)  (apply_bc!, fluid-ops.jl:249-275, fused with the copy into Γtmp)
__device__ inline double copy_apply_bc_at(const Dom &d, const double *__restrict__ in,
                                          const double *__restrict__ rbc, double fac, long t) {
    const int i = imod(t, d.m) + 1, j = idiv(t, d.m) + 1;
    double v = in ? in[t] : 0.0;
    // same order as the reference: bottom, top, left, right
    if (j == 1) v += fac * BC(rbc, (d.Bo + 1) + i);
    if (j == d.ny - 1) v += fac * BC(rbc, (d.To + 1) + i);
    if (i == 1) v += fac * BC(rbc, (d.Lo + 1) + j);
    if (i == d.nx - 1) v += fac * BC(rbc, (d.Ro + 1) + j);
    return v;
}

__global__ void k_copy_apply_bc(Dom d, const double *__restrict__ in, long sin,
                                double *__restrict__ out, long sout,
                                const double *__restrict__ rbc, long sbc, double fac,
                                const int *skip) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const int b = blockIdx.y;
    out[b * sout + t] = copy_apply_bc_at(d, in ? in + b * sin : nullptr, rbc + b * sbc, fac, t);
}

void launch_copy_apply_bc(const Dom &d, const double *in, long sin, double *out, long sout,
                          const double *rbc, long sbc, double fac, int nbatch, const int *skip,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_copy_apply_bc, dim3(nblk(d.ng, 256), nbatch), dim3(256), 0, s, d, in,
                       sin, out, sout, rbc, sbc, fac, skip);
}

// get_bc! and the copy with apply_bc! in one pass: rbc = get_bc!(r) (the coarser level's ψ) and
// out = in + edges(rbc); the edge elements evaluate their boundary value from r themselves, the
// same value the rbc entry receives.  Threads [0, nΓ) copy, [nΓ, nΓ + nbc) write rbc.
// the value element t receives: out[t] for t < nΓ, rbc[t − nΓ] after
__device__ inline double copy_getbc_val(const Dom &d, const double *__restrict__ in,
                                        const double *__restrict__ r, long t) {
    if (t >= d.ng) return get_bc_at(d, r, (int)(t - d.ng));
    const int i = imod(t, d.m) + 1, j = idiv(t, d.m) + 1;
    double v = in[t];
    // as copy_apply_bc_at (fac = 1): bottom, top, left, right
    if (j == 1) v += get_bc_at(d, r, d.Bo + i);
    if (j == d.ny - 1) v += get_bc_at(d, r, d.To + i);
    if (i == 1) v += get_bc_at(d, r, d.Lo + j);
    if (i == d.nx - 1) v += get_bc_at(d, r, d.Ro + j);
    return v;
}
__device__ inline void copy_getbc_put(const Dom &d, double *__restrict__ out,
                                      double *__restrict__ rbc, long t, double v) {
    if (t < d.ng) out[t] = v;
    else rbc[t - d.ng] = v;
}
__device__ inline void copy_getbc_at(const Dom &d, const double *__restrict__ in,
                                     double *__restrict__ out, const double *__restrict__ r,
                                     double *__restrict__ rbc, long t) {
    if (t < d.ng + d.nbc) copy_getbc_put(d, out, rbc, t, copy_getbc_val(d, in, r, t));
}

__global__ void k_copy_getbc(Dom d, const double *__restrict__ in, long sin,
                             double *__restrict__ out, long sout, const double *__restrict__ r,
                             long sr, double *__restrict__ rbc, long sbc, const int *skip) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    copy_getbc_at(d, in + b * sin, out + b * sout, r + b * sr, rbc + b * sbc, t);
}

void launch_copy_getbc(const Dom &d, const double *in, long sin, double *out, long sout,
                       const double *r, long sr, double *rbc, long sbc, int nbatch,
                       const int *skip, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_getbc, dim3(nblk(d.ng + d.nbc, 256), nbatch), dim3(256), 0, s, d, in,
                       sin, out, sout, r, sr, rbc, sbc, skip);
}

void launch_apply_bc(const Dom &d, double *r, long sr, const double *rbc, long sbc, double fac,
                     int nbatch, const int *skip, hipStream_t s) {
    launch_copy_apply_bc(d, r, sr, r, sr, rbc, sbc, fac, nbatch, skip, s);
}

// Nonlinear term at one level: avg_flux! + direct_product! + rot! + 1/hc² (cnab/ops.jl:151-168,
// fluid-ops.jl:277-401).  q0 is the constant base flux (ux·hc, uy·hc) (fluid-ops.jl:403-417).
//
// One NL_TX×NL_TY tile of outputs per 256-thread workgroup.  Γ and ψ are staged in LDS together
// with their boundary ring: position 0 and nx of the x index, 0 and ny of the y index hold the
// boundary values (Γbc·scale, ψbc) that direct_product! and curl! read at the domain edges
// (fluid-ops.jl:316-349, 177-208).  With the ring in place the boundary cases of the flux
// (curl!) and of the face products are the interior formulas, so every thread runs the same
// arithmetic whatever its position.  The one exception is the ψ ring's corners: the x-flux
// reads the left and right columns' values there, the y-flux rows j = 1 and j = ny + 1 the
// bottom and top rows' (they are separate entries of ψbc), so those two y-flux rows read ψbc
// directly.  Γ's corners are never read.
// Tile extents (1-based): Γ, ψ over i ∈ [i0-2, i0+TX-1], j ∈ [j0-2, j0+TY-1]; QX over
// i ∈ [i0-1, i0+TX], j ∈ [j0-1, j0+TY-1]; QY over i ∈ [i0-1, i0+TX-1], j ∈ [j0-1, j0+TY]; zero
// outside the grid and its ring (the products never read those).
struct NLTileMem {
    double qx[NL_TY + 1][NL_TX + 2];
    double qy[NL_TY + 2][NL_TX + 1];
    double g[NL_TY + 2][NL_TX + 2];
};
struct NLTile {
    const NLTileMem *m;
    int i0, j0;
    double U, V;
    __device__ double qx(int i, int j) const { return m->qx[j - j0 + 1][i - i0 + 1]; }
    __device__ double qy(int i, int j) const { return m->qy[j - j0 + 1][i - i0 + 1]; }
    __device__ double gi(int i, int j) const { return m->g[j - j0 + 2][i - i0 + 2]; }
};

__device__ inline double Qx_avg(const NLTile &c, int i, int j) {  // j in 2..ny
    return (c.qx(i, j) + c.qx(i, j - 1) + c.U + c.U) / 2;
}
__device__ inline double Qy_avg(const NLTile &c, int i, int j) {  // i in 2..nx
    return (c.qy(i, j) + c.qy(i - 1, j) + c.V + c.V) / 2;
}
// fqx[i, j], i in 2..nx, j in 1..ny (Γ(i-1, 0) and Γ(i-1, ny) from the ring)
__device__ inline double fqx(const NLTile &c, int i, int j) {
    return (Qy_avg(c, i, j + 1) * c.gi(i - 1, j) + Qy_avg(c, i, j) * c.gi(i - 1, j - 1)) / 2;
}
// fqy[i, j], i in 1..nx, j in 2..ny (Γ(0, j-1) and Γ(nx, j-1) from the ring)
__device__ inline double fqy(const NLTile &c, int i, int j) {
    return -(Qx_avg(c, i + 1, j) * c.gi(i, j - 1) + Qx_avg(c, i, j) * c.gi(i - 1, j - 1)) / 2;
}

// The staging maps threads onto tile rows and columns directly: the first NL_TX columns of a
// block run one row per wave (rows ty, ty + NL_TY, ...), the remaining one or two columns go to
// the first threads.  Every global load of a thread is issued into registers before the first
// LDS store, so a wave has them all in flight together.
// With psi set the flux is curl!(ψ, ψbc) formed from the staged ψ (curl!'s arithmetic: one
// difference per face), and q is not read.
__device__ inline void nonlinear_tile(const Dom &d, const double *__restrict__ q,
                                      const double *__restrict__ g,
                                      const double *__restrict__ gbc, double gbc_scale, double U,
                                      double V, double hc, long tile, double *__restrict__ out,
                                      const double *__restrict__ psi = nullptr,
                                      const double *__restrict__ pbc = nullptr) {
    static_assert(NL_TX == 64 && NL_TX * NL_TY == 256, "a wave per tile row");
    constexpr int NT = NL_TX * NL_TY;
    __shared__ NLTileMem mem;
    __shared__ double ps[NL_TY + 2][NL_TX + 2];
    __shared__ double fx[NL_TY + 1][NL_TX];
    __shared__ double fy[NL_TY][NL_TX + 1];
    const int tid = threadIdx.x, tx = tid % NL_TX, ty = tid / NL_TX;
    const int ntx = (d.m + NL_TX - 1) / NL_TX;
    const int i0 = 2 + imod(tile, ntx) * NL_TX, j0 = 2 + idiv(tile, ntx) * NL_TY;
    const int nx = d.nx, ny = d.ny;

    // slots of a thread in a block of R rows and NL_TX + X columns
    constexpr int GR = NL_TY + 2, GX = 2, GS = (GR + NL_TY - 1) / NL_TY + 1;
    auto slot = [&](int k, int R, int X, int &r, int &cc) {
        const int nm = (R + NL_TY - 1) / NL_TY;
        if (k < nm) {
            r = ty + k * NL_TY;
            cc = tx;
            return r < R;
        }
        r = tid / X;
        cc = NL_TX + tid % X;
        return tid < R * X;
    };
    // Γ and ψ with the ring
    double gv[GS], pv[GS];
#pragma unroll
    for (int k = 0; k < GS; ++k) {
        int r, cc;
        const bool live = slot(k, GR, GX, r, cc);
        const int a = i0 - 2 + cc, b = j0 - 2 + r;
        double x = 0.0, y = 0.0;
        if (live && a >= 1 && a <= d.m && b >= 1 && b <= d.p) {
            x = GI(g, a, b);
            if (psi) y = GI(psi, a, b);
        } else if (live && a >= 0 && a <= nx && b >= 0 && b <= ny) {
            const bool side = a == 0 || a == nx;
            const int kb = a == 0 ? d.Lo + b : a == nx ? d.Ro + b : b == 0 ? d.Bo + a : d.To + a;
            if (!(side && (b == 0 || b == ny))) x = gbc[kb] * gbc_scale;
            if (psi) y = pbc[kb];
        }
        gv[k] = x;
        pv[k] = y;
    }
    constexpr int XR = NL_TY + 1, XS = (XR + NL_TY - 1) / NL_TY + 1;
    constexpr int YR = NL_TY + 2, YS = (YR + NL_TY - 1) / NL_TY + 1;
    double xv[XS], yv[YS];
    if (!psi) {
#pragma unroll
        for (int k = 0; k < XS; ++k) {
            int r, cc;
            const bool live = slot(k, XR, 2, r, cc);
            const int i = i0 - 1 + cc, j = j0 - 1 + r;
            xv[k] = (live && i >= 1 && i <= nx + 1 && j >= 1 && j <= ny) ? QX(q, i, j) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < YS; ++k) {
            int r, cc;
            const bool live = slot(k, YR, 1, r, cc);
            const int i = i0 - 1 + cc, j = j0 - 1 + r;
            yv[k] = (live && i >= 1 && i <= nx && j >= 1 && j <= ny + 1) ? QY(q, i, j) : 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < GS; ++k) {
        int r, cc;
        if (slot(k, GR, GX, r, cc)) {
            mem.g[r][cc] = gv[k];
            if (psi) ps[r][cc] = pv[k];
        }
    }
    if (psi) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < XS; ++k) {  // QX(i, j) = ψ(i-1, j) − ψ(i-1, j-1)
            int r, cc;
            const bool live = slot(k, XR, 2, r, cc);
            const int i = i0 - 1 + cc, j = j0 - 1 + r;
            xv[k] = (live && i >= 1 && i <= nx + 1 && j >= 1 && j <= ny) ? ps[r + 1][cc] - ps[r][cc]
                                                                       : 0.0;
        }
#pragma unroll
        for (int k = 0; k < YS; ++k) {  // QY(i, j) = ψ(i-1, j-1) − ψ(i, j-1)
            int r, cc;
            const bool live = slot(k, YR, 1, r, cc);
            const int i = i0 - 1 + cc, j = j0 - 1 + r;
            double v = 0.0;
            if (live && i >= 1 && i <= nx && j >= 1 && j <= ny + 1) {
                if (j == 1)
                    v = pbc[d.Bo + i - 1] - pbc[d.Bo + i];
                else if (j == ny + 1)
                    v = pbc[d.To + i - 1] - pbc[d.To + i];
                else
                    v = ps[r][cc] - ps[r][cc + 1];
            }
            yv[k] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < XS; ++k) {
        int r, cc;
        if (slot(k, XR, 2, r, cc)) mem.qx[r][cc] = xv[k];
    }
#pragma unroll
    for (int k = 0; k < YS; ++k) {
        int r, cc;
        if (slot(k, YR, 1, r, cc)) mem.qy[r][cc] = yv[k];
    }
    __syncthreads();
    const NLTile c{&mem, i0, j0, U, V};
    const int i = i0 + tx, j = j0 + ty;
    const bool in = i <= d.m + 1 && j <= d.p + 1;
    if (in) {
        fx[ty + 1][tx] = fqx(c, i, j);
        fy[ty][tx + 1] = fqy(c, i, j);
        if (ty == 0) fx[0][tx] = fqx(c, i, j - 1);
        if (tx == 0) fy[ty][0] = fqy(c, i - 1, j);
    }
    __syncthreads();
    if (in) {
        const double v = (fx[ty][tx] - fx[ty + 1][tx]) + (fy[ty][tx + 1] - fy[ty][tx]);
        out[(i - 2) + (long)d.m * (j - 2)] = v * (1 / (hc * hc));
    }
}

__global__ __launch_bounds__(256) void k_nonlinear(Dom d, const double *__restrict__ q, long sq,
                                                   const double *__restrict__ g, long sg,
                                                   const double *__restrict__ gbc, long sbc,
                                                   double gbc_scale, double *__restrict__ out,
                                                   long sout, double ux, double uy, double h,
                                                   int lev0, int q0_given,
                                                   const double *__restrict__ uv,
                                                   const double *__restrict__ psi, long spsi,
                                                   const double *__restrict__ pbc, long spbc) {
    if (uv) {  // freestream from the device parameter block (graph-replayable steps)
        ux = uv[0];
        uy = uv[1];
    }
    // tiles of all levels numbered XCD-contiguously (neighbouring tiles share halo rows)
    const int id = xcd_block();
    const int b = id / (int)gridDim.x;
    const int lev = lev0 + b;
    const double hc = h * (double)(1L << lev);  // gridstep, fluids/fluids.jl:102
    nonlinear_tile(d, q + b * sq, g + b * sg, gbc + b * sbc, gbc_scale,
                   q0_given ? ux : ux * hc, q0_given ? uy : uy * hc, hc, id % (int)gridDim.x,
                   out + b * sout, psi ? psi + b * spsi : nullptr, psi ? pbc + b * spbc : nullptr);
}

void launch_nonlinear(const Dom &d, const double *q, long sq, const double *g, long sg,
                      const double *gbc, long sbc, double gbc_scale, double *out, long sout,
                      double ux, double uy, double h, int lev0, int nbatch, hipStream_t s,
                      int q0_given, const double *uv, const double *psi, long spsi,
                      const double *pbc, long spbc) {
    static_assert(NL_TX * NL_TY == 256, "one thread per tile output");
    hipLaunchKernelGGL(k_nonlinear, dim3((unsigned)nl_tiles(d), nbatch), dim3(256), 0, s, d, q,
                       sq, g, sg, gbc, sbc, gbc_scale, out, sout, ux, uy, h, lev0, q0_given, uv,
                       psi, spsi, pbc, spbc);
}

// accbc[ℓ] = Σ_{l=ℓ}^{L-2} fac[l]·bcraw[l] for ℓ ≤ L-2, accbc[L-1] = 0 (the rhsbc of
// GetTrialState accumulates over levels, timestepping.jl:41-51).
__device__ inline void bc_prefix_at(int nbc, const double *__restrict__ bcraw,
                                    double *__restrict__ acc, const double *__restrict__ fac,
                                    int L, int t) {
    double s = 0.0;
    acc[(long)(L - 1) * nbc + t] = 0.0;
    for (int l = L - 2; l >= 0; --l) {
        s += fac[l] * bcraw[(long)l * nbc + t];
        acc[(long)l * nbc + t] = s;
    }
}

__global__ void k_bc_prefix(int nbc, const double *__restrict__ bcraw, double *__restrict__ acc,
                            const double *__restrict__ fac, int L) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbc) return;
    bc_prefix_at(nbc, bcraw, acc, fac, L, t);
}

void launch_bc_prefix(const Dom &d, const double *bcraw, double *accbc, const double *fac, int L,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_bc_prefix, dim3(nblk(d.nbc, 256)), dim3(256), 0, s, d.nbc, bcraw, accbc,
                       fac, L);
}

// Right-hand side of the trial solve (timestepping.jl:44-57) with A_ℓ applied as its exact
// 5-point stencil: A_ℓ = I − a(−L) with a = dt/(2 Re hc²) (cnab/ops.jl:26; the sine transform
// of fluid-ops.jl:42-56 diagonalises the same operator with zero values beyond the edges),
// then + f1·N1 + f2·N2 + the edge terms of the accumulated boundary vector, in the order of
// the reference's additions.
__device__ inline double rhs_at(const Dom &d, const double *__restrict__ g,
                                const double *__restrict__ N1, const double *__restrict__ N2,
                                const double *__restrict__ bc, double a, double f1, double f2,
                                long t) {
    const int i = imod(t, d.m), j = idiv(t, d.m);  // 0-based
    const double c = g[t];
    const double nsum = (i > 0 ? g[t - 1] : 0.0) + (i < d.m - 1 ? g[t + 1] : 0.0) +
                        (j > 0 ? g[t - d.m] : 0.0) + (j < d.p - 1 ? g[t + d.m] : 0.0);
    double v = c + a * (nsum - 4.0 * c);
    v = v + f1 * N1[t];
    v = v + f2 * N2[t];
    double eb = 0.0;
    if (j == 0) eb += bc[d.Bo + 1 + i];
    if (j == d.p - 1) eb += bc[d.To + 1 + i];
    if (i == 0) eb += bc[d.Lo + 1 + j];
    if (i == d.m - 1) eb += bc[d.Ro + 1 + j];
    return v + eb;
}

__global__ void k_rhs(Dom d, const double *__restrict__ g, const double *__restrict__ N1,
                      const double *__restrict__ N2, const double *__restrict__ bc,
                      double *__restrict__ out, double h, double dt, double Re, double f1,
                      double f2) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const int b = blockIdx.y;  // level
    const double hc = h * (double)(1L << b);
    const double a = dt / (2 * Re * (hc * hc));
    const long o = b * d.ng;
    out[o + t] = rhs_at(d, g + o, N1 + o, N2 + o, bc + (long)b * d.nbc, a, f1, f2, t);
}

void launch_rhs(const Dom &d, const double *g, const double *N1, const double *N2,
                const double *bcacc, double *out, int L, double h, double dt, double Re,
                double f1, double f2, hipStream_t s) {
    hipLaunchKernelGGL(k_rhs, dim3(nblk(d.ng, 256), L), dim3(256), 0, s, d, g, N1, N2, bcacc, out,
                       h, dt, Re, f1, f2);
}

// FP64 residual of a finest-level solve for iterative refinement: r = M·x − rhs with the
// 5-point operator that the sine transform inverts exactly (zero values beyond the edges):
// M = −L (Poisson, a < 0) or M = I + a(−L) (the implicit CN operator, a ≥ 0).
__device__ inline double resid_at(const Dom &d, const double *__restrict__ x,
                                  const double *__restrict__ rhs, double a, long t) {
    const int i = imod(t, d.m), j = idiv(t, d.m);
    const double c = x[t];
    const double nsum = (i > 0 ? x[t - 1] : 0.0) + (i < d.m - 1 ? x[t + 1] : 0.0) +
                        (j > 0 ? x[t - d.m] : 0.0) + (j < d.p - 1 ? x[t + d.m] : 0.0);
    const double lx = 4.0 * c - nsum;  // (−L)x
    return (a < 0 ? lx : c + a * lx) - rhs[t];
}

// operator coefficient of level b: a·ar^b (a < 0, the Poisson marker, is kept as is)
__device__ inline double level_coef(double a, double ar, int b) {
    for (int k = 0; k < b; ++k) a *= ar;
    return a;
}

__global__ void k_resid(Dom d, const double *__restrict__ x, const double *__restrict__ rhs,
                        double a, double ar, double *__restrict__ r) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long o = (long)blockIdx.y * d.ng;
    if (t < d.ng) r[o + t] = resid_at(d, x + o, rhs + o, level_coef(a, ar, blockIdx.y), t);
}

void launch_resid(const Dom &d, const double *x, const double *rhs, double a, double ar,
                  double *r, int nlev, hipStream_t s) {
    hipLaunchKernelGGL(k_resid, dim3(nblk(d.ng, 256), nlev), dim3(256), 0, s, d, x, rhs,
                       a < 0 ? -1.0 : a, a < 0 ? 1.0 : ar, r);
}

__global__ void k_fill(double *p, long n, double v) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// ---- boundary terms of a Poisson solve in the sine basis (used where the interior data is
// sparse and its forward transform skips the zero tiles).  apply_bc! (fluid-ops.jl:249-275)
// adds edge vectors on the first/last rows and columns of the right-hand side: cB at j = 1,
// cT at j = p, cL at i = 1, cR at i = m.  Their transform S_m·E·S_p is rank 4:
//   (S_m cB) ⊗ S_p[0,:] + (S_m cT) ⊗ S_p[p-1,:] + S_m[:,0] ⊗ (S_p cL) + S_m[:,m-1] ⊗ (S_p cR)
// (S symmetric).  k_edge_gemv forms W = [S_m cB | S_m cT | S_p cL | S_p cR] (2m + 2p per
// column), one workgroup per output entry; k_edge_add adds the rank-4 term, times the
// eigenvalue scale D, to the scaled spectral array F = (S_m X S_p) ⊙ D.
// Rows i and n−1−i of S (n = m or p, 0-based; S[i, k] = sin(π(i+1)(k+1)/(n+1))) differ only by
// the sign of the odd columns: S[n−1−i, k] = (−1)^k S[i, k].  So one read of row i gives both outputs, E + O and E − O, with
// E and O the sums over even and odd k.  One wave per such pair (the middle row, i = (n−1)/2, is
// its own mirror: its odd columns are zero); the outputs of pair x of one batch entry (boundary
// values r, W row Wb), x over the pairs of S_m and then of S_p.
__device__ inline int edge_pairs(const Dom &d) { return d.nx / 2 + d.ny / 2; }
__device__ inline void edge_gemv_pair(const Dom &d, const double *__restrict__ Sm,
                                      const double *__restrict__ Sp, const double *__restrict__ r,
                                      double *__restrict__ Wb, int x) {
    const int hx = d.nx / 2;
    if (x >= edge_pairs(d)) return;
    const bool xs = x < hx;                     // S_m pairs first, then S_p pairs
    const int i = xs ? x : x - hx, n = xs ? d.m : d.p;
    const double *S = (xs ? Sm : Sp) + (long)i * n;
    const int o1 = xs ? d.Bo : d.Lo, o2 = xs ? d.To : d.Ro;  // 0-based: entry k at o + 1 + k
    const int lane = threadIdx.x & 63;
    // lane l sums the columns k ≡ l (mod 64): even lanes the even k, odd lanes the odd ones
    double s1 = 0.0, s2 = 0.0;
    for (int k = lane; k < n; k += 64) {
        const double sv = S[k];
        s1 += sv * r[o1 + 1 + k];
        s2 += sv * r[o2 + 1 + k];
    }
    // fixed-order sums over the even lanes (into lane 0) and the odd lanes (into lane 1)
    for (int o = 32; o > 1; o >>= 1) {
        s1 += __shfl_down(s1, o, 64);
        s2 += __shfl_down(s2, o, 64);
    }
    const double o1s = __shfl_down(s1, 1, 64), o2s = __shfl_down(s2, 1, 64);
    if (lane != 0) return;
    const int mi = n - 1 - i;  // the mirror row (sin(π(N − i − 1)(k + 1)/N), N = n + 1)
    double *W1 = Wb + (xs ? 0 : 2 * d.m), *W2 = Wb + (xs ? d.m : 2 * d.m + d.p);
    W1[i] = s1 + o1s;
    W2[i] = s2 + o2s;
    if (mi != i) {
        W1[mi] = s1 - o1s;
        W2[mi] = s2 - o2s;
    }
}

__global__ __launch_bounds__(256) void k_edge_gemv(Dom d, const double *__restrict__ Sm,
                                                   const double *__restrict__ Sp,
                                                   const double *__restrict__ bc, long sbc,
                                                   double *__restrict__ W, long sW,
                                                   const int *skip) {
    SKIP_GUARD
    edge_gemv_pair(d, Sm, Sp, bc + blockIdx.y * sbc, W + blockIdx.y * sW,
                   blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
}

// the rank-4 boundary term at spectral entry (a, c), from one batch entry's W
__device__ inline double edge_term(const Dom &d, const double *__restrict__ Sm,
                                   const double *__restrict__ Sp, const double *__restrict__ Wb,
                                   int a, int c) {
    return Wb[a] * Sp[c] + Wb[d.m + a] * Sp[(long)(d.p - 1) * d.p + c] + Sm[a] * Wb[2 * d.m + c] +
           Sm[(long)(d.m - 1) * d.m + a] * Wb[2 * d.m + d.p + c];
}

__global__ void k_edge_add(Dom d, const double *__restrict__ Sm, const double *__restrict__ Sp,
                           const double *__restrict__ W, long sW, const double *__restrict__ D,
                           double *__restrict__ F, long sF, const int *skip) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const int b = blockIdx.y;
    const int a = imod(t, d.m), c = idiv(t, d.m);
    F[b * sF + t] += D[t] * edge_term(d, Sm, Sp, W + b * sW, a, c);
}

void launch_edge_gemv(const Dom &d, const double *Sm, const double *Sp, const double *bc,
                      long sbc, double *W, int nbatch, const int *skip, hipStream_t s) {
    hipLaunchKernelGGL(k_edge_gemv, dim3(nblk(d.nx / 2 + d.ny / 2, 4), nbatch), dim3(256), 0, s,
                       d, Sm, Sp, bc, sbc, W, 2L * (d.m + d.p), skip);
}

void launch_edge_spectral(const Dom &d, const double *Sm, const double *Sp, const double *bc,
                          long sbc, double *W, const double *D, double *F, long sF, int nbatch,
                          const int *skip, hipStream_t s) {
    const long sW = 2L * (d.m + d.p);
    launch_edge_gemv(d, Sm, Sp, bc, sbc, W, nbatch, skip, s);
    hipLaunchKernelGGL(k_edge_add, dim3(nblk(d.ng, 256), nbatch), dim3(256), 0, s, d, Sm, Sp, W,
                       sW, D, F, sF, skip);
}

// ---- ψ = S_m F S_p on four lines only: the rows i = nx/4, 3nx/4 and the columns j = ny/4,
// 3ny/4 (1-based), the values get_bc! reads (fluid-ops.jl:210-239).  Coarse levels of a
// B-product need nothing else, so this replaces the two inverse contractions there.  Four
// deterministic passes (no atomics): column partial sums v_J = F·S_p[:, J] over LN_CH chunks of
// the column index, their sum, the row sums w_I = S_m[I, :]·F (one workgroup per column), and the
// line values S_m·v_J and w_I·S_p (one workgroup per output).  Work per batch entry:
// 2·LN_CH·m + 2m + 2p doubles.
constexpr int LN_CH = 128;  // two per lane in k_lines_colsum

// With EW (boundary-term GEMVs, 2(m+p) per batch entry) the spectrum read is F + D ⊙ edges.
__device__ inline void lines_colpart(const Dom &d, const double *__restrict__ F, long sF,
                                     const double *__restrict__ Sm, const double *__restrict__ Sp,
                                     const double *__restrict__ EW, const double *__restrict__ D,
                                     double *__restrict__ W, long sW, int a, int ch, int bt) {
    if (a >= d.m) return;
    const double *Fb = F + bt * sF;
    const double *Eb = EW ? EW + bt * 2L * (d.m + d.p) : nullptr;
    const int J1 = d.ny / 4 - 1, J2 = 3 * d.ny / 4 - 1;
    const int b0 = (int)((long)ch * d.p / LN_CH), b1 = (int)((long)(ch + 1) * d.p / LN_CH);
    double s1 = 0.0, s2 = 0.0;
    // four rows' loads issued together; the sums still run over b in order (padding adds 0)
    for (int b = b0; b < b1; b += 4) {
        double f[4], p1[4], p2[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int bb = b + u;
            const bool ok = bb < b1;
            const long t = a + (long)d.m * bb;
            double x = ok ? Fb[t] : 0.0;
            if (Eb && ok) x += D[t] * edge_term(d, Sm, Sp, Eb, a, bb);
            f[u] = x;
            p1[u] = ok ? Sp[(long)J1 * d.p + bb] : 0.0;
            p2[u] = ok ? Sp[(long)J2 * d.p + bb] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s1 += f[u] * p1[u];
            s2 += f[u] * p2[u];
        }
    }
    double *Wb = W + bt * sW;
    Wb[((long)a * LN_CH + ch) * 2] = s1;  // chunks of one entry contiguous: [a][ch][J]
    Wb[((long)a * LN_CH + ch) * 2 + 1] = s2;
}

// one wave per entry a: lane l adds chunks l and l + 64 (coalesced reads), then a wave sum
__device__ inline void lines_colsum(const Dom &d, double *__restrict__ Wb, int a) {
    static_assert(LN_CH == 128, "two chunks per lane");
    const int lane = threadIdx.x & 63;
    if (a >= d.m) return;
    const double *w = Wb + (long)a * LN_CH * 2;
    double s1 = w[lane * 2] + w[(lane + 64) * 2];
    double s2 = w[lane * 2 + 1] + w[(lane + 64) * 2 + 1];
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        Wb[2L * LN_CH * d.m + a] = s1;          // v_J1
        Wb[(2L * LN_CH + 1) * d.m + a] = s2;    // v_J2
    }
}

__global__ void k_lines_colsum(Dom d, double *__restrict__ W, long sW, const int *skip) {
    SKIP_GUARD
    lines_colsum(d, W + blockIdx.y * sW, blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
}

// The row sums and the line values give each output one wave of a 256-thread workgroup
// (LN_OUTS outputs per workgroup).  Lane l sums the indices l, l + 64, ... in order; the loads
// of LN_U of them are issued together before their products are added, and the lane partials
// go through a fixed-order wave sum (deterministic; the single-case and the grouped kernels run
// the same arithmetic).
constexpr int LN_OUTS = 4, LN_U = 8;

// row sums w_I[b] = Σ_a S_m[I, a]·F[a, b] for output column b
__device__ inline void lines_row(const Dom &d, const double *__restrict__ F, long sF,
                                 const double *__restrict__ Sm, const double *__restrict__ Sp,
                                 const double *__restrict__ EW, const double *__restrict__ D,
                                 double *__restrict__ W, long sW, int b, int bt) {
    if (b >= d.p) return;  // whole wave
    const int lane = threadIdx.x & 63;
    const double *Fc = F + bt * sF + (long)d.m * b;
    const double *Dc = D + (long)d.m * b;
    const double *Eb = EW ? EW + bt * 2L * (d.m + d.p) : nullptr;
    const double *S1 = Sm + (long)(d.nx / 4 - 1) * d.m, *S2 = Sm + (long)(3 * d.nx / 4 - 1) * d.m;
    double s1 = 0.0, s2 = 0.0;
    for (int a0 = 0; a0 < d.m; a0 += 64 * LN_U) {
        double f[LN_U], m1[LN_U], m2[LN_U];
#pragma unroll
        for (int u = 0; u < LN_U; ++u) {
            const int a = a0 + 64 * u + lane;
            const bool ok = a < d.m;
            double x = ok ? Fc[a] : 0.0;
            if (Eb && ok) x += Dc[a] * edge_term(d, Sm, Sp, Eb, a, b);
            f[u] = x;
            m1[u] = ok ? S1[a] : 0.0;
            m2[u] = ok ? S2[a] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < LN_U; ++u) {
            s1 += m1[u] * f[u];
            s2 += m2[u] * f[u];
        }
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        double *Wr = W + bt * sW + (2L * LN_CH + 2) * d.m;
        Wr[b] = s1;        // w_I1
        Wr[d.p + b] = s2;  // w_I2
    }
}

// the column chunks (blocks [0, nca·LN_CH)) and the row sums (the blocks after) in one launch
__global__ __launch_bounds__(256) void k_lines_part(Dom d, const double *__restrict__ F, long sF,
                                                    const double *__restrict__ Sm,
                                                    const double *__restrict__ Sp,
                                                    const double *__restrict__ EW,
                                                    const double *__restrict__ D,
                                                    double *__restrict__ W, long sW, int nca,
                                                    const int *skip) {
    SKIP_GUARD
    const int bt = blockIdx.y;
    const int blk = blockIdx.x;
    if (blk < nca * LN_CH) {
        lines_colpart(d, F, sF, Sm, Sp, EW, D, W, sW, (blk % nca) * 256 + threadIdx.x, blk / nca,
                      bt);
    } else {
        lines_row(d, F, sF, Sm, Sp, EW, D, W, sW,
                  (blk - nca * LN_CH) * LN_OUTS + (threadIdx.x >> 6), bt);
    }
}

// The mirrored pairs of outputs: rows o and n−1−o of S share their values up to the sign of the
// odd columns (S[n−1−o, k] = (−1)^k S[o, k]), so one read of row o gives both lines' values at o
// and at its mirror from the even and odd sums, E ± O (lane l of a wave sums the columns of
// parity l, the chunks being 64-aligned).  Pairs [0, nx/2) are the column part
// ψ(o, J) = Σ_a S_m[o, a] v_J[a], then [nx/2, nx/2 + ny/2) the row part ψ(I, c) = Σ_b w_I[b] S_p[b, c].
__device__ inline void lines_out(const Dom &d, const double *__restrict__ Sm,
                                 const double *__restrict__ Sp, const double *__restrict__ Wb,
                                 double *__restrict__ ps, int pr) {
    const int hx = d.nx / 2;
    if (pr >= hx + d.ny / 2) return;  // whole wave
    const int lane = threadIdx.x & 63;
    const bool col = pr < hx;
    const int o = col ? pr : pr - hx;
    const int n = col ? d.m : d.p;
    const double *S = (col ? Sm : Sp) + (long)o * n;
    const double *v1 = col ? Wb + 2L * LN_CH * d.m : Wb + (2L * LN_CH + 2) * d.m;
    const double *v2 = col ? v1 + d.m : v1 + d.p;
    double s1 = 0.0, s2 = 0.0;
    for (int a0 = 0; a0 < n; a0 += 64 * LN_U) {
        double sv[LN_U], x1[LN_U], x2[LN_U];
#pragma unroll
        for (int u = 0; u < LN_U; ++u) {
            const int a = a0 + 64 * u + lane;
            const bool ok = a < n;
            sv[u] = ok ? S[a] : 0.0;
            x1[u] = ok ? v1[a] : 0.0;
            x2[u] = ok ? v2[a] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < LN_U; ++u) {
            s1 += sv[u] * x1[u];
            s2 += sv[u] * x2[u];
        }
    }
    for (int k = 32; k > 1; k >>= 1) {  // even lanes into lane 0, odd lanes into lane 1
        s1 += __shfl_down(s1, k, 64);
        s2 += __shfl_down(s2, k, 64);
    }
    const double O1 = __shfl_down(s1, 1, 64), O2 = __shfl_down(s2, 1, 64);
    if (lane != 0) return;
    const double E1 = s1, E2 = s2;
    const int mo = n - 1 - o;  // the mirror output
    if (col) {
        const long J1 = (long)d.m * (d.ny / 4 - 1), J2 = (long)d.m * (3 * d.ny / 4 - 1);
        ps[o + J1] = E1 + O1;
        ps[o + J2] = E2 + O2;
        if (mo != o) {
            ps[mo + J1] = E1 - O1;
            ps[mo + J2] = E2 - O2;
        }
    } else {
        // the four crossings with the columns belong to the column part (one writer each,
        // so the result does not depend on which workgroup finishes last)
        const int I1 = d.nx / 4 - 1, I2 = 3 * d.nx / 4 - 1;
        auto put = [&](int c, double t1, double t2) {
            if (c == d.ny / 4 - 1 || c == 3 * d.ny / 4 - 1) return;
            ps[I1 + (long)d.m * c] = t1;
            ps[I2 + (long)d.m * c] = t2;
        };
        put(o, E1 + O1, E2 + O2);
        if (mo != o) put(mo, E1 - O1, E2 - O2);
    }
}

__global__ __launch_bounds__(256) void k_lines_out(Dom d, const double *__restrict__ Sm,
                                                   const double *__restrict__ Sp,
                                                   const double *__restrict__ W, long sW,
                                                   double *__restrict__ psi, long spsi,
                                                   const int *skip) {
    SKIP_GUARD
    lines_out(d, Sm, Sp, W + blockIdx.y * sW, psi + blockIdx.y * spsi,
              blockIdx.x * LN_OUTS + (threadIdx.x >> 6));
}

long lines_work_size(const Dom &d) { return 2L * LN_CH * d.m + 2L * d.m + 2L * d.p; }

__global__ void k_spectral_scale(Dom d, const double *__restrict__ A, long sA,
                                 const double *__restrict__ D, double kappa,
                                 const double *__restrict__ Sm, const double *__restrict__ Sp,
                                 const double *__restrict__ EW, double *__restrict__ F, long sF,
                                 const int *skip, const double *__restrict__ sub) {
    SKIP_GUARD
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const int b = blockIdx.y;
    const double a = sub ? A[b * sA + t] - sub[b * sA + t] : A[b * sA + t];
    double v = kappa * a * D[t];
    if (EW) {
        const int a = imod(t, d.m), c = idiv(t, d.m);
        v += D[t] * edge_term(d, Sm, Sp, EW + b * 2L * (d.m + d.p), a, c);
    }
    F[b * sF + t] = v;
}

void launch_spectral_scale(const Dom &d, const double *A, long sA, const double *D, double kappa,
                           double *F, long sF, int nbatch, const int *skip, hipStream_t s,
                           const double *Sm, const double *Sp, const double *EW,
                           const double *sub) {
    hipLaunchKernelGGL(k_spectral_scale, dim3(nblk(d.ng, 256), nbatch), dim3(256), 0, s, d, A, sA,
                       D, kappa, Sm, Sp, EW, F, sF, skip, sub);
}

void launch_lines(const Dom &d, const double *Sm, const double *Sp, const double *F, long sF,
                  double *psi, long spsi, double *W, int nbatch, const int *skip, hipStream_t s,
                  const double *EW, const double *D) {
    const long sW = lines_work_size(d);
    const int nca = (int)nblk(d.m, 256);
    hipLaunchKernelGGL(k_lines_part, dim3(nca * LN_CH + nblk(d.p, LN_OUTS), nbatch), dim3(256), 0, s, d,
                       F, sF, Sm, Sp, EW, D, W, sW, nca, skip);
    hipLaunchKernelGGL(k_lines_colsum, dim3(nblk(d.m, 4), nbatch), dim3(256), 0, s, d, W, sW,
                       skip);
    hipLaunchKernelGGL(k_lines_out, dim3(nblk(d.nx / 2 + d.ny / 2, LN_OUTS), nbatch), dim3(256), 0, s, d, Sm, Sp,
                       W, sW, psi, spsi, skip);
}

__global__ void k_fill_strided(double *p, long n, long stride, double v) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < n) p[blockIdx.y * stride + t] = v;
}

void launch_fill_strided(double *p, long n, long stride, int count, double v, hipStream_t s) {
    if (n <= 0 || count <= 0) return;
    hipLaunchKernelGGL(k_fill_strided, dim3(nblk(n, 256), count), dim3(256), 0, s, p, n, stride, v);
}

void launch_fill(double *p, long n, double v, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, s, p, n, v);
}

// S[i,k] = sin(π i k / n), i, k = 1..n-1, stored (n-1)×(n-1).  The argument is reduced
// exactly in integers (i·k mod 2n) before the sine so every entry is correctly rounded.
__global__ void k_sine_matrix(double *S, int n) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long m = n - 1;
    if (t >= m * m) return;
    const long i = t % m + 1, k = t / m + 1;
    const long r = (i * k) % (2L * n);
    S[t] = sinpi((double)r / (double)n);
}

// Odd and even parts of the DST-I matrix for the split contraction, arguments reduced
// exactly in integers before the sine.
__global__ void k_sine_split(double *So, double *Se, int n) {
    const long h = n / 2;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= h * h) return;
    const long k = t % h, a = t / h;  // element [k + h·a]
    So[t] = sinpi((double)(((a + 1) * (2 * k + 1)) % (2L * n)) / (double)n);
    Se[t] = (a == h - 1 || k == h - 1)
                ? 0.0
                : sinpi((double)((2 * (a + 1) * (k + 1)) % (2L * n)) / (double)n);
}

void launch_sine_split(double *So, double *Se, int n, hipStream_t s) {
    long hh = (long)(n / 2) * (n / 2);
    hipLaunchKernelGGL(k_sine_split, dim3(nblk(hh, 256)), dim3(256), 0, s, So, Se, n);
}

void launch_sine_matrix(double *S, int n, hipStream_t s) {
    long mm = (long)(n - 1) * (n - 1);
    hipLaunchKernelGGL(k_sine_matrix, dim3(nblk(mm, 256)), dim3(256), 0, s, S, n);
}

// Eigenvalue scalings between the two transforms (fluid-ops.jl:37,50,58-61; cnab/ops.jl:26,29):
// sD = s/Λ, sA_ℓ = s/Λexpl_ℓ, sAinv_ℓ = s/Λimpl_ℓ, with s = 1/(4 nx ny) and the 2·2 factors
// of the two RODFT00 dimensions folded into the S matrices' caller (see host code).
__global__ void k_scale_tables(Dom d, double *sD, double *sA, double *sAinv, int L, double h,
                               double dt, double Re, double scale) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const int i = imod(t, d.m) + 1, j = idiv(t, d.m) + 1;
    const double lam = -2 * (cos(M_PI * i / d.nx) + cos(M_PI * j / d.ny) - 2);
    sD[t] = scale / lam;
    for (int l = 0; l < L; ++l) {
        const double hc = h * (double)(1L << l);
        const double lexpl = 1.0 / (1.0 - lam * dt / (2 * Re * (hc * hc)));
        const double limpl = 1.0 + lam * dt / (2 * Re * (hc * hc));
        sA[(long)l * d.ng + t] = scale / lexpl;
        sAinv[(long)l * d.ng + t] = scale / limpl;
    }
}

void launch_scale_tables(const Dom &d, double *sD, double *sA, double *sAinv, int L, double h,
                         double dt, double Re, hipStream_t s) {
    // 16·s: each RODFT00 is 2·S per dimension, so DST∘DST contributes (2·2)² = 16
    const double scale = 16.0 / (4.0 * d.nx * d.ny);
    hipLaunchKernelGGL(k_scale_tables, dim3(nblk(d.ng, 256)), dim3(256), 0, s, d, sD, sA, sAinv,
                       L, h, dt, Re, scale);
}

// ------------------------------------------------------------------------ grouped launches
// The same per-element bodies over a list of problems of different sizes: the workgroup finds
// its problem, then its batch entry b and element t inside that problem.
// grouped kernels take their workgroup id XCD-contiguously (xcd_block)
#define XB xcd_block()
#define GRP_PROLOGUE(TYPE, ELEMS)                                    \
    const int p_ = grp_find(off, n, XB);                     \
    const TYPE A = args[p_];                                         \
    const int lb_ = XB - off[p_];                            \
    const int nbx_ = (int)nblk((ELEMS), GRP_BS);                     \
    const int b = lb_ / nbx_;                                        \
    const long t = (long)(lb_ % nbx_) * GRP_BS + threadIdx.x;        \
    const Dom &d = A.d;                                              \
    (void)d;

// GRP_EPT elements per thread (grp_blocks(ELEMS) workgroups per batch entry): every value is
// formed before any is stored, so the loads of all of them are in flight together.  Defines A,
// d, b and the thread's first element t0; workgroups past the problem's batch entries leave.
#define GRP_PROLOGUE_N(TYPE, ELEMS)                                  \
    const int p_ = grp_find(off, n, XB);                     \
    const TYPE A = args[p_];                                         \
    const int lb_ = XB - off[p_];                            \
    const int nbx_ = (int)grp_blocks(ELEMS);                         \
    const int b = lb_ / nbx_;                                        \
    const Dom &d = A.d;                                              \
    if (b >= A.nbatch) return;                                       \
    const long t0 = (long)(lb_ % nbx_) * GRP_EPT * GRP_BS + threadIdx.x;

#define GRP_ELEMENTS(N, OUT, VALUE)                                  \
    {                                                                \
        double v_[GRP_EPT];                                          \
        _Pragma("unroll") for (int k_ = 0; k_ < GRP_EPT; ++k_) {     \
            const long t = t0 + (long)k_ * GRP_BS;                   \
            v_[k_] = t < (N) ? (VALUE) : 0.0;                        \
        }                                                            \
        _Pragma("unroll") for (int k_ = 0; k_ < GRP_EPT; ++k_) {     \
            const long t = t0 + (long)k_ * GRP_BS;                   \
            if (t < (N)) (OUT)[t] = v_[k_];                          \
        }                                                            \
    }

__global__ __launch_bounds__(GRP_BS) void kg_curl(const CurlA *args, const int *off, int n) {
    GRP_PROLOGUE_N(CurlA, A.d.nq)
    const double *psi = A.psi + b * A.spsi, *pbc = A.pbc + b * A.sbc;
    GRP_ELEMENTS(d.nq, A.q + b * A.sq, curl_at(d, psi, pbc, t))
}

__global__ __launch_bounds__(GRP_BS) void kg_rot(const RotA *args, const int *off, int n) {
    GRP_PROLOGUE(RotA, A.d.ng)
    if (t >= d.ng) return;
    A.g[b * A.sg + t] = rot_at(d, A.q + b * A.sq, t);
}

__global__ __launch_bounds__(GRP_BS) void kg_rot_image(const RotIA *args, const int *off, int n) {
    GRP_PROLOGUE(RotIA, gbox_vertices(A.d, A.box))
    (void)b;  // one batch entry per problem
    if (t >= gbox_vertices(d, A.box)) return;
    const long v = gbox_vertex(d, A.box, t);  // Γ is kept zero outside the box
    A.g[v] = rot_image_at(d, reinterpret_cast<const long long *>(A.img), 1.0 / *A.scale, v);
}

__global__ __launch_bounds__(GRP_BS) void kg_coarsify(const CoarsA *args, const int *off, int n) {
    GRP_PROLOGUE(CoarsA, grp_elems_coarsify(A.d))
    if (t >= grp_elems_coarsify(d)) return;
    coarsify_at(d, A.gc + b * A.sgc, A.g + b * A.sg, t);
}

__global__ __launch_bounds__(GRP_BS) void kg_get_bc(const GetBcA *args, const int *off, int n) {
    GRP_PROLOGUE(GetBcA, A.d.nbc)
    if (t >= d.nbc) return;
    A.rbc[b * A.sbc + t] = get_bc_at(d, A.r + b * A.sr, (int)t);
}

__global__ __launch_bounds__(GRP_BS) void kg_copy_apply_bc(const CopyBcA *args, const int *off,
                                                           int n) {
    GRP_PROLOGUE_N(CopyBcA, A.d.ng)
    const double *in = A.in ? A.in + b * A.sin : nullptr, *rbc = A.rbc + b * A.sbc;
    GRP_ELEMENTS(d.ng, A.out + b * A.sout, copy_apply_bc_at(d, in, rbc, A.fac, t))
}

__global__ __launch_bounds__(GRP_BS) void kg_copy_getbc(const CopyGetBcA *args, const int *off,
                                                        int n) {
    GRP_PROLOGUE_N(CopyGetBcA, A.d.ng + A.d.nbc)
    const long N = d.ng + d.nbc;
    double v_[GRP_EPT];  // every value formed before any is stored
#pragma unroll
    for (int k = 0; k < GRP_EPT; ++k) {
        const long t = t0 + (long)k * GRP_BS;
        v_[k] = t < N ? copy_getbc_val(d, A.in, A.r, t) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < GRP_EPT; ++k) {
        const long t = t0 + (long)k * GRP_BS;
        if (t < N) copy_getbc_put(d, A.out, A.rbc, t, v_[k]);
    }
}

// nl_tiles(d) workgroups per batch entry (level)
__global__ __launch_bounds__(GRP_BS) void kg_nonlinear(const NonlinA *args, const int *off, int n) {
    static_assert(GRP_BS == NL_TX * NL_TY, "one thread per tile output");
    const int p = grp_find(off, n, XB);
    const NonlinA A = args[p];
    const int nt = (int)nl_tiles(A.d);
    const int lb = XB - off[p];
    const int b = lb / nt;
    const double hc = A.h * (double)(1L << (A.lev0 + b));
    const double ux = A.uv[0], uy = A.uv[1];
    nonlinear_tile(A.d, A.q + b * A.sq, A.g + b * A.sg, A.gbc + b * A.sbc, A.gbc_scale, ux * hc,
                   uy * hc, hc, lb % nt, A.out + b * A.sout, A.psi ? A.psi + b * A.spsi : nullptr,
                   A.psi ? A.pbc + b * A.spbc : nullptr);
}

__global__ __launch_bounds__(GRP_BS) void kg_rhs(const RhsA *args, const int *off, int n) {
    GRP_PROLOGUE_N(RhsA, A.d.ng)
    const double hc = A.h * (double)(1L << b);
    const double a = A.dt / (2 * A.Re * (hc * hc));
    const long o = b * d.ng;
    const double *bc = A.bc + (long)b * d.nbc;
    GRP_ELEMENTS(d.ng, A.out + o, rhs_at(d, A.g + o, A.N1 + o, A.N2 + o, bc, a, A.f1, A.f2, t))
}

__global__ __launch_bounds__(GRP_BS) void kg_resid(const ResidA *args, const int *off, int n) {
    GRP_PROLOGUE_N(ResidA, A.d.ng)
    const long o = (long)b * d.ng;
    const double c = level_coef(A.a, A.ar, b);
    GRP_ELEMENTS(d.ng, A.r + o, resid_at(d, A.x + o, A.rhs + o, c, t))
}

__global__ __launch_bounds__(GRP_BS) void kg_bc_prefix(const BcPrefA *args, const int *off, int n) {
    const int p = grp_find(off, n, XB);
    const BcPrefA A = args[p];
    const int t = (XB - off[p]) * GRP_BS + threadIdx.x;
    if (t >= A.nbc) return;
    bc_prefix_at(A.nbc, A.bcraw, A.acc, A.fac, A.L, t);
}

__global__ __launch_bounds__(GRP_BS) void kg_fill(const FillA *args, const int *off, int n) {
    const int p = grp_find(off, n, XB);
    const FillA A = args[p];
    const long t = (long)(XB - off[p]) * GRP_BS + threadIdx.x;
    if (t < A.n) A.p[t] = A.v;
}

#define GRP_LAUNCH(NAME, KERNEL, TYPE)                                                       \
    void NAME(const Grp &g, hipStream_t s) {                                                 \
        if (g.n == 0 || g.blocks == 0) return;                                               \
        hipLaunchKernelGGL(KERNEL, dim3(g.blocks), dim3(GRP_BS), 0, s,                       \
                           static_cast<const TYPE *>(g.args), g.off, g.n);                   \
    }
GRP_LAUNCH(launch_grp_curl, kg_curl, CurlA)
GRP_LAUNCH(launch_grp_rot, kg_rot, RotA)
GRP_LAUNCH(launch_grp_rot_image, kg_rot_image, RotIA)
GRP_LAUNCH(launch_grp_coarsify, kg_coarsify, CoarsA)
GRP_LAUNCH(launch_grp_get_bc, kg_get_bc, GetBcA)
GRP_LAUNCH(launch_grp_copy_apply_bc, kg_copy_apply_bc, CopyBcA)
GRP_LAUNCH(launch_grp_copy_getbc, kg_copy_getbc, CopyGetBcA)
GRP_LAUNCH(launch_grp_nonlinear, kg_nonlinear, NonlinA)
GRP_LAUNCH(launch_grp_bc_prefix, kg_bc_prefix, BcPrefA)
GRP_LAUNCH(launch_grp_rhs, kg_rhs, RhsA)
GRP_LAUNCH(launch_grp_resid, kg_resid, ResidA)
GRP_LAUNCH(launch_grp_fill, kg_fill, FillA)

static_assert(GRP_BS == 256, "the edge and lines passes run 256-thread workgroups");

// m + p workgroups per problem
__global__ __launch_bounds__(GRP_BS) void kg_edge_gemv(const EdgeA *args, const int *off, int n) {
    const int p = grp_find(off, n, XB);
    const EdgeA A = args[p];
    edge_gemv_pair(A.d, A.Sm, A.Sp, A.bc, A.W, (XB - off[p]) * (GRP_BS / 64) + (threadIdx.x >> 6));
}

__global__ __launch_bounds__(GRP_BS) void kg_spectral_scale(const SpecA *args, const int *off,
                                                            int n) {
    GRP_PROLOGUE(SpecA, A.d.ng)
    (void)b;  // one batch entry per problem
    if (t >= d.ng) return;
    const double a = A.sub ? A.A[t] - A.sub[t] : A.A[t];
    double v = A.kappa * a * A.D[t];
    if (A.EW) v += A.D[t] * edge_term(d, A.Sm, A.Sp, A.EW, imod(t, d.m), idiv(t, d.m));
    A.F[t] = v;
}

long grp_lines_blocks(const Dom &d, int pass) {
    if (pass == 0) return nblk(d.m, 256) * (long)LN_CH + nblk(d.p, LN_OUTS);  // row sums
    if (pass == 1) return nblk(d.m, 4);
    return nblk(d.nx / 2 + d.ny / 2, LN_OUTS);  // mirrored output pairs
}

__global__ __launch_bounds__(GRP_BS) void kg_lines_part(const LinesA *args, const int *off, int n) {
    const int p = grp_find(off, n, XB);
    const LinesA A = args[p];
    const int blk = XB - off[p], nca = (int)nblk(A.d.m, 256);
    if (blk < nca * LN_CH)
        lines_colpart(A.d, A.F, 0, A.Sm, A.Sp, A.EW, A.D, A.W, 0, (blk % nca) * 256 + threadIdx.x,
                      blk / nca, 0);
    else
        lines_row(A.d, A.F, 0, A.Sm, A.Sp, A.EW, A.D, A.W, 0,
                  (blk - nca * LN_CH) * LN_OUTS + (threadIdx.x >> 6), 0);
}

__global__ __launch_bounds__(GRP_BS) void kg_lines_colsum(const LinesA *args, const int *off,
                                                          int n) {
    const int p = grp_find(off, n, XB);
    const LinesA A = args[p];
    lines_colsum(A.d, A.W, (XB - off[p]) * (GRP_BS / 64) + (threadIdx.x >> 6));
}

__global__ __launch_bounds__(GRP_BS) void kg_lines_out(const LinesA *args, const int *off, int n) {
    const int p = grp_find(off, n, XB);
    const LinesA A = args[p];
    lines_out(A.d, A.Sm, A.Sp, A.W, A.psi, (XB - off[p]) * LN_OUTS + (threadIdx.x >> 6));
}

__global__ __launch_bounds__(GRP_BS) void kg_axpy_sub(const AxpyA *args, const int *off, int n) {
    const int p = grp_find(off, n, XB);
    const AxpyA A = args[p];
    const long t = (long)(XB - off[p]) * GRP_BS + threadIdx.x;
    if (t < A.n) A.y[t] -= A.x[t] * A.a;
}
#undef XB

GRP_LAUNCH(launch_grp_edge_gemv, kg_edge_gemv, EdgeA)
GRP_LAUNCH(launch_grp_spectral_scale, kg_spectral_scale, SpecA)
GRP_LAUNCH(launch_grp_lines_part, kg_lines_part, LinesA)
GRP_LAUNCH(launch_grp_lines_colsum, kg_lines_colsum, LinesA)
GRP_LAUNCH(launch_grp_lines_out, kg_lines_out, LinesA)
GRP_LAUNCH(launch_grp_axpy_sub, kg_axpy_sub, AxpyA)


// ------------------------------------------------------------------ adjoint transfers
// Adjoints of the multilevel transfers (exact transposes of the maps above, same index sets),
// for the coarse-level part of the B-product and its adjoint (see ibh_api.hip).

// bc = edgesᵀ(f): the transpose of apply_bc!'s edge addition (fluid-ops.jl:249-275): each bc
// entry that apply_bc! adds to a field entry reads that entry; the others are zero.
__global__ void k_edges_adj(Dom d, const double *__restrict__ f, long sf,
                            double *__restrict__ bc, long sbc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.nbc) return;
    const double *F = f + blockIdx.y * sf;
    const int k = t + 1;  // 1-based bc index
    double v = 0.0;
    if (k > d.To) {
        const int i = k - d.To - 1;  // BC((To + 1) + i), i in 1..m
        if (i >= 1 && i <= d.m) v = GI(F, i, d.p);
    } else if (k > d.Bo) {
        const int i = k - d.Bo - 1;
        if (i >= 1 && i <= d.m) v = GI(F, i, 1);
    } else if (k > d.Ro) {
        const int j = k - d.Ro - 1;
        if (j >= 1 && j <= d.p) v = GI(F, d.m, j);
    } else {
        const int j = k - d.Lo - 1;
        if (j >= 1 && j <= d.p) v = GI(F, 1, j);
    }
    bc[blockIdx.y * sbc + t] = v;
}

void launch_edges_adj(const Dom &d, const double *f, long sf, double *bc, long sbc, int nbatch,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_edges_adj, dim3(nblk(d.nbc, 256), nbatch), dim3(256), 0, s, d, f, sf, bc,
                       sbc);
}

// a = get_bcᵀ(bc) on the whole coarse field (zero off the four lines get_bc! reads): a line
// vertex t steps from the line's first vertex receives 0.25·bc(2t+1) from the copied value and
// 0.125 from each averaged neighbour bc(2t), bc(2t+2) (fluid-ops.jl:210-239).
__device__ inline double gbc_adj_line(const double *__restrict__ bc, int off, int t, int half) {
    if (t < 0 || t > half) return 0.0;
    double v = 0.25 * BC(bc, off + 2 * t + 1);
    if (t >= 1) v += 0.125 * BC(bc, off + 2 * t);
    if (t + 1 <= half) v += 0.125 * BC(bc, off + 2 * t + 2);
    return v;
}

__global__ void k_get_bc_adj(Dom d, const double *__restrict__ bc, long sbc,
                             double *__restrict__ a, long sa) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const double *B = bc + blockIdx.y * sbc;
    const int i = imod(t, d.m) + 1, j = idiv(t, d.m) + 1;
    const int nx = d.nx, ny = d.ny;
    double v = 0.0;
    if (j == ny / 4) v += gbc_adj_line(B, d.Bo, i - nx / 4, nx / 2);
    if (j == 3 * ny / 4) v += gbc_adj_line(B, d.To, i - nx / 4, nx / 2);
    if (i == nx / 4) v += gbc_adj_line(B, d.Lo, j - ny / 4, ny / 2);
    if (i == 3 * nx / 4) v += gbc_adj_line(B, d.Ro, j - ny / 4, ny / 2);
    a[blockIdx.y * sa + t] = v;
}

void launch_get_bc_adj(const Dom &d, const double *bc, long sbc, double *a, long sa, int nbatch,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_get_bc_adj, dim3(nblk(d.ng, 256), nbatch), dim3(256), 0, s, d, bc, sbc,
                       a, sa);
}

// out = base + coarsifyᵀ(y): fine vertex i receives the coarse centre values whose stencil
// (weights 1, ½, ¼ as products of 1 and ½ per dimension) covers it.  base may be null.
__device__ inline int coarsify_adj_taps(int i, int n, int *a, double *w) {
    int c = 0;
    if ((i & 1) == 0) {
        const int k = (i - 2) / 2;
        if (k >= 0 && k < n) { a[c] = k; w[c++] = 1.0; }
    } else {
        const int k0 = (i - 3) / 2, k1 = (i - 1) / 2;
        if (i >= 3 && k0 < n) { a[c] = k0; w[c++] = 0.5; }
        if (k1 >= 0 && k1 < n) { a[c] = k1; w[c++] = 0.5; }
    }
    return c;
}

__global__ void k_coarsify_adj(Dom d, const double *__restrict__ y, long sy,
                               const double *__restrict__ base, long sbase,
                               double *__restrict__ out, long sout) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.ng) return;
    const double *Y = y + blockIdx.y * sy;
    const int i = imod(t, d.m) + 1, j = idiv(t, d.m) + 1;
    int ai[2], bj[2];
    double wi[2], wj[2];
    const int ni = coarsify_adj_taps(i, d.nx / 2 - 1, ai, wi);
    const int nj = coarsify_adj_taps(j, d.ny / 2 - 1, bj, wj);
    double v = 0.0;
    for (int q = 0; q < nj; ++q)
        for (int p = 0; p < ni; ++p)
            v += (wi[p] * wj[q]) * GI(Y, d.nx / 4 + 1 + ai[p], d.ny / 4 + 1 + bj[q]);
    out[blockIdx.y * sout + t] = (base ? base[blockIdx.y * sbase + t] : 0.0) + v;
}

void launch_coarsify_adj(const Dom &d, const double *y, long sy, const double *base, long sbase,
                         double *out, long sout, int nbatch, hipStream_t s) {
    hipLaunchKernelGGL(k_coarsify_adj, dim3(nblk(d.ng, 256), nbatch), dim3(256), 0, s, d, y, sy,
                       base, sbase, out, sout);
}

// out = (a − b) on the flux faces with 1-based indices [i0, i1] × [j0, j1] (clipped to each
// component's extent), zero elsewhere; b may be null
__global__ void k_box_diff(Dom d, const double *__restrict__ a, const double *__restrict__ b,
                           long sq, double *__restrict__ out, int i0, int i1, int j0, int j1) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= d.nq) return;
    int i, j;
    if (t < d.nu) {
        i = imod(t, (d.nx + 1)) + 1;
        j = idiv(t, (d.nx + 1)) + 1;
    } else {
        i = (int)((t - d.nu) % d.nx) + 1;
        j = (int)((t - d.nu) / d.nx) + 1;
    }
    const long o = blockIdx.y * sq + t;
    const bool in = i >= i0 && i <= i1 && j >= j0 && j <= j1;
    out[o] = in ? a[o] - (b ? b[o] : 0.0) : 0.0;
}

void launch_box_diff(const Dom &d, const double *a, const double *b, long sq, double *out,
                     int nbatch, int i0, int i1, int j0, int j1, hipStream_t s) {
    hipLaunchKernelGGL(k_box_diff, dim3(nblk(d.nq, 256), nbatch), dim3(256), 0, s, d, a, b, sq,
                       out, i0, i1, j0, j1);
}

// p[b·stride + i] for i < n: a uniform value in [−1, 1) from a counter-based hash of (seed, b, i)
// (splitmix64), on the vertex box [xa, xb] × [ya, yb] (0-based, m-wide rows), zero elsewhere
__global__ void k_fill_random_box(double *p, long stride, int m, int p_, unsigned long long seed,
                                  int xa, int xb, int ya, int yb) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)m * p_) return;
    const int i = imod(t, m), j = idiv(t, m);
    double v = 0.0;
    if (i >= xa && i <= xb && j >= ya && j <= yb) {
        unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (1ull + t + (unsigned long long)blockIdx.y * (unsigned long long)m * p_);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        v = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    p[blockIdx.y * stride + t] = v;
}

void launch_fill_random_box(const Dom &d, double *p, long stride, int count,
                            unsigned long long seed, int xa, int xb, int ya, int yb,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_fill_random_box, dim3(nblk(d.ng, 256), count), dim3(256), 0, s, p,
                       stride, d.m, d.p, seed, xa, xb, ya, yb);
}

}  // namespace ibh
