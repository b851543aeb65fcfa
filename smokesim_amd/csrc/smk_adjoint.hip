// smk_adjoint.hip -- the differentiable step and its hand-written adjoint.
//
// The differentiable step is the reference's step (physics.c:72-101) on FP32
// fields with one density channel, a per-cell density source in place of the
// emitters, and a converged projection:
//   rho1 = rho + src*dt                        (source, on a box of cells)
//   v1   = advect(v; v)                        advect_vx.comp:33-78
//   v2   = v1 + e_y * (-alpha*rho1(face))*dt   apply_buoyancy.comp:39-52
//   w    = curl(v2); v3 = v2 + confine(w)      calculate_vorticity.comp:59-84,
//                                              apply_vorticity_x.comp:43-63
//   v4   = v3 - dt/(rho h) grad A^-1 b(v3)     update_vx.comp:45-63 (FP64 refined)
//   rho2 = advect(rho1; v4)                    advect_smoke.comp:55-72
// The temperature stays at t_ambient, so the beta term of the buoyancy is
// zero.  Each formula is the forward kernel's; these kernels add FP32
// storage of the vorticity and of the density.
//
// Adjoint scatters (the transposed trilinear gathers) accumulate into 64-bit
// fixed-point integers: an atomic integer sum does not depend on the order
// of the atomics, so gradients are bit-reproducible.  Each scatter call gets
// a power-of-two scale from the largest input magnitude (k_maxabs), keeping
// 2^36 of resolution on the largest contribution and 2^27 of headroom.
#include <hip/hip_runtime.h>

#include "smk_device.h"
#include "smk_kernels.h"

namespace smk {

// ---------------------------------------------------------------------------
// fixed point
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fx_add(unsigned long long *acc, uint32_t idx, double v, double s) {
    const long long q = __double2ll_rn(v * s);
    if (q != 0) atomicAdd(acc + idx, (unsigned long long)q);
}

// Scale of one accumulator: the largest contribution is bounded by
// top = max|lambda| * max(c0, c1 * max|field|) (weights <= 1); scale =
// 2^(36 - e) with top < 2^e.  Maxima arrive as float bit patterns.
__global__ void k_fx_scale(const unsigned int *m_lam, const unsigned int *m_field, float c0, float c1,
                           double *scale) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double lam = (double)__uint_as_float(*m_lam);
    const double fld = m_field ? (double)__uint_as_float(*m_field) : 0.0;
    const double top = lam * fmax((double)c0, (double)c1 * fld);
    int e = 0;
    if (top > 0.0) (void)frexp(top, &e);
    *scale = ldexp(1.0, 36 - e);
}

__global__ void __launch_bounds__(256) k_maxabs(const float *x, size_t n, unsigned int *out) {
    __shared__ float red[4];
    float m = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(x[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    // one atomic per block; non-negative floats order like their bit patterns
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        if (m > 0.f) atomicMax(out, __float_as_uint(m));
    }
}

__global__ void k_track_residual(const double *norms, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double r2 = norms[0], b2 = norms[1];
    const double rel = b2 > 0.0 ? sqrt(r2 / b2) : (r2 > 0.0 ? INFINITY : 0.0);
    if (!(rel <= *out)) *out = rel;  // NaN propagates
}

// ---------------------------------------------------------------------------
// source fitting on the device: loss, gradient norm and the projected Polyak
// step.  Sums are per-block partials over a fixed grid of kFitBlocks blocks,
// added in a fixed order, so every figure is reproducible run to run.
// ---------------------------------------------------------------------------
constexpr int kFitBlocks = 256;
static_assert(kFitBlocks == kFitPartials, "one partial sum per block");

// Sum of one double per thread of a 256-thread block, in a fixed order; the
// result is valid in thread 0.
__device__ __forceinline__ double block_sum256(double v) {
    __shared__ double red[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) t = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();
    return t;
}

// d = rho - target in double, rb = (float)(2 d) (dL/drho of L = sum d^2),
// part[block] = the block's sum of d^2.
template <int F>
__global__ void __launch_bounds__(256) k_fit_diff(const void *rho, const float *tgt, size_t n, float *rb,
                                                  double *part) {
    using S = Store<F>;
    const typename S::T *r = static_cast<const typename S::T *>(rho);
    double sq = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double d = (double)S::ld(r[i]) - (double)tgt[i];
        rb[i] = (float)(2.0 * d);
        sq += d * d;
    }
    const double t = block_sum256(sq);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// part[block] = the block's sum of x^2 (x in double)
__global__ void __launch_bounds__(256) k_fit_sumsq(const float *x, size_t n, double *part) {
    double sq = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double v = (double)x[i];
        sq += v * v;
    }
    const double t = block_sum256(sq);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// L = the sum of the loss partials; losses[*cnt] = L and ++*cnt (one thread)
__global__ void k_fit_record(const double *part, double *L, double *losses, int cap, int *cnt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (int b = 0; b < kFitBlocks; ++b) t += part[b];
    *L = t;
    const int c = *cnt;
    if (c < cap) losses[c] = t;
    *cnt = c + 1;
}

// src <- max(0, src - (L / (2 gg)) g), gg = |g|^2 from its partials (summed
// by every block in the same order); no step when gg == 0.
__global__ void __launch_bounds__(256) k_fit_update(float *src, const float *g, size_t n, const double *L,
                                                    const double *part_g) {
    __shared__ double sgg;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int b = 0; b < kFitBlocks; ++b) t += part_g[b];
        sgg = t;
    }
    __syncthreads();
    const double gg = sgg;
    if (!(gg > 0.0)) return;
    const double step = 0.5 * *L / gg;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double v = (double)src[i] - step * (double)g[i];
        src[i] = (float)(v > 0.0 ? v : 0.0);
    }
}

// dst = base + acc / scale (base may alias dst or be null)
__global__ void k_fx_to_f32(const unsigned long long *acc, const double *scale, const float *base,
                            float *dst, size_t n) {
    const double inv = 1.0 / *scale;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = (double)(long long)acc[i] * inv;
        dst[i] = base ? (float)((double)base[i] + v) : (float)v;
    }
}

// ---------------------------------------------------------------------------
// trilinear helpers (GL LINEAR, x then y then z)
// ---------------------------------------------------------------------------
struct Tri {
    int i0[3];
    float a[3];
};

// corner q (bit d = +1 along axis d): texel index validity and weight
__device__ __forceinline__ float tri_w(const Tri &t, int q, int D) {
    float w = (q & 1) ? t.a[0] : 1.0f - t.a[0];
    w *= (q & 2) ? t.a[1] : 1.0f - t.a[1];
    if (D == 3) w *= (q & 4) ? t.a[2] : 1.0f - t.a[2];
    return w;
}

// d(sample)/d(a_d) from the 8 (4) corner values c[q]
__device__ __forceinline__ float tri_da(const Tri &t, const float *c, int d, int D) {
    float s = 0.f;
    const int nq = D == 3 ? 8 : 4;
    for (int q = 0; q < nq; ++q) {
        if (q & (1 << d)) continue;
        float w = 1.0f;
        for (int e = 0; e < D; ++e) {
            if (e == d) continue;
            w *= (q & (1 << e)) ? t.a[e] : 1.0f - t.a[e];
        }
        s += w * (c[q | (1 << d)] - c[q]);
    }
    return s;
}

template <int D, int F>
__device__ __forceinline__ void tri_corners(const Tri &t, const void *texv, int sx, int sy, int sz,
                                            float border, float *c, uint32_t *idx, bool *ok) {
    const typename Store<F>::T *tex = static_cast<const typename Store<F>::T *>(texv);
    const int nq = D == 3 ? 8 : 4;
    for (int q = 0; q < nq; ++q) {
        const int ii = t.i0[0] + (q & 1), jj = t.i0[1] + ((q >> 1) & 1),
                  kk = D == 3 ? t.i0[2] + ((q >> 2) & 1) : 0;
        ok[q] = inside(ii, jj, kk, sx, sy, sz);
        idx[q] = ok[q] ? lin(ii, jj, kk, sx, sy) : 0u;
        c[q] = ok[q] ? Store<F>::ld(tex[idx[q]]) : border;
    }
}

// The face-centred velocity interpolation of interp_velocity_tab as
// explicit weights: component cc at the position of face c (c = -1: cell
// centre) of index id.
template <int D>
__device__ __forceinline__ Tri interp_tri(const CoordTab &T, const int *id, int c, int cc) {
    Tri t;
    t.i0[2] = 0;
    t.a[2] = 0.f;
    for (int d = 0; d < D; ++d) {
        const TC e = ctab(T, d, (d == c ? 2 : 0) + (d == cc ? 1 : 0), id[d]);
        t.i0[d] = e.i0;
        t.a[d] = e.a;
    }
    return t;
}

// ---------------------------------------------------------------------------
// differentiable forward kernels
// ---------------------------------------------------------------------------
template <int F>
__global__ void k_d_source(void *rhov, const float *src, int nx, int ny, Box3 box, float dt) {
    using S = Store<F>;
    typename S::T *rho = static_cast<typename S::T *>(rhov);
    const int sx = box.hi[0] - box.lo[0], sy = box.hi[1] - box.lo[1], sz = box.hi[2] - box.lo[2];
    const size_t n = (size_t)sx * sy * sz;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % sx), j = (int)((c / sx) % sy), k = (int)(c / ((size_t)sx * sy));
        const uint32_t cell = lin(box.lo[0] + i, box.lo[1] + j, box.lo[2] + k, nx, ny);
        rho[cell] = S::st(S::ld(rho[cell]) + src[c] * dt);
    }
}

// density sample of apply_buoyancy.comp at the vy face (i, j, k)
template <int D>
__device__ __forceinline__ Tri buoy_tri(const SimParams &P, int i, int j, int k) {
    Tri t;
    float uv[3] = {((float)i + 0.5f) / (float)P.n[0], ((float)j + 0.0f) / (float)P.n[1],
                   ((float)k + 0.5f) / (float)P.n[2]};
    texcoord(uv[0], P.n[0], t.i0[0], t.a[0]);
    texcoord(uv[1], P.n[1], t.i0[1], t.a[1]);
    t.i0[2] = 0;
    t.a[2] = 0.f;
    if (D == 3) texcoord(uv[2], P.n[2], t.i0[2], t.a[2]);
    return t;
}

template <int D>
__device__ __forceinline__ bool buoy_face(const SimParams &P, const uint8_t *f0, int i, int j, int k) {
    return !(solid_flag(P, f0, i, j, k, false) || solid_flag(P, f0, i, j - 1, k, false));
}

template <int D, int F>
__global__ void k_d_buoyancy(XG X_, SimParams P, const void *vy_inv, void *vy_outv, const uint8_t *f0,
                             const void *rho) {
    using S = Store<F>;
    const typename S::T *vy_in = static_cast<const typename S::T *>(vy_inv);
    typename S::T *vy_out = static_cast<typename S::T *>(vy_outv);
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    int sx, sy, sz;
    vsize<D>(P, 1, sx, sy, sz);
    if (!inside(i, j, k, sx, sy, sz)) return;
    const uint32_t f = lin(i, j, k, sx, sy);
    float nv = S::ld(vy_in[f]);
    if (buoy_face<D>(P, f0, i, j, k)) {
        const Tri t = buoy_tri<D>(P, i, j, k);
        float c[8];
        uint32_t idx[8];
        bool ok[8];
        tri_corners<D, F>(t, rho, P.n[0], P.n[1], P.n[2], 0.0f, c, idx, ok);
        float s;
        if (D == 2) {
            s = lerpf(lerpf(c[0], c[1], t.a[0]), lerpf(c[2], c[3], t.a[0]), t.a[1]);
        } else {
            float y0 = lerpf(lerpf(c[0], c[1], t.a[0]), lerpf(c[2], c[3], t.a[0]), t.a[1]);
            float y1 = lerpf(lerpf(c[4], c[5], t.a[0]), lerpf(c[6], c[7], t.a[0]), t.a[1]);
            s = lerpf(y0, y1, t.a[2]);
        }
        const float b = -P.alpha * s;
        nv = nv + b * P.dt;
    }
    vy_out[f] = S::st(nv);
}

// calculate_vorticity.comp 3D with the vorticity kept in FP32 (x, y, z, 0)
template <int F>
__device__ __forceinline__ float vgetF(const SimParams &P, const VelPtrs &V, int c, int i, int j, int k) {
    using S = Store<F>;
    int sx, sy, sz;
    vsize<3>(P, c, sx, sy, sz);
    const bool ok = inside(i, j, k, sx, sy, sz);
    return S::ld(ld_sel(static_cast<const typename S::T *>(V.c[c]), ok, lin(i, j, k, sx, sy), (typename S::T)0));
}
template <int F>
__device__ __forceinline__ float ccvF(const SimParams &P, const VelPtrs &V, int c, int i, int j, int k) {
    return (vgetF<F>(P, V, c, i, j, k) + vgetF<F>(P, V, c, i + (c == 0), j + (c == 1), k + (c == 2))) * 0.5f;
}

// (2D: calculate_vorticity.comp 2D:51-69, the scalar curl in .x)
template <int D, int F>
__global__ void k_d_vorticity(XG X_, SimParams P, VelPtrs V, const uint8_t *f0, float4 *vort) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    const uint32_t cell = lin(i, j, k, P.n[0], P.n[1]);
    const float h2 = 2.0f * P.h;
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if (D == 2 && !f0[cell]) {
        float z1 = ccvF<F>(P, V, 1, i + 1, j, k) - ccvF<F>(P, V, 1, i - 1, j, k);
        float z2 = ccvF<F>(P, V, 0, i, j + 1, k) - ccvF<F>(P, V, 0, i, j - 1, k);
        o.x = (z1 - z2) / h2;
    } else if (D == 3 && !f0[cell]) {
        float x1 = ccvF<F>(P, V, 2, i, j + 1, k) - ccvF<F>(P, V, 2, i, j - 1, k);
        float x2 = ccvF<F>(P, V, 1, i, j, k + 1) - ccvF<F>(P, V, 1, i, j, k - 1);
        o.x = (x1 - x2) / h2;
        float y1 = ccvF<F>(P, V, 0, i, j, k + 1) - ccvF<F>(P, V, 0, i, j, k - 1);
        float y2 = ccvF<F>(P, V, 2, i + 1, j, k) - ccvF<F>(P, V, 2, i - 1, j, k);
        o.y = (y1 - y2) / h2;
        float z1 = ccvF<F>(P, V, 1, i + 1, j, k) - ccvF<F>(P, V, 1, i - 1, j, k);
        float z2 = ccvF<F>(P, V, 0, i, j + 1, k) - ccvF<F>(P, V, 0, i, j - 1, k);
        o.z = (z1 - z2) / h2;
    }
    vort[cell] = o;
}

__device__ __forceinline__ float4 w_at(const SimParams &P, const float4 *w, int i, int j, int k) {
    const bool ok = inside(i, j, k, P.n[0], P.n[1], P.n[2]);
    return ok ? w[lin(i, j, k, P.n[0], P.n[1])] : make_float4(0.f, 0.f, 0.f, 0.f);
}
__device__ __forceinline__ float w_mag(const SimParams &P, const float4 *w, int i, int j, int k) {
    const float4 o = w_at(P, w, i, j, k);
    return sqrtf(o.x * o.x + o.y * o.y + o.z * o.z);
}

// apply_vorticity_{x,y,z}.comp 3D: the confinement force of cell (i,j,k)
// drives the three faces with index (i,j,k).  2D (apply_vorticity_{x,y}.comp
// 2D:41-60): |w| = fabs(w.x), F = (-n_y w, -n_x w).
template <int D, int F>
__global__ void k_d_confine(XG X_, SimParams P, VelPtrs in, VelPtrs out, const float4 *w) {
    using S = Store<F>;
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (i > P.n[0] || j > P.n[1] || k > P.n[2] - (D == 2 ? 1 : 0)) return;
    const float h2 = 2.0f * P.h;
    float f[3];
    if (D == 2) {
        float dx = fabsf(w_at(P, w, i + 1, j, k).x) - fabsf(w_at(P, w, i - 1, j, k).x);
        float dy = fabsf(w_at(P, w, i, j + 1, k).x) - fabsf(w_at(P, w, i, j - 1, k).x);
        float om = w_at(P, w, i, j, k).x;
        float gx = dx / h2, gy = dy / h2;
        float len = sqrtf(gx * gx + gy * gy) + 1e-20f;
        float nx = gx / len, ny = gy / len;
        f[0] = -ny * om;
        f[1] = -nx * om;
        f[2] = 0.f;
    } else {
        float dx = w_mag(P, w, i + 1, j, k) - w_mag(P, w, i - 1, j, k);
        float dy = w_mag(P, w, i, j + 1, k) - w_mag(P, w, i, j - 1, k);
        float dz = w_mag(P, w, i, j, k + 1) - w_mag(P, w, i, j, k - 1);
        float4 om = w_at(P, w, i, j, k);
        float gx = dx / h2, gy = dy / h2, gz = dz / h2;
        float len = sqrtf(gx * gx + gy * gy + gz * gz) + 1e-20f;
        float nx = gx / len, ny = gy / len, nz = gz / len;
        f[0] = ny * om.z - nz * om.y;
        f[1] = nz * om.x - nx * om.z;
        f[2] = nx * om.y - ny * om.x;
    }
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        if (!inside(i, j, k, sx, sy, sz)) continue;
        const uint32_t fi = lin(i, j, k, sx, sy);
        const float cur = S::ld(static_cast<const typename S::T *>(in.c[c])[fi]);
        static_cast<typename S::T *>(out.c[c])[fi] = S::st(cur + (f[c] * 0.5f * P.eps * P.h * P.dt) / P.rho);
    }
}

// advect_smoke.comp for one FP32 density channel: border 0, solid cells 0.
template <int D, int F>
__device__ __forceinline__ Tri scalar_departure(const SimParams &P, const VelPtrs &V, const CoordTab &T,
                                                const int *id, float *u) {
    interp_velocity_tab<D, F>(P, V, T, id, -1, u);
    Tri t;
    t.i0[2] = 0;
    t.a[2] = 0.f;
    for (int d = 0; d < D; ++d) {
        float world = ((float)id[d] + 0.5f) * P.h;
        float from = world - u[d] * P.dt;
        float gs = from / P.h;
        texcoord(gs / (float)P.n[d], P.n[d], t.i0[d], t.a[d]);
    }
    return t;
}

template <int D, int F>
__global__ void k_d_advect_scalar(XG X_, SimParams P, VelPtrs V, const uint8_t *f0, CoordTab T,
                                  const void *rin, void *rout) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    const uint32_t cell = lin(i, j, k, P.n[0], P.n[1]);
    float o = 0.0f;
    if (!f0[cell]) {
        const int id[3] = {i, j, k};
        float u[3] = {0.f, 0.f, 0.f};
        const Tri t = scalar_departure<D, F>(P, V, T, id, u);
        float c[8];
        uint32_t idx[8];
        bool ok[8];
        tri_corners<D, F>(t, rin, P.n[0], P.n[1], P.n[2], 0.0f, c, idx, ok);
        if (D == 2) {
            o = lerpf(lerpf(c[0], c[1], t.a[0]), lerpf(c[2], c[3], t.a[0]), t.a[1]);
        } else {
            float y0 = lerpf(lerpf(c[0], c[1], t.a[0]), lerpf(c[2], c[3], t.a[0]), t.a[1]);
            float y1 = lerpf(lerpf(c[4], c[5], t.a[0]), lerpf(c[6], c[7], t.a[0]), t.a[1]);
            o = lerpf(y0, y1, t.a[2]);
        }
    }
    static_cast<typename Store<F>::T *>(rout)[cell] = Store<F>::st(o);
}

// ---------------------------------------------------------------------------
// adjoint kernels
// ---------------------------------------------------------------------------
struct Acc3 {
    unsigned long long *c[3];
};

// Departure-point velocity adjoints ubar[cc] per source position, kept for
// the transposed interpolation (k_a_interp_gather) instead of scattering
// them: u[c][cc] for face positions of component c (c = 3: cell centres).
struct UBar {
    float *u[4][3];
};

template <int D>
__device__ __forceinline__ void put_ubar(const UBar &U, int c, uint32_t pos, const float *ubar) {
    for (int cc = 0; cc < D; ++cc) U.u[c < 0 ? 3 : c][cc][pos] = ubar[cc];
}

// Adjoint of the density advection: rbar1 += T^T rbar2 (acc_r), and the
// departure-point term of the velocity: ubar_d = lambda * dS/da_d * (-dt/h),
// scattered to the faces (acc_v).
template <int D, int F>
__global__ void k_a_advect_scalar(XG X_, SimParams P, VelPtrs V, const uint8_t *f0, CoordTab T,
                                  const void *rin, const float *rbar_out, unsigned long long *acc_r,
                                  const double *s_r, UBar U) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    const uint32_t cell = lin(i, j, k, P.n[0], P.n[1]);
    const float lam = rbar_out[cell];
    if (f0[cell] || lam == 0.0f) {
        const float z[3] = {0.f, 0.f, 0.f};
        put_ubar<D>(U, -1, cell, z);
        return;
    }
    const int id[3] = {i, j, k};
    float u[3] = {0.f, 0.f, 0.f};
    const Tri t = scalar_departure<D, F>(P, V, T, id, u);
    float c[8];
    uint32_t idx[8];
    bool ok[8];
    tri_corners<D, F>(t, rin, P.n[0], P.n[1], P.n[2], 0.0f, c, idx, ok);
    const double sr = *s_r;
    const int nq = D == 3 ? 8 : 4;
    for (int q = 0; q < nq; ++q)
        if (ok[q]) fx_add(acc_r, idx[q], (double)(tri_w(t, q, D) * lam), sr);
    float ubar[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < D; ++d) ubar[d] = lam * tri_da(t, c, d, D) * (-P.dt / P.h);
    put_ubar<D>(U, -1, cell, ubar);
}

// Adjoint of the velocity advection (advect_vx.comp:33-78), faces with an
// advected value only (wind, solid and out-of-domain faces are constants).
template <int D, int F>
__global__ void k_a_advect_velocity(XG X_, SimParams P, VelPtrs V, const uint8_t *f0, CoordTab T,
                                    VelPtrs vbar_out, Acc3 acc, const double *s_acc, UBar U) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    const int id[3] = {i, j, k};
    const bool here_in = inside(i, j, k, P.n[0], P.n[1], P.n[2]);
    const bool here_s = here_in ? f0[lin(i, j, k, P.n[0], P.n[1])] != 0 : false;
    const double si = *s_acc;
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        if (!inside(i, j, k, sx, sy, sz)) continue;
        const uint32_t f = lin(i, j, k, sx, sy);
        const float lam = static_cast<const float *>(vbar_out.c[c])[f];
        const int e0 = c == 0, e1 = c == 1, e2 = c == 2;
        if (lam == 0.0f || (id[c] == 0 && P.wind[c] != 0.0f) || !here_in || here_s ||
            solid_flag(P, f0, i - e0, j - e1, k - e2, true)) {
            const float z[3] = {0.f, 0.f, 0.f};
            put_ubar<D>(U, c, f, z);
            continue;
        }
        float v[3] = {0.f, 0.f, 0.f};
        interp_velocity_tab<D, F>(P, V, T, id, c, v);
        Tri t;
        t.i0[2] = 0;
        t.a[2] = 0.f;
        for (int d = 0; d < D; ++d) {
            float world = (d == c ? (float)id[d] : (float)id[d] + 0.5f) * P.h;
            float from = world - v[d] * P.dt;
            float gs = from / P.h;
            float uvd = (gs + (d == c ? 0.5f : 0.0f)) / ((float)P.n[d] + (d == c ? 1.0f : 0.0f));
            texcoord(uvd, P.n[d] + (d == c ? 1 : 0), t.i0[d], t.a[d]);
        }
        float cv[8];
        uint32_t idx[8];
        bool ok[8];
        tri_corners<D, F>(t, V.c[c], sx, sy, sz, 0.0f, cv, idx, ok);
        const int nq = D == 3 ? 8 : 4;
        for (int q = 0; q < nq; ++q)
            if (ok[q]) fx_add(acc.c[c], idx[q], (double)(tri_w(t, q, D) * lam), si);
        float ubar[3] = {0.f, 0.f, 0.f};
        for (int d = 0; d < D; ++d) ubar[d] = lam * tri_da(t, cv, d, D) * (-P.dt / P.h);
        put_ubar<D>(U, c, f, ubar);
    }
}

// ---------------------------------------------------------------------------
// LDS-binned scatters (3D)
//
// A block owns a kBinTX x kBinTY tile of one target lattice (the faces of one
// velocity component, or the cells) and marches kBinKZ planes in z.  A
// departure point within one cell of its position puts its corners within
// one index of it on every axis, so while plane k is processed every corner
// near the tile lies in planes k-1..k+1 of an (kBinTX+2) x (kBinTY+2) window:
// a ring of three planes of int64 bins in LDS.  Corners outside the window
// (faster flow) go straight to the global accumulator.  After plane k the
// bins of plane k-1 can receive nothing more from this block: they are added
// to the global accumulator (one atomic per non-zero bin) and cleared for
// plane k+2.  Every contribution is rounded to fixed point on its own, as in
// the direct scatter, and integer sums do not depend on grouping or order,
// so the accumulators (and the gradients) are bit-identical to it.
// ---------------------------------------------------------------------------
constexpr int kBinTX = 16, kBinTY = 16, kBinKZ = 16;
constexpr int kBinW = kBinTX + 2, kBinH = kBinTY + 2, kBinPlane = kBinW * kBinH;

struct BinTile {
    int x0, y0, sx, sy, sz;
};

// Add q for texel (ii, jj, kk) while plane k is processed.
__device__ __forceinline__ void bin_add(unsigned long long *bins, unsigned long long *acc, const BinTile &B,
                                        int ii, int jj, int kk, int k, long long q) {
    const int bx = ii - B.x0 + 1, by = jj - B.y0 + 1, dz = kk - k + 1;
    if ((unsigned)bx < (unsigned)kBinW && (unsigned)by < (unsigned)kBinH && (unsigned)dz < 3u) {
        const int slot = (kk + 3) % 3;  // kk >= -1
        atomicAdd(bins + slot * kBinPlane + by * kBinW + bx, (unsigned long long)q);
    } else {
        atomicAdd(acc + lin(ii, jj, kk, B.sx, B.sy), (unsigned long long)q);
    }
}

// Flush plane kk's bins to the global accumulator and clear them (whole block).
__device__ __forceinline__ void bin_flush(unsigned long long *bins, unsigned long long *acc, const BinTile &B,
                                          int kk) {
    if (kk < -1) return;
    unsigned long long *pl = bins + ((kk + 3) % 3) * kBinPlane;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
    for (int b = tid; b < kBinPlane; b += nt) {
        const unsigned long long v = pl[b];
        if (v != 0ull) {
            const int ii = B.x0 - 1 + b % kBinW, jj = B.y0 - 1 + b / kBinW;
            // bins outside the texture never receive a contribution
            atomicAdd(acc + lin(ii, jj, kk, B.sx, B.sy), v);
            pl[b] = 0ull;
        }
    }
}

// k_a_advect_velocity for component C, binned.  Blocks: kBinTX x kBinTY
// threads; X_ tiles the component's face lattice with z in runs of kBinKZ.
template <int C, int F>
__global__ void __launch_bounds__(256) k_a_advect_velocity_bin(XG X_, SimParams P, VelPtrs V,
                                                               const uint8_t *f0, CoordTab T,
                                                               VelPtrs vbar_out, unsigned long long *acc,
                                                               const double *s_acc, UBar U) {
    constexpr int D = 3;
    __shared__ unsigned long long bins[3 * kBinPlane];
    XG_BLOCK(X_)
    BinTile B;
    vsize<D>(P, C, B.sx, B.sy, B.sz);
    B.x0 = bx_ * kBinTX;
    B.y0 = by_ * kBinTY;
    const int i = B.x0 + (int)threadIdx.x, j = B.y0 + (int)threadIdx.y;
    const int k0 = bz_ * kBinKZ, kend = min(k0 + kBinKZ, B.sz);
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int b = tid; b < 3 * kBinPlane; b += 256) bins[b] = 0ull;
    __syncthreads();
    const double si = *s_acc;
    constexpr int e0 = C == 0, e1 = C == 1, e2 = C == 2;
    for (int k = k0; k < kend; ++k) {
        if (i < B.sx && j < B.sy) {
            const int id[3] = {i, j, k};
            const uint32_t f = lin(i, j, k, B.sx, B.sy);
            const float lam = static_cast<const float *>(vbar_out.c[C])[f];
            const bool here_in = inside(i, j, k, P.n[0], P.n[1], P.n[2]);
            const bool here_s = here_in ? f0[lin(i, j, k, P.n[0], P.n[1])] != 0 : false;
            if (lam == 0.0f || (id[C] == 0 && P.wind[C] != 0.0f) || !here_in || here_s ||
                solid_flag(P, f0, i - e0, j - e1, k - e2, true)) {
                const float z[3] = {0.f, 0.f, 0.f};
                put_ubar<D>(U, C, f, z);
            } else {
                float v[3] = {0.f, 0.f, 0.f};
                interp_velocity_tab<D, F>(P, V, T, id, C, v);
                Tri t;
                for (int d = 0; d < D; ++d) {
                    float world = (d == C ? (float)id[d] : (float)id[d] + 0.5f) * P.h;
                    float from = world - v[d] * P.dt;
                    float gs = from / P.h;
                    float uvd = (gs + (d == C ? 0.5f : 0.0f)) / ((float)P.n[d] + (d == C ? 1.0f : 0.0f));
                    texcoord(uvd, P.n[d] + (d == C ? 1 : 0), t.i0[d], t.a[d]);
                }
                float cv[8];
                uint32_t idx[8];
                bool ok[8];
                tri_corners<D, F>(t, V.c[C], B.sx, B.sy, B.sz, 0.0f, cv, idx, ok);
                for (int q = 0; q < 8; ++q) {
                    if (!ok[q]) continue;
                    const long long fx = __double2ll_rn((double)(tri_w(t, q, D) * lam) * si);
                    if (fx != 0)
                        bin_add(bins, acc, B, t.i0[0] + (q & 1), t.i0[1] + ((q >> 1) & 1),
                                t.i0[2] + ((q >> 2) & 1), k, fx);
                }
                float ubar[3] = {0.f, 0.f, 0.f};
                for (int d = 0; d < D; ++d) ubar[d] = lam * tri_da(t, cv, d, D) * (-P.dt / P.h);
                put_ubar<D>(U, C, f, ubar);
            }
        }
        __syncthreads();
        bin_flush(bins, acc, B, k - 1);
        __syncthreads();
    }
    bin_flush(bins, acc, B, kend - 1);
    bin_flush(bins, acc, B, kend);
}

// k_a_advect_scalar, binned (same tiling over the cells).
template <int F>
__global__ void __launch_bounds__(256) k_a_advect_scalar_bin(XG X_, SimParams P, VelPtrs V, const uint8_t *f0,
                                                             CoordTab T, const void *rin,
                                                             const float *rbar_out, unsigned long long *acc_r,
                                                             const double *s_r, UBar U) {
    constexpr int D = 3;
    __shared__ unsigned long long bins[3 * kBinPlane];
    XG_BLOCK(X_)
    BinTile B{bx_ * kBinTX, by_ * kBinTY, P.n[0], P.n[1], P.n[2]};
    const int i = B.x0 + (int)threadIdx.x, j = B.y0 + (int)threadIdx.y;
    const int k0 = bz_ * kBinKZ, kend = min(k0 + kBinKZ, B.sz);
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int b = tid; b < 3 * kBinPlane; b += 256) bins[b] = 0ull;
    __syncthreads();
    const double sr = *s_r;
    for (int k = k0; k < kend; ++k) {
        if (i < B.sx && j < B.sy) {
            const uint32_t cell = lin(i, j, k, B.sx, B.sy);
            const float lam = rbar_out[cell];
            if (f0[cell] || lam == 0.0f) {
                const float z[3] = {0.f, 0.f, 0.f};
                put_ubar<D>(U, -1, cell, z);
            } else {
                const int id[3] = {i, j, k};
                float u[3] = {0.f, 0.f, 0.f};
                const Tri t = scalar_departure<D, F>(P, V, T, id, u);
                float c[8];
                uint32_t idx[8];
                bool ok[8];
                tri_corners<D, F>(t, rin, B.sx, B.sy, B.sz, 0.0f, c, idx, ok);
                for (int q = 0; q < 8; ++q) {
                    if (!ok[q]) continue;
                    const long long fx = __double2ll_rn((double)(tri_w(t, q, D) * lam) * sr);
                    if (fx != 0)
                        bin_add(bins, acc_r, B, t.i0[0] + (q & 1), t.i0[1] + ((q >> 1) & 1),
                                t.i0[2] + ((q >> 2) & 1), k, fx);
                }
                float ubar[3] = {0.f, 0.f, 0.f};
                for (int d = 0; d < D; ++d) ubar[d] = lam * tri_da(t, c, d, D) * (-P.dt / P.h);
                put_ubar<D>(U, -1, cell, ubar);
            }
        }
        __syncthreads();
        bin_flush(bins, acc_r, B, k - 1);
        __syncthreads();
    }
    bin_flush(bins, acc_r, B, kend - 1);
    bin_flush(bins, acc_r, B, kend);
}

// Transposed face interpolation as a gather: face t of component cc
// collects, from every source position p whose interpolation stencil
// contains t, the weight of t times u[c][cc][p]; sources are the faces of all
// components (cells = 0) or the cell centres (cells = 1).  The stencil is
// separable, so per axis the sources are the positions whose table entry
// has i0 == t (weight 1 - a) or i0 + 1 == t (weight a); the table's i0
// stays within one of idx + off (off = -1 for a face position sampling a
// component not staggered along the axis), which leaves four candidates per
// axis.  out = base + sum, summed in a fixed order.
template <int D>
__global__ void k_a_interp_gather(XG X_, SimParams P, CoordTab T, UBar U, int cells, VelPtrs base,
                                  VelPtrs out) {
    XG_BLOCK(X_)
    const int t[3] = {(int)(bx_ * blockDim.x + threadIdx.x), (int)(by_ * blockDim.y + threadIdx.y),
                      (int)bz_};
    for (int cc = 0; cc < D; ++cc) {
        int tx, ty, tz;
        vsize<D>(P, cc, tx, ty, tz);
        if (!inside(t[0], t[1], t[2], tx, ty, tz)) continue;
        float sum = 0.0f;
        const int c0 = cells ? 3 : 0, c1 = cells ? 4 : D;
        for (int cs = c0; cs < c1; ++cs) {
            const int c = cs == 3 ? -1 : cs;
            int ss[3] = {P.n[0], P.n[1], P.n[2]};
            if (c >= 0) vsize<D>(P, c, ss[0], ss[1], ss[2]);
            // per axis: up to four source indices and their weights
            int pid[3][4];
            float pw[3][4];
            for (int d = 0; d < 3; ++d)
                for (int m = 0; m < 4; ++m) {
                    pid[d][m] = 0;
                    pw[d][m] = 0.0f;
                }
            for (int d = 0; d < D; ++d) {
                const int off = (d == c && d != cc) ? -1 : 0;
                const int qq = (d == c ? 2 : 0) + (d == cc ? 1 : 0);
                for (int m = 0; m < 4; ++m) {
                    const int pp = t[d] - off - 2 + m;
                    pid[d][m] = pp;
                    if (pp < 0 || pp >= ss[d]) continue;
                    const TC e = ctab(T, d, qq, pp);
                    pw[d][m] = e.i0 == t[d] ? 1.0f - e.a : (e.i0 + 1 == t[d] ? e.a : 0.0f);
                }
            }
            if (D == 2) {
                pw[2][0] = 1.0f;
                pid[2][0] = 0;
            }
            const float *u = U.u[cs][cc];
            for (int mz = 0; mz < (D == 3 ? 4 : 1); ++mz) {
                if (pw[2][mz] == 0.0f) continue;
                for (int my = 0; my < 4; ++my) {
                    if (pw[1][my] == 0.0f) continue;
                    for (int mx = 0; mx < 4; ++mx) {
                        if (pw[0][mx] == 0.0f) continue;
                        const float w = pw[0][mx] * pw[1][my] * pw[2][mz];
                        sum += w * u[lin(pid[0][mx], pid[1][my], pid[2][mz], ss[0], ss[1])];
                    }
                }
            }
        }
        const uint32_t f = lin(t[0], t[1], t[2], tx, ty);
        static_cast<float *>(out.c[cc])[f] = static_cast<const float *>(base.c[cc])[f] + sum;
    }
}

// Adjoint of the buoyancy: rbar1 += -alpha*dt * W^T vbar_y (fluid faces).
template <int D>
__global__ void k_a_buoyancy(XG X_, SimParams P, const uint8_t *f0, const float *vybar,
                             unsigned long long *acc_r, const double *s_r) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    int sx, sy, sz;
    vsize<D>(P, 1, sx, sy, sz);
    if (!inside(i, j, k, sx, sy, sz)) return;
    const float lam = vybar[lin(i, j, k, sx, sy)];
    if (lam == 0.0f || !buoy_face<D>(P, f0, i, j, k)) return;
    const Tri t = buoy_tri<D>(P, i, j, k);
    const float g = -P.alpha * P.dt * lam;
    const double s = *s_r;
    const int nq = D == 3 ? 8 : 4;
    for (int q = 0; q < nq; ++q) {
        const int ii = t.i0[0] + (q & 1), jj = t.i0[1] + ((q >> 1) & 1),
                  kk = D == 3 ? t.i0[2] + ((q >> 2) & 1) : 0;
        if (!inside(ii, jj, kk, P.n[0], P.n[1], P.n[2])) continue;
        fx_add(acc_r, lin(ii, jj, kk, P.n[0], P.n[1]), (double)(tri_w(t, q, D) * g), s);
    }
}

// Confinement adjoint, pass 1 (per cell): with Fbar = K * (vbar faces of
// index (i,j,k)), K = 0.5*eps*h*dt/rho, F = n x w:
//   wbar_direct = Fbar x n,  nbar = w x Fbar,
//   gbar = nbar/len - g (g . nbar) / (|g| len^2),  len = |g| + 1e-20.
// 2D: F = (-n_y w, -n_x w) with w scalar (in .x): wbar_direct = -n_y Fbar_x
// - n_x Fbar_y, nbar = (-w Fbar_y, -w Fbar_x), |w| = fabs(w).
template <int D>
__global__ void k_a_confine1(XG X_, SimParams P, const float4 *w, VelPtrs vbar, float4 *wbar,
                             float4 *gbar) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    const uint32_t cell = lin(i, j, k, P.n[0], P.n[1]);
    const float K = 0.5f * P.eps * P.h * P.dt / P.rho;
    float Fb[3] = {0.f, 0.f, 0.f};
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        Fb[c] = K * static_cast<const float *>(vbar.c[c])[lin(i, j, k, sx, sy)];
    }
    const float h2 = 2.0f * P.h;
    if (D == 2) {
        const float gx = (fabsf(w_at(P, w, i + 1, j, k).x) - fabsf(w_at(P, w, i - 1, j, k).x)) / h2;
        const float gy = (fabsf(w_at(P, w, i, j + 1, k).x) - fabsf(w_at(P, w, i, j - 1, k).x)) / h2;
        const float gl = sqrtf(gx * gx + gy * gy);
        const float len = gl + 1e-20f;
        const float nx = gx / len, ny = gy / len;
        const float om = w[cell].x;
        const float nbx = -om * Fb[1], nby = -om * Fb[0];
        float4 gb;
        if (gl > 0.f) {
            const float gn = (gx * nbx + gy * nby) / (gl * len * len);
            gb = make_float4(nbx / len - gx * gn, nby / len - gy * gn, 0.f, 0.f);
        } else {
            gb = make_float4(nbx / len, nby / len, 0.f, 0.f);
        }
        wbar[cell] = make_float4(-ny * Fb[0] - nx * Fb[1], 0.f, 0.f, 0.f);
        gbar[cell] = gb;
        return;
    }
    const float gx = (w_mag(P, w, i + 1, j, k) - w_mag(P, w, i - 1, j, k)) / h2;
    const float gy = (w_mag(P, w, i, j + 1, k) - w_mag(P, w, i, j - 1, k)) / h2;
    const float gz = (w_mag(P, w, i, j, k + 1) - w_mag(P, w, i, j, k - 1)) / h2;
    const float gl = sqrtf(gx * gx + gy * gy + gz * gz);
    const float len = gl + 1e-20f;
    const float nx = gx / len, ny = gy / len, nz = gz / len;
    const float4 om = w[cell];
    float4 wb;
    wb.x = Fb[1] * nz - Fb[2] * ny;
    wb.y = Fb[2] * nx - Fb[0] * nz;
    wb.z = Fb[0] * ny - Fb[1] * nx;
    wb.w = 0.f;
    const float nbx = om.y * Fb[2] - om.z * Fb[1];
    const float nby = om.z * Fb[0] - om.x * Fb[2];
    const float nbz = om.x * Fb[1] - om.y * Fb[0];
    float4 gb;
    if (gl > 0.f) {
        const float gn = (gx * nbx + gy * nby + gz * nbz) / (gl * len * len);
        gb = make_float4(nbx / len - gx * gn, nby / len - gy * gn, nbz / len - gz * gn, 0.f);
    } else {
        gb = make_float4(nbx / len, nby / len, nbz / len, 0.f);
    }
    wbar[cell] = wb;
    gbar[cell] = gb;
}

// pass 2 (per cell): mbar = sum_d (gbar_d(x - e_d) - gbar_d(x + e_d)) / (2h);
// wbar += mbar * w / |w|.
template <int D>
__global__ void k_a_confine2(XG X_, SimParams P, const float4 *w, const float4 *gbar, float4 *wbar) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    const uint32_t cell = lin(i, j, k, P.n[0], P.n[1]);
    auto G = [&](int ii, int jj, int kk) { return w_at(P, gbar, ii, jj, kk); };
    const float h2 = 2.0f * P.h;
    if (D == 2) {
        const float mb = ((G(i - 1, j, k).x - G(i + 1, j, k).x) + (G(i, j - 1, k).y - G(i, j + 1, k).y)) / h2;
        const float om = w[cell].x;
        const float m = fabsf(om);
        if (m > 0.f && mb != 0.f) {
            float4 wb = wbar[cell];
            wb.x += mb * om / m;
            wbar[cell] = wb;
        }
        return;
    }
    const float mb = ((G(i - 1, j, k).x - G(i + 1, j, k).x) + (G(i, j - 1, k).y - G(i, j + 1, k).y) +
                      (G(i, j, k - 1).z - G(i, j, k + 1).z)) / h2;
    const float4 om = w[cell];
    const float m = sqrtf(om.x * om.x + om.y * om.y + om.z * om.z);
    if (m > 0.f && mb != 0.f) {
        float4 wb = wbar[cell];
        wb.x += mb * om.x / m;
        wb.y += mb * om.y / m;
        wb.z += mb * om.z / m;
        wbar[cell] = wb;
    }
}

// pass 3 (per cell): adjoint of the cell-centred velocities of the curl
// (vorticity of solid cells is the constant 0, so their wbar is dropped):
//   ccbar_x = (wy(x-ez) - wy(x+ez) - wz(x-ey) + wz(x+ey)) / 2h
//   ccbar_y = (-wx(x-ez) + wx(x+ez) + wz(x-ex) - wz(x+ex)) / 2h
//   ccbar_z = (wx(x-ey) - wx(x+ey) - wy(x-ex) + wy(x+ex)) / 2h
// 2D (w in .x): ccbar_x = (w(x+ey) - w(x-ey)) / 2h, ccbar_y = (w(x-ex) -
// w(x+ex)) / 2h.
template <int D>
__global__ void k_a_confine3(XG X_, SimParams P, const uint8_t *f0, const float4 *wbar, float4 *ccbar) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    auto W = [&](int ii, int jj, int kk) -> float4 {
        const bool ok = inside(ii, jj, kk, P.n[0], P.n[1], P.n[2]);
        const uint32_t c = ok ? lin(ii, jj, kk, P.n[0], P.n[1]) : 0u;
        return (ok && !f0[c]) ? wbar[c] : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    const float h2 = 2.0f * P.h;
    float4 o;
    if (D == 2) {
        o.x = (-W(i, j - 1, k).x + W(i, j + 1, k).x) / h2;
        o.y = (W(i - 1, j, k).x - W(i + 1, j, k).x) / h2;
        o.z = 0.f;
        o.w = 0.f;
        ccbar[lin(i, j, k, P.n[0], P.n[1])] = o;
        return;
    }
    o.x = (W(i, j, k - 1).y - W(i, j, k + 1).y - W(i, j - 1, k).z + W(i, j + 1, k).z) / h2;
    o.y = (-W(i, j, k - 1).x + W(i, j, k + 1).x + W(i - 1, j, k).z - W(i + 1, j, k).z) / h2;
    o.z = (W(i, j - 1, k).x - W(i, j + 1, k).x - W(i - 1, j, k).y + W(i + 1, j, k).y) / h2;
    o.w = 0.f;
    ccbar[lin(i, j, k, P.n[0], P.n[1])] = o;
}

// pass 4 (per face): vbar_c(face) = vbar3_c(face) + 0.5 (ccbar_c(face - e_c)
// + ccbar_c(face)), cells inside the domain only.
template <int D>
__global__ void k_a_confine4(XG X_, SimParams P, VelPtrs vin, VelPtrs vout, const float4 *ccbar) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        if (!inside(i, j, k, sx, sy, sz)) continue;
        const uint32_t f = lin(i, j, k, sx, sy);
        const float4 a = w_at(P, ccbar, i - (c == 0), j - (c == 1), k - (c == 2));
        const float4 b = w_at(P, ccbar, i, j, k);
        const float ca = c == 0 ? a.x : (c == 1 ? a.y : a.z);
        const float cb = c == 0 ? b.x : (c == 1 ? b.y : b.z);
        static_cast<float *>(vout.c[c])[f] = static_cast<const float *>(vin.c[c])[f] + 0.5f * (ca + cb);
    }
}

// a face is a projection variable when both adjacent cells are inside the
// domain and fluid
template <int D>
__device__ __forceinline__ bool fluid_face(const SimParams &P, const uint8_t *f0, int c, int i, int j,
                                           int k) {
    const int e0 = c == 0, e1 = c == 1, e2 = c == 2;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2]) ||
        !inside(i - e0, j - e1, k - e2, P.n[0], P.n[1], P.n[2]))
        return false;
    return !f0[lin(i, j, k, P.n[0], P.n[1])] && !f0[lin(i - e0, j - e1, k - e2, P.n[0], P.n[1])];
}

// Projection adjoint, right-hand side: pbar_i = kc * sum_c (vbar_c(i) -
// vbar_c(i + e_c)) over fluid faces, kc = -dt/(rho h), written colour-split
// for the refined solver (which solves A x = rhs).
template <int D>
__global__ void k_a_project_rhs(XG X_, SimParams P, VelPtrs vbar, const uint8_t *f0, Level L0,
                                float *rhs) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    double acc = 0.0;
    if (!f0[lin(i, j, k, P.n[0], P.n[1])]) {
        for (int c = 0; c < D; ++c) {
            int sx, sy, sz;
            vsize<D>(P, c, sx, sy, sz);
            const float *vb = static_cast<const float *>(vbar.c[c]);
            const int e0 = c == 0, e1 = c == 1, e2 = c == 2;
            if (fluid_face<D>(P, f0, c, i, j, k)) acc += (double)vb[lin(i, j, k, sx, sy)];
            if (fluid_face<D>(P, f0, c, i + e0, j + e1, k + e2))
                acc -= (double)vb[lin(i + e0, j + e1, k + e2, sx, sy)];
        }
    }
    const double kc = -(double)P.dt / ((double)P.rho * (double)P.h);
    rhs[sidx(L0, i, j, k)] = (float)(kc * acc);
}

// Projection adjoint, face update: vbar_out = vbar_in + beta (x_minus -
// x_plus) on fluid faces, 0 elsewhere (solid faces are constants);
// beta = rho/(dt h), x = A^-1 rhs (FP64, colour-split).
template <int D>
__global__ void k_a_project_apply(XG X_, SimParams P, VelPtrs vin, VelPtrs vout, const uint8_t *f0,
                                  Level L0, const double *x) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    const double beta = (double)P.rho / ((double)P.dt * (double)P.h);
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        if (!inside(i, j, k, sx, sy, sz)) continue;
        const uint32_t f = lin(i, j, k, sx, sy);
        float o = 0.0f;
        if (fluid_face<D>(P, f0, c, i, j, k)) {
            const double xm = x[sidx(L0, i - (c == 0), j - (c == 1), k - (c == 2))];
            const double xp = x[sidx(L0, i, j, k)];
            o = (float)((double)static_cast<const float *>(vin.c[c])[f] + beta * (xm - xp));
        }
        static_cast<float *>(vout.c[c])[f] = o;
    }
}

// source adjoint: sbar += dt * rbar on the source box
__global__ void k_a_source(const float *rbar, float *sbar, int nx, int ny, Box3 box, float dt) {
    const int sx = box.hi[0] - box.lo[0], sy = box.hi[1] - box.lo[1], sz = box.hi[2] - box.lo[2];
    const size_t n = (size_t)sx * sy * sz;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % sx), j = (int)((c / sx) % sy), k = (int)(c / ((size_t)sx * sy));
        sbar[c] = sbar[c] + dt * rbar[lin(box.lo[0] + i, box.lo[1] + j, box.lo[2] + k, nx, ny)];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned flat_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 65535 * 4 ? 65535 * 4 : (b == 0 ? 1 : b));
}

void launch_d_source(hipStream_t s, const SimParams &P, int fmt, void *rho, const float *src, Box3 box) {
    size_t n = (size_t)(box.hi[0] - box.lo[0]) * (box.hi[1] - box.lo[1]) * (box.hi[2] - box.lo[2]);
    DISPATCH_VF(fmt, k_d_source<VF_><<<flat_blocks(n), 256, 0, s>>>(rho, src, P.n[0], P.n[1], box, P.dt));
}
void launch_d_buoyancy(hipStream_t s, const SimParams &P, int fmt, const void *vy_in, void *vy_out,
                       const uint8_t *f0, const void *rho) {
    Shape sh = shape_for(P, P.n[0], P.n[1] + 1, P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(fmt, k_d_buoyancy<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, vy_in, vy_out, f0, rho)));
}
void launch_d_vorticity(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0, float4 *w) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(fmt, k_d_vorticity<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, v, f0, w)));
}
void launch_d_confine(hipStream_t s, const SimParams &P, int fmt, VelPtrs in, VelPtrs out, const float4 *w) {
    Shape sh = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    DISPATCH_D(P, DISPATCH_VF(fmt, k_d_confine<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, w)));
}
void launch_d_advect_scalar(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0,
                            CoordTab T, const void *rin, void *rout) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(fmt, k_d_advect_scalar<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, v, f0, T, rin, rout)));
}
template <int F>
__global__ void k_state_from_f32(const float *src, void *dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        static_cast<typename Store<F>::T *>(dst)[i] = Store<F>::st(src[i]);
}
template <int F>
__global__ void k_state_to_f32(const void *src, float *dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = Store<F>::ld(static_cast<const typename Store<F>::T *>(src)[i]);
}
void launch_state_from_f32(hipStream_t s, int fmt, const float *src, void *dst, size_t n) {
    const unsigned b = flat_blocks(n);
    DISPATCH_VF(fmt, k_state_from_f32<VF_><<<b > 2048 ? 2048 : (b ? b : 1), 256, 0, s>>>(src, dst, n));
}
void launch_state_to_f32(hipStream_t s, int fmt, const void *src, float *dst, size_t n) {
    const unsigned b = flat_blocks(n);
    DISPATCH_VF(fmt, k_state_to_f32<VF_><<<b > 2048 ? 2048 : (b ? b : 1), 256, 0, s>>>(src, dst, n));
}

void launch_maxabs(hipStream_t s, const float *x, size_t n, unsigned int *out) {
    const unsigned b = flat_blocks(n);
    k_maxabs<<<b > 512 ? 512 : b, 256, 0, s>>>(x, n, out);
}
void launch_fit_loss(hipStream_t s, int fmt, const void *rho, const float *tgt, size_t n, float *rb,
                     double *part, double *L, double *losses, int cap, int *cnt) {
    DISPATCH_VF(fmt, k_fit_diff<VF_><<<kFitBlocks, 256, 0, s>>>(rho, tgt, n, rb, part));
    k_fit_record<<<1, 64, 0, s>>>(part, L, losses, cap, cnt);
}
void launch_fit_update(hipStream_t s, float *src, const float *g, size_t n, const double *L, double *part) {
    k_fit_sumsq<<<kFitBlocks, 256, 0, s>>>(g, n, part);
    const unsigned b = flat_blocks(n);
    k_fit_update<<<b > 64 ? 64 : (b ? b : 1), 256, 0, s>>>(src, g, n, L, part);
}
void launch_track_residual(hipStream_t s, const double *norms, double *out) {
    k_track_residual<<<1, 64, 0, s>>>(norms, out);
}
void launch_fx_scale(hipStream_t s, const unsigned int *m_lam, const unsigned int *m_field, float c0,
                     float c1, double *scale) {
    k_fx_scale<<<1, 64, 0, s>>>(m_lam, m_field, c0, c1, scale);
}
void launch_fx_to_f32(hipStream_t s, const unsigned long long *acc, const double *scale,
                      const float *base, float *dst, size_t n) {
    k_fx_to_f32<<<flat_blocks(n), 256, 0, s>>>(acc, scale, base, dst, n);
}

static UBar ubar_of(float *const ubar[4][3]) {
    UBar U;
    for (int c = 0; c < 4; ++c)
        for (int cc = 0; cc < 3; ++cc) U.u[c][cc] = ubar[c][cc];
    return U;
}

// Launch shape of the binned scatters: kBinTX x kBinTY blocks, z in runs of kBinKZ.
static Shape bin_shape(int sx, int sy, int sz) {
    Shape sh;
    sh.b = dim3(kBinTX, kBinTY, 1);
    sh.x.gx = (sx + kBinTX - 1) / kBinTX;
    sh.x.gy = (sy + kBinTY - 1) / kBinTY;
    sh.x.gz = (sz + kBinKZ - 1) / kBinKZ;
    const long total = (long)sh.x.gx * sh.x.gy * sh.x.gz;
    sh.x.per = (int)((total + kXcds - 1) / kXcds);
    sh.g = dim3((unsigned)(sh.x.per * kXcds), 1, 1);
    return sh;
}

void launch_a_advect_scalar(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0,
                            CoordTab T, const void *rin, const float *rbar_out,
                            unsigned long long *acc_r, const double *s_r, float *const ubar[4][3]) {
    if (P.dim == 3 && !P.gsamp) {  // path 1: the direct scatter below
        Shape sb = bin_shape(P.n[0], P.n[1], P.n[2]);
        DISPATCH_VF(fmt, k_a_advect_scalar_bin<VF_><<<sb.g, sb.b, 0, s>>>(sb.x, P, v, f0, T, rin, rbar_out,
                                                                         acc_r, s_r, ubar_of(ubar)));
        return;
    }
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(fmt, k_a_advect_scalar<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, v, f0, T, rin,
                                                                                  rbar_out, acc_r, s_r,
                                                                                  ubar_of(ubar))));
}
void launch_a_advect_velocity(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0,
                              CoordTab T, VelPtrs vbar_out, unsigned long long *const acc[3],
                              const double *s_acc, float *const ubar[4][3]) {
    if (P.dim == 3 && !P.gsamp) {  // path 1: the direct scatter below
        const UBar U = ubar_of(ubar);
        Shape s0 = bin_shape(P.n[0] + 1, P.n[1], P.n[2]);
        Shape s1 = bin_shape(P.n[0], P.n[1] + 1, P.n[2]);
        Shape s2 = bin_shape(P.n[0], P.n[1], P.n[2] + 1);
        DISPATCH_VF(fmt, {
            k_a_advect_velocity_bin<0, VF_><<<s0.g, s0.b, 0, s>>>(s0.x, P, v, f0, T, vbar_out, acc[0], s_acc, U);
            k_a_advect_velocity_bin<1, VF_><<<s1.g, s1.b, 0, s>>>(s1.x, P, v, f0, T, vbar_out, acc[1], s_acc, U);
            k_a_advect_velocity_bin<2, VF_><<<s2.g, s2.b, 0, s>>>(s2.x, P, v, f0, T, vbar_out, acc[2], s_acc, U);
        });
        return;
    }
    Shape sh = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    Acc3 a{{acc[0], acc[1], acc[2]}};
    DISPATCH_D(P, DISPATCH_VF(fmt, k_a_advect_velocity<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, v, f0, T,
                                                                                    vbar_out, a, s_acc,
                                                                                    ubar_of(ubar))));
}
void launch_a_interp_gather(hipStream_t s, const SimParams &P, CoordTab T, float *const ubar[4][3],
                            int cells, VelPtrs base, VelPtrs out) {
    Shape sh = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    DISPATCH_D(P, k_a_interp_gather<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, T, ubar_of(ubar), cells, base,
                                                             out));
}
void launch_a_buoyancy(hipStream_t s, const SimParams &P, const uint8_t *f0, const float *vybar,
                       unsigned long long *acc_r, const double *s_r) {
    Shape sh = shape_for(P, P.n[0], P.n[1] + 1, P.n[2]);
    DISPATCH_D(P, k_a_buoyancy<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, f0, vybar, acc_r, s_r));
}
void launch_a_confine(hipStream_t s, const SimParams &P, const uint8_t *f0, const float4 *w,
                      VelPtrs vbar_in, VelPtrs vbar_out, float4 *wbar, float4 *gbar, float4 *ccbar) {
    Shape sc = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, k_a_confine1<D_><<<sc.g, sc.b, 0, s>>>(sc.x, P, w, vbar_in, wbar, gbar));
    DISPATCH_D(P, k_a_confine2<D_><<<sc.g, sc.b, 0, s>>>(sc.x, P, w, gbar, wbar));
    DISPATCH_D(P, k_a_confine3<D_><<<sc.g, sc.b, 0, s>>>(sc.x, P, f0, wbar, ccbar));
    Shape sf = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    DISPATCH_D(P, k_a_confine4<D_><<<sf.g, sf.b, 0, s>>>(sf.x, P, vbar_in, vbar_out, ccbar));
}
void launch_a_project_rhs(hipStream_t s, const SimParams &P, VelPtrs vbar, const uint8_t *f0,
                          const Level &L0, float *rhs) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, k_a_project_rhs<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, vbar, f0, L0, rhs));
}
void launch_a_project_apply(hipStream_t s, const SimParams &P, VelPtrs vin, VelPtrs vout,
                            const uint8_t *f0, const Level &L0, const double *x) {
    Shape sh = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    DISPATCH_D(P, k_a_project_apply<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, vin, vout, f0, L0, x));
}
void launch_a_source(hipStream_t s, const SimParams &P, const float *rbar, float *sbar, Box3 box) {
    size_t n = (size_t)(box.hi[0] - box.lo[0]) * (box.hi[1] - box.lo[1]) * (box.hi[2] - box.lo[2]);
    k_a_source<<<flat_blocks(n), 256, 0, s>>>(rbar, sbar, P.n[0], P.n[1], box, P.dt);
}

}  // namespace smk
