// smk_kernels.hip -- gfx950 kernels of the smoke simulation step.
//
// Every stage is a bandwidth-bound stencil or gather; no MFMA.  Kernels are
// templated on the dimension D (2 or 3) and on the storage formats of the
// velocity faces (VF) and of smoke (SF).  The arithmetic of each per-cell
// update follows the reference GLSL expression by expression (compiled with
// -ffp-contract=off) so that results reproduce the reference step:
//   shaders{2d,3d}/compute/*.comp and pressure_solve/*.comp.
// Layout in HBM: every field is a dense x-fastest array (GL texel order);
// staggered components have one extra face along their own axis.  Smoke,
// 3D vorticity and the solid pyramid are 4 x fp16 (8 B) per cell, pressure
// and residual fp32.
#include "smk_device.h"

SMK_NS_BEGIN

// ---------------------------------------------------------------------------
// emitters: update_emitters_status.comp 3D:49-60, handle_emitters.comp 3D:49-76
// ---------------------------------------------------------------------------
__device__ __forceinline__ void emitter_status_at(int *em, int e, const float *time) {
    int *s = em + e * 12 + 8;  // delay, last emit, flag, falloff
    float t = *time;
    if (s[2] == 1 && s[0] != 0) {
        s[2] = 0;
        s[1] = (int)t;
        return;
    }
    if (t - (float)s[1] >= (float)s[0]) s[2] = 1;
}
__global__ void k_emitter_status(int *em, int ne, const float *time) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    emitter_status_at(em, e, time);
}

template <int D>
__device__ __forceinline__ float emitter_dist(int i, int j, int k, const int *E) {
    float dx = (float)i - (float)E[0];
    float dy = (float)j - (float)E[1];
    float dd = dx * dx + dy * dy;
    if (D == 3) {
        float dz = (float)k - (float)E[2];
        dd = dd + dz * dz;
    }
    return sqrtf(dd);
}

// Emission in place over the emitters' bounding box.  The reference copies
// the whole smoke texture (smoke_write = smoke_read + emission) and swaps;
// texels outside every emitter are unchanged by that copy, so only the box
// is touched here.  The same thread also applies the inflow extension to
// the vy face with the cell's index.
template <int D, int VF, int SF>
__device__ __forceinline__ void emit_body(int bx_, int by_, int bz_, const SimParams &P, uint16_t *smoke, void *vyv, const int *em, const float *inflow, int ne, const Box3 &box) {
    int i = box.lo[0] + bx_ * blockDim.x + threadIdx.x;
    int j = box.lo[1] + by_ * blockDim.y + threadIdx.y;
    int k = box.lo[2] + bz_;
    if (i >= box.hi[0] || j >= box.hi[1] || k >= box.hi[2]) return;
    const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
    if (i < nx && j < ny && k < nz) {
        uint32_t cell = lin(i, j, k, nx, ny);
        float4 s = ld4<SF>(smoke, cell);
        for (int e = 0; e < ne; ++e) {
            const int *E = em + e * 12;
            if (!(E[10] > 0)) continue;
            float dist = emitter_dist<D>(i, j, k, E);
            if (dist < (float)E[3]) {
                float fall = 1.0f;
                if (E[11] > 0) {
                    float e0 = (float)E[3], e1 = 0.0f;
                    float t = (dist - e0) / (e1 - e0);
                    t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
                    fall = t * t * (3.0f - 2.0f * t);
                }
                s.x = s.x + (float)E[4] * fall * P.dt;
                s.y = s.y + (float)E[5] * fall * P.dt;
                s.z = s.z + (float)E[6] * fall * P.dt;
                s.w = s.w + (float)E[7] * fall * P.dt;
            }
        }
        st4<SF>(smoke, cell, s);
    }
    if (inflow && i < nx && j <= ny && k < nz) {
        using S = Store<VF>;
        typename S::T *vy = static_cast<typename S::T *>(vyv);
        for (int e = 0; e < ne; ++e) {
            const int *E = em + e * 12;
            if (!(E[10] > 0) || inflow[e] == 0.0f) continue;
            if (emitter_dist<D>(i, j, k, E) < (float)E[3]) vy[lin(i, j, k, nx, ny + 1)] = S::st(inflow[e]);
        }
    }
}
template <int D, int VF, int SF>
__global__ void k_emit(XG X_, SimParams P, uint16_t *smoke, void *vyv, const int *em, const float *inflow, int ne, Box3 box) {
    XG_BLOCK(X_)
    emit_body<D, VF, SF>(bx_, by_, bz_, P, smoke, vyv, em, inflow, ne, box);
}

// ---------------------------------------------------------------------------
// advection: advect_v{x,y,z}.comp 3D:54-78 / 2D:50-75 (+ apply_buoyancy)
// ---------------------------------------------------------------------------
// One thread per (i,j,k) of the (n+1)^D box handles the x, y and z faces
// with that index.  With fuse_buoy the vy result also receives the buoyancy
// update of apply_buoyancy.comp 3D:39-52: faces next to a solid cell (OOB =
// fluid there) keep the pre-advection value, which is what the reference's
// ping-pong write texture holds at those faces.
// The block's slice of the coordinate table (the x entries of its columns,
// the y entries of its rows, the z entries of its plane) is staged in LDS
// first: each face's interpolation then starts from LDS instead of a
// dependent global load.
struct TabSlice {
    TC x[4][256], y[4][256], z[4];
};
__device__ __forceinline__ void load_tab_slice(TabSlice &S, const SimParams &P, const CoordTab &T,
                                               int i0, int j0, int k, int D) {
    const int t = threadIdx.y * blockDim.x + threadIdx.x, nt = blockDim.x * blockDim.y;
    for (int e = t; e < 4 * (int)blockDim.x; e += nt)
        S.x[e / blockDim.x][e % blockDim.x] = ctab(T, 0, e / blockDim.x, min(i0 + (int)(e % blockDim.x), P.n[0]));
    for (int e = t; e < 4 * (int)blockDim.y; e += nt)
        S.y[e / blockDim.y][e % blockDim.y] = ctab(T, 1, e / blockDim.y, min(j0 + (int)(e % blockDim.y), P.n[1]));
    if (t < 4) S.z[t] = D == 3 ? ctab(T, 2, t, min(k, P.n[2])) : TC{0, 0.0f};
    __syncthreads();
}

// BUF: every LINEAR sample goes through the range-checked buffer samplers
// (same texels, same arithmetic, border texels read as 0 by the buffer unit).
template <int D, int VF, int SF, bool FUSE_BUOY, bool BUF>
__device__ __forceinline__ void advect_velocity_body(int bx_, int by_, int bz_, const SimParams &P, const VelPtrs &in, const VelPtrs &out, const uint8_t *f0, const CoordTab &T, const uint16_t *smoke) {
    using S = Store<VF>;
    __shared__ TabSlice tab;
    Rsrc R[3], RS;
    if constexpr (BUF) {
#pragma unroll
        for (int cc = 0; cc < D; ++cc) R[cc] = make_rsrc(in.c[cc], vbytes<D, VF>(P, cc));
        RS = make_rsrc(smoke, (uint32_t)P.n[0] * (uint32_t)P.n[1] * (uint32_t)(D == 3 ? P.n[2] : 1) * 8u);
    }
    const int i0 = bx_ * blockDim.x, j0 = by_ * blockDim.y;
    int i = i0 + threadIdx.x;
    int j = j0 + threadIdx.y;
    int k = bz_;
    load_tab_slice(tab, P, T, i0, j0, k, D);
    const int id[3] = {i, j, k};
    // is_solid(idx) with OOB = solid (advection) and OOB = fluid (buoyancy)
    const bool here_in = inside(i, j, k, P.n[0], P.n[1], P.n[2]);
    const bool here_s = here_in ? f0[lin(i, j, k, P.n[0], P.n[1])] != 0 : false;
#pragma unroll
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        if (!inside(i, j, k, sx, sy, sz)) continue;
        uint32_t f = lin(i, j, k, sx, sy);
        const int e0 = c == 0, e1 = c == 1, e2 = c == 2;
        float nv;
        if (id[c] == 0 && P.wind[c] != 0.0f) {
            nv = P.wind[c];
        } else if (!here_in || here_s || solid_flag(P, f0, i - e0, j - e1, k - e2, true)) {
            nv = 0.0f;
        } else {
            // get_interpolated_velocity at the face: the table coordinates
            // of position kind (face along c) and component cc
            float v[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int cc = 0; cc < D; ++cc) {
                const TC tx = tab.x[(0 == c ? 2 : 0) + (0 == cc ? 1 : 0)][threadIdx.x];
                const TC ty = tab.y[(1 == c ? 2 : 0) + (1 == cc ? 1 : 0)][threadIdx.y];
                const TC tz = tab.z[(2 == c ? 2 : 0) + (2 == cc ? 1 : 0)];
                if constexpr (BUF)
                    v[cc] = sample_vel_buf<D, VF>(P, R[cc], cc, tx, ty, tz);
                else
                    v[cc] = sample_vel_tc<D, VF>(P, in.c[cc], cc, tx, ty, tz);
            }
            TC t[3] = {TC{0, 0.0f}, TC{0, 0.0f}, TC{0, 0.0f}};
#pragma unroll
            for (int d = 0; d < D; ++d) {
                float world = (d == c ? (float)id[d] : (float)id[d] + 0.5f) * P.h;
                float from = world - v[d] * P.dt;
                float gs = div_rn(from, P.h, P.rh, P.fdiv);
                float uv = div_rn(gs + (d == c ? 0.5f : 0.0f), (float)P.n[d] + (d == c ? 1.0f : 0.0f),
                                  d == c ? P.rn1[d] : P.rn[d], P.fdiv);
                t[d] = texcoord_tc(uv, P.n[d] + (d == c ? 1 : 0));
            }
            if constexpr (BUF)
                nv = sample_vel_buf<D, VF>(P, R[c], c, t[0], t[1], t[2]);
            else
                nv = sample_vel_tc<D, VF>(P, in.c[c], c, t[0], t[1], t[2]);
        }
        if (FUSE_BUOY && c == 1) {
            if (here_s || solid_flag(P, f0, i, j - 1, k, false)) {
                nv = S::ld(static_cast<const typename S::T *>(in.c[1])[f]);
            } else {
                float uv[3] = {div_rn((float)i + 0.5f, (float)P.n[0], P.rn[0], P.fdiv),
                               div_rn((float)j + 0.0f, (float)P.n[1], P.rn[1], P.fdiv),
                               div_rn((float)k + 0.5f, (float)P.n[2], P.rn[2], P.fdiv)};
                float4 sm;
                if constexpr (BUF)
                    sm = sample_smoke_buf<D, SF>(P, RS, uv);
                else
                    sm = sample_smoke<D, SF>(P, smoke, uv);
                float total = sm.x + sm.y + sm.z;
                float b = -P.alpha * total + P.beta * (sm.w - P.tamb);
                // the reference stores the advected value, then reads it back
                nv = S::ld(S::st(nv)) + b * P.dt;
            }
        }
        static_cast<typename S::T *>(out.c[c])[f] = S::st(nv);
    }
}
template <int D, int VF, int SF, bool FUSE_BUOY, bool BUF>
__global__ void k_advect_velocity(XG X_, SimParams P, VelPtrs in, VelPtrs out, const uint8_t *f0, CoordTab T, const uint16_t *smoke) {
    XG_BLOCK(X_)
    advect_velocity_body<D, VF, SF, FUSE_BUOY, BUF>(bx_, by_, bz_, P, in, out, f0, T, smoke);
}

// apply_buoyancy.comp 3D:39-52 / 2D:37-50 alone (vy_in -> vy_out, solid faces
// untouched).
template <int D, int VF, int SF>
__global__ void k_buoyancy(XG X_, SimParams P, const void *vy_in, void *vy_out, const uint8_t *f0,
                           const uint16_t *smoke) {
    XG_BLOCK(X_)
    using S = Store<VF>;
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    int sx, sy, sz;
    vsize<D>(P, 1, sx, sy, sz);
    if (!inside(i, j, k, sx, sy, sz)) return;
    if (solid_flag(P, f0, i, j, k, false) || solid_flag(P, f0, i, j - 1, k, false)) return;
    float uv[3] = {div_rn((float)i + 0.5f, (float)P.n[0], P.rn[0], P.fdiv),
                   div_rn((float)j + 0.0f, (float)P.n[1], P.rn[1], P.fdiv),
                   div_rn((float)k + 0.5f, (float)P.n[2], P.rn[2], P.fdiv)};
    float4 sm = sample_smoke<D, SF>(P, smoke, uv);
    float total = sm.x + sm.y + sm.z;
    float b = -P.alpha * total + P.beta * (sm.w - P.tamb);
    uint32_t f = lin(i, j, k, sx, sy);
    float nv = S::ld(static_cast<const typename S::T *>(vy_in)[f]) + b * P.dt;
    static_cast<typename S::T *>(vy_out)[f] = S::st(nv);
}

// ---------------------------------------------------------------------------
// vorticity: calculate_vorticity.comp 3D:59-84 / 2D:51-69,
//            apply_vorticity_{x,y,z}.comp 3D:43-63 / 2D:41-60
// ---------------------------------------------------------------------------
template <int D, int VF>
__device__ __forceinline__ float ccv(const SimParams &P, const VelPtrs &V, int c, int i, int j,
                                     int k) {
    return (vget<D, VF>(P, V, c, i, j, k) +
            vget<D, VF>(P, V, c, i + (c == 0), j + (c == 1), k + (c == 2))) *
           0.5f;
}

template <int D, int VF>
__device__ __forceinline__ void vorticity_at(int ti_, int tj_, int tk_, const SimParams &P, const VelPtrs &V, const uint8_t *f0, void *vort) {
    int i = ti_;
    int j = tj_;
    int k = tk_;
    const int nx = P.n[0], ny = P.n[1];
    if (!inside(i, j, k, nx, ny, P.n[2])) return;
    uint32_t cell = lin(i, j, k, nx, ny);
    const float h2 = 2.0f * P.h;
    bool solid = f0[cell] != 0;
    if (D == 3) {
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!solid) {
            float x1 = ccv<D, VF>(P, V, 2, i, j + 1, k) - ccv<D, VF>(P, V, 2, i, j - 1, k);
            float x2 = ccv<D, VF>(P, V, 1, i, j, k + 1) - ccv<D, VF>(P, V, 1, i, j, k - 1);
            o.x = (x1 - x2) / h2;
            float y1 = ccv<D, VF>(P, V, 0, i, j, k + 1) - ccv<D, VF>(P, V, 0, i, j, k - 1);
            float y2 = ccv<D, VF>(P, V, 2, i + 1, j, k) - ccv<D, VF>(P, V, 2, i - 1, j, k);
            o.y = (y1 - y2) / h2;
            float z1 = ccv<D, VF>(P, V, 1, i + 1, j, k) - ccv<D, VF>(P, V, 1, i - 1, j, k);
            float z2 = ccv<D, VF>(P, V, 0, i, j + 1, k) - ccv<D, VF>(P, V, 0, i, j - 1, k);
            o.z = (z1 - z2) / h2;
        }
        st4<FMT_F16>(static_cast<uint16_t *>(vort), cell, o);
    } else {
        float o = 0.0f;
        if (!solid) {
            float z1 = ccv<D, VF>(P, V, 1, i + 1, j, k) - ccv<D, VF>(P, V, 1, i - 1, j, k);
            float z2 = ccv<D, VF>(P, V, 0, i, j + 1, k) - ccv<D, VF>(P, V, 0, i, j - 1, k);
            o = (z1 - z2) / h2;
        }
        static_cast<float *>(vort)[cell] = o;
    }
}

template <int D, int VF>
__device__ __forceinline__ void vorticity_body(int bx_, int by_, int bz_, const SimParams &P, const VelPtrs &V, const uint8_t *f0, void *vort) {
    vorticity_at<D, VF>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, V, f0, vort);
}
template <int D, int VF>
__global__ void k_vorticity(XG X_, SimParams P, VelPtrs V, const uint8_t *f0, void *vort) {
    XG_BLOCK(X_)
    vorticity_body<D, VF>(bx_, by_, bz_, P, V, f0, vort);
}

template <int D>
__device__ __forceinline__ float4 vort_at(const SimParams &P, const void *vort, int i, int j, int k) {
    const bool ok = inside(i, j, k, P.n[0], P.n[1], P.n[2]);
    const uint32_t cell = lin(i, j, k, P.n[0], P.n[1]);
    if (D == 3)
        return ld4_sel<FMT_F16>(static_cast<const uint16_t *>(vort), ok, cell, make_float4(0.f, 0.f, 0.f, 0.f));
    return make_float4(ld_sel(static_cast<const float *>(vort), ok, cell, 0.0f), 0.f, 0.f, 0.f);
}
template <int D>
__device__ __forceinline__ float vmag(const SimParams &P, const void *vort, int i, int j, int k) {
    float4 o = vort_at<D>(P, vort, i, j, k);
    if (D == 3) return sqrtf(o.x * o.x + o.y * o.y + o.z * o.z);
    return fabsf(o.x);
}

// One thread per (i,j,k): the confinement force at cell (i,j,k) drives the
// x, y and z faces with that index (apply_vorticity_x.comp:47-56 samples the
// cell idx for face idx).
template <int D, int VF>
__device__ __forceinline__ void apply_vorticity_at(int ti_, int tj_, int tk_, const SimParams &P, const VelPtrs &in, const VelPtrs &out, const void *vort) {
    using S = Store<VF>;
    int i = ti_;
    int j = tj_;
    int k = tk_;
    if (i > P.n[0] || j > P.n[1] || k > P.n[2] - (D == 2 ? 1 : 0)) return;
    // the faces' current values load with the vorticity (one memory round
    // trip instead of one more per component)
    bool valid[3] = {false, false, false};
    uint32_t fi[3] = {0u, 0u, 0u};
    float cur[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        valid[c] = inside(i, j, k, sx, sy, sz);
        fi[c] = lin(i, j, k, sx, sy);
        cur[c] = S::ld(ld_sel(static_cast<const typename S::T *>(in.c[c]), valid[c], fi[c], typename S::T(0)));
    }
    const float h2 = 2.0f * P.h;
    float dx = vmag<D>(P, vort, i + 1, j, k) - vmag<D>(P, vort, i - 1, j, k);
    float dy = vmag<D>(P, vort, i, j + 1, k) - vmag<D>(P, vort, i, j - 1, k);
    float4 om = vort_at<D>(P, vort, i, j, k);
    float f[3];
    if (D == 3) {
        float dz = vmag<D>(P, vort, i, j, k + 1) - vmag<D>(P, vort, i, j, k - 1);
        const float rh2 = 0.5f * P.rh;  // RN(1/(2h)) = RN(1/h)/2
        float gx = div_rn(dx, h2, rh2, P.fdiv), gy = div_rn(dy, h2, rh2, P.fdiv);
        float gz = div_rn(dz, h2, rh2, P.fdiv);
        float len = sqrtf(gx * gx + gy * gy + gz * gz) + 1e-20f;
        float nx = gx / len, ny = gy / len, nz = gz / len;
        f[0] = ny * om.z - nz * om.y;
        f[1] = nz * om.x - nx * om.z;
        f[2] = nx * om.y - ny * om.x;
    } else {
        const float rh2 = 0.5f * P.rh;
        float gx = div_rn(dx, h2, rh2, P.fdiv), gy = div_rn(dy, h2, rh2, P.fdiv);
        float len = sqrtf(gx * gx + gy * gy) + 1e-20f;
        float nx = gx / len, ny = gy / len;
        f[0] = -ny * om.x;  // 2D apply_vorticity_x.comp:54
        f[1] = -nx * om.x;  // 2D apply_vorticity_y.comp:54
        f[2] = 0.f;
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
        if (!valid[c]) continue;
        float nv = cur[c] + div_rn(f[c] * 0.5f * P.eps * P.h * P.dt, P.rho, P.rrho, P.fdiv);
        static_cast<typename S::T *>(out.c[c])[fi[c]] = S::st(nv);
    }
}

template <int D, int VF>
__device__ __forceinline__ void apply_vorticity_body(int bx_, int by_, int bz_, const SimParams &P, const VelPtrs &in, const VelPtrs &out, const void *vort) {
    apply_vorticity_at<D, VF>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, in, out, vort);
}
template <int D, int VF>
__global__ void k_apply_vorticity(XG X_, SimParams P, VelPtrs in, VelPtrs out, const void *vort) {
    XG_BLOCK(X_)
    apply_vorticity_body<D, VF>(bx_, by_, bz_, P, in, out, vort);
}

// ---------------------------------------------------------------------------
// projection: update_v{x,y,z}.comp 3D:45-63 / 2D:43-61
// ---------------------------------------------------------------------------
// P64: the pressure is the FP64 refined solution, the gradient is taken in
// double.
template <int D, int VF, bool P64>
__device__ __forceinline__ void update_velocity_at(int ti_, int tj_, int tk_, const SimParams &P, const VelPtrs &in, const VelPtrs &out, const uint16_t *s0, const uint8_t *f0, const Level &L0, const double *p64) {
    using S = Store<VF>;
    int i = ti_;
    int j = tj_;
    int k = tk_;
    const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
    const float kc = -P.dt / (P.rho * P.h);
    const double kd = -(double)P.dt / ((double)P.rho * (double)P.h);
    // every load of the three faces first (branch free, before the mask
    // test: one memory round trip), then the updates
    const bool here_in = inside(i, j, k, nx, ny, nz);
    const float p_here = P64 ? 0.0f : ld_sel(L0.p, here_in, sidx(L0, i, j, k), 0.0f);
    const double p_here64 = P64 ? ld_sel(p64, here_in, sidx(L0, i, j, k), 0.0) : 0.0;
    bool valid[3] = {false, false, false}, solid[3] = {false, false, false};
    uint32_t f[3] = {0u, 0u, 0u};
    float vold[3] = {0.f, 0.f, 0.f}, pl[3] = {0.f, 0.f, 0.f};
    double pl64[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < D; ++c) {
        int sx, sy, sz;
        vsize<D>(P, c, sx, sy, sz);
        const int e0 = c == 0, e1 = c == 1, e2 = c == 2;
        valid[c] = inside(i, j, k, sx, sy, sz);
        f[c] = lin(i, j, k, sx, sy);
        vold[c] = S::ld(ld_sel(static_cast<const typename S::T *>(in.c[c]), valid[c], f[c],
                               typename S::T(0)));
        if (P64)
            pl64[c] = ld_sel(p64, inside(i - e0, j - e1, k - e2, nx, ny, nz),
                             sidx(L0, i - e0, j - e1, k - e2), 0.0);
        else
            pl[c] = ld_sel(L0.p, inside(i - e0, j - e1, k - e2, nx, ny, nz),
                           sidx(L0, i - e0, j - e1, k - e2), 0.0f);
    }
    // solid tests from the level-0 neighbour mask of the cell (solid bit,
    // fluid bits of the -x/-y/-z neighbours); cells on the domain boundary
    // or next to a fractional solid read the flags.  The boundary is known
    // from the indices alone, so those cells load their flags with the
    // first batch; only cells next to a fractional solid wait for the mask
    // before loading theirs.
    const uint32_t msk = ld_sel(L0.mask, here_in, sidx(L0, i, j, k), (uint8_t)0);
    const bool bnd = here_in && (i == 0 || j == 0 || i == nx - 1 || j == ny - 1 ||
                                 (D == 3 && (k == 0 || k == nz - 1)));
    bool fl[4] = {false, false, false, false};
    if (bnd) {
        fl[3] = solid_flag(P, f0, i, j, k, false);
#pragma unroll
        for (int c = 0; c < D; ++c)
            fl[c] = solid_flag(P, f0, i - (c == 0), j - (c == 1), k - (c == 2), false);
    }
    const bool fast = here_in && !(msk & 0x80u);
    bool s_here;
    bool nsolid[3] = {false, false, false};
    if (fast) {
        s_here = (msk & 0x01u) != 0;
        nsolid[0] = !(msk & 0x10u);  // -x neighbour not fluid
        nsolid[1] = !(msk & 0x04u);  // -y
        nsolid[2] = !(msk & 0x40u);  // -z
    } else if (bnd) {
        s_here = fl[3];
#pragma unroll
        for (int c = 0; c < D; ++c) nsolid[c] = fl[c];
    } else {
        s_here = solid_flag(P, f0, i, j, k, false);
#pragma unroll
        for (int c = 0; c < D; ++c)
            nsolid[c] = solid_flag(P, f0, i - (c == 0), j - (c == 1), k - (c == 2), false);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) solid[c] = s_here || nsolid[c];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        if (!valid[c]) continue;
        float nv = P64 ? (float)((double)vold[c] + kd * (p_here64 - pl64[c]))
                       : vold[c] + kc * (p_here - pl[c]);
        if (solid[c]) {
            // imageLoad(solid_map, idx).{x,y,z}: zero outside the image
            nv = f16_to_f32(ld_sel(s0, here_in, lin(i, j, k, nx, ny) * 4 + c, (uint16_t)0));
        }
        static_cast<typename S::T *>(out.c[c])[f[c]] = S::st(nv);
    }
}

template <int D, int VF, bool P64>
__device__ __forceinline__ void update_velocity_body(int bx_, int by_, int bz_, const SimParams &P, const VelPtrs &in, const VelPtrs &out, const uint16_t *s0, const uint8_t *f0, const Level &L0, const double *p64) {
    update_velocity_at<D, VF, P64>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, in, out, s0, f0, L0, p64);
}
template <int D, int VF, bool P64>
__global__ void k_update_velocity(XG X_, SimParams P, VelPtrs in, VelPtrs out, const uint16_t *s0, const uint8_t *f0, Level L0, const double *p64) {
    XG_BLOCK(X_)
    update_velocity_body<D, VF, P64>(bx_, by_, bz_, P, in, out, s0, f0, L0, p64);
}

// ---------------------------------------------------------------------------
// advect_smoke.comp 3D:55-72 / 2D:51-68
// ---------------------------------------------------------------------------
template <int D, int VF, int SF, bool BUF>
__device__ __forceinline__ void advect_smoke_body(int bx_, int by_, int bz_, const SimParams &P, const VelPtrs &V, const uint8_t *f0, const CoordTab &T, const uint16_t *sin, uint16_t *sout) {
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    const int nx = P.n[0], ny = P.n[1];
    if (!inside(i, j, k, nx, ny, P.n[2])) return;
    uint32_t cell = lin(i, j, k, nx, ny);
    float4 o;
    if (f0[cell]) {
        o = make_float4(0.f, 0.f, 0.f, P.tamb);
    } else {
        const int id[3] = {i, j, k};
        float v[3] = {0.f, 0.f, 0.f}, uv[3] = {0.f, 0.f, 0.f};
        Rsrc R[3];
        if constexpr (BUF) {
#pragma unroll
            for (int cc = 0; cc < D; ++cc) R[cc] = make_rsrc(V.c[cc], vbytes<D, VF>(P, cc));
            interp_velocity_buf<D, VF>(P, R, T, id, -1, v);
        } else {
            interp_velocity_tab<D, VF>(P, V, T, id, -1, v);
        }
        for (int d = 0; d < D; ++d) {
            float world = ((float)id[d] + 0.5f) * P.h;
            float from = world - v[d] * P.dt;
            float gs = div_rn(from, P.h, P.rh, P.fdiv);
            uv[d] = div_rn(gs, (float)P.n[d], P.rn[d], P.fdiv);
        }
        if constexpr (BUF)
            o = sample_smoke_buf<D, SF>(
                P, make_rsrc(sin, (uint32_t)nx * (uint32_t)ny * (uint32_t)(D == 3 ? P.n[2] : 1) * 8u), uv);
        else
            o = sample_smoke<D, SF>(P, sin, uv);
    }
    st4<SF>(sout, cell, o);
}
template <int D, int VF, int SF, bool BUF>
__global__ void k_advect_smoke(XG X_, SimParams P, VelPtrs V, const uint8_t *f0, CoordTab T, const uint16_t *sin, uint16_t *sout) {
    XG_BLOCK(X_)
    advect_smoke_body<D, VF, SF, BUF>(bx_, by_, bz_, P, V, f0, T, sin, sout);
}

// Coordinate table of the face-centred interpolation (see CoordTab).
template <int D>
__global__ void k_coord_table(SimParams P, TC *e, int stride) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int dq = blockIdx.y;  // d * 4 + q
    int d = dq >> 2, q = dq & 3;
    if (d >= D || idx > P.n[d]) return;
    const int s = q >> 1, t = q & 1;
    float world = (s ? (float)idx : (float)idx + 0.5f) * P.h;
    float gs = world / P.h;
    float uv = (gs + (t ? 0.5f : 0.0f)) / ((float)P.n[d] + (t ? 1.0f : 0.0f));
    e[(size_t)dq * stride + idx] = texcoord_tc(uv, P.n[d] + t);
}

// Level-0 solid flags (a >= 1).
__global__ void k_solid_flag(const uint16_t *s0, uint8_t *f0, size_t n) {
    for (uint32_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x)
        f0[c] = solid_a(s0, c) >= 1.0f ? 1 : 0;
}

// ---------------------------------------------------------------------------
// obstacles / masks: obstacle_update.comp 3D:51-134 / 2D:49-130,
// update_solid_map.comp 3D:46-82 / 2D:44-80, init_solid_map.comp 3D:34-47
// ---------------------------------------------------------------------------
// One obstacle per thread; all obstacles of the input buffer are visible to
// every thread (the reference tiles them through shared memory).
template <int D>
__device__ __forceinline__ void obstacle_update_at(const SimParams &P, const float *in, float *out, int no, int o) {
    const float *cur = in + o * 8;
    float pos[3] = {cur[0], cur[1], D == 3 ? cur[2] : 0.f};
    float vel[3] = {cur[4], cur[5], D == 3 ? cur[6] : 0.f};
    float r = cur[3], anch = cur[7], mov = 1.0f - anch;
    if (!(anch == 1.0f)) {
        const float grav[3] = {0.f, -9.81f, 0.f};
        for (int d = 0; d < D; ++d) vel[d] = vel[d] + grav[d] * P.dt * mov;
        for (int d = 0; d < D; ++d) pos[d] = pos[d] + vel[d] * P.dt * mov;
        const float bounce = 0.999f;
        for (int d = 0; d < D; ++d) {
            float mx = (float)P.n[d];
            if (pos[d] - r < 0.0f) {
                pos[d] = 0.0f + r;
                vel[d] = vel[d] * -bounce;
            }
            if (pos[d] + r > mx) {
                pos[d] = mx - r;
                vel[d] = vel[d] * -bounce;
            }
        }
        for (int q = 0; q < no; ++q) {
            if (q == o) continue;
            const float *oth = in + q * 8;
            float dd[3] = {0.f, 0.f, 0.f};
            for (int d = 0; d < D; ++d) dd[d] = pos[d] - oth[d];
            float d2 = dd[0] * dd[0] + dd[1] * dd[1];
            if (D == 3) d2 = d2 + dd[2] * dd[2];
            float dist = sqrtf(d2);
            float mind = r + oth[3];
            if (dist < mind && dist > 0.0001f) {
                float nn[3] = {0.f, 0.f, 0.f};
                for (int d = 0; d < D; ++d) nn[d] = dd[d] / dist;
                for (int d = 0; d < D; ++d) pos[d] = pos[d] + nn[d] * (mind - dist) * 0.5f * mov;
                float rel[3] = {0.f, 0.f, 0.f};
                for (int d = 0; d < D; ++d) rel[d] = vel[d] - oth[4 + d];
                float vn = rel[0] * nn[0] + rel[1] * nn[1];
                if (D == 3) vn = vn + rel[2] * nn[2];
                if (vn < 0.0f)
                    for (int d = 0; d < D; ++d) vel[d] = vel[d] - vn * nn[d] * mov;
            }
        }
    }
    float *w = out + o * 8;
    w[0] = pos[0];
    w[1] = pos[1];
    w[2] = D == 3 ? pos[2] : 0.f;
    w[3] = r;
    for (int d = 0; d < 3; ++d) w[4 + d] = d < D ? vel[d] * 0.999f : 0.f;
    w[7] = anch;
}
template <int D>
__global__ void k_obstacle_update(SimParams P, const float *in, float *out, int no) {
    for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < no; o += gridDim.x * blockDim.x)
        obstacle_update_at<D>(P, in, out, no, o);
}

template <int D>
__global__ void k_update_solid(XG X_, SimParams P, uint16_t *s0, const float *ob, int no, const int *box,
                               int nb) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
    bool valid = !(i <= 1 || j <= 1 || i >= nx - 1 || j >= ny - 1);
    if (D == 3) valid = valid && !(k <= 1 || k >= nz - 1);
    if (!valid || !inside(i, j, k, nx, ny, nz)) return;
    float4 nv = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int o = 0; o < no; ++o) {
        const float *c = ob + o * 8;
        float dx = (float)i - c[0], dy = (float)j - c[1];
        float d2 = dx * dx + dy * dy;
        if (D == 3) {
            float dz = (float)k - c[2];
            d2 = d2 + dz * dz;
        }
        if (sqrtf(d2) < c[3]) {
            float m = 1.0f - c[7];
            nv = make_float4(c[4] * m, c[5] * m, c[6] * m, 1.0f);
        }
    }
    for (int b = 0; b < nb; ++b) {
        const int *B = box + b * 6;
        bool in = i >= B[0] && i < B[3] && j >= B[1] && j < B[4];
        if (D == 3) in = in && k >= B[2] && k < B[5];
        if (in) nv = make_float4(0.f, 0.f, 0.f, 1.0f);
    }
    st4<FMT_F16>(s0, lin(i, j, k, nx, ny), nv);
}

template <int D>
__global__ void k_init_solid(XG X_, SimParams P, uint16_t *s0) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
    if (!inside(i, j, k, nx, ny, nz)) return;
    bool border = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
    if (D == 3) border = border || k == 0 || k == nz - 1;
    if (border) st4<FMT_F16>(s0, lin(i, j, k, nx, ny), make_float4(0.f, 0.f, 0.f, 1.0f));
}

// ---------------------------------------------------------------------------
// utilities
// ---------------------------------------------------------------------------
__global__ void k_fill_f32(float *p, size_t n, float v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u16(uint16_t *p, size_t n, uint16_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_set_f32(float *p, float v) { *p = v; }

// Plain streaming copy (16-byte loads and stores), the calibration kernel for
// HBM counters and the practical bandwidth ceiling.
__global__ void k_copy16(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_to_f64(const void *src, int fmt, double *dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        float v;
        if (fmt == FMT_F32) v = static_cast<const float *>(src)[i];
        else if (fmt == FMT_BF16) v = bf16_to_f32(static_cast<const uint16_t *>(src)[i]);
        else v = f16_to_f32(static_cast<const uint16_t *>(src)[i]);
        dst[i] = (double)v;
    }
}
__global__ void k_from_f64(const double *src, void *dst, int fmt, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        float v = (float)src[i];
        if (fmt == FMT_F32) static_cast<float *>(dst)[i] = v;
        else if (fmt == FMT_BF16) static_cast<uint16_t *>(dst)[i] = f32_to_bf16(v);
        else static_cast<uint16_t *>(dst)[i] = f32_to_f16(v);
    }
}

template <int D, int VF>
__global__ void k_divergence_field(XG X_, SimParams P, VelPtrs V, const uint16_t *s0, double *out) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    uint32_t cell = lin(i, j, k, P.n[0], P.n[1]);
    out[cell] = solid_a(s0, cell) >= 1.0f
                    ? 0.0
                    : (double)(level0_dv<D, VF>(P, V, s0, i, j, k, 0.0f) / P.h);
}

// ---------------------------------------------------------------------------
// multi-scene kernels (fused batches): the single-scene bodies over every
// scene of a launch, each block on its scene's record
// ---------------------------------------------------------------------------
__global__ void k_emit_status_multi(const int *start, int n, const MEmitStatus *E) {
    int s_, l_;
    if (!multi_block(start, n, s_, l_)) return;
    const MEmitStatus &a = E[s_];
    const int e = l_ * blockDim.x + threadIdx.x;
    if (e < a.ne) emitter_status_at(a.em, e, a.time);
}
template <int D, int VF, int SF>
__global__ void k_emit_multi(const int *start, int n, const MEmit *E) {
    MULTI_BLOCK(start, n, E)
    emit_body<D, VF, SF>(bx_, by_, bz_, a_.P, a_.smoke, a_.vy, a_.em, a_.inflow, a_.ne, a_.box);
}
template <int D, int VF, int SF, bool FUSE_BUOY, bool BUF>
__global__ void k_advect_velocity_multi(const int *start, int n, const MAdvectVel *E) {
    MULTI_BLOCK(start, n, E)
    advect_velocity_body<D, VF, SF, FUSE_BUOY, BUF>(bx_, by_, bz_, a_.P, a_.in, a_.out, a_.f0, a_.T, a_.smoke);
}
template <int D, int VF>
__global__ void k_vorticity_multi(const int *start, int n, const MVort *E) {
    MULTI_FLAT_BLOCK(start, n, E)
    vorticity_at<D, VF>(x_, y_, z_, a_.P, a_.V, a_.f0, a_.vort);
}
template <int D, int VF>
__global__ void k_apply_vorticity_multi(const int *start, int n, const MApplyVort *E) {
    MULTI_FLAT_BLOCK(start, n, E)
    apply_vorticity_at<D, VF>(x_, y_, z_, a_.P, a_.in, a_.out, a_.vort);
}
template <int D, int VF>
__global__ void k_update_velocity_multi(const int *start, int n, const MUpdateVel *E) {
    MULTI_FLAT_BLOCK(start, n, E)
    update_velocity_at<D, VF, false>(x_, y_, z_, a_.P, a_.in, a_.out, a_.s0, a_.f0, a_.L0, nullptr);
}
template <int D, int VF, int SF, bool BUF>
__global__ void k_advect_smoke_multi(const int *start, int n, const MAdvectSmoke *E) {
    MULTI_BLOCK(start, n, E)
    advect_smoke_body<D, VF, SF, BUF>(bx_, by_, bz_, a_.P, a_.V, a_.f0, a_.T, a_.sin, a_.sout);
}
template <int D>
__global__ void k_obstacle_update_multi(const int *start, int n, const MObstacles *E) {
    int s_, l_;
    if (!multi_block(start, n, s_, l_)) return;
    const MObstacles &a = E[s_];
    const int o = l_ * blockDim.x + threadIdx.x;
    if (o < a.no) obstacle_update_at<D>(a.P, a.in, a.out, a.no, o);
}
__global__ void k_set_times(float *const *times, int n, float t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) *times[i] = t;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
void launch_emitter_status(hipStream_t s, int *em, int ne, const float *time) {
    if (ne <= 0) return;
    k_emitter_status<<<(ne + 63) / 64, 64, 0, s>>>(em, ne, time);
}

void launch_emit(hipStream_t s, const SimParams &P, int vfmt, int sfmt, void *smoke, void *vy,
                 const int *em, const float *inflow, int ne, Box3 box) {
    if (ne <= 0) return;
    int ex = box.hi[0] - box.lo[0], ey = box.hi[1] - box.lo[1], ez = box.hi[2] - box.lo[2];
    if (ex <= 0 || ey <= 0 || ez <= 0) return;
    Shape sh = shape_for(P, ex, ey, ez);
    DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, k_emit<D_, VF_, SF_><<<sh.g, sh.b, 0, s>>>(sh.x, 
                                                         P, (uint16_t *)smoke, vy, em, inflow, ne, box))));
}

void launch_advect_velocity(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs in,
                            VelPtrs out, const uint8_t *flag0, CoordTab T, const void *smoke,
                            int fuse_buoy) {
    Shape sh = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    const bool buf = tex_buf_ok(P);
    DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, {
        if (fuse_buoy && buf)
            k_advect_velocity<D_, VF_, SF_, true, true>
                <<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, flag0, T, (const uint16_t *)smoke);
        else if (fuse_buoy)
            k_advect_velocity<D_, VF_, SF_, true, false>
                <<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, flag0, T, (const uint16_t *)smoke);
        else if (buf)
            k_advect_velocity<D_, VF_, SF_, false, true>
                <<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, flag0, T, (const uint16_t *)smoke);
        else
            k_advect_velocity<D_, VF_, SF_, false, false>
                <<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, flag0, T, (const uint16_t *)smoke);
    })));
}

void launch_coord_table(hipStream_t s, const SimParams &P, TC *e, int stride) {
    dim3 g((unsigned)((stride + 63) / 64), 12, 1);
    DISPATCH_D(P, k_coord_table<D_><<<g, 64, 0, s>>>(P, e, stride));
}

void launch_solid_flag(hipStream_t s, const uint16_t *solid0, uint8_t *flag0, size_t n) {
    size_t b = (n + 255) / 256;
    unsigned blocks = (unsigned)(b > 65535 * 4 ? 65535 * 4 : (b == 0 ? 1 : b));
    k_solid_flag<<<blocks, 256, 0, s>>>(solid0, flag0, n);
}

void launch_buoyancy(hipStream_t s, const SimParams &P, int vfmt, int sfmt, const void *vy_in,
                     void *vy_out, const uint8_t *flag0, const void *smoke) {
    Shape sh = shape_for(P, P.n[0], P.n[1] + 1, P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, k_buoyancy<D_, VF_, SF_><<<sh.g, sh.b, 0, s>>>(sh.x, 
                                                         P, vy_in, vy_out, flag0,
                                                         (const uint16_t *)smoke))));
}

void launch_vorticity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in,
                      const uint8_t *flag0, void *vort) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_vorticity<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, in, flag0, vort)));
}

void launch_apply_vorticity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in, VelPtrs out,
                            const void *vort) {
    Shape sh = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_apply_vorticity<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, vort)));
}

void launch_update_velocity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in, VelPtrs out,
                            const uint16_t *solid0, const uint8_t *flag0, const Level &L0,
                            const double *p64) {
    Shape sh = shape_for(P, P.n[0] + 1, P.n[1] + 1, P.n[2] + (P.dim == 3 ? 1 : 0));
    if (p64)
        DISPATCH_D(P, DISPATCH_VF(vfmt, k_update_velocity<D_, VF_, true><<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, solid0, flag0, L0, p64)));
    else
        DISPATCH_D(P, DISPATCH_VF(vfmt, k_update_velocity<D_, VF_, false><<<sh.g, sh.b, 0, s>>>(sh.x, P, in, out, solid0, flag0, L0, p64)));
}

void launch_advect_smoke(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs vel,
                         const uint8_t *flag0, CoordTab T, const void *smoke_in, void *smoke_out) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    if (tex_buf_ok(P))
        DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, k_advect_smoke<D_, VF_, SF_, true><<<sh.g, sh.b, 0, s>>>(
                                                             sh.x, P, vel, flag0, T, (const uint16_t *)smoke_in,
                                                             (uint16_t *)smoke_out))));
    else
        DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, k_advect_smoke<D_, VF_, SF_, false><<<sh.g, sh.b, 0, s>>>(
                                                             sh.x, P, vel, flag0, T, (const uint16_t *)smoke_in,
                                                             (uint16_t *)smoke_out))));
}

void launch_obstacle_update(hipStream_t s, const SimParams &P, const float *ob_in, float *ob_out,
                            int no) {
    if (no <= 0) return;
    int blocks = (no + 255) / 256;
    DISPATCH_D(P, k_obstacle_update<D_><<<blocks, 256, 0, s>>>(P, ob_in, ob_out, no));
}

void launch_update_solid(hipStream_t s, const SimParams &P, uint16_t *solid0, const float *ob,
                         int no, const int *boxes, int nb) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, k_update_solid<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, solid0, ob, no, boxes, nb));
}

void launch_init_solid(hipStream_t s, const SimParams &P, uint16_t *solid0) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, k_init_solid<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, solid0));
}

static unsigned flat_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 65535 * 4 ? 65535 * 4 : (b == 0 ? 1 : b));
}

void launch_fill_f32(hipStream_t s, float *p, size_t n, float v) {
    k_fill_f32<<<flat_blocks(n), 256, 0, s>>>(p, n, v);
}
void launch_fill_u16(hipStream_t s, uint16_t *p, size_t n, uint16_t v) {
    k_fill_u16<<<flat_blocks(n), 256, 0, s>>>(p, n, v);
}
void launch_set_f32(hipStream_t s, float *p, float v) { k_set_f32<<<1, 1, 0, s>>>(p, v); }
void launch_copy16(hipStream_t s, const void *src, void *dst, size_t bytes) {
    size_t n = bytes / 16;
    k_copy16<<<2048, 256, 0, s>>>(static_cast<const float4 *>(src), static_cast<float4 *>(dst), n);
}
void launch_to_f64(hipStream_t s, const void *src, int fmt, double *dst, size_t n) {
    k_to_f64<<<flat_blocks(n), 256, 0, s>>>(src, fmt, dst, n);
}
void launch_from_f64(hipStream_t s, const double *src, void *dst, int fmt, size_t n) {
    k_from_f64<<<flat_blocks(n), 256, 0, s>>>(src, dst, fmt, n);
}
void launch_divergence_field(hipStream_t s, const SimParams &P, int vfmt, VelPtrs vel,
                             const uint16_t *solid0, double *out) {
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_divergence_field<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, vel, solid0, out)));
}

// multi-scene launchers
void launch_emit_status_multi(hipStream_t s, const MultiLaunch &m, const MEmitStatus *E) {
    if (m.n <= 0) return;
    k_emit_status_multi<<<multi_grid(m), m.bx, 0, s>>>(m.start, m.n, E);
}
void launch_emit_multi(hipStream_t s, int dim, int vfmt, int sfmt, const MultiLaunch &m, const MEmit *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, k_emit_multi<D_, VF_, SF_>
                                                          <<<multi_grid(m), multi_block_dim(m), 0, s>>>(m.start, m.n, E))));
}
void launch_advect_velocity_multi(hipStream_t s, int dim, int vfmt, int sfmt, bool fuse_buoy, bool buf,
                                  const MultiLaunch &m, const MAdvectVel *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    const dim3 g = multi_grid(m), b = multi_block_dim(m);
    DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, {
        if (fuse_buoy && buf) k_advect_velocity_multi<D_, VF_, SF_, true, true><<<g, b, 0, s>>>(m.start, m.n, E);
        else if (fuse_buoy) k_advect_velocity_multi<D_, VF_, SF_, true, false><<<g, b, 0, s>>>(m.start, m.n, E);
        else if (buf) k_advect_velocity_multi<D_, VF_, SF_, false, true><<<g, b, 0, s>>>(m.start, m.n, E);
        else k_advect_velocity_multi<D_, VF_, SF_, false, false><<<g, b, 0, s>>>(m.start, m.n, E);
    })));
}
void launch_vorticity_multi(hipStream_t s, int dim, int vfmt, const MultiLaunch &m, const MVort *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_vorticity_multi<D_, VF_><<<multi_grid(m), multi_block_dim(m), 0, s>>>(m.start, m.n, E)));
}
void launch_apply_vorticity_multi(hipStream_t s, int dim, int vfmt, const MultiLaunch &m, const MApplyVort *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_apply_vorticity_multi<D_, VF_>
                                        <<<multi_grid(m), multi_block_dim(m), 0, s>>>(m.start, m.n, E)));
}
void launch_update_velocity_multi(hipStream_t s, int dim, int vfmt, const MultiLaunch &m, const MUpdateVel *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_update_velocity_multi<D_, VF_>
                                        <<<multi_grid(m), multi_block_dim(m), 0, s>>>(m.start, m.n, E)));
}
void launch_advect_smoke_multi(hipStream_t s, int dim, int vfmt, int sfmt, bool buf, const MultiLaunch &m,
                               const MAdvectSmoke *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    const dim3 g = multi_grid(m), b = multi_block_dim(m);
    DISPATCH_D(P, DISPATCH_VF(vfmt, DISPATCH_SF(sfmt, {
        if (buf) k_advect_smoke_multi<D_, VF_, SF_, true><<<g, b, 0, s>>>(m.start, m.n, E);
        else k_advect_smoke_multi<D_, VF_, SF_, false><<<g, b, 0, s>>>(m.start, m.n, E);
    })));
}
void launch_obstacle_update_multi(hipStream_t s, int dim, const MultiLaunch &m, const MObstacles *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    DISPATCH_D(P, k_obstacle_update_multi<D_><<<multi_grid(m), m.bx, 0, s>>>(m.start, m.n, E));
}
bool buf_samplers_ok(const SimParams &P) { return tex_buf_ok(P); }
void launch_set_times(hipStream_t s, float *const *times, int n, float t) {
    if (n <= 0) return;
    k_set_times<<<(n + 63) / 64, 64, 0, s>>>(times, n, t);
}
SMK_NS_END
