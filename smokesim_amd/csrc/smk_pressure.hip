// smk_pressure.hip -- red-black multigrid pressure solve (gfx950).
//
// Restates pressure_solve/{smooth,compute_residual,restrict_residual,
// restrict_solid,prolong_pressure}.comp (3D and 2D) and the V-cycle data
// flow of grid.c:173-266 on an MI355X-shaped layout:
//   - pressure and right-hand side are colour-split (struct Level), so a
//     half sweep streams the other colour's pressure, its own right-hand
//     side and mask, and writes its own pressure, all with unit stride;
//   - the seven solid-fraction loads per cell of the reference become one
//     mask byte per cell (k_build_mask), exact whenever the cell and its
//     neighbours are fully solid or fully fluid; cells with a fractional
//     neighbour (coarse levels near obstacles) or on the domain boundary
//     take the general path, which reads the solid map like the reference;
//   - the residual of a level is computed and restricted in one pass, and
//     the coarse pressure is cleared in the same pass (grid.c:193-200).
// Every update keeps the reference's operation order (no FMA contraction).
#include "smk_device.h"

namespace smk {

size_t split_half(const Level &L) { return (size_t)L.hx * (size_t)L.n[1] * (size_t)L.n[2]; }

// Mask bits.  A neighbour bit is set when that neighbour's solid fraction is
// exactly 0 (its stencil weight f = 1 - a is then exactly 1).
enum : uint32_t {
    M_SOLID = 1u,    // a >= 1: the cell is skipped (smooth.comp:84)
    M_TOP = 2u,      // +y
    M_BOT = 4u,      // -y
    M_RIGHT = 8u,    // +x
    M_LEFT = 16u,    // -x
    M_FRONT = 32u,   // +z
    M_BACK = 64u,    // -z
    M_GENERAL = 128u // fractional a nearby or domain boundary: general path
};

template <int D>
__device__ __forceinline__ float a_at(const Level &L, int i, int j, int k) {
    return solid_a(L.s, lin(i, j, k, L.n[0], L.n[1]));
}

template <int D>
__global__ void k_build_mask(Level L) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    if (!inside(i, j, k, sx, sy, sz)) return;
    float a = a_at<D>(L, i, j, k);
    uint32_t m = a >= 1.0f ? M_SOLID : 0u;
    bool general = !(a == 0.0f || a == 1.0f);
    general |= i == 0 || j == 0 || i == sx - 1 || j == sy - 1;
    if (D == 3) general |= k == 0 || k == sz - 1;
    const int off[6][3] = {{0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int q = 0; q < 2 * D; ++q) {
        int ii = i + off[q][0], jj = j + off[q][1], kk = k + off[q][2];
        if (!inside(ii, jj, kk, sx, sy, sz)) continue;  // boundary cell: general
        float an = a_at<D>(L, ii, jj, kk);
        if (!(an == 0.0f || an == 1.0f)) general = true;
        if (an == 0.0f) m |= (M_TOP << q);
    }
    if (general) m |= M_GENERAL;
    L.mask[sidx(L, i, j, k)] = (uint8_t)m;
}

// Level-0 divergence sum of the smoother (OOB neighbour = fluid).
template <int D, int VF>
__global__ void k_divergence0(SimParams P, VelPtrs V, const uint16_t *s0, Level L0, float *dv0) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    dv0[sidx(L0, i, j, k)] = level0_dv<D, VF>(P, V, s0, i, j, k, 0.0f);
}

// Pressure of a neighbour with the OOB value 0 (get_pressure()).
__device__ __forceinline__ float p_at(const Level &L, int i, int j, int k) {
    return inside(i, j, k, L.n[0], L.n[1], L.n[2]) ? L.p[sidx(L, i, j, k)] : 0.0f;
}

// One red-black half sweep (smooth.comp 3D:79-135 / 2D:74-122).  Thread
// (m, j, k) owns cell i = 2m + ((parity + j + k) & 1) of the sweep's colour.
template <int D>
__global__ void k_sweep(SimParams P, Level L, int level, int parity, const float *rhs, float hl) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2], hx = L.hx;
    if (m >= hx || j >= sy || k >= sz) return;
    const int o = (parity + j + k) & 1;
    const int i = 2 * m + o;
    if (i >= sx) return;
    const size_t half = split_half_d(L);
    const size_t row = ((size_t)k * sy + j) * hx;
    const size_t own = (size_t)parity * half + row + m;
    const uint32_t msk = L.mask[own];
    if (msk & M_SOLID) return;
    float ft, fb, fr, fl, ff = 0.f, fk = 0.f, fsum, pt, pb, pr, pl, pf = 0.f, pk = 0.f;
    if (!(msk & M_GENERAL)) {
        // interior cell, all fractions 0 or 1: neighbours are the other
        // colour at fixed offsets
        const float *po = L.p + (size_t)(1 - parity) * half + row;
        ft = (msk & M_TOP) ? 1.0f : 0.0f;
        fb = (msk & M_BOT) ? 1.0f : 0.0f;
        fr = (msk & M_RIGHT) ? 1.0f : 0.0f;
        fl = (msk & M_LEFT) ? 1.0f : 0.0f;
        pt = po[m + hx];
        pb = po[(ptrdiff_t)m - hx];
        pr = po[m + o];
        pl = po[m - 1 + o];
        if (D == 3) {
            const ptrdiff_t plane = (ptrdiff_t)sy * hx;
            ff = (msk & M_FRONT) ? 1.0f : 0.0f;
            fk = (msk & M_BACK) ? 1.0f : 0.0f;
            pf = po[m + plane];
            pk = po[(ptrdiff_t)m - plane];
        }
    } else {
        // smooth.comp solid_data(): OOB texel (0,0,0,0) -> f = 1, p = 0
        auto A = [&](int ii, int jj, int kk) -> float {
            return inside(ii, jj, kk, sx, sy, sz) ? a_at<D>(L, ii, jj, kk) : 0.0f;
        };
        ft = 1.0f - A(i, j + 1, k);
        fb = 1.0f - A(i, j - 1, k);
        fr = 1.0f - A(i + 1, j, k);
        fl = 1.0f - A(i - 1, j, k);
        pt = p_at(L, i, j + 1, k);
        pb = p_at(L, i, j - 1, k);
        pr = p_at(L, i + 1, j, k);
        pl = p_at(L, i - 1, j, k);
        if (D == 3) {
            ff = 1.0f - A(i, j, k + 1);
            fk = 1.0f - A(i, j, k - 1);
            pf = p_at(L, i, j, k + 1);
            pk = p_at(L, i, j, k - 1);
        }
    }
    if (D == 3) fsum = ft + fb + fr + fl + ff + fk;
    else fsum = ft + fb + fr + fl;
    if (fsum == 0.0f) return;
    const float dv = rhs[own];
    float psum = pt * ft + pb * fb;
    psum = psum + pr * fr;
    psum = psum + pl * fl;
    if (D == 3) {
        psum = psum + pf * ff;
        psum = psum + pk * fk;
    }
    float np;
    if (level == 0) np = (psum - P.rho * hl * dv / P.dt) / fsum;
    else np = (psum - dv * (hl * hl)) / fsum;
    L.p[own] = np;
}

// r = b - A p at one cell (compute_residual.comp 3D:77-130 / 2D:72-117):
// solid cells 0; OOB neighbours solid (f = 0).  On level 0 b comes from the
// face divergence: interior cells reuse dv0 (same value under both OOB
// conventions), boundary cells recompute it with this kernel's convention.
template <int D, int VF>
__device__ float residual_at(const SimParams &P, const Level &L, int level, const float *rhs,
                             const VelPtrs &V, const uint16_t *s0, int i, int j, int k, float hl) {
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    const size_t own = sidx(L, i, j, k);
    const uint32_t msk = L.mask[own];
    if (msk & M_SOLID) return 0.0f;
    float psum, fsum;
    if (!(msk & M_GENERAL)) {
        const float ft = (msk & M_TOP) ? 1.0f : 0.0f, fb = (msk & M_BOT) ? 1.0f : 0.0f;
        const float fr = (msk & M_RIGHT) ? 1.0f : 0.0f, fl = (msk & M_LEFT) ? 1.0f : 0.0f;
        psum = L.p[sidx(L, i, j + 1, k)] * ft + L.p[sidx(L, i, j - 1, k)] * fb;
        psum = psum + L.p[sidx(L, i + 1, j, k)] * fr;
        psum = psum + L.p[sidx(L, i - 1, j, k)] * fl;
        if (D == 3) {
            const float ff = (msk & M_FRONT) ? 1.0f : 0.0f, fk = (msk & M_BACK) ? 1.0f : 0.0f;
            psum = psum + L.p[sidx(L, i, j, k + 1)] * ff;
            psum = psum + L.p[sidx(L, i, j, k - 1)] * fk;
            fsum = ft + fb + fr + fl + ff + fk;
        } else {
            fsum = ft + fb + fr + fl;
        }
    } else {
        auto A = [&](int ii, int jj, int kk) -> float {
            return inside(ii, jj, kk, sx, sy, sz) ? a_at<D>(L, ii, jj, kk) : 1.0f;
        };
        const float ft = 1.0f - A(i, j + 1, k), fb = 1.0f - A(i, j - 1, k);
        const float fr = 1.0f - A(i + 1, j, k), fl = 1.0f - A(i - 1, j, k);
        psum = p_at(L, i, j + 1, k) * ft + p_at(L, i, j - 1, k) * fb;
        psum = psum + p_at(L, i + 1, j, k) * fr;
        psum = psum + p_at(L, i - 1, j, k) * fl;
        if (D == 3) {
            const float ff = 1.0f - A(i, j, k + 1), fk = 1.0f - A(i, j, k - 1);
            psum = psum + p_at(L, i, j, k + 1) * ff;
            psum = psum + p_at(L, i, j, k - 1) * fk;
            fsum = ft + fb + fr + fl + ff + fk;
        } else {
            fsum = ft + fb + fr + fl;
        }
    }
    float old;
    if (level == 0) {
        const bool edge = i == 0 || j == 0 || i == sx - 1 || j == sy - 1 ||
                          (D == 3 && (k == 0 || k == sz - 1));
        const float dv = edge ? level0_dv<D, VF>(P, V, s0, i, j, k, 1.0f) : rhs[own];
        old = P.rho * dv / (P.dt * hl);
    } else {
        old = rhs[own];
    }
    return old - (psum - fsum * L.p[own]) / (hl * hl);
}

// Residual of level `level` for the 2x2(x2) block of fine cells under one
// coarse cell, stored in place (the reference's residual image), then the
// restriction of restrict_residual.comp (3D sum order: the four children at
// y = 2j, then the four at y = 2j+1) and the clear of the coarse pressure.
// Threads past the coarse size cover trailing fine cells of odd sizes.
template <int D, int VF>
__global__ void k_residual_restrict(SimParams P, Level F, Level C, int level, const float *rhs,
                                    VelPtrs V, const uint16_t *s0, float hl, int has_coarse) {
    const int ci = blockIdx.x * blockDim.x + threadIdx.x;
    const int cj = blockIdx.y * blockDim.y + threadIdx.y;
    const int ck = blockIdx.z;
    const int I = 2 * ci, J = 2 * cj, K = D == 3 ? 2 * ck : 0;
    if (I >= F.n[0] || J >= F.n[1] || K >= F.n[2]) return;
    float r[2][2][2] = {{{0.f, 0.f}, {0.f, 0.f}}, {{0.f, 0.f}, {0.f, 0.f}}};  // [dk][dj][di]
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int i = I + di, j = J + dj, k = K + dk;
                if (!inside(i, j, k, F.n[0], F.n[1], F.n[2])) continue;
                const float v = residual_at<D, VF>(P, F, level, rhs, V, s0, i, j, k, hl);
                F.r[sidx(F, i, j, k)] = v;
                r[dk][dj][di] = v;
            }
    if (!has_coarse || !inside(ci, cj, ck, C.n[0], C.n[1], C.n[2])) return;
    float v;
    if (D == 3) {
        float bsum = r[0][0][0] + r[1][0][0];
        bsum = bsum + r[0][0][1];
        bsum = bsum + r[1][0][1];
        float tsum = r[0][1][0] + r[1][1][0];
        tsum = tsum + r[0][1][1];
        tsum = tsum + r[1][1][1];
        v = (bsum + tsum) / 8.0f;
    } else {
        float sum = r[0][0][0] + r[0][0][1];
        sum = sum + r[0][1][0];
        sum = sum + r[0][1][1];
        v = sum / 4.0f;
    }
    const size_t c = sidx(C, ci, cj, ck);
    C.r[c] = v;
    C.p[c] = 0.0f;
}

// restrict_solid.comp 3D:53-78 / 2D:53-73 (natural layout, 4 x fp16).
template <int D>
__global__ void k_restrict_solid(Level F, Level C) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (!inside(i, j, k, C.n[0], C.n[1], C.n[2])) return;
    const int fx = F.n[0], fy = F.n[1];
    int I = 2 * i, J = 2 * j, K = D == 3 ? 2 * k : 0;
    float4 o;
    auto S = [&](int a, int b, int c) { return ld4<FMT_F16>(F.s, lin(a, b, c, fx, fy)); };
    auto add = [](float4 a, float4 b) {
        return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    };
    if (D == 3) {
        float4 b = add(add(add(S(I, J, K), S(I, J, K + 1)), S(I + 1, J, K)), S(I + 1, J, K + 1));
        float4 t = add(add(add(S(I, J + 1, K), S(I, J + 1, K + 1)), S(I + 1, J + 1, K)),
                       S(I + 1, J + 1, K + 1));
        float4 s = add(b, t);
        o = make_float4(s.x / 8.0f, s.y / 8.0f, s.z / 8.0f, s.w / 8.0f);
    } else {
        float4 s = add(add(add(S(I, J, 0), S(I + 1, J, 0)), S(I, J + 1, 0)), S(I + 1, J + 1, 0));
        o = make_float4(s.x / 4.0f, s.y / 4.0f, s.z / 4.0f, s.w / 4.0f);
    }
    float m = o.w >= 0.5f ? 1.0f : 0.0f;
    o = make_float4(o.x * m, o.y * m, o.z * m, o.w * m);
    st4<FMT_F16>(C.s, lin(i, j, k, C.n[0], C.n[1]), o);
}

// CLAMP_TO_EDGE LINEAR sample of a colour-split pressure level.
template <int D>
__device__ float sample_edge_split(const Level &C, const float *uv) {
    const int sx = C.n[0], sy = C.n[1], sz = C.n[2];
    int i0, j0, k0 = 0;
    float a, b, g = 0.f;
    texcoord(uv[0], sx, i0, a);
    texcoord(uv[1], sy, j0, b);
    if (D == 3) texcoord(uv[2], sz, k0, g);
    auto F = [&](int i, int j, int k) -> float {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        k = min(max(k, 0), sz - 1);
        return C.p[sidx(C, i, j, k)];
    };
    if (D == 2) {
        float x0 = lerpf(F(i0, j0, 0), F(i0 + 1, j0, 0), a);
        float x1 = lerpf(F(i0, j0 + 1, 0), F(i0 + 1, j0 + 1, 0), a);
        return lerpf(x0, x1, b);
    }
    float x00 = lerpf(F(i0, j0, k0), F(i0 + 1, j0, k0), a);
    float x10 = lerpf(F(i0, j0 + 1, k0), F(i0 + 1, j0 + 1, k0), a);
    float x01 = lerpf(F(i0, j0, k0 + 1), F(i0 + 1, j0, k0 + 1), a);
    float x11 = lerpf(F(i0, j0 + 1, k0 + 1), F(i0 + 1, j0 + 1, k0 + 1), a);
    float y0 = lerpf(x00, x10, b);
    float y1 = lerpf(x01, x11, b);
    return lerpf(y0, y1, g);
}

// prolong_pressure.comp 3D:39-51 / 2D:39-51.
template <int D>
__global__ void k_prolong(Level F, Level C) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    const int sx = F.n[0], sy = F.n[1], sz = F.n[2];
    if (!inside(i, j, k, sx, sy, sz)) return;
    float uv[3] = {((float)i + 0.5f) / (float)sx, ((float)j + 0.5f) / (float)sy,
                   ((float)k + 0.5f) / (float)sz};
    float c = sample_edge_split<D>(C, uv);
    size_t cell = sidx(F, i, j, k);
    F.p[cell] = F.p[cell] + c;
}

// Red-black iterations of a small level by one workgroup with pressure and
// solid fraction in LDS (natural order).  Each half sweep updates one
// colour from the other, so the result equals the separate dispatches of
// grid.c:228-233.
constexpr int kCoarseMaxCells = 6144;  // 6144 * (4 B p + 4 B a) = 48 KiB LDS
template <int D>
__global__ void __launch_bounds__(1024) k_coarse_smooth(SimParams P, Level L, int level,
                                                        int iterations, const float *rhs,
                                                        float hl) {
    __shared__ float sp[kCoarseMaxCells];
    __shared__ float sa[kCoarseMaxCells];
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    const int n = sx * sy * sz;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        sp[c] = L.p[sidx(L, i, j, k)];
        sa[c] = solid_a(L.s, c);
    }
    __syncthreads();
    auto A = [&](int ii, int jj, int kk) -> float {
        return inside(ii, jj, kk, sx, sy, sz) ? sa[lin(ii, jj, kk, sx, sy)] : 0.0f;
    };
    auto Pn = [&](int ii, int jj, int kk) -> float {
        return inside(ii, jj, kk, sx, sy, sz) ? sp[lin(ii, jj, kk, sx, sy)] : 0.0f;
    };
    int parity = 0;
    for (int it = 0; it < iterations * 2; ++it) {
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
            if ((i + j + k) % 2 != parity) continue;
            if (sa[c] >= 1.0f) continue;
            float ft = 1.0f - A(i, j + 1, k);
            float fb = 1.0f - A(i, j - 1, k);
            float fr = 1.0f - A(i + 1, j, k);
            float fl = 1.0f - A(i - 1, j, k);
            float ff = 0.f, fk = 0.f, fsum;
            if (D == 3) {
                ff = 1.0f - A(i, j, k + 1);
                fk = 1.0f - A(i, j, k - 1);
                fsum = ft + fb + fr + fl + ff + fk;
            } else {
                fsum = ft + fb + fr + fl;
            }
            if (fsum == 0.0f) continue;
            float dv = rhs[sidx(L, i, j, k)];
            float psum = Pn(i, j + 1, k) * ft + Pn(i, j - 1, k) * fb;
            psum = psum + Pn(i + 1, j, k) * fr;
            psum = psum + Pn(i - 1, j, k) * fl;
            if (D == 3) {
                psum = psum + Pn(i, j, k + 1) * ff;
                psum = psum + Pn(i, j, k - 1) * fk;
            }
            float np;
            if (level == 0) np = (psum - P.rho * hl * dv / P.dt) / fsum;
            else np = (psum - dv * (hl * hl)) / fsum;
            sp[c] = np;
        }
        __syncthreads();
        parity = 1 - parity;
    }
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        L.p[sidx(L, i, j, k)] = sp[c];
    }
}

__global__ void k_split_to_f64(Level L, const float *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(c % L.n[0]), j = (int)((c / L.n[0]) % L.n[1]), k = (int)(c / ((size_t)L.n[0] * L.n[1]));
        dst[c] = (double)src[sidx(L, i, j, k)];
    }
}
__global__ void k_split_from_f64(Level L, const double *src, float *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(c % L.n[0]), j = (int)((c / L.n[0]) % L.n[1]), k = (int)(c / ((size_t)L.n[0] * L.n[1]));
        dst[sidx(L, i, j, k)] = (float)src[c];
    }
}

// Level-0 residual norm: the residual stencil accumulated in FP64.
template <int D, int VF>
__global__ void k_residual_norm(SimParams P, Level L, VelPtrs V, double *acc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    double r2 = 0.0;
    if (inside(i, j, k, sx, sy, sz) && a_at<D>(L, i, j, k) < 1.0f) {
        auto A = [&](int ii, int jj, int kk) -> double {
            return inside(ii, jj, kk, sx, sy, sz) ? (double)a_at<D>(L, ii, jj, kk) : 1.0;
        };
        const int off[6][3] = {{0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
        double psum = 0.0, fsum = 0.0;
        for (int q = 0; q < 2 * D; ++q) {
            int ii = i + off[q][0], jj = j + off[q][1], kk = k + off[q][2];
            double f = 1.0 - A(ii, jj, kk);
            psum += (double)p_at(L, ii, jj, kk) * f;
            fsum += f;
        }
        double h = (double)P.h;
        double dv = (double)level0_dv<D, VF>(P, V, L.s, i, j, k, 1.0f);
        double b = (double)P.rho * dv / ((double)P.dt * h);
        double r = b - (psum - fsum * (double)p_at(L, i, j, k)) / (h * h);
        r2 = r * r;
    }
    for (int o = 32; o > 0; o >>= 1) r2 += __shfl_down(r2, o, 64);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0 && r2 != 0.0) atomicAdd(acc, r2);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static float level_h(const SimParams &P, int level) { return P.h * (float)(1 << level); }

static unsigned flat_blocks_p(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 65535 * 4 ? 65535 * 4 : (b == 0 ? 1 : b));
}

void launch_build_mask(hipStream_t s, const SimParams &P, const Level &L) {
    dim3 g = grid_for(L.n[0], L.n[1], L.n[2], kBlk);
    DISPATCH_D(P, k_build_mask<D_><<<g, kBlk, 0, s>>>(L));
}

void launch_divergence0(hipStream_t s, const SimParams &P, int vfmt, VelPtrs vel,
                        const uint16_t *solid0, const Level &L0, float *dv0) {
    dim3 g = grid_for(P.n[0], P.n[1], P.n[2], kBlk);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_divergence0<D_, VF_><<<g, kBlk, 0, s>>>(P, vel, solid0, L0, dv0)));
}

void launch_sweep(hipStream_t s, const SimParams &P, const Level &L, int level, int parity,
                  const float *rhs) {
    dim3 g = grid_for(L.hx, L.n[1], L.n[2], kBlk);
    float hl = level_h(P, level);
    DISPATCH_D(P, k_sweep<D_><<<g, kBlk, 0, s>>>(P, L, level, parity, rhs, hl));
}

void launch_residual_restrict(hipStream_t s, const SimParams &P, const Level &fine,
                              const Level &coarse, int level, const float *rhs, int vfmt,
                              VelPtrs vel, const uint16_t *solid0) {
    dim3 g = grid_for((fine.n[0] + 1) / 2, (fine.n[1] + 1) / 2,
                      P.dim == 3 ? (fine.n[2] + 1) / 2 : 1, kBlk);
    float hl = level_h(P, level);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_restrict<D_, VF_><<<g, kBlk, 0, s>>>(
                                       P, fine, coarse, level, rhs, vel, solid0, hl, 1)));
}

void launch_residual(hipStream_t s, const SimParams &P, const Level &L, int level,
                     const float *rhs, int vfmt, VelPtrs vel, const uint16_t *solid0) {
    dim3 g = grid_for((L.n[0] + 1) / 2, (L.n[1] + 1) / 2, P.dim == 3 ? (L.n[2] + 1) / 2 : 1, kBlk);
    float hl = level_h(P, level);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_restrict<D_, VF_><<<g, kBlk, 0, s>>>(
                                       P, L, L, level, rhs, vel, solid0, hl, 0)));
}

void launch_restrict_solid(hipStream_t s, const SimParams &P, const Level &fine,
                           const Level &coarse) {
    dim3 g = grid_for(coarse.n[0], coarse.n[1], coarse.n[2], kBlk);
    DISPATCH_D(P, k_restrict_solid<D_><<<g, kBlk, 0, s>>>(fine, coarse));
}

void launch_prolong(hipStream_t s, const SimParams &P, const Level &fine, const Level &coarse) {
    dim3 g = grid_for(fine.n[0], fine.n[1], fine.n[2], kBlk);
    DISPATCH_D(P, k_prolong<D_><<<g, kBlk, 0, s>>>(fine, coarse));
}

bool coarse_fits(const SimParams &P, const Level &L) {
    (void)P;
    return (long)L.n[0] * L.n[1] * L.n[2] <= kCoarseMaxCells;
}

void launch_coarse_smooth(hipStream_t s, const SimParams &P, const Level &L, int level,
                          int iterations, const float *rhs) {
    float hl = level_h(P, level);
    int n = L.n[0] * L.n[1] * L.n[2];
    int threads = n >= 1024 ? 1024 : ((n + 63) / 64) * 64;
    DISPATCH_D(P, k_coarse_smooth<D_><<<1, threads, 0, s>>>(P, L, level, iterations, rhs, hl));
}

void launch_split_to_f64(hipStream_t s, const Level &L, const float *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    k_split_to_f64<<<flat_blocks_p(n), 256, 0, s>>>(L, src, dst);
}
void launch_split_from_f64(hipStream_t s, const Level &L, const double *src, float *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    k_split_from_f64<<<flat_blocks_p(n), 256, 0, s>>>(L, src, dst);
}

void launch_residual_norm(hipStream_t s, const SimParams &P, const Level &L0, int vfmt,
                          VelPtrs vel, double *acc) {
    dim3 g = grid_for(L0.n[0], L0.n[1], L0.n[2], kBlk);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_norm<D_, VF_><<<g, kBlk, 0, s>>>(P, L0, vel, acc)));
}

}  // namespace smk
