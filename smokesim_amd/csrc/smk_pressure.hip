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
#include <hip/hip_ext.h>

#include <cmath>

#include "smk_device.h"

SMK_NS_BEGIN

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
    M_GENERAL = 128u // fluid cell with a fractional a nearby or on the domain
                     // boundary: general path
};

template <int D>
__device__ __forceinline__ float a_at(const Level &L, int i, int j, int k) {
    return solid_a(L.s, lin(i, j, k, L.n[0], L.n[1]));
}
// solid fraction with an OOB value, branch free
__device__ __forceinline__ float a_sel(const Level &L, bool ok, int i, int j, int k, float oob) {
    const uint16_t b = ld_sel(L.s, ok, lin(i, j, k, L.n[0], L.n[1]) * 4 + 3, (uint16_t)0);
    return ok ? f16_to_f32(b) : oob;
}

template <int D>
__global__ void k_build_mask(XG X_, Level L) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
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
    // solid cells are skipped by the smoother and have residual 0: they never
    // need the general path
    if (general && !(m & M_SOLID)) m |= M_GENERAL;
    L.mask[sidx(L, i, j, k)] = (uint8_t)m;
}

// Level-0 divergence sum of the smoother (OOB neighbour = fluid), and the
// smoother's right-hand-side term rho*h*dv/dt (smooth.comp:129), which is
// the same for every sweep of a solve.  Needs the level-0 neighbour masks.
template <int D, int VF>
__device__ __forceinline__ void divergence0_at(int ti_, int tj_, int tk_, const SimParams &P, const VelPtrs &V, const uint16_t *s0, const Level &L0, float *dv0, float *b0) {
    int i = ti_;
    int j = tj_;
    int k = tk_;
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return;
    const uint32_t c = sidx(L0, i, j, k);
    const uint32_t m = L0.mask[c];
    float dv;
    if (m & M_SOLID) {
        // never read: the smoother skips solid cells, their residual is 0
        dv = 0.0f;
    } else if (m & M_GENERAL) {
        dv = level0_dv<D, VF>(P, V, s0, i, j, k, 0.0f);
    } else {
        // interior cell with binary neighbours: the mask tells which
        // neighbours are solid, and only their faces read the obstacle
        // velocity (the solid texel's channel of the face's axis)
        const uint32_t nx = P.n[0], ny = P.n[1];
        const uint32_t cell = lin(i, j, k, nx, ny);
        auto solid_v = [&](uint32_t ncell, int ch) { return f16_to_f32(s0[ncell * 4 + ch]); };
        const float ft = vget<D, VF>(P, V, 1, i, j + 1, k), fb = vget<D, VF>(P, V, 1, i, j, k);
        const float fr = vget<D, VF>(P, V, 0, i + 1, j, k), fl = vget<D, VF>(P, V, 0, i, j, k);
        float vtop = ft, vbot = fb, vr = fr, vl = fl;
        if (!(m & M_TOP)) vtop = solid_v(cell + nx, 1);
        if (!(m & M_BOT)) vbot = solid_v(cell - nx, 1);
        if (!(m & M_RIGHT)) vr = solid_v(cell + 1, 0);
        if (!(m & M_LEFT)) vl = solid_v(cell - 1, 0);
        dv = vr - vl + vtop - vbot;
        if (D == 3) {
            const float ff = vget<D, VF>(P, V, 2, i, j, k + 1), fk = vget<D, VF>(P, V, 2, i, j, k);
            float vf = ff, vb = fk;
            if (!(m & M_FRONT)) vf = solid_v(cell + nx * ny, 2);
            if (!(m & M_BACK)) vb = solid_v(cell - nx * ny, 2);
            dv = dv + vf - vb;
        }
    }
    dv0[c] = dv;
    b0[c] = P.rho * P.h * dv / P.dt;
}

template <int D, int VF>
__device__ __forceinline__ void divergence0_body(int bx_, int by_, int bz_, const SimParams &P, const VelPtrs &V, const uint16_t *s0, const Level &L0, float *dv0, float *b0) {
    divergence0_at<D, VF>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, V, s0, L0, dv0, b0);
}
template <int D, int VF>
__global__ void k_divergence0(XG X_, SimParams P, VelPtrs V, const uint16_t *s0, Level L0, float *dv0, float *b0) {
    XG_BLOCK(X_)
    divergence0_body<D, VF>(bx_, by_, bz_, P, V, s0, L0, dv0, b0);
}

// k_divergence0 for 3D FP32 faces with nx % 4 == 0: four consecutive cells
// of a row per thread, 16-byte loads of the y and z faces, the row's five x
// faces, two bytes of mask and one float2 of dv0 / b0 per colour; the same
// arithmetic per cell.
// four consecutive faces of one row (16-byte FP32 or 8-byte 16-bit load)
template <int VF>
__device__ __forceinline__ float4 ld_row4(const void *base, uint32_t idx) {
    if constexpr (VF == FMT_F32) {
        return *reinterpret_cast<const float4 *>(static_cast<const float *>(base) + idx);
    } else {
        const q16 q = *reinterpret_cast<const q16 *>(static_cast<const uint16_t *>(base) + idx);
        using S = Store<VF>;
        return make_float4(S::ld(q.x), S::ld(q.y), S::ld(q.z), S::ld(q.w));
    }
}

template <int VF>
__device__ __forceinline__ void divergence0_x4_at(int ti_, int tj_, int tk_, const SimParams &P, const VelPtrs &V, const uint16_t *s0, const Level &L0, float *dv0, float *b0) {
    const int i0 = 4 * ti_;
    const int j = tj_;
    const int k = tk_;
    const int nx = P.n[0], ny = P.n[1], nz = P.n[2];
    if (i0 >= nx || j >= ny || k >= nz) return;
    using S = Store<VF>;
    const typename S::T *vx = static_cast<const typename S::T *>(V.c[0]);
    const uint32_t rx = ((uint32_t)k * ny + j) * (uint32_t)(nx + 1) + i0;
    float fx[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) fx[q] = S::ld(vx[rx + q]);
    const uint32_t ry = ((uint32_t)k * (ny + 1) + j) * (uint32_t)nx + i0;
    const float4 yb = ld_row4<VF>(V.c[1], ry);
    const float4 yt = ld_row4<VF>(V.c[1], ry + nx);
    const uint32_t rz = ((uint32_t)k * ny + j) * (uint32_t)nx + i0;
    const float4 zk = ld_row4<VF>(V.c[2], rz);
    const float4 zf = ld_row4<VF>(V.c[2], rz + (uint32_t)nx * ny);
    const float fb4[4] = {yb.x, yb.y, yb.z, yb.w}, ft4[4] = {yt.x, yt.y, yt.z, yt.w};
    const float fk4[4] = {zk.x, zk.y, zk.z, zk.w}, ff4[4] = {zf.x, zf.y, zf.z, zf.w};
    // cells i0, i0+2 have colour c0 (row slots m, m+1), i0+1, i0+3 the other
    const int c0 = (j + k) & 1;
    const uint32_t half = split_half_d(L0);
    const uint32_t row = ((uint32_t)k * ny + j) * (uint32_t)L0.hx + (uint32_t)(i0 >> 1);
    const uint32_t a0 = (uint32_t)c0 * half + row, a1 = (uint32_t)(1 - c0) * half + row;
    const uint16_t mk0 = *reinterpret_cast<const uint16_t *>(L0.mask + a0);
    const uint16_t mk1 = *reinterpret_cast<const uint16_t *>(L0.mask + a1);
    const uint32_t mq[4] = {mk0 & 0xffu, mk1 & 0xffu, (uint32_t)mk0 >> 8, (uint32_t)mk1 >> 8};
    float dvq[4], bq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t m = mq[q];
        const int i = i0 + q;
        float dv;
        if (m & M_SOLID) {
            dv = 0.0f;
        } else if (m & M_GENERAL) {
            dv = level0_dv<3, VF>(P, V, s0, i, j, k, 0.0f);
        } else {
            const uint32_t cell = lin(i, j, k, nx, ny);
            auto solid_v = [&](uint32_t ncell, int ch) { return f16_to_f32(s0[ncell * 4 + ch]); };
            float vtop = ft4[q], vbot = fb4[q], vr = fx[q + 1], vl = fx[q];
            if (!(m & M_TOP)) vtop = solid_v(cell + nx, 1);
            if (!(m & M_BOT)) vbot = solid_v(cell - nx, 1);
            if (!(m & M_RIGHT)) vr = solid_v(cell + 1, 0);
            if (!(m & M_LEFT)) vl = solid_v(cell - 1, 0);
            dv = vr - vl + vtop - vbot;
            float vf = ff4[q], vb = fk4[q];
            if (!(m & M_FRONT)) vf = solid_v(cell + (uint32_t)nx * ny, 2);
            if (!(m & M_BACK)) vb = solid_v(cell - (uint32_t)nx * ny, 2);
            dv = dv + vf - vb;
        }
        dvq[q] = dv;
        bq[q] = P.rho * P.h * dv / P.dt;
    }
    *reinterpret_cast<float2 *>(dv0 + a0) = make_float2(dvq[0], dvq[2]);
    *reinterpret_cast<float2 *>(dv0 + a1) = make_float2(dvq[1], dvq[3]);
    *reinterpret_cast<float2 *>(b0 + a0) = make_float2(bq[0], bq[2]);
    *reinterpret_cast<float2 *>(b0 + a1) = make_float2(bq[1], bq[3]);
}

template <int VF>
__device__ __forceinline__ void divergence0_x4_body(int bx_, int by_, int bz_, const SimParams &P, const VelPtrs &V, const uint16_t *s0, const Level &L0, float *dv0, float *b0) {
    divergence0_x4_at<VF>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, V, s0, L0, dv0, b0);
}
template <int VF>
__global__ void k_divergence0_x4(XG X_, SimParams P, VelPtrs V, const uint16_t *s0, Level L0, float *dv0, float *b0) {
    XG_BLOCK(X_)
    divergence0_x4_body<VF>(bx_, by_, bz_, P, V, s0, L0, dv0, b0);
}

// Pressure of a neighbour with the OOB value 0 (get_pressure()).
__device__ __forceinline__ float p_at(const Level &L, int i, int j, int k) {
    return ld_sel(L.p, inside(i, j, k, L.n[0], L.n[1], L.n[2]), sidx(L, i, j, k), 0.0f);
}

// One red-black half sweep (smooth.comp 3D:79-135 / 2D:74-122).  Thread
// (m, j, k) owns cell i = 2m + ((parity + j + k) & 1) of the sweep's colour.
template <int D>
__device__ __forceinline__ void sweep_at(int m, int j, int k, const SimParams &P, const Level &L, int level,
                                         int parity, const float *rhs, float hl) {
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2], hx = L.hx;
    if (m >= hx || j >= sy || k >= sz) return;
    const int o = (parity + j + k) & 1;
    const int i = 2 * m + o;
    if (i >= sx) return;
    const uint32_t half = split_half_d(L);
    const uint32_t row = ((uint32_t)k * sy + j) * hx;
    const uint32_t own = (size_t)parity * half + row + m;
    // the mask, the right-hand side and the six interior-stencil neighbours
    // load together (one memory round trip: the small levels are latency
    // bound); off the interior the neighbour loads read element 0 and are
    // discarded, those cells being M_GENERAL
    const uint32_t msk = L.mask[own];
    const float dv = rhs[own];
    const bool gi = j > 0 && j < sy - 1 && (D == 2 || (k > 0 && k < sz - 1)) && m - 1 + o >= 0 &&
                    m + o < hx;
    const uint32_t ob = (uint32_t)(1 - parity) * half + row + m;
    const uint32_t plane = (uint32_t)sy * hx;
    const float qt = ld_sel(L.p, gi, ob + hx, 0.0f), qb = ld_sel(L.p, gi, ob - hx, 0.0f);
    const float qr = ld_sel(L.p, gi, ob + o, 0.0f), ql = ld_sel(L.p, gi, ob - 1 + o, 0.0f);
    const float qf = D == 3 ? ld_sel(L.p, gi, ob + plane, 0.0f) : 0.0f;
    const float qk = D == 3 ? ld_sel(L.p, gi, ob - plane, 0.0f) : 0.0f;
    if (msk & M_SOLID) return;
    float ft, fb, fr, fl, ff = 0.f, fk = 0.f, fsum, pt, pb, pr, pl, pf = 0.f, pk = 0.f;
    if (!(msk & M_GENERAL)) {
        // interior cell, all fractions 0 or 1: neighbours are the other
        // colour at fixed offsets
        ft = (msk & M_TOP) ? 1.0f : 0.0f;
        fb = (msk & M_BOT) ? 1.0f : 0.0f;
        fr = (msk & M_RIGHT) ? 1.0f : 0.0f;
        fl = (msk & M_LEFT) ? 1.0f : 0.0f;
        pt = qt;
        pb = qb;
        pr = qr;
        pl = ql;
        if (D == 3) {
            ff = (msk & M_FRONT) ? 1.0f : 0.0f;
            fk = (msk & M_BACK) ? 1.0f : 0.0f;
            pf = qf;
            pk = qk;
        }
    } else {
        // smooth.comp solid_data(): OOB texel (0,0,0,0) -> f = 1, p = 0;
        // OOB solid (f = 0) in the correction solve of the refined mode
        const float oob = P.corr ? 1.0f : 0.0f;
        auto A = [&](int ii, int jj, int kk) -> float {
            return a_sel(L, inside(ii, jj, kk, sx, sy, sz), ii, jj, kk, oob);
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
    float psum = pt * ft + pb * fb;
    psum = psum + pr * fr;
    psum = psum + pl * fl;
    if (D == 3) {
        psum = psum + pf * ff;
        psum = psum + pk * fk;
    }
    // level 0: rhs holds rho*h*dv/dt; coarser (and level 0 of the
    // correction solve): r * h_l^2 (smooth.comp:128-132)
    const float term = level == 0 && !P.corr ? dv : dv * (hl * hl);
    L.p[own] = (psum - term) / fsum;
}
template <int D>
__device__ __forceinline__ void sweep_body(int bx_, int by_, int bz_, const SimParams &P, const Level &L, int level, int parity, const float *rhs, float hl) {
    sweep_at<D>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, L, level, parity, rhs, hl);
}
template <int D>
__global__ void k_sweep(XG X_, SimParams P, Level L, int level, int parity, const float *rhs, float hl) {
    XG_BLOCK(X_)
    sweep_body<D>(bx_, by_, bz_, P, L, level, parity, rhs, hl);
}

// Four cells of one colour per thread (m = 4t .. 4t+3 of a row), with 16-byte
// loads of the right-hand side and of the other colour's rows; the left and
// right neighbours come from the same-row vector plus one scalar.  Cells
// flagged M_GENERAL (every fluid cell of a boundary row among them) go
// through the scalar code, each on its own.  Requires hx % 4 == 0.
template <int D>
__device__ __forceinline__ float sweep_general(const SimParams &P, const Level &L, int i, int j,
                                               int k, float term, bool &skip) {
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    const float oob = P.corr ? 1.0f : 0.0f;
    auto A = [&](int ii, int jj, int kk) -> float {
        return a_sel(L, inside(ii, jj, kk, sx, sy, sz), ii, jj, kk, oob);
    };
    float ft = 1.0f - A(i, j + 1, k), fb = 1.0f - A(i, j - 1, k);
    float fr = 1.0f - A(i + 1, j, k), fl = 1.0f - A(i - 1, j, k);
    float ff = 0.f, fk = 0.f, fsum;
    if (D == 3) {
        ff = 1.0f - A(i, j, k + 1);
        fk = 1.0f - A(i, j, k - 1);
        fsum = ft + fb + fr + fl + ff + fk;
    } else {
        fsum = ft + fb + fr + fl;
    }
    skip = fsum == 0.0f;
    float psum = p_at(L, i, j + 1, k) * ft + p_at(L, i, j - 1, k) * fb;
    psum = psum + p_at(L, i + 1, j, k) * fr;
    psum = psum + p_at(L, i - 1, j, k) * fl;
    if (D == 3) {
        psum = psum + p_at(L, i, j, k + 1) * ff;
        psum = psum + p_at(L, i, j, k - 1) * fk;
    }
    return (psum - term) / fsum;
}

template <int D>
__device__ __forceinline__ void sweep4_at(int t, int j, int k, const SimParams &P, const Level &L, int level,
                                          int parity, const float *rhs, float hl) {
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2], hx = L.hx;
    const int m0 = 4 * t;
    if (m0 >= hx || j >= sy || k >= sz) return;
    const int o = (parity + j + k) & 1;
    const uint32_t half = split_half_d(L);
    const uint32_t row = ((uint32_t)k * sy + j) * hx;
    const uint32_t own = (uint32_t)parity * half + row + m0;
    // Every load of the thread issues before anything waits on one (one
    // memory round trip): no early exit on the mask, and the neighbour rows
    // load unconditionally, edge rows reading their own row instead (in
    // bounds, and unused: their fluid cells are all M_GENERAL).
    const bool edge_row = j == 0 || j == sy - 1 || (D == 3 && (k == 0 || k == sz - 1));
    const uint32_t mask4 = *reinterpret_cast<const uint32_t *>(L.mask + own);
    const float4 rv = *reinterpret_cast<const float4 *>(rhs + own);
    const float *po = L.p + (uint32_t)(1 - parity) * half + row;
    const int dy = edge_row ? 0 : hx, dz = edge_row || D == 2 ? 0 : sy * hx;
    const float4 C = *reinterpret_cast<const float4 *>(po + m0);
    const float4 T = *reinterpret_cast<const float4 *>(po + m0 + dy);
    const float4 B = *reinterpret_cast<const float4 *>(po + m0 - dy);
    float4 F = make_float4(0.f, 0.f, 0.f, 0.f), K = F;
    if (D == 3) {
        F = *reinterpret_cast<const float4 *>(po + m0 + dz);
        K = *reinterpret_cast<const float4 *>(po + m0 - dz);
    }
    // left/right neighbour beyond the vector: index m0-1 (o = 0) or m0+4
    // (o = 1), clamped into the row and only used when inside it
    const int xi = min(max(m0 - 1 + 5 * o, 0), hx - 1);
    const float extra = po[xi];
    float term[4] = {rv.x, rv.y, rv.z, rv.w};
    if (level != 0 || P.corr)
        for (int q = 0; q < 4; ++q) term[q] = term[q] * (hl * hl);
    float np[4];
    bool wr[4];
    // One straight-line path: a cell flagged M_GENERAL (every fluid cell of
    // an edge row, the first and last cell of a row, fractional solid
    // nearby) takes the general update alone; solid cells and the padding
    // past sx are not written.
    const float c[4] = {C.x, C.y, C.z, C.w}, tt[4] = {T.x, T.y, T.z, T.w},
                bb[4] = {B.x, B.y, B.z, B.w}, ff4[4] = {F.x, F.y, F.z, F.w},
                kk4[4] = {K.x, K.y, K.z, K.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t mk = (mask4 >> (8 * q)) & 0xffu;
        const int i = 2 * (m0 + q) + o;
        const float pr = o ? (q < 3 ? c[q + 1] : extra) : c[q];
        const float pl = o ? c[q] : (q > 0 ? c[q - 1] : extra);
        const float ft = (mk & M_TOP) ? 1.0f : 0.0f, fb = (mk & M_BOT) ? 1.0f : 0.0f;
        const float fr = (mk & M_RIGHT) ? 1.0f : 0.0f, fl = (mk & M_LEFT) ? 1.0f : 0.0f;
        float fsum;
        float psum = tt[q] * ft + bb[q] * fb;
        psum = psum + pr * fr;
        psum = psum + pl * fl;
        if (D == 3) {
            const float f5 = (mk & M_FRONT) ? 1.0f : 0.0f, f6 = (mk & M_BACK) ? 1.0f : 0.0f;
            psum = psum + ff4[q] * f5;
            psum = psum + kk4[q] * f6;
            fsum = ft + fb + fr + fl + f5 + f6;
        } else {
            fsum = ft + fb + fr + fl;
        }
        wr[q] = !(mk & M_SOLID) && i < sx && fsum != 0.0f;
        np[q] = (psum - term[q]) / fsum;
        if ((mk & M_GENERAL) && i < sx) {
            bool skip;
            np[q] = sweep_general<D>(P, L, i, j, k, term[q], skip);
            wr[q] = !skip;
        }
    }
    float *dst = L.p + own;
    if (wr[0] && wr[1] && wr[2] && wr[3]) {
        *reinterpret_cast<float4 *>(dst) = make_float4(np[0], np[1], np[2], np[3]);
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (wr[q]) dst[q] = np[q];
    }
}
template <int D>
__device__ __forceinline__ void sweep4_body(int bx_, int by_, int bz_, const SimParams &P, const Level &L, int level, int parity, const float *rhs, float hl) {
    sweep4_at<D>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, L, level, parity, rhs,
                 hl);
}
template <int D>
__global__ void k_sweep4(XG X_, SimParams P, Level L, int level, int parity, const float *rhs, float hl) {
    XG_BLOCK(X_)
    sweep4_body<D>(bx_, by_, bz_, P, L, level, parity, rhs, hl);
}

// r = b - A p at one cell (compute_residual.comp 3D:77-130 / 2D:72-117):
// solid cells 0; OOB neighbours solid (f = 0).  On level 0 b comes from the
// face divergence: interior cells reuse dv0 (same value under both OOB
// conventions), boundary cells recompute it with this kernel's convention.
template <int D, int VF>
__device__ float residual_at(const SimParams &P, const Level &L, int level, const float *rhs,
                             const VelPtrs &V, const uint16_t *s0, int i, int j, int k, float hl) {
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    const uint32_t own = sidx(L, i, j, k);
    const uint32_t msk = L.mask[own];
    if (msk & M_SOLID) return 0.0f;
    float psum, fsum;
    if (!(msk & M_GENERAL)) {
        // interior cell: the six neighbours are the other colour at fixed
        // offsets from its row (see k_sweep)
        const int hx = L.hx, m = i >> 1, o = i & 1;
        const float *po = L.p + (uint32_t)(1 - ((i + j + k) & 1)) * split_half_d(L) +
                          ((uint32_t)k * sy + j) * hx;
        const float ft = (msk & M_TOP) ? 1.0f : 0.0f, fb = (msk & M_BOT) ? 1.0f : 0.0f;
        const float fr = (msk & M_RIGHT) ? 1.0f : 0.0f, fl = (msk & M_LEFT) ? 1.0f : 0.0f;
        psum = po[m + hx] * ft + po[m - hx] * fb;
        psum = psum + po[m + o] * fr;
        psum = psum + po[m - 1 + o] * fl;
        if (D == 3) {
            const int plane = sy * hx;
            const float ff = (msk & M_FRONT) ? 1.0f : 0.0f, fk = (msk & M_BACK) ? 1.0f : 0.0f;
            psum = psum + po[m + plane] * ff;
            psum = psum + po[m - plane] * fk;
            fsum = ft + fb + fr + fl + ff + fk;
        } else {
            fsum = ft + fb + fr + fl;
        }
    } else {
        auto A = [&](int ii, int jj, int kk) -> float {
            return a_sel(L, inside(ii, jj, kk, sx, sy, sz), ii, jj, kk, 1.0f);
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
    if (level == 0 && !P.corr) {
        const bool edge = i == 0 || j == 0 || i == sx - 1 || j == sy - 1 ||
                          (D == 3 && (k == 0 || k == sz - 1));
        const float dv = edge ? level0_dv<D, VF>(P, V, s0, i, j, k, 1.0f) : rhs[own];
        old = P.rho * dv / (P.dt * hl);
    } else {
        old = rhs[own];
    }
    return old - (psum - fsum * L.p[own]) / (hl * hl);
}

// residual_at in two halves, so that a thread with several cells issues
// all their loads before the first use: residual_load reads the mask, the
// right-hand side, the own pressure and the six interior-stencil neighbours
// (index 0 in place of a neighbour when the cell is not strictly inside, or
// of the cell itself when `in` is false); residual_finish forms residual_at's
// value for a solid or interior cell and calls residual_at for an M_GENERAL
// one.  A fluid cell that is not M_GENERAL is never on the domain boundary,
// so its level-0 right-hand side is dv0 (rhs) as in residual_at.
struct ResNb {
    uint32_t msk;
    float rv, own, q[6];
};
template <int D>
__device__ __forceinline__ void residual_load(const Level &L, const float *rhs, int i, int j, int k, bool in,
                                              ResNb &R) {
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2], hx = L.hx, m = i >> 1, o = i & 1;
    const uint32_t own = in ? sidx(L, i, j, k) : 0u;
    R.msk = L.mask[own];
    R.rv = rhs[own];
    R.own = L.p[own];
    const bool gi = in && i > 0 && i < sx - 1 && j > 0 && j < sy - 1 && (D == 2 || (k > 0 && k < sz - 1));
    const uint32_t ob = (uint32_t)(1 - ((i + j + k) & 1)) * split_half_d(L) + ((uint32_t)k * sy + j) * hx + m;
    const uint32_t plane = (uint32_t)sy * hx;
    R.q[0] = L.p[gi ? ob + hx : 0u];
    R.q[1] = L.p[gi ? ob - hx : 0u];
    R.q[2] = L.p[gi ? ob + o : 0u];
    R.q[3] = L.p[gi ? ob - 1 + o : 0u];
    if (D == 3) {
        R.q[4] = L.p[gi ? ob + plane : 0u];
        R.q[5] = L.p[gi ? ob - plane : 0u];
    }
}
template <int D, int VF>
__device__ __forceinline__ float residual_finish(const SimParams &P, const Level &L, int level, const float *rhs,
                                                 const VelPtrs &V, const uint16_t *s0, int i, int j, int k,
                                                 float hl, const ResNb &R) {
    const uint32_t msk = R.msk;
    if (msk & M_SOLID) return 0.0f;
    if (msk & M_GENERAL) return residual_at<D, VF>(P, L, level, rhs, V, s0, i, j, k, hl);
    const float ft = (msk & M_TOP) ? 1.0f : 0.0f, fb = (msk & M_BOT) ? 1.0f : 0.0f;
    const float fr = (msk & M_RIGHT) ? 1.0f : 0.0f, fl = (msk & M_LEFT) ? 1.0f : 0.0f;
    float psum = R.q[0] * ft + R.q[1] * fb;
    psum = psum + R.q[2] * fr;
    psum = psum + R.q[3] * fl;
    float fsum;
    if (D == 3) {
        const float ff = (msk & M_FRONT) ? 1.0f : 0.0f, fk = (msk & M_BACK) ? 1.0f : 0.0f;
        psum = psum + R.q[4] * ff;
        psum = psum + R.q[5] * fk;
        fsum = ft + fb + fr + fl + ff + fk;
    } else {
        fsum = ft + fb + fr + fl;
    }
    const float old = level == 0 && !P.corr ? P.rho * R.rv / (P.dt * hl) : R.rv;
    return old - (psum - fsum * R.own) / (hl * hl);
}

// Residual of level `level` for the 2x2(x2) block of fine cells under one
// coarse cell, stored in place (the reference's residual image), then the
// restriction of restrict_residual.comp (3D sum order: the four children at
// y = 2j, then the four at y = 2j+1) and the clear of the coarse pressure.
// Threads past the coarse size cover trailing fine cells of odd sizes.
template <int D, int VF>
__device__ __forceinline__ void residual_restrict_at(int ci, int cj, int ck, const SimParams &P, const Level &F, const Level &C, int level, const float *rhs, const VelPtrs &V, const uint16_t *s0, float hl, int has_coarse) {
    const int I = 2 * ci, J = 2 * cj, K = D == 3 ? 2 * ck : 0;
    if (I >= F.n[0] || J >= F.n[1] || K >= F.n[2]) return;
    // all loads first, then the stores: the residual is written in place
    // over the right-hand side it reads on coarse levels
    float r[2][2][2] = {{{0.f, 0.f}, {0.f, 0.f}}, {{0.f, 0.f}, {0.f, 0.f}}};  // [dk][dj][di]
    // per plane of children (four cells), every load in flight before the
    // first use
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk) {
        ResNb R[2][2];
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int i = I + di, j = J + dj, k = K + dk;
                residual_load<D>(F, rhs, i, j, k, inside(i, j, k, F.n[0], F.n[1], F.n[2]), R[dj][di]);
            }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int i = I + di, j = J + dj, k = K + dk;
                if (!inside(i, j, k, F.n[0], F.n[1], F.n[2])) continue;
                r[dk][dj][di] = residual_finish<D, VF>(P, F, level, rhs, V, s0, i, j, k, hl, R[dj][di]);
            }
    }
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int i = I + di, j = J + dj, k = K + dk;
                if (inside(i, j, k, F.n[0], F.n[1], F.n[2])) F.r[sidx(F, i, j, k)] = r[dk][dj][di];
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
    const uint32_t c = sidx(C, ci, cj, ck);
    C.r[c] = v;
    C.p[c] = 0.0f;
}
template <int D, int VF>
__device__ __forceinline__ void residual_restrict_body(int bx_, int by_, int bz_, const SimParams &P, const Level &F, const Level &C, int level, const float *rhs, const VelPtrs &V, const uint16_t *s0, float hl, int has_coarse) {
    residual_restrict_at<D, VF>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, P, F, C, level,
                                rhs, V, s0, hl, has_coarse);
}
template <int D, int VF>
__global__ void k_residual_restrict(XG X_, SimParams P, Level F, Level C, int level, const float *rhs, VelPtrs V, const uint16_t *s0, float hl, int has_coarse) {
    XG_BLOCK(X_)
    residual_restrict_body<D, VF>(bx_, by_, bz_, P, F, C, level, rhs, V, s0, hl, has_coarse);
}

// restrict_solid.comp 3D:53-78 / 2D:53-73 (natural layout, 4 x fp16).
template <int D>
__global__ void k_restrict_solid(XG X_, Level F, Level C) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
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

// The six neighbours of cell (i,j,k) for a general residual: solid fractions
// and pressures read at coordinates clamped into the level (always in
// bounds, 32-bit byte offsets), which of them are inside returned as bits.
// Several cells' loads then issue before the first use, with only the values
// and one word per cell live until then.
template <int D>
__device__ __forceinline__ uint32_t nb6_load(const Level &L, int i, int j, int k, uint16_t a[6], float pn[6]) {
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    const int off[6][3] = {{0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
    uint32_t in = 0;
#pragma unroll
    for (int e = 0; e < 2 * D; ++e) {
        const int ii = i + off[e][0], jj = j + off[e][1], kk = k + off[e][2];
        in |= (inside(ii, jj, kk, sx, sy, sz) ? 1u : 0u) << e;
        const int ci = min(max(ii, 0), sx - 1), cj = min(max(jj, 0), sy - 1),
                  ck = D == 3 ? min(max(kk, 0), sz - 1) : 0;
        a[e] = *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(L.s) +
                                                   (lin(ci, cj, ck, sx, sy) * 8u + 6u));
        pn[e] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(L.p) + sidx(L, ci, cj, ck) * 4u);
    }
    return in;
}

// Residual + restriction for levels with even sizes (fine = 2 x coarse on
// every axis): the thread of coarse cell (I,J,K) owns fine cells
// (2I+di, 2J+dj, 2K+dk), all at index I of their colour rows, so every
// neighbour is the other colour at index I-1+di (left), I+di (right) or I
// (rows 2J+dj+-1, planes 2K+dk+-1).  Blocks touching the domain boundary and
// cells flagged M_GENERAL use residual_at().
template <int D, int VF, bool GB = false>
__device__ __forceinline__ void residual_restrict_even_body(int bx_, int by_, int bz_, const SimParams &P, const Level &F, const Level &C, int level, const float *rhs, const VelPtrs &V, const uint16_t *s0, float hl, int wfine) {
    const int I = bx_ * blockDim.x + threadIdx.x;
    const int J = by_ * blockDim.y + threadIdx.y;
    const int K = bz_;
    const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
    if (I >= cx || J >= cy || K >= cz) return;
    const int fy = F.n[1], fhx = F.hx;
    const uint32_t half = split_half_d(F);
    float r[2][2][2];
    uint32_t msk[2][2][2];
    bool general = false;
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int j = 2 * J + dj, k = D == 3 ? 2 * K + dk : 0;
                const int c = (di + dj + dk) & 1;
                msk[dk][dj][di] = F.mask[(uint32_t)c * half + ((uint32_t)k * fy + j) * fhx + I];
                general |= (msk[dk][dj][di] & M_GENERAL) != 0;
            }
    if (!general) {
        // straight-line code: all loads of the eight cells issue together
#pragma unroll
        for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int j = 2 * J + dj, k = D == 3 ? 2 * K + dk : 0;
                    const int c = (di + dj + dk) & 1;
                    const uint32_t row = ((uint32_t)k * fy + j) * fhx;
                    const uint32_t own = (uint32_t)c * half + row + I;
                    const uint32_t m = msk[dk][dj][di];
                    // non-solid cells here are interior (boundary fluid cells
                    // are M_GENERAL); the clamps only keep the discarded
                    // neighbour loads of solid boundary cells in bounds
                    const float *po = F.p + (uint32_t)(1 - c) * half;
                    const uint32_t rup = ((uint32_t)k * fy + min(j + 1, fy - 1)) * fhx;
                    const uint32_t rdn = ((uint32_t)k * fy + max(j - 1, 0)) * fhx;
                    const int il = max(I - 1 + di, 0), ir = min(I + di, fhx - 1);
                    const float ft = (m & M_TOP) ? 1.0f : 0.0f, fb = (m & M_BOT) ? 1.0f : 0.0f;
                    const float fr = (m & M_RIGHT) ? 1.0f : 0.0f, fl = (m & M_LEFT) ? 1.0f : 0.0f;
                    float psum = po[rup + I] * ft + po[rdn + I] * fb;
                    psum = psum + po[row + ir] * fr;
                    psum = psum + po[row + il] * fl;
                    float fsum;
                    if (D == 3) {
                        const int fz = F.n[2];
                        const uint32_t rfr = ((uint32_t)min(k + 1, fz - 1) * fy + j) * fhx;
                        const uint32_t rbk = ((uint32_t)max(k - 1, 0) * fy + j) * fhx;
                        const float ff = (m & M_FRONT) ? 1.0f : 0.0f, fk = (m & M_BACK) ? 1.0f : 0.0f;
                        psum = psum + po[rfr + I] * ff;
                        psum = psum + po[rbk + I] * fk;
                        fsum = ft + fb + fr + fl + ff + fk;
                    } else {
                        fsum = ft + fb + fr + fl;
                    }
                    const float b = level == 0 && !P.corr ? P.rho * rhs[own] / (P.dt * hl) : rhs[own];
                    const float v = b - (psum - fsum * F.p[own]) / (hl * hl);
                    r[dk][dj][di] = (m & M_SOLID) ? 0.0f : v;
                }
    } else if constexpr (GB) {
        // (GB: only for levels >= 1 and the correction solve, whose
        // right-hand side is the stored one) residual_at for each cell, its
        // loads issued for a row pair of cells at a time (neighbour
        // fractions and pressures, right-hand side, own pressure) before
        // either is used: one memory round trip per pair instead of two or
        // more per cell
#pragma unroll
        for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                uint16_t a[2][6];
                float pn[2][6], rv[2], pw[2];
                uint32_t in[2];
                const int j = 2 * J + dj, k = D == 3 ? 2 * K + dk : 0;
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int i = 2 * I + di;
                    const uint32_t own = sidx(F, i, j, k);
                    in[di] = nb6_load<D>(F, i, j, k, a[di], pn[di]);
                    rv[di] = rhs[own];
                    pw[di] = F.p[own];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const uint32_t m = msk[dk][dj][di];
                    float f[6], q[6];
#pragma unroll
                    for (int e = 0; e < 2 * D; ++e) {
                        const bool ok = (in[di] >> e) & 1u;
                        // compute_residual.comp: OOB neighbours solid (f = 0)
                        f[e] = (m & M_GENERAL) ? 1.0f - (ok ? f16_to_f32(a[di][e]) : 1.0f)
                                               : ((m & (M_TOP << e)) ? 1.0f : 0.0f);
                        q[e] = ok ? pn[di][e] : 0.0f;
                    }
                    float psum = q[0] * f[0] + q[1] * f[1];
                    psum = psum + q[2] * f[2];
                    psum = psum + q[3] * f[3];
                    float fsum;
                    if (D == 3) {
                        psum = psum + q[4] * f[4];
                        psum = psum + q[5] * f[5];
                        fsum = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
                    } else {
                        fsum = f[0] + f[1] + f[2] + f[3];
                    }
                    const float v = rv[di] - (psum - fsum * pw[di]) / (hl * hl);
                    r[dk][dj][di] = (m & M_SOLID) ? 0.0f : v;
                }
            }
    } else {
#pragma unroll
        for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int i = 2 * I + di, j = 2 * J + dj, k = D == 3 ? 2 * K + dk : 0;
                    r[dk][dj][di] = residual_at<D, VF>(P, F, level, rhs, V, s0, i, j, k, hl);
                }
    }
    // the fine residual image: coarse levels smooth against it; on level 0
    // it is only the observable residual texture (wfine = 0 skips it)
    if (wfine) {
#pragma unroll
        for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int j = 2 * J + dj, k = D == 3 ? 2 * K + dk : 0;
                    const int c = (di + dj + dk) & 1;
                    F.r[(uint32_t)c * half + ((uint32_t)k * fy + j) * fhx + I] = r[dk][dj][di];
                }
    }
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
    const uint32_t cc = sidx(C, I, J, K);
    C.r[cc] = v;
    C.p[cc] = 0.0f;
}
template <int D, int VF>
__global__ void k_residual_restrict_even(XG X_, SimParams P, Level F, Level C, int level, const float *rhs, VelPtrs V, const uint16_t *s0, float hl, int wfine) {
    XG_BLOCK(X_)
    residual_restrict_even_body<D, VF>(bx_, by_, bz_, P, F, C, level, rhs, V, s0, hl, wfine);
}
// The same with the general cells' loads batched, for levels >= 1 (and the
// correction solve); 256-thread blocks (every launch shape here), so the
// batch has the registers it needs (about 170) without spilling.
template <int D, int VF>
__global__ void __launch_bounds__(256) k_residual_restrict_even_gb(XG X_, SimParams P, Level F, Level C, int level, const float *rhs, VelPtrs V, const uint16_t *s0, float hl, int wfine) {
    XG_BLOCK(X_)
    residual_restrict_even_body<D, VF, true>(bx_, by_, bz_, P, F, C, level, rhs, V, s0, hl, wfine);
}

// k_residual_restrict_even for 3D FP32 levels with nx % 4 == 0, two coarse
// cells per thread: thread (I2, J, K) owns coarse cells 2*I2 and 2*I2+1,
// i.e. fine x = 4*I2 .. 4*I2+3 of rows 2J+dj, planes 2K+dk.  In each fine
// row each colour holds two of these cells, at indices 2*I2 and 2*I2+1, so
// the pressure and right-hand side of the own rows and of the eight
// neighbour rows outside the 4x2x2 block come in 8-byte loads and the masks
// in 2-byte loads; the x neighbours beyond the four cells are one load per
// row.  Every other neighbour is an own cell.  Per cell and per coarse cell
// the arithmetic of k_residual_restrict_even; its two divisions by launch
// constants, (dt*hl) for the level-0 right-hand side and hl^2, use div_rn()
// with the reciprocals in RD.
static bool div_range(float y) { return std::fabs(y) >= 0x1p-30f && std::fabs(y) <= 0x1p30f; }
ResDiv res_div(const SimParams &P, float hl) {
    ResDiv R;
    R.d_b = P.dt * hl;
    R.d_h2 = hl * hl;
    R.r_b = (float)(1.0 / (double)R.d_b);
    R.r_h2 = (float)(1.0 / (double)R.d_h2);
    // the host checked dt*h and h*h once per physics change (P.fdiv_res);
    // hl = h*2^l makes these the same significands
    R.ok = P.fdiv_res && div_range(R.d_b) && div_range(R.d_h2) ? 1 : 0;
    return R;
}

template <int VF>
__device__ __forceinline__ void residual_restrict_even_x2_body(int bx_, int by_, int bz_, const SimParams &P, const Level &F, const Level &C, int level, const float *rhs, const VelPtrs &V, const uint16_t *s0, float hl, int wfine, const ResDiv &RD) {
    const int I2 = bx_ * blockDim.x + threadIdx.x;
    const int J = by_ * blockDim.y + threadIdx.y;
    const int K = bz_;
    const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
    if (2 * I2 >= cx || J >= cy || K >= cz) return;
    const int fy = F.n[1], fz = F.n[2], fhx = F.hx;
    const uint32_t half = split_half_d(F);
    const int m0 = 2 * I2;
    auto rowof = [&](int j, int k) -> uint32_t { return ((uint32_t)k * fy + j) * fhx; };
    uint32_t mk[2][2][2];
    bool general = false;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const uint32_t at = (uint32_t)c * half + rowof(2 * J + dj, 2 * K + dk) + m0;
                mk[dk][dj][c] = *reinterpret_cast<const uint16_t *>(F.mask + at);
                general |= (mk[dk][dj][c] & 0x8080u) != 0;
            }
    float r[2][2][4];  // [dk][dj][q], fine x = 4*I2 + q
    if (!general) {
        auto ld2 = [&](int c, uint32_t row) -> float2 {
            return *reinterpret_cast<const float2 *>(F.p + (uint32_t)c * half + row + m0);
        };
        // non-solid cells here are interior (boundary fluid cells are
        // M_GENERAL); the clamps keep the discarded loads of solid boundary
        // cells in bounds
        const int jd = max(2 * J - 1, 0), ju = min(2 * J + 2, fy - 1);
        const int kb = max(2 * K - 1, 0), kf = min(2 * K + 2, fz - 1);
        float2 OP[2][2][2], RH[2][2][2], DN[2][2], UP[2][2], BK[2][2], FR[2][2];
        float XL[2][2], XR[2][2];
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                const uint32_t row = rowof(2 * J + dj, 2 * K + dk);
                const int s = (dj + dk) & 1;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    OP[dk][dj][c] = ld2(c, row);
                    RH[dk][dj][c] = *reinterpret_cast<const float2 *>(rhs + (uint32_t)c * half + row + m0);
                }
                // x = 4*I2 - 1 has the colour of q = 1, x = 4*I2 + 4 that of q = 2
                XL[dk][dj] = F.p[(uint32_t)((1 + s) & 1) * half + row + max(m0 - 1, 0)];
                XR[dk][dj] = F.p[(uint32_t)(s & 1) * half + row + min(m0 + 2, fhx - 1)];
            }
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                DN[d][c] = ld2(c, rowof(jd, 2 * K + d));
                UP[d][c] = ld2(c, rowof(ju, 2 * K + d));
                BK[d][c] = ld2(c, rowof(2 * J + d, kb));
                FR[d][c] = ld2(c, rowof(2 * J + d, kf));
            }
        auto el = [](float2 v, int e) { return e ? v.y : v.x; };
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int s = (dj + dk) & 1, c = (q + s) & 1, oc = 1 - c, e = q >> 1;
                    const uint32_t m = (mk[dk][dj][c] >> (8 * e)) & 0xffu;
                    const float pu = dj == 0 ? el(OP[dk][1][oc], e) : el(UP[dk][oc], e);
                    const float pd = dj == 1 ? el(OP[dk][0][oc], e) : el(DN[dk][oc], e);
                    const float pf = dk == 0 ? el(OP[1][dj][oc], e) : el(FR[dj][oc], e);
                    const float pk = dk == 1 ? el(OP[0][dj][oc], e) : el(BK[dj][oc], e);
                    const float pr = q < 3 ? el(OP[dk][dj][oc], (q + 1) >> 1) : XR[dk][dj];
                    const float pl = q > 0 ? el(OP[dk][dj][oc], (q - 1) >> 1) : XL[dk][dj];
                    const float ft = (m & M_TOP) ? 1.0f : 0.0f, fb = (m & M_BOT) ? 1.0f : 0.0f;
                    const float fr = (m & M_RIGHT) ? 1.0f : 0.0f, fl = (m & M_LEFT) ? 1.0f : 0.0f;
                    const float ff = (m & M_FRONT) ? 1.0f : 0.0f, fk = (m & M_BACK) ? 1.0f : 0.0f;
                    float psum = pu * ft + pd * fb;
                    psum = psum + pr * fr;
                    psum = psum + pl * fl;
                    psum = psum + pf * ff;
                    psum = psum + pk * fk;
                    const float fsum = ft + fb + fr + fl + ff + fk;
                    const float rv = el(RH[dk][dj][c], e);
                    const float b = level == 0 && !P.corr ? div_rn(P.rho * rv, RD.d_b, RD.r_b, RD.ok) : rv;
                    const float v = b - div_rn(psum - fsum * el(OP[dk][dj][c], e), RD.d_h2, RD.r_h2, RD.ok);
                    r[dk][dj][q] = (m & M_SOLID) ? 0.0f : v;
                }
    } else {
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    r[dk][dj][q] = residual_at<3, VF>(P, F, level, rhs, V, s0, 4 * I2 + q,
                                                          2 * J + dj, 2 * K + dk, hl);
    }
    if (wfine) {
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const int q0 = (c + dj + dk) & 1;  // the colour's cells: q0, q0 + 2
                    *reinterpret_cast<float2 *>(F.r + (uint32_t)c * half + rowof(2 * J + dj, 2 * K + dk) + m0) =
                        make_float2(r[dk][dj][q0], r[dk][dj][q0 + 2]);
                }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        // coarse cell 2*I2 + h: children q = 2h + di
        float bsum = r[0][0][2 * h] + r[1][0][2 * h];
        bsum = bsum + r[0][0][2 * h + 1];
        bsum = bsum + r[1][0][2 * h + 1];
        float tsum = r[0][1][2 * h] + r[1][1][2 * h];
        tsum = tsum + r[0][1][2 * h + 1];
        tsum = tsum + r[1][1][2 * h + 1];
        const float v = (bsum + tsum) / 8.0f;
        const uint32_t cc = sidx(C, 2 * I2 + h, J, K);
        C.r[cc] = v;
        C.p[cc] = 0.0f;
    }
}
template <int VF>
__global__ void k_residual_restrict_even_x2(XG X_, SimParams P, Level F, Level C, int level, const float *rhs, VelPtrs V, const uint16_t *s0, float hl, int wfine, ResDiv RD) {
    XG_BLOCK(X_)
    residual_restrict_even_x2_body<VF>(bx_, by_, bz_, P, F, C, level, rhs, V, s0, hl, wfine, RD);
}

// k_residual_restrict_even in 3D with two lanes per coarse cell (small,
// latency-bound levels): lane pair
// (2I, 2I+1) of a row splits the eight fine cells by dj; each lane computes
// the residual of its four cells and its half of the restriction sum
// (bsum: dj = 0, tsum: dj = 1, in restrict_residual.comp's order), and the
// dj = 0 lane adds the partner's half (one lane shuffle).
template <int VF, bool GB>
__device__ __forceinline__ void residual_restrict_even_dj_body(int bx_, int by_, int bz_, const SimParams &P,
                                                               const Level &F, const Level &C, int level,
                                                               const float *rhs, const VelPtrs &V,
                                                               const uint16_t *s0, float hl, int wfine) {
    const int gx = bx_ * blockDim.x + threadIdx.x;
    const int I = gx >> 1, dj = gx & 1;
    const int J = by_ * blockDim.y + threadIdx.y;
    const int K = bz_;
    const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
    // both lanes of a pair exist (blockDim.x even); out-of-range pairs still
    // join the shuffle below
    const bool act = I < cx && J < cy && K < cz;
    const int fy = F.n[1], fhx = F.hx, fz = F.n[2];
    const uint32_t half = split_half_d(F);
    const int j = 2 * J + dj;
    float r[2][2] = {{0.f, 0.f}, {0.f, 0.f}};  // [dk][di]
    if (act) {
        uint32_t msk[2][2];
        bool general = false;
#pragma unroll
        for (int dk = 0; dk < 2; ++dk)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int k = 2 * K + dk;
                const int c = (di + dj + dk) & 1;
                msk[dk][di] = F.mask[(uint32_t)c * half + ((uint32_t)k * fy + j) * fhx + I];
                general |= (msk[dk][di] & M_GENERAL) != 0;
            }
        if (!general) {
#pragma unroll
            for (int dk = 0; dk < 2; ++dk)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int k = 2 * K + dk;
                    const int c = (di + dj + dk) & 1;
                    const uint32_t row = ((uint32_t)k * fy + j) * fhx;
                    const uint32_t own = (uint32_t)c * half + row + I;
                    const uint32_t m = msk[dk][di];
                    const float *po = F.p + (uint32_t)(1 - c) * half;
                    const uint32_t rup = ((uint32_t)k * fy + min(j + 1, fy - 1)) * fhx;
                    const uint32_t rdn = ((uint32_t)k * fy + max(j - 1, 0)) * fhx;
                    const int il = max(I - 1 + di, 0), ir = min(I + di, fhx - 1);
                    const float ft = (m & M_TOP) ? 1.0f : 0.0f, fb = (m & M_BOT) ? 1.0f : 0.0f;
                    const float fr = (m & M_RIGHT) ? 1.0f : 0.0f, fl = (m & M_LEFT) ? 1.0f : 0.0f;
                    float psum = po[rup + I] * ft + po[rdn + I] * fb;
                    psum = psum + po[row + ir] * fr;
                    psum = psum + po[row + il] * fl;
                    const uint32_t rfr = ((uint32_t)min(k + 1, fz - 1) * fy + j) * fhx;
                    const uint32_t rbk = ((uint32_t)max(k - 1, 0) * fy + j) * fhx;
                    const float ff = (m & M_FRONT) ? 1.0f : 0.0f, fk = (m & M_BACK) ? 1.0f : 0.0f;
                    psum = psum + po[rfr + I] * ff;
                    psum = psum + po[rbk + I] * fk;
                    const float fsum = ft + fb + fr + fl + ff + fk;
                    const float b = level == 0 && !P.corr ? P.rho * rhs[own] / (P.dt * hl) : rhs[own];
                    const float v = b - (psum - fsum * F.p[own]) / (hl * hl);
                    r[dk][di] = (m & M_SOLID) ? 0.0f : v;
                }
        } else if constexpr (GB) {
            // (levels >= 1 and the correction solve) residual_at for the four
            // cells with all their loads issued before the first use
            uint16_t a[2][2][6];
            float pn[2][2][6], rv[2][2], pw[2][2];
            uint32_t in[2][2];
#pragma unroll
            for (int dk = 0; dk < 2; ++dk)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int i = 2 * I + di, k = 2 * K + dk;
                    const uint32_t own = sidx(F, i, j, k);
                    in[dk][di] = nb6_load<3>(F, i, j, k, a[dk][di], pn[dk][di]);
                    rv[dk][di] = rhs[own];
                    pw[dk][di] = F.p[own];
                }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int dk = 0; dk < 2; ++dk)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const uint32_t m = msk[dk][di];
                    float f[6], q[6];
#pragma unroll
                    for (int e = 0; e < 6; ++e) {
                        const bool ok = (in[dk][di] >> e) & 1u;
                        // compute_residual.comp: OOB neighbours solid (f = 0)
                        f[e] = (m & M_GENERAL) ? 1.0f - (ok ? f16_to_f32(a[dk][di][e]) : 1.0f)
                                               : ((m & (M_TOP << e)) ? 1.0f : 0.0f);
                        q[e] = ok ? pn[dk][di][e] : 0.0f;
                    }
                    float psum = q[0] * f[0] + q[1] * f[1];
                    psum = psum + q[2] * f[2];
                    psum = psum + q[3] * f[3];
                    psum = psum + q[4] * f[4];
                    psum = psum + q[5] * f[5];
                    const float fsum = f[0] + f[1] + f[2] + f[3] + f[4] + f[5];
                    const float v = rv[dk][di] - (psum - fsum * pw[dk][di]) / (hl * hl);
                    r[dk][di] = (m & M_SOLID) ? 0.0f : v;
                }
        } else {
#pragma unroll
            for (int dk = 0; dk < 2; ++dk)
#pragma unroll
                for (int di = 0; di < 2; ++di)
                    r[dk][di] = residual_at<3, VF>(P, F, level, rhs, V, s0, 2 * I + di, j, 2 * K + dk, hl);
        }
        if (wfine) {
#pragma unroll
            for (int dk = 0; dk < 2; ++dk)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    const int k = 2 * K + dk;
                    const int c = (di + dj + dk) & 1;
                    F.r[(uint32_t)c * half + ((uint32_t)k * fy + j) * fhx + I] = r[dk][di];
                }
        }
    }
    // bsum (dj = 0) or tsum (dj = 1): r[0][0][0] + r[1][0][0] + r[0][0][1] + r[1][0][1]
    float part = r[0][0] + r[1][0];
    part = part + r[0][1];
    part = part + r[1][1];
    const float other = __shfl_xor(part, 1);
    if (!act || dj) return;
    const float v = (part + other) / 8.0f;
    const uint32_t cc = sidx(C, I, J, K);
    C.r[cc] = v;
    C.p[cc] = 0.0f;
}
template <int VF>
__global__ void k_residual_restrict_even_dj(XG X_, SimParams P, Level F, Level C, int level,
                                            const float *rhs, VelPtrs V, const uint16_t *s0,
                                            float hl, int wfine) {
    XG_BLOCK(X_)
    residual_restrict_even_dj_body<VF, false>(bx_, by_, bz_, P, F, C, level, rhs, V, s0, hl, wfine);
}
// levels >= 1 (and the correction solve): the general cells' loads batched
// (256-thread blocks, so the batch gets its registers)
template <int VF>
__global__ void __launch_bounds__(256) k_residual_restrict_even_dj_gb(XG X_, SimParams P, Level F, Level C,
                                                                     int level, const float *rhs, VelPtrs V,
                                                                     const uint16_t *s0, float hl, int wfine) {
    XG_BLOCK(X_)
    residual_restrict_even_dj_body<VF, true>(bx_, by_, bz_, P, F, C, level, rhs, V, s0, hl, wfine);
}

// Prolongation for levels with even sizes: the thread of coarse cell
// (I,J,K) loads the 3x3x3 coarse neighbourhood once (CLAMP_TO_EDGE) and
// updates its 8 fine cells, whose samples fall inside it (the host checks
// that the table entry of fine cell 2I+d has i0 = I-1+d on every axis).
template <int D>
__device__ __forceinline__ void prolong_even_body(int bx_, int by_, int bz_, const Level &F, const Level &C, const TC *tx, const TC *ty, const TC *tz) {
    const int I = bx_ * blockDim.x + threadIdx.x;
    const int J = by_ * blockDim.y + threadIdx.y;
    const int K = bz_;
    const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
    if (I >= cx || J >= cy || K >= cz) return;
    float q[3][3][3];  // [c][b][a] = p(clamp(I-1+a), clamp(J-1+b), clamp(K-1+c))
#pragma unroll
    for (int c = 0; c < (D == 3 ? 3 : 1); ++c)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ii = min(max(I - 1 + a, 0), cx - 1);
                const int jj = min(max(J - 1 + b, 0), cy - 1);
                const int kk = D == 3 ? min(max(K - 1 + c, 0), cz - 1) : 0;
                q[c][b][a] = C.p[sidx(C, ii, jj, kk)];
            }
    const int fy = F.n[1], fhx = F.hx;
    const uint32_t half = split_half_d(F);
    // the coordinate entries and the thread's own fine cells, all read
    // before the first store (each read after the previous store would be a
    // memory round trip of its own)
    TC ta[2], tb[2], tg[2] = {TC{0, 0.0f}, TC{0, 0.0f}};
    float old[2][2][2];  // [dk][dj][di]
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        ta[d] = tx[2 * I + d];
        tb[d] = ty[2 * J + d];
        if (D == 3) tg[d] = tz[2 * K + d];
    }
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int j = 2 * J + dj, k = D == 3 ? 2 * K + dk : 0;
                old[dk][dj][di] = F.p[(uint32_t)((di + dj + dk) & 1) * half + ((uint32_t)k * fy + j) * fhx + I];
            }
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int di = 0; di < 2; ++di) {
                const int j = 2 * J + dj, k = D == 3 ? 2 * K + dk : 0;
                // the host checked i0 = I-1+di (and likewise on y, z), so the
                // sample's texels are q[dk..dk+1][dj..dj+1][di..di+1]
                float v;
                if (D == 2) {
                    float x0 = lerpf(q[0][dj][di], q[0][dj][di + 1], ta[di].a);
                    float x1 = lerpf(q[0][dj + 1][di], q[0][dj + 1][di + 1], ta[di].a);
                    v = lerpf(x0, x1, tb[dj].a);
                } else {
                    float x00 = lerpf(q[dk][dj][di], q[dk][dj][di + 1], ta[di].a);
                    float x10 = lerpf(q[dk][dj + 1][di], q[dk][dj + 1][di + 1], ta[di].a);
                    float x01 = lerpf(q[dk + 1][dj][di], q[dk + 1][dj][di + 1], ta[di].a);
                    float x11 = lerpf(q[dk + 1][dj + 1][di], q[dk + 1][dj + 1][di + 1], ta[di].a);
                    float y0 = lerpf(x00, x10, tb[dj].a);
                    float y1 = lerpf(x01, x11, tb[dj].a);
                    v = lerpf(y0, y1, tg[dk].a);
                }
                const uint32_t cell =
                    (uint32_t)((di + dj + dk) & 1) * half + ((uint32_t)k * fy + j) * fhx + I;
                F.p[cell] = old[dk][dj][di] + v;
            }
}
template <int D>
__global__ void k_prolong_even(XG X_, Level F, Level C, const TC *tx, const TC *ty, const TC *tz) {
    XG_BLOCK(X_)
    prolong_even_body<D>(bx_, by_, bz_, F, C, tx, ty, tz);
}

// k_prolong_even in 3D for nx % 4 == 0 with two coarse cells per thread:
// thread (I2, J, K) adds the correction to fine x = 4*I2 .. 4*I2+3 of rows
// 2J+dj, planes 2K+dk.  The coarse window is x = 2*I2-1 .. 2*I2+2 (clamped),
// 36 texels instead of 2 x 27; fine x = 4*I2+q samples window texels
// a0 = (q >> 1) + (q & 1) and a0 + 1.  Each fine row is read and written as
// one 8-byte pair per colour.  Per fine cell the arithmetic of
// k_prolong_even.
__device__ __forceinline__ void prolong_even_x2_body(int bx_, int by_, int bz_, const Level &F, const Level &C, const TC *tx, const TC *ty, const TC *tz) {
    const int I2 = bx_ * blockDim.x + threadIdx.x;
    const int J = by_ * blockDim.y + threadIdx.y;
    const int K = bz_;
    const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
    if (2 * I2 >= cx || J >= cy || K >= cz) return;
    float q[3][3][4];  // [c][b][a] = p(clamp(2*I2-1+a), clamp(J-1+b), clamp(K-1+c))
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int ii = min(max(2 * I2 - 1 + a, 0), cx - 1);
                const int jj = min(max(J - 1 + b, 0), cy - 1);
                const int kk = min(max(K - 1 + c, 0), cz - 1);
                q[c][b][a] = C.p[sidx(C, ii, jj, kk)];
            }
    const int fy = F.n[1], fhx = F.hx;
    const uint32_t half = split_half_d(F);
    const int m0 = 2 * I2;
    TC ta[4], tb[2], tg[2];
#pragma unroll
    for (int u = 0; u < 4; ++u) ta[u] = tx[4 * I2 + u];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        tb[d] = ty[2 * J + d];
        tg[d] = tz[2 * K + d];
    }
    // the thread's own fine pairs, all read before the first store (each
    // read after the previous store would be a memory round trip of its own)
    float2 old[2][2][2];  // [dk][dj][colour]
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int c = 0; c < 2; ++c)
                old[dk][dj][c] = *reinterpret_cast<const float2 *>(
                    F.p + (uint32_t)c * half + ((uint32_t)(2 * K + dk) * fy + 2 * J + dj) * fhx + m0);
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const int j = 2 * J + dj, k = 2 * K + dk;
            const uint32_t row = ((uint32_t)k * fy + j) * fhx;
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int a0 = (u >> 1) + (u & 1);
                float x00 = lerpf(q[dk][dj][a0], q[dk][dj][a0 + 1], ta[u].a);
                float x10 = lerpf(q[dk][dj + 1][a0], q[dk][dj + 1][a0 + 1], ta[u].a);
                float x01 = lerpf(q[dk + 1][dj][a0], q[dk + 1][dj][a0 + 1], ta[u].a);
                float x11 = lerpf(q[dk + 1][dj + 1][a0], q[dk + 1][dj + 1][a0 + 1], ta[u].a);
                float y0 = lerpf(x00, x10, tb[dj].a);
                float y1 = lerpf(x01, x11, tb[dj].a);
                v[u] = lerpf(y0, y1, tg[dk].a);
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int u0 = (c + dj + dk) & 1;  // the colour's cells: u0, u0 + 2
                float2 nv = old[dk][dj][c];
                nv.x = nv.x + v[u0];
                nv.y = nv.y + v[u0 + 2];
                *reinterpret_cast<float2 *>(F.p + (uint32_t)c * half + row + m0) = nv;
            }
        }
}
__global__ void k_prolong_even_x2(XG X_, Level F, Level C, const TC *tx, const TC *ty, const TC *tz) {
    XG_BLOCK(X_)
    prolong_even_x2_body(bx_, by_, bz_, F, C, tx, ty, tz);
}

// Coordinates of the prolongation sample of fine cell idx along one axis:
// texcoord((idx + 0.5) / n_fine, n_coarse) (prolong_pressure.comp:45), built
// once per level with the kernel's own operation sequence.
__global__ void k_prolong_table(int nf, int nc, TC *e) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nf) return;
    float uv = ((float)idx + 0.5f) / (float)nf;
    e[idx] = texcoord_tc(uv, nc);
}

// prolong_pressure.comp 3D:39-51 / 2D:39-51: p_fine += LINEAR sample of the
// coarse pressure (CLAMP_TO_EDGE).  tx/ty/tz: per-axis coordinate tables.
template <int D>
__device__ __forceinline__ void prolong_at(int i, int j, int k, const Level &F, const Level &C, const TC *tx, const TC *ty, const TC *tz) {
    if (!inside(i, j, k, F.n[0], F.n[1], F.n[2])) return;
    const int cx = C.n[0], cy = C.n[1], cz = C.n[2];
    // the coordinate entries and the fine value load together (the coarse
    // texels depend on the entries: two memory round trips in all)
    const uint32_t cell = sidx(F, i, j, k);
    const TC a = tx[i], b = ty[j], g = D == 3 ? tz[k] : TC{0, 0.0f};
    const float pf = F.p[cell];
    __builtin_amdgcn_sched_barrier(0);
    const int i0 = min(max(a.i0, 0), cx - 1), i1 = min(max(a.i0 + 1, 0), cx - 1);
    const int j0 = min(max(b.i0, 0), cy - 1), j1 = min(max(b.i0 + 1, 0), cy - 1);
    const int k0 = min(max(g.i0, 0), cz - 1), k1 = min(max(g.i0 + 1, 0), cz - 1);
    auto F_ = [&](int ii, int jj, int kk) -> float { return C.p[sidx(C, ii, jj, kk)]; };
    float c;
    if (D == 2) {
        float x0 = lerpf(F_(i0, j0, 0), F_(i1, j0, 0), a.a);
        float x1 = lerpf(F_(i0, j1, 0), F_(i1, j1, 0), a.a);
        c = lerpf(x0, x1, b.a);
    } else {
        float x00 = lerpf(F_(i0, j0, k0), F_(i1, j0, k0), a.a);
        float x10 = lerpf(F_(i0, j1, k0), F_(i1, j1, k0), a.a);
        float x01 = lerpf(F_(i0, j0, k1), F_(i1, j0, k1), a.a);
        float x11 = lerpf(F_(i0, j1, k1), F_(i1, j1, k1), a.a);
        float y0 = lerpf(x00, x10, b.a);
        float y1 = lerpf(x01, x11, b.a);
        c = lerpf(y0, y1, g.a);
    }
    F.p[cell] = pf + c;
}
template <int D>
__device__ __forceinline__ void prolong_body(int bx_, int by_, int bz_, const Level &F, const Level &C, const TC *tx, const TC *ty, const TC *tz) {
    prolong_at<D>(bx_ * blockDim.x + threadIdx.x, by_ * blockDim.y + threadIdx.y, bz_, F, C, tx, ty, tz);
}
template <int D>
__global__ void k_prolong(XG X_, Level F, Level C, const TC *tx, const TC *ty, const TC *tz) {
    XG_BLOCK(X_)
    prolong_body<D>(bx_, by_, bz_, F, C, tx, ty, tz);
}

// Red-black iterations of a small level by one workgroup with pressure,
// solid fraction and the right-hand-side term in LDS (natural order).  Each
// half sweep updates one colour from the other, so the result equals the
// separate dispatches of grid.c:228-233.  The right-hand-side term is the
// reference's subexpression (rho*h*dv/dt on level 0, r*h^2 above) evaluated
// once, which leaves every update bit-identical.
constexpr int kCoarseMaxCells = 4096;       // any level (coarse_fits)
constexpr int kCoarsestMaxCells = 12288;    // the coarsest level: 12 B/cell = 144 KiB LDS
template <int D>
__device__ __forceinline__ void coarse_smooth_body(const SimParams &P, const Level &L, int level, int iterations,
                                                   const float *rhs, float hl) {
    // pressure, solid fraction and right-hand-side term of every cell
    extern __shared__ float coarse_lds[];
    const int ncell = L.n[0] * L.n[1] * L.n[2];
    float *sp = coarse_lds, *sa = coarse_lds + ncell, *sb = coarse_lds + 2 * ncell;
    const int sx = L.n[0], sy = L.n[1], sz = L.n[2];
    const int n = sx * sy * sz;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        uint32_t si = sidx(L, i, j, k);
        float dv = rhs[si];
        sp[c] = L.p[si];
        sa[c] = solid_a(L.s, c);
        sb[c] = level == 0 && !P.corr ? P.rho * hl * dv / P.dt : dv * (hl * hl);
    }
    __syncthreads();
    const float oob = P.corr ? 1.0f : 0.0f;
    auto A = [&](int ii, int jj, int kk) -> float {
        return inside(ii, jj, kk, sx, sy, sz) ? sa[lin(ii, jj, kk, sx, sy)] : oob;
    };
    auto Pn = [&](int ii, int jj, int kk) -> float {
        return inside(ii, jj, kk, sx, sy, sz) ? sp[lin(ii, jj, kk, sx, sy)] : 0.0f;
    };
    int parity = 0;
    for (int it = 0; it < iterations * 2; ++it) {
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
            if (((i + j + k) & 1) != parity) continue;
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
            float psum = Pn(i, j + 1, k) * ft + Pn(i, j - 1, k) * fb;
            psum = psum + Pn(i + 1, j, k) * fr;
            psum = psum + Pn(i - 1, j, k) * fl;
            if (D == 3) {
                psum = psum + Pn(i, j, k + 1) * ff;
                psum = psum + Pn(i, j, k - 1) * fk;
            }
            sp[c] = (psum - sb[c]) / fsum;
        }
        __syncthreads();
        parity = 1 - parity;
    }
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        L.p[sidx(L, i, j, k)] = sp[c];
    }
}
template <int D>
__global__ void __launch_bounds__(1024) k_coarse_smooth(SimParams P, Level L, int level,
                                                        int iterations, const float *rhs,
                                                        float hl) {
    coarse_smooth_body<D>(P, L, level, iterations, rhs, hl);
}
// one workgroup per scene of a multi launch
template <int D>
__global__ void __launch_bounds__(1024) k_coarse_smooth_multi(const MCoarse *E) {
    const MCoarse &a = E[blockIdx.x];
    coarse_smooth_body<D>(a.P, a.L, a.level, a.iterations, a.rhs, a.hl);
}

// ---------------------------------------------------------------------------
// The bottom of the V-cycle in one workgroup.  Levels lb..L-1 (each at most a
// few thousand cells) run v_cycle(lb) of grid.c:177-209 entirely in LDS:
// pre-smoothing, residual (in place), restriction, coarse clear, the
// coarsest-level iterations, prolongation and post-smoothing, with a
// __syncthreads() between dependent passes.  Each thread keeps the stencil
// of the cells it owns (neighbour LDS slots, weights, diagonal, right-hand
// side term) in registers for the whole smoothing call, so a half sweep is
// six LDS loads and the update.  Every arithmetic step is the reference's.
// ---------------------------------------------------------------------------
constexpr int kBottomThreads = 1024;
// cells a thread owns per level in a half sweep (half of each colour): 4
// covers levels of up to 4096 cells, 6 up to 6144 (kBottomMaxCells)

template <int D, int kBottomOwn>
__device__ __forceinline__ void bottom_smooth(const BottomLevels &B, int m, int iterations, float *sp,
                              const float *sr, const float *sa, int zero_slot, float oob_f) {
    const int sx = B.n[m][0], sy = B.n[m][1], sz = B.n[m][2];
    const int n = sx * sy * sz, base = B.off[m], n0 = B.n0[m];
    const float hl = B.hl[m];
    // owned cells q: q < kBottomOwn/2 of colour 0, the others of colour 1,
    // from the colour-sorted list, so every lane of a half sweep has a cell
    // of that colour while the level has enough of them
    constexpr int kHalf = kBottomOwn / 2;
    int own[kBottomOwn], nb[kBottomOwn][6];
    float f[kBottomOwn][6], fsum[kBottomOwn], bterm[kBottomOwn];
#pragma unroll
    for (int q = 0; q < kBottomOwn; ++q) {
        const int par = q / kHalf;
        const int idx = threadIdx.x + (q % kHalf) * kBottomThreads;
        const int cnt = par ? n - n0 : n0;
        own[q] = -1;
        fsum[q] = 0.f;
        bterm[q] = 0.f;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            nb[q][e] = zero_slot;
            f[q][e] = 0.f;
        }
        if (idx >= cnt) continue;
        const int c = B.list[base + (par ? n0 : 0) + idx];
        const int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        const int off[6][3] = {{0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
        for (int e = 0; e < 2 * D; ++e) {
            const int ii = i + off[e][0], jj = j + off[e][1], kk = k + off[e][2];
            // smooth.comp solid_data(): OOB -> a = 0 (f = 1), pressure 0;
            // the correction solve of the refined mode has OOB solid (f = 0)
            if (inside(ii, jj, kk, sx, sy, sz)) {
                nb[q][e] = base + lin(ii, jj, kk, sx, sy);
                f[q][e] = 1.0f - sa[nb[q][e]];
            } else {
                f[q][e] = oob_f;
            }
        }
        fsum[q] = D == 3 ? f[q][0] + f[q][1] + f[q][2] + f[q][3] + f[q][4] + f[q][5]
                         : f[q][0] + f[q][1] + f[q][2] + f[q][3];
        // skip solid cells (smooth.comp:84) and f_sum == 0 (smooth.comp:102)
        if (!(sa[base + c] >= 1.0f) && fsum[q] != 0.0f) own[q] = base + c;
        bterm[q] = sr[base + c] * (hl * hl);
    }
    for (int it = 0; it < iterations * 2; ++it) {
        const int parity = it & 1;
#pragma unroll
        for (int q = 0; q < kBottomOwn; ++q) {
            if (q / kHalf != parity || own[q] < 0) continue;
            float psum = sp[nb[q][0]] * f[q][0] + sp[nb[q][1]] * f[q][1];
            psum = psum + sp[nb[q][2]] * f[q][2];
            psum = psum + sp[nb[q][3]] * f[q][3];
            if (D == 3) {
                psum = psum + sp[nb[q][4]] * f[q][4];
                psum = psum + sp[nb[q][5]] * f[q][5];
            }
            sp[own[q]] = (psum - bterm[q]) / fsum[q];
        }
        __syncthreads();
    }
}

// The coarsest level of the refined mode in FP64: the update of bottom_smooth
// (OOB solid) with pressure, right-hand-side term and weights in double,
// pressure staged in `dp` and stored back rounded to float.
template <int D, int kBottomOwn>
__device__ void bottom_smooth64(const BottomLevels &B, int m, int iterations, float *sp,
                                const float *sr, const float *sa, double *dp) {
    const int sx = B.n[m][0], sy = B.n[m][1], sz = B.n[m][2];
    const int n = sx * sy * sz, base = B.off[m];
    const double hl = (double)B.hl[m];
    for (int c = threadIdx.x; c < n; c += kBottomThreads) dp[c] = (double)sp[base + c];
    // the stencil of each owned cell in registers (neighbour slot n holds
    // 0); owned cells from the colour-sorted list as in bottom_smooth
    constexpr int kHalf = kBottomOwn / 2;
    const int n0 = B.n0[m];
    int own[kBottomOwn], nb[kBottomOwn][6];
    double f[kBottomOwn][6], fsum[kBottomOwn], term[kBottomOwn];
    const int off[6][3] = {{0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
    for (int q = 0; q < kBottomOwn; ++q) {
        const int pq = q / kHalf;
        const int idx = threadIdx.x + (q % kHalf) * kBottomThreads;
        own[q] = -1;
        fsum[q] = 0.0;
        term[q] = 0.0;
        for (int e = 0; e < 6; ++e) {
            nb[q][e] = n;
            f[q][e] = 0.0;
        }
        if (idx >= (pq ? n - n0 : n0)) continue;
        const int c = B.list[base + (pq ? n0 : 0) + idx];
        const int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        for (int e = 0; e < 2 * D; ++e) {
            const int ii = i + off[e][0], jj = j + off[e][1], kk = k + off[e][2];
            if (inside(ii, jj, kk, sx, sy, sz)) {
                nb[q][e] = lin(ii, jj, kk, sx, sy);
                f[q][e] = 1.0 - (double)sa[base + nb[q][e]];
            }
        }
        fsum[q] = D == 3 ? f[q][0] + f[q][1] + f[q][2] + f[q][3] + f[q][4] + f[q][5]
                         : f[q][0] + f[q][1] + f[q][2] + f[q][3];
        if (!(sa[base + c] >= 1.0f) && fsum[q] != 0.0) own[q] = c;
        term[q] = (double)sr[base + c] * (hl * hl);
    }
    if (threadIdx.x == 0) dp[n] = 0.0;
    __syncthreads();
    int parity = 0;
    for (int it = 0; it < iterations * 2; ++it) {
#pragma unroll
        for (int q = 0; q < kBottomOwn; ++q) {
            if (own[q] < 0 || q / kHalf != parity) continue;
            double psum = dp[nb[q][0]] * f[q][0] + dp[nb[q][1]] * f[q][1];
            psum = psum + dp[nb[q][2]] * f[q][2];
            psum = psum + dp[nb[q][3]] * f[q][3];
            if (D == 3) {
                psum = psum + dp[nb[q][4]] * f[q][4];
                psum = psum + dp[nb[q][5]] * f[q][5];
            }
            dp[own[q]] = (psum - term[q]) / fsum[q];
        }
        __syncthreads();
        parity = 1 - parity;
    }
    for (int c = threadIdx.x; c < n; c += kBottomThreads) sp[base + c] = (float)dp[c];
    __syncthreads();
}

// compute_residual.comp on level m > 0 (OOB neighbours solid, f = 0).
template <int D>
__device__ void bottom_residual(const BottomLevels &B, int m, const float *sp, float *sr,
                                const float *sa) {
    const int sx = B.n[m][0], sy = B.n[m][1], sz = B.n[m][2];
    const int n = sx * sy * sz, base = B.off[m];
    const float hl = B.hl[m];
    const int off[6][3] = {{0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int c = threadIdx.x; c < n; c += kBottomThreads) {
        const int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        float r = 0.0f;
        if (!(sa[base + c] >= 1.0f)) {
            float pv[6], fv[6];
            for (int e = 0; e < 2 * D; ++e) {
                const int ii = i + off[e][0], jj = j + off[e][1], kk = k + off[e][2];
                const bool in = inside(ii, jj, kk, sx, sy, sz);
                const int s = base + (in ? lin(ii, jj, kk, sx, sy) : 0);
                fv[e] = 1.0f - (in ? sa[s] : 1.0f);
                pv[e] = in ? sp[s] : 0.0f;
            }
            float psum = pv[0] * fv[0] + pv[1] * fv[1];
            psum = psum + pv[2] * fv[2];
            psum = psum + pv[3] * fv[3];
            float fsum;
            if (D == 3) {
                psum = psum + pv[4] * fv[4];
                psum = psum + pv[5] * fv[5];
                fsum = fv[0] + fv[1] + fv[2] + fv[3] + fv[4] + fv[5];
            } else {
                fsum = fv[0] + fv[1] + fv[2] + fv[3];
            }
            r = sr[base + c] - (psum - fsum * sp[base + c]) / (hl * hl);
        }
        sr[base + c] = r;
    }
}

// restrict_residual.comp + the coarse clear: level m -> m+1.
template <int D>
__device__ void bottom_restrict(const BottomLevels &B, int m, float *sp, float *sr) {
    const int fx = B.n[m][0], fy = B.n[m][1];
    const int cx = B.n[m + 1][0], cy = B.n[m + 1][1], cz = B.n[m + 1][2];
    const int fb = B.off[m], cb = B.off[m + 1];
    for (int c = threadIdx.x; c < cx * cy * cz; c += kBottomThreads) {
        const int i = c % cx, j = (c / cx) % cy, k = c / (cx * cy);
        const int I = 2 * i, J = 2 * j, K = D == 3 ? 2 * k : 0;
        auto R = [&](int a, int b, int d) { return sr[fb + lin(a, b, d, fx, fy)]; };
        float v;
        if (D == 3) {
            float bsum = R(I, J, K) + R(I, J, K + 1);
            bsum = bsum + R(I + 1, J, K);
            bsum = bsum + R(I + 1, J, K + 1);
            float tsum = R(I, J + 1, K) + R(I, J + 1, K + 1);
            tsum = tsum + R(I + 1, J + 1, K);
            tsum = tsum + R(I + 1, J + 1, K + 1);
            v = (bsum + tsum) / 8.0f;
        } else {
            float sum = R(I, J, 0) + R(I + 1, J, 0);
            sum = sum + R(I, J + 1, 0);
            sum = sum + R(I + 1, J + 1, 0);
            v = sum / 4.0f;
        }
        sr[cb + c] = v;
        sp[cb + c] = 0.0f;
    }
}

// prolong_pressure.comp: level m+1 -> m.
template <int D>
__device__ void bottom_prolong(const BottomLevels &B, int m, float *sp) {
    const int sx = B.n[m][0], sy = B.n[m][1], sz = B.n[m][2];
    const int cx = B.n[m + 1][0], cy = B.n[m + 1][1], cz = B.n[m + 1][2];
    const int fb = B.off[m], cb = B.off[m + 1];
    for (int c = threadIdx.x; c < sx * sy * sz; c += kBottomThreads) {
        const int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
        TC a = texcoord_tc(((float)i + 0.5f) / (float)sx, cx);
        TC b = texcoord_tc(((float)j + 0.5f) / (float)sy, cy);
        TC g = D == 3 ? texcoord_tc(((float)k + 0.5f) / (float)sz, cz) : TC{0, 0.0f};
        const int i0 = min(max(a.i0, 0), cx - 1), i1 = min(max(a.i0 + 1, 0), cx - 1);
        const int j0 = min(max(b.i0, 0), cy - 1), j1 = min(max(b.i0 + 1, 0), cy - 1);
        const int k0 = min(max(g.i0, 0), cz - 1), k1 = min(max(g.i0 + 1, 0), cz - 1);
        auto Q = [&](int ii, int jj, int kk) { return sp[cb + lin(ii, jj, kk, cx, cy)]; };
        float v;
        if (D == 2) {
            float x0 = lerpf(Q(i0, j0, 0), Q(i1, j0, 0), a.a);
            float x1 = lerpf(Q(i0, j1, 0), Q(i1, j1, 0), a.a);
            v = lerpf(x0, x1, b.a);
        } else {
            float x00 = lerpf(Q(i0, j0, k0), Q(i1, j0, k0), a.a);
            float x10 = lerpf(Q(i0, j1, k0), Q(i1, j1, k0), a.a);
            float x01 = lerpf(Q(i0, j0, k1), Q(i1, j0, k1), a.a);
            float x11 = lerpf(Q(i0, j1, k1), Q(i1, j1, k1), a.a);
            float y0 = lerpf(x00, x10, b.a);
            float y1 = lerpf(x01, x11, b.a);
            v = lerpf(y0, y1, g.a);
        }
        sp[fb + c] = sp[fb + c] + v;
    }
}

// corr: the correction solve of the refined mode (OOB solid in the
// smoother); C64: the coarsest level in FP64.
template <int D, bool C64, int OWN>
__device__ __forceinline__ void bottom_vcycle_body(const BottomLevels &B, int pre, int post, int coarse, int corr) {
    extern __shared__ double lds64[];
    const int total = B.off[B.nb - 1] + B.n[B.nb - 1][0] * B.n[B.nb - 1][1] * B.n[B.nb - 1][2];
    const int last = B.n[B.nb - 1][0] * B.n[B.nb - 1][1] * B.n[B.nb - 1][2];
    double *dp = lds64;            // FP64 coarsest pressure and a zero slot (C64 only)
    float *sp = reinterpret_cast<float *>(lds64 + (C64 ? last + 1 : 0));  // pressure, then a zero slot
    float *sr = sp + total + 1;    // residual / right-hand side
    float *sa = sr + total + 1;    // solid fraction
    const int zero_slot = total;
    const float oob_f = corr ? 0.0f : 1.0f;
    for (int m = 0; m < B.nb; ++m) {
        const int sx = B.n[m][0], sy = B.n[m][1], n = sx * sy * B.n[m][2];
        Level L;
        L.n[0] = sx;
        L.n[1] = sy;
        L.n[2] = B.n[m][2];
        L.hx = B.hx[m];
        for (int c = threadIdx.x; c < n; c += kBottomThreads) {
            const int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
            const uint32_t si = sidx(L, i, j, k);
            sa[B.off[m] + c] = solid_a(B.s[m], c);
            sp[B.off[m] + c] = m == 0 ? B.p[m][si] : 0.0f;
            sr[B.off[m] + c] = m == 0 ? B.r[m][si] : 0.0f;
        }
    }
    if (threadIdx.x == 0) sp[zero_slot] = 0.0f;
    __syncthreads();
    for (int m = 0; m + 1 < B.nb; ++m) {
        bottom_smooth<D, OWN>(B, m, pre, sp, sr, sa, zero_slot, oob_f);
        bottom_residual<D>(B, m, sp, sr, sa);
        __syncthreads();
        bottom_restrict<D>(B, m, sp, sr);
        __syncthreads();
    }
    if (C64) bottom_smooth64<D, OWN>(B, B.nb - 1, coarse, sp, sr, sa, dp);
    else bottom_smooth<D, OWN>(B, B.nb - 1, coarse, sp, sr, sa, zero_slot, oob_f);
    for (int m = B.nb - 2; m >= 0; --m) {
        bottom_prolong<D>(B, m, sp);
        __syncthreads();
        bottom_smooth<D, OWN>(B, m, post, sp, sr, sa, zero_slot, oob_f);
    }
    // write back pressure and residual of every bottom level
    for (int m = 0; m < B.nb; ++m) {
        const int sx = B.n[m][0], sy = B.n[m][1], n = sx * sy * B.n[m][2];
        Level L;
        L.n[0] = sx;
        L.n[1] = sy;
        L.n[2] = B.n[m][2];
        L.hx = B.hx[m];
        for (int c = threadIdx.x; c < n; c += kBottomThreads) {
            const int i = c % sx, j = (c / sx) % sy, k = c / (sx * sy);
            const uint32_t si = sidx(L, i, j, k);
            B.p[m][si] = sp[B.off[m] + c];
            B.r[m][si] = sr[B.off[m] + c];
        }
    }
}
template <int D, bool C64, int OWN>
__global__ void __launch_bounds__(kBottomThreads) k_bottom_vcycle(BottomLevels B, int pre,
                                                                  int post, int coarse, int corr) {
    bottom_vcycle_body<D, C64, OWN>(B, pre, post, coarse, corr);
}
// one workgroup per scene of a multi launch
template <int D, int OWN>
__global__ void __launch_bounds__(kBottomThreads) k_bottom_vcycle_multi(const MBottom *__restrict__ E) {
    const MBottom &a = E[blockIdx.x];
    bottom_vcycle_body<D, false, OWN>(a.B, a.pre, a.post, a.coarse, a.corr);
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement of the pressure solve.  The refined level-0
// pressure p64 is kept in double (colour-split like p).  One residual pass:
// p64' = p64 + e (the correction the last V-cycle left in the level-0 p,
// when add_e), r = b - A p64' with b = bnum*src/bden (the pressure solve:
// rho*dv/(dt*h) from the face divergence) and the operator of
// compute_residual.comp (OOB neighbours solid), all in double.  r rounded to
// float is the right-hand side of the next correction V-cycle.  Each block
// writes its partial |r|^2 and |b|^2 to part[2*block], part[2*block+1]
// (block = logical block index); k_sum_partials adds them in a fixed order.
// Reads p64 and e of the seven stencil points, so it writes the other p64
// buffer.
// ---------------------------------------------------------------------------
// r and b of one cell of the refinement residual (pc: its refined pressure).
template <int D>
__device__ __forceinline__ void refine_cell(const SimParams &P, const Level &L0, const double *pin,
                                            const float *e, int add_e, int i, int j, int k, int c,
                                            int m, int o, uint32_t row, uint32_t half, uint32_t msk,
                                            double pc, double bnum, float bs, double bden, double &r,
                                            double &b) {
    const int sx = L0.n[0], sy = L0.n[1], sz = L0.n[2], hx = L0.hx;
    r = 0.0;
    b = 0.0;
    if (msk & M_SOLID) return;
    double ft, fb, fr, fl, ff = 0.0, fk = 0.0;
    double qt, qb, qr, ql, qf = 0.0, qk = 0.0;
    if (!(msk & M_GENERAL)) {
        // interior cell: the neighbours are the other colour at fixed
        // offsets from its row
        const uint32_t ob = (uint32_t)(1 - c) * half + row;
        auto Q = [&](uint32_t s) -> double { return add_e ? pin[ob + s] + (double)e[ob + s] : pin[ob + s]; };
        ft = (msk & M_TOP) ? 1.0 : 0.0;
        fb = (msk & M_BOT) ? 1.0 : 0.0;
        fr = (msk & M_RIGHT) ? 1.0 : 0.0;
        fl = (msk & M_LEFT) ? 1.0 : 0.0;
        qt = Q(m + hx);
        qb = Q(m - hx);
        qr = Q(m + o);
        ql = Q(m - 1 + o);
        if (D == 3) {
            const uint32_t plane = (uint32_t)sy * hx;
            ff = (msk & M_FRONT) ? 1.0 : 0.0;
            fk = (msk & M_BACK) ? 1.0 : 0.0;
            qf = Q(m + plane);
            qk = Q(m - plane);
        }
    } else {
        auto Q = [&](int ii, int jj, int kk) -> double {
            const bool ok = inside(ii, jj, kk, sx, sy, sz);
            const uint32_t s = ok ? sidx(L0, ii, jj, kk) : 0u;
            double v = ok ? pin[s] : 0.0;
            if (add_e) v = v + (double)(ok ? e[s] : 0.0f);
            return ok ? v : 0.0;
        };
        auto A = [&](int ii, int jj, int kk) -> double {
            return (double)a_sel(L0, inside(ii, jj, kk, sx, sy, sz), ii, jj, kk, 1.0f);
        };
        ft = 1.0 - A(i, j + 1, k);
        fb = 1.0 - A(i, j - 1, k);
        fr = 1.0 - A(i + 1, j, k);
        fl = 1.0 - A(i - 1, j, k);
        qt = Q(i, j + 1, k);
        qb = Q(i, j - 1, k);
        qr = Q(i + 1, j, k);
        ql = Q(i - 1, j, k);
        if (D == 3) {
            ff = 1.0 - A(i, j, k + 1);
            fk = 1.0 - A(i, j, k - 1);
            qf = Q(i, j, k + 1);
            qk = Q(i, j, k - 1);
        }
    }
    double psum = qt * ft + qb * fb;
    psum = psum + qr * fr;
    psum = psum + ql * fl;
    double fsum;
    if (D == 3) {
        psum = psum + qf * ff;
        psum = psum + qk * fk;
        fsum = ft + fb + fr + fl + ff + fk;
    } else {
        fsum = ft + fb + fr + fl;
    }
    const double h = (double)P.h;
    b = bnum * (double)bs / bden;
    r = b - (psum - fsum * pc) / (h * h);
}

// block partial sums of r^2 and b^2: wave reduction, then the waves in
// order; written to part[2*block], part[2*block+1] (logical block index)
__device__ __forceinline__ void refine_partial(const XG &X_, int bx_, int by_, int bz_, double r2,
                                               double b2, double *part) {
    __shared__ double red[2][4];
    for (int w = 32; w > 0; w >>= 1) {
        r2 += __shfl_xor(r2, w);
        b2 += __shfl_xor(b2, w);
    }
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0) {
        red[0][t >> 6] = r2;
        red[1][t >> 6] = b2;
    }
    __syncthreads();
    if (t == 0) {
        const int nw = (blockDim.x * blockDim.y + 63) / 64;
        double sr = 0.0, sb = 0.0;
        for (int w = 0; w < nw; ++w) {
            sr += red[0][w];
            sb += red[1][w];
        }
        const uint32_t blk = ((uint32_t)bz_ * X_.gy + by_) * X_.gx + bx_;
        part[2 * blk] = sr;
        part[2 * blk + 1] = sb;
    }
}

template <int D>
__global__ void k_refine_residual(XG X_, SimParams P, Level L0, const double *pin, double *pout,
                                  const float *e, const float *bsrc, double bnum, double bden,
                                  float *rc, double *part, int add_e) {
    XG_BLOCK(X_)
    // thread (m, j, k, colour c): cell i = 2m + o of colour c in row (j, k),
    // visited in storage order
    const int m = bx_ * blockDim.x + threadIdx.x;
    const int j = by_ * blockDim.y + threadIdx.y;
    const int c = bz_ & 1, k = bz_ >> 1;
    const int sx = L0.n[0], sy = L0.n[1], hx = L0.hx;
    const int o = (c + j + k) & 1;
    const int i = 2 * m + o;
    double r = 0.0, b = 0.0;
    if (m < hx && j < sy && i < sx) {
        const uint32_t half = split_half_d(L0);
        const uint32_t row = ((uint32_t)k * sy + j) * hx;
        const uint32_t own = (uint32_t)c * half + row + m;
        const double pc = add_e ? pin[own] + (double)e[own] : pin[own];
        pout[own] = pc;
        refine_cell<D>(P, L0, pin, e, add_e, i, j, k, c, m, o, row, half, L0.mask[own], pc, bnum,
                       bsrc[own], bden, r, b);
        rc[own] = (float)r;
    }
    refine_partial(X_, bx_, by_, bz_, r * r, b * b, part);
}

// k_refine_residual in 3D for nx % 8 == 0, four cells of a colour row per
// thread (m = 4t .. 4t+3): the own row and the other colour's five
// neighbour rows come in 16-byte loads (two per four doubles), as in
// k_sweep4; rows on the domain boundary and M_GENERAL cells take
// refine_cell.  Per cell the arithmetic of k_refine_residual; a thread's
// four r^2 (b^2) are added in order before the block reduction.
__global__ void k_refine_residual4(XG X_, SimParams P, Level L0, const double *pin, double *pout,
                                   const float *e, const float *bsrc, double bnum, double bden,
                                   float *rc, double *part, int add_e) {
    XG_BLOCK(X_)
    const int t = bx_ * blockDim.x + threadIdx.x;
    const int j = by_ * blockDim.y + threadIdx.y;
    const int c = bz_ & 1, k = bz_ >> 1;
    const int sy = L0.n[1], sz = L0.n[2], hx = L0.hx;
    const int m0 = 4 * t;
    double r2 = 0.0, b2 = 0.0;
    if (m0 < hx && j < sy) {
        const int o = (c + j + k) & 1;
        const uint32_t half = split_half_d(L0);
        const uint32_t row = ((uint32_t)k * sy + j) * hx;
        const uint32_t own = (uint32_t)c * half + row + m0;
        // Every load of the thread before the first use (one memory round
        // trip): p64 (and e) of the own four and of the five neighbour rows'
        // fours (edge rows read their own row instead: in bounds, unused),
        // the x neighbour beyond the four (index m0-1 for o = 0, m0+4 for
        // o = 1, clamped into the row and only used inside it), the mask and
        // the source.  pout is the other p64 buffer, so its stores can wait.
        const bool edge_row = j == 0 || j == sy - 1 || k == 0 || k == sz - 1;
        const uint32_t ob = (uint32_t)(1 - c) * half + row;
        const uint32_t plane = (uint32_t)sy * hx;
        const uint32_t dy = edge_row ? 0u : (uint32_t)hx, dz = edge_row ? 0u : plane;
        const uint32_t src[6] = {own, ob + m0, ob + m0 + dy, ob + m0 - dy, ob + m0 + dz, ob + m0 - dz};
        const uint32_t xi = ob + (uint32_t)min(max(m0 - 1 + 5 * o, 0), hx - 1);
        double2 pa[6], pb[6];
        float4 ev[6];
        float ex = 0.0f;
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            pa[v] = *reinterpret_cast<const double2 *>(pin + src[v]);
            pb[v] = *reinterpret_cast<const double2 *>(pin + src[v] + 2);
            ev[v] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const double px = pin[xi];
        if (add_e) {
#pragma unroll
            for (int v = 0; v < 6; ++v) ev[v] = *reinterpret_cast<const float4 *>(e + src[v]);
            ex = e[xi];
        }
        const uint32_t mask4 = *reinterpret_cast<const uint32_t *>(L0.mask + own);
        const float4 bs4 = *reinterpret_cast<const float4 *>(bsrc + own);
        __builtin_amdgcn_sched_barrier(0);
        // the refined pressures p + e of the six fours
        double Q[6][4];
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            Q[v][0] = pa[v].x;
            Q[v][1] = pa[v].y;
            Q[v][2] = pb[v].x;
            Q[v][3] = pb[v].y;
            if (add_e) {
                Q[v][0] = Q[v][0] + (double)ev[v].x;
                Q[v][1] = Q[v][1] + (double)ev[v].y;
                Q[v][2] = Q[v][2] + (double)ev[v].z;
                Q[v][3] = Q[v][3] + (double)ev[v].w;
            }
        }
        const double *pc = Q[0];
        *reinterpret_cast<double2 *>(pout + own) = make_double2(pc[0], pc[1]);
        *reinterpret_cast<double2 *>(pout + own + 2) = make_double2(pc[2], pc[3]);
        const float bs[4] = {bs4.x, bs4.y, bs4.z, bs4.w};
        double r[4], b[4];
        if (!edge_row && (mask4 & 0x80808080u) == 0u) {
            const double *C = Q[1], *T = Q[2], *B = Q[3], *F = Q[4], *K = Q[5];
            const double qx = add_e ? px + (double)ex : px;
            const double h = (double)P.h;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t mk = (mask4 >> (8 * q)) & 0xffu;
                r[q] = 0.0;
                b[q] = 0.0;
                if (mk & M_SOLID) continue;
                const double pr = o ? (q < 3 ? C[q + 1] : qx) : C[q];
                const double pl = o ? C[q] : (q > 0 ? C[q - 1] : qx);
                const double ft = (mk & M_TOP) ? 1.0 : 0.0, fb = (mk & M_BOT) ? 1.0 : 0.0;
                const double fr = (mk & M_RIGHT) ? 1.0 : 0.0, fl = (mk & M_LEFT) ? 1.0 : 0.0;
                const double ff = (mk & M_FRONT) ? 1.0 : 0.0, fk = (mk & M_BACK) ? 1.0 : 0.0;
                double psum = T[q] * ft + B[q] * fb;
                psum = psum + pr * fr;
                psum = psum + pl * fl;
                psum = psum + F[q] * ff;
                psum = psum + K[q] * fk;
                const double fsum = ft + fb + fr + fl + ff + fk;
                b[q] = bnum * (double)bs[q] / bden;
                r[q] = b[q] - (psum - fsum * pc[q]) / (h * h);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + q;
                refine_cell<3>(P, L0, pin, e, add_e, 2 * m + o, j, k, c, m, o, row, half,
                               (mask4 >> (8 * q)) & 0xffu, pc[q], bnum, bs[q], bden, r[q], b[q]);
            }
        }
        *reinterpret_cast<float4 *>(rc + own) = make_float4((float)r[0], (float)r[1], (float)r[2], (float)r[3]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            r2 += r[q] * r[q];
            b2 += b[q] * b[q];
        }
    }
    refine_partial(X_, bx_, by_, bz_, r2, b2, part);
}

__global__ void __launch_bounds__(1024) k_sum_partials(const double *part, int n, double *out) {
    __shared__ double s[2][1024];
    const double2 *pp = reinterpret_cast<const double2 *>(part);
    double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + 3 * 1024 < n; i += 4 * 1024) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double2 v = pp[i + u * 1024];
            a[u] += v.x;
            b[u] += v.y;
        }
    }
    for (; i < n; i += 1024) {
        const double2 v = pp[i];
        a[0] += v.x;
        b[0] += v.y;
    }
    s[0][threadIdx.x] = (a[0] + a[1]) + (a[2] + a[3]);
    s[1][threadIdx.x] = (b[0] + b[1]) + (b[2] + b[3]);
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s[0][0];
        out[1] = s[1][0];
    }
}

// split <-> natural conversion of the FP64 refined pressure
__global__ void k_split64_to_f64(Level L, const double *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    for (uint32_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(c % L.n[0]), j = (int)((c / L.n[0]) % L.n[1]), k = (int)(c / ((size_t)L.n[0] * L.n[1]));
        dst[c] = src[sidx(L, i, j, k)];
    }
}
__global__ void k_split64_from_f64(Level L, const double *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    for (uint32_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(c % L.n[0]), j = (int)((c / L.n[0]) % L.n[1]), k = (int)(c / ((size_t)L.n[0] * L.n[1]));
        dst[sidx(L, i, j, k)] = src[c];
    }
}

__global__ void k_split_to_f64(Level L, const float *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    for (uint32_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(c % L.n[0]), j = (int)((c / L.n[0]) % L.n[1]), k = (int)(c / ((size_t)L.n[0] * L.n[1]));
        dst[c] = (double)src[sidx(L, i, j, k)];
    }
}
__global__ void k_split_from_f64(Level L, const double *src, float *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    for (uint32_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(c % L.n[0]), j = (int)((c / L.n[0]) % L.n[1]), k = (int)(c / ((size_t)L.n[0] * L.n[1]));
        dst[sidx(L, i, j, k)] = (float)src[c];
    }
}

// Level-0 residual norm: the residual stencil accumulated in FP64, one
// partial sum per block (refine_partial), added in a fixed order by
// k_sum_partials, so the norm is the same run to run.
template <int D, int VF>
__global__ void k_residual_norm(XG X_, SimParams P, Level L, VelPtrs V, double *part) {
    XG_BLOCK(X_)
    int i = bx_ * blockDim.x + threadIdx.x;
    int j = by_ * blockDim.y + threadIdx.y;
    int k = bz_;
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
    refine_partial(X_, bx_, by_, bz_, r2, 0.0, part);
}

// ---------------------------------------------------------------------------
// multi-scene kernels (fused batches)
// ---------------------------------------------------------------------------
template <int D, int VF, bool FOUR>
__global__ void k_divergence0_multi(const int *start, int n, const MDiv0 *E) {
    MULTI_FLAT_BLOCK(start, n, E)
    if constexpr (FOUR && D == 3) divergence0_x4_at<VF>(x_, y_, z_, a_.P, a_.V, a_.s0, a_.L0, a_.dv0, a_.b0);
    else divergence0_at<D, VF>(x_, y_, z_, a_.P, a_.V, a_.s0, a_.L0, a_.dv0, a_.b0);
}
// Each scene's threads (hx/4 threads per colour row with four cells each,
// or hx with one, then rows, then planes) taken as one linear range, 256
// threads per block: no lane idles at the end of a short row.
template <int D, bool FOUR>
__global__ void k_sweep_multi(const int *start, int n, const MSweep *E) {
    int s_, l_;
    if (!multi_block(start, n, s_, l_)) return;
    const MSweep &a_ = E[s_];
    const uint32_t idx = (uint32_t)l_ * 256u + threadIdx.x;
    const uint32_t tw = (uint32_t)a_.L.hx / (FOUR ? 4u : 1u), sy = (uint32_t)a_.L.n[1];
    const uint32_t row = idx / tw, k = row / sy;
    const int t = (int)(idx - row * tw), j = (int)(row - k * sy);
    if (FOUR) sweep4_at<D>(t, j, (int)k, a_.P, a_.L, a_.level, a_.parity, a_.rhs, a_.hl);
    else sweep_at<D>(t, j, (int)k, a_.P, a_.L, a_.level, a_.parity, a_.rhs, a_.hl);
}
template <int D, int VF, int VAR>
__global__ void __launch_bounds__(256) k_residual_restrict_multi(const int *start, int n, const MResRestrict *E) {
    MULTI_BLOCK(start, n, E)
    if (VAR == kRrX2) {
        if constexpr (D == 3)
            residual_restrict_even_x2_body<VF>(bx_, by_, bz_, a_.P, a_.F, a_.C, a_.level, a_.rhs, a_.V, a_.s0,
                                               a_.hl, a_.wf, a_.RD);
    } else if (VAR == kRrEven) {
        residual_restrict_even_body<D, VF>(bx_, by_, bz_, a_.P, a_.F, a_.C, a_.level, a_.rhs, a_.V, a_.s0, a_.hl,
                                           a_.wf);
    } else {
        MULTI_FLAT(a_, x_, y_, z_)
        residual_restrict_at<D, VF>(x_, y_, z_, a_.P, a_.F, a_.C, a_.level, a_.rhs, a_.V, a_.s0, a_.hl, 1);
    }
}
// kRrEvenB: k_residual_restrict_even_gb per scene (levels >= 1)
template <int D, int VF>
__global__ void __launch_bounds__(256) k_residual_restrict_even_gb_multi(const int *start, int n,
                                                                        const MResRestrict *E) {
    MULTI_BLOCK(start, n, E)
    residual_restrict_even_body<D, VF, true>(bx_, by_, bz_, a_.P, a_.F, a_.C, a_.level, a_.rhs, a_.V, a_.s0,
                                             a_.hl, a_.wf);
}
template <int D, int VAR>
__global__ void k_prolong_multi(const int *start, int n, const MProlong *E) {
    MULTI_BLOCK(start, n, E)
    if (VAR == kPrX2) {
        if constexpr (D == 3) prolong_even_x2_body(bx_, by_, bz_, a_.F, a_.C, a_.tx, a_.ty, a_.tz);
    } else if (VAR == kPrEven) {
        prolong_even_body<D>(bx_, by_, bz_, a_.F, a_.C, a_.tx, a_.ty, a_.tz);
    } else {
        MULTI_FLAT(a_, x_, y_, z_)
        prolong_at<D>(x_, y_, z_, a_.F, a_.C, a_.tx, a_.ty, a_.tz);
    }
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
    Shape sh = shape_for(P, L.n[0], L.n[1], L.n[2]);
    DISPATCH_D(P, k_build_mask<D_><<<sh.g, sh.b, 0, s>>>(sh.x, L));
}

void launch_divergence0(hipStream_t s, const SimParams &P, int vfmt, VelPtrs vel,
                        const uint16_t *solid0, const Level &L0, float *dv0, float *b0) {
    if (P.dim == 3 && P.n[0] % 4 == 0) {
        Shape s4 = shape_for(P, P.n[0] / 4, P.n[1], P.n[2]);
        DISPATCH_VF(vfmt, k_divergence0_x4<VF_><<<s4.g, s4.b, 0, s>>>(s4.x, P, vel, solid0, L0, dv0, b0));
        return;
    }
    Shape sh = shape_for(P, P.n[0], P.n[1], P.n[2]);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_divergence0<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, vel, solid0, L0, dv0, b0)));
}

void launch_sweep(hipStream_t s, const SimParams &P, const Level &L, int level, int parity,
                  const float *rhs, hipEvent_t e0, hipEvent_t e1) {
    float hl = level_h(P, level);
    // four cells per thread pays off only on large levels (small ones need
    // every thread they can get to hide latency): from 2^20 cells of a colour
    // on level 0, where walls and rasterised obstacles have solid fraction 0
    // or 1 so next to nothing takes the general update, from 2^22 on the
    // coarser levels, whose fractional solid fractions along walls and
    // obstacles send more cells through the general update (one memory round
    // trip each)
    if (L.hx % 4 == 0 && (long)L.hx * L.n[1] * L.n[2] >= (level == 0 ? (1L << 20) : (1L << 22))) {
        Shape sh = shape_for(P, L.hx / 4, L.n[1], L.n[2]);
        if (e0 || e1)
            DISPATCH_D(P, hipExtLaunchKernelGGL(k_sweep4<D_>, sh.g, sh.b, 0, s, e0, e1, 0u, sh.x, P, L, level,
                                                parity, rhs, hl));
        else
            DISPATCH_D(P, k_sweep4<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, L, level, parity, rhs, hl));
        return;
    }
    Shape sh = shape_for(P, L.hx, L.n[1], L.n[2]);
    if (e0 || e1)
        DISPATCH_D(P, hipExtLaunchKernelGGL(k_sweep<D_>, sh.g, sh.b, 0, s, e0, e1, 0u, sh.x, P, L, level, parity,
                                            rhs, hl));
    else
        DISPATCH_D(P, k_sweep<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, L, level, parity, rhs, hl));
}

// Two coarse cells per thread pays off only when that still leaves about
// half a million threads; smaller levels are latency bound.
static bool x2_level(const Level &coarse) {
    return (long)coarse.n[0] * coarse.n[1] * coarse.n[2] >= (1L << 20);
}

static bool even_pair(const SimParams &P, const Level &fine, const Level &coarse) {
    return fine.n[0] == 2 * coarse.n[0] && fine.n[1] == 2 * coarse.n[1] &&
           (P.dim == 2 || fine.n[2] == 2 * coarse.n[2]);
}

void launch_residual_restrict(hipStream_t s, const SimParams &P, const Level &fine,
                              const Level &coarse, int level, const float *rhs, int vfmt,
                              VelPtrs vel, const uint16_t *solid0, bool write_fine) {
    if (even_pair(P, fine, coarse)) {
        Shape sh = shape_for(P, coarse.n[0], coarse.n[1], coarse.n[2]);
        float hl = level_h(P, level);
        const int wf = (write_fine || level > 0) ? 1 : 0;
        // small levels are latency bound: two lanes per coarse cell halve the
        // dependent work per lane (level 0 and large levels: one lane)
        if (P.dim == 3 && fine.n[0] % 4 == 0 && x2_level(coarse)) {
            Shape s2 = shape_for(P, coarse.n[0] / 2, coarse.n[1], coarse.n[2]);
            const ResDiv rd = res_div(P, hl);
            DISPATCH_VF(vfmt, k_residual_restrict_even_x2<VF_><<<s2.g, s2.b, 0, s>>>(s2.x, P, fine, coarse,
                                                                  level, rhs, vel, solid0, hl, wf, rd));
            return;
        }
        if (P.dim == 3 && (long)coarse.n[0] * coarse.n[1] * coarse.n[2] < 64L * 64 * 64) {
            Shape s2 = shape_for(P, 2 * coarse.n[0], coarse.n[1], coarse.n[2]);
            if (level > 0 || P.corr)
                DISPATCH_VF(vfmt, k_residual_restrict_even_dj_gb<VF_><<<s2.g, s2.b, 0, s>>>(s2.x, P, fine, coarse,
                                                                     level, rhs, vel, solid0, hl, wf));
            else
                DISPATCH_VF(vfmt, k_residual_restrict_even_dj<VF_><<<s2.g, s2.b, 0, s>>>(s2.x, P, fine, coarse,
                                                                  level, rhs, vel, solid0, hl, wf));
            return;
        }
        // levels >= 1: the general cells' loads batched (level 0 keeps the
        // leaner kernel: its general cells are rare and it needs occupancy)
        if (level > 0 || P.corr)
            DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_restrict_even_gb<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x,
                                               P, fine, coarse, level, rhs, vel, solid0, hl, wf)));
        else
            DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_restrict_even<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x,
                                               P, fine, coarse, level, rhs, vel, solid0, hl, wf)));
        return;
    }
    Shape sh = shape_for(P, (fine.n[0] + 1) / 2, (fine.n[1] + 1) / 2,
                      P.dim == 3 ? (fine.n[2] + 1) / 2 : 1);
    float hl = level_h(P, level);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_restrict<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, 
                                       P, fine, coarse, level, rhs, vel, solid0, hl, 1)));
}

void launch_residual(hipStream_t s, const SimParams &P, const Level &L, int level,
                     const float *rhs, int vfmt, VelPtrs vel, const uint16_t *solid0) {
    Shape sh = shape_for(P, (L.n[0] + 1) / 2, (L.n[1] + 1) / 2, P.dim == 3 ? (L.n[2] + 1) / 2 : 1);
    float hl = level_h(P, level);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_restrict<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, 
                                       P, L, L, level, rhs, vel, solid0, hl, 0)));
}

void launch_restrict_solid(hipStream_t s, const SimParams &P, const Level &fine,
                           const Level &coarse) {
    Shape sh = shape_for(P, coarse.n[0], coarse.n[1], coarse.n[2]);
    DISPATCH_D(P, k_restrict_solid<D_><<<sh.g, sh.b, 0, s>>>(sh.x, fine, coarse));
}

void launch_prolong_table(hipStream_t s, int nf, int nc, TC *e) {
    k_prolong_table<<<(nf + 63) / 64, 64, 0, s>>>(nf, nc, e);
}

void launch_prolong(hipStream_t s, const SimParams &P, const Level &fine, const Level &coarse,
                    const TC *tx, const TC *ty, const TC *tz, bool even_ok) {
    if (even_ok && even_pair(P, fine, coarse)) {
        if (P.dim == 3 && fine.n[0] % 4 == 0 && x2_level(coarse)) {
            Shape s2 = shape_for(P, coarse.n[0] / 2, coarse.n[1], coarse.n[2]);
            k_prolong_even_x2<<<s2.g, s2.b, 0, s>>>(s2.x, fine, coarse, tx, ty, tz);
            return;
        }
        Shape sh = shape_for(P, coarse.n[0], coarse.n[1], coarse.n[2]);
        DISPATCH_D(P, k_prolong_even<D_><<<sh.g, sh.b, 0, s>>>(sh.x, fine, coarse, tx, ty, tz));
        return;
    }
    Shape sh = shape_for(P, fine.n[0], fine.n[1], fine.n[2]);
    DISPATCH_D(P, k_prolong<D_><<<sh.g, sh.b, 0, s>>>(sh.x, fine, coarse, tx, ty, tz));
}

bool coarse_fits(const SimParams &P, const Level &L) {
    (void)P;
    return (long)L.n[0] * L.n[1] * L.n[2] <= kCoarseMaxCells;
}
bool coarsest_fits(const Level &L) { return (long)L.n[0] * L.n[1] * L.n[2] <= kCoarsestMaxCells; }
size_t coarse_lds_bytes(const Level &L) { return (size_t)L.n[0] * L.n[1] * L.n[2] * 3 * sizeof(float); }
// dynamic LDS above 64 KiB needs the kernel's limit raised once
template <class K>
static void allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

void launch_coarse_smooth(hipStream_t s, const SimParams &P, const Level &L, int level,
                          int iterations, const float *rhs) {
    float hl = level_h(P, level);
    int n = L.n[0] * L.n[1] * L.n[2];
    int threads = n >= 1024 ? 1024 : ((n + 63) / 64) * 64;
    const size_t lds = coarse_lds_bytes(L);
    DISPATCH_D(P, {
        allow_lds(k_coarse_smooth<D_>, lds);
        k_coarse_smooth<D_><<<1, threads, lds, s>>>(P, L, level, iterations, rhs, hl);
    });
}

size_t bottom_lds_bytes(const BottomLevels &B, bool c64) {
    int last = B.n[B.nb - 1][0] * B.n[B.nb - 1][1] * B.n[B.nb - 1][2];
    int total = B.off[B.nb - 1] + last;
    return (size_t)(total + 1) * 3 * sizeof(float) + (c64 ? (size_t)(last + 1) * sizeof(double) : 0);
}

int bottom_own(const BottomLevels &B) {
    int most = 0;
    for (int m = 0; m < B.nb; ++m) most = std::max(most, B.n[m][0] * B.n[m][1] * B.n[m][2]);
    return most <= 4 * kBottomThreads ? 4 : 6;
}
void launch_bottom_vcycle(hipStream_t s, const SimParams &P, const BottomLevels &B, int pre,
                          int post, int coarse, bool c64) {
    size_t lds = bottom_lds_bytes(B, c64);
    const bool six = bottom_own(B) == 6;
    DISPATCH_D(P, {
        if (c64 && six) {
            allow_lds(k_bottom_vcycle<D_, true, 6>, lds);
            k_bottom_vcycle<D_, true, 6><<<1, kBottomThreads, lds, s>>>(B, pre, post, coarse, P.corr);
        } else if (c64) {
            allow_lds(k_bottom_vcycle<D_, true, 4>, lds);
            k_bottom_vcycle<D_, true, 4><<<1, kBottomThreads, lds, s>>>(B, pre, post, coarse, P.corr);
        } else if (six) {
            allow_lds(k_bottom_vcycle<D_, false, 6>, lds);
            k_bottom_vcycle<D_, false, 6><<<1, kBottomThreads, lds, s>>>(B, pre, post, coarse, P.corr);
        } else {
            allow_lds(k_bottom_vcycle<D_, false, 4>, lds);
            k_bottom_vcycle<D_, false, 4><<<1, kBottomThreads, lds, s>>>(B, pre, post, coarse, P.corr);
        }
    });
}

// four cells per thread for 3D levels with nx % 8 == 0
static bool refine_x4(const Level &L0) { return L0.n[2] > 1 && L0.n[0] % 8 == 0; }
static Shape refine_shape(const Level &L0) {
    return shape_for(refine_x4(L0) ? L0.hx / 4 : L0.hx, L0.n[1], 2 * L0.n[2]);
}
size_t refine_partials(const Level &L0) {
    Shape sh = refine_shape(L0);
    return (size_t)sh.x.gx * sh.x.gy * sh.x.gz;
}

void launch_refine_residual(hipStream_t s, const SimParams &P, const Level &L0, const double *pin,
                            double *pout, const float *e, const float *bsrc, double bnum,
                            double bden, float *rc, double *part, double *acc, int add_e) {
    Shape sh = refine_shape(L0);
    if (P.dim == 3 && refine_x4(L0))
        k_refine_residual4<<<sh.g, sh.b, 0, s>>>(sh.x, P, L0, pin, pout, e, bsrc, bnum, bden, rc, part, add_e);
    else
    DISPATCH_D(P, k_refine_residual<D_><<<sh.g, sh.b, 0, s>>>(sh.x, P, L0, pin, pout, e, bsrc, bnum,
                                                             bden, rc, part, add_e));
    const int n = (int)((size_t)sh.x.gx * sh.x.gy * sh.x.gz);
    k_sum_partials<<<1, 1024, 0, s>>>(part, n, acc);
}
void launch_split64_to_f64(hipStream_t s, const Level &L, const double *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    k_split64_to_f64<<<flat_blocks_p(n), 256, 0, s>>>(L, src, dst);
}
void launch_split64_from_f64(hipStream_t s, const Level &L, const double *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    k_split64_from_f64<<<flat_blocks_p(n), 256, 0, s>>>(L, src, dst);
}

void launch_split_to_f64(hipStream_t s, const Level &L, const float *src, double *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    k_split_to_f64<<<flat_blocks_p(n), 256, 0, s>>>(L, src, dst);
}
void launch_split_from_f64(hipStream_t s, const Level &L, const double *src, float *dst) {
    size_t n = (size_t)L.n[0] * L.n[1] * L.n[2];
    k_split_from_f64<<<flat_blocks_p(n), 256, 0, s>>>(L, src, dst);
}

size_t residual_norm_partials(const SimParams &P, const Level &L0) {
    const Shape sh = shape_for(P, L0.n[0], L0.n[1], L0.n[2]);
    return (size_t)sh.x.gx * sh.x.gy * sh.x.gz;
}
void launch_residual_norm(hipStream_t s, const SimParams &P, const Level &L0, int vfmt,
                          VelPtrs vel, double *part, double *acc) {
    Shape sh = shape_for(P, L0.n[0], L0.n[1], L0.n[2]);
    DISPATCH_D(P, DISPATCH_VF(vfmt, k_residual_norm<D_, VF_><<<sh.g, sh.b, 0, s>>>(sh.x, P, L0, vel, part)));
    k_sum_partials<<<1, 1024, 0, s>>>(part, (int)((size_t)sh.x.gx * sh.x.gy * sh.x.gz), acc);
}


// multi-scene launchers
void launch_divergence0_multi(hipStream_t s, int dim, int vfmt, bool four, const MultiLaunch &m, const MDiv0 *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    const dim3 g = multi_grid(m), b = multi_block_dim(m);
    DISPATCH_D(P, DISPATCH_VF(vfmt, {
        if (four) k_divergence0_multi<D_, VF_, true><<<g, b, 0, s>>>(m.start, m.n, E);
        else k_divergence0_multi<D_, VF_, false><<<g, b, 0, s>>>(m.start, m.n, E);
    }));
}
void launch_sweep_multi(hipStream_t s, int dim, bool four, const MultiLaunch &m, const MSweep *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    const dim3 g = multi_grid(m), b = multi_block_dim(m);
    DISPATCH_D(P, {
        if (four) k_sweep_multi<D_, true><<<g, b, 0, s>>>(m.start, m.n, E);
        else k_sweep_multi<D_, false><<<g, b, 0, s>>>(m.start, m.n, E);
    });
}
void launch_residual_restrict_multi(hipStream_t s, int dim, int vfmt, int variant, const MultiLaunch &m,
                                    const MResRestrict *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    const dim3 g = multi_grid(m), b = multi_block_dim(m);
    DISPATCH_D(P, DISPATCH_VF(vfmt, {
        if (variant == kRrX2) k_residual_restrict_multi<D_, VF_, kRrX2><<<g, b, 0, s>>>(m.start, m.n, E);
        else if (variant == kRrEven) k_residual_restrict_multi<D_, VF_, kRrEven><<<g, b, 0, s>>>(m.start, m.n, E);
        else if (variant == kRrEvenB) k_residual_restrict_even_gb_multi<D_, VF_><<<g, b, 0, s>>>(m.start, m.n, E);
        else k_residual_restrict_multi<D_, VF_, kRrGeneral><<<g, b, 0, s>>>(m.start, m.n, E);
    }));
}
void launch_prolong_multi(hipStream_t s, int dim, int variant, const MultiLaunch &m, const MProlong *E) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    const dim3 g = multi_grid(m), b = multi_block_dim(m);
    DISPATCH_D(P, {
        if (variant == kPrX2) k_prolong_multi<D_, kPrX2><<<g, b, 0, s>>>(m.start, m.n, E);
        else if (variant == kPrEven) k_prolong_multi<D_, kPrEven><<<g, b, 0, s>>>(m.start, m.n, E);
        else k_prolong_multi<D_, kPrGeneral><<<g, b, 0, s>>>(m.start, m.n, E);
    });
}
// The variant of a fused launch for one scene, and the size of the block
// grid it tiles: the single-scene launchers' choices without their
// launch-size thresholds (a fused launch has the threads of every scene).
FusedVariant fused_variant_residual_restrict(const SimParams &P, const Level &F, const Level &C, int level,
                                             bool write_fine) {
    FusedVariant v;
    if (even_pair(P, F, C)) {
        v.wf = (write_fine || level > 0) ? 1 : 0;
        v.RD = res_div(P, level_h(P, level));
        if (P.dim == 3 && F.n[0] % 4 == 0) {
            v.variant = kRrX2;
            v.s[0] = C.n[0] / 2, v.s[1] = C.n[1], v.s[2] = C.n[2];
            return v;
        }
        v.variant = level > 0 || P.corr ? kRrEvenB : kRrEven;
        v.s[0] = C.n[0], v.s[1] = C.n[1], v.s[2] = C.n[2];
        return v;
    }
    v.variant = kRrGeneral;
    v.wf = 1;
    v.s[0] = (F.n[0] + 1) / 2, v.s[1] = (F.n[1] + 1) / 2, v.s[2] = P.dim == 3 ? (F.n[2] + 1) / 2 : 1;
    return v;
}
FusedVariant fused_variant_prolong(const SimParams &P, const Level &F, const Level &C, bool even_ok) {
    FusedVariant v;
    if (even_ok && even_pair(P, F, C)) {
        if (P.dim == 3 && F.n[0] % 4 == 0) {
            v.variant = kPrX2;
            v.s[0] = C.n[0] / 2, v.s[1] = C.n[1], v.s[2] = C.n[2];
            return v;
        }
        v.variant = kPrEven;
        v.s[0] = C.n[0], v.s[1] = C.n[1], v.s[2] = C.n[2];
        return v;
    }
    v.variant = kPrGeneral;
    v.s[0] = F.n[0], v.s[1] = F.n[1], v.s[2] = F.n[2];
    return v;
}
void launch_coarse_smooth_multi(hipStream_t s, int dim, const MultiLaunch &m, const MCoarse *E, size_t lds) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    DISPATCH_D(P, {
        allow_lds(k_coarse_smooth_multi<D_>, lds);
        k_coarse_smooth_multi<D_><<<(unsigned)m.n, 1024, lds, s>>>(E);
    });
}
void launch_bottom_vcycle_multi(hipStream_t s, int dim, const MultiLaunch &m, const MBottom *E, size_t lds,
                                int own) {
    if (m.n <= 0) return;
    const SimParams P{{0, 0, 0}, dim};
    DISPATCH_D(P, {
        if (own == 6) {
            allow_lds(k_bottom_vcycle_multi<D_, 6>, lds);
            k_bottom_vcycle_multi<D_, 6><<<(unsigned)m.n, kBottomThreads, lds, s>>>(E);
        } else {
            allow_lds(k_bottom_vcycle_multi<D_, 4>, lds);
            k_bottom_vcycle_multi<D_, 4><<<(unsigned)m.n, kBottomThreads, lds, s>>>(E);
        }
    });
}
SMK_NS_END
