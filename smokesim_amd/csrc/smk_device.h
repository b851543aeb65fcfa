// smk_device.h -- device helpers shared by the kernel translation units:
// storage formats, texel addressing, the GL LINEAR sampler restatement and
// launch shapes.
#pragma once

#include <hip/hip_runtime.h>

#include "smk_kernels.h"

// Built with SMK_GL_ARITH=1, the kernel translation units compile a second
// time into namespace smk::gl with the two pieces of arithmetic that the GL
// specification leaves to the implementation made as Mesa llvmpipe makes
// them: binary16 stores round toward zero (f32_to_f16) and the linear
// filter's lerp is one fused multiply-add (lerpf).  The default build rounds
// to nearest even and lerps with a multiply and an add.  (SMK_NS_BEGIN /
// SMK_NS_END come from smk_kernels.h.)

SMK_NS_BEGIN

// ---------------------------------------------------------------------------
// storage formats
// ---------------------------------------------------------------------------
__device__ __forceinline__ float f16_to_f32(uint16_t b) {
    return (float)__builtin_bit_cast(_Float16, b);
}
__device__ __forceinline__ uint16_t f32_to_f16(float x) {
#if SMK_GL_ARITH
    // toward zero: the binary16 value of largest magnitude not above |x|
    // (65504 when |x| is larger), formed in float and converted exactly
    const uint32_t u = __builtin_bit_cast(uint32_t, x), ax = u & 0x7fffffffu;
    float t;
    if (ax >= 0x7f800000u) t = x;                                   // inf, nan
    else if (ax >= 0x477fe000u) t = __builtin_copysignf(65504.0f, x);
    else if (ax >= 0x38800000u) t = __builtin_bit_cast(float, u & ~0x1fffu);  // normal: 10 bits
    else t = __builtin_copysignf(floorf(__builtin_fabsf(x) * 0x1p24f) * 0x1p-24f, x);  // subnormal
    return __builtin_bit_cast(uint16_t, (_Float16)t);
#else
    return __builtin_bit_cast(uint16_t, (_Float16)x);  // v_cvt_f16_f32, RTNE
#endif
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
    return __builtin_bit_cast(float, (uint32_t)b << 16);
}
__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u | 0x00400000u) >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <int F>
struct Store;
template <>
struct Store<FMT_F32> {
    using T = float;
    static __device__ __forceinline__ float ld(T v) { return v; }
    static __device__ __forceinline__ T st(float v) { return v; }
};
template <>
struct Store<FMT_BF16> {
    using T = uint16_t;
    static __device__ __forceinline__ float ld(T v) { return bf16_to_f32(v); }
    static __device__ __forceinline__ T st(float v) { return f32_to_bf16(v); }
};
template <>
struct Store<FMT_F16> {
    using T = uint16_t;
    static __device__ __forceinline__ float ld(T v) { return f16_to_f32(v); }
    static __device__ __forceinline__ T st(float v) { return f32_to_f16(v); }
};

// 4 x 16-bit texel (smoke, 3D vorticity, solid map).
struct alignas(8) q16 {
    uint16_t x, y, z, w;
};

template <int F>
__device__ __forceinline__ float4 ld4(const uint16_t *base, uint32_t cell) {
    q16 q = reinterpret_cast<const q16 *>(base)[cell];
    using S = Store<F == FMT_F32 ? FMT_F16 : F>;
    return make_float4(S::ld(q.x), S::ld(q.y), S::ld(q.z), S::ld(q.w));
}
template <int F>
__device__ __forceinline__ void st4(uint16_t *base, uint32_t cell, float4 v) {
    using S = Store<F == FMT_F32 ? FMT_F16 : F>;
    q16 q;
    q.x = S::st(v.x);
    q.y = S::st(v.y);
    q.z = S::st(v.z);
    q.w = S::st(v.w);
    reinterpret_cast<q16 *>(base)[cell] = q;
}

// solid fraction a (channel w) of an fp16x4 solid texel
__device__ __forceinline__ float solid_a(const uint16_t *s, uint32_t cell) {
    return f16_to_f32(reinterpret_cast<const q16 *>(s)[cell].w);
}

// ---------------------------------------------------------------------------
// geometry helpers
// ---------------------------------------------------------------------------
// Element index of (i,j,k) in an x-fastest array.  32-bit: the host limits
// every field to fewer than 2^32 elements (1024^3 grids fit).
__device__ __forceinline__ uint32_t lin(int i, int j, int k, int sx, int sy) {
    return ((uint32_t)k * (uint32_t)sy + (uint32_t)j) * (uint32_t)sx + (uint32_t)i;
}
__device__ __forceinline__ bool inside(int i, int j, int k, int sx, int sy, int sz) {
    return (unsigned)i < (unsigned)sx && (unsigned)j < (unsigned)sy && (unsigned)k < (unsigned)sz;
}

// Bounds-checked gather without a branch: the load always executes (at
// element 0 when out of range) and the result is selected afterwards.  A
// load under a branch makes the compiler wait for it before the branch
// joins, which serialises the independent gathers of a stencil.
template <class T>
__device__ __forceinline__ T ld_sel(const T *p, bool ok, uint32_t idx, T other) {
    T v = p[ok ? idx : 0u];
    return ok ? v : other;
}
template <int F>
__device__ __forceinline__ float4 ld4_sel(const uint16_t *base, bool ok, uint32_t cell, float4 other) {
    float4 v = ld4<F>(base, ok ? cell : 0u);
    return ok ? v : other;
}

// Sizes of velocity component c.
template <int D>
__device__ __forceinline__ void vsize(const SimParams &P, int c, int &sx, int &sy, int &sz) {
    sx = P.n[0] + (c == 0);
    sy = P.n[1] + (c == 1);
    sz = P.n[2] + (D == 3 && c == 2);
}

// GL LINEAR texel coordinate: u' = u*size - 0.5, i0 = floor(u'), a = u'-i0.
// Coordinates far outside are clamped to [-4, size+4]: CLAMP_TO_BORDER and
// CLAMP_TO_EDGE both give the same value there as at the clamp point.
__device__ __forceinline__ void texcoord(float uv, int n, int &i0, float &a) {
    float u = uv * (float)n - 0.5f;
    if (!(u > -4.0f)) u = -4.0f;
    if (u > (float)(n + 4)) u = (float)(n + 4);
    float fl = floorf(u);
    i0 = (int)fl;
    a = u - fl;
}
__device__ __forceinline__ float lerpf(float v0, float v1, float a) {
#if SMK_GL_ARITH
    return __builtin_fmaf(a, v1 - v0, v0);
#else
    return v0 + a * (v1 - v0);
#endif
}

// x / y rounded to nearest even, from r = RN(1/y): q = RN(x*r), the exact
// remainder rem = x - q*y (one fma), then RN(q + rem*r) (Markstein's
// correction), three instructions instead of the IEEE division sequence.
// The host sets `fast` only after comparing this with IEEE division for
// every significand of one binade of x, for each divisor it is used with;
// every step scales exactly with x's exponent while all intermediates stay
// normal.  x = +-0 gives q = x*r, the signed zero IEEE division returns
// (sign of x xor sign of y; r is finite and has the sign of y).  Lanes with
// other x outside [2^-60, 2^60] (tiny, huge, non-finite) take the IEEE
// division, behind a branch that is uniform over the wave.
__device__ __forceinline__ float div_rn(float x, float y, float r, bool fast) {
    if (!fast) return x / y;
    const float q = x * r;
    const float rem = __builtin_fmaf(-q, y, x);
    float res = __builtin_fmaf(rem, r, q);
    const float ax = fabsf(x);
    res = ax == 0.0f ? q : res;
    const bool ok = (ax >= 0x1p-60f && ax <= 0x1p60f) || ax == 0.0f;
    if (__builtin_amdgcn_ballot_w64(!ok) != 0) res = ok ? res : x / y;
    return res;
}

// CLAMP_TO_BORDER (border 0) LINEAR sample of one velocity component.
template <int D, int VF>
__device__ float sample_vel(const SimParams &P, const void *tex, int c, const float *uv) {
    using S = Store<VF>;
    const typename S::T *t = static_cast<const typename S::T *>(tex);
    int sx, sy, sz;
    vsize<D>(P, c, sx, sy, sz);
    int i0, j0, k0 = 0;
    float a, b, g = 0.f;
    texcoord(uv[0], sx, i0, a);
    texcoord(uv[1], sy, j0, b);
    if (D == 3) texcoord(uv[2], sz, k0, g);
    auto F = [&](int i, int j, int k) -> float {
        const bool ok = inside(i, j, k, sx, sy, sz);
        return S::ld(ld_sel(t, ok, lin(i, j, k, sx, sy), typename S::T(0)));
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

// CLAMP_TO_BORDER LINEAR sample of smoke with border (0,0,0,t_ambient).
template <int D, int SF>
__device__ float4 sample_smoke(const SimParams &P, const uint16_t *sm, const float *uv) {
    const int sx = P.n[0], sy = P.n[1], sz = P.n[2];
    int i0, j0, k0 = 0;
    float a, b, g = 0.f;
    texcoord(uv[0], sx, i0, a);
    texcoord(uv[1], sy, j0, b);
    if (D == 3) texcoord(uv[2], sz, k0, g);
    const float4 border = make_float4(0.f, 0.f, 0.f, P.tamb);
    auto F = [&](int i, int j, int k) -> float4 {
        return ld4_sel<SF>(sm, inside(i, j, k, sx, sy, sz), lin(i, j, k, sx, sy), border);
    };
    auto L4 = [](float4 u, float4 v, float w) {
        return make_float4(lerpf(u.x, v.x, w), lerpf(u.y, v.y, w), lerpf(u.z, v.z, w),
                           lerpf(u.w, v.w, w));
    };
    if (D == 2) {
        float4 x0 = L4(F(i0, j0, 0), F(i0 + 1, j0, 0), a);
        float4 x1 = L4(F(i0, j0 + 1, 0), F(i0 + 1, j0 + 1, 0), a);
        return L4(x0, x1, b);
    }
    float4 x00 = L4(F(i0, j0, k0), F(i0 + 1, j0, k0), a);
    float4 x10 = L4(F(i0, j0 + 1, k0), F(i0 + 1, j0 + 1, k0), a);
    float4 x01 = L4(F(i0, j0, k0 + 1), F(i0 + 1, j0, k0 + 1), a);
    float4 x11 = L4(F(i0, j0 + 1, k0 + 1), F(i0 + 1, j0 + 1, k0 + 1), a);
    float4 y0 = L4(x00, x10, b);
    float4 y1 = L4(x01, x11, b);
    return L4(y0, y1, g);
}

// CLAMP_TO_EDGE LINEAR sample of a pressure level.
template <int D>
__device__ float sample_edge(const float *t, int sx, int sy, int sz, const float *uv) {
    int i0, j0, k0 = 0;
    float a, b, g = 0.f;
    texcoord(uv[0], sx, i0, a);
    texcoord(uv[1], sy, j0, b);
    if (D == 3) texcoord(uv[2], sz, k0, g);
    auto F = [&](int i, int j, int k) -> float {
        i = min(max(i, 0), sx - 1);
        j = min(max(j, 0), sy - 1);
        k = min(max(k, 0), sz - 1);
        return t[lin(i, j, k, sx, sy)];
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

__device__ __forceinline__ TC texcoord_tc(float uv, int n) {
    TC t;
    texcoord(uv, n, t.i0, t.a);
    return t;
}

// Coordinates of the face-centred velocity interpolation, which depend only
// on cell indices: for axis d, q = 2*s + t with s = 1 when the position is a
// face position along d (world = idx*h, else (idx+0.5)*h) and t = 1 when the
// sampled component is staggered along d (uv offset 0.5, size n+1).  Entry
// (d, q, idx) holds texcoord(((world/h) + 0.5t) / (n + t), n + t), the exact
// operation sequence of advect_vx.comp:36-43.
__device__ __forceinline__ TC ctab(const CoordTab &T, int d, int q, int idx) {
    return T.e[(uint32_t)(d * 4 + q) * (uint32_t)T.stride + (uint32_t)idx];
}

// LINEAR sample (border 0) of a velocity component at precomputed texel
// coordinates.  Branch free: every texel index is clamped into the texture,
// all loads issue together, and texels outside it are replaced by the
// border value 0 afterwards.
template <int D, int VF>
__device__ __forceinline__ float sample_vel_tc(const SimParams &P, const void *tex, int c, TC x,
                                               TC y, TC z) {
    using S = Store<VF>;
    const typename S::T *t = static_cast<const typename S::T *>(tex);
    int sx, sy, sz;
    vsize<D>(P, c, sx, sy, sz);
    const int xa = min(max(x.i0, 0), sx - 1), xb = min(max(x.i0 + 1, 0), sx - 1);
    const int ya = min(max(y.i0, 0), sy - 1), yb = min(max(y.i0 + 1, 0), sy - 1);
    const bool vxa = (unsigned)x.i0 < (unsigned)sx, vxb = (unsigned)(x.i0 + 1) < (unsigned)sx;
    const bool vya = (unsigned)y.i0 < (unsigned)sy, vyb = (unsigned)(y.i0 + 1) < (unsigned)sy;
    const uint32_t ra = (uint32_t)ya * sx, rb = (uint32_t)yb * sx;
    auto V = [&](uint32_t plane, uint32_t row, int xx, bool ok) -> float {
        float v = S::ld(t[plane + row + (uint32_t)xx]);
        return ok ? v : 0.0f;
    };
    if (D == 2) {
        float x0 = lerpf(V(0, ra, xa, vxa && vya), V(0, ra, xb, vxb && vya), x.a);
        float x1 = lerpf(V(0, rb, xa, vxa && vyb), V(0, rb, xb, vxb && vyb), x.a);
        return lerpf(x0, x1, y.a);
    }
    const int za = min(max(z.i0, 0), sz - 1), zb = min(max(z.i0 + 1, 0), sz - 1);
    const bool vza = (unsigned)z.i0 < (unsigned)sz, vzb = (unsigned)(z.i0 + 1) < (unsigned)sz;
    const uint32_t sxy = (uint32_t)sx * sy;
    const uint32_t pa = (uint32_t)za * sxy, pb = (uint32_t)zb * sxy;
    const float v000 = V(pa, ra, xa, vxa && vya && vza), v100 = V(pa, ra, xb, vxb && vya && vza);
    const float v010 = V(pa, rb, xa, vxa && vyb && vza), v110 = V(pa, rb, xb, vxb && vyb && vza);
    const float v001 = V(pb, ra, xa, vxa && vya && vzb), v101 = V(pb, ra, xb, vxb && vya && vzb);
    const float v011 = V(pb, rb, xa, vxa && vyb && vzb), v111 = V(pb, rb, xb, vxb && vyb && vzb);
    float x00 = lerpf(v000, v100, x.a);
    float x10 = lerpf(v010, v110, x.a);
    float x01 = lerpf(v001, v101, x.a);
    float x11 = lerpf(v011, v111, x.a);
    float y0 = lerpf(x00, x10, y.a);
    float y1 = lerpf(x01, x11, y.a);
    return lerpf(y0, y1, z.a);
}

// get_interpolated_velocity at the face of component c with index id (c = -1:
// at the cell centre), from the coordinate table.
template <int D, int VF>
__device__ __forceinline__ void interp_velocity_tab(const SimParams &P, const VelPtrs &V,
                                                    const CoordTab &T, const int *id, int c,
                                                    float *v) {
#pragma unroll
    for (int cc = 0; cc < D; ++cc) {
        TC t[3];
#pragma unroll
        for (int d = 0; d < D; ++d) t[d] = ctab(T, d, (d == c ? 2 : 0) + (d == cc ? 1 : 0), id[d]);
        if (D == 2) t[2] = TC{0, 0.0f};
        v[cc] = sample_vel_tc<D, VF>(P, V.c[cc], cc, t[0], t[1], t[2]);
    }
}

// ---------------------------------------------------------------------------
// Range-checked buffer sampling (CLAMP_TO_BORDER through the buffer unit)
// ---------------------------------------------------------------------------
// A raw buffer resource (stride 0) over a field of `bytes` bytes: a buffer
// load at a byte offset at or beyond num_records returns 0 and reads no
// memory.  The LINEAR samplers below send every texel outside the texture
// to offset kOob, so a border-0 sample needs no clamped indices, no select
// after the load and no 64-bit address; the row and plane products are
// 24-bit multiplies.  Texel offsets are computed modulo 2^32 from the
// unclamped corner (x0, y0, z0), exact for every texel inside the field.
// Usable when tex_buf_ok() holds for the grid (fields under 2 GiB, the
// 24-bit products in range); the launchers fall back to the clamped
// global-load samplers otherwise.
constexpr uint32_t kOob = 0x80000000u;
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ Rsrc make_rsrc(const void *p, uint32_t bytes) {
    // third descriptor word 0x00020000: 32-bit data format, as for every
    // gfx9 raw buffer
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             0x00020000);
}
template <int F>
__device__ __forceinline__ float bload(Rsrc r, uint32_t off) {
    if constexpr (F == FMT_F32)
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    else
        return Store<F>::ld(__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0));
}
// Bytes of velocity component c.
template <int D, int VF>
__device__ __forceinline__ uint32_t vbytes(const SimParams &P, int c) {
    int sx, sy, sz;
    vsize<D>(P, c, sx, sy, sz);
    return (uint32_t)sx * (uint32_t)sy * (uint32_t)sz * (uint32_t)sizeof(typename Store<VF>::T);
}

// sample_vel_tc through the range-checked buffer R of component c.
template <int D, int VF>
__device__ __forceinline__ float sample_vel_buf(const SimParams &P, Rsrc R, int c, TC x, TC y,
                                                TC z) {
    constexpr uint32_t E = sizeof(typename Store<VF>::T);
    int sx, sy, sz;
    vsize<D>(P, c, sx, sy, sz);
    const bool vxa = (unsigned)x.i0 < (unsigned)sx, vxb = (unsigned)(x.i0 + 1) < (unsigned)sx;
    const bool vya = (unsigned)y.i0 < (unsigned)sy, vyb = (unsigned)(y.i0 + 1) < (unsigned)sy;
    const uint32_t row = (uint32_t)sx * E;
    if (D == 2) {
        const uint32_t o = (uint32_t)(__mul24(y.i0, sx) + x.i0) * E;
        const float v00 = bload<VF>(R, vxa && vya ? o : kOob);
        const float v10 = bload<VF>(R, vxb && vya ? o + E : kOob);
        const float v01 = bload<VF>(R, vxa && vyb ? o + row : kOob);
        const float v11 = bload<VF>(R, vxb && vyb ? o + row + E : kOob);
        float x0 = lerpf(v00, v10, x.a);
        float x1 = lerpf(v01, v11, x.a);
        return lerpf(x0, x1, y.a);
    }
    const bool vza = (unsigned)z.i0 < (unsigned)sz, vzb = (unsigned)(z.i0 + 1) < (unsigned)sz;
    const uint32_t pl = (uint32_t)sx * (uint32_t)sy * E;
    const uint32_t o = (uint32_t)(__mul24(__mul24(z.i0, sy) + y.i0, sx) + x.i0) * E;
    const uint32_t o1 = o + pl;
    const float v000 = bload<VF>(R, vxa && vya && vza ? o : kOob);
    const float v100 = bload<VF>(R, vxb && vya && vza ? o + E : kOob);
    const float v010 = bload<VF>(R, vxa && vyb && vza ? o + row : kOob);
    const float v110 = bload<VF>(R, vxb && vyb && vza ? o + row + E : kOob);
    const float v001 = bload<VF>(R, vxa && vya && vzb ? o1 : kOob);
    const float v101 = bload<VF>(R, vxb && vya && vzb ? o1 + E : kOob);
    const float v011 = bload<VF>(R, vxa && vyb && vzb ? o1 + row : kOob);
    const float v111 = bload<VF>(R, vxb && vyb && vzb ? o1 + row + E : kOob);
    float x00 = lerpf(v000, v100, x.a);
    float x10 = lerpf(v010, v110, x.a);
    float x01 = lerpf(v001, v101, x.a);
    float x11 = lerpf(v011, v111, x.a);
    float y0 = lerpf(x00, x10, y.a);
    float y1 = lerpf(x01, x11, y.a);
    return lerpf(y0, y1, z.a);
}

// interp_velocity_tab through range-checked buffers.
template <int D, int VF>
__device__ __forceinline__ void interp_velocity_buf(const SimParams &P, const Rsrc *R,
                                                    const CoordTab &T, const int *id, int c,
                                                    float *v) {
#pragma unroll
    for (int cc = 0; cc < D; ++cc) {
        TC t[3];
#pragma unroll
        for (int d = 0; d < D; ++d) t[d] = ctab(T, d, (d == c ? 2 : 0) + (d == cc ? 1 : 0), id[d]);
        if (D == 2) t[2] = TC{0, 0.0f};
        v[cc] = sample_vel_buf<D, VF>(P, R[cc], cc, t[0], t[1], t[2]);
    }
}

// sample_smoke through the range-checked buffer R of the smoke field: the
// border's x, y, z channels are the 0 the buffer returns, only w (t_ambient)
// is selected.
template <int D, int SF>
__device__ float4 sample_smoke_buf(const SimParams &P, Rsrc R, const float *uv) {
    using S = Store<SF == FMT_F32 ? FMT_F16 : SF>;
    const int sx = P.n[0], sy = P.n[1], sz = P.n[2];
    int i0, j0, k0 = 0;
    float a, b, g = 0.f;
    texcoord(uv[0], sx, i0, a);
    texcoord(uv[1], sy, j0, b);
    if (D == 3) texcoord(uv[2], sz, k0, g);
    const bool vxa = (unsigned)i0 < (unsigned)sx, vxb = (unsigned)(i0 + 1) < (unsigned)sx;
    const bool vya = (unsigned)j0 < (unsigned)sy, vyb = (unsigned)(j0 + 1) < (unsigned)sy;
    const bool vza = D == 3 ? (unsigned)k0 < (unsigned)sz : true;
    const bool vzb = D == 3 ? (unsigned)(k0 + 1) < (unsigned)sz : true;
    const uint32_t row = (uint32_t)sx * 8u, pl = (uint32_t)sx * (uint32_t)sy * 8u;
    const uint32_t o = (uint32_t)(__mul24(__mul24(k0, sy) + j0, sx) + i0) * 8u;
    auto F = [&](uint32_t off, bool ok) -> float4 {
        const u32x2 q = __builtin_amdgcn_raw_buffer_load_b64(R, (int)(ok ? off : kOob), 0, 0);
        float4 v = make_float4(S::ld((uint16_t)(q.x & 0xffffu)), S::ld((uint16_t)(q.x >> 16)),
                               S::ld((uint16_t)(q.y & 0xffffu)), S::ld((uint16_t)(q.y >> 16)));
        v.w = ok ? v.w : P.tamb;
        return v;
    };
    auto L4 = [](float4 u, float4 v, float w) {
        return make_float4(lerpf(u.x, v.x, w), lerpf(u.y, v.y, w), lerpf(u.z, v.z, w),
                           lerpf(u.w, v.w, w));
    };
    if (D == 2) {
        float4 x0 = L4(F(o, vxa && vya), F(o + 8u, vxb && vya), a);
        float4 x1 = L4(F(o + row, vxa && vyb), F(o + row + 8u, vxb && vyb), a);
        return L4(x0, x1, b);
    }
    const uint32_t o1 = o + pl;
    float4 x00 = L4(F(o, vxa && vya && vza), F(o + 8u, vxb && vya && vza), a);
    float4 x10 = L4(F(o + row, vxa && vyb && vza), F(o + row + 8u, vxb && vyb && vza), a);
    float4 x01 = L4(F(o1, vxa && vya && vzb), F(o1 + 8u, vxb && vya && vzb), a);
    float4 x11 = L4(F(o1 + row, vxa && vyb && vzb), F(o1 + row + 8u, vxb && vyb && vzb), a);
    float4 y0 = L4(x00, x10, b);
    float4 y1 = L4(x01, x11, b);
    return L4(y0, y1, g);
}

// Host: whether the range-checked samplers are exact for this grid.  Every
// field (velocity components at 4 bytes, smoke at 8) stays under 2^31
// bytes, below kOob; the corner index (z0*sy + y0)*sx + x0 with texel
// coordinates in [-4, size+4] keeps both 24-bit multiplies in range.
static inline bool tex_buf_ok(const SimParams &P) {
    if (P.gsamp) return false;
    const long lim = 1L << 23;
    const int D = P.dim;
    for (int c = 0; c < D; ++c) {
        long s[3] = {P.n[0] + (c == 0), P.n[1] + (c == 1), D == 3 ? P.n[2] + (c == 2) : 1};
        if (s[0] + 8 >= lim || s[1] + 8 >= lim) return false;
        if (D == 3 && (s[2] + 8) * (s[1] + 8) >= lim) return false;
        if (s[0] * s[1] * s[2] * 4 >= (1L << 31)) return false;
    }
    const long cells = (long)P.n[0] * P.n[1] * (D == 3 ? P.n[2] : 1);
    return cells * 8 < (1L << 31);
}

// Level-0 solid flag (a >= 1) with an OOB value: 1 byte per cell, natural
// order, rebuilt whenever the level-0 solid map changes.
__device__ __forceinline__ bool solid_flag(const SimParams &P, const uint8_t *f0, int i, int j,
                                           int k, bool oob) {
    const bool ok = inside(i, j, k, P.n[0], P.n[1], P.n[2]);
    const uint8_t v = ld_sel(f0, ok, lin(i, j, k, P.n[0], P.n[1]), (uint8_t)(oob ? 1 : 0));
    return v != 0;
}

// Velocity at a world position: advect_vx.comp:33-46 get_interpolated_velocity.
template <int D, int VF>
__device__ void interp_velocity(const SimParams &P, const VelPtrs &V, const float *world,
                                float *v) {
    float gs[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < D; ++d) gs[d] = world[d] / P.h;
    for (int c = 0; c < D; ++c) {
        float uv[3] = {0.f, 0.f, 0.f};
        for (int d = 0; d < D; ++d)
            uv[d] = (gs[d] + (d == c ? 0.5f : 0.0f)) / ((float)P.n[d] + (d == c ? 1.0f : 0.0f));
        v[c] = sample_vel<D, VF>(P, V.c[c], c, uv);
    }
}

// is_solid(): level-0 solid fraction with a given OOB value.
template <int D>
__device__ __forceinline__ float solid_oob(const SimParams &P, const uint16_t *s0, int i, int j,
                                           int k, float oob) {
    if (!inside(i, j, k, P.n[0], P.n[1], P.n[2])) return oob;
    return solid_a(s0, lin(i, j, k, P.n[0], P.n[1]));
}

// 3D launch shape: 64 x 4 threads per workgroup, one wave per x-row segment.
static inline dim3 grid_for(int sx, int sy, int sz, dim3 blk) {
    return dim3((unsigned)((sx + (int)blk.x - 1) / (int)blk.x),
                (unsigned)((sy + (int)blk.y - 1) / (int)blk.y), (unsigned)sz);
}
static const dim3 kBlk(64, 4, 1);

// Launch shape for an sx x sy x sz thread domain: 256 threads per
// workgroup, x extent the largest power of two <= 64 that does not exceed
// sx (so narrow rows do not leave most of a wave idle), the rest along y.
//
// XCD-aware placement: the command processor hands workgroups to the 8 XCDs
// round-robin, so consecutive blocks land on different L2 caches.  Kernels
// are launched on a 1D grid and remap their workgroup id so that XCD x
// processes the contiguous range [x*per, (x+1)*per) of block ids, i.e. a
// slab of rows and planes whose stencil neighbours mostly stay in its own
// L2.  The remap is uniform per workgroup (scalar ALU).
struct XG {
    int gx, gy, gz;  // grid in blocks
    int per;         // blocks per XCD
};
constexpr int kXcds = 8;
struct Shape {
    dim3 g, b;
    XG x;
};
static inline Shape shape_for(int sx, int sy, int sz) {
    int bx = 64;
    while (bx > 1 && bx / 2 >= sx) bx >>= 1;
    int by = 256 / bx;
    Shape s;
    s.b = dim3((unsigned)bx, (unsigned)by, 1);
    s.x.gx = (sx + bx - 1) / bx;
    s.x.gy = (sy + by - 1) / by;
    s.x.gz = sz;
    long total = (long)s.x.gx * s.x.gy * s.x.gz;
    s.x.per = (int)((total + kXcds - 1) / kXcds);
    s.g = dim3((unsigned)(s.x.per * kXcds), 1, 1);
    return s;
}
// The launch shape of a grid whose tile width was chosen by the autotuner
// (P.tile_bx > 0: blocks of tile_bx x 256/tile_bx threads, narrowed to the
// extent in x).
static inline Shape shape_for(const SimParams &P, int sx, int sy, int sz) {
    if (P.tile_bx <= 0) return shape_for(sx, sy, sz);
    int bx = P.tile_bx;
    while (bx > 1 && bx / 2 >= sx) bx >>= 1;
    int by = 256 / bx;
    Shape s;
    s.b = dim3((unsigned)bx, (unsigned)by, 1);
    s.x.gx = (sx + bx - 1) / bx;
    s.x.gy = (sy + by - 1) / by;
    s.x.gz = sz;
    long total = (long)s.x.gx * s.x.gy * s.x.gz;
    s.x.per = (int)((total + kXcds - 1) / kXcds);
    s.g = dim3((unsigned)(s.x.per * kXcds), 1, 1);
    return s;
}
__device__ __forceinline__ bool xcd_block(const XG &X, int &bx, int &by, int &bz) {
    const int b = blockIdx.x;
    const int logical = (b % kXcds) * X.per + b / kXcds;
    const int gxy = X.gx * X.gy;
    if (logical >= gxy * X.gz) return false;
    bz = logical / gxy;
    const int rem = logical - bz * gxy;
    by = rem / X.gx;
    bx = rem - by * X.gx;
    return true;
}
#define XG_BLOCK(X)                                  \
    int bx_, by_, bz_;                               \
    if (!xcd_block((X), bx_, by_, bz_)) return;

// Multi-scene launches (MultiLaunch): logical block of this block, dealt to
// the XCDs in contiguous runs as xcd_block does (per = ceil(blocks / 8)),
// and the scene s whose range [start[s], start[s+1]) holds it (binary search
// over the n + 1 prefix sums; uniform over the block).
__device__ __forceinline__ bool multi_block(const int *start, int n, int &s, int &local) {
    const int total = start[n];
    const int per = (total + kXcds - 1) / kXcds;
    const int b = blockIdx.x;
    const int logical = (b % kXcds) * per + b / kXcds;
    if (logical >= total) return false;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= logical) lo = mid;
        else hi = mid;
    }
    s = lo;
    local = logical - start[lo];
    return true;
}
// Scene record a_ and tile (bx_, by_, bz_) of this block of a multi launch.
#define MULTI_BLOCK(start, n, E)                                                         \
    int s_, l_;                                                                          \
    if (!multi_block((start), (n), s_, l_)) return;                                     \
    const auto &a_ = (E)[s_];                                                            \
    const int bx_ = l_ % a_.t.gx, by_ = (l_ / a_.t.gx) % a_.t.gy, bz_ = l_ / (a_.t.gx * a_.t.gy);
// Thread (x_, y_, z_) of a flat multi launch: the scene's threads taken as
// one linear range of 256-thread blocks, the record's MTile holding the
// thread extents (fused_add_flat); the threads past the range return.
#define MULTI_FLAT(a, x_, y_, z_)                                                        \
    const uint32_t fi_ = (uint32_t)l_ * 256u + threadIdx.x;                              \
    const uint32_t fr_ = fi_ / (uint32_t)(a).t.gx;                                       \
    const int x_ = (int)(fi_ - fr_ * (uint32_t)(a).t.gx);                                \
    const int z_ = (int)(fr_ / (uint32_t)(a).t.gy);                                      \
    const int y_ = (int)(fr_ - (uint32_t)z_ * (uint32_t)(a).t.gy);                       \
    if (z_ >= (a).t.gz) return;
// Scene record a_ and thread (x_, y_, z_) of this block of a flat multi
// launch.
#define MULTI_FLAT_BLOCK(start, n, E)                                                    \
    int s_, l_;                                                                          \
    if (!multi_block((start), (n), s_, l_)) return;                                     \
    const auto &a_ = (E)[s_];                                                            \
    MULTI_FLAT(a_, x_, y_, z_)
// grid and block of a multi launch
static inline dim3 multi_grid(const MultiLaunch &m) {
    return dim3((unsigned)(((m.blocks + kXcds - 1) / kXcds) * kXcds), 1, 1);
}
static inline dim3 multi_block_dim(const MultiLaunch &m) { return dim3((unsigned)m.bx, (unsigned)(256 / m.bx), 1); }


// Colour-split address of cell (i,j,k) of a level (see struct Level).
__device__ __forceinline__ uint32_t split_half_d(const Level &L) {
    return (uint32_t)L.hx * (uint32_t)L.n[1] * (uint32_t)L.n[2];
}
__device__ __forceinline__ uint32_t sidx(const Level &L, int i, int j, int k) {
    uint32_t c = (uint32_t)((i + j + k) & 1);
    return c * split_half_d(L) + ((uint32_t)k * (uint32_t)L.n[1] + (uint32_t)j) * (uint32_t)L.hx +
           (uint32_t)(i >> 1);
}

// Face velocity of component c, zero outside its texture (texelFetch).
template <int D, int VF>
__device__ __forceinline__ float vget(const SimParams &P, const VelPtrs &V, int c, int i, int j,
                                      int k) {
    using S = Store<VF>;
    int sx, sy, sz;
    vsize<D>(P, c, sx, sy, sz);
    const bool ok = inside(i, j, k, sx, sy, sz);
    return S::ld(ld_sel(static_cast<const typename S::T *>(V.c[c]), ok, lin(i, j, k, sx, sy),
                        typename S::T(0)));
}
// solid_data() of the pressure kernels on level 0 with the OOB texel
// (0,0,0,oob_a): smooth.comp uses a = 0, compute_residual.comp a = 1.
template <int D>
__device__ __forceinline__ float4 solid_data0(const SimParams &P, const uint16_t *s0, int i, int j,
                                              int k, float oob_a) {
    return ld4_sel<FMT_F16>(s0, inside(i, j, k, P.n[0], P.n[1], P.n[2]),
                            lin(i, j, k, P.n[0], P.n[1]), make_float4(0.f, 0.f, 0.f, oob_a));
}

// Level-0 face divergence sum (smooth.comp 3D:105-113, compute_residual.comp
// 3D:116-123): a face whose neighbour is solid (a >= 1) takes that cell's
// obstacle velocity.
template <int D, int VF>
__device__ float level0_dv(const SimParams &P, const VelPtrs &V, const uint16_t *s0, int i, int j,
                           int k, float oob_a) {
    float4 sr = solid_data0<D>(P, s0, i + 1, j, k, oob_a);
    float4 sl = solid_data0<D>(P, s0, i - 1, j, k, oob_a);
    float4 st = solid_data0<D>(P, s0, i, j + 1, k, oob_a);
    float4 sb = solid_data0<D>(P, s0, i, j - 1, k, oob_a);
    // all face loads first (no branch), then the solid substitutions
    const float ft = vget<D, VF>(P, V, 1, i, j + 1, k), fb = vget<D, VF>(P, V, 1, i, j, k);
    const float fr = vget<D, VF>(P, V, 0, i + 1, j, k), fl = vget<D, VF>(P, V, 0, i, j, k);
    float vtop = st.w >= 1.0f ? st.y : ft;
    float vbot = sb.w >= 1.0f ? sb.y : fb;
    float vr = sr.w >= 1.0f ? sr.x : fr;
    float vl = sl.w >= 1.0f ? sl.x : fl;
    float dv = vr - vl + vtop - vbot;
    if (D == 3) {
        float4 sf = solid_data0<D>(P, s0, i, j, k + 1, oob_a);
        float4 sk = solid_data0<D>(P, s0, i, j, k - 1, oob_a);
        const float ff = vget<D, VF>(P, V, 2, i, j, k + 1), fk = vget<D, VF>(P, V, 2, i, j, k);
        float vf = sf.w >= 1.0f ? sf.z : ff;
        float vb = sk.w >= 1.0f ? sk.z : fk;
        dv = dv + vf - vb;
    }
    return dv;
}

// Template dispatch of launch wrappers on dimension / storage formats.
#define DISPATCH_D(P, ...)                  \
    do {                                    \
        if ((P).dim == 3) {                 \
            constexpr int D_ = 3;           \
            __VA_ARGS__;                    \
        } else {                            \
            constexpr int D_ = 2;           \
            __VA_ARGS__;                    \
        }                                   \
    } while (0)

#define DISPATCH_VF(vf, ...)                        \
    do {                                            \
        if ((vf) == FMT_F32) {                      \
            constexpr int VF_ = FMT_F32;            \
            __VA_ARGS__;                            \
        } else if ((vf) == FMT_BF16) {              \
            constexpr int VF_ = FMT_BF16;           \
            __VA_ARGS__;                            \
        } else {                                    \
            constexpr int VF_ = FMT_F16;            \
            __VA_ARGS__;                            \
        }                                           \
    } while (0)

#define DISPATCH_SF(sf, ...)                        \
    do {                                            \
        if ((sf) == FMT_BF16) {                     \
            constexpr int SF_ = FMT_BF16;           \
            __VA_ARGS__;                            \
        } else {                                    \
            constexpr int SF_ = FMT_F16;            \
            __VA_ARGS__;                            \
        }                                           \
    } while (0)


SMK_NS_END
