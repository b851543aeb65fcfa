/*
 * smk_oracle.c -- CPU restatement of SmokeSim's simulation step (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * One function per reference compute shader; every function cites the GLSL
 * file and lines it restates (paths relative to the reference repository).
 * "3D" cites shaders3d/compute, "2D" cites shaders2d/compute.
 *
 * Semantics kept from the reference on purpose:
 *   - ping-pong read/write textures with the same swap points as
 *     common/physics/grid.c (a stage that writes only some texels leaves
 *     the write texture's previous content elsewhere: apply_buoyancy);
 *   - the out-of-bounds convention of each kernel (solid / fluid / zero);
 *   - RGBA16F storage (smoke, 3D vorticity, solid pyramid) rounded to FP16
 *     round-to-nearest-even on every image store;
 *   - GL LINEAR filtering: u' = u*size - 0.5, i0 = floor(u'), a = u' - i0,
 *     lerp per axis x,y,z as v0 + a*(v1-v0); CLAMP_TO_BORDER returns the
 *     border colour for texels outside the image, CLAMP_TO_EDGE clamps the
 *     texel index (texture.c:54-78, 3d/grid3d.c:25-68);
 *   - image loads / texel fetches outside the image return 0.
 * Compiled with -ffp-contract=off so every expression rounds as written.
 */
#include "smk_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL R;

#define ORC_IS_DOUBLE (sizeof(R) == sizeof(double))

/* FMT_F16Z: binary16 rounded toward zero (see orc_set_f16_rounding) */
enum { FMT_F32 = 0, FMT_F16 = 1, FMT_BF16 = 2, FMT_F16Z = 3 };
enum { WRAP_BORDER = 0, WRAP_EDGE = 1 };

typedef struct {
    int n[3];
    int ch;
    int fmt;
    int wrap;
    int fused; /* linear filter lerps as one fused multiply-add */
    R border[4];
    R *d;
} otex;

#define MAXL 16

struct orc_grid {
    int dim;
    int n[3];
    R h;
    R tamb;
    otex vx[2], vy[2], vz[2], sm[2], vort;
    int vxi, vyi, vzi, smi;
    int L;
    int ls[MAXL][3];
    otex p[MAXL], r[MAXL], s[MAXL];
    R set[12];
    int ne;
    int *em;
    R *inflow;
    int no;
    R *ob[2];
    int obi;
    int nb;
    int *box;
    float time;
    int pre, post, coarse, vcycles;
    /* storage (smk_precision codes) and the temperature offset of 16-bit
     * smoke: with BF16 smoke the alpha channel holds T - toff */
    int storage;
    R toff;
    /* FP64 iterative refinement of the pressure solve (0: off) */
    int refine, coarse64, corr;
    double *p64;     /* level-0 pressure, natural order */
    double *norms;   /* |r|^2 after each residual pass, then |b|^2 */
    otex rc;         /* level-0 right-hand side of the correction solve */
};

/* ------------------------------------------------------------------ */
/* storage rounding                                                   */
/* ------------------------------------------------------------------ */

/* float -> binary16 -> float, round to nearest even (IEEE 754). */
float orc_round_f16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    uint32_t sign = u & 0x80000000u;
    uint32_t ax = u & 0x7fffffffu;
    uint16_t h;
    if (ax >= 0x7f800000u) {                   /* inf / nan */
        h = (uint16_t)((sign >> 16) | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    } else if (ax >= 0x477ff000u) {            /* >= 65520 rounds to inf */
        h = (uint16_t)((sign >> 16) | 0x7c00u);
    } else if (ax >= 0x38800000u) {            /* normal half */
        uint32_t mant = ax & 0x7fffffu;
        int32_t e = (int32_t)(ax >> 23) - 127 + 15;
        uint32_t hm = mant >> 13;
        uint32_t rem = mant & 0x1fffu;
        uint32_t v = ((uint32_t)e << 10) | hm;
        if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) v += 1u;
        h = (uint16_t)((sign >> 16) | v);
    } else {                                    /* subnormal half or zero */
        /* value = ax as float; half subnormal unit = 2^-24 */
        float f;
        memcpy(&f, &ax, 4);
        /* scale to units of 2^-24 and round to nearest even */
        double q = (double)f * 16777216.0;      /* exact */
        double fl = floor(q);
        double rem = q - fl;
        uint32_t v = (uint32_t)fl;
        if (rem > 0.5 || (rem == 0.5 && (v & 1u))) v += 1u;
        h = (uint16_t)((sign >> 16) | v);       /* v may carry into exp: ok */
    }
    /* back to float */
    uint32_t hs = (uint32_t)(h & 0x8000u) << 16;
    uint32_t he = (h >> 10) & 0x1fu;
    uint32_t hm = h & 0x3ffu;
    uint32_t out;
    if (he == 0x1fu) {
        out = hs | 0x7f800000u | (hm << 13);
    } else if (he == 0) {
        float f = (float)hm * (1.0f / 16777216.0f);
        memcpy(&out, &f, 4);
        out |= hs;
    } else {
        out = hs | ((he - 15u + 127u) << 23) | (hm << 13);
    }
    float r;
    memcpy(&r, &out, 4);
    return r;
}

/* float -> binary16 -> float, rounded toward zero: the magnitude is
 * truncated to the binary16 value at or below it (overflow gives 65504). */
float orc_round_f16_rtz(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    uint32_t ax = u & 0x7fffffffu;
    if (ax >= 0x7f800000u) return x;            /* inf / nan unchanged */
    float a;
    memcpy(&a, &ax, 4);
    float r;
    if (a >= 65504.0f) {
        r = 65504.0f;
    } else if (ax >= 0x38800000u) {             /* normal half: drop 13 bits */
        uint32_t t = ax & ~0x1fffu;
        memcpy(&r, &t, 4);
    } else {                                    /* subnormal half: units of 2^-24 */
        r = (float)(floor((double)a * 16777216.0) / 16777216.0);
    }
    return (u & 0x80000000u) ? -r : r;
}

/* float -> bfloat16 -> float, round to nearest even. */
float orc_round_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) {
        u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        uint32_t lsb = (u >> 16) & 1u;
        u += 0x7fffu + lsb;
        u &= 0xffff0000u;
    }
    float r;
    memcpy(&r, &u, 4);
    return r;
}

int orc_is_double(void) { return ORC_IS_DOUBLE ? 1 : 0; }

/* threads the stage loops run on (1 without OpenMP) */
int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* The per-cell loops of the stages run on OpenMP threads when the library
 * is built with -fopenmp: every cell (face) writes only its own value from
 * inputs that the loop does not write, or, in a red-black half sweep, from
 * cells of the other colour, so results do not depend on the thread count.
 * The FP64 residual norms of the refined solve stay sequential sums. */
#ifdef _OPENMP
#define ORC_PARALLEL_KJ _Pragma("omp parallel for collapse(2) schedule(static)")
#else
#define ORC_PARALLEL_KJ
#endif

static R rnd(int fmt, R v) {
    if (ORC_IS_DOUBLE) return v;
    if (fmt == FMT_F16) return (R)orc_round_f16((float)v);
    if (fmt == FMT_F16Z) return (R)orc_round_f16_rtz((float)v);
    if (fmt == FMT_BF16) return (R)orc_round_bf16((float)v);
    return v;
}

/* ------------------------------------------------------------------ */
/* textures                                                            */
/* ------------------------------------------------------------------ */

static int tex_alloc(otex *t, int nx, int ny, int nz, int ch, int fmt, int wrap,
                     const R *border) {
    t->n[0] = nx;
    t->n[1] = ny;
    t->n[2] = nz;
    t->ch = ch;
    t->fmt = fmt;
    t->wrap = wrap;
    for (int c = 0; c < 4; ++c) t->border[c] = border ? border[c] : (R)0;
    t->d = (R *)calloc((size_t)nx * ny * nz * ch, sizeof(R));
    return t->d != NULL;
}

static void tex_free(otex *t) {
    free(t->d);
    t->d = NULL;
}

static void tex_fill(otex *t, const R *val) {
    size_t cnt = (size_t)t->n[0] * t->n[1] * t->n[2];
    for (size_t i = 0; i < cnt; ++i)
        for (int c = 0; c < t->ch; ++c) t->d[i * t->ch + c] = rnd(t->fmt, val ? val[c] : (R)0);
}

static int inb(const otex *t, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < t->n[0] && j < t->n[1] && k < t->n[2];
}

static size_t off(const otex *t, int i, int j, int k) {
    return (((size_t)k * t->n[1] + j) * t->n[0] + i) * t->ch;
}

/* imageLoad / texelFetch: zero outside the image. */
static R ld(const otex *t, int i, int j, int k, int c) {
    if (!inb(t, i, j, k)) return (R)0;
    return t->d[off(t, i, j, k) + c];
}

/* imageStore: ignored outside the image, rounded to the storage format. */
static void st(otex *t, int i, int j, int k, int c, R v) {
    if (!inb(t, i, j, k)) return;
    t->d[off(t, i, j, k) + c] = rnd(t->fmt, v);
}

static R fetch_wrap(const otex *t, int i, int j, int k, int c) {
    int id[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        if (t->wrap == WRAP_EDGE) {
            if (id[d] < 0) id[d] = 0;
            if (id[d] > t->n[d] - 1) id[d] = t->n[d] - 1;
        } else if (id[d] < 0 || id[d] >= t->n[d]) {
            return t->border[c];
        }
    }
    return t->d[off(t, id[0], id[1], id[2]) + c];
}

static R lerp(R v0, R v1, R a) { return v0 + a * (v1 - v0); }
static R lerp_fused(R v0, R v1, R a) {
    return ORC_IS_DOUBLE ? (R)fma((double)a, (double)(v1 - v0), (double)v0)
                         : (R)fmaf((float)a, (float)(v1 - v0), (float)v0);
}
#define LERP(v0, v1, a) (t->fused ? lerp_fused(v0, v1, a) : lerp(v0, v1, a))

/* GL LINEAR sampling of a `dim`-dimensional texture at normalised uvw. */
static void tex_sample(const otex *t, int dim, const R *uvw, R *out) {
    int i0[3] = {0, 0, 0};
    R a[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
        R u = uvw[d] * (R)t->n[d] - (R)0.5;
        if (!(u > (R)-4)) u = (R)-4;
        if (u > (R)(t->n[d] + 4)) u = (R)(t->n[d] + 4);
        R fl = (R)floor((double)u);
        i0[d] = (int)fl;
        a[d] = u - fl;
    }
    for (int c = 0; c < t->ch; ++c) {
        if (dim == 2) {
            R v00 = fetch_wrap(t, i0[0], i0[1], 0, c);
            R v10 = fetch_wrap(t, i0[0] + 1, i0[1], 0, c);
            R v01 = fetch_wrap(t, i0[0], i0[1] + 1, 0, c);
            R v11 = fetch_wrap(t, i0[0] + 1, i0[1] + 1, 0, c);
            R x0 = LERP(v00, v10, a[0]);
            R x1 = LERP(v01, v11, a[0]);
            out[c] = LERP(x0, x1, a[1]);
        } else {
            R v000 = fetch_wrap(t, i0[0], i0[1], i0[2], c);
            R v100 = fetch_wrap(t, i0[0] + 1, i0[1], i0[2], c);
            R v010 = fetch_wrap(t, i0[0], i0[1] + 1, i0[2], c);
            R v110 = fetch_wrap(t, i0[0] + 1, i0[1] + 1, i0[2], c);
            R v001 = fetch_wrap(t, i0[0], i0[1], i0[2] + 1, c);
            R v101 = fetch_wrap(t, i0[0] + 1, i0[1], i0[2] + 1, c);
            R v011 = fetch_wrap(t, i0[0], i0[1] + 1, i0[2] + 1, c);
            R v111 = fetch_wrap(t, i0[0] + 1, i0[1] + 1, i0[2] + 1, c);
            R x00 = LERP(v000, v100, a[0]);
            R x10 = LERP(v010, v110, a[0]);
            R x01 = LERP(v001, v101, a[0]);
            R x11 = LERP(v011, v111, a[0]);
            R y0 = LERP(x00, x10, a[1]);
            R y1 = LERP(x01, x11, a[1]);
            out[c] = LERP(y0, y1, a[2]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* grid                                                               */
/* ------------------------------------------------------------------ */

static otex *vread(orc_grid *g, int c) {
    return c == 0 ? &g->vx[g->vxi] : c == 1 ? &g->vy[g->vyi] : &g->vz[g->vzi];
}
static otex *vwrite(orc_grid *g, int c) {
    return c == 0 ? &g->vx[1 - g->vxi] : c == 1 ? &g->vy[1 - g->vyi] : &g->vz[1 - g->vzi];
}
static void swap_v(orc_grid *g, int c) {
    if (c == 0) g->vxi = 1 - g->vxi;
    if (c == 1) g->vyi = 1 - g->vyi;
    if (c == 2) g->vzi = 1 - g->vzi;
}
static void swap_all_v(orc_grid *g) {
    for (int c = 0; c < g->dim; ++c) swap_v(g, c);
}

/* Pyramid rule of 2d/grid2d.c:45-63 and 3d/grid3d.c:45-68, with the
 * reference's fixed threshold 4 made a parameter. */
static int pyramid(orc_grid *g, int minsz) {
    int d[3] = {g->n[0], g->n[1], g->dim == 3 ? g->n[2] : minsz};
    int L = 0;
    while (d[0] >= minsz && d[1] >= minsz && d[2] >= minsz && L < MAXL) {
        g->ls[L][0] = g->n[0] >> L;
        g->ls[L][1] = g->n[1] >> L;
        g->ls[L][2] = g->dim == 3 ? (g->n[2] >> L) : 1;
        d[0] >>= 1;
        d[1] >>= 1;
        if (g->dim == 3) d[2] >>= 1;
        ++L;
    }
    return L;
}

orc_grid *orc_create(int nx, int ny, int nz, float cell_size, float t_ambient,
                     int min_level_size) {
    if (nx < 1 || ny < 1 || nz < 0 || min_level_size < 1) return NULL;
    orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
    if (!g) return NULL;
    g->dim = nz == 0 ? 2 : 3;
    g->n[0] = nx;
    g->n[1] = ny;
    g->n[2] = nz == 0 ? 1 : nz;
    g->h = (R)cell_size;
    g->tamb = (R)t_ambient;
    g->L = pyramid(g, min_level_size);
    if (g->L < 1) {
        free(g);
        return NULL;
    }
    int z1 = g->dim == 3 ? 1 : 0;
    R smoke_init[4] = {0, 0, 0, g->tamb};
    int ok = 1;
    for (int b = 0; b < 2; ++b) {
        ok &= tex_alloc(&g->vx[b], nx + 1, ny, g->n[2], 1, FMT_F32, WRAP_BORDER, NULL);
        ok &= tex_alloc(&g->vy[b], nx, ny + 1, g->n[2], 1, FMT_F32, WRAP_BORDER, NULL);
        if (g->dim == 3)
            ok &= tex_alloc(&g->vz[b], nx, ny, g->n[2] + z1, 1, FMT_F32, WRAP_BORDER, NULL);
        ok &= tex_alloc(&g->sm[b], nx, ny, g->n[2], 4, FMT_F16, WRAP_BORDER, smoke_init);
        if (ok) tex_fill(&g->sm[b], smoke_init);
    }
    ok &= tex_alloc(&g->vort, nx, ny, g->n[2], g->dim == 3 ? 4 : 1,
                    g->dim == 3 ? FMT_F16 : FMT_F32, WRAP_BORDER, NULL);
    for (int l = 0; l < g->L; ++l) {
        ok &= tex_alloc(&g->p[l], g->ls[l][0], g->ls[l][1], g->ls[l][2], 1, FMT_F32, WRAP_EDGE, NULL);
        ok &= tex_alloc(&g->r[l], g->ls[l][0], g->ls[l][1], g->ls[l][2], 1, FMT_F32, WRAP_EDGE, NULL);
        ok &= tex_alloc(&g->s[l], g->ls[l][0], g->ls[l][1], g->ls[l][2], 4, FMT_F16, WRAP_EDGE, NULL);
    }
    if (!ok) {
        orc_destroy(g);
        return NULL;
    }
    /* physics.c:7-20 defaults, t_ambient from the caller */
    R def[12] = {0, 0, 0, 0, (R)(1 / 120.0f), 1, g->tamb, (R)0.07f, 1, 1, 0, 0};
    memcpy(g->set, def, sizeof(def));
    g->pre = 16;
    g->post = 16;
    g->coarse = 80;
    g->vcycles = 1;
    if (!tex_alloc(&g->rc, g->ls[0][0], g->ls[0][1], g->ls[0][2], 1, FMT_F32, WRAP_EDGE, NULL)) {
        orc_destroy(g);
        return NULL;
    }
    return g;
}

/* Storage of velocity faces and smoke (smk_precision codes): 0 = FP32 faces
 * and FP16 smoke (the reference's formats), 1 = BF16 faces and BF16 smoke
 * holding T - t_ambient in the alpha channel, 2 = FP16 faces and smoke.
 * Call right after orc_create; the smoke is re-initialised. */
int orc_set_storage(orc_grid *g, int storage) {
    if (storage < 0 || storage > 2) return -1;
    int vf = storage == 0 ? FMT_F32 : (storage == 1 ? FMT_BF16 : FMT_F16);
    int sf = storage == 1 ? FMT_BF16 : FMT_F16;
    g->storage = storage;
    g->toff = storage == 1 ? g->tamb : (R)0;
    R init[4] = {0, 0, 0, g->tamb - g->toff};
    for (int b = 0; b < 2; ++b) {
        g->vx[b].fmt = vf;
        g->vy[b].fmt = vf;
        if (g->dim == 3) g->vz[b].fmt = vf;
        g->sm[b].fmt = sf;
        for (int c = 0; c < 4; ++c) g->sm[b].border[c] = init[c];
        tex_fill(&g->sm[b], init);
    }
    return 0;
}

/* Rounding of every binary16 store: 0 = to nearest even (the default),
 * 1 = toward zero.  The GL specification lets an implementation return
 * either neighbour when a float is stored to a 16-bit float format; Mesa's
 * llvmpipe truncates, so the oracle replays llvmpipe runs of the reference
 * shaders with mode 1.  Call after orc_set_storage. */
int orc_set_f16_rounding(orc_grid *g, int mode) {
    if (mode < 0 || mode > 1) return -1;
    int from = mode ? FMT_F16 : FMT_F16Z, to = mode ? FMT_F16Z : FMT_F16;
    otex *ts[4 + 3 * MAXL];
    int n = 0;
    for (int b = 0; b < 2; ++b) {
        ts[n++] = &g->vx[b];
        ts[n++] = &g->vy[b];
        if (g->dim == 3) ts[n++] = &g->vz[b];
        ts[n++] = &g->sm[b];
    }
    ts[n++] = &g->vort;
    for (int l = 0; l < g->L; ++l) ts[n++] = &g->s[l];
    for (int i = 0; i < n; ++i)
        if (ts[i]->fmt == from) ts[i]->fmt = to;
    return 0;
}

/* The linear filter's lerp v0 + a*(v1 - v0): 0 = a multiply and an add,
 * each rounded (the default, and what the HIP kernels compute), 1 = one
 * fused multiply-add, as Mesa llvmpipe's texture unit evaluates it.  The GL
 * specification leaves the filter's rounding to the implementation. */
int orc_set_lerp(orc_grid *g, int fused) {
    if (fused < 0 || fused > 1) return -1;
    for (int b = 0; b < 2; ++b) {
        g->vx[b].fused = fused;
        g->vy[b].fused = fused;
        g->vz[b].fused = fused;
        g->sm[b].fused = fused;
    }
    g->vort.fused = fused;
    for (int l = 0; l < g->L; ++l) {
        g->p[l].fused = fused;
        g->r[l].fused = fused;
        g->s[l].fused = fused;
    }
    return 0;
}

/* FP64 iterative refinement of the pressure solve: `outer` correction
 * V-cycles per solve, each followed by an FP64 residual pass; coarse64 runs
 * the coarsest level's smoothing in FP64.  outer = 0 restores the plain
 * solve of grid.c:173-209. */
int orc_set_refinement(orc_grid *g, int outer, int coarse64) {
    if (outer < 0) return -1;
    size_t n = (size_t)g->n[0] * g->n[1] * g->n[2];
    if (outer > 0 && !g->p64) {
        g->p64 = (double *)calloc(n, sizeof(double));
        if (!g->p64) return -1;
        /* start from the current level-0 pressure */
        for (size_t i = 0; i < n; ++i) g->p64[i] = (double)g->p[0].d[i];
    }
    free(g->norms);
    g->norms = (double *)calloc((size_t)outer + 2, sizeof(double));
    g->refine = outer;
    g->coarse64 = coarse64 != 0;
    return 0;
}

/* |r|^2 after each of the outer+1 residual passes of the last solve, then
 * |b|^2 (fluid cells). */
void orc_get_refine_norms(const orc_grid *g, double *out) {
    for (int i = 0; i < g->refine + 2; ++i) out[i] = g->norms ? g->norms[i] : 0.0;
}

void orc_destroy(orc_grid *g) {
    if (!g) return;
    tex_free(&g->rc);
    free(g->p64);
    free(g->norms);
    for (int b = 0; b < 2; ++b) {
        tex_free(&g->vx[b]);
        tex_free(&g->vy[b]);
        tex_free(&g->vz[b]);
        tex_free(&g->sm[b]);
    }
    tex_free(&g->vort);
    for (int l = 0; l < g->L; ++l) {
        tex_free(&g->p[l]);
        tex_free(&g->r[l]);
        tex_free(&g->s[l]);
    }
    free(g->em);
    free(g->inflow);
    free(g->ob[0]);
    free(g->ob[1]);
    free(g->box);
    free(g);
}

int orc_levels(const orc_grid *g) { return g->L; }
void orc_level_size(const orc_grid *g, int level, int out[3]) {
    out[0] = g->ls[level][0];
    out[1] = g->ls[level][1];
    out[2] = g->ls[level][2];
}

void orc_set_physics(orc_grid *g, const float *s) {
    for (int i = 0; i < 12; ++i) g->set[i] = (R)s[i];
}
void orc_set_emitters(orc_grid *g, const int *em, int n) {
    free(g->em);
    free(g->inflow);
    g->ne = n;
    g->em = (int *)calloc((size_t)(n > 0 ? n : 1) * 12, sizeof(int));
    g->inflow = (R *)calloc((size_t)(n > 0 ? n : 1), sizeof(R));
    if (n > 0) memcpy(g->em, em, (size_t)n * 12 * sizeof(int));
}
void orc_set_emitter_inflow(orc_grid *g, const float *inflow, int n) {
    for (int i = 0; i < n && i < g->ne; ++i) g->inflow[i] = (R)inflow[i];
}
void orc_get_emitters(const orc_grid *g, int *em) {
    if (g->ne > 0) memcpy(em, g->em, (size_t)g->ne * 12 * sizeof(int));
}
void orc_set_obstacles(orc_grid *g, const float *ob, int n) {
    free(g->ob[0]);
    free(g->ob[1]);
    g->no = n;
    g->obi = 0;
    for (int b = 0; b < 2; ++b) {
        g->ob[b] = (R *)calloc((size_t)(n > 0 ? n : 1) * 8, sizeof(R));
        for (int i = 0; i < n * 8; ++i) g->ob[b][i] = (R)ob[i];
    }
}
void orc_get_obstacles(const orc_grid *g, float *ob) {
    for (int i = 0; i < g->no * 8; ++i) ob[i] = (float)g->ob[g->obi][i];
}
void orc_set_boxes(orc_grid *g, const int *boxes, int n) {
    free(g->box);
    g->nb = n;
    g->box = (int *)calloc((size_t)(n > 0 ? n : 1) * 6, sizeof(int));
    if (n > 0) memcpy(g->box, boxes, (size_t)n * 6 * sizeof(int));
}
void orc_set_time(orc_grid *g, float t) { g->time = t; }
void orc_set_mg(orc_grid *g, int pre, int post, int coarse, int vcycles) {
    g->pre = pre;
    g->post = post;
    g->coarse = coarse;
    g->vcycles = vcycles;
}

#define WIND(c) (g->set[(c)])
#define DT (g->set[4])
#define RHO (g->set[5])
#define TAMB (g->set[6] - g->toff)
#define ALPHA (g->set[7])
#define BETA (g->set[8])
#define EPS (g->set[9])

/* ------------------------------------------------------------------ */
/* field I/O                                                           */
/* ------------------------------------------------------------------ */

static otex *field_tex(orc_grid *g, int field, int level) {
    switch (field) {
    case ORC_VX: return &g->vx[g->vxi];
    case ORC_VY: return &g->vy[g->vyi];
    case ORC_VZ: return g->dim == 3 ? &g->vz[g->vzi] : NULL;
    case ORC_SMOKE: return &g->sm[g->smi];
    case ORC_VORTICITY: return &g->vort;
    case ORC_PRESSURE: return level >= 0 && level < g->L ? &g->p[level] : NULL;
    case ORC_RESIDUAL: return level >= 0 && level < g->L ? &g->r[level] : NULL;
    case ORC_SOLID: return level >= 0 && level < g->L ? &g->s[level] : NULL;
    default: return NULL;
    }
}

long orc_field_count(const orc_grid *g, int field, int level) {
    if (field == ORC_DIVERGENCE) return (long)g->n[0] * g->n[1] * g->n[2];
    otex *t = field_tex((orc_grid *)g, field, level);
    if (!t) return -1;
    return (long)t->n[0] * t->n[1] * t->n[2] * t->ch;
}

static R level0_dv(orc_grid *g, int i, int j, int k, int oob_solid);

long orc_get_field(const orc_grid *g_, int field, int level, double *out) {
    orc_grid *g = (orc_grid *)g_;
    if (field == ORC_DIVERGENCE) {
        long n = 0;
        for (int k = 0; k < g->n[2]; ++k)
            for (int j = 0; j < g->n[1]; ++j)
                for (int i = 0; i < g->n[0]; ++i) {
                    R a = ld(&g->s[0], i, j, k, 3);
                    out[n++] = a >= (R)1 ? 0.0 : (double)(level0_dv(g, i, j, k, 0) / g->h);
                }
        return n;
    }
    if (field == ORC_PRESSURE && level == 0 && g->refine > 0) {
        long n = (long)g->n[0] * g->n[1] * g->n[2];
        for (long i = 0; i < n; ++i) out[i] = g->p64[i];
        return n;
    }
    otex *t = field_tex(g, field, level);
    if (!t) return -1;
    long cnt = (long)t->n[0] * t->n[1] * t->n[2] * t->ch;
    for (long i = 0; i < cnt; ++i) out[i] = (double)t->d[i];
    if (field == ORC_SMOKE && g->toff != (R)0)
        for (long i = 3; i < cnt; i += 4) out[i] = out[i] + (double)g->toff;
    return cnt;
}

long orc_set_field(orc_grid *g, int field, int level, const double *in) {
    if (field == ORC_PRESSURE && level == 0 && g->refine > 0) {
        long n = (long)g->n[0] * g->n[1] * g->n[2];
        for (long i = 0; i < n; ++i) g->p64[i] = in[i];
        return n;
    }
    otex *t = field_tex(g, field, level);
    if (!t) return -1;
    long cnt = (long)t->n[0] * t->n[1] * t->n[2] * t->ch;
    int off_a = field == ORC_SMOKE && g->toff != (R)0;
    for (long i = 0; i < cnt; ++i) {
        double v = in[i];
        if (off_a && i % 4 == 3) v = v - (double)g->toff;
        t->d[i] = rnd(t->fmt, (R)v);
    }
    return cnt;
}

void orc_sample(const orc_grid *g, int field, int level, const float *uvw, double *out) {
    otex *t = field_tex((orc_grid *)g, field, level);
    R u[3] = {(R)uvw[0], (R)uvw[1], (R)uvw[2]};
    R o[4] = {0, 0, 0, 0};
    if (!t) return;
    tex_sample(t, g->dim, u, o);
    for (int c = 0; c < t->ch; ++c) out[c] = (double)o[c];
}

/* ------------------------------------------------------------------ */
/* emitters: update_emitters_status.comp 3D:49-60 / 2D:47-58,          */
/*           handle_emitters.comp 3D:49-76 / 2D:47-74                  */
/* ------------------------------------------------------------------ */

void orc_update_emitters_status(orc_grid *g) {
    for (int e = 0; e < g->ne; ++e) {
        int *s = g->em + e * 12 + 8; /* x delay, y last emit, z flag, w falloff */
        if (s[2] == 1 && s[0] != 0) {
            s[2] = 0;
            s[1] = (int)g->time;
            continue;
        }
        if (g->time - (float)s[1] >= (float)s[0]) s[2] = 1;
    }
}

static R dist_cell(orc_grid *g, int i, int j, int k, const int *pos) {
    R dx = (R)i - (R)pos[0];
    R dy = (R)j - (R)pos[1];
    R dd = dx * dx + dy * dy;
    if (g->dim == 3) {
        R dz = (R)k - (R)pos[2];
        dd = dd + dz * dz;
    }
    return (R)sqrt((double)dd);
}

/* GLSL smoothstep(e0, e1, x) = t*t*(3-2t), t = clamp((x-e0)/(e1-e0), 0, 1). */
static R smoothstep_r(R e0, R e1, R x) {
    R t = (x - e0) / (e1 - e0);
    if (t < (R)0) t = (R)0;
    if (t > (R)1) t = (R)1;
    return t * t * ((R)3 - (R)2 * t);
}

void orc_handle_emitters(orc_grid *g) {
    otex *rd = &g->sm[g->smi];
    otex *wr = &g->sm[1 - g->smi];
    ORC_PARALLEL_KJ
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                R s[4];
                for (int c = 0; c < 4; ++c) s[c] = ld(rd, i, j, k, c);
                for (int e = 0; e < g->ne; ++e) {
                    const int *E = g->em + e * 12;
                    if (!(E[10] > 0)) continue;
                    R dist = dist_cell(g, i, j, k, E);
                    if (dist < (R)E[3]) {
                        R fall = E[11] > 0 ? smoothstep_r((R)E[3], (R)0, dist) : (R)1;
                        for (int c = 0; c < 3; ++c) s[c] = s[c] + (R)E[4 + c] * fall * DT;
                        s[3] = s[3] + (R)E[7] * fall * DT;
                    }
                }
                for (int c = 0; c < 4; ++c) st(wr, i, j, k, c, s[c]);
            }
    g->smi = 1 - g->smi; /* grid.c:86 swap_smoke_buffer */
}

/* Extension (no reference counterpart): per-emitter upward inflow.  After
 * emission, every vy face (i,j,k) whose index satisfies the emitter's own
 * cell test distance((i,j,k), pos) < r is set to the inflow value in the
 * current vy read texture. */
void orc_apply_inflow(orc_grid *g) {
    otex *vy = &g->vy[g->vyi];
    for (int e = 0; e < g->ne; ++e) {
        const int *E = g->em + e * 12;
        if (!(E[10] > 0) || g->inflow[e] == (R)0) continue;
        ORC_PARALLEL_KJ
        for (int k = 0; k < vy->n[2]; ++k)
            for (int j = 0; j < vy->n[1]; ++j)
                for (int i = 0; i < vy->n[0]; ++i)
                    if (dist_cell(g, i, j, k, E) < (R)E[3]) st(vy, i, j, k, 0, g->inflow[e]);
    }
}

/* ------------------------------------------------------------------ */
/* advection: advect_v{x,y,z}.comp 3D:33-78 / 2D:31-75,                */
/*            advect_smoke.comp 3D:33-72 / 2D:31-68                    */
/* ------------------------------------------------------------------ */

/* is_solid with OOB = solid (advect_vx.comp:48-52) */
static R solid_oob1(orc_grid *g, int i, int j, int k) {
    if (!inb(&g->s[0], i, j, k)) return (R)1;
    return ld(&g->s[0], i, j, k, 3);
}
/* is_solid with OOB = fluid (apply_buoyancy.comp:33-37, update_vx.comp:33-37) */
static R solid_oob0(orc_grid *g, int i, int j, int k) {
    if (!inb(&g->s[0], i, j, k)) return (R)0;
    return ld(&g->s[0], i, j, k, 3);
}

static void interp_velocity(orc_grid *g, const R *world, R *v) {
    R gs[3] = {0, 0, 0};
    for (int d = 0; d < g->dim; ++d) gs[d] = world[d] / g->h;
    for (int c = 0; c < g->dim; ++c) {
        R uv[3] = {0, 0, 0};
        for (int d = 0; d < g->dim; ++d)
            uv[d] = (gs[d] + (d == c ? (R)0.5 : (R)0)) / ((R)g->n[d] + (d == c ? (R)1 : (R)0));
        tex_sample(vread(g, c), g->dim, uv, &v[c]);
    }
}

void orc_advect_velocities(orc_grid *g) {
    for (int c = 0; c < g->dim; ++c) {
        otex *w = vwrite(g, c);
        otex *own = vread(g, c);
        int e[3] = {c == 0, c == 1, c == 2};
        ORC_PARALLEL_KJ
        for (int k = 0; k < w->n[2]; ++k)
            for (int j = 0; j < w->n[1]; ++j)
                for (int i = 0; i < w->n[0]; ++i) {
                    int id[3] = {i, j, k};
                    if (id[c] == 0 && WIND(c) != (R)0) {
                        st(w, i, j, k, 0, WIND(c));
                        continue;
                    }
                    if (solid_oob1(g, i, j, k) >= (R)1 ||
                        solid_oob1(g, i - e[0], j - e[1], k - e[2]) >= (R)1) {
                        st(w, i, j, k, 0, (R)0);
                        continue;
                    }
                    R world[3] = {0, 0, 0}, v[3] = {0, 0, 0}, uv[3] = {0, 0, 0}, out;
                    for (int d = 0; d < g->dim; ++d)
                        world[d] = (d == c ? (R)id[d] : (R)id[d] + (R)0.5) * g->h;
                    interp_velocity(g, world, v);
                    for (int d = 0; d < g->dim; ++d) {
                        R from = world[d] - v[d] * DT;
                        R gs = from / g->h;
                        uv[d] = (gs + (d == c ? (R)0.5 : (R)0)) / ((R)g->n[d] + (d == c ? (R)1 : (R)0));
                    }
                    tex_sample(own, g->dim, uv, &out);
                    st(w, i, j, k, 0, out);
                }
    }
    swap_all_v(g); /* grid.c:75 */
}

void orc_advect_smoke(orc_grid *g) {
    otex *rd = &g->sm[g->smi];
    otex *wr = &g->sm[1 - g->smi];
    ORC_PARALLEL_KJ
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                int id[3] = {i, j, k};
                if (solid_oob1(g, i, j, k) >= (R)1) {
                    st(wr, i, j, k, 0, 0);
                    st(wr, i, j, k, 1, 0);
                    st(wr, i, j, k, 2, 0);
                    st(wr, i, j, k, 3, TAMB);
                    continue;
                }
                R world[3] = {0, 0, 0}, v[3] = {0, 0, 0}, uv[3] = {0, 0, 0}, o[4];
                for (int d = 0; d < g->dim; ++d) world[d] = ((R)id[d] + (R)0.5) * g->h;
                interp_velocity(g, world, v);
                for (int d = 0; d < g->dim; ++d) {
                    R from = world[d] - v[d] * DT;
                    R gs = from / g->h;
                    uv[d] = gs / (R)g->n[d];
                }
                tex_sample(rd, g->dim, uv, o);
                for (int c = 0; c < 4; ++c) st(wr, i, j, k, c, o[c]);
            }
    g->smi = 1 - g->smi; /* grid.c:103 */
}

/* ------------------------------------------------------------------ */
/* buoyancy: apply_buoyancy.comp 3D:39-52 / 2D:37-50                   */
/* ------------------------------------------------------------------ */

void orc_apply_buoyancy(orc_grid *g) {
    otex *rd = vread(g, 1), *wr = vwrite(g, 1), *sm = &g->sm[g->smi];
    ORC_PARALLEL_KJ
    for (int k = 0; k < wr->n[2]; ++k)
        for (int j = 0; j < wr->n[1]; ++j)
            for (int i = 0; i < wr->n[0]; ++i) {
                if (solid_oob0(g, i, j, k) >= (R)1 || solid_oob0(g, i, j - 1, k) >= (R)1) continue;
                R uv[3] = {((R)i + (R)0.5) / (R)g->n[0], ((R)j + (R)0) / (R)g->n[1],
                           ((R)k + (R)0.5) / (R)g->n[2]};
                R s[4];
                tex_sample(sm, g->dim, uv, s);
                R total = s[0] + s[1] + s[2];
                R b = -ALPHA * total + BETA * (s[3] - TAMB);
                R nv = ld(rd, i, j, k, 0) + b * DT;
                st(wr, i, j, k, 0, nv);
            }
    swap_v(g, 1); /* grid.c:114 */
}

/* ------------------------------------------------------------------ */
/* vorticity: calculate_vorticity.comp 3D:35-84 / 2D:33-69,            */
/*            apply_vorticity_{x,y,z}.comp 3D:33-63 / 2D:33-60         */
/* ------------------------------------------------------------------ */

/* get_v* of calculate_vorticity.comp:40-53: zero outside, texelFetch
 * outside the texture (the `>` checks let index N+1 through) is zero too. */
static R vget(orc_grid *g, int c, int i, int j, int k) {
    return ld(vread(g, c), i, j, k, 0);
}
static R cc(orc_grid *g, int c, int i, int j, int k) {
    int e[3] = {c == 0, c == 1, c == 2};
    return (vget(g, c, i, j, k) + vget(g, c, i + e[0], j + e[1], k + e[2])) * (R)0.5;
}

void orc_calculate_vorticity(orc_grid *g) {
    otex *w = &g->vort;
    R h2 = (R)2 * g->h;
    ORC_PARALLEL_KJ
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                if (solid_oob1(g, i, j, k) >= (R)1) {
                    for (int c = 0; c < w->ch; ++c) st(w, i, j, k, c, (R)0);
                    continue;
                }
                if (g->dim == 3) {
                    R x1 = cc(g, 2, i, j + 1, k) - cc(g, 2, i, j - 1, k);
                    R x2 = cc(g, 1, i, j, k + 1) - cc(g, 1, i, j, k - 1);
                    R ox = (x1 - x2) / h2;
                    R y1 = cc(g, 0, i, j, k + 1) - cc(g, 0, i, j, k - 1);
                    R y2 = cc(g, 2, i + 1, j, k) - cc(g, 2, i - 1, j, k);
                    R oy = (y1 - y2) / h2;
                    R z1 = cc(g, 1, i + 1, j, k) - cc(g, 1, i - 1, j, k);
                    R z2 = cc(g, 0, i, j + 1, k) - cc(g, 0, i, j - 1, k);
                    R oz = (z1 - z2) / h2;
                    st(w, i, j, k, 0, ox);
                    st(w, i, j, k, 1, oy);
                    st(w, i, j, k, 2, oz);
                    st(w, i, j, k, 3, (R)0);
                } else {
                    R z1 = cc(g, 1, i + 1, j, k) - cc(g, 1, i - 1, j, k);
                    R z2 = cc(g, 0, i, j + 1, k) - cc(g, 0, i, j - 1, k);
                    st(w, i, j, k, 0, (z1 - z2) / h2);
                }
            }
}

static void vort_at(orc_grid *g, int i, int j, int k, R *o) {
    o[0] = o[1] = o[2] = (R)0;
    if (!inb(&g->vort, i, j, k)) return;
    if (g->dim == 3) {
        for (int c = 0; c < 3; ++c) o[c] = ld(&g->vort, i, j, k, c);
    } else {
        o[0] = ld(&g->vort, i, j, k, 0);
    }
}
static R vmag(orc_grid *g, int i, int j, int k) {
    R o[3];
    vort_at(g, i, j, k, o);
    if (g->dim == 3) return (R)sqrt((double)(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
    return (R)fabs((double)o[0]);
}

void orc_apply_vorticity(orc_grid *g) {
    R h2 = (R)2 * g->h;
    for (int c = 0; c < g->dim; ++c) {
        otex *rd = vread(g, c), *wr = vwrite(g, c);
        ORC_PARALLEL_KJ
        for (int k = 0; k < wr->n[2]; ++k)
            for (int j = 0; j < wr->n[1]; ++j)
                for (int i = 0; i < wr->n[0]; ++i) {
                    R dx = vmag(g, i + 1, j, k) - vmag(g, i - 1, j, k);
                    R dy = vmag(g, i, j + 1, k) - vmag(g, i, j - 1, k);
                    R f;
                    R om[3];
                    vort_at(g, i, j, k, om);
                    if (g->dim == 3) {
                        R dz = vmag(g, i, j, k + 1) - vmag(g, i, j, k - 1);
                        R gx = dx / h2, gy = dy / h2, gz = dz / h2;
                        R len = (R)sqrt((double)(gx * gx + gy * gy + gz * gz)) + (R)1e-20f;
                        R nx = gx / len, ny = gy / len, nz = gz / len;
                        if (c == 0) f = ny * om[2] - nz * om[1];
                        else if (c == 1) f = nz * om[0] - nx * om[2];
                        else f = nx * om[1] - ny * om[0];
                    } else {
                        R gx = dx / h2, gy = dy / h2;
                        R len = (R)sqrt((double)(gx * gx + gy * gy)) + (R)1e-20f;
                        R nx = gx / len, ny = gy / len;
                        /* 2D apply_vorticity_x.comp:54 F_x = -n.y*w, _y.comp:54 F_y = -n.x*w */
                        f = c == 0 ? -ny * om[0] : -nx * om[0];
                    }
                    R cur = ld(rd, i, j, k, 0);
                    R nv = cur + (f * (R)0.5 * EPS * g->h * DT) / RHO;
                    st(wr, i, j, k, 0, nv);
                }
    }
    swap_all_v(g); /* grid.c:144 */
}

/* ------------------------------------------------------------------ */
/* pressure solve: pressure_solve/{smooth,compute_residual,            */
/* restrict_residual,restrict_solid,prolong_pressure}.comp, grid.c:173-266 */
/* ------------------------------------------------------------------ */

/* solid_data() of the pressure kernels: OOB returns (0,0,0,0) in
 * smooth.comp:50-54 and (0,0,0,1) in compute_residual.comp:47-51. */
static void solid_data(orc_grid *g, int i, int j, int k, int oob_solid, R *o) {
    otex *s = &g->s[0];
    if (!inb(s, i, j, k)) {
        o[0] = o[1] = o[2] = (R)0;
        o[3] = oob_solid ? (R)1 : (R)0;
        return;
    }
    for (int c = 0; c < 4; ++c) o[c] = ld(s, i, j, k, c);
}

/* Level-0 face divergence sum of smooth.comp 3D:105-113 / 2D:97-103 and
 * compute_residual.comp 3D:116-123: a face whose neighbour cell is solid
 * (a >= 1) takes that cell's solid velocity, otherwise the face velocity. */
static R level0_dv(orc_grid *g, int i, int j, int k, int oob_solid) {
    R sr[4], sl[4], st_[4], sb[4], sf[4] = {0, 0, 0, 0}, sk[4] = {0, 0, 0, 0};
    solid_data(g, i + 1, j, k, oob_solid, sr);
    solid_data(g, i - 1, j, k, oob_solid, sl);
    solid_data(g, i, j + 1, k, oob_solid, st_);
    solid_data(g, i, j - 1, k, oob_solid, sb);
    if (g->dim == 3) {
        solid_data(g, i, j, k + 1, oob_solid, sf);
        solid_data(g, i, j, k - 1, oob_solid, sk);
    }
    R vtop = st_[3] >= (R)1 ? st_[1] : vget(g, 1, i, j + 1, k);
    R vbot = sb[3] >= (R)1 ? sb[1] : vget(g, 1, i, j, k);
    R vr = sr[3] >= (R)1 ? sr[0] : vget(g, 0, i + 1, j, k);
    R vl = sl[3] >= (R)1 ? sl[0] : vget(g, 0, i, j, k);
    R dv = vr - vl + vtop - vbot;
    if (g->dim == 3) {
        R vf = sf[3] >= (R)1 ? sf[2] : vget(g, 2, i, j, k + 1);
        R vb = sk[3] >= (R)1 ? sk[2] : vget(g, 2, i, j, k);
        dv = dv + vf - vb;
    }
    return dv;
}

/* smooth.comp 3D:79-135 / 2D:74-122: one red-black half sweep. */
void orc_smooth_sweep(orc_grid *g, int l, int parity) {
    otex *p = &g->p[l], *s = &g->s[l], *r = &g->r[l];
    R hl = g->h * (R)(1 << l); /* grid.c:341 */
    ORC_PARALLEL_KJ
    for (int k = 0; k < p->n[2]; ++k)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) {
                if ((i + j + k) % 2 != parity) continue;
                if (ld(s, i, j, k, 3) >= (R)1) continue; /* OOB=0 impossible here */
                /* solid_data(): OOB = (0,0,0,0) -> f = 1, pressure 0.  The
                 * correction solve of the refined mode treats OOB as solid
                 * (f = 0), the convention of compute_residual.comp. */
#define AX(ii, jj, kk) (g->corr && !inb(s, ii, jj, kk) ? (R)1 : ld(s, ii, jj, kk, 3))
                R ft = (R)1 - AX(i, j + 1, k);
                R fb = (R)1 - AX(i, j - 1, k);
                R fr = (R)1 - AX(i + 1, j, k);
                R fl = (R)1 - AX(i - 1, j, k);
                R fsum;
                R ff = 0, fk = 0;
                if (g->dim == 3) {
                    ff = (R)1 - AX(i, j, k + 1);
                    fk = (R)1 - AX(i, j, k - 1);
#undef AX
                    fsum = ft + fb + fr + fl + ff + fk;
                } else {
                    fsum = ft + fb + fr + fl;
                }
                if (fsum == (R)0) continue;
                int div0 = l == 0 && !g->corr;
                R dv = div0 ? level0_dv(g, i, j, k, 0) : ld(l == 0 ? &g->rc : r, i, j, k, 0);
                R pt = ld(p, i, j + 1, k, 0) * ft;
                R pb = ld(p, i, j - 1, k, 0) * fb;
                R pr = ld(p, i + 1, j, k, 0) * fr;
                R pl = ld(p, i - 1, j, k, 0) * fl;
                R psum = pt + pb + pr + pl;
                if (g->dim == 3) {
                    R pf = ld(p, i, j, k + 1, 0) * ff;
                    R pk = ld(p, i, j, k - 1, 0) * fk;
                    psum = psum + pf + pk;
                }
                R np;
                if (div0) np = (psum - RHO * hl * dv / DT) / fsum;
                else np = (psum - dv * (hl * hl)) / fsum;
                st(p, i, j, k, 0, np);
            }
}

void orc_smooth(orc_grid *g, int l, int iterations) {
    int parity = 0;
    for (int i = 0; i < iterations * 2; ++i) {
        orc_smooth_sweep(g, l, parity);
        parity = 1 - parity;
    }
}

/* compute_residual.comp 3D:77-130 / 2D:72-117 (in place; OOB = solid). */
void orc_compute_residual(orc_grid *g, int l) {
    otex *p = &g->p[l], *s = &g->s[l], *r = &g->r[l];
    R hl = g->h * (R)(1 << l); /* grid.c:341 */
    ORC_PARALLEL_KJ
    for (int k = 0; k < p->n[2]; ++k)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) {
                if (ld(s, i, j, k, 3) >= (R)1) {
                    st(r, i, j, k, 0, (R)0);
                    continue;
                }
                /* solid_data(): OOB = (0,0,0,1) -> f = 0 */
#define AOOB1(ii, jj, kk) (inb(s, ii, jj, kk) ? ld(s, ii, jj, kk, 3) : (R)1)
                R ft = (R)1 - AOOB1(i, j + 1, k);
                R fb = (R)1 - AOOB1(i, j - 1, k);
                R fr = (R)1 - AOOB1(i + 1, j, k);
                R fl = (R)1 - AOOB1(i - 1, j, k);
                R pt = ld(p, i, j + 1, k, 0) * ft;
                R pb = ld(p, i, j - 1, k, 0) * fb;
                R pr = ld(p, i + 1, j, k, 0) * fr;
                R pl = ld(p, i - 1, j, k, 0) * fl;
                R psum, fsum;
                if (g->dim == 3) {
                    R ff = (R)1 - AOOB1(i, j, k + 1);
                    R fk = (R)1 - AOOB1(i, j, k - 1);
                    R pf = ld(p, i, j, k + 1, 0) * ff;
                    R pk = ld(p, i, j, k - 1, 0) * fk;
                    psum = pt + pb + pr + pl + pf + pk;
                    fsum = ft + fb + fr + fl + ff + fk;
                } else {
                    psum = pt + pb + pr + pl;
                    fsum = ft + fb + fr + fl;
                }
#undef AOOB1
                R old = ld(l == 0 && g->corr ? &g->rc : r, i, j, k, 0);
                if (l == 0 && !g->corr) {
                    /* in this kernel an OOB neighbour is solid, velocity 0 */
                    R dv = level0_dv(g, i, j, k, 1);
                    old = RHO * dv / (DT * hl);
                }
                R nr = old - (psum - fsum * ld(p, i, j, k, 0)) / (hl * hl);
                st(r, i, j, k, 0, nr);
            }
}

/* restrict_residual.comp 3D:40-64 / 2D:40-58 and restrict_solid.comp
 * 3D:53-78 / 2D:53-73 (grid.c:243-257). */
void orc_restrict(orc_grid *g, int l) {
    otex *rf = &g->r[l], *sf = &g->s[l], *rc = &g->r[l + 1], *sc = &g->s[l + 1];
    ORC_PARALLEL_KJ
    for (int k = 0; k < rc->n[2]; ++k)
        for (int j = 0; j < rc->n[1]; ++j)
            for (int i = 0; i < rc->n[0]; ++i) {
                int I = 2 * i, J = 2 * j, K = g->dim == 3 ? 2 * k : 0;
                if (g->dim == 3) {
                    R bsum = ld(rf, I, J, K, 0) + ld(rf, I, J, K + 1, 0) + ld(rf, I + 1, J, K, 0) +
                             ld(rf, I + 1, J, K + 1, 0);
                    R tsum = ld(rf, I, J + 1, K, 0) + ld(rf, I, J + 1, K + 1, 0) +
                             ld(rf, I + 1, J + 1, K, 0) + ld(rf, I + 1, J + 1, K + 1, 0);
                    st(rc, i, j, k, 0, (bsum + tsum) / (R)8);
                    R sv[4];
                    for (int c = 0; c < 4; ++c) {
                        R b = ld(sf, I, J, K, c) + ld(sf, I, J, K + 1, c) + ld(sf, I + 1, J, K, c) +
                              ld(sf, I + 1, J, K + 1, c);
                        R t = ld(sf, I, J + 1, K, c) + ld(sf, I, J + 1, K + 1, c) +
                              ld(sf, I + 1, J + 1, K, c) + ld(sf, I + 1, J + 1, K + 1, c);
                        sv[c] = (b + t) / (R)8;
                    }
                    R m = sv[3] >= (R)0.5 ? (R)1 : (R)0;
                    for (int c = 0; c < 4; ++c) st(sc, i, j, k, c, sv[c] * m);
                } else {
                    R sum = ld(rf, I, J, 0, 0) + ld(rf, I + 1, J, 0, 0) + ld(rf, I, J + 1, 0, 0) +
                            ld(rf, I + 1, J + 1, 0, 0);
                    st(rc, i, j, 0, 0, sum / (R)4);
                    R sv[4];
                    for (int c = 0; c < 4; ++c)
                        sv[c] = (ld(sf, I, J, 0, c) + ld(sf, I + 1, J, 0, c) + ld(sf, I, J + 1, 0, c) +
                                 ld(sf, I + 1, J + 1, 0, c)) /
                                (R)4;
                    R m = sv[3] >= (R)0.5 ? (R)1 : (R)0;
                    for (int c = 0; c < 4; ++c) st(sc, i, j, 0, c, sv[c] * m);
                }
            }
}

/* prolong_pressure.comp 3D:39-51 / 2D:39-51: p_l += LINEAR sample of
 * p_{l+1} (CLAMP_TO_EDGE) at (idx+0.5)/size_l. */
void orc_prolong(orc_grid *g, int l) {
    otex *pf = &g->p[l], *pc = &g->p[l + 1];
    ORC_PARALLEL_KJ
    for (int k = 0; k < pf->n[2]; ++k)
        for (int j = 0; j < pf->n[1]; ++j)
            for (int i = 0; i < pf->n[0]; ++i) {
                R uv[3] = {((R)i + (R)0.5) / (R)pf->n[0], ((R)j + (R)0.5) / (R)pf->n[1],
                           ((R)k + (R)0.5) / (R)pf->n[2]};
                R c;
                tex_sample(pc, g->dim, uv, &c);
                st(pf, i, j, k, 0, ld(pf, i, j, k, 0) + c);
            }
}

/* Red-black smoothing of the coarsest level in FP64 (refinement mode):
 * the update of smooth.comp on level l > 0 with the correction solve's OOB
 * convention (solid), p, the right-hand side and the fractions held in
 * double; the result is stored back rounded. */
static void smooth64(orc_grid *g, int l, int iterations) {
    /* a one-level pyramid smooths level 0, whose correction right-hand side
     * is rc (as in orc_smooth_sweep) */
    otex *p = &g->p[l], *s = &g->s[l], *r = l == 0 ? &g->rc : &g->r[l];
    const int nx = p->n[0], ny = p->n[1], nz = p->n[2];
    const size_t n = (size_t)nx * ny * nz;
    double *pd = (double *)malloc(n * sizeof(double));
    if (!pd) return;
    for (size_t c = 0; c < n; ++c) pd[c] = (double)p->d[c];
    const float hlf = (float)(g->h * (R)(1 << l));
    const double hl = (double)hlf;
#define A64(ii, jj, kk) (inb(s, ii, jj, kk) ? (double)ld(s, ii, jj, kk, 3) : 1.0)
#define P64(ii, jj, kk) (inb(s, ii, jj, kk) ? pd[((size_t)(kk) * ny + (jj)) * nx + (ii)] : 0.0)
    int parity = 0;
    for (int it = 0; it < iterations * 2; ++it) {
        ORC_PARALLEL_KJ
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    if ((i + j + k) % 2 != parity) continue;
                    if (ld(s, i, j, k, 3) >= (R)1) continue;
                    double ft = 1.0 - A64(i, j + 1, k), fb = 1.0 - A64(i, j - 1, k);
                    double fr = 1.0 - A64(i + 1, j, k), fl = 1.0 - A64(i - 1, j, k);
                    double ff = 0.0, fk = 0.0, fsum;
                    if (g->dim == 3) {
                        ff = 1.0 - A64(i, j, k + 1);
                        fk = 1.0 - A64(i, j, k - 1);
                        fsum = ft + fb + fr + fl + ff + fk;
                    } else {
                        fsum = ft + fb + fr + fl;
                    }
                    if (fsum == 0.0) continue;
                    double psum = P64(i, j + 1, k) * ft + P64(i, j - 1, k) * fb;
                    psum = psum + P64(i + 1, j, k) * fr;
                    psum = psum + P64(i - 1, j, k) * fl;
                    if (g->dim == 3) {
                        psum = psum + P64(i, j, k + 1) * ff;
                        psum = psum + P64(i, j, k - 1) * fk;
                    }
                    double term = (double)ld(r, i, j, k, 0) * (hl * hl);
                    pd[((size_t)k * ny + j) * nx + i] = (psum - term) / fsum;
                }
        parity = 1 - parity;
    }
#undef A64
#undef P64
    for (size_t c = 0; c < n; ++c) p->d[c] = rnd(p->fmt, (R)pd[c]);
    free(pd);
}

/* grid.c:177-209 */
void orc_v_cycle(orc_grid *g, int l) {
    if (l == g->L - 1) {
        if (g->refine > 0 && g->coarse64) smooth64(g, l, g->coarse);
        else orc_smooth(g, l, g->coarse);
        return;
    }
    orc_smooth(g, l, g->pre);
    orc_compute_residual(g, l);
    orc_restrict(g, l);
    tex_fill(&g->p[l + 1], NULL); /* grid.c:196-200 glClearTexImage */
    orc_v_cycle(g, l + 1);
    orc_prolong(g, l);
    orc_smooth(g, l, g->post);
}

static double *g_brhs; /* right-hand side override (the rollout adjoint's solve) */

/* One FP64 residual pass of the refined solve: optionally p64 += e (the
 * level-0 correction), then r = b - A p64 with b = rho*dv/(dt*h) and the
 * operator of compute_residual.comp (OOB neighbours solid), all in double;
 * the residual rounded to float becomes the correction right-hand side. */
static void refine_residual(orc_grid *g, int add_e, double *rr, double *bb) {
    otex *s = &g->s[0];
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    if (add_e)
        for (size_t c = 0; c < (size_t)nx * ny * nz; ++c) g->p64[c] = g->p64[c] + (double)g->p[0].d[c];
    const double h = (double)g->h;
    const double rho = (double)RHO, dt = (double)DT;
    double r2 = 0.0, b2 = 0.0;
#define A1(ii, jj, kk) (inb(s, ii, jj, kk) ? (double)ld(s, ii, jj, kk, 3) : 1.0)
#define Q(ii, jj, kk) (inb(s, ii, jj, kk) ? g->p64[((size_t)(kk) * ny + (jj)) * nx + (ii)] : 0.0)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double r = 0.0;
                if (!(ld(s, i, j, k, 3) >= (R)1)) {
                    double ft = 1.0 - A1(i, j + 1, k), fb = 1.0 - A1(i, j - 1, k);
                    double fr = 1.0 - A1(i + 1, j, k), fl = 1.0 - A1(i - 1, j, k);
                    double psum = Q(i, j + 1, k) * ft + Q(i, j - 1, k) * fb;
                    psum = psum + Q(i + 1, j, k) * fr;
                    psum = psum + Q(i - 1, j, k) * fl;
                    double fsum;
                    if (g->dim == 3) {
                        double ff = 1.0 - A1(i, j, k + 1), fk = 1.0 - A1(i, j, k - 1);
                        psum = psum + Q(i, j, k + 1) * ff;
                        psum = psum + Q(i, j, k - 1) * fk;
                        fsum = ft + fb + fr + fl + ff + fk;
                    } else {
                        fsum = ft + fb + fr + fl;
                    }
                    double b = g_brhs ? g_brhs[((size_t)k * ny + j) * nx + i]
                                      : rho * (double)level0_dv(g, i, j, k, 0) / (dt * h);
                    r = b - (psum - fsum * Q(i, j, k)) / (h * h);
                    b2 += b * b;
                }
                r2 += r * r;
                st(&g->rc, i, j, k, 0, (R)r);
            }
#undef A1
#undef Q
    *rr = r2;
    *bb = b2;
}

void orc_pressure_solve(orc_grid *g) {
    if (g->refine <= 0) {
        for (int v = 0; v < g->vcycles; ++v) orc_v_cycle(g, 0);
        return;
    }
    double r2, b2;
    refine_residual(g, 0, &r2, &b2);
    g->norms[0] = r2;
    g->corr = 1;
    for (int it = 1; it <= g->refine; ++it) {
        tex_fill(&g->p[0], NULL);
        orc_v_cycle(g, 0);
        refine_residual(g, 1, &r2, &b2);
        g->norms[it] = r2;
    }
    g->corr = 0;
    g->norms[g->refine + 1] = b2;
}

/* ------------------------------------------------------------------ */
/* projection: update_v{x,y,z}.comp 3D:45-63 / 2D:43-61                */
/* ------------------------------------------------------------------ */

void orc_update_velocities(orc_grid *g) {
    otex *s = &g->s[0], *p = &g->p[0];
    R kc = -DT / (RHO * g->h);
    for (int c = 0; c < g->dim; ++c) {
        otex *rd = vread(g, c), *wr = vwrite(g, c);
        int e[3] = {c == 0, c == 1, c == 2};
        ORC_PARALLEL_KJ
        for (int k = 0; k < wr->n[2]; ++k)
            for (int j = 0; j < wr->n[1]; ++j)
                for (int i = 0; i < wr->n[0]; ++i) {
                    if (solid_oob0(g, i, j, k) >= (R)1 ||
                        solid_oob0(g, i - e[0], j - e[1], k - e[2]) >= (R)1) {
                        st(wr, i, j, k, 0, ld(s, i, j, k, c));
                        continue;
                    }
                    R nv;
                    if (g->refine > 0) {
                        /* refined pressure: the gradient in double */
                        const int ii = i - e[0], jj = j - e[1], kk = k - e[2];
                        double pr = inb(p, i, j, k) ? g->p64[((size_t)k * g->n[1] + j) * g->n[0] + i] : 0.0;
                        double pl = inb(p, ii, jj, kk) ? g->p64[((size_t)kk * g->n[1] + jj) * g->n[0] + ii] : 0.0;
                        double kd = -(double)DT / ((double)RHO * (double)g->h);
                        nv = (R)((double)ld(rd, i, j, k, 0) + kd * (pr - pl));
                    } else {
                        R pr = ld(p, i, j, k, 0);
                        R pl = ld(p, i - e[0], j - e[1], k - e[2], 0);
                        nv = ld(rd, i, j, k, 0) + kc * (pr - pl);
                    }
                    st(wr, i, j, k, 0, nv);
                }
    }
    swap_all_v(g); /* grid.c:54 */
}

/* ------------------------------------------------------------------ */
/* obstacles: obstacle_update.comp 3D:51-134 / 2D:49-130,              */
/*            update_solid_map.comp 3D:46-82 / 2D:44-80,               */
/*            init_solid_map.comp 3D:34-47 / 2D:32-44                  */
/* ------------------------------------------------------------------ */

void orc_obstacle_update_step(orc_grid *g) {
    R *in = g->ob[g->obi], *out = g->ob[1 - g->obi];
    int D = g->dim;
    for (int o = 0; o < g->no; ++o) {
        const R *cur = in + o * 8;
        R pos[3] = {cur[0], cur[1], D == 3 ? cur[2] : (R)0};
        R vel[3] = {cur[4], cur[5], D == 3 ? cur[6] : (R)0};
        R r = cur[3], anch = cur[7], mov = (R)1 - anch;
        int anchored = anch == (R)1;
        if (!anchored) {
            R grav[3] = {0, (R)-9.81f, 0};
            for (int d = 0; d < D; ++d) vel[d] = vel[d] + grav[d] * DT * mov;
            for (int d = 0; d < D; ++d) pos[d] = pos[d] + vel[d] * DT * mov;
            const R bounce = (R)0.999f;
            for (int d = 0; d < D; ++d) {
                R mx = (R)g->n[d];
                if (pos[d] - r < (R)0) {
                    pos[d] = (R)0 + r;
                    vel[d] = vel[d] * -bounce;
                }
                if (pos[d] + r > mx) {
                    pos[d] = mx - r;
                    vel[d] = vel[d] * -bounce;
                }
            }
            for (int q = 0; q < g->no; ++q) {
                if (q == o) continue;
                const R *oth = in + q * 8;
                R dd[3] = {0, 0, 0};
                R d2 = 0;
                for (int d = 0; d < D; ++d) dd[d] = pos[d] - oth[d];
                d2 = dd[0] * dd[0] + dd[1] * dd[1];
                if (D == 3) d2 = d2 + dd[2] * dd[2];
                R dist = (R)sqrt((double)d2);
                R mind = r + oth[3];
                if (dist < mind && dist > (R)0.0001f) {
                    R n[3] = {0, 0, 0};
                    for (int d = 0; d < D; ++d) n[d] = dd[d] / dist;
                    for (int d = 0; d < D; ++d) pos[d] = pos[d] + n[d] * (mind - dist) * (R)0.5f * mov;
                    R rel[3] = {0, 0, 0};
                    for (int d = 0; d < D; ++d) rel[d] = vel[d] - oth[4 + d];
                    R vn = rel[0] * n[0] + rel[1] * n[1];
                    if (D == 3) vn = vn + rel[2] * n[2];
                    if (vn < (R)0)
                        for (int d = 0; d < D; ++d) vel[d] = vel[d] - vn * n[d] * mov;
                }
            }
        }
        R *w = out + o * 8;
        w[0] = pos[0];
        w[1] = pos[1];
        w[2] = D == 3 ? pos[2] : (R)0;
        w[3] = r;
        for (int d = 0; d < 3; ++d) w[4 + d] = d < D ? vel[d] * (R)0.999f : (R)0;
        w[7] = anch;
    }
    g->obi = 1 - g->obi; /* grid.c:436-440 */
}

void orc_update_solid_map(orc_grid *g) {
    otex *s = &g->s[0];
    const R *ob = g->ob[g->obi];
    ORC_PARALLEL_KJ
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                int valid = !(i <= 1 || j <= 1 || i >= g->n[0] - 1 || j >= g->n[1] - 1);
                if (g->dim == 3) valid = valid && !(k <= 1 || k >= g->n[2] - 1);
                if (!valid) continue;
                R nv[4] = {0, 0, 0, 0};
                for (int o = 0; o < g->no; ++o) {
                    const R *c = ob + o * 8;
                    R dx = (R)i - c[0], dy = (R)j - c[1];
                    R d2 = dx * dx + dy * dy;
                    if (g->dim == 3) {
                        R dz = (R)k - c[2];
                        d2 = d2 + dz * dz;
                    }
                    R dist = (R)sqrt((double)d2);
                    if (dist < c[3]) {
                        R m = (R)1 - c[7];
                        nv[0] = c[4] * m;
                        nv[1] = c[5] * m;
                        nv[2] = c[6] * m;
                        nv[3] = (R)1;
                    }
                }
                /* Extension: axis-aligned boxes (static, zero velocity),
                 * rasterised after the spheres over the same interior. */
                for (int b = 0; b < g->nb; ++b) {
                    const int *B = g->box + b * 6;
                    int in = i >= B[0] && i < B[3] && j >= B[1] && j < B[4];
                    if (g->dim == 3) in = in && k >= B[2] && k < B[5];
                    if (in) {
                        nv[0] = nv[1] = nv[2] = (R)0;
                        nv[3] = (R)1;
                    }
                }
                for (int c = 0; c < 4; ++c) st(s, i, j, k, c, nv[c]);
            }
}

void orc_init_solid_map(orc_grid *g) {
    otex *s = &g->s[0];
    ORC_PARALLEL_KJ
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                int border = i == 0 || i == g->n[0] - 1 || j == 0 || j == g->n[1] - 1;
                if (g->dim == 3) border = border || k == 0 || k == g->n[2] - 1;
                if (border) {
                    st(s, i, j, k, 0, 0);
                    st(s, i, j, k, 1, 0);
                    st(s, i, j, k, 2, 0);
                    st(s, i, j, k, 3, 1);
                }
            }
}

/* clear_grid3d (3d/grid3d.c:164-184) / clear_grid2d (2d/grid2d.c:146-159):
 * the 2D variant does not clear the pressure pyramid. */
void orc_clear_grid(orc_grid *g) {
    R smoke[4] = {0, 0, 0, g->tamb - g->toff};
    for (int b = 0; b < 2; ++b) {
        tex_fill(&g->vx[b], NULL);
        tex_fill(&g->vy[b], NULL);
        if (g->dim == 3) tex_fill(&g->vz[b], NULL);
        tex_fill(&g->sm[b], smoke);
    }
    tex_fill(&g->vort, NULL);
    if (g->dim == 3) {
        for (int l = 0; l < g->L; ++l) tex_fill(&g->p[l], NULL);
        if (g->p64) memset(g->p64, 0, sizeof(double) * (size_t)g->n[0] * g->n[1] * g->n[2]);
    }
}

/* physics.c:72-101, with the inflow extension after emission. */
void orc_run_physics_step(orc_grid *g, const int *t) {
    static const int all[7] = {1, 1, 1, 1, 1, 1, 1};
    if (!t) t = all;
    /* toggles: 0 advect_smoke, 1 advect_velocities, 2 apply_vorticity,
     * 3 apply_buoyancy, 4 resolve_pressure, 5 update_velocities,
     * 6 handle_emitters */
    if (t[6]) {
        orc_update_emitters_status(g);
        orc_handle_emitters(g);
        orc_apply_inflow(g);
    }
    if (t[1]) orc_advect_velocities(g);
    if (t[3]) orc_apply_buoyancy(g);
    if (t[2]) {
        orc_calculate_vorticity(g);
        orc_apply_vorticity(g);
    }
    if (t[4]) orc_pressure_solve(g);
    if (t[5]) orc_update_velocities(g);
    if (t[0]) orc_advect_smoke(g);
    orc_obstacle_update_step(g);
    orc_update_solid_map(g);
}

/* ================================================================== */
/* Differentiable rollout and its adjoint (2D and 3D): the FP32-field */
/* step of the GPU rollout (smokesim_amd/csrc/smk_adjoint.hip) in the */
/* oracle's arithmetic, every step state kept, the adjoint of every    */
/* stage applied in reverse with plain sequential sums.  Built with    */
/* ORC_REAL=double this is the FP64 C adjoint the GPU one is checked   */
/* against.                                                            */
/* ================================================================== */

typedef struct {
    int i0[3];
    R a[3];
} otri;

/* the coordinate arithmetic of tex_sample() for one axis */
static void tcoord(R uv, int n, int *i0, R *a) {
    R u = uv * (R)n - (R)0.5;
    if (!(u > (R)-4)) u = (R)-4;
    if (u > (R)(n + 4)) u = (R)(n + 4);
    R fl = (R)floor((double)u);
    *i0 = (int)fl;
    *a = u - fl;
}

static R tri_w(const otri *t, int q) {
    R w = (q & 1) ? t->a[0] : (R)1 - t->a[0];
    w *= (q & 2) ? t->a[1] : (R)1 - t->a[1];
    w *= (q & 4) ? t->a[2] : (R)1 - t->a[2];
    return w;
}

/* d(sample)/d(a_d) from the 8 corner values */
static R tri_da(const otri *t, const R *c, int d) {
    R s = 0;
    for (int q = 0; q < 8; ++q) {
        if (q & (1 << d)) continue;
        R w = 1;
        for (int e = 0; e < 3; ++e) {
            if (e == d) continue;
            w *= (q & (1 << e)) ? t->a[e] : (R)1 - t->a[e];
        }
        s += w * (c[q | (1 << d)] - c[q]);
    }
    return s;
}

static int in3(const int *s, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < s[0] && j < s[1] && k < s[2];
}
static size_t at3(const int *s, int i, int j, int k) { return ((size_t)k * s[1] + j) * s[0] + i; }

static void tri_corner(const otri *t, int q, int *ii, int *jj, int *kk) {
    *ii = t->i0[0] + (q & 1);
    *jj = t->i0[1] + ((q >> 1) & 1);
    *kk = t->i0[2] + ((q >> 2) & 1);
}

/* rollout state arrays */
typedef struct {
    int fs[3][3]; /* face sizes */
    int cs[3];    /* cell sizes */
    size_t nf[3], nc;
} olay;

static olay layout(const orc_grid *g) {
    olay L;
    for (int c = 0; c < 3; ++c) {
        for (int d = 0; d < 3; ++d) L.fs[c][d] = g->n[d] + (d == c && c < g->dim);
        L.nf[c] = c < g->dim ? (size_t)L.fs[c][0] * L.fs[c][1] * L.fs[c][2] : 0;
        L.cs[c] = g->n[c];
    }
    L.nc = (size_t)g->n[0] * g->n[1] * g->n[2];
    return L;
}

static void put_state(orc_grid *g, const olay *L, R *const *v, const R *rho) {
    for (int c = 0; c < g->dim; ++c) {
        otex *t = vread(g, c);
        for (size_t i = 0; i < L->nf[c]; ++i) t->d[i] = rnd(t->fmt, v[c][i]);
    }
    otex *s = &g->sm[g->smi];
    for (size_t i = 0; i < L->nc; ++i) {
        s->d[4 * i] = rnd(s->fmt, rho[i]);
        s->d[4 * i + 1] = 0;
        s->d[4 * i + 2] = 0;
        s->d[4 * i + 3] = rnd(s->fmt, g->tamb - g->toff);
    }
}

static void get_v(orc_grid *g, const olay *L, R *const *v) {
    for (int c = 0; c < g->dim; ++c) memcpy(v[c], vread(g, c)->d, L->nf[c] * sizeof(R));
}
static void get_rho(orc_grid *g, const olay *L, R *rho) {
    otex *s = &g->sm[g->smi];
    for (size_t i = 0; i < L->nc; ++i) rho[i] = s->d[4 * i];
}

/* fluid face of the projection: both adjacent cells inside and fluid */
static int fluid_face(orc_grid *g, int c, int i, int j, int k) {
    int e[3] = {c == 0, c == 1, c == 2};
    otex *s = &g->s[0];
    if (!inb(s, i, j, k) || !inb(s, i - e[0], j - e[1], k - e[2])) return 0;
    return ld(s, i, j, k, 3) < (R)1 && ld(s, i - e[0], j - e[1], k - e[2], 3) < (R)1;
}
static int buoy_face(orc_grid *g, int i, int j, int k) {
    return !(solid_oob0(g, i, j, k) >= (R)1 || solid_oob0(g, i, j - 1, k) >= (R)1);
}

/* velocity of component cc at face position (c, id) (c = -1: cell centre)
 * as an explicit trilinear stencil: the coordinates interp_velocity() uses */
static otri interp_tri(orc_grid *g, const int *id, int c, int cc) {
    otri t;
    t.i0[2] = 0;
    t.a[2] = 0;
    for (int d = 0; d < g->dim; ++d) {
        R world = (d == c ? (R)id[d] : (R)id[d] + (R)0.5) * g->h;
        R gs = world / g->h;
        R uv = (gs + (d == cc ? (R)0.5 : (R)0)) / ((R)g->n[d] + (d == cc ? (R)1 : (R)0));
        tcoord(uv, g->n[d] + (d == cc), &t.i0[d], &t.a[d]);
    }
    return t;
}

/* adjoint of the interpolation: vbar_cc += w * ubar_cc */
static void scatter_interp(orc_grid *g, const olay *L, const int *id, int c, const R *ubar, R *const *vb) {
    for (int cc = 0; cc < g->dim; ++cc) {
        if (ubar[cc] == (R)0) continue;
        otri t = interp_tri(g, id, c, cc);
        for (int q = 0; q < 8; ++q) {
            int ii, jj, kk;
            tri_corner(&t, q, &ii, &jj, &kk);
            if (!in3(L->fs[cc], ii, jj, kk)) continue;
            vb[cc][at3(L->fs[cc], ii, jj, kk)] += tri_w(&t, q) * ubar[cc];
        }
    }
}

/* curl of the cell-centred velocities (calculate_vorticity.comp 3D) */
static void vort64(orc_grid *g, const olay *L, R *const *v, R *w) {
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
#define VG(c, i, j, k) (in3(L->fs[c], i, j, k) ? v[c][at3(L->fs[c], i, j, k)] : (R)0)
#define CC(c, i, j, k) ((VG(c, i, j, k) + VG(c, (i) + ((c) == 0), (j) + ((c) == 1), (k) + ((c) == 2))) * (R)0.5)
    const R h2 = 2 * g->h;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = at3(L->cs, i, j, k);
                if (ld(&g->s[0], i, j, k, 3) >= (R)1) {
                    w[3 * c] = w[3 * c + 1] = w[3 * c + 2] = 0;
                    continue;
                }
                if (g->dim == 2) {
                    /* calculate_vorticity.comp 2D: the scalar curl in w[3c] */
                    R z1 = CC(1, i + 1, j, k) - CC(1, i - 1, j, k), z2 = CC(0, i, j + 1, k) - CC(0, i, j - 1, k);
                    w[3 * c] = (z1 - z2) / h2;
                    w[3 * c + 1] = w[3 * c + 2] = 0;
                    continue;
                }
                R x1 = CC(2, i, j + 1, k) - CC(2, i, j - 1, k), x2 = CC(1, i, j, k + 1) - CC(1, i, j, k - 1);
                R y1 = CC(0, i, j, k + 1) - CC(0, i, j, k - 1), y2 = CC(2, i + 1, j, k) - CC(2, i - 1, j, k);
                R z1 = CC(1, i + 1, j, k) - CC(1, i - 1, j, k), z2 = CC(0, i, j + 1, k) - CC(0, i, j - 1, k);
                w[3 * c] = (x1 - x2) / h2;
                w[3 * c + 1] = (y1 - y2) / h2;
                w[3 * c + 2] = (z1 - z2) / h2;
            }
#undef CC
#undef VG
}

/* The projection adjoint's solve A x = rhs: the refined solver with a
 * given right-hand side, from x = 0.  Returns x in g->p64. */
static void solve_rhs(orc_grid *g, const R *rhs, int cycles);


int orc_rollout(orc_grid *g, int steps, const int *lo, const int *sz, const double *src,
                const double *vx0, const double *vy0, const double *vz0, const double *rho0, int cycles,
                const double *rho_bar, double *rho_final, double *src_bar, double *vx_bar, double *vy_bar,
                double *vz_bar) {
    if (steps < 1 || cycles < 1) return -1;
    const int D = g->dim;
    olay L = layout(g);
    const size_t nsrc = (size_t)sz[0] * sz[1] * sz[2];
    const int keep = rho_bar != NULL;
    /* saved states per step: v_in, v2, v4 (3 each), rho1 */
    R **sv = (R **)calloc((size_t)steps * 10, sizeof(R *));
    R *v[3], *rho = (R *)malloc(L.nc * sizeof(R));
    for (int c = 0; c < 3; ++c) v[c] = (R *)malloc((L.nf[c] ? L.nf[c] : 1) * sizeof(R));
    if (!sv || !rho) return -1;
    for (int c = 0; c < D; ++c) {
        const double *h0 = c == 0 ? vx0 : (c == 1 ? vy0 : vz0);
        for (size_t i = 0; i < L.nf[c]; ++i) v[c][i] = (R)h0[i];
    }
    for (size_t i = 0; i < L.nc; ++i) rho[i] = (R)rho0[i];
    int saved_refine = g->refine, saved_c64 = g->coarse64;
    if (orc_set_refinement(g, cycles, 1) != 0) return -1;
    put_state(g, &L, v, rho);
    for (int s = 0; s < steps; ++s) {
        R **S = sv + (size_t)s * 10;
        if (keep) {
            for (int c = 0; c < D; ++c) S[c] = (R *)malloc(L.nf[c] * sizeof(R));
            for (int c = 0; c < D; ++c) S[3 + c] = (R *)malloc(L.nf[c] * sizeof(R));
            for (int c = 0; c < D; ++c) S[6 + c] = (R *)malloc(L.nf[c] * sizeof(R));
            S[9] = (R *)malloc(L.nc * sizeof(R));
            get_v(g, &L, S);
        }
        /* source */
        otex *sm = &g->sm[g->smi];
        for (int k = 0; k < sz[2]; ++k)
            for (int j = 0; j < sz[1]; ++j)
                for (int i = 0; i < sz[0]; ++i) {
                    int I = lo[0] + i, J = lo[1] + j, K = lo[2] + k;
                    R val = ld(sm, I, J, K, 0) + (R)src[((size_t)k * sz[1] + j) * sz[0] + i] * DT;
                    st(sm, I, J, K, 0, val);
                }
        if (keep) get_rho(g, &L, S[9]);
        orc_advect_velocities(g);
        /* buoyancy (temperature at ambient: the beta term is zero); faces
         * next to a solid keep their value */
        {
            otex *vy = vread(g, 1);
            otex *sm2 = &g->sm[g->smi];
            for (int k = 0; k < vy->n[2]; ++k)
                for (int j = 0; j < vy->n[1]; ++j)
                    for (int i = 0; i < vy->n[0]; ++i) {
                        if (!buoy_face(g, i, j, k)) continue;
                        R uv[3] = {((R)i + (R)0.5) / (R)g->n[0], ((R)j + (R)0) / (R)g->n[1],
                                   ((R)k + (R)0.5) / (R)g->n[2]};
                        R o[4];
                        tex_sample(sm2, D, uv, o);
                        R b = -ALPHA * o[0];
                        st(vy, i, j, k, 0, ld(vy, i, j, k, 0) + b * DT);
                    }
        }
        if (keep) get_v(g, &L, S + 3);
        orc_calculate_vorticity(g);
        orc_apply_vorticity(g);
        memset(g->p64, 0, L.nc * sizeof(double));
        orc_pressure_solve(g);
        orc_update_velocities(g);
        if (keep) get_v(g, &L, S + 6);
        orc_advect_smoke(g);
    }
    get_rho(g, &L, rho);
    if (rho_final)
        for (size_t i = 0; i < L.nc; ++i) rho_final[i] = (double)rho[i];
    if (keep) {
        /* adjoint state */
        R *vb[3], *vb2[3], *rb = (R *)malloc(L.nc * sizeof(R)), *rb1 = (R *)malloc(L.nc * sizeof(R));
        R *w = (R *)malloc(3 * L.nc * sizeof(R)), *wb = (R *)malloc(3 * L.nc * sizeof(R));
        R *gb = (R *)malloc(3 * L.nc * sizeof(R)), *ccb = (R *)malloc(3 * L.nc * sizeof(R));
        R *rhs = (R *)malloc(L.nc * sizeof(R));
        R *sb = (R *)calloc(nsrc ? nsrc : 1, sizeof(R));
        for (int c = 0; c < 3; ++c) {
            vb[c] = (R *)calloc(L.nf[c] ? L.nf[c] : 1, sizeof(R));
            vb2[c] = (R *)calloc(L.nf[c] ? L.nf[c] : 1, sizeof(R));
        }
        for (size_t i = 0; i < L.nc; ++i) rb[i] = (R)rho_bar[i];
        const R hh = g->h, dt = DT;
        const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
        for (int s = steps - 1; s >= 0; --s) {
            R **S = sv + (size_t)s * 10;
            R *vin[3] = {S[0], S[1], S[2]}, *v2[3] = {S[3], S[4], S[5]}, *v4[3] = {S[6], S[7], S[8]};
            R *rho1 = S[9];
            /* (a) density advection */
            memset(rb1, 0, L.nc * sizeof(R));
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        size_t cell = at3(L.cs, i, j, k);
                        R lam = rb[cell];
                        if (ld(&g->s[0], i, j, k, 3) >= (R)1 || lam == (R)0) continue;
                        int id[3] = {i, j, k};
                        R u[3] = {0, 0, 0};
                        for (int cc = 0; cc < D; ++cc) {
                            otri t = interp_tri(g, id, -1, cc);
                            R s2 = 0;
                            for (int q = 0; q < 8; ++q) {
                                int ii, jj, kk;
                                tri_corner(&t, q, &ii, &jj, &kk);
                                if (in3(L.fs[cc], ii, jj, kk)) s2 += tri_w(&t, q) * v4[cc][at3(L.fs[cc], ii, jj, kk)];
                            }
                            u[cc] = s2;
                        }
                        otri t = {{0, 0, 0}, {0, 0, 0}};
                        for (int d = 0; d < D; ++d) {
                            R world = ((R)id[d] + (R)0.5) * hh;
                            R from = world - u[d] * dt;
                            tcoord(from / hh / (R)g->n[d], g->n[d], &t.i0[d], &t.a[d]);
                        }
                        R cv[8];
                        for (int q = 0; q < 8; ++q) {
                            int ii, jj, kk;
                            tri_corner(&t, q, &ii, &jj, &kk);
                            int ok = in3(L.cs, ii, jj, kk);
                            cv[q] = ok ? rho1[at3(L.cs, ii, jj, kk)] : (R)0;
                            if (ok) rb1[at3(L.cs, ii, jj, kk)] += tri_w(&t, q) * lam;
                        }
                        R ub[3] = {0, 0, 0};
                        for (int d = 0; d < D; ++d) ub[d] = lam * tri_da(&t, cv, d) * (-dt / hh);
                        scatter_interp(g, &L, id, -1, ub, vb);
                    }
            /* (b) projection: rhs_i = kc sum_c (vb_c(i) - vb_c(i+e_c)) on fluid faces */
            {
                const double kc = -(double)dt / ((double)RHO * (double)hh);
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i) {
                            double acc = 0;
                            if (ld(&g->s[0], i, j, k, 3) < (R)1)
                                for (int c = 0; c < D; ++c) {
                                    int e[3] = {c == 0, c == 1, c == 2};
                                    if (fluid_face(g, c, i, j, k)) acc += (double)vb[c][at3(L.fs[c], i, j, k)];
                                    if (fluid_face(g, c, i + e[0], j + e[1], k + e[2]))
                                        acc -= (double)vb[c][at3(L.fs[c], i + e[0], j + e[1], k + e[2])];
                                }
                            rhs[at3(L.cs, i, j, k)] = (R)(kc * acc);
                        }
                solve_rhs(g, rhs, cycles);
                const double beta = (double)RHO / ((double)dt * (double)hh);
                for (int c = 0; c < D; ++c) {
                    int e[3] = {c == 0, c == 1, c == 2};
                    for (int k = 0; k < L.fs[c][2]; ++k)
                        for (int j = 0; j < L.fs[c][1]; ++j)
                            for (int i = 0; i < L.fs[c][0]; ++i) {
                                size_t f = at3(L.fs[c], i, j, k);
                                R o = 0;
                                if (fluid_face(g, c, i, j, k)) {
                                    double xm = g->p64[at3(L.cs, i - e[0], j - e[1], k - e[2])];
                                    double xp = g->p64[at3(L.cs, i, j, k)];
                                    o = (R)((double)vb[c][f] + beta * (xm - xp));
                                }
                                vb2[c][f] = o;
                            }
                }
            }
            /* (c) confinement: vb2 = dL/dv3 -> vb = dL/dv2 */
            vort64(g, &L, v2, w);
            if (D == 2) {
                /* 2D: F = (-n_y w, -n_x w) (apply_vorticity_{x,y}.comp 2D:54),
                 * n = g/(|g| + 1e-20), g = grad |w| / (2h) */
                const R K = (R)0.5 * EPS * hh * dt / RHO;
                const R h2 = 2 * hh;
#define WM2(i, j) (in3(L.cs, i, j, 0) ? (R)fabs((double)w[3 * at3(L.cs, i, j, 0)]) : (R)0)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        size_t cell = at3(L.cs, i, j, 0);
                        R Fb0 = K * vb2[0][at3(L.fs[0], i, j, 0)], Fb1 = K * vb2[1][at3(L.fs[1], i, j, 0)];
                        R gx = (WM2(i + 1, j) - WM2(i - 1, j)) / h2;
                        R gy = (WM2(i, j + 1) - WM2(i, j - 1)) / h2;
                        R gl = (R)sqrt((double)(gx * gx + gy * gy));
                        R len = gl + (R)1e-20;
                        R nxv = gx / len, nyv = gy / len;
                        R om = w[3 * cell];
                        wb[3 * cell] = -nyv * Fb0 - nxv * Fb1;
                        R nbx = -om * Fb1, nby = -om * Fb0;
                        if (gl > 0) {
                            R gn = (gx * nbx + gy * nby) / (gl * len * len);
                            gb[3 * cell] = nbx / len - gx * gn;
                            gb[3 * cell + 1] = nby / len - gy * gn;
                        } else {
                            gb[3 * cell] = nbx / len;
                            gb[3 * cell + 1] = nby / len;
                        }
                    }
#undef WM2
#define GB2(i, j, d) (in3(L.cs, i, j, 0) ? gb[3 * at3(L.cs, i, j, 0) + (d)] : (R)0)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        size_t cell = at3(L.cs, i, j, 0);
                        R mb = ((GB2(i - 1, j, 0) - GB2(i + 1, j, 0)) + (GB2(i, j - 1, 1) - GB2(i, j + 1, 1))) / h2;
                        R om = w[3 * cell];
                        R m = (R)fabs((double)om);
                        if (m > 0 && mb != 0) wb[3 * cell] += mb * om / m;
                    }
#undef GB2
#define WB2(i, j) ((in3(L.cs, i, j, 0) && ld(&g->s[0], i, j, 0, 3) < (R)1) ? wb[3 * at3(L.cs, i, j, 0)] : (R)0)
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        size_t cell = at3(L.cs, i, j, 0);
                        ccb[3 * cell] = (-WB2(i, j - 1) + WB2(i, j + 1)) / h2;
                        ccb[3 * cell + 1] = (WB2(i - 1, j) - WB2(i + 1, j)) / h2;
                    }
#undef WB2
                for (int c = 0; c < 2; ++c) {
                    int e[2] = {c == 0, c == 1};
                    for (int j = 0; j < L.fs[c][1]; ++j)
                        for (int i = 0; i < L.fs[c][0]; ++i) {
                            R ca = in3(L.cs, i - e[0], j - e[1], 0) ? ccb[3 * at3(L.cs, i - e[0], j - e[1], 0) + c] : (R)0;
                            R cb = in3(L.cs, i, j, 0) ? ccb[3 * at3(L.cs, i, j, 0) + c] : (R)0;
                            size_t f = at3(L.fs[c], i, j, 0);
                            vb[c][f] = vb2[c][f] + (R)0.5 * (ca + cb);
                        }
                }
            } else {
                const R K = (R)0.5 * EPS * hh * dt / RHO;
                const R h2 = 2 * hh;
#define WM(i, j, k) (in3(L.cs, i, j, k) ? (R)sqrt((double)(w[3 * at3(L.cs, i, j, k)] * w[3 * at3(L.cs, i, j, k)] + \
                                              w[3 * at3(L.cs, i, j, k) + 1] * w[3 * at3(L.cs, i, j, k) + 1] + \
                                              w[3 * at3(L.cs, i, j, k) + 2] * w[3 * at3(L.cs, i, j, k) + 2])) : (R)0)
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i) {
                            size_t cell = at3(L.cs, i, j, k);
                            R Fb[3];
                            for (int c = 0; c < 3; ++c) Fb[c] = K * vb2[c][at3(L.fs[c], i, j, k)];
                            R gx = (WM(i + 1, j, k) - WM(i - 1, j, k)) / h2;
                            R gy = (WM(i, j + 1, k) - WM(i, j - 1, k)) / h2;
                            R gz = (WM(i, j, k + 1) - WM(i, j, k - 1)) / h2;
                            R gl = (R)sqrt((double)(gx * gx + gy * gy + gz * gz));
                            R len = gl + (R)1e-20;
                            R nxv = gx / len, nyv = gy / len, nzv = gz / len;
                            R *om = w + 3 * cell;
                            wb[3 * cell] = Fb[1] * nzv - Fb[2] * nyv;
                            wb[3 * cell + 1] = Fb[2] * nxv - Fb[0] * nzv;
                            wb[3 * cell + 2] = Fb[0] * nyv - Fb[1] * nxv;
                            R nbx = om[1] * Fb[2] - om[2] * Fb[1];
                            R nby = om[2] * Fb[0] - om[0] * Fb[2];
                            R nbz = om[0] * Fb[1] - om[1] * Fb[0];
                            if (gl > 0) {
                                R gn = (gx * nbx + gy * nby + gz * nbz) / (gl * len * len);
                                gb[3 * cell] = nbx / len - gx * gn;
                                gb[3 * cell + 1] = nby / len - gy * gn;
                                gb[3 * cell + 2] = nbz / len - gz * gn;
                            } else {
                                gb[3 * cell] = nbx / len;
                                gb[3 * cell + 1] = nby / len;
                                gb[3 * cell + 2] = nbz / len;
                            }
                        }
#undef WM
#define GB(i, j, k, d) (in3(L.cs, i, j, k) ? gb[3 * at3(L.cs, i, j, k) + (d)] : (R)0)
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i) {
                            size_t cell = at3(L.cs, i, j, k);
                            R mb = ((GB(i - 1, j, k, 0) - GB(i + 1, j, k, 0)) + (GB(i, j - 1, k, 1) - GB(i, j + 1, k, 1)) +
                                    (GB(i, j, k - 1, 2) - GB(i, j, k + 1, 2))) / h2;
                            R *om = w + 3 * cell;
                            R m = (R)sqrt((double)(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]));
                            if (m > 0 && mb != 0)
                                for (int d = 0; d < 3; ++d) wb[3 * cell + d] += mb * om[d] / m;
                        }
#undef GB
#define WB(i, j, k, d) ((in3(L.cs, i, j, k) && ld(&g->s[0], i, j, k, 3) < (R)1) ? wb[3 * at3(L.cs, i, j, k) + (d)] : (R)0)
                for (int k = 0; k < nz; ++k)
                    for (int j = 0; j < ny; ++j)
                        for (int i = 0; i < nx; ++i) {
                            size_t cell = at3(L.cs, i, j, k);
                            ccb[3 * cell] = (WB(i, j, k - 1, 1) - WB(i, j, k + 1, 1) - WB(i, j - 1, k, 2) + WB(i, j + 1, k, 2)) / h2;
                            ccb[3 * cell + 1] = (-WB(i, j, k - 1, 0) + WB(i, j, k + 1, 0) + WB(i - 1, j, k, 2) - WB(i + 1, j, k, 2)) / h2;
                            ccb[3 * cell + 2] = (WB(i, j - 1, k, 0) - WB(i, j + 1, k, 0) - WB(i - 1, j, k, 1) + WB(i + 1, j, k, 1)) / h2;
                        }
#undef WB
                for (int c = 0; c < 3; ++c) {
                    int e[3] = {c == 0, c == 1, c == 2};
                    for (int k = 0; k < L.fs[c][2]; ++k)
                        for (int j = 0; j < L.fs[c][1]; ++j)
                            for (int i = 0; i < L.fs[c][0]; ++i) {
                                R ca = in3(L.cs, i - e[0], j - e[1], k - e[2]) ? ccb[3 * at3(L.cs, i - e[0], j - e[1], k - e[2]) + c] : (R)0;
                                R cb = in3(L.cs, i, j, k) ? ccb[3 * at3(L.cs, i, j, k) + c] : (R)0;
                                size_t f = at3(L.fs[c], i, j, k);
                                vb[c][f] = vb2[c][f] + (R)0.5 * (ca + cb);
                            }
                }
            }
            /* (d) buoyancy: rbar1 += -alpha*dt W^T vb_y */
            for (int k = 0; k < L.fs[1][2]; ++k)
                for (int j = 0; j < L.fs[1][1]; ++j)
                    for (int i = 0; i < L.fs[1][0]; ++i) {
                        R lam = vb[1][at3(L.fs[1], i, j, k)];
                        if (lam == (R)0 || !buoy_face(g, i, j, k)) continue;
                        R uv[3] = {((R)i + (R)0.5) / (R)nx, ((R)j + (R)0) / (R)ny, ((R)k + (R)0.5) / (R)nz};
                        otri t = {{0, 0, 0}, {0, 0, 0}};
                        for (int d = 0; d < D; ++d) tcoord(uv[d], g->n[d], &t.i0[d], &t.a[d]);
                        for (int q = 0; q < 8; ++q) {
                            int ii, jj, kk;
                            tri_corner(&t, q, &ii, &jj, &kk);
                            if (in3(L.cs, ii, jj, kk)) rb1[at3(L.cs, ii, jj, kk)] += tri_w(&t, q) * (-ALPHA * dt * lam);
                        }
                    }
            /* (e) velocity advection: vb = dL/dv1 -> vb2 = dL/dv_in */
            for (int c = 0; c < 3; ++c) memset(vb2[c], 0, L.nf[c] * sizeof(R));
            for (int c = 0; c < D; ++c) {
                int e[3] = {c == 0, c == 1, c == 2};
                for (int k = 0; k < L.fs[c][2]; ++k)
                    for (int j = 0; j < L.fs[c][1]; ++j)
                        for (int i = 0; i < L.fs[c][0]; ++i) {
                            int id[3] = {i, j, k};
                            R lam = vb[c][at3(L.fs[c], i, j, k)];
                            if (lam == (R)0) continue;
                            if (id[c] == 0 && WIND(c) != (R)0) continue;
                            if (solid_oob1(g, i, j, k) >= (R)1 || solid_oob1(g, i - e[0], j - e[1], k - e[2]) >= (R)1)
                                continue;
                            R u[3] = {0, 0, 0};
                            for (int cc = 0; cc < D; ++cc) {
                                otri t = interp_tri(g, id, c, cc);
                                R s2 = 0;
                                for (int q = 0; q < 8; ++q) {
                                    int ii, jj, kk;
                                    tri_corner(&t, q, &ii, &jj, &kk);
                                    if (in3(L.fs[cc], ii, jj, kk)) s2 += tri_w(&t, q) * vin[cc][at3(L.fs[cc], ii, jj, kk)];
                                }
                                u[cc] = s2;
                            }
                            otri t = {{0, 0, 0}, {0, 0, 0}};
                            for (int d = 0; d < D; ++d) {
                                R world = (d == c ? (R)id[d] : (R)id[d] + (R)0.5) * hh;
                                R from = world - u[d] * dt;
                                R gs = from / hh;
                                R uv = (gs + (d == c ? (R)0.5 : (R)0)) / ((R)g->n[d] + (d == c ? (R)1 : (R)0));
                                tcoord(uv, g->n[d] + (d == c), &t.i0[d], &t.a[d]);
                            }
                            R cv[8];
                            for (int q = 0; q < 8; ++q) {
                                int ii, jj, kk;
                                tri_corner(&t, q, &ii, &jj, &kk);
                                int ok = in3(L.fs[c], ii, jj, kk);
                                cv[q] = ok ? vin[c][at3(L.fs[c], ii, jj, kk)] : (R)0;
                                if (ok) vb2[c][at3(L.fs[c], ii, jj, kk)] += tri_w(&t, q) * lam;
                            }
                            R ub[3] = {0, 0, 0};
                            for (int d = 0; d < D; ++d) ub[d] = lam * tri_da(&t, cv, d) * (-dt / hh);
                            scatter_interp(g, &L, id, c, ub, vb2);
                        }
            }
            for (int c = 0; c < 3; ++c) memcpy(vb[c], vb2[c], L.nf[c] * sizeof(R));
            /* (f) source */
            for (int k = 0; k < sz[2]; ++k)
                for (int j = 0; j < sz[1]; ++j)
                    for (int i = 0; i < sz[0]; ++i)
                        sb[((size_t)k * sz[1] + j) * sz[0] + i] += dt * rb1[at3(L.cs, lo[0] + i, lo[1] + j, lo[2] + k)];
            memcpy(rb, rb1, L.nc * sizeof(R));
        }
        if (src_bar)
            for (size_t i = 0; i < nsrc; ++i) src_bar[i] = (double)sb[i];
        double *outs[3] = {vx_bar, vy_bar, vz_bar};
        for (int c = 0; c < 3; ++c) {
            if (outs[c])
                for (size_t i = 0; i < L.nf[c]; ++i) outs[c][i] = (double)vb[c][i];
            free(vb[c]);
            free(vb2[c]);
        }
        free(rb);
        free(rb1);
        free(w);
        free(wb);
        free(gb);
        free(ccb);
        free(rhs);
        free(sb);
    }
    for (size_t i = 0; i < (size_t)steps * 10; ++i) free(sv[i]);
    free(sv);
    for (int c = 0; c < 3; ++c) free(v[c]);
    free(rho);
    return orc_set_refinement(g, saved_refine, saved_c64);
}

static void solve_rhs(orc_grid *g, const R *rhs, int cycles) {
    const size_t n = (size_t)g->n[0] * g->n[1] * g->n[2];
    double *b = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) b[i] = (double)rhs[i];
    g_brhs = b;
    int saved = g->refine;
    g->refine = cycles;
    memset(g->p64, 0, n * sizeof(double));
    double r2, b2;
    refine_residual(g, 0, &r2, &b2);
    g->norms[0] = r2;
    g->corr = 1;
    for (int it = 1; it <= cycles; ++it) {
        tex_fill(&g->p[0], NULL);
        orc_v_cycle(g, 0);
        refine_residual(g, 1, &r2, &b2);
        g->norms[it] = r2;
    }
    g->corr = 0;
    g->norms[cycles + 1] = b2;
    g->refine = saved;
    g_brhs = NULL;
    free(b);
}
