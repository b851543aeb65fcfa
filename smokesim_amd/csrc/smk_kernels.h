// smk_kernels.h -- launch wrappers of the gfx950 kernels (internal to
// libsmoke_hip.so; the public surface is include/smoke_abi.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

namespace smk {

// Storage format of velocity faces and smoke.
enum Fmt { FMT_F32 = 0, FMT_BF16 = 1, FMT_F16 = 2 };

// Scalars every stage reads (std140 block of advect_vx.comp:19-31 made a
// by-value kernel argument).
struct SimParams {
    int n[3];        // cells; n[2] == 1 for 2D
    int dim;         // 2 or 3
    float h;         // cell size
    float dt, rho, tamb, alpha, beta, eps;
    float wind[4];
    int corr;        // 1: correction solve of the refined pressure mode
    int tile_bx;     // > 0: block width picked by the autotuner
    int gsamp;       // 1: clamped global-load samplers even where buffers fit
    // Reciprocals RN(1/y) of the constant divisors of the advection
    // coordinates (cell size, n[d], n[d] + 1) and of the confinement force
    // (density) for div_rn(); fdiv = 1 when
    // the host has checked the corrected quotient against IEEE division for
    // every one of them (see smk_device.h).
    float rh, rrho, rn[3], rn1[3];
    int fdiv;
    // 1 when div_rn() was checked against IEEE division for dt*h and h*h:
    // the residual's divisors dt*h_l and h_l^2 at level l are those times a
    // power of two, for which the check carries over (every step of div_rn
    // scales exactly) while they stay in [2^-30, 2^30]
    int fdiv_res;
};

// Host: whether div_rn() with r = RN(1/y) equals IEEE division by y for
// every x it accepts (checked once per divisor, cached).
bool fast_div_ok(float y);

// Velocity faces: component c has sizes n + e_c (2D: no z component).
struct VelPtrs {
    void *c[3];
};

// Texel coordinate of one axis of a LINEAR sample: i0 = floor(u'), a = u'-i0.
struct TC {
    int i0;
    float a;
};
// Table of the face-centred interpolation coordinates (see smk_device.h).
struct CoordTab {
    const TC *e;
    int stride;  // max(n) + 1
};

struct Box3 {
    int lo[3];
    int hi[3];  // exclusive
};

// One multigrid level.  Pressure, right-hand side and the neighbour mask are
// stored colour-split: all cells with (i+j+k) even first, then the odd ones,
// each colour as rows of hx slots (ceil(nx/2) cells, padded to a multiple of
// four with solid slots); cell (i,j,k) of colour c sits at
// c*half + (k*ny + j)*hx + i/2.  A red-black half sweep then reads
// the other colour and writes its own with unit stride.  The solid map keeps
// the reference's 4 x fp16 texel in natural (x fastest) order.
struct Level {
    int n[3];
    int hx;
    float *p;        // pressure (split)
    float *r;        // residual / right-hand side (split)
    uint16_t *s;     // solid map, 4 x fp16 per cell (natural)
    uint8_t *mask;   // neighbour mask (split), see k_build_mask
};
// The bottom levels lb..L-1 of the pyramid, solved in one workgroup.
constexpr int kBottomMaxLevels = 8;
constexpr int kBottomMaxCells = 6144;        // per level (up to 6 cells per thread)
constexpr size_t kBottomMaxLds = 144 * 1024;  // bytes (MI355X: 160 KiB of LDS per CU)
struct BottomLevels {
    int nb;                        // number of levels (>= 1)
    int n[kBottomMaxLevels][3];
    int hx[kBottomMaxLevels];
    int off[kBottomMaxLevels];     // LDS offset (cells) of each level
    float hl[kBottomMaxLevels];    // cell size of each level
    float *p[kBottomMaxLevels];
    float *r[kBottomMaxLevels];
    const uint16_t *s[kBottomMaxLevels];
    // cells of each level sorted by colour (parity 0 first, n0[m] of them),
    // level m's list starting at off[m]
    const int *list;
    int n0[kBottomMaxLevels];
};

// The level-l residual's two divisions by launch constants, dt*h_l and h_l^2,
// with their reciprocals for div_rn() (ok: both checked against IEEE
// division on the host).
struct ResDiv {
    float d_b, r_b, d_h2, r_h2;
    int ok;
};

// -- several scenes per launch (fused batches) --------------------------------
// A multi-scene launch runs one stage kernel over the tiles of n scenes: scene
// s owns the logical blocks [start[s], start[s+1]) (`start` in device memory,
// n + 1 entries) and its record below carries its block grid (gx, gy: the
// tile of a logical block is (l % gx, l / gx % gy, l / (gx * gy))) and the
// arguments the single-scene launcher passes.  Every block of a launch has
// the same shape (bx, 256 / bx).
struct MTile {
    int gx, gy;  // tiled launches: tiles along x and y; flat ones: thread extents
    int gz;      // flat launches: thread extent along z
};
struct MultiLaunch {
    const int *start;  // device, n + 1 prefix sums of the scenes' blocks
    int n;             // scenes
    int blocks;        // start[n]
    int bx;            // block width
};
struct MEmitStatus { int *em; int ne; const float *time; };
struct MEmit { MTile t; SimParams P; uint16_t *smoke; void *vy; const int *em; const float *inflow; int ne; Box3 box; };
struct MAdvectVel { MTile t; SimParams P; VelPtrs in, out; const uint8_t *f0; CoordTab T; const uint16_t *smoke; };
struct MVort { MTile t; SimParams P; VelPtrs V; const uint8_t *f0; void *vort; };
struct MApplyVort { MTile t; SimParams P; VelPtrs in, out; const void *vort; };
struct MUpdateVel { MTile t; SimParams P; VelPtrs in, out; const uint16_t *s0; const uint8_t *f0; Level L0; };
struct MAdvectSmoke { MTile t; SimParams P; VelPtrs V; const uint8_t *f0; CoordTab T; const uint16_t *sin; uint16_t *sout; };
struct MDiv0 { MTile t; SimParams P; VelPtrs V; const uint16_t *s0; Level L0; float *dv0, *b0; };
struct MSweep { MTile t; SimParams P; Level L; int level, parity; const float *rhs; float hl; };
struct MResRestrict { MTile t; SimParams P; Level F, C; int level; const float *rhs; VelPtrs V; const uint16_t *s0; float hl; int wf; ResDiv RD; };
struct MProlong { MTile t; Level F, C; const TC *tx, *ty, *tz; };
struct MBottom { BottomLevels B; int pre, post, coarse, corr; };
// kernel variants of the residual + restriction and prolongation launches
// kRrEvenB: the even kernel with batched general-cell loads (levels >= 1)
enum { kRrGeneral = 0, kRrEven = 1, kRrX2 = 2, kRrEvenB = 3 };
enum { kPrGeneral = 0, kPrEven = 1, kPrX2 = 2 };
struct FusedVariant {
    int variant = 0;
    int s[3] = {0, 0, 0};  // extent of the block grid (threads along x, y; z planes)
    int wf = 1;
    ResDiv RD{};
};
struct MCoarse { SimParams P; Level L; int level, iterations; const float *rhs; float hl; };
struct MObstacles { SimParams P; const float *in; float *out; int no; };

}  // namespace smk

// The launchers are declared in the namespace of the build that defines
// them: smk, or smk::gl for the translation units built with SMK_GL_ARITH
// (llvmpipe-compatible binary16 rounding and lerp, see smk_device.h), so
// that a call made inside either build finds only its own launchers.
#ifndef SMK_GL_ARITH
#define SMK_GL_ARITH 0
#endif
#if SMK_GL_ARITH
#define SMK_NS_BEGIN namespace smk { namespace gl {
#define SMK_NS_END } }
#else
#define SMK_NS_BEGIN namespace smk {
#define SMK_NS_END }
#endif

SMK_NS_BEGIN

// -- emitters --------------------------------------------------------------
void launch_emitter_status(hipStream_t s, int *em, int ne, const float *time);
void launch_emit(hipStream_t s, const SimParams &P, int vfmt, int sfmt, void *smoke, void *vy,
                 const int *em, const float *inflow, int ne, Box3 box);

// -- advection / forces ------------------------------------------------------
// flag0: level-0 solid flags (a >= 1), 1 byte per cell, natural order.
void launch_coord_table(hipStream_t s, const SimParams &P, TC *e, int stride);
void launch_solid_flag(hipStream_t s, const uint16_t *solid0, uint8_t *flag0, size_t n);
void launch_advect_velocity(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs in,
                            VelPtrs out, const uint8_t *flag0, CoordTab T, const void *smoke,
                            int fuse_buoy);
void launch_buoyancy(hipStream_t s, const SimParams &P, int vfmt, int sfmt, const void *vy_in,
                     void *vy_out, const uint8_t *flag0, const void *smoke);
void launch_vorticity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in,
                      const uint8_t *flag0, void *vort);
void launch_apply_vorticity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in, VelPtrs out,
                            const void *vort);
// p64: the FP64 refined pressure (colour-split), or null for L0.p
void launch_update_velocity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in, VelPtrs out,
                            const uint16_t *solid0, const uint8_t *flag0, const Level &L0,
                            const double *p64);
void launch_advect_smoke(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs vel,
                         const uint8_t *flag0, CoordTab T, const void *smoke_in, void *smoke_out);

// -- pressure solve -------------------------------------------------------
size_t split_half(const Level &L);  // cells per colour (padded rows)

void launch_build_mask(hipStream_t s, const SimParams &P, const Level &L);
// dv0: level-0 divergence sum; b0: the smoother's term rho*h*dv0/dt
void launch_divergence0(hipStream_t s, const SimParams &P, int vfmt, VelPtrs vel,
                        const uint16_t *solid0, const Level &L0, float *dv0, float *b0);
// rhs: b0 on level 0, the residual r_l above
// e0, e1 (may be null): events stamped at the kernel's start and end
void launch_sweep(hipStream_t s, const SimParams &P, const Level &L, int level, int parity,
                  const float *rhs, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// r_l = b - A p_l on level l (written in place), restricted into r_{l+1};
// p_{l+1} cleared.  rhs is dv0 on level 0 and r_l otherwise.
// write_fine = false skips storing the level-0 residual image (levels > 0
// always store it: their post-smoothing reads it as right-hand side)
void launch_residual_restrict(hipStream_t s, const SimParams &P, const Level &fine,
                              const Level &coarse, int level, const float *rhs, int vfmt,
                              VelPtrs vel, const uint16_t *solid0, bool write_fine = true);
void launch_residual(hipStream_t s, const SimParams &P, const Level &L, int level,
                     const float *rhs, int vfmt, VelPtrs vel, const uint16_t *solid0);
void launch_restrict_solid(hipStream_t s, const SimParams &P, const Level &fine,
                           const Level &coarse);
// prolongation coordinates of one axis of a level: nf fine cells, nc coarse
void launch_prolong_table(hipStream_t s, int nf, int nc, TC *e);
// even_ok: the host verified that the tables of an even-sized pair map fine
// cell 2I+d to coarse texels I-1+d and I+d
void launch_prolong(hipStream_t s, const SimParams &P, const Level &fine, const Level &coarse,
                    const TC *tx, const TC *ty, const TC *tz, bool even_ok);
// Single-workgroup red-black smoothing of a level that fits in LDS.
bool coarse_fits(const SimParams &P, const Level &L);
// the coarsest level of a pyramid may take the single-workgroup smoother up
// to a larger size (its many iterations dominate); LDS bytes it needs
bool coarsest_fits(const Level &L);
size_t coarse_lds_bytes(const Level &L);
void launch_coarse_smooth(hipStream_t s, const SimParams &P, const Level &L, int level,
                          int iterations, const float *rhs);
size_t bottom_lds_bytes(const BottomLevels &B, bool c64);
// v_cycle(lb) of grid.c:177-209 for levels lb..L-1 (lb >= 1): pre/post
// iterations on every level but the last, `coarse` iterations on the last
// (in FP64 when c64).
void launch_bottom_vcycle(hipStream_t s, const SimParams &P, const BottomLevels &B, int pre,
                          int post, int coarse, bool c64);
// FP64 residual pass of the refined solve (see k_refine_residual); part
// holds 2 * refine_partials(L0) doubles, acc receives |r|^2 and |b|^2
size_t refine_partials(const Level &L0);
void launch_refine_residual(hipStream_t s, const SimParams &P, const Level &L0, const double *pin,
                            double *pout, const float *e, const float *bsrc, double bnum,
                            double bden, float *rc, double *part, double *acc, int add_e);
void launch_split64_to_f64(hipStream_t s, const Level &L, const double *src, double *dst);
void launch_split64_from_f64(hipStream_t s, const Level &L, const double *src, double *dst);
// split <-> natural conversion of a level field (pressure or residual)
void launch_split_to_f64(hipStream_t s, const Level &L, const float *src, double *dst);
void launch_split_from_f64(hipStream_t s, const Level &L, const double *src, float *dst);
// acc[0] = |b - A p|^2 on level 0 (acc[1] = 0); part holds
// 2 * residual_norm_partials() doubles of per-block sums
size_t residual_norm_partials(const SimParams &P, const Level &L0);
void launch_residual_norm(hipStream_t s, const SimParams &P, const Level &L0, int vfmt,
                          VelPtrs vel, double *part, double *acc);

// -- obstacles / masks --------------------------------------------------------
void launch_obstacle_update(hipStream_t s, const SimParams &P, const float *ob_in, float *ob_out,
                            int no);
void launch_update_solid(hipStream_t s, const SimParams &P, uint16_t *solid0, const float *ob,
                         int no, const int *boxes, int nb);
void launch_init_solid(hipStream_t s, const SimParams &P, uint16_t *solid0);

// -- utilities ------------------------------------------------------------
void launch_fill_f32(hipStream_t s, float *p, size_t n, float v);
void launch_fill_u16(hipStream_t s, uint16_t *p, size_t n, uint16_t v);
void launch_set_f32(hipStream_t s, float *p, float v);
void launch_copy16(hipStream_t s, const void *src, void *dst, size_t bytes);
// Convert between storage (fmt 0 f32, 1 bf16, 2 f16) and a double/float host
// staging buffer on the device.
void launch_to_f64(hipStream_t s, const void *src, int fmt, double *dst, size_t n);
void launch_from_f64(hipStream_t s, const double *src, void *dst, int fmt, size_t n);
void launch_divergence_field(hipStream_t s, const SimParams &P, int vfmt, VelPtrs vel,
                             const uint16_t *solid0, double *out);

// -- multi-scene launchers (fused batches; see MultiLaunch) -------------------
// Each runs the single-scene kernel's arithmetic on every scene of the
// launch; E is the device array of the n scenes' records.  dim, vfmt and
// sfmt are common to the scenes of a launch.
void launch_emit_status_multi(hipStream_t s, const MultiLaunch &m, const MEmitStatus *E);
void launch_emit_multi(hipStream_t s, int dim, int vfmt, int sfmt, const MultiLaunch &m, const MEmit *E);
void launch_advect_velocity_multi(hipStream_t s, int dim, int vfmt, int sfmt, bool fuse_buoy, bool buf,
                                  const MultiLaunch &m, const MAdvectVel *E);
void launch_vorticity_multi(hipStream_t s, int dim, int vfmt, const MultiLaunch &m, const MVort *E);
void launch_apply_vorticity_multi(hipStream_t s, int dim, int vfmt, const MultiLaunch &m, const MApplyVort *E);
void launch_update_velocity_multi(hipStream_t s, int dim, int vfmt, const MultiLaunch &m, const MUpdateVel *E);
void launch_advect_smoke_multi(hipStream_t s, int dim, int vfmt, int sfmt, bool buf, const MultiLaunch &m,
                               const MAdvectSmoke *E);
void launch_obstacle_update_multi(hipStream_t s, int dim, const MultiLaunch &m, const MObstacles *E);
// four: 4 cells per thread (k_divergence0_x4 for 3D rows of a multiple of 4,
// k_sweep4 for colour rows of a multiple of 4)
void launch_divergence0_multi(hipStream_t s, int dim, int vfmt, bool four, const MultiLaunch &m, const MDiv0 *E);
void launch_sweep_multi(hipStream_t s, int dim, bool four, const MultiLaunch &m, const MSweep *E);
void launch_residual_restrict_multi(hipStream_t s, int dim, int vfmt, int variant, const MultiLaunch &m,
                                    const MResRestrict *E);
void launch_prolong_multi(hipStream_t s, int dim, int variant, const MultiLaunch &m, const MProlong *E);
FusedVariant fused_variant_residual_restrict(const SimParams &P, const Level &F, const Level &C, int level,
                                             bool write_fine);
FusedVariant fused_variant_prolong(const SimParams &P, const Level &F, const Level &C, bool even_ok);
// one workgroup per scene (m.blocks == m.n); lds: the largest scene's bytes
void launch_coarse_smooth_multi(hipStream_t s, int dim, const MultiLaunch &m, const MCoarse *E, size_t lds);
// own: cells per thread per level, bottom_own() of the largest scene
void launch_bottom_vcycle_multi(hipStream_t s, int dim, const MultiLaunch &m, const MBottom *E, size_t lds,
                                int own);
int bottom_own(const BottomLevels &B);
// launch_set_f32 of every scene's time at once
void launch_set_times(hipStream_t s, float *const *times, int n, float t);
// whether the advection kernels take the range-checked buffer samplers
bool buf_samplers_ok(const SimParams &P);

SMK_NS_END

#if !SMK_GL_ARITH
namespace smk {
// -- llvmpipe-compatible arithmetic -------------------------------------------
// The launchers of smk_kernels.hip and smk_pressure.hip built a second time
// with SMK_GL_ARITH (binary16 stores rounded toward zero, fused-lerp linear
// filter; see smk_device.h).  Only those whose results depend on the two
// choices are declared: emission, advection, buoyancy, 3D vorticity, solid
// maps and their restriction, prolongation, the bottom V-cycle and uploads.
namespace gl {
void launch_emit(hipStream_t s, const SimParams &P, int vfmt, int sfmt, void *smoke, void *vy,
                 const int *em, const float *inflow, int ne, Box3 box);
void launch_advect_velocity(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs in,
                            VelPtrs out, const uint8_t *flag0, CoordTab T, const void *smoke,
                            int fuse_buoy);
void launch_buoyancy(hipStream_t s, const SimParams &P, int vfmt, int sfmt, const void *vy_in,
                     void *vy_out, const uint8_t *flag0, const void *smoke);
void launch_vorticity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in,
                      const uint8_t *flag0, void *vort);
void launch_advect_smoke(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs vel,
                         const uint8_t *flag0, CoordTab T, const void *smoke_in, void *smoke_out);
void launch_update_solid(hipStream_t s, const SimParams &P, uint16_t *solid0, const float *ob,
                         int no, const int *boxes, int nb);
void launch_restrict_solid(hipStream_t s, const SimParams &P, const Level &fine,
                           const Level &coarse);
void launch_prolong(hipStream_t s, const SimParams &P, const Level &fine, const Level &coarse,
                    const TC *tx, const TC *ty, const TC *tz, bool even_ok);
void launch_bottom_vcycle(hipStream_t s, const SimParams &P, const BottomLevels &B, int pre,
                          int post, int coarse, bool c64);
void launch_from_f64(hipStream_t s, const double *src, void *dst, int fmt, size_t n);
}  // namespace gl
}  // namespace smk
#endif

namespace smk {
// -- fixed staging array of the no-allocation download (smk_noalloc.hip) ----
// Fields of at most kNoAllocMaxValues values; the array holds them as doubles
// (noalloc_stage_bytes()) followed by the same count of floats.
constexpr size_t kNoAllocMaxValues = (size_t)4 << 20;
// Device address of the array on `device`, which must be current (looked up
// on the first call per device, then kept).
hipError_t noalloc_scratch(int device, unsigned char **ptr);
// Held for the whole of a no-allocation call: calls on one GPU share the array.
std::mutex &noalloc_lock(int device);
size_t noalloc_stage_bytes();
// doubles -> float (out_f32) or double, alpha_off added to every 4th value
void launch_stage_to_out(hipStream_t s, const double *src, void *dst, int out_f32, size_t n,
                         double alpha_off);

// -- differentiable step and adjoint (smk_adjoint.hip) ------------------------
// `fmt` is the storage format (FMT_*) of the rollout's state: velocities and
// density; adjoint fields, the source and all arithmetic are FP32.
void launch_d_source(hipStream_t s, const SimParams &P, int fmt, void *rho, const float *src, Box3 box);
void launch_d_buoyancy(hipStream_t s, const SimParams &P, int fmt, const void *vy_in, void *vy_out,
                       const uint8_t *f0, const void *rho);
void launch_d_vorticity(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0, float4 *w);
void launch_d_confine(hipStream_t s, const SimParams &P, int fmt, VelPtrs in, VelPtrs out, const float4 *w);
void launch_d_advect_scalar(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0,
                            CoordTab T, const void *rin, void *rout);
// conversions of state arrays (n elements) between FP32 and `fmt`
void launch_state_from_f32(hipStream_t s, int fmt, const float *src, void *dst, size_t n);
void launch_state_to_f32(hipStream_t s, int fmt, const void *src, float *dst, size_t n);
void launch_maxabs(hipStream_t s, const float *x, size_t n, unsigned int *out);
// *out = max(*out, sqrt(norms[0] / norms[1])) for the squared residual and
// right-hand-side norms of a refinement pass (one thread, stream ordered).
void launch_track_residual(hipStream_t s, const double *norms, double *out);
// Source fitting: part (kFitPartials doubles) receives per-block sums.
// launch_fit_loss: rb = 2 (rho - tgt), *L = sum (rho - tgt)^2 (FP64),
// losses[*cnt] = *L when *cnt < cap, ++*cnt.  launch_fit_update: the
// projected Polyak step src <- max(0, src - (*L / (2 |g|^2)) g).
constexpr int kFitPartials = 256;
void launch_fit_loss(hipStream_t s, int fmt, const void *rho, const float *tgt, size_t n, float *rb,
                     double *part, double *L, double *losses, int cap, int *cnt);
void launch_fit_update(hipStream_t s, float *src, const float *g, size_t n, const double *L, double *part);
void launch_fx_scale(hipStream_t s, const unsigned int *m_lam, const unsigned int *m_field, float c0,
                     float c1, double *scale);
void launch_fx_to_f32(hipStream_t s, const unsigned long long *acc, const double *scale,
                      const float *base, float *dst, size_t n);
// ubar receives the departure-point velocity adjoints of every source
// position, transposed through the face interpolation by
// launch_a_interp_gather
void launch_a_advect_scalar(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0,
                            CoordTab T, const void *rin, const float *rbar_out,
                            unsigned long long *acc_r, const double *s_r, float *const ubar[4][3]);
void launch_a_advect_velocity(hipStream_t s, const SimParams &P, int fmt, VelPtrs v, const uint8_t *f0,
                              CoordTab T, VelPtrs vbar_out, unsigned long long *const acc[3],
                              const double *s_acc, float *const ubar[4][3]);
void launch_a_interp_gather(hipStream_t s, const SimParams &P, CoordTab T, float *const ubar[4][3],
                            int cells, VelPtrs base, VelPtrs out);
void launch_a_buoyancy(hipStream_t s, const SimParams &P, const uint8_t *f0, const float *vybar,
                       unsigned long long *acc_r, const double *s_r);
void launch_a_confine(hipStream_t s, const SimParams &P, const uint8_t *f0, const float4 *w,
                      VelPtrs vbar_in, VelPtrs vbar_out, float4 *wbar, float4 *gbar, float4 *ccbar);
void launch_a_project_rhs(hipStream_t s, const SimParams &P, VelPtrs vbar, const uint8_t *f0,
                          const Level &L0, float *rhs);
void launch_a_project_apply(hipStream_t s, const SimParams &P, VelPtrs vin, VelPtrs vout,
                            const uint8_t *f0, const Level &L0, const double *x);
void launch_a_source(hipStream_t s, const SimParams &P, const float *rbar, float *sbar, Box3 box);

}  // namespace smk
