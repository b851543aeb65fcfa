// smk_kernels.h -- launch wrappers of the gfx950 kernels (internal to
// libsmoke_hip.so; the public surface is include/smoke_abi.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

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
};

// Velocity faces: component c has sizes n + e_c (2D: no z component).
struct VelPtrs {
    void *c[3];
};

struct Box3 {
    int lo[3];
    int hi[3];  // exclusive
};

// -- emitters --------------------------------------------------------------
void launch_emitter_status(hipStream_t s, int *em, int ne, const float *time);
void launch_emit(hipStream_t s, const SimParams &P, int vfmt, int sfmt, void *smoke, void *vy,
                 const int *em, const float *inflow, int ne, Box3 box);

// -- advection / forces ------------------------------------------------------
void launch_advect_velocity(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs in,
                            VelPtrs out, const uint16_t *solid0, const void *smoke, int fuse_buoy);
void launch_buoyancy(hipStream_t s, const SimParams &P, int vfmt, int sfmt, const void *vy_in,
                     void *vy_out, const uint16_t *solid0, const void *smoke);
void launch_vorticity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in,
                      const uint16_t *solid0, void *vort);
void launch_apply_vorticity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in, VelPtrs out,
                            const void *vort);
void launch_update_velocity(hipStream_t s, const SimParams &P, int vfmt, VelPtrs in, VelPtrs out,
                            const uint16_t *solid0, const float *p0);
void launch_advect_smoke(hipStream_t s, const SimParams &P, int vfmt, int sfmt, VelPtrs vel,
                         const uint16_t *solid0, const void *smoke_in, void *smoke_out);

// -- pressure solve -------------------------------------------------------
struct Level {
    int n[3];
    float *p;
    float *r;
    uint16_t *s;  // 4 x fp16 per cell
};
void launch_divergence0(hipStream_t s, const SimParams &P, int vfmt, VelPtrs vel,
                        const uint16_t *solid0, float *dv0);
void launch_smooth(hipStream_t s, const SimParams &P, const Level &L, int level, int parity,
                   const float *dv0);
void launch_residual(hipStream_t s, const SimParams &P, const Level &L, int level,
                     const float *dv0, int vfmt, VelPtrs vel);
void launch_restrict_residual(hipStream_t s, const SimParams &P, const Level &fine,
                              const Level &coarse);
void launch_restrict_solid(hipStream_t s, const SimParams &P, const Level &fine,
                           const Level &coarse);
void launch_prolong(hipStream_t s, const SimParams &P, const Level &fine, const Level &coarse);
// Single-workgroup red-black solve of a level that fits in LDS; returns
// false when the level is too large for it.
bool coarse_fits(const SimParams &P, const Level &L);
void launch_coarse_smooth(hipStream_t s, const SimParams &P, const Level &L, int level,
                          int iterations, const float *dv0);

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
// Convert between storage (fmt 0 f32, 1 bf16, 2 f16) and a double/float host
// staging buffer on the device.
void launch_to_f64(hipStream_t s, const void *src, int fmt, double *dst, size_t n);
void launch_from_f64(hipStream_t s, const double *src, void *dst, int fmt, size_t n);
void launch_divergence_field(hipStream_t s, const SimParams &P, int vfmt, VelPtrs vel,
                             const uint16_t *solid0, double *out);
void launch_residual_norm(hipStream_t s, const SimParams &P, const Level &L0, int vfmt,
                          VelPtrs vel, double *acc);

}  // namespace smk
