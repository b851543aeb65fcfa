/*
 * smoke_abi.h -- C-ABI of libsmoke_hip.so, the MI355X (gfx950) smoke
 * simulation step.
 *
 * This header replaces the per-step operator surface of SmokeSim's
 * common/physics/grid.h:31-84 and physics.h:50-67.  The reference passes all
 * data through bound GL textures, UBOs and SSBOs (opengl_bindings.h:8-42);
 * here every datum is an explicit argument or lives in an opaque grid handle
 * that owns its device memory and one HIP stream.
 *
 * Conventions
 *   - Every function returns smk_status (SMK_OK == 0); none exits or aborts.
 *     smk_last_error() returns a thread-local message for the last failure.
 *   - Work is asynchronous on the handle's stream; smk_sync() waits for it.
 *     Host pointers are borrowed for the duration of the call only.
 *   - A handle is not thread safe; distinct handles may be used from
 *     distinct host threads.
 *   - Field transfers use GL texel order (x fastest, then y, then z) with the
 *     staggered sizes of the reference's MAC textures (3d/grid3d.c:25-43):
 *     vx (nx+1)*ny*nz, vy nx*(ny+1)*nz, vz nx*ny*(nz+1); smoke 4 channels
 *     (r,g,b density, a temperature); solid 4 channels (obstacle velocity,
 *     solid fraction a).  2D grids have nz == 1 internally and no vz.
 * Only plain C types appear in the signatures.
 */
#ifndef SMOKE_ABI_H
#define SMOKE_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMK_ABI_VERSION 1

/* Opaque handles. */
struct smk_grid;
struct smk_rollout;
struct smk_batch;

typedef enum smk_status {
    SMK_OK = 0,
    SMK_ERR_INVALID_ARG = 1,
    SMK_ERR_OUT_OF_MEMORY = 2,
    SMK_ERR_HIP = 3,
    SMK_ERR_NOT_CONVERGED = 4,
    SMK_ERR_NO_DEVICE = 5,
    SMK_ERR_UNSUPPORTED = 6
} smk_status;

/* Storage format of velocity faces and smoke.  FP32: velocity faces FP32 and
 * smoke FP16 exactly as the reference's R32F / RGBA16F textures
 * (3d/grid3d.c:26-38).  BF16 / FP16: velocity faces and smoke 16-bit,
 * arithmetic in FP32. */
typedef enum smk_precision {
    SMK_PREC_FP32 = 0,
    SMK_PREC_BF16 = 1,
    SMK_PREC_FP16 = 2
} smk_precision;

typedef enum smk_field {
    SMK_FIELD_VX = 0,
    SMK_FIELD_VY = 1,
    SMK_FIELD_VZ = 2,
    SMK_FIELD_SMOKE = 3,
    SMK_FIELD_VORTICITY = 4,  /* 3D: 4 channels (x,y,z,0); 2D: 1 channel */
    SMK_FIELD_PRESSURE = 5,   /* per multigrid level */
    SMK_FIELD_RESIDUAL = 6,   /* per multigrid level */
    SMK_FIELD_SOLID = 7,      /* per multigrid level, 4 channels */
    SMK_FIELD_DIVERGENCE = 8  /* derived, level 0: face divergence / h, 0 in solids */
} smk_field;

typedef enum smk_dtype { SMK_DTYPE_F32 = 0, SMK_DTYPE_F64 = 1 } smk_dtype;

/* Replaces get_grid2d/get_grid3d + gen_grid_textures (grid.c:25-34,
 * 2d/grid2d.c:10-64, 3d/grid3d.c:10-68) and gui.c:117-135. */
typedef struct smk_grid_desc {
    int size[3];        /* nx, ny, nz; nz == 0 selects a 2D grid            */
    float cell_size;    /* grid.h:23 cell_size                               */
    float t_ambient;    /* smoke clear / border temperature                  */
    int storage;        /* smk_precision                                     */
    int min_level_size; /* pyramid keeps levels with every dim >= this; the
                           reference rule (2d/grid2d.c:45-49) is 4           */
    int enclosed;       /* 1: write the border ring (init_solid_map)         */
    int device;         /* HIP device ordinal                                */
} smk_grid_desc;

/* Byte-identical to physics_settings (common/physics/physics.h:22-32, 48 B).
 * cell_size comes from smk_grid_desc. */
typedef struct smk_physics_settings {
    float wind[4];
    float time_step;
    float density;
    float t_ambient;
    float t_weight_coeff;        /* alpha: weight of density in buoyancy */
    float t_temp_coeff;          /* beta: weight of temperature          */
    float vorticity_confinement; /* epsilon                               */
    float _pad0;
    float _pad1;
} smk_physics_settings;

/* Byte-identical to physics_step_settings (physics.h:10-18). */
typedef struct smk_step_settings {
    int advect_smoke;
    int advect_velocities;
    int apply_vorticity;
    int apply_buoyancy;
    int resolve_pressure;
    int update_velocities;
    int handle_emitters;
} smk_step_settings;

/* Byte-identical to emitter (common/physics/emitters.h:10-14, 48 B):
 * pos_radius = (x, y, z, r) cells; density_temp = (r, g, b, T);
 * emitter_settings = (delay, last emit time, emit flag, falloff flag). */
typedef struct smk_emitter {
    int pos_radius[4];
    int density_temp[4];
    int emitter_settings[4];
} smk_emitter;

/* Byte-identical to obstacle (common/physics/obstacles.h:10-13, 32 B). */
typedef struct smk_obstacle {
    float pos_radius[4];
    float velocities_anchored[4];
} smk_obstacle;

/* Extension: static axis-aligned box obstacle, half-open cell ranges. */
typedef struct smk_box {
    int lo[3];
    int hi[3];
} smk_box;

/* V-cycle schedule.  The reference hard-codes 16 pre, 16 post, 80 coarsest
 * iterations and one V-cycle per step (grid.c:173-209). */
typedef struct smk_mg_params {
    int pre_iterations;    /* red-black iterations before restriction   */
    int post_iterations;   /* after prolongation                         */
    int coarse_iterations; /* on the coarsest level                      */
    int v_cycles;          /* V-cycles per pressure solve                */
} smk_mg_params;

/* ---- handles (grid.c:9-34, gui.c:117-135) ---------------------------- */
smk_status smk_grid_create(const smk_grid_desc *desc, struct smk_grid **out);
smk_status smk_grid_destroy(struct smk_grid *g);
/* clear_grid (grid.c:442-448, 3d/grid3d.c:164-184, 2d/grid2d.c:146-159) */
smk_status smk_grid_clear(struct smk_grid *g);
smk_status smk_grid_levels(struct smk_grid *g, int *levels);
smk_status smk_grid_level_size(struct smk_grid *g, int level, int size_out[3]);

/* ---- settings and scene (physics.c:52-70, 142-198; emitters.c:60-83;
 *      obstacles.c:67-105) ---------------------------------------------- */
smk_status smk_set_physics(struct smk_grid *g, const smk_physics_settings *s);
smk_status smk_set_step_toggles(struct smk_grid *g, const smk_step_settings *s);
smk_status smk_set_mg_params(struct smk_grid *g, const smk_mg_params *p);
smk_status smk_set_emitters(struct smk_grid *g, const smk_emitter *e, int n);
smk_status smk_get_emitters(struct smk_grid *g, smk_emitter *e, int n);
/* Extension: upward inflow (vy) forced on faces inside emitting emitters. */
smk_status smk_set_emitter_inflow(struct smk_grid *g, const float *inflow, int n);
smk_status smk_set_obstacles(struct smk_grid *g, const smk_obstacle *o, int n);
/* extract_obstacles_data (obstacles.c:100-105) */
smk_status smk_get_obstacles(struct smk_grid *g, smk_obstacle *o, int n);
smk_status smk_set_box_obstacles(struct smk_grid *g, const smk_box *b, int n);
/* Replaces the time UBO (physics.c:182-198, main.c:132-135). */
smk_status smk_set_time(struct smk_grid *g, float t);

/* ---- stages, 1:1 with grid.h:36-84 ------------------------------------- */
smk_status smk_update_emitters_status(struct smk_grid *g); /* grid.c:89-93   */
smk_status smk_handle_emitters(struct smk_grid *g);        /* grid.c:78-87 (+inflow) */
smk_status smk_advect_velocities(struct smk_grid *g);      /* grid.c:57-76   */
smk_status smk_apply_buoyancy(struct smk_grid *g);         /* grid.c:106-115 */
smk_status smk_calculate_vorticity(struct smk_grid *g);    /* grid.c:117-124 */
smk_status smk_apply_vorticity(struct smk_grid *g);        /* grid.c:126-145 */
smk_status smk_pressure_solve(struct smk_grid *g);         /* grid.c:173-209 */
smk_status smk_update_velocities(struct smk_grid *g);      /* grid.c:36-55   */
smk_status smk_advect_smoke(struct smk_grid *g);           /* grid.c:95-104  */
smk_status smk_obstacle_update_step(struct smk_grid *g);   /* grid.c:156-162 */
smk_status smk_update_solid_map(struct smk_grid *g);       /* grid.c:164-171 */
smk_status smk_init_solid_map(struct smk_grid *g);         /* grid.c:147-154 */
/* run_physics_step (physics.c:72-101) honouring the step toggles. */
smk_status smk_run_physics_step(struct smk_grid *g);
/* n consecutive run_physics_step calls with time = t0 + i*dt_time set
 * before step i; replayed from a captured HIP graph when possible. */
smk_status smk_run_steps(struct smk_grid *g, int n, float t0, float dt_time);

/* ---- field I/O, diagnostics ------------------------------------------- */
smk_status smk_field_count(struct smk_grid *g, int field, int level, size_t *count);
smk_status smk_download_field(struct smk_grid *g, int field, int level, void *host,
                              size_t bytes, int dtype);
smk_status smk_upload_field(struct smk_grid *g, int field, int level, const void *host,
                            size_t bytes, int dtype);
smk_status smk_sync(struct smk_grid *g);

/* ---- mixed precision: FP64 iterative refinement ------------------------
 * No counterpart in the reference (its solve is one fixed V-cycle per step,
 * grid.c:173-209).  With outer_cycles > 0 each pressure solve keeps the
 * level-0 pressure in FP64 and runs: an FP64 residual pass r = b - A p
 * (compute_residual.comp's operator), then outer_cycles times a correction
 * V-cycle on A e = r (FP32 smoothing, out-of-domain neighbours solid,
 * coarsest level in FP64 when coarse_fp64) followed by p += e and the next
 * residual pass.  update_velocities then takes the FP64 gradient.  Needs an
 * enclosed grid.  outer_cycles = 0 restores the reference solve. */
smk_status smk_set_refinement(struct smk_grid *g, int outer_cycles, int coarse_fp64);
/* |r| after each of the outer_cycles + 1 residual passes of the last solve,
 * then |b|; n >= outer_cycles + 2.  Synchronises the stream. */
smk_status smk_refinement_norms(struct smk_grid *g, double *out, int n);
/* Measurement helper (no reference counterpart): `reps` copies of `bytes`
 * bytes between two device buffers by a plain streaming kernel; reports the
 * average time of one copy.  Used to calibrate HBM counters and as the
 * practical bandwidth ceiling. */
smk_status smk_calibrate_copy(int device, size_t bytes, int reps, double *ms_per_copy);
/* hipDeviceSynchronize on `device` through the library's own HIP runtime. */
smk_status smk_device_sync(int device);
/* Level-0 residual L2 norm of the current pressure (FP64 accumulation). */
smk_status smk_residual_norm(struct smk_grid *g, double *out);
/* Use a caller stream instead of the handle's own (NULL restores it). */
smk_status smk_set_stream(struct smk_grid *g, void *hip_stream);
smk_status smk_get_stream(struct smk_grid *g, void **hip_stream);
/* 1: capture each step's launch sequence in a HIP graph and replay it. */
smk_status smk_set_graph_mode(struct smk_grid *g, int enable);
/* Per-kernel timing with HIP events on the handle's stream: kernel ids are
 * smk_kernel_id.  smk_profile_read returns the total time and the number of
 * launches since profiling was enabled or last read for that kernel.  While
 * profiling is on, steps launch eagerly even in graph mode. */
smk_status smk_profile_enable(struct smk_grid *g, int enable);
smk_status smk_profile_read(struct smk_grid *g, int kernel_id, double *total_ms,
                            int *launches);
typedef enum smk_kernel_id {
    SMK_K_SMOOTH_L0 = 0,   /* red-black sweep kernel on level 0 */
    SMK_K_ADVECT_VEL = 1,
    SMK_K_ADVECT_SMOKE = 2,
    SMK_K_RESIDUAL_L0 = 3,
    SMK_K_COUNT = 4
} smk_kernel_id;

/* ---- differentiable rollouts ----------------------------------------------
 * No counterpart in the reference (it has no gradients).  A rollout runs
 * `steps` differentiable steps on a 2D or 3D enclosed grid (in 2D:
 * src_lo[2] = 0, src_size[2] = 1, vz and vz_bar unused).  Its state
 * (velocities, density, checkpoints, saved step states) is stored in the
 * grid's format (FP32, BF16 or FP16); arithmetic, adjoint fields and the
 * source are FP32.  A step is:
 * source (rho += src*dt on a box of cells), velocity advection, buoyancy
 * (one density channel, temperature at ambient), vorticity confinement, a
 * projection refined with `refine_cycles` correction cycles from a cold
 * start, and density advection -- the formulas of grid.c's stages.  The
 * rollout borrows the grid's solid maps, physics settings and pressure
 * solver; do not step the grid between a forward and its backward.
 * Checkpoints are kept every checkpoint_every steps; backward recomputes
 * each segment but the last from its checkpoint and runs the hand-written
 * adjoint of every stage in reverse.  The adjoint projection solves are
 * refined with `adjoint_cycles` correction cycles (0: refine_cycles), each
 * warm-started from the previous adjoint step's solution; the first starts
 * from zero and gets max(refine_cycles, adjoint_cycles).  The sequence of
 * adjoint solves does not depend on checkpoint_every, so neither do the
 * gradients.  Host arrays are FP32 in GL texel order; the
 * source is src_size[0]*src_size[1]*src_size[2] values. */
typedef struct smk_rollout_desc {
    int steps;
    int checkpoint_every;
    int src_lo[3];
    int src_size[3];
    int refine_cycles;
    int adjoint_cycles;
} smk_rollout_desc;
smk_status smk_rollout_create(struct smk_grid *g, const smk_rollout_desc *d,
                              struct smk_rollout **out);
smk_status smk_rollout_destroy(struct smk_rollout *r);
smk_status smk_rollout_set_initial(struct smk_rollout *r, const float *vx, const float *vy,
                                   const float *vz, const float *rho);
smk_status smk_rollout_set_source(struct smk_rollout *r, const float *src);
/* Runs the rollout from the initial state; rho_final (optional) receives
 * the final density. */
smk_status smk_rollout_forward(struct smk_rollout *r, float *rho_final);
/* Given dL/d(final density), writes dL/d(source) and dL/d(initial velocity)
 * (any output may be null). */
smk_status smk_rollout_backward(struct smk_rollout *r, const float *rho_bar, float *src_bar,
                                float *vx_bar, float *vy_bar, float *vz_bar);
/* The largest relative residual |b - A p| / |b| (FP64, after the last
 * correction cycle) over the projection solves of the last forward
 * (recomputations during backward included) and over the adjoint projection
 * solves of the last backward; either output may be null. */
smk_status smk_rollout_residuals(struct smk_rollout *r, double *forward_max, double *adjoint_max);
/* Source fitting on the device (no counterpart in the reference).  One
 * iteration is the forward rollout from the initial state, the loss
 * L = sum (rho_final - target)^2 (FP64, fixed-order sums), the backward pass
 * from rho_bar = 2 (rho_final - target), and the projected Polyak step
 * src <- max(0, src - (L / (2 |dL/dsrc|^2)) dL/dsrc) (no step when the
 * gradient is 0), with no host synchronisation.  smk_rollout_set_target
 * starts a fit.  smk_rollout_fit captures the iteration once as a HIP graph
 * on the grid's stream (again after the grid's settings or buffers change)
 * and queues `iterations` more replays; it returns at once.  Queue one
 * iteration per rollout in turn to keep several streams busy.
 * smk_rollout_fit_result waits for the stream and copies the loss before
 * each step since smk_rollout_set_target (min(n, recorded) values, at most
 * 4096 are kept), the number recorded, and the current source (any output
 * may be null).  Rollouts on grids with different streams fit
 * concurrently. */
smk_status smk_rollout_set_target(struct smk_rollout *r, const float *target);
smk_status smk_rollout_fit(struct smk_rollout *r, int iterations);
smk_status smk_rollout_fit_result(struct smk_rollout *r, double *losses, int n, int *recorded,
                                  float *src);

/* ---- launch tiles and batches of uneven scenes ----------------------------
 * No counterpart in the reference (one scene, fixed 8x8x8 workgroups,
 * grid.c:41).  smk_set_tile fixes the block width of the grid's kernels (a
 * power of two <= 256, 256 threads per block; 0 restores the built-in
 * choice).  smk_autotune times every candidate width on a scratch grid with
 * the same description and settings (`reps` steps each) and keeps the
 * fastest.  A batch owns n grids of any sizes and `nstreams` streams (scene
 * i runs on stream i mod nstreams); smk_batch_step replays the scenes'
 * captured steps (graphs cached per ping-pong state), either one graph per
 * scene or all scenes in one graph with one parallel branch per stream.
 * Configure scene i through smk_batch_grid (the grid stays owned by the
 * batch). */
smk_status smk_set_tile(struct smk_grid *g, int tile_bx);
smk_status smk_get_tile(struct smk_grid *g, int *tile_bx);
smk_status smk_autotune(struct smk_grid *g, int reps, int *tile_bx, double *ms_per_step);
/* ---- sampler implementation ------------------------------------------------
 * No counterpart in the reference.  The advection kernels sample through
 * range-checked buffer loads (a texel outside the texture reads the border
 * 0 from the buffer unit) when every field of the grid is under 2 GiB, and
 * through clamped global loads otherwise.  path 0 keeps that automatic
 * choice; path 1 takes the clamped global loads on any grid.  In 3D
 * rollouts path 1 also makes the adjoint advection scatter every
 * contribution with its own global atomic instead of through LDS bins.
 * Results are identical either way; path 1 lets small grids exercise the
 * samplers that large grids use, and tests compare the two scatters. */
smk_status smk_set_sampler_path(struct smk_grid *g, int path);
smk_status smk_batch_create(const smk_grid_desc *descs, int n, int nstreams,
                            struct smk_batch **out);
smk_status smk_batch_destroy(struct smk_batch *b);
smk_status smk_batch_grid(struct smk_batch *b, int i, struct smk_grid **g);
smk_status smk_batch_autotune(struct smk_batch *b, int reps, int *tiles);
/* SMK_BATCH_PER_SCENE (default): every scene replays its own captured step
 * on its stream.  SMK_BATCH_SINGLE_GRAPH: all scenes in one graph.
 * SMK_BATCH_FUSED: one launch per stage, multigrid level and half sweep
 * runs that stage for every scene (per dimension), the whole batch step
 * captured as one graph on the batch's first stream; fields equal the
 * per-scene modes' bit for bit.  It needs scenes with one storage format,
 * one multigrid schedule, all step toggles on, no refinement and static
 * obstacles (smk_batch_set_mode and smk_batch_step return
 * SMK_ERR_UNSUPPORTED, naming the scene, otherwise). */
enum { SMK_BATCH_PER_SCENE = 0, SMK_BATCH_SINGLE_GRAPH = 1, SMK_BATCH_FUSED = 2 };
smk_status smk_batch_set_mode(struct smk_batch *b, int mode);
/* Fused batches: block width (blocks of width x 256/width threads) of each
 * kind of fused launch.  By default a launch takes the power of two that
 * leaves the fewest idle lanes over its scenes' thread extents.
 * smk_batch_autotune in SMK_BATCH_FUSED mode instead times every candidate
 * width (8 to 128) per launch kind on a scratch batch with the same scenes,
 * each launch bracketed by its own events, and keeps the fastest per kind
 * (tiles[i] then reports the scenes' own grid widths, unchanged).  The
 * red-black sweeps, the curl, the confinement, the level-0 divergence and
 * the projection update always run 256-thread blocks over each scene's
 * threads taken as one linear range; the autotuner records 256 for them.
 * smk_batch_fused_widths copies the widths in force, in SMK_FUSED_* order
 * (0: the default rule); results never depend on them. */
enum {
    SMK_FUSED_EMIT = 0, SMK_FUSED_ADVECT_VELOCITY, SMK_FUSED_VORTICITY, SMK_FUSED_APPLY_VORTICITY,
    SMK_FUSED_DIVERGENCE, SMK_FUSED_SWEEP, SMK_FUSED_RESTRICT, SMK_FUSED_PROLONG,
    SMK_FUSED_UPDATE_VELOCITY, SMK_FUSED_ADVECT_SMOKE, SMK_FUSED_KINDS
};
smk_status smk_batch_fused_widths(struct smk_batch *b, int *widths, int n);
smk_status smk_batch_step(struct smk_batch *b, int steps, float t0, float dt_time);
smk_status smk_batch_sync(struct smk_batch *b);

/* ---- the GL choices of the shaders' arithmetic -------------------------------
 * No counterpart in the reference, whose results depend on the GL driver.
 * The GL specification leaves two pieces of the shaders' arithmetic to the
 * implementation: the rounding of values stored to 16-bit float images
 * (RGBA16F smoke, 3D vorticity and the solid maps, 3d/grid3d.c:37-40,64) and
 * the rounding inside the linear texture filter (texture() in
 * advect_vx.comp:73, advect_smoke.comp:68, apply_buoyancy.comp:44,
 * prolong_pressure.comp:45).  SMK_ARITH_DEFAULT rounds binary16 stores to
 * nearest even and lerps v0 + a*(v1 - v0) with two roundings;
 * SMK_ARITH_LLVMPIPE rounds them toward zero and lerps with one fused
 * multiply-add, as Mesa llvmpipe does, so that results can be compared with
 * the reference shaders run there.  Divisions are IEEE in both modes. */
enum { SMK_ARITH_DEFAULT = 0, SMK_ARITH_LLVMPIPE = 1 };
smk_status smk_set_arithmetic(struct smk_grid *g, int mode);

/* ---- download without allocation ------------------------------------------
 * No counterpart in the reference (its only readback, extract_obstacles_data,
 * obstacles.c:100-105, reads a buffer in place).  For callers that cannot
 * allocate device memory -- tight serving loops and harnesses that forbid
 * hipMalloc -- smk_download_field with the same arguments and results,
 * staged through a fixed array per GPU instead of a device buffer allocated
 * per call (no device or host allocation).  Limit: fields of at most
 * SMK_NOALLOC_MAX_VALUES values; larger ones return SMK_ERR_INVALID_ARG
 * (use smk_download_field).  The array (48 MiB) is held on every GPU the
 * entry point runs on until the process ends.  Calls on one GPU share it;
 * each call is synchronous and holds a per-GPU lock, so calls from several
 * host threads on one GPU wait for each other instead of overlapping. */
#define SMK_NOALLOC_MAX_VALUES (4u << 20)
smk_status smk_download_field_noalloc(struct smk_grid *g, int field, int level, void *host,
                                      size_t bytes, int dtype);

const char *smk_last_error(void);
int smk_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
