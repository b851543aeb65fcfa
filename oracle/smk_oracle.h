/*
 * smk_oracle.h -- CPU restatement of SmokeSim's per-step simulation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product in smokesim_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Each stage restates one GLSL compute shader of the reference (file:line in
 * smk_oracle.c) including its ping-pong buffer semantics, its per-kernel
 * out-of-bounds conventions, FP16 storage rounding of RGBA16F images and the
 * GL linear-filter sampler (u*size-0.5, floor/frac, CLAMP_TO_BORDER or
 * CLAMP_TO_EDGE).  Built twice: ORC_REAL=float (parity oracle, 16-bit storage
 * rounding on) and ORC_REAL=double (the FP64 oracle: every field double, no
 * storage rounding).
 *
 * All field transfers use float arrays in GL texel order (x fastest).
 */
#ifndef SMK_ORACLE_H
#define SMK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;

enum orc_field {
    ORC_VX = 0, ORC_VY = 1, ORC_VZ = 2,
    ORC_SMOKE = 3,      /* 4 channels: r,g,b density, a temperature       */
    ORC_VORTICITY = 4,  /* 3D: 4 channels (xyz,0); 2D: 1 channel           */
    ORC_PRESSURE = 5,   /* per level                                        */
    ORC_RESIDUAL = 6,   /* per level                                        */
    ORC_SOLID = 7,      /* per level, 4 channels (obstacle velocity, a)    */
    ORC_DIVERGENCE = 8  /* derived: level-0 face divergence (read only)    */
};

/* nz == 0 makes a 2D grid.  min_level_size: pyramid levels are kept while
 * every dimension is >= this value (the reference uses 4). */
orc_grid *orc_create(int nx, int ny, int nz, float cell_size, float t_ambient,
                     int min_level_size);
void orc_destroy(orc_grid *g);

int orc_levels(const orc_grid *g);
void orc_level_size(const orc_grid *g, int level, int out[3]);

/* 12 floats: wind[4], time_step, density, t_ambient, t_weight_coeff,
 * t_temp_coeff, vorticity_confinement, pad, pad (physics.h:22-32). */
void orc_set_physics(orc_grid *g, const float *settings12);
/* 12 ints per emitter: pos_radius[4], density_temp[4], settings[4]. */
void orc_set_emitters(orc_grid *g, const int *em, int n);
/* Extension: per-emitter upward inflow (vy) applied after emission. */
void orc_set_emitter_inflow(orc_grid *g, const float *inflow, int n);
/* 8 floats per obstacle: pos_radius[4], velocities_anchored[4]. */
void orc_set_obstacles(orc_grid *g, const float *ob, int n);
/* Extension: 6 ints per box: lo[3], hi[3] (half-open cell ranges). */
void orc_set_boxes(orc_grid *g, const int *boxes, int n);
void orc_get_emitters(const orc_grid *g, int *em);
void orc_get_obstacles(const orc_grid *g, float *ob);
void orc_set_time(orc_grid *g, float t);
void orc_set_mg(orc_grid *g, int pre, int post, int coarse, int vcycles);
/* Storage (0 FP32 faces + FP16 smoke, 1 BF16 faces + BF16 smoke holding
 * T - t_ambient, 2 FP16 faces + FP16 smoke); call right after orc_create. */
int orc_set_storage(orc_grid *g, int storage);
/* FP64 iterative refinement of the pressure solve: `outer` correction
 * V-cycles, each followed by an FP64 residual pass; coarse64 smooths the
 * coarsest level in FP64.  outer = 0: the plain solve. */
int orc_set_refinement(orc_grid *g, int outer, int coarse64);
/* binary16 store rounding: 0 nearest even (default), 1 toward zero (llvmpipe) */
int orc_set_f16_rounding(orc_grid *g, int mode);
/* linear filter lerp: 0 multiply then add (default), 1 fused (llvmpipe) */
int orc_set_lerp(orc_grid *g, int fused);
/* |r|^2 of each of the outer+1 residual passes of the last solve, then |b|^2. */
void orc_get_refine_norms(const orc_grid *g, double *out);

/* Copy a field (current read buffer) out / in.  Returns element count. */
long orc_get_field(const orc_grid *g, int field, int level, double *out);
long orc_set_field(orc_grid *g, int field, int level, const double *in);
long orc_field_count(const orc_grid *g, int field, int level);

/* Stages, 1:1 with common/physics/grid.h:31-84. */
void orc_update_emitters_status(orc_grid *g);
void orc_handle_emitters(orc_grid *g);
void orc_apply_inflow(orc_grid *g);
void orc_advect_velocities(orc_grid *g);
void orc_apply_buoyancy(orc_grid *g);
void orc_calculate_vorticity(orc_grid *g);
void orc_apply_vorticity(orc_grid *g);
void orc_pressure_solve(orc_grid *g);
void orc_update_velocities(orc_grid *g);
void orc_advect_smoke(orc_grid *g);
void orc_obstacle_update_step(orc_grid *g);
void orc_update_solid_map(orc_grid *g);
void orc_init_solid_map(orc_grid *g);
void orc_clear_grid(orc_grid *g);

/* V-cycle pieces for unit tests. */
void orc_smooth(orc_grid *g, int level, int iterations);
void orc_smooth_sweep(orc_grid *g, int level, int parity);
void orc_compute_residual(orc_grid *g, int level);
void orc_restrict(orc_grid *g, int level);
void orc_prolong(orc_grid *g, int level);
void orc_v_cycle(orc_grid *g, int level);

/* toggles: advect_smoke, advect_velocities, apply_vorticity, apply_buoyancy,
 * resolve_pressure, update_velocities, handle_emitters (physics.h:10-18);
 * NULL = all on. */
void orc_run_physics_step(orc_grid *g, const int *toggles7);

/* Differentiable rollout (2D and 3D): `steps` steps of source (rho +=
 * src*dt on the box lo..lo+sz), velocity advection, buoyancy (temperature
 * at ambient), vorticity confinement, the refined projection (`cycles`
 * correction cycles from zero) and density advection, from the given
 * initial state.  With rho_bar, also the adjoint: dL/dsrc and dL/d(initial
 * velocity) for dL/d(final density) = rho_bar.  Outputs may be NULL; in 2D
 * vz0 and vz_bar are unused (sz[2] = 1, lo[2] = 0).  Returns 0 on
 * success. */
int orc_rollout(orc_grid *g, int steps, const int *lo, const int *sz, const double *src,
                const double *vx0, const double *vy0, const double *vz0, const double *rho0,
                int cycles, const double *rho_bar, double *rho_final, double *src_bar,
                double *vx_bar, double *vy_bar, double *vz_bar);

/* Sampler exposed for closed-form tests: sample field `field` (read buffer)
 * at normalised coordinates uvw, all channels into out. */
void orc_sample(const orc_grid *g, int field, int level, const float *uvw, double *out);

/* Storage rounding helpers exposed for tests. */
float orc_round_f16(float x);
float orc_round_f16_rtz(float x);
float orc_round_bf16(float x);
int orc_is_double(void);

#ifdef __cplusplus
}
#endif
#endif
