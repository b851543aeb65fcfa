/*
 * smk_driver.c -- a C host of libsmoke_hip.so: the simulation loop of
 * SmokeSim's main.c (main.c:92-141: run_physics_step once per frame, then the
 * time update) without the window, GUI and renderer.
 *
 *   smk_driver [nx ny nz] [steps] [storage: fp32|bf16|fp16]
 *
 * Builds one scene (an emitter with upward inflow near the floor and an
 * anchored sphere above it, enclosed domain), runs 3 + `steps` steps and
 * prints the milliseconds per step and field checksums.  nz = 0 selects 2D.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "smoke_abi.h"

#define CHECK(call)                                                               \
    do {                                                                          \
        smk_status s_ = (call);                                                   \
        if (s_ != SMK_OK) {                                                       \
            fprintf(stderr, "%s failed (%d): %s\n", #call, (int)s_, smk_last_error()); \
            return 1;                                                             \
        }                                                                         \
    } while (0)

static double now_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int main(int argc, char **argv) {
    int nx = 128, ny = 128, nz = 128, steps = 50;
    smk_precision prec = SMK_PREC_FP32;
    if (argc >= 4) {
        nx = atoi(argv[1]);
        ny = atoi(argv[2]);
        nz = atoi(argv[3]);
    }
    if (argc >= 5) steps = atoi(argv[4]);
    if (argc >= 6) prec = !strcmp(argv[5], "bf16") ? SMK_PREC_BF16 : (!strcmp(argv[5], "fp16") ? SMK_PREC_FP16 : SMK_PREC_FP32);
    const int dim3 = nz > 0;

    /* gui.c:117-135: create the grid, then init_solid_map when enclosed */
    smk_grid_desc d = {{nx, ny, nz}, 0.1f, 293.0f, prec, 8, 1, 0};
    struct smk_grid *g = NULL;
    CHECK(smk_grid_create(&d, &g));

    /* physics.c:7-20 defaults with the benchmark coefficients */
    smk_physics_settings ps = {{0, 0, 0, 0}, 0.1f, 1.0f, 293.0f, 0.05f, 0.1f, 0.3f, 0, 0};
    CHECK(smk_set_physics(g, &ps));
    smk_mg_params mg = {2, 2, 80, 4};
    CHECK(smk_set_mg_params(g, &mg));

    /* emitters.c:52-83 layout: pos+radius, density+temperature, settings */
    const int r = nx / 10 > 3 ? nx / 10 : 3;
    smk_emitter em = {{nx / 2, ny / 8 + r, dim3 ? nz / 2 : 0, r}, {1, 0, 0, 0}, {0, 0, 1, 0}};
    CHECK(smk_set_emitters(g, &em, 1));
    float inflow = 2.0f;
    CHECK(smk_set_emitter_inflow(g, &inflow, 1));
    smk_obstacle ob = {{nx / 2.0f, ny * 0.6f, dim3 ? nz / 2.0f : 0.0f, (float)r}, {0, 0, 0, 1}};
    CHECK(smk_set_obstacles(g, &ob, 1));
    CHECK(smk_update_solid_map(g));
    CHECK(smk_set_graph_mode(g, 1));

    /* main.c:114-116 run_physics_step per frame, main.c:133-135 time */
    CHECK(smk_run_steps(g, 3, 0.0f, 0.1f));
    CHECK(smk_sync(g));
    double t0 = now_ms();
    CHECK(smk_run_steps(g, steps, 0.3f, 0.1f));
    CHECK(smk_sync(g));
    double ms = (now_ms() - t0) / steps;

    size_t n = 0;
    CHECK(smk_field_count(g, SMK_FIELD_SMOKE, 0, &n));
    float *sm = (float *)malloc(n * sizeof(float));
    CHECK(smk_download_field(g, SMK_FIELD_SMOKE, 0, sm, n * sizeof(float), SMK_DTYPE_F32));
    double dens = 0.0;
    for (size_t i = 0; i < n; i += 4) dens += sm[i];
    free(sm);
    CHECK(smk_field_count(g, SMK_FIELD_VY, 0, &n));
    float *vy = (float *)malloc(n * sizeof(float));
    CHECK(smk_download_field(g, SMK_FIELD_VY, 0, vy, n * sizeof(float), SMK_DTYPE_F32));
    double vmax = 0.0;
    for (size_t i = 0; i < n; ++i) vmax = fmax(vmax, fabs(vy[i]));
    free(vy);

    const double cells = (double)nx * ny * (dim3 ? nz : 1);
    /* checksums at full precision: the density channel summed in texel
     * order in double, and the largest |vy| */
    printf("grid %dx%dx%d storage %d: %.3f ms/step, %.3f Gcell-steps/s, total density %.17g, max |vy| %.9g\n",
           nx, ny, dim3 ? nz : 1, (int)prec, ms, cells / (ms * 1e-3) / 1e9, dens, vmax);
    CHECK(smk_grid_destroy(g));
    return 0;
}
