// smk_runtime.cpp -- host runtime behind include/smoke_abi.h.
//
// Replaces the host half of SmokeSim's step: the grid/texture bookkeeping of
// 2d/grid2d.c and 3d/grid3d.c, the stage dispatchers and V-cycle recursion of
// common/physics/grid.c, the step order of physics.c:72-101 and the scene
// uploads of emitters.c / obstacles.c.  GL bindings become explicit kernel
// arguments; ping-pong texture pairs become pairs of device buffers with a
// "current" index per field.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/smoke_abi.h"
#include "smk_kernels.h"

using namespace smk;

namespace {

thread_local std::string g_err;

smk_status fail(smk_status s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
smk_status fail(smk_status s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(SMK_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                               \
    } while (0)

constexpr int kMaxLevels = 16;

struct StepState {
    int vcur[3];
    int scur;
    int obcur;
    bool pyr_dirty;
    bool raster_valid;
    int p64cur;  // current buffer of the FP64 refined pressure
};

}  // namespace

struct smk_grid {
    int dev = 0;
    int D = 3;
    int n[3] = {0, 0, 0};
    float h = 0.1f, tamb = 293.f;
    int vfmt = FMT_F32, sfmt = FMT_F16;
    int min_level = 4, enclosed = 1;
    hipStream_t own = nullptr, stream = nullptr;

    // One device allocation holds every field and level buffer below (v,
    // smoke, vort, p/r/s/mask per level, dv0, b0, flag0), each piece aligned
    // to 2 MiB (4 KiB when smaller).  Grids created one after another then
    // get the same memory back from the runtime in one piece; with one
    // hipMalloc per buffer a later grid ran 8 % slower than the first at
    // 256^3 on the same addresses.
    char *arena = nullptr;
    size_t arena_bytes = 0;
    void *v[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    uint16_t *smoke[2] = {nullptr, nullptr};
    void *vort = nullptr;
    int L = 0;
    int ls[kMaxLevels][3];
    float *p[kMaxLevels] = {};
    float *r[kMaxLevels] = {};
    uint16_t *s[kMaxLevels] = {};
    uint8_t *mask[kMaxLevels] = {};
    TC *ptab[kMaxLevels][3] = {};  // prolongation coordinates (level l from l+1)
    bool prolong_even[kMaxLevels] = {};  // tables fit the 3x3x3 coarse-cell kernel
    int bottom = -1;               // first level solved in LDS, -1: none
    int *bottom_list = nullptr;    // colour-sorted cells of the bottom levels
    int bottom_n0[kBottomMaxLevels] = {};  // parity-0 cells per bottom level
    float *dv0 = nullptr;
    float *b0 = nullptr;  // level-0 smoother term rho*h*dv0/dt
    uint8_t *flag0 = nullptr;  // level-0 solid flags (a >= 1)
    TC *ctab = nullptr;        // face-centred interpolation coordinates
    int ctab_stride = 0;
    float *d_time = nullptr;
    double *d_acc = nullptr;    // two doubles: a norm and its companion sum
    double *d_rpart = nullptr;  // per-block partial sums of smk_residual_norm
    size_t rpart_cap = 0;
    StepState st{{0, 0, 0}, 0, 0, true, false, 0};
    // FP64 iterative refinement of the pressure solve
    int refine = 0;            // correction V-cycles per solve (0: the plain solve)
    bool coarse64 = false;     // coarsest level smoothed in FP64
    bool corr = false;         // inside a correction V-cycle
    bool skip_r0 = false;      // this V-cycle need not store the level-0 residual image
    double *p64[2] = {nullptr, nullptr};
    double *d_norms = nullptr;  // per residual pass: |r|^2, |b|^2
    double *d_part = nullptr;   // per-block partial norms of one pass
    int norms_cap = 0;
    float toff = 0.f;          // BF16 smoke stores T - toff
    int tile_bx = 0;           // launch tile width (0: the default heuristic)
    int sampler = 0;           // 1: clamped global-load samplers on any grid
    int arith = 0;             // SMK_ARITH_LLVMPIPE: the smk::gl launchers
    unsigned gen = 0;          // bumped whenever the captured step changes

    smk_physics_settings phys{};
    // div_rn() checks of the constant divisors, refreshed by refresh_div()
    // whenever the physics settings change (grid creation included)
    int fdiv = 0, fdiv_res = 0;
    smk_step_settings tog{1, 1, 1, 1, 1, 1, 1};
    smk_mg_params mg{16, 16, 80, 1};

    std::vector<smk_emitter> em;
    std::vector<float> inflow;
    int *d_em = nullptr;
    float *d_inflow = nullptr;
    int em_cap = 0;
    bool has_inflow = false;
    Box3 embox{{0, 0, 0}, {0, 0, 0}};

    std::vector<smk_obstacle> ob;
    float *d_ob[2] = {nullptr, nullptr};
    int ob_cap = 0;
    std::vector<smk_box> box;
    int *d_box = nullptr;
    int box_cap = 0;

    bool graph_mode = false;
    std::map<std::string, std::pair<hipGraphExec_t, StepState>> graphs;

    bool prof = false;
    std::vector<hipEvent_t> ev[SMK_K_COUNT][2];
    int ev_used[SMK_K_COUNT] = {};
};

// Whether div_rn() (q = RN(x*r), one Markstein correction) equals IEEE
// division by y for every significand of x in [1, 2).  The quotient and
// every intermediate scale exactly with x's exponent while they stay normal
// (div_rn restricts |x| to [2^-60, 2^60]), and the sequence is odd in x, so
// this one binade covers every x the fast path accepts.  About 10 ms per
// divisor significand; results are cached for the process (at most 4096).
bool smk::fast_div_ok(float y) {
    static std::mutex mu;
    static std::map<uint32_t, bool> cache;
    uint32_t key;
    std::memcpy(&key, &y, 4);
    bool ok = std::isfinite(y) && std::fabs(y) >= 0x1p-30f && std::fabs(y) <= 0x1p30f;
    // only a power-of-two multiple of y needs a check of its own, so key
    // the cache on the significand (the exponent is in range by the test
    // above)
    uint32_t sig = key & 0x807fffffu;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(sig);
        if (it != cache.end()) return ok && it->second;
    }
    const float r = (float)(1.0 / (double)y);  // RN(1/y): no double rounding for float y
    for (uint32_t m = 0; ok && m < (1u << 23); ++m) {
        const uint32_t u = 0x3f800000u | m;
        float x;
        std::memcpy(&x, &u, 4);
        const float q = x * r;
        const float rem = std::fmaf(-q, y, x);
        const float q2 = std::fmaf(rem, r, q);
        const float ref = x / y;
        ok = std::memcmp(&q2, &ref, 4) == 0;
    }
    std::lock_guard<std::mutex> lock(mu);
    // bounded: a caller that changes dt every step adds a divisor per step
    if (cache.size() >= 4096) cache.clear();
    if (std::isfinite(y) && std::fabs(y) >= 0x1p-30f && std::fabs(y) <= 0x1p30f) cache[sig] = ok;
    return ok;
}

// The launcher `fn` of the grid's arithmetic mode: smk::fn, or smk::gl::fn
// (llvmpipe-compatible binary16 rounding and lerp).
#define KL(g, fn) ((g)->arith ? &smk::gl::fn : &smk::fn)

namespace {

SimParams params(const smk_grid *g) {
    SimParams P;
    P.n[0] = g->n[0];
    P.n[1] = g->n[1];
    P.n[2] = g->n[2];
    P.dim = g->D;
    P.h = g->h;
    P.dt = g->phys.time_step;
    P.rho = g->phys.density;
    P.tamb = g->phys.t_ambient - g->toff;
    P.alpha = g->phys.t_weight_coeff;
    P.beta = g->phys.t_temp_coeff;
    P.eps = g->phys.vorticity_confinement;
    for (int c = 0; c < 4; ++c) P.wind[c] = g->phys.wind[c];
    P.corr = g->corr ? 1 : 0;
    P.tile_bx = g->tile_bx;
    P.gsamp = g->sampler;
    P.rh = (float)(1.0 / (double)P.h);
    P.rrho = (float)(1.0 / (double)P.rho);
    for (int d = 0; d < 3; ++d) {
        P.rn[d] = (float)(1.0 / (double)(float)P.n[d]);
        P.rn1[d] = (float)(1.0 / (double)((float)P.n[d] + 1.0f));
    }
    P.fdiv = g->fdiv;
    P.fdiv_res = g->fdiv_res;
    return P;
}

// The host checks of div_rn() for every constant divisor the kernels use:
// cell size, density, n and n+1 per axis, and the residual's dt*h and h*h.
void refresh_div(smk_grid *g) {
    bool ok = fast_div_ok(g->h) && fast_div_ok(g->phys.density);
    for (int d = 0; d < g->D; ++d)
        ok = ok && fast_div_ok((float)g->n[d]) && fast_div_ok((float)g->n[d] + 1.0f);
    g->fdiv = ok ? 1 : 0;
    const float db = g->phys.time_step * g->h, dh2 = g->h * g->h;
    g->fdiv_res = fast_div_ok(db) && fast_div_ok(dh2) ? 1 : 0;
}

size_t cells(const smk_grid *g) { return (size_t)g->n[0] * g->n[1] * g->n[2]; }
size_t level_cells(const smk_grid *g, int l) {
    return (size_t)g->ls[l][0] * g->ls[l][1] * g->ls[l][2];
}
// colour-split storage size of a level (two colours of padded rows)
// Slots per colour-split row of a level nx cells wide: ceil(nx/2) rounded up
// to a multiple of four, so every row starts 16-byte aligned and every level
// can take the four-cell kernels (k_sweep4 reads and writes rows in float4);
// the extra slots are padding (mask byte 0x01: solid, skipped everywhere).
int split_hx(int nx) { return ((nx + 1) / 2 + 3) & ~3; }
size_t level_split(const smk_grid *g, int l) {
    return 2 * (size_t)split_hx(g->ls[l][0]) * g->ls[l][1] * g->ls[l][2];
}
size_t vcount(const smk_grid *g, int c) {
    return (size_t)(g->n[0] + (c == 0)) * (g->n[1] + (c == 1)) * (g->n[2] + (g->D == 3 && c == 2));
}
size_t vbytes(const smk_grid *g) { return g->vfmt == FMT_F32 ? 4 : 2; }

VelPtrs vel_cur(const smk_grid *g) {
    VelPtrs V{{nullptr, nullptr, nullptr}};
    for (int c = 0; c < g->D; ++c) V.c[c] = g->v[c][g->st.vcur[c]];
    return V;
}
VelPtrs vel_other(const smk_grid *g) {
    VelPtrs V{{nullptr, nullptr, nullptr}};
    for (int c = 0; c < g->D; ++c) V.c[c] = g->v[c][1 - g->st.vcur[c]];
    return V;
}
void flip_all(smk_grid *g) {
    for (int c = 0; c < g->D; ++c) g->st.vcur[c] ^= 1;
}
CoordTab coord_tab(const smk_grid *g) { return CoordTab{g->ctab, g->ctab_stride}; }

Level level(const smk_grid *g, int l) {
    Level L;
    L.n[0] = g->ls[l][0];
    L.n[1] = g->ls[l][1];
    L.n[2] = g->ls[l][2];
    L.hx = split_hx(g->ls[l][0]);
    L.p = g->p[l];
    L.r = g->r[l];
    L.s = g->s[l];
    L.mask = g->mask[l];
    return L;
}

smk_status check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SMK_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return SMK_OK;
}

// Profiling brackets around instrumented launches.  attach: the launcher
// hands the events to the dispatch itself (start(), stop() into
// hipExtLaunchKernelGGL), which stamps them at the kernel's own start and
// end; otherwise they are recorded on the stream around the launch, which
// adds the command processor's time for the two event packets.
struct ProfScope {
    smk_grid *g;
    int id;
    int slot = -1;
    bool attach = false;
    hipEvent_t start() const { return slot >= 0 && attach ? g->ev[id][0][slot] : nullptr; }
    hipEvent_t stop() const { return slot >= 0 && attach ? g->ev[id][1][slot] : nullptr; }
    ProfScope(smk_grid *g_, int id_, bool attach_ = false) : g(g_), id(id_), attach(attach_) {
        if (!g->prof || id < 0) return;
        slot = g->ev_used[id]++;
        while ((int)g->ev[id][0].size() <= slot) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
                slot = -1;
                return;
            }
            g->ev[id][0].push_back(a);
            g->ev[id][1].push_back(b);
        }
        if (!attach) (void)hipEventRecord(g->ev[id][0][slot], g->stream);
    }
    ~ProfScope() {
        if (slot >= 0 && !attach) (void)hipEventRecord(g->ev[id][1][slot], g->stream);
    }
};

void invalidate_graphs(smk_grid *g) {
    for (auto &kv : g->graphs) (void)hipGraphExecDestroy(kv.second.first);
    g->graphs.clear();
    ++g->gen;
}

// Union of the emitters' cell boxes (cells and the vy faces with the same
// index), clipped to the grid.
void compute_embox(smk_grid *g) {
    Box3 b{{1 << 30, 1 << 30, 1 << 30}, {-1, -1, -1}};
    for (const auto &e : g->em) {
        int r = e.pos_radius[3];
        for (int d = 0; d < 3; ++d) {
            int lo = e.pos_radius[d] - r, hi = e.pos_radius[d] + r + 1;
            b.lo[d] = std::min(b.lo[d], lo);
            b.hi[d] = std::max(b.hi[d], hi);
        }
    }
    int lim[3] = {g->n[0], g->n[1] + 1, g->n[2]};
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = std::max(b.lo[d], 0);
        b.hi[d] = std::min(b.hi[d], lim[d]);
    }
    if (g->D == 2) {
        b.lo[2] = 0;
        b.hi[2] = 1;
    }
    if (g->em.empty()) b = Box3{{0, 0, 0}, {0, 0, 0}};
    g->embox = b;
}

bool all_anchored(const smk_grid *g) {
    for (const auto &o : g->ob)
        if (!(o.velocities_anchored[3] == 1.0f)) return false;
    return true;
}

template <class T>
smk_status ensure_dev(T **ptr, int *cap, int need) {
    if (need <= *cap && *ptr) return SMK_OK;
    if (*ptr) HIPCHK(hipFree(*ptr));
    *ptr = nullptr;
    int c = std::max(need, 1);
    HIPCHK(hipMalloc(ptr, sizeof(T) * (size_t)c));
    *cap = c;
    return SMK_OK;
}

// ---------------------------------------------------------------- stages
smk_status st_emitters(smk_grid *g) {
    int ne = (int)g->em.size();
    if (ne == 0) return SMK_OK;
    SimParams P = params(g);
    launch_emitter_status(g->stream, g->d_em, ne, g->d_time);
    KL(g, launch_emit)(g->stream, P, g->vfmt, g->sfmt, g->smoke[g->st.scur], g->v[1][g->st.vcur[1]],
                       g->d_em, g->has_inflow ? g->d_inflow : nullptr, ne, g->embox);
    return check_launch("emit");
}

smk_status st_advect(smk_grid *g, bool fuse_buoy) {
    SimParams P = params(g);
    {
        ProfScope ps(g, SMK_K_ADVECT_VEL);
        KL(g, launch_advect_velocity)(g->stream, P, g->vfmt, g->sfmt, vel_cur(g), vel_other(g), g->flag0,
                                      coord_tab(g), g->smoke[g->st.scur], fuse_buoy ? 1 : 0);
    }
    flip_all(g);
    return check_launch("advect_velocity");
}

smk_status st_buoyancy(smk_grid *g) {
    SimParams P = params(g);
    KL(g, launch_buoyancy)(g->stream, P, g->vfmt, g->sfmt, g->v[1][g->st.vcur[1]],
                    g->v[1][1 - g->st.vcur[1]], g->flag0, g->smoke[g->st.scur]);
    g->st.vcur[1] ^= 1;
    return check_launch("buoyancy");
}

smk_status st_vorticity(smk_grid *g) {
    SimParams P = params(g);
    KL(g, launch_vorticity)(g->stream, P, g->vfmt, vel_cur(g), g->flag0, g->vort);
    return check_launch("vorticity");
}

smk_status st_apply_vorticity(smk_grid *g) {
    SimParams P = params(g);
    launch_apply_vorticity(g->stream, P, g->vfmt, vel_cur(g), vel_other(g), g->vort);
    flip_all(g);
    return check_launch("apply_vorticity");
}

// right-hand side as the residual / LDS kernels read it (dv0 on level 0)
// (in a correction V-cycle of the refined mode level 0 reads the rounded
// FP64 residual, kept in b0)
const float *level_rhs(const smk_grid *g, int l) {
    return l == 0 ? (g->corr ? g->b0 : g->dv0) : g->r[l];
}
// right-hand side of the sweep kernels (the precomputed term on level 0)
const float *sweep_rhs(const smk_grid *g, int l) { return l == 0 ? g->b0 : g->r[l]; }

void smooth_level(smk_grid *g, int l, int iterations) {
    if (iterations <= 0) return;
    SimParams P = params(g);
    Level Lv = level(g, l);
    if (coarse_fits(P, Lv) || (l == g->L - 1 && coarsest_fits(Lv))) {
        launch_coarse_smooth(g->stream, P, Lv, l, iterations, level_rhs(g, l));
        return;
    }
    int parity = 0;
    for (int i = 0; i < iterations * 2; ++i) {
        if (l == 0) {
            ProfScope ps(g, SMK_K_SMOOTH_L0, true);
            launch_sweep(g->stream, P, Lv, l, parity, sweep_rhs(g, l), ps.start(), ps.stop());
        } else {
            launch_sweep(g->stream, P, Lv, l, parity, sweep_rhs(g, l));
        }
        parity = 1 - parity;
    }
}

// grid.c:177-209
BottomLevels bottom_levels(const smk_grid *g) {
    BottomLevels B{};
    B.nb = g->L - g->bottom;
    B.list = g->bottom_list;
    int off = 0;
    for (int m = 0; m < B.nb; ++m) {
        int l = g->bottom + m;
        for (int d = 0; d < 3; ++d) B.n[m][d] = g->ls[l][d];
        B.hx[m] = split_hx(g->ls[l][0]);
        B.off[m] = off;
        off += (int)level_cells(g, l);
        B.hl[m] = g->h * (float)(1 << l);
        B.p[m] = g->p[l];
        B.r[m] = g->r[l];
        B.s[m] = g->s[l];
        B.n0[m] = g->bottom_n0[m];
    }
    return B;
}

// First level of the pyramid whose remaining levels all fit one workgroup's
// LDS (k_bottom_vcycle); level 0 always stays on the multi-workgroup path.
int choose_bottom(const smk_grid *g) {
    for (int lb = 1; lb < g->L; ++lb) {
        if (g->L - lb > kBottomMaxLevels) continue;
        bool fits = true;
        size_t total = 0;
        for (int l = lb; l < g->L; ++l) {
            fits &= level_cells(g, l) <= (size_t)kBottomMaxCells;
            total += level_cells(g, l);
        }
        if (fits && (total + 1) * 3 * sizeof(float) <= kBottomMaxLds) return lb;
    }
    return -1;
}

void v_cycle(smk_grid *g, int l) {
    SimParams P = params(g);
    if (l == g->bottom) {
        KL(g, launch_bottom_vcycle)(g->stream, P, bottom_levels(g), g->mg.pre_iterations,
                             g->mg.post_iterations, g->mg.coarse_iterations,
                             g->corr && g->coarse64);
        return;
    }
    if (l == g->L - 1) {
        smooth_level(g, l, g->mg.coarse_iterations);
        return;
    }
    smooth_level(g, l, g->mg.pre_iterations);
    {
        // compute_residual + compute_restrict + coarse clear (grid.c:193-200)
        ProfScope ps(g, l == 0 ? SMK_K_RESIDUAL_L0 : -1);
        launch_residual_restrict(g->stream, P, level(g, l), level(g, l + 1), l, level_rhs(g, l),
                                 g->vfmt, vel_cur(g), g->s[0], !g->skip_r0);
    }
    v_cycle(g, l + 1);
    KL(g, launch_prolong)(g->stream, P, level(g, l), level(g, l + 1), g->ptab[l][0], g->ptab[l][1],
                   g->ptab[l][2], g->prolong_even[l]);
    smooth_level(g, l, g->mg.post_iterations);
}

// FP64 iterative refinement: residual pass with b = bnum*bsrc/bden, then
// `cycles` times a correction V-cycle (level-0 p = e from zero, right-hand
// side = the rounded residual) followed by p64 += e and the next residual
// pass.
void refine_loop(smk_grid *g, const SimParams &P, const float *bsrc, double bnum, double bden,
                 int cycles) {
    const Level L0 = level(g, 0);
    int &cur = g->st.p64cur;
    launch_refine_residual(g->stream, P, L0, g->p64[cur], g->p64[1 - cur], g->p[0], bsrc, bnum, bden,
                           g->b0, g->d_part, g->d_norms, 0);
    cur ^= 1;
    for (int it = 1; it <= cycles; ++it) {
        launch_fill_f32(g->stream, g->p[0], level_split(g, 0), 0.f);
        g->corr = true;
        g->skip_r0 = it < cycles;  // only the last cycle's residual image survives
        v_cycle(g, 0);
        g->skip_r0 = false;
        g->corr = false;
        launch_refine_residual(g->stream, P, L0, g->p64[cur], g->p64[1 - cur], g->p[0], bsrc, bnum,
                               bden, g->b0, g->d_part, g->d_norms + 2 * it, 1);
        cur ^= 1;
    }
}

void build_masks_if_dirty(smk_grid *g, const SimParams &P) {
    if (!g->st.pyr_dirty) return;
    // The reference re-restricts the solid masks in every V-cycle
    // (grid.c:254-256); they only change when level 0 changes.
    for (int l = 0; l + 1 < g->L; ++l) KL(g, launch_restrict_solid)(g->stream, P, level(g, l), level(g, l + 1));
    for (int l = 0; l < g->L; ++l) launch_build_mask(g->stream, P, level(g, l));
    g->st.pyr_dirty = false;
}

smk_status st_pressure(smk_grid *g) {
    SimParams P = params(g);
    build_masks_if_dirty(g, P);
    launch_divergence0(g->stream, P, g->vfmt, vel_cur(g), g->s[0], level(g, 0), g->dv0, g->b0);
    if (g->refine <= 0) {
        // the level-0 residual image is observable state: only the last
        // V-cycle's survives the solve, so the earlier ones skip storing it
        for (int v = 0; v < g->mg.v_cycles; ++v) {
            g->skip_r0 = v + 1 < g->mg.v_cycles;
            v_cycle(g, 0);
        }
        g->skip_r0 = false;
        return check_launch("pressure_solve");
    }
    refine_loop(g, P, g->dv0, (double)P.rho, (double)P.dt * (double)P.h, g->refine);
    return check_launch("pressure_solve");
}

smk_status st_update_velocities(smk_grid *g) {
    SimParams P = params(g);
    build_masks_if_dirty(g, P);  // the update reads the level-0 neighbour masks
    launch_update_velocity(g->stream, P, g->vfmt, vel_cur(g), vel_other(g), g->s[0], g->flag0,
                           level(g, 0), g->refine > 0 ? g->p64[g->st.p64cur] : nullptr);
    flip_all(g);
    return check_launch("update_velocity");
}

smk_status st_advect_smoke(smk_grid *g) {
    SimParams P = params(g);
    {
        ProfScope ps(g, SMK_K_ADVECT_SMOKE);
        KL(g, launch_advect_smoke)(g->stream, P, g->vfmt, g->sfmt, vel_cur(g), g->flag0, coord_tab(g),
                            g->smoke[g->st.scur], g->smoke[1 - g->st.scur]);
    }
    g->st.scur ^= 1;
    return check_launch("advect_smoke");
}

smk_status st_obstacles(smk_grid *g) {
    int no = (int)g->ob.size();
    if (no > 0) {
        SimParams P = params(g);
        launch_obstacle_update(g->stream, P, g->d_ob[g->st.obcur], g->d_ob[1 - g->st.obcur], no);
    }
    g->st.obcur ^= 1;  // grid.c:436-440 swaps even with no obstacles
    return check_launch("obstacle_update");
}

// update_solid_map: rasterisation of a set of anchored obstacles and static
// boxes does not change from step to step, so it is re-run only when the
// obstacle set changes or an obstacle can move.
smk_status st_update_solid(smk_grid *g, bool force) {
    if (!force && g->st.raster_valid && all_anchored(g)) return SMK_OK;
    SimParams P = params(g);
    KL(g, launch_update_solid)(g->stream, P, g->s[0], g->d_ob[g->st.obcur], (int)g->ob.size(), g->d_box,
                        (int)g->box.size());
    launch_solid_flag(g->stream, g->s[0], g->flag0, cells(g));
    g->st.raster_valid = all_anchored(g);
    g->st.pyr_dirty = true;
    return check_launch("update_solid_map");
}

bool default_toggles(const smk_grid *g) {
    const smk_step_settings &t = g->tog;
    return t.advect_smoke && t.advect_velocities && t.apply_vorticity && t.apply_buoyancy &&
           t.resolve_pressure && t.update_velocities && t.handle_emitters;
}

// physics.c:72-101
smk_status step_impl(smk_grid *g) {
    smk_status s;
    const smk_step_settings &t = g->tog;
    if (t.handle_emitters && (s = st_emitters(g)) != SMK_OK) return s;
    if (t.advect_velocities && (s = st_advect(g, t.apply_buoyancy != 0)) != SMK_OK) return s;
    if (!t.advect_velocities && t.apply_buoyancy && (s = st_buoyancy(g)) != SMK_OK) return s;
    if (t.apply_vorticity) {
        if ((s = st_vorticity(g)) != SMK_OK) return s;
        if ((s = st_apply_vorticity(g)) != SMK_OK) return s;
    }
    if (t.resolve_pressure && (s = st_pressure(g)) != SMK_OK) return s;
    if (t.update_velocities && (s = st_update_velocities(g)) != SMK_OK) return s;
    if (t.advect_smoke && (s = st_advect_smoke(g)) != SMK_OK) return s;
    if ((s = st_obstacles(g)) != SMK_OK) return s;
    return st_update_solid(g, false);
}

std::string state_key(const smk_grid *g) {
    char buf[128];
    snprintf(buf, sizeof(buf), "%d%d%d%d%d%d%d%d%d", g->st.vcur[0], g->st.vcur[1], g->st.vcur[2],
             g->st.scur, g->st.obcur, (int)g->st.pyr_dirty, (int)g->st.raster_valid, (int)g->prof,
             g->st.p64cur);
    return buf;
}

smk_status step_graph(smk_grid *g) {
    std::string key = state_key(g);
    auto it = g->graphs.find(key);
    if (it == g->graphs.end()) {
        StepState before = g->st;
        HIPCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        smk_status s = step_impl(g);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(g->stream, &graph);
        if (s != SMK_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            g->st = before;
            return s;
        }
        if (e != hipSuccess) {
            g->st = before;
            return fail(SMK_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g->st = before;
            return fail(SMK_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
        }
        it = g->graphs.emplace(key, std::make_pair(exec, g->st)).first;
        g->st = before;
    }
    HIPCHK(hipGraphLaunch(it->second.first, g->stream));
    g->st = it->second.second;
    return SMK_OK;
}

// Field table: device pointer, storage format, element count.
constexpr int kFmtF64Split = -1;  // the FP64 refined pressure
struct FieldRef {
    void *ptr;
    int fmt;
    size_t count;
    int split_level;  // >= 0: colour-split level field
};

smk_status field_ref(smk_grid *g, int field, int lvl, FieldRef *out) {
    switch (field) {
    case SMK_FIELD_VX:
    case SMK_FIELD_VY:
    case SMK_FIELD_VZ: {
        int c = field - SMK_FIELD_VX;
        if (c >= g->D) return fail(SMK_ERR_INVALID_ARG, "2D grid has no vz");
        *out = FieldRef{g->v[c][g->st.vcur[c]], g->vfmt, vcount(g, c), -1};
        return SMK_OK;
    }
    case SMK_FIELD_SMOKE:
        *out = FieldRef{g->smoke[g->st.scur], g->sfmt, cells(g) * 4, -1};
        return SMK_OK;
    case SMK_FIELD_VORTICITY:
        *out = g->D == 3 ? FieldRef{g->vort, FMT_F16, cells(g) * 4, -1}
                         : FieldRef{g->vort, FMT_F32, cells(g), -1};
        return SMK_OK;
    case SMK_FIELD_PRESSURE:
    case SMK_FIELD_RESIDUAL:
    case SMK_FIELD_SOLID:
        if (lvl < 0 || lvl >= g->L) return fail(SMK_ERR_INVALID_ARG, "level %d out of range", lvl);
        if (field == SMK_FIELD_PRESSURE && lvl == 0 && g->refine > 0)
            *out = FieldRef{g->p64[g->st.p64cur], kFmtF64Split, level_cells(g, 0), 0};
        else if (field == SMK_FIELD_PRESSURE) *out = FieldRef{g->p[lvl], FMT_F32, level_cells(g, lvl), lvl};
        else if (field == SMK_FIELD_RESIDUAL) *out = FieldRef{g->r[lvl], FMT_F32, level_cells(g, lvl), lvl};
        else *out = FieldRef{g->s[lvl], FMT_F16, level_cells(g, lvl) * 4, -1};
        return SMK_OK;
    case SMK_FIELD_DIVERGENCE:
        *out = FieldRef{nullptr, FMT_F32, cells(g), -1};
        return SMK_OK;
    default:
        return fail(SMK_ERR_INVALID_ARG, "unknown field %d", field);
    }
}

#define GRID_CHECK(g)                                                       \
    do {                                                                    \
        if (!(g)) return fail(SMK_ERR_INVALID_ARG, "null grid handle");     \
        HIPCHK(hipSetDevice((g)->dev));                                     \
    } while (0)

void free_all(smk_grid *g) {
    // the field and level buffers are pieces of the arena
    if (g->arena) (void)hipFree(g->arena);
    g->arena = nullptr;
    for (int l = 0; l < g->L; ++l)
        for (int d = 0; d < 3; ++d)
            if (g->ptab[l][d]) (void)hipFree(g->ptab[l][d]);
    if (g->ctab) (void)hipFree(g->ctab);
    if (g->bottom_list) (void)hipFree(g->bottom_list);
    if (g->d_time) (void)hipFree(g->d_time);
    if (g->d_acc) (void)hipFree(g->d_acc);
    if (g->d_rpart) (void)hipFree(g->d_rpart);
    for (int b = 0; b < 2; ++b)
        if (g->p64[b]) (void)hipFree(g->p64[b]);
    if (g->d_norms) (void)hipFree(g->d_norms);
    if (g->d_part) (void)hipFree(g->d_part);
    if (g->d_em) (void)hipFree(g->d_em);
    if (g->d_inflow) (void)hipFree(g->d_inflow);
    for (int b = 0; b < 2; ++b)
        if (g->d_ob[b]) (void)hipFree(g->d_ob[b]);
    if (g->d_box) (void)hipFree(g->d_box);
    invalidate_graphs(g);
    for (int k = 0; k < SMK_K_COUNT; ++k)
        for (int b = 0; b < 2; ++b)
            for (auto e : g->ev[k][b]) (void)hipEventDestroy(e);
    if (g->own) (void)hipStreamDestroy(g->own);
}

}  // namespace

// Smoke texture clear to (0,0,0,t_ambient) (3d/grid3d.c:35-38): a host
// pattern converted to the storage format on the device.
static smk_status fill_smoke(smk_grid *g, int b) {
    size_t n = cells(g) * 4;
    std::vector<double> host(n, 0.0);
    for (size_t i = 3; i < n; i += 4) host[i] = (double)(g->tamb - g->toff);
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, n * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, host.data(), n * sizeof(double), hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess) {
        launch_from_f64(g->stream, d, g->smoke[b], g->sfmt, n);
        e = hipStreamSynchronize(g->stream);
    }
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SMK_ERR_HIP, "smoke fill: %s", hipGetErrorString(e));
    return SMK_OK;
}

extern "C" {

const char *smk_last_error(void) { return g_err.c_str(); }
int smk_abi_version(void) { return SMK_ABI_VERSION; }

smk_status smk_grid_create(const smk_grid_desc *desc, smk_grid **out) {
    if (!desc || !out) return fail(SMK_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const int *sz = desc->size;
    int D = sz[2] == 0 ? 2 : 3;
    if (sz[0] < 1 || sz[1] < 1 || sz[2] < 0)
        return fail(SMK_ERR_INVALID_ARG, "invalid size %d x %d x %d", sz[0], sz[1], sz[2]);
    if (sz[0] > 65536 || sz[1] > 65536 || sz[2] > 65536)
        return fail(SMK_ERR_INVALID_ARG, "size above 65536 (grid.h:18 MAX_GRID_SIZE)");
    // kernels index fields with 32-bit element offsets
    if ((double)(sz[0] + 1) * (sz[1] + 1) * (sz[2] + 1) >= 4294967296.0)
        return fail(SMK_ERR_INVALID_ARG, "grid of %d x %d x %d exceeds 2^32 faces", sz[0], sz[1], sz[2]);
    if (!(desc->cell_size > 0.0f)) return fail(SMK_ERR_INVALID_ARG, "cell_size must be > 0");
    if (desc->storage < SMK_PREC_FP32 || desc->storage > SMK_PREC_FP16)
        return fail(SMK_ERR_INVALID_ARG, "unknown storage %d", desc->storage);
    int minl = desc->min_level_size > 0 ? desc->min_level_size : 4;
    if (sz[0] < minl || sz[1] < minl || (D == 3 && sz[2] < minl))
        return fail(SMK_ERR_INVALID_ARG, "grid smaller than min_level_size %d", minl);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(SMK_ERR_NO_DEVICE, "no HIP device");
    if (desc->device < 0 || desc->device >= ndev)
        return fail(SMK_ERR_INVALID_ARG, "device %d out of range", desc->device);

    smk_grid *g = new (std::nothrow) smk_grid();
    if (!g) return fail(SMK_ERR_OUT_OF_MEMORY, "host allocation failed");
    g->dev = desc->device;
    g->D = D;
    g->n[0] = sz[0];
    g->n[1] = sz[1];
    g->n[2] = D == 3 ? sz[2] : 1;
    g->h = desc->cell_size;
    g->tamb = desc->t_ambient;
    g->vfmt = desc->storage == SMK_PREC_FP32 ? FMT_F32 : (desc->storage == SMK_PREC_BF16 ? FMT_BF16 : FMT_F16);
    g->sfmt = desc->storage == SMK_PREC_BF16 ? FMT_BF16 : FMT_F16;
    // BF16 cannot resolve temperatures near ambient (2 K steps at 293 K):
    // its smoke alpha channel holds T - t_ambient instead
    g->toff = g->sfmt == FMT_BF16 ? g->tamb : 0.f;
    g->min_level = minl;
    g->enclosed = desc->enclosed;
    // physics.c:7-20 defaults (t_ambient from the descriptor)
    g->phys = smk_physics_settings{{0, 0, 0, 0}, 1.0f / 120.0f, 1.0f, desc->t_ambient, 0.07f, 1.0f, 1.0f, 0, 0};
    refresh_div(g);

    // pyramid: levels while every dimension >= min_level (2d/grid2d.c:45-63)
    {
        int d[3] = {g->n[0], g->n[1], D == 3 ? g->n[2] : minl};
        g->L = 0;
        while (d[0] >= minl && d[1] >= minl && d[2] >= minl && g->L < kMaxLevels) {
            g->ls[g->L][0] = g->n[0] >> g->L;
            g->ls[g->L][1] = g->n[1] >> g->L;
            g->ls[g->L][2] = D == 3 ? g->n[2] >> g->L : 1;
            d[0] >>= 1;
            d[1] >>= 1;
            if (D == 3) d[2] >>= 1;
            ++g->L;
        }
    }

    auto cleanup = [&](smk_status s) {
        free_all(g);
        delete g;
        return s;
    };
#define ALLOC(ptr, bytes)                                                             \
    do {                                                                              \
        hipError_t e_ = hipMalloc((void **)&(ptr), (bytes));                          \
        if (e_ != hipSuccess)                                                         \
            return cleanup(fail(e_ == hipErrorOutOfMemory ? SMK_ERR_OUT_OF_MEMORY : SMK_ERR_HIP, \
                                "hipMalloc(%zu): %s", (size_t)(bytes), hipGetErrorString(e_)));   \
    } while (0)

    if (hipSetDevice(g->dev) != hipSuccess) return cleanup(fail(SMK_ERR_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&g->own, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(SMK_ERR_HIP, "stream create failed"));
    g->stream = g->own;
    {
        // the arena: sizes first, then one allocation, then the pieces in
        // the same order
        auto piece = [](size_t bytes) {
            const size_t a = bytes >= ((size_t)1 << 21) ? ((size_t)1 << 21) : 4096;
            return (bytes + a - 1) / a * a;
        };
        std::vector<size_t> sizes;
        for (int c = 0; c < D; ++c)
            for (int b = 0; b < 2; ++b) sizes.push_back(vcount(g, c) * vbytes(g));
        for (int b = 0; b < 2; ++b) sizes.push_back(cells(g) * 8);
        sizes.push_back(D == 3 ? cells(g) * 8 : cells(g) * 4);
        for (int l = 0; l < g->L; ++l) {
            sizes.push_back(level_split(g, l) * 4);
            sizes.push_back(level_split(g, l) * 4);
            sizes.push_back(level_cells(g, l) * 8);
            sizes.push_back(level_split(g, l));
        }
        sizes.push_back(level_split(g, 0) * 4);
        sizes.push_back(level_split(g, 0) * 4);
        sizes.push_back(cells(g));
        size_t total = 0;
        for (size_t b : sizes) total += piece(b);
        ALLOC(g->arena, total);
        g->arena_bytes = total;
        size_t at = 0, next = 0;
        auto take = [&]() {
            char *p = g->arena + at;
            at += piece(sizes[next++]);
            return (void *)p;
        };
        for (int c = 0; c < D; ++c)
            for (int b = 0; b < 2; ++b) g->v[c][b] = take();
        for (int b = 0; b < 2; ++b) g->smoke[b] = (uint16_t *)take();
        g->vort = take();
        for (int l = 0; l < g->L; ++l) {
            g->p[l] = (float *)take();
            g->r[l] = (float *)take();
            g->s[l] = (uint16_t *)take();
            g->mask[l] = (uint8_t *)take();
        }
        g->dv0 = (float *)take();
        g->b0 = (float *)take();
        g->flag0 = (uint8_t *)take();
    }
    for (int l = 0; l + 1 < g->L; ++l)
        for (int d = 0; d < 3; ++d) ALLOC(g->ptab[l][d], sizeof(TC) * (size_t)g->ls[l][d]);
    g->bottom = choose_bottom(g);
    if (g->bottom >= 0) {
        // colour-sorted cells of every bottom level (parity 0 first), at the
        // level's LDS offset: k_bottom_vcycle gives each thread cells of one
        // colour per half sweep
        std::vector<int> lst;
        for (int l = g->bottom; l < g->L; ++l) {
            const int sx = g->ls[l][0], sy = g->ls[l][1], sz = g->ls[l][2];
            int n0 = 0;
            for (int par = 0; par < 2; ++par)
                for (int k = 0; k < sz; ++k)
                    for (int j = 0; j < sy; ++j)
                        for (int i = 0; i < sx; ++i)
                            if (((i + j + k) & 1) == par) {
                                lst.push_back(i + sx * (j + sy * k));
                                n0 += par == 0;
                            }
            g->bottom_n0[l - g->bottom] = n0;
        }
        ALLOC(g->bottom_list, lst.size() * sizeof(int));
        if (hipMemcpy(g->bottom_list, lst.data(), lst.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(SMK_ERR_HIP, "bottom cell lists upload failed"));
    }
    g->ctab_stride = std::max(g->n[0], std::max(g->n[1], g->n[2])) + 1;
    ALLOC(g->ctab, sizeof(TC) * 12 * (size_t)g->ctab_stride);
    ALLOC(g->d_time, sizeof(float));
    ALLOC(g->d_acc, 2 * sizeof(double));
    ALLOC(g->d_ob[0], 32);
    ALLOC(g->d_ob[1], 32);
    g->ob_cap = 1;
    ALLOC(g->d_box, 24);
    g->box_cap = 1;
    ALLOC(g->d_em, 48);
    ALLOC(g->d_inflow, 4);
    g->em_cap = 1;
#undef ALLOC

    // zero everything, then smoke = (0,0,0,T_ambient) (texture.c:54-78 clear)
    for (int c = 0; c < D; ++c)
        for (int b = 0; b < 2; ++b) {
            if (g->vfmt == FMT_F32) launch_fill_f32(g->stream, (float *)g->v[c][b], vcount(g, c), 0.f);
            else launch_fill_u16(g->stream, (uint16_t *)g->v[c][b], vcount(g, c), 0);
        }
    if (D == 3) launch_fill_u16(g->stream, (uint16_t *)g->vort, cells(g) * 4, 0);
    else launch_fill_f32(g->stream, (float *)g->vort, cells(g), 0.f);
    for (int l = 0; l < g->L; ++l) {
        launch_fill_f32(g->stream, g->p[l], level_split(g, l), 0.f);
        launch_fill_f32(g->stream, g->r[l], level_split(g, l), 0.f);
        launch_fill_u16(g->stream, g->s[l], level_cells(g, l) * 4, 0);
    }
    launch_fill_f32(g->stream, g->dv0, level_split(g, 0), 0.f);
    launch_fill_f32(g->stream, g->b0, level_split(g, 0), 0.f);
    for (int l = 0; l < g->L; ++l)  // padding slots of odd rows stay defined
        launch_fill_u16(g->stream, (uint16_t *)g->mask[l], level_split(g, l) / 2, 0x0101);
    launch_set_f32(g->stream, g->d_time, 0.f);
    for (int b = 0; b < 2; ++b) {
        smk_status s = fill_smoke(g, b);
        if (s != SMK_OK) return cleanup(s);
    }
    if (g->enclosed) launch_init_solid(g->stream, params(g), g->s[0]);
    launch_solid_flag(g->stream, g->s[0], g->flag0, cells(g));
    launch_coord_table(g->stream, params(g), g->ctab, g->ctab_stride);
    for (int l = 0; l + 1 < g->L; ++l)
        for (int d = 0; d < 3; ++d)
            launch_prolong_table(g->stream, g->ls[l][d], g->ls[l + 1][d], g->ptab[l][d]);
    // The coarse-cell prolongation kernel needs fine cell 2I+d to sample
    // coarse texels starting at I-1+d on every axis.
    for (int l = 0; l + 1 < g->L; ++l) {
        bool ok = true;
        for (int d = 0; d < D; ++d) {
            int nf = g->ls[l][d];
            if (nf != 2 * g->ls[l + 1][d]) ok = false;
            if (!ok) break;
            std::vector<TC> t((size_t)nf);
            if (hipMemcpy(t.data(), g->ptab[l][d], sizeof(TC) * (size_t)nf, hipMemcpyDeviceToHost) != hipSuccess)
                return cleanup(fail(SMK_ERR_HIP, "prolongation table copy failed"));
            for (int i = 0; i < nf && ok; ++i) ok = t[(size_t)i].i0 == i / 2 - 1 + (i & 1);
        }
        g->prolong_even[l] = ok;
    }
    g->st.pyr_dirty = true;
    g->st.raster_valid = false;
    if (hipStreamSynchronize(g->stream) != hipSuccess)
        return cleanup(fail(SMK_ERR_HIP, "initialisation failed: %s", hipGetErrorString(hipGetLastError())));
    *out = g;
    return SMK_OK;
}

smk_status smk_grid_destroy(smk_grid *g) {
    if (!g) return SMK_OK;
    (void)hipSetDevice(g->dev);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    free_all(g);
    delete g;
    return SMK_OK;
}

smk_status smk_grid_clear(smk_grid *g) {
    GRID_CHECK(g);
    invalidate_graphs(g);
    for (int c = 0; c < g->D; ++c)
        for (int b = 0; b < 2; ++b) {
            if (g->vfmt == FMT_F32) launch_fill_f32(g->stream, (float *)g->v[c][b], vcount(g, c), 0.f);
            else launch_fill_u16(g->stream, (uint16_t *)g->v[c][b], vcount(g, c), 0);
        }
    for (int b = 0; b < 2; ++b) {
        smk_status s = fill_smoke(g, b);
        if (s != SMK_OK) return s;
    }
    if (g->D == 3) {
        launch_fill_u16(g->stream, (uint16_t *)g->vort, cells(g) * 4, 0);
        // 3d/grid3d.c:181-183 clears the pressure pyramid; 2D does not
        for (int l = 0; l < g->L; ++l) launch_fill_f32(g->stream, g->p[l], level_split(g, l), 0.f);
        for (int b = 0; b < 2; ++b)
            if (g->p64[b]) launch_fill_f32(g->stream, reinterpret_cast<float *>(g->p64[b]), 2 * level_split(g, 0), 0.f);
    } else {
        launch_fill_f32(g->stream, (float *)g->vort, cells(g), 0.f);
    }
    return check_launch("clear");
}

smk_status smk_grid_levels(smk_grid *g, int *levels) {
    if (!g || !levels) return fail(SMK_ERR_INVALID_ARG, "null argument");
    *levels = g->L;
    return SMK_OK;
}

smk_status smk_grid_level_size(smk_grid *g, int l, int size_out[3]) {
    if (!g || !size_out) return fail(SMK_ERR_INVALID_ARG, "null argument");
    if (l < 0 || l >= g->L) return fail(SMK_ERR_INVALID_ARG, "level %d out of range", l);
    for (int d = 0; d < 3; ++d) size_out[d] = g->ls[l][d];
    return SMK_OK;
}

smk_status smk_set_physics(smk_grid *g, const smk_physics_settings *s) {
    if (!g || !s) return fail(SMK_ERR_INVALID_ARG, "null argument");
    if (!(s->density != 0.0f) || !(s->time_step != 0.0f))
        return fail(SMK_ERR_INVALID_ARG, "density and time_step must be non-zero");
    g->phys = *s;
    refresh_div(g);
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_set_step_toggles(smk_grid *g, const smk_step_settings *s) {
    if (!g || !s) return fail(SMK_ERR_INVALID_ARG, "null argument");
    g->tog = *s;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_set_mg_params(smk_grid *g, const smk_mg_params *p) {
    if (!g || !p) return fail(SMK_ERR_INVALID_ARG, "null argument");
    if (p->pre_iterations < 0 || p->post_iterations < 0 || p->coarse_iterations < 0 || p->v_cycles < 0)
        return fail(SMK_ERR_INVALID_ARG, "negative multigrid parameter");
    g->mg = *p;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_set_emitters(smk_grid *g, const smk_emitter *e, int n) {
    GRID_CHECK(g);
    if (n < 0 || (n > 0 && !e)) return fail(SMK_ERR_INVALID_ARG, "bad emitter array");
    smk_status s;
    if ((s = ensure_dev(&g->d_em, &g->em_cap, n * 12)) != SMK_OK) return s;
    int fcap = (int)g->inflow.size();
    (void)fcap;
    g->em.assign(e, e + n);
    g->inflow.assign((size_t)n, 0.0f);
    g->has_inflow = false;
    if (g->d_inflow) HIPCHK(hipFree(g->d_inflow));
    g->d_inflow = nullptr;
    HIPCHK(hipMalloc(&g->d_inflow, sizeof(float) * (size_t)std::max(n, 1)));
    if (n > 0) {
        HIPCHK(hipMemcpyAsync(g->d_em, e, sizeof(smk_emitter) * (size_t)n, hipMemcpyHostToDevice, g->stream));
        HIPCHK(hipMemcpyAsync(g->d_inflow, g->inflow.data(), sizeof(float) * (size_t)n,
                              hipMemcpyHostToDevice, g->stream));
    }
    HIPCHK(hipStreamSynchronize(g->stream));
    compute_embox(g);
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_get_emitters(smk_grid *g, smk_emitter *e, int n) {
    GRID_CHECK(g);
    if (n < 0 || n > (int)g->em.size() || (n > 0 && !e)) return fail(SMK_ERR_INVALID_ARG, "bad emitter count");
    if (n == 0) return SMK_OK;
    HIPCHK(hipMemcpyAsync(e, g->d_em, sizeof(smk_emitter) * (size_t)n, hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return SMK_OK;
}

smk_status smk_set_emitter_inflow(smk_grid *g, const float *inflow, int n) {
    GRID_CHECK(g);
    if (n < 0 || n > (int)g->em.size() || (n > 0 && !inflow))
        return fail(SMK_ERR_INVALID_ARG, "inflow count %d exceeds emitter count %zu", n, g->em.size());
    for (int i = 0; i < n; ++i) g->inflow[i] = inflow[i];
    g->has_inflow = false;
    for (float f : g->inflow) g->has_inflow |= (f != 0.0f);
    if (!g->inflow.empty())
        HIPCHK(hipMemcpyAsync(g->d_inflow, g->inflow.data(), sizeof(float) * g->inflow.size(),
                              hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_set_obstacles(smk_grid *g, const smk_obstacle *o, int n) {
    GRID_CHECK(g);
    if (n < 0 || (n > 0 && !o)) return fail(SMK_ERR_INVALID_ARG, "bad obstacle array");
    if (n > g->ob_cap) {
        for (int b = 0; b < 2; ++b) {
            HIPCHK(hipFree(g->d_ob[b]));
            g->d_ob[b] = nullptr;
            HIPCHK(hipMalloc(&g->d_ob[b], sizeof(smk_obstacle) * (size_t)n));
        }
        g->ob_cap = n;
    }
    g->ob.assign(o, o + n);
    if (n > 0)
        for (int b = 0; b < 2; ++b)  // obstacles.c:67-87 uploads both buffers
            HIPCHK(hipMemcpyAsync(g->d_ob[b], o, sizeof(smk_obstacle) * (size_t)n, hipMemcpyHostToDevice,
                                  g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    g->st.raster_valid = false;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_get_obstacles(smk_grid *g, smk_obstacle *o, int n) {
    GRID_CHECK(g);
    if (n < 0 || n > (int)g->ob.size() || (n > 0 && !o)) return fail(SMK_ERR_INVALID_ARG, "bad obstacle count");
    if (n == 0) return SMK_OK;
    HIPCHK(hipMemcpyAsync(o, g->d_ob[g->st.obcur], sizeof(smk_obstacle) * (size_t)n, hipMemcpyDeviceToHost,
                          g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return SMK_OK;
}

smk_status smk_set_box_obstacles(smk_grid *g, const smk_box *b, int n) {
    GRID_CHECK(g);
    if (n < 0 || (n > 0 && !b)) return fail(SMK_ERR_INVALID_ARG, "bad box array");
    smk_status s;
    if ((s = ensure_dev(&g->d_box, &g->box_cap, n * 6)) != SMK_OK) return s;
    g->box.assign(b, b + n);
    if (n > 0)
        HIPCHK(hipMemcpyAsync(g->d_box, b, sizeof(smk_box) * (size_t)n, hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    g->st.raster_valid = false;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_set_time(smk_grid *g, float t) {
    GRID_CHECK(g);
    launch_set_f32(g->stream, g->d_time, t);
    return check_launch("set_time");
}

smk_status smk_update_emitters_status(smk_grid *g) {
    GRID_CHECK(g);
    launch_emitter_status(g->stream, g->d_em, (int)g->em.size(), g->d_time);
    return check_launch("emitter_status");
}

smk_status smk_handle_emitters(smk_grid *g) {
    GRID_CHECK(g);
    if (g->em.empty()) return SMK_OK;
    KL(g, launch_emit)(g->stream, params(g), g->vfmt, g->sfmt, g->smoke[g->st.scur], g->v[1][g->st.vcur[1]], g->d_em,
                g->has_inflow ? g->d_inflow : nullptr, (int)g->em.size(), g->embox);
    return check_launch("emit");
}

smk_status smk_advect_velocities(smk_grid *g) {
    GRID_CHECK(g);
    return st_advect(g, false);
}
smk_status smk_apply_buoyancy(smk_grid *g) {
    GRID_CHECK(g);
    return st_buoyancy(g);
}
smk_status smk_calculate_vorticity(smk_grid *g) {
    GRID_CHECK(g);
    return st_vorticity(g);
}
smk_status smk_apply_vorticity(smk_grid *g) {
    GRID_CHECK(g);
    return st_apply_vorticity(g);
}
smk_status smk_pressure_solve(smk_grid *g) {
    GRID_CHECK(g);
    return st_pressure(g);
}
smk_status smk_update_velocities(smk_grid *g) {
    GRID_CHECK(g);
    return st_update_velocities(g);
}
smk_status smk_advect_smoke(smk_grid *g) {
    GRID_CHECK(g);
    return st_advect_smoke(g);
}
smk_status smk_obstacle_update_step(smk_grid *g) {
    GRID_CHECK(g);
    invalidate_graphs(g);
    return st_obstacles(g);
}
smk_status smk_update_solid_map(smk_grid *g) {
    GRID_CHECK(g);
    return st_update_solid(g, true);
}
smk_status smk_init_solid_map(smk_grid *g) {
    GRID_CHECK(g);
    launch_init_solid(g->stream, params(g), g->s[0]);
    launch_solid_flag(g->stream, g->s[0], g->flag0, cells(g));
    g->st.pyr_dirty = true;
    return check_launch("init_solid_map");
}

// Profiled steps launch eagerly: HIP events recorded inside a captured
// graph cannot be timed.
smk_status smk_run_physics_step(smk_grid *g) {
    GRID_CHECK(g);
    if (g->graph_mode && !g->prof) return step_graph(g);
    return step_impl(g);
}

smk_status smk_run_steps(smk_grid *g, int n, float t0, float dt_time) {
    GRID_CHECK(g);
    if (n < 0) return fail(SMK_ERR_INVALID_ARG, "negative step count");
    for (int i = 0; i < n; ++i) {
        launch_set_f32(g->stream, g->d_time, t0 + (float)i * dt_time);
        smk_status s = (g->graph_mode && !g->prof) ? step_graph(g) : step_impl(g);
        if (s != SMK_OK) return s;
    }
    return SMK_OK;
}

smk_status smk_field_count(smk_grid *g, int field, int lvl, size_t *count) {
    if (!g || !count) return fail(SMK_ERR_INVALID_ARG, "null argument");
    FieldRef f;
    smk_status s = field_ref(g, field, lvl, &f);
    if (s != SMK_OK) return s;
    *count = f.count;
    return SMK_OK;
}

smk_status smk_download_field(smk_grid *g, int field, int lvl, void *host, size_t bytes, int dtype) {
    GRID_CHECK(g);
    FieldRef f;
    smk_status s = field_ref(g, field, lvl, &f);
    if (s != SMK_OK) return s;
    size_t esz = dtype == SMK_DTYPE_F64 ? 8 : (dtype == SMK_DTYPE_F32 ? 4 : 0);
    if (esz == 0) return fail(SMK_ERR_INVALID_ARG, "unknown dtype %d", dtype);
    if (!host || bytes != f.count * esz)
        return fail(SMK_ERR_INVALID_ARG, "buffer of %zu bytes, field needs %zu", bytes, f.count * esz);
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, f.count * sizeof(double)));
    if (field == SMK_FIELD_DIVERGENCE)
        launch_divergence_field(g->stream, params(g), g->vfmt, vel_cur(g), g->s[0], d);
    else if (f.fmt == kFmtF64Split)
        launch_split64_to_f64(g->stream, level(g, 0), (const double *)f.ptr, d);
    else if (f.split_level >= 0)
        launch_split_to_f64(g->stream, level(g, f.split_level), (const float *)f.ptr, d);
    else
        launch_to_f64(g->stream, f.ptr, f.fmt, d, f.count);
    std::vector<double> tmp;
    void *dst = host;
    if (dtype == SMK_DTYPE_F32) {
        tmp.resize(f.count);
        dst = tmp.data();
    }
    hipError_t e = hipMemcpyAsync(dst, d, f.count * sizeof(double), hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SMK_ERR_HIP, "download: %s", hipGetErrorString(e));
    if (field == SMK_FIELD_SMOKE && g->toff != 0.f) {
        double *v = static_cast<double *>(dst);
        for (size_t i = 3; i < f.count; i += 4) v[i] = v[i] + (double)g->toff;
    }
    if (dtype == SMK_DTYPE_F32) {
        float *o = static_cast<float *>(host);
        for (size_t i = 0; i < f.count; ++i) o[i] = (float)tmp[i];
    }
    return SMK_OK;
}

// smk_download_field through the fixed staging array of smk_noalloc.hip:
// the same conversion kernels write the field as doubles into the array,
// k_stage_to_out applies the BF16 smoke offset and the FP32 conversion in a
// second region of it, and one copy brings the result to the host.  No
// device or host memory is allocated.
smk_status smk_download_field_noalloc(smk_grid *g, int field, int lvl, void *host, size_t bytes, int dtype) {
    GRID_CHECK(g);
    FieldRef f;
    smk_status s = field_ref(g, field, lvl, &f);
    if (s != SMK_OK) return s;
    size_t esz = dtype == SMK_DTYPE_F64 ? 8 : (dtype == SMK_DTYPE_F32 ? 4 : 0);
    if (esz == 0) return fail(SMK_ERR_INVALID_ARG, "unknown dtype %d", dtype);
    if (!host || bytes != f.count * esz)
        return fail(SMK_ERR_INVALID_ARG, "buffer of %zu bytes, field needs %zu", bytes, f.count * esz);
    if (f.count > kNoAllocMaxValues)
        return fail(SMK_ERR_INVALID_ARG,
                    "field of %zu values exceeds the %zu-value limit of smk_download_field_noalloc "
                    "(smk_download_field has no limit)",
                    f.count, kNoAllocMaxValues);
    unsigned char *base = nullptr;  // this device's copy of the array
    HIPCHK(noalloc_scratch(g->dev, &base));
    std::lock_guard<std::mutex> lock(noalloc_lock(g->dev));
    double *d = reinterpret_cast<double *>(base);
    if (field == SMK_FIELD_DIVERGENCE)
        launch_divergence_field(g->stream, params(g), g->vfmt, vel_cur(g), g->s[0], d);
    else if (f.fmt == kFmtF64Split)
        launch_split64_to_f64(g->stream, level(g, 0), (const double *)f.ptr, d);
    else if (f.split_level >= 0)
        launch_split_to_f64(g->stream, level(g, f.split_level), (const float *)f.ptr, d);
    else
        launch_to_f64(g->stream, f.ptr, f.fmt, d, f.count);
    const double off = field == SMK_FIELD_SMOKE ? (double)g->toff : 0.0;
    void *out = d;
    if (dtype == SMK_DTYPE_F32) {
        out = base + noalloc_stage_bytes();
        launch_stage_to_out(g->stream, d, out, 1, f.count, off);
    } else if (off != 0.0) {
        launch_stage_to_out(g->stream, d, d, 0, f.count, off);  // in place, element by element
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, out, bytes, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) return fail(SMK_ERR_HIP, "download (no allocation): %s", hipGetErrorString(e));
    return SMK_OK;
}

smk_status smk_upload_field(smk_grid *g, int field, int lvl, const void *host, size_t bytes, int dtype) {
    GRID_CHECK(g);
    if (field == SMK_FIELD_DIVERGENCE) return fail(SMK_ERR_INVALID_ARG, "divergence is read only");
    FieldRef f;
    smk_status s = field_ref(g, field, lvl, &f);
    if (s != SMK_OK) return s;
    size_t esz = dtype == SMK_DTYPE_F64 ? 8 : (dtype == SMK_DTYPE_F32 ? 4 : 0);
    if (esz == 0) return fail(SMK_ERR_INVALID_ARG, "unknown dtype %d", dtype);
    if (!host || bytes != f.count * esz)
        return fail(SMK_ERR_INVALID_ARG, "buffer of %zu bytes, field needs %zu", bytes, f.count * esz);
    std::vector<double> tmp;
    const void *src = host;
    const bool offset = field == SMK_FIELD_SMOKE && g->toff != 0.f;
    if (dtype == SMK_DTYPE_F32 || offset) {
        tmp.resize(f.count);
        if (dtype == SMK_DTYPE_F32) {
            const float *i = static_cast<const float *>(host);
            for (size_t k = 0; k < f.count; ++k) tmp[k] = (double)i[k];
        } else {
            memcpy(tmp.data(), host, f.count * sizeof(double));
        }
        if (offset)
            for (size_t k = 3; k < f.count; k += 4) tmp[k] = tmp[k] - (double)g->toff;
        src = tmp.data();
    }
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, f.count * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, src, f.count * sizeof(double), hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess) {
        if (f.fmt == kFmtF64Split)
            launch_split64_from_f64(g->stream, level(g, 0), d, (double *)f.ptr);
        else if (f.split_level >= 0)
            launch_split_from_f64(g->stream, level(g, f.split_level), d, (float *)f.ptr);
        else
            KL(g, launch_from_f64)(g->stream, d, f.ptr, f.fmt, f.count);
        e = hipStreamSynchronize(g->stream);
    }
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SMK_ERR_HIP, "upload: %s", hipGetErrorString(e));
    if (field == SMK_FIELD_SOLID) {
        if (lvl == 0) {
            g->st.raster_valid = false;
            launch_solid_flag(g->stream, g->s[0], g->flag0, cells(g));
            HIPCHK(hipStreamSynchronize(g->stream));
        }
        g->st.pyr_dirty = true;
    }
    invalidate_graphs(g);
    return SMK_OK;
}

// FP64 refined pressure, per-block partial norms and the norm slots of
// `cycles` correction cycles.
static smk_status alloc_refine(smk_grid *g, int cycles) {
    if (!g->p64[0]) {
        size_t n = level_split(g, 0);
        for (int b = 0; b < 2; ++b) {
            if (hipMalloc(&g->p64[b], n * sizeof(double)) != hipSuccess)
                return fail(SMK_ERR_OUT_OF_MEMORY, "refined pressure: %zu doubles", n);
        }
        if (hipMalloc(&g->d_part, 2 * refine_partials(level(g, 0)) * sizeof(double)) != hipSuccess)
            return fail(SMK_ERR_OUT_OF_MEMORY, "refinement partial sums");
        // warm start from the current level-0 pressure
        launch_to_f64(g->stream, g->p[0], FMT_F32, g->p64[g->st.p64cur], n);
        launch_to_f64(g->stream, g->p[0], FMT_F32, g->p64[1 - g->st.p64cur], n);
    }
    if (cycles + 1 > g->norms_cap) {
        if (g->d_norms) (void)hipFree(g->d_norms);
        g->d_norms = nullptr;
        g->norms_cap = 0;
        if (hipMalloc(&g->d_norms, sizeof(double) * 2 * (size_t)(cycles + 1)) != hipSuccess)
            return fail(SMK_ERR_OUT_OF_MEMORY, "refinement norms");
        g->norms_cap = cycles + 1;
        launch_fill_f32(g->stream, reinterpret_cast<float *>(g->d_norms), 4 * (size_t)g->norms_cap, 0.f);
    }
    return check_launch("refinement buffers");
}

smk_status smk_set_refinement(smk_grid *g, int outer_cycles, int coarse_fp64) {
    GRID_CHECK(g);
    if (outer_cycles < 0 || outer_cycles > 256)
        return fail(SMK_ERR_INVALID_ARG, "outer cycles %d outside [0, 256]", outer_cycles);
    if (outer_cycles > 0 && !g->enclosed)
        return fail(SMK_ERR_UNSUPPORTED, "refinement needs an enclosed domain");
    if (outer_cycles > 0 && coarse_fp64) {
        if (g->bottom < 0)
            return fail(SMK_ERR_UNSUPPORTED, "FP64 coarsest level needs the LDS bottom solve");
        if (bottom_lds_bytes(bottom_levels(g), true) > kBottomMaxLds)
            return fail(SMK_ERR_UNSUPPORTED, "FP64 coarsest level does not fit in LDS");
    }
    invalidate_graphs(g);
    if (outer_cycles > 0) {
        smk_status s = alloc_refine(g, outer_cycles);
        if (s != SMK_OK) return s;
    }
    g->refine = outer_cycles;
    g->coarse64 = coarse_fp64 != 0;
    return SMK_OK;
}

smk_status smk_refinement_norms(smk_grid *g, double *out, int n) {
    GRID_CHECK(g);
    if (g->refine <= 0) return fail(SMK_ERR_INVALID_ARG, "refinement is off");
    if (!out || n < g->refine + 2) return fail(SMK_ERR_INVALID_ARG, "need %d values", g->refine + 2);
    std::vector<double> h(2 * (size_t)(g->refine + 1));
    HIPCHK(hipMemcpyAsync(h.data(), g->d_norms, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    for (int i = 0; i <= g->refine; ++i) out[i] = std::sqrt(h[2 * (size_t)i]);
    out[g->refine + 1] = std::sqrt(h[2 * (size_t)g->refine + 1]);
    return SMK_OK;
}

smk_status smk_sync(smk_grid *g) {
    GRID_CHECK(g);
    HIPCHK(hipStreamSynchronize(g->stream));
    return SMK_OK;
}

smk_status smk_calibrate_copy(int device, size_t bytes, int reps, double *ms_per_copy) {
    if (!ms_per_copy || reps < 1 || bytes < 16) return fail(SMK_ERR_INVALID_ARG, "bad calibration request");
    bytes &= ~(size_t)15;
    HIPCHK(hipSetDevice(device));
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    smk_status st = SMK_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        st = fail(SMK_ERR_OUT_OF_MEMORY, "calibration buffers");
    } else {
        launch_fill_u16(s, (uint16_t *)a, bytes / 2, 0x3c00);
        launch_copy16(s, a, b, bytes);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) launch_copy16(s, a, b, bytes);
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            st = fail(SMK_ERR_HIP, "calibration timing failed");
        *ms_per_copy = (double)ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return st;
}

smk_status smk_device_sync(int device) {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceSynchronize());
    return SMK_OK;
}

smk_status smk_residual_norm(smk_grid *g, double *out) {
    GRID_CHECK(g);
    if (!out) return fail(SMK_ERR_INVALID_ARG, "null output");
    const SimParams P = params(g);
    const size_t np = 2 * residual_norm_partials(P, level(g, 0));
    if (np > g->rpart_cap) {
        if (g->d_rpart) HIPCHK(hipFree(g->d_rpart));
        g->d_rpart = nullptr;
        g->rpart_cap = 0;
        HIPCHK(hipMalloc(&g->d_rpart, np * sizeof(double)));
        g->rpart_cap = np;
    }
    launch_residual_norm(g->stream, P, level(g, 0), g->vfmt, vel_cur(g), g->d_rpart, g->d_acc);
    double v = 0.0;
    HIPCHK(hipMemcpyAsync(&v, g->d_acc, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    *out = std::sqrt(v);
    return SMK_OK;
}

smk_status smk_set_stream(smk_grid *g, void *stream) {
    GRID_CHECK(g);
    HIPCHK(hipStreamSynchronize(g->stream));
    g->stream = stream ? static_cast<hipStream_t>(stream) : g->own;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_get_stream(smk_grid *g, void **stream) {
    if (!g || !stream) return fail(SMK_ERR_INVALID_ARG, "null argument");
    *stream = g->stream;
    return SMK_OK;
}

smk_status smk_set_graph_mode(smk_grid *g, int enable) {
    GRID_CHECK(g);
    g->graph_mode = enable != 0;
    if (!g->graph_mode) invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_profile_enable(smk_grid *g, int enable) {
    GRID_CHECK(g);
    HIPCHK(hipStreamSynchronize(g->stream));
    g->prof = enable != 0;
    for (int k = 0; k < SMK_K_COUNT; ++k) g->ev_used[k] = 0;
    return SMK_OK;
}

smk_status smk_profile_read(smk_grid *g, int id, double *total_ms, int *launches) {
    GRID_CHECK(g);
    if (id < 0 || id >= SMK_K_COUNT || !total_ms || !launches)
        return fail(SMK_ERR_INVALID_ARG, "bad profile query");
    HIPCHK(hipStreamSynchronize(g->stream));
    double tot = 0.0;
    int n = g->ev_used[id];
    for (int i = 0; i < n; ++i) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, g->ev[id][0][i], g->ev[id][1][i]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = n;
    g->ev_used[id] = 0;  // read and reset
    return SMK_OK;
}

}  // extern "C"

// ===========================================================================
// Differentiable rollouts: the FP32 differentiable step of smk_adjoint.hip,
// checkpointed every `checkpoint_every` steps, and its adjoint.  The rollout
// borrows its grid's solid maps, coordinate table, physics settings and
// pressure solver (refined, cold-started so that recomputing a segment from
// its checkpoint reproduces the forward bit for bit).
// ===========================================================================
struct smk_rollout {
    smk_grid *g = nullptr;
    smk_rollout_desc d{};
    Box3 box{};
    size_t nf[3] = {0, 0, 0};
    size_t nc = 0, nsrc = 0;
    int nseg = 0;
    // state (velocities, density) in the grid's storage format: FP32, BF16
    // or FP16, es bytes per element; arithmetic and adjoint fields are FP32
    int fmt = FMT_F32;
    size_t es = 4;
    void *v[3][2] = {};
    void *rho[2] = {nullptr, nullptr};
    float4 *w = nullptr;
    float *src = nullptr;
    // checkpoints: state (v, rho) at the start of every segment
    std::vector<void *> cp_v[3], cp_rho;
    // per step of the segment being differentiated: v_in, v2 (before the
    // confinement), v4 (after the projection), rho1 (after the source)
    std::vector<void *> sv_in[3], sv2[3], sv4[3], srho1;
    int saved_seg = -1;
    // adjoint state
    float *vb[3][2] = {};
    float *rb[2] = {nullptr, nullptr};
    float *sbar = nullptr;
    unsigned long long *acc_v[3] = {};
    unsigned long long *acc_r[2] = {nullptr, nullptr};
    float4 *wbar = nullptr, *gbar = nullptr, *ccbar = nullptr;
    float *rhs = nullptr;  // colour-split right-hand side of the adjoint solve
    double *lam = nullptr; // the last adjoint solve's solution (FP64 colour-split): next warm start
    double *d_resid = nullptr;  // largest relative residual: [0] forward solves, [1] adjoint solves
    bool adj_cold = true;       // the next adjoint solve starts from zero
    // source fitting on the device
    float *target = nullptr;     // target final density
    double *fit_part = nullptr;  // partial sums: loss, then |dL/dsrc|^2
    double *fit_L = nullptr;     // the current iteration's loss
    double *losses = nullptr;    // loss before each step of the current fit call
    int *fit_cnt = nullptr;      // iterations recorded in losses
    hipGraphExec_t fit_exec = nullptr;  // one captured fit iteration
    unsigned fit_gen = 0;        // the grid's generation it was captured at
    // departure-point velocity adjoints per source position
    float *ubar[4][3] = {};
    unsigned int *d_max = nullptr;  // 8 maxima (float bits)
    double *d_scale = nullptr;      // 8 scales
    std::vector<void *> owned;
};

namespace {

template <class T>
smk_status ro_alloc(smk_rollout *r, T **p, size_t n) {
    if (hipMalloc(p, sizeof(T) * (n ? n : 1)) != hipSuccess)
        return fail(SMK_ERR_OUT_OF_MEMORY, "rollout: %zu elements", n);
    r->owned.push_back(*p);
    return SMK_OK;
}

// n elements of state (r->es bytes each)
smk_status ro_alloc_state(smk_rollout *r, void **p, size_t n) {
    if (hipMalloc(p, r->es * (n ? n : 1)) != hipSuccess)
        return fail(SMK_ERR_OUT_OF_MEMORY, "rollout: %zu elements", n);
    r->owned.push_back(*p);
    return SMK_OK;
}

VelPtrs vp(float *const *v) { return VelPtrs{{v[0], v[1], v[2]}}; }
VelPtrs vp(void *const *v) { return VelPtrs{{v[0], v[1], v[2]}}; }

// copy n elements of state
smk_status ro_copy(smk_rollout *r, void *dst, const void *src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * r->es, hipMemcpyDeviceToDevice, r->g->stream));
    return SMK_OK;
}

void ro_zero_state(smk_rollout *r, void *p, size_t n) {
    if (r->es == 4) launch_fill_f32(r->g->stream, static_cast<float *>(p), n, 0.f);
    else launch_fill_u16(r->g->stream, static_cast<uint16_t *>(p), n, 0);
}

// projection of the differentiable step: vin -> vout (refined, cold start)
void ro_project(smk_rollout *r, const SimParams &P, VelPtrs vin, VelPtrs vout) {
    smk_grid *g = r->g;
    build_masks_if_dirty(g, P);
    launch_divergence0(g->stream, P, r->fmt, vin, g->s[0], level(g, 0), g->dv0, g->b0);
    launch_fill_f32(g->stream, reinterpret_cast<float *>(g->p64[g->st.p64cur]), 2 * level_split(g, 0), 0.f);
    refine_loop(g, P, g->dv0, (double)P.rho, (double)P.dt * (double)P.h, r->d.refine_cycles);
    launch_track_residual(g->stream, g->d_norms + 2 * r->d.refine_cycles, r->d_resid + 0);
    launch_update_velocity(g->stream, P, r->fmt, vin, vout, g->s[0], g->flag0, level(g, 0),
                           g->p64[g->st.p64cur]);
}

// One differentiable step on the working state; slot >= 0 saves the states
// the adjoint needs.  Buffers: v[.][0] -> v[.][1] (advect, buoyancy in
// place) -> v[.][0] (confinement) -> v[.][1] (projection); rho in place
// (source) -> rho[1] (advection); then both swap back to index 0.
smk_status ro_forward_step(smk_rollout *r, int slot) {
    smk_grid *g = r->g;
    SimParams P = params(g);
    smk_status s;
    const int D = g->D;
    if (slot >= 0)
        for (int c = 0; c < D; ++c)
            if ((s = ro_copy(r, r->sv_in[c][slot], r->v[c][0], r->nf[c])) != SMK_OK) return s;
    if (r->nsrc) launch_d_source(g->stream, P, r->fmt, r->rho[0], r->src, r->box);
    if (slot >= 0 && (s = ro_copy(r, r->srho1[slot], r->rho[0], r->nc)) != SMK_OK) return s;
    void *a[3] = {r->v[0][0], r->v[1][0], r->v[2][0]};
    void *b[3] = {r->v[0][1], r->v[1][1], r->v[2][1]};
    launch_advect_velocity(g->stream, P, r->fmt, FMT_F16, vp(a), vp(b), g->flag0, coord_tab(g), nullptr, 0);
    launch_d_buoyancy(g->stream, P, r->fmt, b[1], b[1], g->flag0, r->rho[0]);
    if (slot >= 0)
        for (int c = 0; c < D; ++c)
            if ((s = ro_copy(r, r->sv2[c][slot], b[c], r->nf[c])) != SMK_OK) return s;
    launch_d_vorticity(g->stream, P, r->fmt, vp(b), g->flag0, r->w);
    launch_d_confine(g->stream, P, r->fmt, vp(b), vp(a), r->w);
    ro_project(r, P, vp(a), vp(b));
    if (slot >= 0)
        for (int c = 0; c < D; ++c)
            if ((s = ro_copy(r, r->sv4[c][slot], b[c], r->nf[c])) != SMK_OK) return s;
    launch_d_advect_scalar(g->stream, P, r->fmt, vp(b), g->flag0, coord_tab(g), r->rho[0], r->rho[1]);
    for (int c = 0; c < 3; ++c) std::swap(r->v[c][0], r->v[c][1]);
    std::swap(r->rho[0], r->rho[1]);
    return check_launch("rollout forward step");
}

void ro_max(smk_rollout *r, int slot, const float *x, size_t n) {
    launch_maxabs(r->g->stream, x, n, r->d_max + slot);
}

// Adjoint of one step (slot of the saved segment).  In: vb[.][0] = dL/dv4,
// rb[0] = dL/drho2.  Out: vb[.][0] = dL/dv_in, rb[0] = dL/drho_in; sbar
// accumulates dL/dsrc.
smk_status ro_adjoint_step(smk_rollout *r, int slot) {
    smk_grid *g = r->g;
    SimParams P = params(g);
    hipStream_t st = g->stream;
    void *sv4[3] = {r->sv4[0][slot], r->sv4[1][slot], r->sv4[2][slot]};
    void *sv2[3] = {r->sv2[0][slot], r->sv2[1][slot], r->sv2[2][slot]};
    void *svin[3] = {r->sv_in[0][slot], r->sv_in[1][slot], r->sv_in[2][slot]};
    float *A[3] = {r->vb[0][0], r->vb[1][0], r->vb[2][0]};
    float *B[3] = {r->vb[0][1], r->vb[1][1], r->vb[2][1]};
    launch_fill_f32(st, reinterpret_cast<float *>(r->d_max), 8, 0.f);
    // (a) density advection: rbar1 (acc_r[0]); the departure-point velocity
    // term goes through the transposed interpolation into A
    ro_max(r, 0, r->rb[0], r->nc);
    launch_fx_scale(st, r->d_max + 0, nullptr, 1.0f, 0.0f, r->d_scale + 0);
    launch_fill_f32(st, reinterpret_cast<float *>(r->acc_r[0]), 2 * r->nc, 0.f);
    launch_a_advect_scalar(st, P, r->fmt, vp(sv4), g->flag0, coord_tab(g), r->srho1[slot], r->rb[0], r->acc_r[0],
                           r->d_scale + 0, r->ubar);
    launch_a_interp_gather(st, P, coord_tab(g), r->ubar, 1, vp(A), vp(A));
    // (b) projection: A = dL/dv4 -> B = dL/dv3
    const Level L0 = level(g, 0);
    launch_a_project_rhs(st, P, vp(A), g->flag0, L0, r->rhs);
    // warm start from the previous adjoint step's solution (zero before the
    // first): the adjoint solves run in the same order whatever the
    // checkpoint interval
    // (zero before the first, which gets at least refine_cycles cycles)
    const size_t n64 = level_split(g, 0);
    HIPCHK(hipMemcpyAsync(g->p64[g->st.p64cur], r->lam, n64 * sizeof(double), hipMemcpyDeviceToDevice, st));
    const int cyc = r->adj_cold ? std::max(r->d.refine_cycles, r->d.adjoint_cycles) : r->d.adjoint_cycles;
    r->adj_cold = false;
    refine_loop(g, P, r->rhs, 1.0, 1.0, cyc);
    launch_track_residual(st, g->d_norms + 2 * cyc, r->d_resid + 1);
    HIPCHK(hipMemcpyAsync(r->lam, g->p64[g->st.p64cur], n64 * sizeof(double), hipMemcpyDeviceToDevice, st));
    launch_a_project_apply(st, P, vp(A), vp(B), g->flag0, L0, g->p64[g->st.p64cur]);
    // (c) confinement: B = dL/dv3 -> A = dL/dv2 (vorticity of v2 recomputed)
    launch_d_vorticity(st, P, r->fmt, vp(sv2), g->flag0, r->w);
    launch_a_confine(st, P, g->flag0, r->w, vp(B), vp(A), r->wbar, r->gbar, r->ccbar);
    // (d) buoyancy: rbar1 += -alpha*dt W^T vbar_y (acc_r[1]); dL/dv1 = A
    ro_max(r, 2, A[1], r->nf[1]);
    launch_fx_scale(st, r->d_max + 2, nullptr, fabsf(P.alpha * P.dt), 0.0f, r->d_scale + 2);
    launch_fill_f32(st, reinterpret_cast<float *>(r->acc_r[1]), 2 * r->nc, 0.f);
    launch_a_buoyancy(st, P, g->flag0, A[1], r->acc_r[1], r->d_scale + 2);
    launch_fx_to_f32(st, r->acc_r[0], r->d_scale + 0, nullptr, r->rb[1], r->nc);
    launch_fx_to_f32(st, r->acc_r[1], r->d_scale + 2, r->rb[1], r->rb[1], r->nc);
    // (e) velocity advection: A = dL/dv1 -> B = dL/dv_in
    const int D = g->D;
    for (int c = 0; c < D; ++c) ro_max(r, 3, A[c], r->nf[c]);
    launch_fx_scale(st, r->d_max + 3, nullptr, 1.0f, 0.0f, r->d_scale + 3);
    for (int c = 0; c < D; ++c) launch_fill_f32(st, reinterpret_cast<float *>(r->acc_v[c]), 2 * r->nf[c], 0.f);
    launch_a_advect_velocity(st, P, r->fmt, vp(svin), g->flag0, coord_tab(g), vp(A), r->acc_v, r->d_scale + 3,
                             r->ubar);
    for (int c = 0; c < D; ++c) launch_fx_to_f32(st, r->acc_v[c], r->d_scale + 3, nullptr, B[c], r->nf[c]);
    launch_a_interp_gather(st, P, coord_tab(g), r->ubar, 0, vp(B), vp(B));
    for (int c = 0; c < 3; ++c) std::swap(r->vb[c][0], r->vb[c][1]);
    // (f) source: sbar += dt * rbar1; dL/drho_in = rbar1
    if (r->nsrc) launch_a_source(st, P, r->rb[1], r->sbar, r->box);
    std::swap(r->rb[0], r->rb[1]);
    return check_launch("rollout adjoint step");
}

smk_status ro_load_checkpoint(smk_rollout *r, int seg) {
    smk_status s;
    for (int c = 0; c < r->g->D; ++c)
        if ((s = ro_copy(r, r->v[c][0], r->cp_v[c][seg], r->nf[c])) != SMK_OK) return s;
    return ro_copy(r, r->rho[0], r->cp_rho[seg], r->nc);
}

// The forward rollout from the working state (the initial state), on the
// grid's stream with no host synchronisation: the initial state becomes
// checkpoint 0, later checkpoints are taken every checkpoint_every steps,
// the last segment keeps its step states, and rho[0] ends holding the final
// density (the working state holds the final state).
smk_status ro_forward_run(smk_rollout *r) {
    smk_grid *g = r->g;
    smk_status s;
    launch_fill_f32(g->stream, reinterpret_cast<float *>(r->d_resid), 2, 0.f);
    for (int c = 0; c < g->D; ++c)
        if ((s = ro_copy(r, r->cp_v[c][0], r->v[c][0], r->nf[c])) != SMK_OK) return s;
    if ((s = ro_copy(r, r->cp_rho[0], r->rho[0], r->nc)) != SMK_OK) return s;
    const int ce = r->d.checkpoint_every, last = r->nseg - 1;
    for (int step = 0; step < r->d.steps; ++step) {
        const int seg = step / ce;
        if (step > 0 && step % ce == 0) {
            for (int c = 0; c < g->D; ++c)
                if ((s = ro_copy(r, r->cp_v[c][seg], r->v[c][0], r->nf[c])) != SMK_OK) return s;
            if ((s = ro_copy(r, r->cp_rho[seg], r->rho[0], r->nc)) != SMK_OK) return s;
        }
        // the last segment keeps its step states: its adjoint needs no recompute
        if ((s = ro_forward_step(r, seg == last ? step - seg * ce : -1)) != SMK_OK) return s;
    }
    r->saved_seg = last;
    return check_launch("rollout forward");
}

// The backward pass from rb[0] = dL/d(final density), on the grid's stream
// with no host synchronisation: vb[.][0] = dL/d(initial velocity), sbar =
// dL/dsrc, and the working state is the initial state again.  Needs the
// saved step states of the last forward (saved_seg).
smk_status ro_backward_run(smk_rollout *r) {
    smk_grid *g = r->g;
    smk_status s;
    for (int c = 0; c < g->D; ++c) launch_fill_f32(g->stream, r->vb[c][0], r->nf[c], 0.f);
    launch_fill_f32(g->stream, r->sbar, r->nsrc ? r->nsrc : 1, 0.f);
    launch_fill_f32(g->stream, reinterpret_cast<float *>(r->lam), 2 * level_split(g, 0), 0.f);
    launch_fill_f32(g->stream, reinterpret_cast<float *>(r->d_resid + 1), 2, 0.f);
    r->adj_cold = true;
    const int ce = r->d.checkpoint_every;
    for (int seg = r->nseg - 1; seg >= 0; --seg) {
        const int s0 = seg * ce, s1 = std::min(r->d.steps, s0 + ce);
        if (seg != r->saved_seg) {
            if ((s = ro_load_checkpoint(r, seg)) != SMK_OK) return s;
            for (int step = s0; step < s1; ++step)
                if ((s = ro_forward_step(r, step - s0)) != SMK_OK) return s;
            r->saved_seg = seg;
        }
        for (int step = s1 - 1; step >= s0; --step)
            if ((s = ro_adjoint_step(r, step - s0)) != SMK_OK) return s;
    }
    // the step states of segment 0 are the saved ones now
    r->saved_seg = 0;
    if ((s = ro_load_checkpoint(r, 0)) != SMK_OK) return s;
    return check_launch("rollout backward");
}

// One iteration of the source fit on the grid's stream: forward rollout,
// loss L = sum (rho - target)^2 and rho_bar = 2 (rho - target), backward,
// and the projected Polyak step src <- max(0, src - (L / (2 |g|^2)) g).
constexpr int kFitMaxLosses = 4096;
smk_status ro_fit_iteration(smk_rollout *r) {
    smk_grid *g = r->g;
    smk_status s;
    if ((s = ro_forward_run(r)) != SMK_OK) return s;
    launch_fit_loss(g->stream, r->fmt, r->rho[0], r->target, r->nc, r->rb[0], r->fit_part, r->fit_L, r->losses,
                    kFitMaxLosses, r->fit_cnt);
    if ((s = ro_backward_run(r)) != SMK_OK) return s;
    launch_fit_update(g->stream, r->src, r->sbar, r->nsrc, r->fit_L, r->fit_part + kFitPartials);
    return check_launch("rollout fit iteration");
}

// The fit iteration captured as one graph on the grid's stream (again when
// the grid's buffers or settings changed since the capture).
smk_status ro_fit_graph(smk_rollout *r) {
    smk_grid *g = r->g;
    if (r->fit_exec && r->fit_gen == g->gen) return SMK_OK;
    if (r->fit_exec) (void)hipGraphExecDestroy(r->fit_exec);
    r->fit_exec = nullptr;
    // masks outside the graph, so that the graph does not rebuild them
    build_masks_if_dirty(g, params(g));
    const StepState before = g->st;
    HIPCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    smk_status s = ro_fit_iteration(r);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    g->st = before;
    if (s != SMK_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return s != SMK_OK ? s : fail(SMK_ERR_HIP, "fit capture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&r->fit_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        r->fit_exec = nullptr;
        return fail(SMK_ERR_HIP, "fit graph instantiate failed: %s", hipGetErrorString(e));
    }
    r->fit_gen = g->gen;
    return SMK_OK;
}

}  // namespace

extern "C" {

smk_status smk_rollout_create(smk_grid *g, const smk_rollout_desc *d, smk_rollout **out) {
    GRID_CHECK(g);
    if (!d || !out) return fail(SMK_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!g->enclosed) return fail(SMK_ERR_UNSUPPORTED, "differentiable rollouts need an enclosed grid");
    if (d->steps < 1 || d->checkpoint_every < 1 || d->refine_cycles < 1 || d->refine_cycles > 256 ||
        d->adjoint_cycles < 0 || d->adjoint_cycles > 256)
        return fail(SMK_ERR_INVALID_ARG, "bad rollout description");
    for (int a = 0; a < 3; ++a)
        if (d->src_size[a] < 0 || d->src_lo[a] < 0 || d->src_lo[a] + d->src_size[a] > g->n[a])
            return fail(SMK_ERR_INVALID_ARG, "source box outside the grid");
    smk_rollout *r = new smk_rollout;
    r->g = g;
    r->d = *d;
    r->fmt = g->vfmt;
    r->es = g->vfmt == FMT_F32 ? 4 : 2;
    if (r->d.adjoint_cycles == 0) r->d.adjoint_cycles = d->refine_cycles;
    smk_status s = alloc_refine(g, std::max(r->d.refine_cycles, r->d.adjoint_cycles));
    if (s != SMK_OK) {
        delete r;
        return s;
    }
    for (int a = 0; a < 3; ++a) {
        r->box.lo[a] = d->src_lo[a];
        r->box.hi[a] = d->src_lo[a] + d->src_size[a];
    }
    r->nsrc = (size_t)d->src_size[0] * d->src_size[1] * d->src_size[2];
    r->nc = cells(g);
    for (int c = 0; c < 3; ++c) r->nf[c] = c < g->D ? vcount(g, c) : 0;
    r->nseg = (d->steps + d->checkpoint_every - 1) / d->checkpoint_every;
    const int ce = std::min(d->checkpoint_every, d->steps);
#define RA(ptr, n)                                   \
    if ((s = ro_alloc(r, ptr, n)) != SMK_OK) {       \
        smk_rollout_destroy(r);                      \
        return s;                                    \
    }
#define RS(ptr, n)                                   \
    if ((s = ro_alloc_state(r, ptr, n)) != SMK_OK) { \
        smk_rollout_destroy(r);                      \
        return s;                                    \
    }
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 2; ++b) {
            RS(&r->v[c][b], r->nf[c]);
            RA(&r->vb[c][b], r->nf[c]);
        }
    for (int b = 0; b < 2; ++b) {
        RS(&r->rho[b], r->nc);
        RA(&r->rb[b], r->nc);
        RA(&r->acc_r[b], r->nc);
    }
    RA(&r->w, r->nc);
    RA(&r->wbar, r->nc);
    RA(&r->gbar, r->nc);
    RA(&r->ccbar, r->nc);
    RA(&r->src, r->nsrc);
    RA(&r->sbar, r->nsrc);
    RA(&r->rhs, level_split(g, 0));
    RA(&r->lam, level_split(g, 0));
    RA(&r->d_resid, 2);
    launch_fill_f32(g->stream, reinterpret_cast<float *>(r->d_resid), 4, 0.f);
    for (int c = 0; c < 4; ++c)
        for (int cc = 0; cc < 3; ++cc) RA(&r->ubar[c][cc], c < 3 ? r->nf[c] : r->nc);
    RA(&r->d_max, 8);
    RA(&r->d_scale, 8);
    for (int c = 0; c < 3; ++c) {
        RA(&r->acc_v[c], r->nf[c]);
        r->cp_v[c].resize(r->nseg);
        for (int k = 0; k < r->nseg; ++k) RS(&r->cp_v[c][k], r->nf[c]);
        r->sv_in[c].resize(ce);
        r->sv2[c].resize(ce);
        r->sv4[c].resize(ce);
        for (int k = 0; k < ce; ++k) {
            RS(&r->sv_in[c][k], r->nf[c]);
            RS(&r->sv2[c][k], r->nf[c]);
            RS(&r->sv4[c][k], r->nf[c]);
        }
    }
    r->cp_rho.resize(r->nseg);
    for (int k = 0; k < r->nseg; ++k) RS(&r->cp_rho[k], r->nc);
    r->srho1.resize(ce);
    for (int k = 0; k < ce; ++k) RS(&r->srho1[k], r->nc);
#undef RA
#undef RS
    for (int c = 0; c < g->D; ++c) ro_zero_state(r, r->v[c][0], r->nf[c]);
    ro_zero_state(r, r->rho[0], r->nc);
    launch_fill_f32(g->stream, r->src, r->nsrc ? r->nsrc : 1, 0.f);
    invalidate_graphs(g);
    if ((s = check_launch("rollout create")) != SMK_OK) {
        smk_rollout_destroy(r);
        return s;
    }
    *out = r;
    return SMK_OK;
}

smk_status smk_rollout_destroy(smk_rollout *r) {
    if (!r) return SMK_OK;
    (void)hipSetDevice(r->g->dev);
    (void)hipStreamSynchronize(r->g->stream);
    if (r->fit_exec) (void)hipGraphExecDestroy(r->fit_exec);
    for (void *p : r->owned) (void)hipFree(p);
    delete r;
    return SMK_OK;
}

smk_status smk_rollout_set_initial(smk_rollout *r, const float *vx, const float *vy, const float *vz,
                                   const float *rho) {
    if (!r || !vx || !vy || (!vz && r->g->D == 3) || !rho) return fail(SMK_ERR_INVALID_ARG, "null argument");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    const float *h[3] = {vx, vy, vz};
    // 16-bit state: through the FP32 adjoint buffers, rounded on the device
    for (int c = 0; c < g->D; ++c) {
        float *stage = r->es == 4 ? static_cast<float *>(r->v[c][0]) : r->vb[c][1];
        HIPCHK(hipMemcpyAsync(stage, h[c], r->nf[c] * sizeof(float), hipMemcpyHostToDevice, g->stream));
        if (r->es != 4) launch_state_from_f32(g->stream, r->fmt, stage, r->v[c][0], r->nf[c]);
    }
    float *stage = r->es == 4 ? static_cast<float *>(r->rho[0]) : r->rb[1];
    HIPCHK(hipMemcpyAsync(stage, rho, r->nc * sizeof(float), hipMemcpyHostToDevice, g->stream));
    if (r->es != 4) launch_state_from_f32(g->stream, r->fmt, stage, r->rho[0], r->nc);
    HIPCHK(hipStreamSynchronize(g->stream));
    r->saved_seg = -1;
    return SMK_OK;
}

smk_status smk_rollout_set_source(smk_rollout *r, const float *src) {
    if (!r || (!src && r->nsrc)) return fail(SMK_ERR_INVALID_ARG, "null argument");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    if (r->nsrc)
        HIPCHK(hipMemcpyAsync(r->src, src, r->nsrc * sizeof(float), hipMemcpyHostToDevice, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    r->saved_seg = -1;
    return SMK_OK;
}

smk_status smk_rollout_forward(smk_rollout *r, float *rho_final) {
    if (!r) return fail(SMK_ERR_INVALID_ARG, "null rollout");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    smk_status s;
    if ((s = ro_forward_run(r)) != SMK_OK) return s;
    if (rho_final) {
        const float *out = static_cast<const float *>(r->rho[0]);
        if (r->es != 4) {
            launch_state_to_f32(g->stream, r->fmt, r->rho[0], r->rb[1], r->nc);
            out = r->rb[1];
        }
        HIPCHK(hipMemcpyAsync(rho_final, out, r->nc * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    }
    // put the initial state back as the working state
    if ((s = ro_load_checkpoint(r, 0)) != SMK_OK) return s;
    HIPCHK(hipStreamSynchronize(g->stream));
    return SMK_OK;
}

smk_status smk_rollout_backward(smk_rollout *r, const float *rho_bar, float *src_bar, float *vx_bar,
                                float *vy_bar, float *vz_bar) {
    if (!r || !rho_bar) return fail(SMK_ERR_INVALID_ARG, "null argument");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    if (r->saved_seg < 0) return fail(SMK_ERR_INVALID_ARG, "backward before forward");
    smk_status s;
    HIPCHK(hipMemcpyAsync(r->rb[0], rho_bar, r->nc * sizeof(float), hipMemcpyHostToDevice, g->stream));
    if ((s = ro_backward_run(r)) != SMK_OK) return s;
    float *host[3] = {vx_bar, vy_bar, vz_bar};
    for (int c = 0; c < g->D; ++c)
        if (host[c])
            HIPCHK(hipMemcpyAsync(host[c], r->vb[c][0], r->nf[c] * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    if (src_bar && r->nsrc)
        HIPCHK(hipMemcpyAsync(src_bar, r->sbar, r->nsrc * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    return SMK_OK;
}

smk_status smk_rollout_set_target(smk_rollout *r, const float *target) {
    if (!r || !target) return fail(SMK_ERR_INVALID_ARG, "null argument");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    smk_status s;
    if (!r->target) {
        if ((s = ro_alloc(r, &r->target, r->nc)) != SMK_OK) return s;
        if ((s = ro_alloc(r, &r->fit_part, 2 * (size_t)kFitPartials)) != SMK_OK) return s;
        if ((s = ro_alloc(r, &r->fit_L, 1)) != SMK_OK) return s;
        if ((s = ro_alloc(r, &r->losses, kFitMaxLosses)) != SMK_OK) return s;
        if ((s = ro_alloc(r, &r->fit_cnt, 1)) != SMK_OK) return s;
        launch_fill_f32(g->stream, reinterpret_cast<float *>(r->fit_cnt), 1, 0.f);
    }
    HIPCHK(hipMemcpyAsync(r->target, target, r->nc * sizeof(float), hipMemcpyHostToDevice, g->stream));
    launch_fill_f32(g->stream, reinterpret_cast<float *>(r->fit_cnt), 1, 0.f);  // a new fit
    HIPCHK(hipStreamSynchronize(g->stream));
    return SMK_OK;
}

smk_status smk_rollout_fit(smk_rollout *r, int iterations) {
    if (!r) return fail(SMK_ERR_INVALID_ARG, "null rollout");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    if (!r->target) return fail(SMK_ERR_INVALID_ARG, "fit without a target (smk_rollout_set_target)");
    if (iterations < 0 || iterations > kFitMaxLosses)
        return fail(SMK_ERR_INVALID_ARG, "iterations %d outside [0, %d]", iterations, kFitMaxLosses);
    smk_status s;
    if (g->prof) {
        // profiling brackets kernels with events: eager iterations
        for (int it = 0; it < iterations; ++it)
            if ((s = ro_fit_iteration(r)) != SMK_OK) return s;
        return SMK_OK;
    }
    if ((s = ro_fit_graph(r)) != SMK_OK) return s;
    for (int it = 0; it < iterations; ++it) HIPCHK(hipGraphLaunch(r->fit_exec, g->stream));
    r->saved_seg = 0;
    return SMK_OK;
}

smk_status smk_rollout_fit_result(smk_rollout *r, double *losses, int n, int *recorded, float *src) {
    if (!r || n < 0 || (n > 0 && !losses)) return fail(SMK_ERR_INVALID_ARG, "bad argument");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    if (!r->target) return fail(SMK_ERR_INVALID_ARG, "no fit (smk_rollout_set_target)");
    int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, r->fit_cnt, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    const int m = std::min(std::min(cnt, n), kFitMaxLosses);
    if (m > 0) HIPCHK(hipMemcpyAsync(losses, r->losses, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    if (src && r->nsrc)
        HIPCHK(hipMemcpyAsync(src, r->src, r->nsrc * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    if (recorded) *recorded = cnt;
    return SMK_OK;
}

smk_status smk_rollout_residuals(smk_rollout *r, double *forward_max, double *adjoint_max) {
    if (!r) return fail(SMK_ERR_INVALID_ARG, "null rollout");
    smk_grid *g = r->g;
    GRID_CHECK(g);
    double h[2] = {0.0, 0.0};
    HIPCHK(hipMemcpyAsync(h, r->d_resid, sizeof h, hipMemcpyDeviceToHost, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
    if (forward_max) *forward_max = h[0];
    if (adjoint_max) *adjoint_max = h[1];
    return SMK_OK;
}

}  // extern "C"

// ===========================================================================
// Launch-tile autotuner and batches of uneven scenes.
// ===========================================================================
namespace {

// The description a grid was created from.
smk_grid_desc desc_of(const smk_grid *g) {
    smk_grid_desc d{};
    for (int a = 0; a < 3; ++a) d.size[a] = g->D == 2 && a == 2 ? 0 : g->n[a];
    d.cell_size = g->h;
    d.t_ambient = g->tamb;
    d.storage = g->vfmt == FMT_F32 ? SMK_PREC_FP32 : (g->vfmt == FMT_BF16 ? SMK_PREC_BF16 : SMK_PREC_FP16);
    d.min_level_size = g->min_level;
    d.enclosed = g->enclosed;
    d.device = g->dev;
    return d;
}

// Milliseconds of `reps` steps of a scratch grid with g's description,
// physics and solver settings at tile width bx.
smk_status time_tile(smk_grid *g, int bx, int reps, double *ms) {
    const smk_grid_desc d = desc_of(g);
    smk_grid *t = nullptr;
    smk_status s = smk_grid_create(&d, &t);
    if (s != SMK_OK) return s;
    t->phys = g->phys;
    refresh_div(t);
    t->mg = g->mg;
    t->tile_bx = bx;
    t->arith = g->arith;
    if (g->refine > 0) s = smk_set_refinement(t, g->refine, g->coarse64 ? 1 : 0);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s == SMK_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        s = fail(SMK_ERR_HIP, "event create failed");
    if (s == SMK_OK) s = step_impl(t);  // warm-up (first-touch, masks)
    // the faster of two timed rounds (eager launches vary by a few percent)
    float best = 0.f;
    for (int round = 0; round < 2 && s == SMK_OK; ++round) {
        if (hipEventRecord(e0, t->stream) != hipSuccess) s = fail(SMK_ERR_HIP, "event record");
        for (int i = 0; i < reps && s == SMK_OK; ++i) s = step_impl(t);
        float f = 0.f;
        if (s == SMK_OK && (hipEventRecord(e1, t->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&f, e0, e1) != hipSuccess))
            s = fail(SMK_ERR_HIP, "autotune timing failed");
        if (s == SMK_OK && (round == 0 || f < best)) best = f;
    }
    *ms = best;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    smk_grid_destroy(t);
    return s;
}

}  // namespace

struct smk_batch {
    int mode = SMK_BATCH_PER_SCENE;
    std::vector<smk_grid *> g;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    // order a single-graph launch on streams[0] after the work already
    // queued on the other streams (pre), and their later work after it (post)
    std::vector<hipEvent_t> pre;
    hipEvent_t post = nullptr;
    std::map<std::string, std::pair<hipGraphExec_t, std::vector<StepState>>> graphs;
    // fused mode: per ping-pong state, the graph of one batch step and the
    // device arena holding its launches' scene tables
    struct Fused {
        hipGraphExec_t exec = nullptr;
        void *arena = nullptr;
        std::vector<StepState> after;
    };
    std::map<std::string, Fused> fused;
    float **d_times = nullptr;  // every scene's time slot (device array)
    // fused mode: block width per public launch kind (SMK_FUSED_*), 0 for
    // the idle-lane rule of fused_tiles(); set by smk_batch_autotune
    int fused_bx[SMK_FUSED_KINDS] = {};
    // fused mode: the scenes fall into groups (kFusedGroups per dimension,
    // balanced by cells) whose launches run on streams of their own, forked
    // from and joined back into streams[0] within the captured step (no two
    // groups share a scene)
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_join;
    hipEvent_t side_fork = nullptr;
};

namespace {

std::string batch_key(const smk_batch *b) {
    std::string k;
    for (const smk_grid *g : b->g) {
        k += state_key(g);
        k += ':';
        k += std::to_string(g->gen);
        k += ';';
    }
    return k;
}

void batch_free_graphs(smk_batch *b) {
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second.first);
    b->graphs.clear();
    for (auto &kv : b->fused) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.arena) (void)hipFree(kv.second.arena);
    }
    b->fused.clear();
}

// ---------------------------------------------------------------------------
// Fused batch step.  One launch per stage (and per multigrid level and half
// sweep) covers every scene of a dimension class: the planner walks the
// step of step_impl() for all scenes at once, advancing each scene's
// ping-pong state as step_impl does, and records for each launch the table
// of per-scene arguments (the single-scene launchers' arguments) and the
// prefix sums of their block counts.  The V-cycles are aligned at level 0:
// level l of the descent and ascent runs for the scenes whose LDS bottom
// starts below l, and one launch runs every scene's bottom V-cycle (one
// workgroup each) once all descents are done.  The general kernels serve
// every level (the special-size variants give the same results), so fused
// fields equal single-scene ones bit for bit.
enum FusedKind {
    FK_EMIT_STATUS, FK_EMIT, FK_ADVECT_VEL, FK_VORT, FK_APPLY_VORT, FK_DIV0, FK_SWEEP, FK_RESTRICT,
    FK_BOTTOM, FK_PROLONG, FK_UPDATE_VEL, FK_ADVECT_SMOKE, FK_OBSTACLES, FK_COARSE
};

constexpr int kFusedBx = 32;  // block width of the multi-scene launches (256 threads)
constexpr int kFusedGroups = 3;  // scene groups per dimension, each on its own stream

// The public kind (SMK_FUSED_*) of a tiled fused launch, -1 for the launches
// whose shape is fixed (emitter status, bottom V-cycles, coarse smoother,
// obstacles).
int public_kind(int fk) {
    switch (fk) {
    case FK_EMIT: return SMK_FUSED_EMIT;
    case FK_ADVECT_VEL: return SMK_FUSED_ADVECT_VELOCITY;
    case FK_VORT: return SMK_FUSED_VORTICITY;
    case FK_APPLY_VORT: return SMK_FUSED_APPLY_VORTICITY;
    case FK_DIV0: return SMK_FUSED_DIVERGENCE;
    case FK_SWEEP: return SMK_FUSED_SWEEP;
    case FK_RESTRICT: return SMK_FUSED_RESTRICT;
    case FK_PROLONG: return SMK_FUSED_PROLONG;
    case FK_UPDATE_VEL: return SMK_FUSED_UPDATE_VELOCITY;
    case FK_ADVECT_SMOKE: return SMK_FUSED_ADVECT_SMOKE;
    default: return -1;
    }
}

struct FusedLaunch {
    int kind, dim, variant = 0;
    int group = 0;  // the scene group (its own stream in the captured step)
    int n = 0, blocks = 0, bx = kFusedBx;
    size_t start_off = 0, rec_off = 0, lds = 0;
    int own = 4;  // bottom launches: cells per thread (bottom_own)
    std::vector<int> start{0};
    std::vector<unsigned char> recs;
    // tiled launches (records that start with an MTile): each scene's thread
    // extent and record offset; fused_tiles() picks the block width
    bool tiled = false;
    int min_bx = 16;
    std::vector<std::array<int, 3>> ext;
    std::vector<size_t> rec_at;
};

struct FusedPlan {
    int vfmt = 0, sfmt = 0;
    bool buf = true;
    std::vector<FusedLaunch> launches;
};


template <class R>
void fused_add(FusedLaunch &L, const R &r, int blocks) {
    const size_t o = L.recs.size();
    L.recs.resize(o + sizeof(R));
    std::memcpy(L.recs.data() + o, &r, sizeof(R));
    L.start.push_back(L.start.back() + blocks);
    ++L.n;
    L.blocks += blocks;
}
// A scene of a tiled launch: its record (whose leading MTile fused_tiles()
// fills in) and the extent of its thread grid (x, y threads, z planes).
template <class R>
void fused_add_tiled(FusedLaunch &L, const R &r, int sx, int sy, int sz) {
    static_assert(offsetof(R, t) == 0, "the tile leads the record");
    L.tiled = true;
    L.rec_at.push_back(L.recs.size());
    L.ext.push_back({sx, sy, sz});
    fused_add(L, r, 0);
}
// A scene of a flat launch: the record's MTile takes its thread extents
// and the scene gets ceil(sx*sy*sz / 256) blocks of 256 threads.
template <class R>
void fused_add_flat(FusedLaunch &L, R r, int sx, int sy, int sz) {
    static_assert(offsetof(R, t) == 0, "the tile leads the record");
    L.bx = 256;
    r.t = MTile{sx, sy, sz};
    fused_add(L, r, (int)(((long)sx * sy * sz + 255) / 256));
}
// A scene of a sweep launch: its threads (hx/4 per colour row with four
// cells each, else hx) as one linear range, 256 per block (k_sweep_multi).
void fused_add_sweep(FusedLaunch &L, const MSweep &r, bool four) {
    L.bx = 256;
    const long threads = (long)(r.L.hx / (four ? 4 : 1)) * r.L.n[1] * r.L.n[2];
    fused_add(L, r, (int)((threads + 255) / 256));
}
// Block shape of a tiled launch once all its scenes are in: force_bx when
// the autotuner set one, else the width bx (blocks of bx x 256/bx threads,
// bx a power of two from min_bx to 64) that leaves the fewest idle lanes
// over the scenes' extents, the wider on a tie; then every scene's tile and
// the prefix sums of its blocks.
void fused_tiles(FusedLaunch &L, int force_bx) {
    long best = -1;
    int bx = kFusedBx;
    for (int w = 64; w >= L.min_bx && force_bx <= 0; w >>= 1) {
        const int h = 256 / w;
        long lanes = 0;
        for (const auto &e : L.ext) lanes += (long)((e[0] + w - 1) / w) * w * ((e[1] + h - 1) / h) * h * e[2];
        if (best < 0 || lanes < best) {
            best = lanes;
            bx = w;
        }
    }
    if (force_bx > 0) bx = force_bx;
    const int by = 256 / bx;
    L.bx = bx;
    L.start.assign(1, 0);
    L.blocks = 0;
    for (size_t i = 0; i < L.ext.size(); ++i) {
        const MTile t{(L.ext[i][0] + bx - 1) / bx, (L.ext[i][1] + by - 1) / by};
        std::memcpy(L.recs.data() + L.rec_at[i], &t, sizeof(MTile));
        const int nb = t.gx * t.gy * L.ext[i][2];
        L.start.push_back(L.start.back() + nb);
        L.blocks += nb;
    }
}

// Whether the batch can run fused, with a reason when it cannot.
bool fused_ok(const smk_batch *b, std::string *why) {
    const smk_grid *g0 = b->g[0];
    for (const smk_grid *g : b->g) {
        char buf[160];
        auto no = [&](const char *r) {
            snprintf(buf, sizeof(buf), "scene %dx%dx%d: %s", g->n[0], g->n[1], g->n[2], r);
            *why = buf;
            return false;
        };
        if (!default_toggles(g)) return no("step toggles other than all on");
        if (g->refine > 0) return no("FP64 refinement");
        if (g->prof) return no("profiling on");
        if (g->arith != g0->arith || g->arith != SMK_ARITH_DEFAULT) return no("llvmpipe arithmetic");
        if (g->vfmt != g0->vfmt || g->sfmt != g0->sfmt) return no("storage differs from scene 0");
        if (std::memcmp(&g->mg, &g0->mg, sizeof(smk_mg_params)) != 0) return no("multigrid schedule differs");
        if (!all_anchored(g)) return no("an obstacle that moves (per-step rasterisation)");
    }
    return true;
}

void fused_plan(smk_batch *b, FusedPlan &plan) {
    plan.vfmt = b->g[0]->vfmt;
    plan.sfmt = b->g[0]->sfmt;
    for (smk_grid *g : b->g) plan.buf = plan.buf && buf_samplers_ok(params(g));
    const smk_mg_params mg = b->g[0]->mg;
    // one launch per dimension (and per kernel variant, for the stages that
    // have variants); add() records a scene into L when it takes part
    // scene groups: per dimension, kFusedGroups of them with about equal
    // cells (largest scene first into the lightest group)
    std::unordered_map<const smk_grid *, int> group_of;
    for (int dim = 2; dim <= 3; ++dim) {
        std::vector<smk_grid *> sc;
        for (smk_grid *g : b->g)
            if (g->D == dim) sc.push_back(g);
        std::stable_sort(sc.begin(), sc.end(), [](const smk_grid *x, const smk_grid *y) { return cells(x) > cells(y); });
        size_t load[kFusedGroups] = {};
        for (smk_grid *g : sc) {
            int q = 0;
            for (int c = 1; c < kFusedGroups; ++c)
                if (load[c] < load[q]) q = c;
            group_of[g] = q;
            load[q] += cells(g);
        }
    }
    auto for_class = [&](int kind, int dim, int variant, auto &&add) {
        for (int q = 0; q < kFusedGroups; ++q) {
            FusedLaunch L;
            L.kind = kind;
            L.dim = dim;
            L.variant = variant;
            L.group = q;
            for (smk_grid *g : b->g)
                if (g->D == dim && group_of[g] == q) add(L, g);
            if (L.tiled) {
                const int pk = public_kind(kind);
                fused_tiles(L, pk >= 0 ? b->fused_bx[pk] : 0);
            }
            if (L.n > 0) plan.launches.push_back(std::move(L));
        }
    };
    auto stage = [&](int kind, auto &&add) {
        for (int dim = 2; dim <= 3; ++dim) for_class(kind, dim, 0, add);
    };
    auto stage_v = [&](int kind, int nvariants, auto &&add) {
        for (int dim = 2; dim <= 3; ++dim)
            for (int v = 0; v < nvariants; ++v) for_class(kind, dim, v, add);
    };
    auto zf = [](const smk_grid *g) { return g->D == 3 ? 1 : 0; };
    // emitters (st_emitters)
    stage(FK_EMIT_STATUS, [&](FusedLaunch &L, smk_grid *g) {
        if (g->em.empty()) return;
        L.bx = 64;
        fused_add(L, MEmitStatus{g->d_em, (int)g->em.size(), g->d_time}, ((int)g->em.size() + 63) / 64);
    });
    stage(FK_EMIT, [&](FusedLaunch &L, smk_grid *g) {
        if (g->em.empty()) return;
        const Box3 &bx = g->embox;
        const int ex = bx.hi[0] - bx.lo[0], ey = bx.hi[1] - bx.lo[1], ez = bx.hi[2] - bx.lo[2];
        if (ex <= 0 || ey <= 0 || ez <= 0) return;
        fused_add_tiled(L, MEmit{MTile{}, params(g), g->smoke[g->st.scur], g->v[1][g->st.vcur[1]], g->d_em,
                           g->has_inflow ? g->d_inflow : nullptr, (int)g->em.size(), bx}, ex, ey, ez);
    });
    // velocity advection with the fused buoyancy (st_advect)
    stage(FK_ADVECT_VEL, [&](FusedLaunch &L, smk_grid *g) {
        fused_add_tiled(L, MAdvectVel{MTile{}, params(g), vel_cur(g), vel_other(g), g->flag0, coord_tab(g), g->smoke[g->st.scur]},
            g->n[0] + 1, g->n[1] + 1, g->n[2] + zf(g));
        flip_all(g);
    });
    // vorticity confinement (st_vorticity, st_apply_vorticity)
    stage(FK_VORT, [&](FusedLaunch &L, smk_grid *g) {
        fused_add_flat(L, MVort{MTile{}, params(g), vel_cur(g), g->flag0, g->vort}, g->n[0], g->n[1], g->n[2]);
    });
    stage(FK_APPLY_VORT, [&](FusedLaunch &L, smk_grid *g) {
        fused_add_flat(L, MApplyVort{MTile{}, params(g), vel_cur(g), vel_other(g), g->vort},
            g->n[0] + 1, g->n[1] + 1, g->n[2] + zf(g));
        flip_all(g);
    });
    // pressure solve (st_pressure, v_cycle)
    stage_v(FK_DIV0, 2, [&](FusedLaunch &L, smk_grid *g) {
        const bool four = g->D == 3 && g->n[0] % 4 == 0;  // k_divergence0_x4
        if ((int)four != L.variant) return;
        fused_add_flat(L, MDiv0{MTile{}, params(g), vel_cur(g), g->s[0], level(g, 0), g->dv0, g->b0},
            four ? g->n[0] / 4 : g->n[0], g->n[1], g->n[2]);
    });
    // the level where a scene's V-cycle bottoms out: its LDS bottom, or its
    // coarsest level, smoothed by plain half sweeps (v_cycle at l == L - 1)
    auto bot = [](const smk_grid *g) { return g->bottom >= 1 ? g->bottom : g->L - 1; };
    int maxb = 0;
    for (smk_grid *g : b->g) maxb = std::max(maxb, bot(g));
    auto sweeps = [&](int l, int iterations) {
        for (int it = 0; it < 2 * iterations; ++it)
            stage_v(FK_SWEEP, 2, [&](FusedLaunch &L, smk_grid *g) {
                if (l >= bot(g)) return;
                const Level Lv = level(g, l);
                // k_sweep4 on level 0 only: the coarser levels' fractional
                // solid fractions send more cells through its general update
                const bool four = Lv.hx % 4 == 0 && l == 0;
                if ((int)four != L.variant) return;
                SimParams P = params(g);
                fused_add_sweep(L, MSweep{MTile{}, P, Lv, l, it & 1, sweep_rhs(g, l), P.h * (float)(1 << l)}, four);
            });
    };
    for (int v = 0; v < mg.v_cycles; ++v) {
        // the level-0 residual image is observable: only the last V-cycle
        // must store it (as in st_pressure)
        const bool write_fine = v + 1 == mg.v_cycles;
        for (int l = 0; l < maxb; ++l) {
            sweeps(l, mg.pre_iterations);
            stage_v(FK_RESTRICT, 4, [&](FusedLaunch &L, smk_grid *g) {
                if (l >= bot(g)) return;
                const Level F = level(g, l), C = level(g, l + 1);
                SimParams P = params(g);
                const FusedVariant fv = fused_variant_residual_restrict(P, F, C, l, write_fine);
                if (fv.variant != L.variant) return;
                const MResRestrict rec{MTile{},
                    P, F, C, l, level_rhs(g, l), vel_cur(g), g->s[0], P.h * (float)(1 << l), fv.wf, fv.RD};
                // the general kernel runs flat (one thread per coarse cell,
                // no tile edges), the even ones tiled
                if (fv.variant == kRrGeneral) fused_add_flat(L, rec, fv.s[0], fv.s[1], fv.s[2]);
                else fused_add_tiled(L, rec, fv.s[0], fv.s[1], fv.s[2]);
            });
        }
        // coarsest levels without an LDS bottom: the single-workgroup
        // smoother (one workgroup per scene), or 2 * coarse half sweeps
        size_t coarse_lds[4] = {0, 0, 0, 0};
        stage(FK_COARSE, [&](FusedLaunch &L, smk_grid *g) {
            const int l = g->L - 1;
            const Level Lv = level(g, l);
            if (g->bottom >= 1 || !(coarse_fits(params(g), Lv) || coarsest_fits(Lv))) return;
            coarse_lds[g->D] = std::max(coarse_lds[g->D], coarse_lds_bytes(Lv));
            SimParams P = params(g);
            fused_add(L, MCoarse{P, Lv, l, mg.coarse_iterations, level_rhs(g, l), P.h * (float)(1 << l)}, 1);
        });
        for (auto &L : plan.launches)
            if (L.kind == FK_COARSE && L.lds == 0) L.lds = coarse_lds[L.dim];
        for (int it = 0; it < 2 * mg.coarse_iterations; ++it)
            stage(FK_SWEEP, [&](FusedLaunch &L, smk_grid *g) {
                if (g->bottom >= 1) return;
                if (coarse_fits(params(g), level(g, g->L - 1)) || coarsest_fits(level(g, g->L - 1))) return;
                const int l = g->L - 1;
                const Level Lv = level(g, l);
                SimParams P = params(g);
                fused_add_sweep(L, MSweep{MTile{}, P, Lv, l, it & 1, sweep_rhs(g, l), P.h * (float)(1 << l)}, false);
            });
        size_t lds_max[4] = {0, 0, 0, 0};
        int own_max[4] = {4, 4, 4, 4};
        stage(FK_BOTTOM, [&](FusedLaunch &L, smk_grid *g) {
            if (g->bottom < 1) return;
            const BottomLevels B = bottom_levels(g);
            lds_max[g->D] = std::max(lds_max[g->D], bottom_lds_bytes(B, false));
            own_max[g->D] = std::max(own_max[g->D], bottom_own(B));
            fused_add(L, MBottom{B, mg.pre_iterations, mg.post_iterations, mg.coarse_iterations, 0}, 1);
        });
        for (auto &L : plan.launches)
            if (L.kind == FK_BOTTOM && L.lds == 0) {
                L.lds = lds_max[L.dim];
                L.own = own_max[L.dim];
            }
        for (int l = maxb - 1; l >= 0; --l) {
            stage_v(FK_PROLONG, 3, [&](FusedLaunch &L, smk_grid *g) {
                if (l >= bot(g)) return;
                const Level F = level(g, l), C = level(g, l + 1);
                const FusedVariant fv = fused_variant_prolong(params(g), F, C, g->prolong_even[l]);
                if (fv.variant != L.variant) return;
                const MProlong rec{MTile{}, F, C, g->ptab[l][0], g->ptab[l][1], g->ptab[l][2]};
                if (fv.variant == kPrGeneral) fused_add_flat(L, rec, fv.s[0], fv.s[1], fv.s[2]);
                else fused_add_tiled(L, rec, fv.s[0], fv.s[1], fv.s[2]);
            });
            sweeps(l, mg.post_iterations);
        }
    }
    // projection (st_update_velocities), smoke advection, obstacles
    stage(FK_UPDATE_VEL, [&](FusedLaunch &L, smk_grid *g) {
        fused_add_flat(L, MUpdateVel{MTile{}, params(g), vel_cur(g), vel_other(g), g->s[0], g->flag0, level(g, 0)},
            g->n[0] + 1, g->n[1] + 1, g->n[2] + zf(g));
        flip_all(g);
    });
    stage(FK_ADVECT_SMOKE, [&](FusedLaunch &L, smk_grid *g) {
        fused_add_tiled(L, MAdvectSmoke{MTile{},
            params(g), vel_cur(g), g->flag0, coord_tab(g), g->smoke[g->st.scur],
                                  g->smoke[1 - g->st.scur]}, g->n[0], g->n[1], g->n[2]);
        g->st.scur ^= 1;
    });
    stage(FK_OBSTACLES, [&](FusedLaunch &L, smk_grid *g) {
        if (!g->ob.empty()) {
            L.bx = 256;
            fused_add(L, MObstacles{params(g), g->d_ob[g->st.obcur], g->d_ob[1 - g->st.obcur], (int)g->ob.size()},
                      ((int)g->ob.size() + 255) / 256);
        }
    });
    for (smk_grid *g : b->g) g->st.obcur ^= 1;  // grid.c:436-440 swaps even with no obstacles
}

// Issue the launches of a plan whose tables sit in `arena` (device); ev,
// when given, holds an event pair per launch recorded around it.
// With `side` streams (and fork/join events), each scene group (dimension,
// group) issues its launches, in their order, on a stream of its own: the
// 3D group 0 on s0, the others on side[0..] between a fork from and joins
// back into s0.  No two groups touch the same scene, and the groups' small,
// latency-bound launches (coarse levels, bottom) then overlap.
void fused_issue(hipStream_t s0, const FusedPlan &plan, const unsigned char *arena,
                 const std::vector<std::pair<hipEvent_t, hipEvent_t>> *ev = nullptr,
                 const std::vector<hipStream_t> *side = nullptr, hipEvent_t fork = nullptr,
                 const std::vector<hipEvent_t> *join = nullptr) {
    // stream slot of a launch: 0 for (3D, group 0), else 1 + index
    auto slot = [](const FusedLaunch &L) {
        const int q = (L.dim == 3 ? 0 : kFusedGroups) + L.group;
        return q;
    };
    std::vector<char> used(2 * kFusedGroups, 0);
    for (const FusedLaunch &L : plan.launches) used[slot(L)] = 1;
    const bool split = side && fork && join && (int)side->size() >= 2 * kFusedGroups - 1 &&
                       (int)join->size() >= 2 * kFusedGroups - 1;
    auto stream_of = [&](int q) { return q == 0 || !split ? s0 : (*side)[q - 1]; };
    if (split) {
        (void)hipEventRecord(fork, s0);
        for (int q = 1; q < 2 * kFusedGroups; ++q)
            if (used[q]) (void)hipStreamWaitEvent(stream_of(q), fork, 0);
    }
    for (size_t li = 0; li < plan.launches.size(); ++li) {
        const FusedLaunch &L = plan.launches[li];
        const hipStream_t s = stream_of(slot(L));
        if (ev) (void)hipEventRecord((*ev)[li].first, s);
        MultiLaunch m{reinterpret_cast<const int *>(arena + L.start_off), L.n, L.blocks, L.bx};
        const void *E = arena + L.rec_off;
        const int d = L.dim, vf = plan.vfmt, sf = plan.sfmt;
        switch (L.kind) {
        case FK_EMIT_STATUS: launch_emit_status_multi(s, m, (const MEmitStatus *)E); break;
        case FK_EMIT: launch_emit_multi(s, d, vf, sf, m, (const MEmit *)E); break;
        case FK_ADVECT_VEL: launch_advect_velocity_multi(s, d, vf, sf, true, plan.buf, m, (const MAdvectVel *)E); break;
        case FK_VORT: launch_vorticity_multi(s, d, vf, m, (const MVort *)E); break;
        case FK_APPLY_VORT: launch_apply_vorticity_multi(s, d, vf, m, (const MApplyVort *)E); break;
        case FK_DIV0: launch_divergence0_multi(s, d, vf, L.variant == 1, m, (const MDiv0 *)E); break;
        case FK_SWEEP: launch_sweep_multi(s, d, L.variant == 1, m, (const MSweep *)E); break;
        case FK_RESTRICT: launch_residual_restrict_multi(s, d, vf, L.variant, m, (const MResRestrict *)E); break;
        case FK_BOTTOM: launch_bottom_vcycle_multi(s, d, m, (const MBottom *)E, L.lds, L.own); break;
        case FK_PROLONG: launch_prolong_multi(s, d, L.variant, m, (const MProlong *)E); break;
        case FK_UPDATE_VEL: launch_update_velocity_multi(s, d, vf, m, (const MUpdateVel *)E); break;
        case FK_ADVECT_SMOKE: launch_advect_smoke_multi(s, d, vf, sf, plan.buf, m, (const MAdvectSmoke *)E); break;
        case FK_OBSTACLES: launch_obstacle_update_multi(s, d, m, (const MObstacles *)E); break;
        case FK_COARSE: launch_coarse_smooth_multi(s, d, m, (const MCoarse *)E, L.lds); break;
        }
        if (ev) (void)hipEventRecord((*ev)[li].second, s);
    }
    if (split)
        for (int q = 1; q < 2 * kFusedGroups; ++q)
            if (used[q]) {
                (void)hipEventRecord((*join)[q - 1], stream_of(q));
                (void)hipStreamWaitEvent(s0, (*join)[q - 1], 0);
            }
}

// A plan of one batch step from the scenes' current ping-pong state (left
// unchanged), its tables laid out in one 256-B aligned arena image.
void fused_build(smk_batch *b, FusedPlan &plan, std::vector<StepState> &after, std::vector<unsigned char> &host) {
    std::vector<StepState> before;
    for (smk_grid *g : b->g) before.push_back(g->st);
    fused_plan(b, plan);
    after.clear();
    for (size_t i = 0; i < b->g.size(); ++i) {
        after.push_back(b->g[i]->st);
        b->g[i]->st = before[i];
    }
    size_t bytes = 0;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    for (FusedLaunch &L : plan.launches) {
        L.start_off = bytes;
        bytes = align(bytes + L.start.size() * sizeof(int));
        L.rec_off = bytes;
        bytes = align(bytes + L.recs.size());
    }
    host.assign(std::max<size_t>(bytes, 256), 0);
    for (const FusedLaunch &L : plan.launches) {
        std::memcpy(host.data() + L.start_off, L.start.data(), L.start.size() * sizeof(int));
        std::memcpy(host.data() + L.rec_off, L.recs.data(), L.recs.size());
    }
}

// One fused batch step on streams[0] (the time slots already set).
smk_status fused_step(smk_batch *b) {
    std::string key = batch_key(b);
    auto it = b->fused.find(key);
    if (it == b->fused.end()) {
        // one arena: every launch's prefix sums and records
        FusedPlan plan;
        std::vector<StepState> after;
        std::vector<unsigned char> host;
        fused_build(b, plan, after, host);
        smk_batch::Fused F;
        F.after = after;
        HIPCHK(hipMalloc(&F.arena, host.size()));
        hipError_t e = hipMemcpy(F.arena, host.data(), host.size(), hipMemcpyHostToDevice);
        hipStream_t s0 = b->streams[0];
        hipGraph_t graph = nullptr;
        while (e == hipSuccess && (int)b->side.size() < 2 * kFusedGroups - 1) {
            hipStream_t st = nullptr;
            hipEvent_t ej = nullptr;
            e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
            if (e == hipSuccess) b->side.push_back(st);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ej, hipEventDisableTiming);
            if (e == hipSuccess) b->side_join.push_back(ej);
        }
        if (e == hipSuccess && !b->side_fork) e = hipEventCreateWithFlags(&b->side_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            fused_issue(s0, plan, static_cast<const unsigned char *>(F.arena), nullptr, &b->side, b->side_fork,
                        &b->side_join);
            e = hipStreamEndCapture(s0, &graph);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&F.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            (void)hipFree(F.arena);
            return fail(SMK_ERR_HIP, "fused batch capture failed: %s", hipGetErrorString(e));
        }
        it = b->fused.emplace(key, F).first;
    }
    HIPCHK(hipGraphLaunch(it->second.exec, b->streams[0]));
    for (size_t i = 0; i < b->g.size(); ++i) b->g[i]->st = it->second.after[i];
    return SMK_OK;
}

// The fused mode's autotuner.  A scratch batch holds grids of b's
// descriptions with b's physics, multigrid schedule and scene (emitters,
// inflow, spheres, boxes; zero fields), so every fused launch has b's
// scenes and shapes.  For each candidate width every launch of one batch
// step is issued eagerly between its own pair of events, `reps` times;
// each public launch kind keeps the width with the least time summed over
// its launches (the widths change no result).
smk_status fused_autotune(smk_batch *b, int reps) {
    std::vector<smk_grid_desc> descs;
    for (const smk_grid *g : b->g) descs.push_back(desc_of(g));
    smk_batch *t = nullptr;
    smk_status s = smk_batch_create(descs.data(), (int)descs.size(), 1, &t);
    if (s != SMK_OK) return s;
    for (size_t i = 0; i < b->g.size() && s == SMK_OK; ++i) {
        const smk_grid *g = b->g[i];
        smk_grid *h = t->g[i];
        s = smk_set_physics(h, &g->phys);
        if (s == SMK_OK) s = smk_set_mg_params(h, &g->mg);
        if (s == SMK_OK && !g->em.empty()) s = smk_set_emitters(h, g->em.data(), (int)g->em.size());
        if (s == SMK_OK && g->has_inflow) s = smk_set_emitter_inflow(h, g->inflow.data(), (int)g->inflow.size());
        if (s == SMK_OK && !g->ob.empty()) s = smk_set_obstacles(h, g->ob.data(), (int)g->ob.size());
        if (s == SMK_OK && !g->box.empty()) s = smk_set_box_obstacles(h, g->box.data(), (int)g->box.size());
    }
    if (s == SMK_OK) s = smk_batch_set_mode(t, SMK_BATCH_FUSED);
    if (s == SMK_OK) s = smk_batch_step(t, 1, 0.0f, 0.1f);  // masks, first touch
    if (s == SMK_OK) s = smk_batch_sync(t);
    static const int cand[] = {8, 16, 32, 64, 128};
    double best_ms[SMK_FUSED_KINDS];
    int best_bx[SMK_FUSED_KINDS];
    for (int k = 0; k < SMK_FUSED_KINDS; ++k) {
        best_ms[k] = -1.0;
        best_bx[k] = 0;
    }
    hipStream_t s0 = t->streams[0];
    for (int w : cand) {
        if (s != SMK_OK) break;
        for (int k = 0; k < SMK_FUSED_KINDS; ++k) t->fused_bx[k] = w;
        FusedPlan plan;
        std::vector<StepState> after;
        std::vector<unsigned char> host;
        fused_build(t, plan, after, host);
        void *arena = nullptr;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> ev(plan.launches.size(), {nullptr, nullptr});
        if (hipMalloc(&arena, host.size()) != hipSuccess ||
            hipMemcpy(arena, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
            s = fail(SMK_ERR_HIP, "fused autotune: arena");
        for (auto &e : ev)
            if (s == SMK_OK && (hipEventCreate(&e.first) != hipSuccess || hipEventCreate(&e.second) != hipSuccess))
                s = fail(SMK_ERR_HIP, "fused autotune: events");
        double ms[SMK_FUSED_KINDS] = {};
        for (int r = 0; r <= reps && s == SMK_OK; ++r) {  // round 0 warms up
            fused_issue(s0, plan, static_cast<const unsigned char *>(arena), &ev);
            if (hipStreamSynchronize(s0) != hipSuccess) {
                s = fail(SMK_ERR_HIP, "fused autotune: launch failed");
                break;
            }
            for (size_t li = 0; r > 0 && li < plan.launches.size(); ++li) {
                const int pk = public_kind(plan.launches[li].kind);
                float f = 0.f;
                if (pk >= 0 && hipEventElapsedTime(&f, ev[li].first, ev[li].second) == hipSuccess) ms[pk] += f;
            }
        }
        for (auto &e : ev) {
            if (e.first) (void)hipEventDestroy(e.first);
            if (e.second) (void)hipEventDestroy(e.second);
        }
        if (arena) (void)hipFree(arena);
        for (size_t i = 0; i < t->g.size(); ++i) t->g[i]->st = after[i];
        for (int k = 0; k < SMK_FUSED_KINDS && s == SMK_OK; ++k)
            if (ms[k] > 0.0 && (best_ms[k] < 0.0 || ms[k] < best_ms[k])) {
                best_ms[k] = ms[k];
                best_bx[k] = w;
            }
    }
    smk_batch_destroy(t);
    if (s != SMK_OK) return s;
    for (int k = 0; k < SMK_FUSED_KINDS; ++k) b->fused_bx[k] = best_bx[k];
    // the launches over each scene's linear thread range run 256-thread
    // blocks whatever the width (fused_add_flat, fused_add_sweep)
    for (int k : {SMK_FUSED_SWEEP, SMK_FUSED_VORTICITY, SMK_FUSED_APPLY_VORTICITY, SMK_FUSED_DIVERGENCE,
                  SMK_FUSED_UPDATE_VELOCITY})
        b->fused_bx[k] = 256;
    batch_free_graphs(b);
    return SMK_OK;
}


}  // namespace

extern "C" {

smk_status smk_set_sampler_path(smk_grid *g, int path) {
    GRID_CHECK(g);
    if (path != 0 && path != 1) return fail(SMK_ERR_INVALID_ARG, "sampler path %d is not 0 or 1", path);
    g->sampler = path;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_set_arithmetic(smk_grid *g, int mode) {
    GRID_CHECK(g);
    if (mode != SMK_ARITH_DEFAULT && mode != SMK_ARITH_LLVMPIPE)
        return fail(SMK_ERR_INVALID_ARG, "arithmetic mode %d is not SMK_ARITH_DEFAULT or SMK_ARITH_LLVMPIPE", mode);
    g->arith = mode;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_set_tile(smk_grid *g, int tile_bx) {
    GRID_CHECK(g);
    if (tile_bx != 0 && (tile_bx < 1 || tile_bx > 256 || (tile_bx & (tile_bx - 1))))
        return fail(SMK_ERR_INVALID_ARG, "tile width %d is not a power of two in [1, 256]", tile_bx);
    g->tile_bx = tile_bx;
    invalidate_graphs(g);
    return SMK_OK;
}

smk_status smk_get_tile(smk_grid *g, int *tile_bx) {
    if (!g || !tile_bx) return fail(SMK_ERR_INVALID_ARG, "null argument");
    *tile_bx = g->tile_bx;
    return SMK_OK;
}

smk_status smk_autotune(smk_grid *g, int reps, int *tile_bx, double *ms_per_step) {
    GRID_CHECK(g);
    if (reps < 1) return fail(SMK_ERR_INVALID_ARG, "reps must be >= 1");
    static const int cand[] = {16, 32, 64, 128, 256};
    int best = 0;
    double best_ms = 1e300;
    for (int bx : cand) {
        if (bx / 2 >= g->n[0] + 1 && bx > 16) break;  // wider than the grid: same shape as narrower
        double ms = 0.0;
        smk_status s = time_tile(g, bx, reps, &ms);
        if (s != SMK_OK) return s;
        if (ms < best_ms) {
            best_ms = ms;
            best = bx;
        }
    }
    g->tile_bx = best;
    invalidate_graphs(g);
    if (tile_bx) *tile_bx = best;
    if (ms_per_step) *ms_per_step = best_ms / reps;
    return SMK_OK;
}

smk_status smk_batch_create(const smk_grid_desc *descs, int n, int nstreams, smk_batch **out) {
    if (!descs || !out || n < 1 || nstreams < 1 || nstreams > 16)
        return fail(SMK_ERR_INVALID_ARG, "bad batch description");
    *out = nullptr;
    smk_batch *b = new smk_batch;
    smk_status s = SMK_OK;
    for (int i = 0; i < n && s == SMK_OK; ++i) {
        smk_grid *g = nullptr;
        s = smk_grid_create(&descs[i], &g);
        if (s == SMK_OK) b->g.push_back(g);
    }
    if (s == SMK_OK) {
        HIPCHK(hipSetDevice(descs[0].device));
        for (int k = 0; k < nstreams && s == SMK_OK; ++k) {
            hipStream_t st = nullptr;
            hipEvent_t ev = nullptr, pv = nullptr;
            if (hipStreamCreate(&st) != hipSuccess || hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&pv, hipEventDisableTiming) != hipSuccess)
                s = fail(SMK_ERR_HIP, "batch stream create failed");
            b->streams.push_back(st);
            b->join.push_back(ev);
            b->pre.push_back(pv);
        }
        if (s == SMK_OK && (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
                            hipEventCreateWithFlags(&b->post, hipEventDisableTiming) != hipSuccess))
            s = fail(SMK_ERR_HIP, "batch event create failed");
    }
    if (s != SMK_OK) {
        smk_batch_destroy(b);
        return s;
    }
    for (size_t i = 0; i < b->g.size(); ++i) {
        smk_grid *g = b->g[i];
        (void)hipStreamSynchronize(g->stream);
        g->stream = b->streams[i % b->streams.size()];
        invalidate_graphs(g);
    }
    *out = b;
    return SMK_OK;
}

smk_status smk_batch_destroy(smk_batch *b) {
    if (!b) return SMK_OK;
    for (hipStream_t st : b->streams)
        if (st) (void)hipStreamSynchronize(st);
    batch_free_graphs(b);
    for (smk_grid *g : b->g) {
        g->stream = g->own;
        smk_grid_destroy(g);
    }
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->pre)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->post) (void)hipEventDestroy(b->post);
    for (hipStream_t st : b->side) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (hipEvent_t e : b->side_join)
        if (e) (void)hipEventDestroy(e);
    if (b->side_fork) (void)hipEventDestroy(b->side_fork);
    if (b->d_times) (void)hipFree(b->d_times);
    for (hipStream_t st : b->streams)
        if (st) (void)hipStreamDestroy(st);
    delete b;
    return SMK_OK;
}

smk_status smk_batch_grid(smk_batch *b, int i, smk_grid **g) {
    if (!b || !g || i < 0 || i >= (int)b->g.size()) return fail(SMK_ERR_INVALID_ARG, "bad batch index");
    *g = b->g[(size_t)i];
    return SMK_OK;
}

smk_status smk_batch_autotune(smk_batch *b, int reps, int *tiles) {
    if (!b) return fail(SMK_ERR_INVALID_ARG, "null batch");
    if (reps < 1) return fail(SMK_ERR_INVALID_ARG, "reps must be >= 1");
    if (b->mode == SMK_BATCH_FUSED) {
        smk_status s = fused_autotune(b, reps);
        for (size_t i = 0; s == SMK_OK && tiles && i < b->g.size(); ++i) tiles[i] = b->g[i]->tile_bx;
        return s;
    }
    std::map<std::string, int> seen;  // one tuning per distinct description
    for (size_t i = 0; i < b->g.size(); ++i) {
        smk_grid *g = b->g[i];
        char key[96];
        snprintf(key, sizeof(key), "%d,%d,%d,%d,%d,%d", g->n[0], g->n[1], g->n[2], g->vfmt, g->refine,
                 g->mg.coarse_iterations);
        auto it = seen.find(key);
        if (it == seen.end()) {
            int bx = 0;
            smk_status s = smk_autotune(g, reps, &bx, nullptr);
            if (s != SMK_OK) return s;
            it = seen.emplace(key, bx).first;
        } else {
            smk_status s = smk_set_tile(g, it->second);
            if (s != SMK_OK) return s;
        }
        if (tiles) tiles[i] = it->second;
    }
    batch_free_graphs(b);
    return SMK_OK;
}

smk_status smk_batch_fused_widths(smk_batch *b, int *widths, int n) {
    if (!b || !widths || n < 0) return fail(SMK_ERR_INVALID_ARG, "bad fused widths query");
    for (int k = 0; k < n && k < SMK_FUSED_KINDS; ++k) widths[k] = b->fused_bx[k];
    return SMK_OK;
}

// All scenes' steps in one graph: the scenes are spread over the batch's
// streams, forked from and joined to the first one inside the capture, so
// the graph holds one independent branch per stream.
smk_status smk_batch_set_mode(smk_batch *b, int mode) {
    if (!b || (mode != SMK_BATCH_PER_SCENE && mode != SMK_BATCH_SINGLE_GRAPH && mode != SMK_BATCH_FUSED))
        return fail(SMK_ERR_INVALID_ARG, "bad batch mode");
    std::string why;
    if (mode == SMK_BATCH_FUSED && !fused_ok(b, &why))
        return fail(SMK_ERR_UNSUPPORTED, "fused batch: %s", why.c_str());
    b->mode = mode;
    batch_free_graphs(b);
    return SMK_OK;
}

smk_status smk_batch_step(smk_batch *b, int steps, float t0, float dt_time) {
    if (!b || steps < 0) return fail(SMK_ERR_INVALID_ARG, "bad batch step");
    hipStream_t s0 = b->streams[0];
    if (b->mode == SMK_BATCH_FUSED) {
        std::string why;
        if (!fused_ok(b, &why)) return fail(SMK_ERR_UNSUPPORTED, "fused batch: %s", why.c_str());
        // the static solid maps and multigrid masks, which the fused step
        // leaves out (no-ops once built)
        for (smk_grid *g : b->g) {
            HIPCHK(hipSetDevice(g->dev));
            smk_status s = st_update_solid(g, false);
            if (s != SMK_OK) return s;
            build_masks_if_dirty(g, params(g));
        }
        if (!b->d_times) {
            std::vector<float *> t;
            for (smk_grid *g : b->g) t.push_back(g->d_time);
            HIPCHK(hipMalloc(&b->d_times, t.size() * sizeof(float *)));
            HIPCHK(hipMemcpy(b->d_times, t.data(), t.size() * sizeof(float *), hipMemcpyHostToDevice));
        }
        for (int n = 0; n < steps; ++n) {
            for (size_t k = 1; k < b->streams.size(); ++k) {
                HIPCHK(hipEventRecord(b->pre[k], b->streams[k]));
                HIPCHK(hipStreamWaitEvent(s0, b->pre[k], 0));
            }
            launch_set_times(s0, b->d_times, (int)b->g.size(), t0 + (float)n * dt_time);
            smk_status s = fused_step(b);
            if (s != SMK_OK) return s;
            HIPCHK(hipEventRecord(b->post, s0));
            for (size_t k = 1; k < b->streams.size(); ++k) HIPCHK(hipStreamWaitEvent(b->streams[k], b->post, 0));
        }
        return check_launch("fused batch step");
    }
    if (b->mode == SMK_BATCH_PER_SCENE) {
        // each scene replays its own captured step on its stream; the
        // streams run concurrently
        for (int n = 0; n < steps; ++n)
            for (smk_grid *g : b->g) {
                launch_set_f32(g->stream, g->d_time, t0 + (float)n * dt_time);
                smk_status s = step_graph(g);
                if (s != SMK_OK) return s;
            }
        return SMK_OK;
    }
    for (int n = 0; n < steps; ++n) {
        // the graph runs on s0 but touches every scene: it (and the time
        // updates before it) start after the work queued on the scenes'
        // other streams, and their later work (stage calls, smk_sync, field
        // transfers) waits for it
        for (size_t k = 1; k < b->streams.size(); ++k) {
            HIPCHK(hipEventRecord(b->pre[k], b->streams[k]));
            HIPCHK(hipStreamWaitEvent(s0, b->pre[k], 0));
        }
        for (smk_grid *g : b->g) launch_set_f32(s0, g->d_time, t0 + (float)n * dt_time);
        std::string key = batch_key(b);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            std::vector<StepState> before;
            for (smk_grid *g : b->g) before.push_back(g->st);
            HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
            smk_status s = SMK_OK;
            if (hipEventRecord(b->fork, s0) != hipSuccess) s = fail(SMK_ERR_HIP, "fork record");
            for (size_t k = 1; k < b->streams.size() && s == SMK_OK; ++k)
                if (hipStreamWaitEvent(b->streams[k], b->fork, 0) != hipSuccess) s = fail(SMK_ERR_HIP, "fork wait");
            for (smk_grid *g : b->g)
                if (s == SMK_OK) s = step_impl(g);
            for (size_t k = 1; k < b->streams.size() && s == SMK_OK; ++k) {
                if (hipEventRecord(b->join[k], b->streams[k]) != hipSuccess ||
                    hipStreamWaitEvent(s0, b->join[k], 0) != hipSuccess)
                    s = fail(SMK_ERR_HIP, "join failed");
            }
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(s0, &graph);
            std::vector<StepState> after;
            for (size_t i = 0; i < b->g.size(); ++i) {
                after.push_back(b->g[i]->st);
                b->g[i]->st = before[i];
            }
            if (s != SMK_OK || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return s != SMK_OK ? s : fail(SMK_ERR_HIP, "batch capture failed: %s", hipGetErrorString(e));
            }
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(SMK_ERR_HIP, "batch instantiate failed: %s", hipGetErrorString(e));
            it = b->graphs.emplace(key, std::make_pair(exec, after)).first;
        }
        HIPCHK(hipGraphLaunch(it->second.first, s0));
        HIPCHK(hipEventRecord(b->post, s0));
        for (size_t k = 1; k < b->streams.size(); ++k) HIPCHK(hipStreamWaitEvent(b->streams[k], b->post, 0));
        for (size_t i = 0; i < b->g.size(); ++i) b->g[i]->st = it->second.second[i];
    }
    return SMK_OK;
}

smk_status smk_batch_sync(smk_batch *b) {
    if (!b) return fail(SMK_ERR_INVALID_ARG, "null batch");
    for (hipStream_t st : b->streams) HIPCHK(hipStreamSynchronize(st));
    return SMK_OK;
}

}  // extern "C"
