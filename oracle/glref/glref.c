/*
 * glref.c -- runs SmokeSim's own GLSL compute shaders on Mesa's llvmpipe
 * software renderer, with no window and no display, and writes every field
 * after every simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/golden/make_gl_fixtures.py runs it, in a
 * container that holds the reference's shader sources, to write golden
 * vectors; the fixtures travel, this program and the shaders do not.
 *
 * What is restated (the shaders themselves are compiled unchanged):
 *   - texture formats, filters, wrap modes, border colours and clear values:
 *     3d/grid3d.c:23-66, 2d/grid2d.c:24-62, common/util/texture.c:28-78;
 *   - texture/image/UBO/SSBO binding points: common/graphics/opengl_bindings.h;
 *   - UBO layouts: physics.c:139-159 (grid size, physics_settings, cell
 *     size), grid.c:266-296 (pyramid sizes + level count), grid.c:319-331
 *     (current level, cell size), physics.c:175-186 (time); SSBO layouts:
 *     emitters.c:60-75, obstacles.c:68-86 (count, 16-byte header, array);
 *   - the step: physics.c:72-101 (run_physics_step) and the dispatch sizes,
 *     barriers and ping-pong swaps of grid.c:36-258, 3d/grid3d.c:113-166;
 *   - the set-up of main.c:89-90 (bind_physics_buffers, init_solid_map).
 *
 * usage: glref <shader_root> <case.bin> <out.bin>
 *   shader_root holds shaders3d/ and shaders2d/.  The case and output
 *   formats are described at read_case() and dump_step().  With
 *   GLREF_TIME=1 it times each step instead (llvmpipe threads:
 *   LP_NUM_THREADS) and writes no fields.  GLREF_PRE, GLREF_POST,
 *   GLREF_COARSE, GLREF_VCYCLES and GLREF_MIN_LEVEL replace the reference's
 *   multigrid constants (16, 16, 80, one V-cycle, levels down to 4 cells)
 *   for timing other schedules.
 *
 * The context is created through Mesa's DRI software-rasteriser interface
 * (swrast_dri.so): a core-profile 4.5 context bound with no drawable, GL
 * entry points resolved through libglapi.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* a multigrid schedule constant from the environment, or the reference's */
static int sched(const char *name, int dflt) {
    const char *v = getenv(name);
    return v && atoi(v) > 0 ? atoi(v) : dflt;
}

#include <GL/glcorearb.h>
#include <GL/internal/dri_interface.h>

/* ------------------------------------------------------------------ */
/* GL entry points                                                     */
/* ------------------------------------------------------------------ */

#define GL_FUNCS(X)                                                         \
    X(PFNGLGETSTRINGPROC, glGetString)                                      \
    X(PFNGLGETERRORPROC, glGetError)                                        \
    X(PFNGLGENTEXTURESPROC, glGenTextures)                                  \
    X(PFNGLBINDTEXTUREPROC, glBindTexture)                                  \
    X(PFNGLTEXPARAMETERIPROC, glTexParameteri)                              \
    X(PFNGLTEXPARAMETERFVPROC, glTexParameterfv)                            \
    X(PFNGLTEXIMAGE2DPROC, glTexImage2D)                                    \
    X(PFNGLTEXIMAGE3DPROC, glTexImage3D)                                    \
    X(PFNGLTEXSUBIMAGE2DPROC, glTexSubImage2D)                              \
    X(PFNGLTEXSUBIMAGE3DPROC, glTexSubImage3D)                              \
    X(PFNGLCLEARTEXIMAGEPROC, glClearTexImage)                              \
    X(PFNGLGETTEXIMAGEPROC, glGetTexImage)                                  \
    X(PFNGLPIXELSTOREIPROC, glPixelStorei)                                  \
    X(PFNGLBINDTEXTUREUNITPROC, glBindTextureUnit)                          \
    X(PFNGLBINDIMAGETEXTUREPROC, glBindImageTexture)                        \
    X(PFNGLGENBUFFERSPROC, glGenBuffers)                                    \
    X(PFNGLBINDBUFFERPROC, glBindBuffer)                                    \
    X(PFNGLBUFFERDATAPROC, glBufferData)                                    \
    X(PFNGLBUFFERSUBDATAPROC, glBufferSubData)                              \
    X(PFNGLGETBUFFERSUBDATAPROC, glGetBufferSubData)                        \
    X(PFNGLBINDBUFFERBASEPROC, glBindBufferBase)                            \
    X(PFNGLCREATESHADERPROC, glCreateShader)                                \
    X(PFNGLSHADERSOURCEPROC, glShaderSource)                                \
    X(PFNGLCOMPILESHADERPROC, glCompileShader)                              \
    X(PFNGLGETSHADERIVPROC, glGetShaderiv)                                  \
    X(PFNGLGETSHADERINFOLOGPROC, glGetShaderInfoLog)                        \
    X(PFNGLCREATEPROGRAMPROC, glCreateProgram)                              \
    X(PFNGLATTACHSHADERPROC, glAttachShader)                                \
    X(PFNGLLINKPROGRAMPROC, glLinkProgram)                                  \
    X(PFNGLGETPROGRAMIVPROC, glGetProgramiv)                                \
    X(PFNGLGETPROGRAMINFOLOGPROC, glGetProgramInfoLog)                      \
    X(PFNGLDELETESHADERPROC, glDeleteShader)                                \
    X(PFNGLUSEPROGRAMPROC, glUseProgram)                                    \
    X(PFNGLGETUNIFORMLOCATIONPROC, glGetUniformLocation)                    \
    X(PFNGLUNIFORM1IPROC, glUniform1i)                                      \
    X(PFNGLDISPATCHCOMPUTEPROC, glDispatchCompute)                          \
    X(PFNGLMEMORYBARRIERPROC, glMemoryBarrier)                              \
    X(PFNGLFINISHPROC, glFinish)

#define DECL(T, n) static T p_##n;
GL_FUNCS(DECL)
#undef DECL

#define glGetString p_glGetString
#define glGetError p_glGetError
#define glGenTextures p_glGenTextures
#define glBindTexture p_glBindTexture
#define glTexParameteri p_glTexParameteri
#define glTexParameterfv p_glTexParameterfv
#define glTexImage2D p_glTexImage2D
#define glTexImage3D p_glTexImage3D
#define glTexSubImage2D p_glTexSubImage2D
#define glTexSubImage3D p_glTexSubImage3D
#define glClearTexImage p_glClearTexImage
#define glGetTexImage p_glGetTexImage
#define glPixelStorei p_glPixelStorei
#define glBindTextureUnit p_glBindTextureUnit
#define glBindImageTexture p_glBindImageTexture
#define glGenBuffers p_glGenBuffers
#define glBindBuffer p_glBindBuffer
#define glBufferData p_glBufferData
#define glBufferSubData p_glBufferSubData
#define glGetBufferSubData p_glGetBufferSubData
#define glBindBufferBase p_glBindBufferBase
#define glCreateShader p_glCreateShader
#define glShaderSource p_glShaderSource
#define glCompileShader p_glCompileShader
#define glGetShaderiv p_glGetShaderiv
#define glGetShaderInfoLog p_glGetShaderInfoLog
#define glCreateProgram p_glCreateProgram
#define glAttachShader p_glAttachShader
#define glLinkProgram p_glLinkProgram
#define glGetProgramiv p_glGetProgramiv
#define glGetProgramInfoLog p_glGetProgramInfoLog
#define glDeleteShader p_glDeleteShader
#define glUseProgram p_glUseProgram
#define glGetUniformLocation p_glGetUniformLocation
#define glUniform1i p_glUniform1i
#define glDispatchCompute p_glDispatchCompute
#define glMemoryBarrier p_glMemoryBarrier
#define glFinish p_glFinish

static void die(const char *msg) {
    fprintf(stderr, "glref: %s\n", msg);
    exit(2);
}

static void check_gl(const char *where) {
    GLenum e = glGetError();
    if (e != GL_NO_ERROR) {
        fprintf(stderr, "glref: GL error 0x%x at %s\n", e, where);
        exit(2);
    }
}

/* ------------------------------------------------------------------ */
/* headless context                                                    */
/* ------------------------------------------------------------------ */

static void drawable_info(__DRIdrawable *d, int *x, int *y, int *w, int *h, void *p) {
    (void)d, (void)p;
    *x = *y = 0;
    *w = *h = 1;
}
static void put_image(__DRIdrawable *d, int op, int x, int y, int w, int h, char *data, void *p) {
    (void)d, (void)op, (void)x, (void)y, (void)w, (void)h, (void)data, (void)p;
}
static void get_image(__DRIdrawable *d, int x, int y, int w, int h, char *data, void *p) {
    (void)d, (void)x, (void)y, (void)w, (void)h, (void)data, (void)p;
}
static const __DRIswrastLoaderExtension swrast_loader = {
    .base = {__DRI_SWRAST_LOADER, 1},
    .getDrawableInfo = drawable_info,
    .putImage = put_image,
    .getImage = get_image};
static const __DRIextension *loader_exts[] = {&swrast_loader.base, NULL};

static void make_context(void) {
    /* GLSL 4.60 (#version 460 core in every shader) */
    setenv("MESA_GL_VERSION_OVERRIDE", "4.6", 0);
    setenv("MESA_GLSL_VERSION_OVERRIDE", "460", 0);
    void *glapi = dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL);
    if (!glapi) die(dlerror());
    const char *drv_path = getenv("GLREF_DRI_DRIVER");
    if (!drv_path) drv_path = "/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so";
    void *drv = dlopen(drv_path, RTLD_NOW | RTLD_GLOBAL);
    if (!drv) die(dlerror());
    const __DRIextension **(*get_ext)(void) =
        (const __DRIextension **(*)(void))dlsym(drv, "__driDriverGetExtensions_swrast");
    if (!get_ext) die("driver has no __driDriverGetExtensions_swrast");
    const __DRIextension **ext = get_ext();
    const __DRIcoreExtension *core = NULL;
    const __DRIswrastExtension *sw = NULL;
    for (int i = 0; ext[i]; ++i) {
        if (!strcmp(ext[i]->name, __DRI_CORE)) core = (const __DRIcoreExtension *)ext[i];
        if (!strcmp(ext[i]->name, __DRI_SWRAST)) sw = (const __DRIswrastExtension *)ext[i];
    }
    if (!core || !sw) die("driver lacks the core or swrast extension");
    const __DRIconfig **configs = NULL;
    __DRIscreen *scr = sw->createNewScreen2(0, loader_exts, ext, &configs, NULL);
    if (!scr) die("createNewScreen2 failed");
    uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 5};
    unsigned err = 0;
    __DRIcontext *ctx = sw->createContextAttribs(scr, __DRI_API_OPENGL_CORE,
                                                 configs ? configs[0] : NULL, NULL, 2, attribs,
                                                 &err, NULL);
    if (!ctx) die("createContextAttribs failed");
    if (!core->bindContext(ctx, NULL, NULL)) die("bindContext failed");
    void *(*gpa)(const char *) = (void *(*)(const char *))dlsym(glapi, "_glapi_get_proc_address");
    if (!gpa) die("no _glapi_get_proc_address");
#define LOAD(T, n)                                  \
    p_##n = (T)gpa(#n);                             \
    if (!p_##n) die("missing GL entry point " #n);
    GL_FUNCS(LOAD)
#undef LOAD
    fprintf(stderr, "glref: %s / %s / GLSL %s\n", (const char *)glGetString(GL_VERSION),
            (const char *)glGetString(GL_RENDERER),
            (const char *)glGetString(GL_SHADING_LANGUAGE_VERSION));
}

/* ------------------------------------------------------------------ */
/* shaders                                                             */
/* ------------------------------------------------------------------ */

static char g_root[4096];

static GLuint program(int is_2d, const char *name) {
    char path[4608];
    snprintf(path, sizeof path, "%s/%s/compute/%s.comp", g_root, is_2d ? "shaders2d" : "shaders3d",
             name);
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "glref: cannot open %s\n", path);
        exit(2);
    }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *src = (char *)malloc((size_t)n + 1);
    if (!src || fread(src, 1, (size_t)n, f) != (size_t)n) die("read shader");
    src[n] = 0;
    fclose(f);
    GLuint s = glCreateShader(GL_COMPUTE_SHADER);
    const char *srcs[1] = {src};
    glShaderSource(s, 1, srcs, NULL);
    glCompileShader(s);
    GLint ok = 0;
    char log[4096];
    glGetShaderiv(s, GL_COMPILE_STATUS, &ok);
    if (!ok) {
        glGetShaderInfoLog(s, sizeof log, NULL, log);
        fprintf(stderr, "glref: %s failed to compile:\n%s\n", path, log);
        exit(2);
    }
    GLuint p = glCreateProgram();
    glAttachShader(p, s);
    glLinkProgram(p);
    glGetProgramiv(p, GL_LINK_STATUS, &ok);
    if (!ok) {
        glGetProgramInfoLog(p, sizeof log, NULL, log);
        fprintf(stderr, "glref: %s failed to link:\n%s\n", path, log);
        exit(2);
    }
    glDeleteShader(s);
    free(src);
    return p;
}

enum {
    S_UPDATE_VX, S_UPDATE_VY, S_UPDATE_VZ, S_ADVECT_VX, S_ADVECT_VY, S_ADVECT_VZ,
    S_ADVECT_SMOKE, S_BUOYANCY, S_EMITTERS, S_EMITTER_STATUS, S_VORTICITY, S_APPLY_VORT_X,
    S_APPLY_VORT_Y, S_APPLY_VORT_Z, S_INIT_SOLID, S_OBSTACLE_UPDATE, S_UPDATE_SOLID,
    S_RESIDUAL, S_PROLONG, S_RESTRICT_RESIDUAL, S_RESTRICT_SOLID, S_SMOOTH, S_COUNT
};

/* shader file per program (shader.c: compile_physics_shaders3d/2d) */
static const char *const SHADER_FILES[S_COUNT] = {
    "update_vx", "update_vy", "update_vz", "advect_vx", "advect_vy", "advect_vz",
    "advect_smoke", "apply_buoyancy", "handle_emitters", "update_emitters_status",
    "calculate_vorticity", "apply_vorticity_x", "apply_vorticity_y", "apply_vorticity_z",
    "init_solid_map", "obstacle_update", "update_solid_map",
    "pressure_solve/compute_residual", "pressure_solve/prolong_pressure",
    "pressure_solve/restrict_residual", "pressure_solve/restrict_solid",
    "pressure_solve/smooth"};

/* ------------------------------------------------------------------ */
/* grid state                                                          */
/* ------------------------------------------------------------------ */

/* opengl_bindings.h */
enum { U_VX = 0, U_VY = 1, U_VZ = 2, U_SMOKE = 3, U_PRESSURE = 4, U_SOLID = 5, U_VORT = 6 };
enum { MG_CUR_P = 0, MG_CUR_S = 1, MG_CUR_R = 2, MG_NEXT_P = 3, MG_NEXT_S = 4, MG_NEXT_R = 5,
       MG_NEXT_P_TEX = 5 };
enum { B_PHYS_UBO = 1, B_MG_CONST_UBO = 2, B_MG_VAR_UBO = 3, B_EMITTERS = 4, B_OBST_READ = 5,
       B_OBST_WRITE = 6, B_TIME_UBO = 7 };

#define MAXL 16

typedef struct {
    int is_2d;
    int n[3];      /* nz = 1 in 2D */
    float cell_size;
    float t_amb_grid;
    float settings[12];
    int toggles[7]; /* physics.h:10-18 order */
    int steps, n_em, n_ob, enclosed, pre_raster;
    float *time;
    int32_t *emitters;
    float *obstacles;
    /* textures */
    GLuint vx[2], vy[2], vz[2], smoke[2], vort;
    int vx_i, vy_i, vz_i, smoke_i;
    int L;
    int ls[MAXL][3];
    GLuint p[MAXL], s[MAXL], r[MAXL];
    /* buffers */
    GLuint ubo_phys, ubo_mgc, ubo_mgv, ubo_time, ssbo_em, ssbo_ob[2];
    int ob_i;
    GLuint prog[S_COUNT];
    GLint parity_loc;
} grid;

static GLenum target(const grid *g) { return g->is_2d ? GL_TEXTURE_2D : GL_TEXTURE_3D; }

/* texture.c:28-78: create, set filter / wrap / border, allocate, clear */
static GLuint make_tex(const grid *g, int w, int h, int d, GLenum ifmt, GLenum fmt, GLenum filter,
                       const float *init, GLenum wrap, const float *border) {
    GLuint t;
    GLenum tg = target(g);
    glGenTextures(1, &t);
    glBindTexture(tg, t);
    glTexParameteri(tg, GL_TEXTURE_MIN_FILTER, (GLint)filter);
    glTexParameteri(tg, GL_TEXTURE_MAG_FILTER, (GLint)filter);
    glTexParameteri(tg, GL_TEXTURE_WRAP_S, (GLint)wrap);
    glTexParameteri(tg, GL_TEXTURE_WRAP_T, (GLint)wrap);
    if (!g->is_2d) glTexParameteri(tg, GL_TEXTURE_WRAP_R, (GLint)wrap);
    if (wrap == GL_CLAMP_TO_BORDER && border) glTexParameterfv(tg, GL_TEXTURE_BORDER_COLOR, border);
    if (g->is_2d)
        glTexImage2D(tg, 0, (GLint)ifmt, w, h, 0, fmt, GL_FLOAT, NULL);
    else
        glTexImage3D(tg, 0, (GLint)ifmt, w, h, d, 0, fmt, GL_FLOAT, NULL);
    glClearTexImage(t, 0, fmt, GL_FLOAT, init);
    check_gl("make_tex");
    return t;
}

static void upload(const grid *g, GLuint t, int w, int h, int d, GLenum fmt, GLenum type,
                   const void *data) {
    glBindTexture(target(g), t);
    glPixelStorei(GL_UNPACK_ALIGNMENT, 1);
    if (g->is_2d)
        glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, w, h, fmt, type, data);
    else
        glTexSubImage3D(GL_TEXTURE_3D, 0, 0, 0, 0, w, h, d, fmt, type, data);
    check_gl("upload");
}

static void download(const grid *g, GLuint t, GLenum fmt, GLenum type, void *out) {
    glBindTexture(target(g), t);
    glPixelStorei(GL_PACK_ALIGNMENT, 1);
    glGetTexImage(target(g), 0, fmt, type, out);
    check_gl("download");
}

static GLuint make_ubo(GLuint binding, size_t size, const void *data, GLenum usage) {
    GLuint b;
    glGenBuffers(1, &b);
    glBindBuffer(GL_UNIFORM_BUFFER, b);
    glBufferData(GL_UNIFORM_BUFFER, (GLsizeiptr)size, data, usage);
    glBindBufferBase(GL_UNIFORM_BUFFER, binding, b);
    glBindBuffer(GL_UNIFORM_BUFFER, 0);
    return b;
}

/* 3d/grid3d.c:23-66, 2d/grid2d.c:24-62 */
static void gen_textures(grid *g) {
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const float smoke_init[4] = {0, 0, 0, g->t_amb_grid};
    for (int b = 0; b < 2; ++b) {
        g->vx[b] = make_tex(g, nx + 1, ny, nz, GL_R32F, GL_RED, GL_LINEAR, NULL, GL_CLAMP_TO_BORDER, NULL);
        g->vy[b] = make_tex(g, nx, ny + 1, nz, GL_R32F, GL_RED, GL_LINEAR, NULL, GL_CLAMP_TO_BORDER, NULL);
        if (!g->is_2d)
            g->vz[b] = make_tex(g, nx, ny, nz + 1, GL_R32F, GL_RED, GL_LINEAR, NULL, GL_CLAMP_TO_BORDER, NULL);
        g->smoke[b] = make_tex(g, nx, ny, nz, GL_RGBA16F, GL_RGBA, GL_LINEAR, smoke_init,
                               GL_CLAMP_TO_BORDER, smoke_init);
    }
    if (g->is_2d)
        g->vort = make_tex(g, nx, ny, 1, GL_R32F, GL_RED, GL_LINEAR, NULL, GL_CLAMP_TO_BORDER, NULL);
    else
        g->vort = make_tex(g, nx, ny, nz, GL_RGBA16F, GL_RGBA, GL_LINEAR, NULL, GL_CLAMP_TO_BORDER, NULL);
    /* pyramid: halve while every extent is >= 4 (grid3d.c:44-52); the
       schedule knobs below can raise that floor */
    const int minl = sched("GLREF_MIN_LEVEL", 4);
    int dims[3] = {nx, ny, g->is_2d ? minl : nz};
    g->L = 0;
    while (dims[0] >= minl && dims[1] >= minl && dims[2] >= minl) {
        dims[0] >>= 1, dims[1] >>= 1;
        if (!g->is_2d) dims[2] >>= 1;
        g->L++;
    }
    if (g->L < 1 || g->L > MAXL) die("grid too small or too large for the pyramid");
    int cur[3] = {nx, ny, nz};
    for (int l = 0; l < g->L; ++l) {
        memcpy(g->ls[l], cur, sizeof cur);
        g->p[l] = make_tex(g, cur[0], cur[1], cur[2], GL_R32F, GL_RED, GL_LINEAR, NULL, GL_CLAMP_TO_EDGE, NULL);
        g->s[l] = make_tex(g, cur[0], cur[1], cur[2], GL_RGBA16, GL_RGBA, GL_NEAREST, NULL, GL_CLAMP_TO_EDGE, NULL);
        g->r[l] = make_tex(g, cur[0], cur[1], cur[2], GL_R32F, GL_RED, GL_LINEAR, NULL, GL_CLAMP_TO_EDGE, NULL);
        cur[0] /= 2, cur[1] /= 2;
        if (!g->is_2d) cur[2] /= 2;
    }
}

/* grid.c:266-296 and grid.c:305-331 */
static void upload_mg_constants(grid *g) {
    unsigned char buf[16 * 16 + 16];
    memset(buf, 0, sizeof buf);
    for (int l = 0; l < g->L; ++l) {
        int32_t v[4] = {g->ls[l][0], g->ls[l][1], g->is_2d ? 0 : g->ls[l][2], 0};
        memcpy(buf + 16 * l, v, sizeof v);
    }
    int32_t levels = g->L;
    memcpy(buf + 16 * 16, &levels, sizeof levels);
    g->ubo_mgc = make_ubo(B_MG_CONST_UBO, 16 * 16 + sizeof(int32_t), buf, GL_STATIC_DRAW);
}

static void set_mg_level(const grid *g, int level) {
    int32_t lv = level;
    float cs = g->cell_size * (float)(1 << level);
    glBindBuffer(GL_UNIFORM_BUFFER, g->ubo_mgv);
    glBufferSubData(GL_UNIFORM_BUFFER, 0, sizeof lv, &lv);
    glBufferSubData(GL_UNIFORM_BUFFER, sizeof lv, sizeof cs, &cs);
}

/* physics.c:139-159: ivec4 grid size, physics_settings (48 bytes), cell size */
static void upload_physics(grid *g) {
    unsigned char buf[16 + 48 + 4];
    int32_t gs[4] = {g->n[0], g->n[1], g->is_2d ? 0 : g->n[2], 0};
    memcpy(buf, gs, 16);
    memcpy(buf + 16, g->settings, 48);
    memcpy(buf + 64, &g->cell_size, 4);
    g->ubo_phys = make_ubo(B_PHYS_UBO, sizeof buf, buf, GL_STATIC_DRAW);
}

static void set_time(const grid *g, float t) {
    glBindBuffer(GL_UNIFORM_BUFFER, g->ubo_time);
    glBufferSubData(GL_UNIFORM_BUFFER, 0, sizeof t, &t);
}

/* emitters.c:60-75, obstacles.c:68-86 */
static GLuint make_ssbo(GLuint binding, int count, const void *items, size_t item_size) {
    GLuint b;
    glGenBuffers(1, &b);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, b);
    size_t sz = 16 + (size_t)count * item_size;
    glBufferData(GL_SHADER_STORAGE_BUFFER, (GLsizeiptr)sz, NULL, GL_DYNAMIC_DRAW);
    int32_t c = count;
    glBufferSubData(GL_SHADER_STORAGE_BUFFER, 0, sizeof c, &c);
    if (count) glBufferSubData(GL_SHADER_STORAGE_BUFFER, 16, (GLsizeiptr)(count * item_size), items);
    glBindBufferBase(GL_SHADER_STORAGE_BUFFER, binding, b);
    return b;
}

/*
 * The reference binds 3D images with layered = GL_FALSE (3d/grid3d.c:122-128,
 * 132-140, 143-166).  The GL specification (4.6, 8.26) then binds only layer 0
 * as a 2D image, which Mesa implements: image3D stores beyond z = 0 are
 * dropped.  The driver of the reference's published runs exposes the whole 3D
 * level, and the shaders are written for that, so this host binds 3D levels
 * with layered = GL_TRUE.  2D textures have no layers; the flag is ignored.
 */
#define LAYERED(g) ((g)->is_2d ? GL_FALSE : GL_TRUE)

/* 3d/grid3d.c:113-128, 2d/grid2d.c:96-109 */
static void bind_physics_buffers(const grid *g) {
    glBindTextureUnit(U_VX, g->vx[g->vx_i]);
    glBindTextureUnit(U_VY, g->vy[g->vy_i]);
    if (!g->is_2d) glBindTextureUnit(U_VZ, g->vz[g->vz_i]);
    glBindTextureUnit(U_SMOKE, g->smoke[g->smoke_i]);
    glBindImageTexture(U_VX, g->vx[1 - g->vx_i], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    glBindImageTexture(U_VY, g->vy[1 - g->vy_i], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    if (!g->is_2d) glBindImageTexture(U_VZ, g->vz[1 - g->vz_i], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    glBindImageTexture(U_SMOKE, g->smoke[1 - g->smoke_i], 0, LAYERED(g), 0, GL_READ_WRITE, GL_RGBA16F);
    glBindImageTexture(U_PRESSURE, g->p[0], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    glBindImageTexture(U_SOLID, g->s[0], 0, LAYERED(g), 0, GL_READ_WRITE, GL_RGBA16F);
    glBindImageTexture(U_VORT, g->vort, 0, LAYERED(g), 0, GL_READ_WRITE, g->is_2d ? GL_R32F : GL_RGBA16F);
    check_gl("bind_physics_buffers");
}

/* 3d/grid3d.c:130-141 */
static void bind_mg(const grid *g, int level) {
    int next = level + 1 < g->L - 1 ? level + 1 : g->L - 1;
    glBindTextureUnit(MG_NEXT_P_TEX, g->p[next]);
    glBindImageTexture(MG_CUR_P, g->p[level], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    glBindImageTexture(MG_CUR_S, g->s[level], 0, LAYERED(g), 0, GL_READ_WRITE, GL_RGBA16F);
    glBindImageTexture(MG_CUR_R, g->r[level], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    glBindImageTexture(MG_NEXT_P, g->p[next], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    glBindImageTexture(MG_NEXT_S, g->s[next], 0, LAYERED(g), 0, GL_READ_WRITE, GL_RGBA16F);
    glBindImageTexture(MG_NEXT_R, g->r[next], 0, LAYERED(g), 0, GL_READ_WRITE, GL_R32F);
    check_gl("bind_mg");
}

/* 3d/grid3d.c:143-166: flip the read index and rebind that unit */
static void swap_field(const grid *g, int unit, GLuint *tex, int *idx, GLenum fmt) {
    (void)g;
    *idx = 1 - *idx;
    glBindTextureUnit((GLuint)unit, tex[*idx]);
    glBindImageTexture((GLuint)unit, tex[1 - *idx], 0, LAYERED(g), 0, GL_READ_WRITE, fmt);
}
static void swap_velocities(grid *g) {
    swap_field(g, U_VX, g->vx, &g->vx_i, GL_R32F);
    swap_field(g, U_VY, g->vy, &g->vy_i, GL_R32F);
    if (!g->is_2d) swap_field(g, U_VZ, g->vz, &g->vz_i, GL_R32F);
}

/* the group count of every grid dispatch: (int)ceilf((float)(n + 7) / 8) */
static int groups(int n) { return (int)ceilf((float)(n + 7) / 8); }

static void dispatch(const grid *g, int s, int sx, int sy, int sz, GLbitfield barrier) {
    glUseProgram(g->prog[s]);
    glDispatchCompute((GLuint)groups(sx), (GLuint)groups(sy), (GLuint)groups(sz));
    glMemoryBarrier(barrier);
    check_gl(SHADER_FILES[s]);
}

#define IMG_BARRIER (GL_SHADER_IMAGE_ACCESS_BARRIER_BIT | GL_TEXTURE_FETCH_BARRIER_BIT)


/* grid.c:184-201: `iterations` red-black sweep pairs, parity 0 first */
static void smooth(const grid *g, int iterations, int level) {
    glUseProgram(g->prog[S_SMOOTH]);
    int parity = 0;
    for (int i = 0; i < iterations * 2; ++i) {
        glUniform1i(g->parity_loc, parity);
        glDispatchCompute((GLuint)groups(g->ls[level][0]), (GLuint)groups(g->ls[level][1]),
                          (GLuint)groups(g->is_2d ? 1 : g->ls[level][2]));
        glMemoryBarrier(IMG_BARRIER);
        parity = 1 - parity;
    }
    check_gl("smooth");
}

/* grid.c:151-181: recursive V-cycle, 16 sweeps per side, 80 at the coarsest
   (GLREF_PRE, GLREF_POST, GLREF_COARSE override them, for timing other
   schedules; the shaders are unchanged) */
static void v_cycle(grid *g, int level) {
    const int LAST = sched("GLREF_COARSE", 80), PRE = sched("GLREF_PRE", 16), POST = sched("GLREF_POST", 16);
    set_mg_level(g, level);
    bind_mg(g, level);
    const int *sz = g->ls[level];
    const int z = g->is_2d ? 1 : 0;
    if (level == g->L - 1) {
        smooth(g, LAST, level);
        return;
    }
    smooth(g, PRE, level);
    dispatch(g, S_RESIDUAL, sz[0], sz[1], z ? 1 : sz[2], IMG_BARRIER);
    const int *nz_ = g->ls[level + 1];
    dispatch(g, S_RESTRICT_RESIDUAL, nz_[0], nz_[1], z ? 1 : nz_[2], IMG_BARRIER);
    dispatch(g, S_RESTRICT_SOLID, nz_[0], nz_[1], z ? 1 : nz_[2], IMG_BARRIER);
    glClearTexImage(g->p[level + 1], 0, GL_RED, GL_FLOAT, NULL);
    v_cycle(g, level + 1);
    set_mg_level(g, level);
    bind_mg(g, level);
    dispatch(g, S_PROLONG, sz[0], sz[1], z ? 1 : sz[2], IMG_BARRIER);
    smooth(g, POST, level);
}

/* physics.c:72-101 */
static void run_physics_step(grid *g) {
    const int nx = g->n[0], ny = g->n[1], nz = g->is_2d ? 1 : g->n[2];
    const int *t = g->toggles; /* advect_smoke, advect_velocities, apply_vorticity,
                                  apply_buoyancy, resolve_pressure, update_velocities,
                                  handle_emitters */
    if (t[6]) {
        /* grid.c:76-80 */
        glUseProgram(g->prog[S_EMITTER_STATUS]);
        glDispatchCompute((GLuint)(int)ceilf((float)(g->n_em + 127) / 128), 1, 1);
        glMemoryBarrier(IMG_BARRIER);
        dispatch(g, S_EMITTERS, nx, ny, nz, IMG_BARRIER);
        swap_field(g, U_SMOKE, g->smoke, &g->smoke_i, GL_RGBA16F);
    }
    if (t[1]) {
        dispatch(g, S_ADVECT_VX, nx + 1, ny, nz, IMG_BARRIER);
        dispatch(g, S_ADVECT_VY, nx, ny + 1, nz, IMG_BARRIER);
        if (!g->is_2d) dispatch(g, S_ADVECT_VZ, nx, ny, nz + 1, IMG_BARRIER);
        swap_velocities(g);
    }
    if (t[3]) {
        dispatch(g, S_BUOYANCY, nx, ny + 1, nz, IMG_BARRIER);
        swap_field(g, U_VY, g->vy, &g->vy_i, GL_R32F);
    }
    if (t[2]) {
        dispatch(g, S_VORTICITY, nx, ny, nz, IMG_BARRIER);
        dispatch(g, S_APPLY_VORT_X, nx + 1, ny, nz, IMG_BARRIER);
        dispatch(g, S_APPLY_VORT_Y, nx, ny + 1, nz, IMG_BARRIER);
        if (!g->is_2d) dispatch(g, S_APPLY_VORT_Z, nx, ny, nz + 1, IMG_BARRIER);
        swap_velocities(g);
    }
    /* grid.c:173-175 runs one V-cycle per step (GLREF_VCYCLES for others) */
    if (t[4])
        for (int v = sched("GLREF_VCYCLES", 1); v > 0; --v) v_cycle(g, 0);
    bind_physics_buffers(g);
    if (t[5]) {
        dispatch(g, S_UPDATE_VX, nx + 1, ny, nz, IMG_BARRIER);
        dispatch(g, S_UPDATE_VY, nx, ny + 1, nz, IMG_BARRIER);
        if (!g->is_2d) dispatch(g, S_UPDATE_VZ, nx, ny, nz + 1, IMG_BARRIER);
        swap_velocities(g);
    }
    if (t[0]) {
        dispatch(g, S_ADVECT_SMOKE, nx, ny, nz, IMG_BARRIER);
        swap_field(g, U_SMOKE, g->smoke, &g->smoke_i, GL_RGBA16F);
    }
    /* grid.c:138-144: obstacle update, then flip the obstacle buffers */
    glUseProgram(g->prog[S_OBSTACLE_UPDATE]);
    glDispatchCompute((GLuint)(int)ceilf((float)(g->n_ob + 127) / 128), 1, 1);
    glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT | GL_TEXTURE_FETCH_BARRIER_BIT);
    g->ob_i = 1 - g->ob_i;
    glBindBufferBase(GL_SHADER_STORAGE_BUFFER, B_OBST_READ, g->ssbo_ob[g->ob_i]);
    glBindBufferBase(GL_SHADER_STORAGE_BUFFER, B_OBST_WRITE, g->ssbo_ob[1 - g->ob_i]);
    dispatch(g, S_UPDATE_SOLID, nx, ny, nz, IMG_BARRIER);
}

/* ------------------------------------------------------------------ */
/* case input and field output                                         */
/* ------------------------------------------------------------------ */

static void rd(FILE *f, void *p, size_t n) {
    if (n && fread(p, 1, n, f) != n) die("short case file");
}

static size_t cells(const grid *g) { return (size_t)g->n[0] * g->n[1] * g->n[2]; }

/*
 * case.bin, little-endian:
 *   int32  magic 0x524c4753, version 1
 *   int32  nx, ny, nz (nz = 0: 2D)
 *   int32  steps, n_emitters, n_obstacles, enclosed, pre_raster
 *   int32  toggles[7] (physics_step_settings order, physics.h:10-18)
 *   float  cell_size, t_ambient of the grid (smoke clear value and border)
 *   float  physics_settings[12] (physics.h:22-32)
 *   float  time[steps] (the time UBO before each step)
 *   int32  emitters[n_emitters][12] (emitters.h layout)
 *   float  obstacles[n_obstacles][8] (obstacles.h layout)
 *   float  vx, vy[, vz] (read buffers, x fastest)
 *   uint16 smoke[cells][4] (FP16 bit patterns)
 *   float  pressure level 0 (the warm start of the first solve)
 * enclosed runs init_solid_map once (main.c:90); pre_raster then runs
 * update_solid_map once, so obstacles are solid in the first step.
 */
static void read_case(grid *g, const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) die("cannot open the case file");
    int32_t hdr[2], sz[3], cnt[5];
    rd(f, hdr, sizeof hdr);
    if (hdr[0] != 0x524c4753 || hdr[1] != 1) die("bad case header");
    rd(f, sz, sizeof sz);
    rd(f, cnt, sizeof cnt);
    rd(f, g->toggles, sizeof g->toggles);
    g->is_2d = sz[2] == 0;
    g->n[0] = sz[0], g->n[1] = sz[1], g->n[2] = g->is_2d ? 1 : sz[2];
    g->steps = cnt[0], g->n_em = cnt[1], g->n_ob = cnt[2], g->enclosed = cnt[3], g->pre_raster = cnt[4];
    if (g->n[0] < 4 || g->n[1] < 4 || g->n[2] < 1 || g->n[0] > 512 || g->n[1] > 512 || g->n[2] > 512 ||
        g->steps < 0 || g->steps > 100000 || g->n_em < 0 || g->n_em > 4096 || g->n_ob < 0 || g->n_ob > 4096)
        die("case sizes out of range");
    rd(f, &g->cell_size, 4);
    rd(f, &g->t_amb_grid, 4);
    rd(f, g->settings, sizeof g->settings);
    g->time = (float *)calloc((size_t)g->steps + 1, sizeof(float));
    g->emitters = (int32_t *)calloc((size_t)g->n_em * 12 + 1, sizeof(int32_t));
    g->obstacles = (float *)calloc((size_t)g->n_ob * 8 + 1, sizeof(float));
    rd(f, g->time, (size_t)g->steps * 4);
    rd(f, g->emitters, (size_t)g->n_em * 48);
    rd(f, g->obstacles, (size_t)g->n_ob * 32);

    gen_textures(g);
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    float *buf = (float *)malloc((size_t)(nx + 1) * (ny + 1) * (nz + 1) * 4 * sizeof(float));
    rd(f, buf, (size_t)(nx + 1) * ny * nz * 4);
    upload(g, g->vx[0], nx + 1, ny, nz, GL_RED, GL_FLOAT, buf);
    rd(f, buf, (size_t)nx * (ny + 1) * nz * 4);
    upload(g, g->vy[0], nx, ny + 1, nz, GL_RED, GL_FLOAT, buf);
    if (!g->is_2d) {
        rd(f, buf, (size_t)nx * ny * (nz + 1) * 4);
        upload(g, g->vz[0], nx, ny, nz + 1, GL_RED, GL_FLOAT, buf);
    }
    rd(f, buf, cells(g) * 4 * 2);
    upload(g, g->smoke[0], nx, ny, nz, GL_RGBA, GL_HALF_FLOAT, buf);
    rd(f, buf, cells(g) * 4);
    upload(g, g->p[0], nx, ny, nz, GL_RED, GL_FLOAT, buf);
    free(buf);
    fclose(f);
}

static void wr(FILE *f, const void *p, size_t n) {
    if (n && fwrite(p, 1, n, f) != n) die("write failed");
}

static void dump_tex(FILE *f, const grid *g, GLuint t, size_t count, GLenum fmt, GLenum type, size_t esz) {
    void *buf = malloc(count * esz);
    download(g, t, fmt, type, buf);
    wr(f, buf, count * esz);
    free(buf);
}

/*
 * Per step, appended to out.bin:
 *   float  vx, vy[, vz] (the read buffers after the step)
 *   uint16 smoke[cells][4] (FP16 bits)
 *   vorticity: 3D uint16[cells][4] (FP16 bits); 2D float[cells]
 *   per pyramid level: float pressure, float residual, uint16 solid[.][4]
 *     (the solid textures are RGBA16 written through RGBA16F image views,
 *     so their raw 16-bit words are FP16 bit patterns)
 *   int32  emitters[n][12], float obstacles[n][8] (current read buffer)
 */
static void dump_step(FILE *f, grid *g) {
    glMemoryBarrier(GL_ALL_BARRIER_BITS);
    glFinish();
    const int nx = g->n[0], ny = g->n[1], nz = g->n[2];
    dump_tex(f, g, g->vx[g->vx_i], (size_t)(nx + 1) * ny * nz, GL_RED, GL_FLOAT, 4);
    dump_tex(f, g, g->vy[g->vy_i], (size_t)nx * (ny + 1) * nz, GL_RED, GL_FLOAT, 4);
    if (!g->is_2d) dump_tex(f, g, g->vz[g->vz_i], (size_t)nx * ny * (nz + 1), GL_RED, GL_FLOAT, 4);
    dump_tex(f, g, g->smoke[g->smoke_i], cells(g) * 4, GL_RGBA, GL_HALF_FLOAT, 2);
    if (g->is_2d)
        dump_tex(f, g, g->vort, cells(g), GL_RED, GL_FLOAT, 4);
    else
        dump_tex(f, g, g->vort, cells(g) * 4, GL_RGBA, GL_HALF_FLOAT, 2);
    for (int l = 0; l < g->L; ++l) {
        size_t c = (size_t)g->ls[l][0] * g->ls[l][1] * g->ls[l][2];
        dump_tex(f, g, g->p[l], c, GL_RED, GL_FLOAT, 4);
        dump_tex(f, g, g->r[l], c, GL_RED, GL_FLOAT, 4);
        dump_tex(f, g, g->s[l], c * 4, GL_RGBA, GL_UNSIGNED_SHORT, 2);
    }
    if (g->n_em) {
        int32_t *e = (int32_t *)malloc((size_t)g->n_em * 48);
        glBindBuffer(GL_SHADER_STORAGE_BUFFER, g->ssbo_em);
        glGetBufferSubData(GL_SHADER_STORAGE_BUFFER, 16, (GLsizeiptr)g->n_em * 48, e);
        wr(f, e, (size_t)g->n_em * 48);
        free(e);
    }
    if (g->n_ob) {
        float *o = (float *)malloc((size_t)g->n_ob * 32);
        glBindBuffer(GL_SHADER_STORAGE_BUFFER, g->ssbo_ob[g->ob_i]);
        glGetBufferSubData(GL_SHADER_STORAGE_BUFFER, 16, (GLsizeiptr)g->n_ob * 32, o);
        wr(f, o, (size_t)g->n_ob * 32);
        free(o);
    }
    check_gl("dump_step");
}

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: glref <shader_root> <case.bin> <out.bin>\n");
        return 2;
    }
    snprintf(g_root, sizeof g_root, "%s", argv[1]);
    make_context();
    static grid g;
    read_case(&g, argv[2]);
    for (int s = 0; s < S_COUNT; ++s) {
        if (g.is_2d && (s == S_UPDATE_VZ || s == S_ADVECT_VZ || s == S_APPLY_VORT_Z)) continue;
        g.prog[s] = program(g.is_2d, SHADER_FILES[s]);
    }
    g.parity_loc = glGetUniformLocation(g.prog[S_SMOOTH], "parity");
    if (g.parity_loc < 0) die("smooth has no parity uniform");

    /* grid.c:25-34, physics.c:55-60 */
    upload_mg_constants(&g);
    int32_t zero_var[4] = {0, 0, 0, 0};
    float cs0 = g.cell_size;
    memcpy(&zero_var[1], &cs0, 4);
    g.ubo_mgv = make_ubo(B_MG_VAR_UBO, 16, zero_var, GL_DYNAMIC_DRAW);
    upload_physics(&g);
    float t0 = 0.0f;
    g.ubo_time = make_ubo(B_TIME_UBO, sizeof t0, &t0, GL_DYNAMIC_DRAW);
    g.ssbo_em = make_ssbo(B_EMITTERS, g.n_em, g.emitters, 48);
    g.ssbo_ob[0] = make_ssbo(B_OBST_READ, g.n_ob, g.obstacles, 32);
    g.ssbo_ob[1] = make_ssbo(B_OBST_WRITE, g.n_ob, g.obstacles, 32);
    g.ob_i = 0;
    check_gl("buffers");

    /* main.c:89-90 */
    bind_physics_buffers(&g);
    const int nz = g.is_2d ? 1 : g.n[2];
    if (g.enclosed) dispatch(&g, S_INIT_SOLID, g.n[0], g.n[1], nz, IMG_BARRIER);
    if (g.pre_raster) dispatch(&g, S_UPDATE_SOLID, g.n[0], g.n[1], nz, IMG_BARRIER);

    /* GLREF_TIME=1: time each step (glFinish before and after) on stderr and
       write no fields */
    const int timing = getenv("GLREF_TIME") && atoi(getenv("GLREF_TIME")) > 0;
    if (timing) {
        glFinish();
        for (int s = 0; s < g.steps; ++s) {
            struct timespec t0, t1;
            clock_gettime(CLOCK_MONOTONIC, &t0);
            bind_physics_buffers(&g);
            set_time(&g, g.time[s]);
            run_physics_step(&g);
            glFinish();
            clock_gettime(CLOCK_MONOTONIC, &t1);
            fprintf(stderr, "glref: step %d %.3f ms\n", s,
                    (t1.tv_sec - t0.tv_sec) * 1e3 + (t1.tv_nsec - t0.tv_nsec) * 1e-6);
        }
        return 0;
    }
    FILE *out = fopen(argv[3], "wb");
    if (!out) die("cannot open the output file");
    for (int s = 0; s < g.steps; ++s) {
        /* main.c:107-110 binds the grid's buffers every frame; main.c:133-135
           sets the time UBO between frames */
        bind_physics_buffers(&g);
        set_time(&g, g.time[s]);
        run_physics_step(&g);
        dump_step(out, &g);
    }
    fclose(out);
    return 0;
}
