"""ctypes wrapper of the CPU oracle (oracle/smk_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module; the product package
(smokesim_amd) never does.

The oracle restates the reference's GLSL compute shaders stage by stage; see
the citations in smk_oracle.c.  Two builds exist: ``Oracle(fp64=False)`` is
the parity oracle (float arithmetic, FP16 storage rounding of RGBA16F
images), ``Oracle(fp64=True)`` is the FP64 oracle (double everywhere, no
storage rounding).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_LIBS = {}

FIELD = {"vx": 0, "vy": 1, "vz": 2, "smoke": 3, "vorticity": 4,
         "pressure": 5, "residual": 6, "solid": 7, "divergence": 8}


def build(force=False):
    """Compile both oracle libraries with make (gcc)."""
    targets = [os.path.join(_BUILD, n) for n in ("liborc_f32.so", "liborc_f64.so")]
    if force or not all(os.path.exists(t) for t in targets):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return targets


def _lib(fp64):
    key = bool(fp64)
    if key in _LIBS:
        return _LIBS[key]
    path = os.path.join(_BUILD, "liborc_f64.so" if fp64 else "liborc_f32.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I = ctypes.c_int
    F = ctypes.c_float
    L.orc_create.restype = P
    L.orc_create.argtypes = [I, I, I, F, F, I]
    L.orc_destroy.argtypes = [P]
    L.orc_levels.argtypes = [P]
    L.orc_levels.restype = I
    L.orc_level_size.argtypes = [P, I, ctypes.POINTER(I)]
    L.orc_set_physics.argtypes = [P, ctypes.POINTER(F)]
    L.orc_set_emitters.argtypes = [P, ctypes.POINTER(I), I]
    L.orc_set_emitter_inflow.argtypes = [P, ctypes.POINTER(F), I]
    L.orc_set_obstacles.argtypes = [P, ctypes.POINTER(F), I]
    L.orc_set_boxes.argtypes = [P, ctypes.POINTER(I), I]
    L.orc_get_emitters.argtypes = [P, ctypes.POINTER(I)]
    L.orc_get_obstacles.argtypes = [P, ctypes.POINTER(F)]
    L.orc_set_time.argtypes = [P, F]
    L.orc_set_mg.argtypes = [P, I, I, I, I]
    L.orc_set_storage.argtypes = [P, I]
    L.orc_set_storage.restype = I
    L.orc_set_f16_rounding.argtypes = [P, I]
    L.orc_set_f16_rounding.restype = I
    L.orc_set_lerp.argtypes = [P, I]
    L.orc_set_lerp.restype = I
    L.orc_set_refinement.argtypes = [P, I, I]
    L.orc_set_refinement.restype = I
    L.orc_get_refine_norms.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    L.orc_rollout.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(I)] + [P] * 5 + [I] + [P] * 6
    L.orc_rollout.restype = I
    for n in ("orc_get_field", "orc_set_field"):
        getattr(L, n).argtypes = [P, I, I, ctypes.POINTER(ctypes.c_double)]
        getattr(L, n).restype = ctypes.c_long
    L.orc_field_count.argtypes = [P, I, I]
    L.orc_field_count.restype = ctypes.c_long
    for n in ("orc_update_emitters_status", "orc_handle_emitters", "orc_apply_inflow",
              "orc_advect_velocities", "orc_apply_buoyancy", "orc_calculate_vorticity",
              "orc_apply_vorticity", "orc_pressure_solve", "orc_update_velocities",
              "orc_advect_smoke", "orc_obstacle_update_step", "orc_update_solid_map",
              "orc_init_solid_map", "orc_clear_grid"):
        getattr(L, n).argtypes = [P]
    L.orc_smooth.argtypes = [P, I, I]
    L.orc_smooth_sweep.argtypes = [P, I, I]
    L.orc_compute_residual.argtypes = [P, I]
    L.orc_restrict.argtypes = [P, I]
    L.orc_prolong.argtypes = [P, I]
    L.orc_v_cycle.argtypes = [P, I]
    L.orc_run_physics_step.argtypes = [P, ctypes.POINTER(I)]
    L.orc_sample.argtypes = [P, I, I, ctypes.POINTER(F), ctypes.POINTER(ctypes.c_double)]
    L.orc_round_f16.argtypes = [F]
    L.orc_round_f16.restype = F
    L.orc_round_f16_rtz.argtypes = [F]
    L.orc_round_f16_rtz.restype = F
    L.orc_threads.restype = I
    L.orc_round_bf16.argtypes = [F]
    L.orc_round_bf16.restype = F
    _LIBS[key] = L
    return L


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def round_f16(x, fp64=False):
    return _lib(fp64).orc_round_f16(float(x))


def round_f16_rtz(x, fp64=False):
    return _lib(fp64).orc_round_f16_rtz(float(x))


def threads(fp64=False):
    """OpenMP threads the oracle's stage loops run on."""
    return _lib(fp64).orc_threads()


def round_bf16(x, fp64=False):
    return _lib(fp64).orc_round_bf16(float(x))


class Oracle:
    """One simulation grid on the CPU.  ``size`` is (nx, ny) or (nx, ny, nz)."""

    STORAGE = {"fp32": 0, "bf16": 1, "fp16": 2}

    def __init__(self, size, cell_size=0.1, t_ambient=293.0, min_level_size=4, fp64=False,
                 storage="fp32"):
        self.L = _lib(fp64)
        self.refine = 0
        self.fp64 = fp64
        self.size = tuple(int(s) for s in size)
        self.dim = len(self.size)
        nz = self.size[2] if self.dim == 3 else 0
        self.g = self.L.orc_create(self.size[0], self.size[1], nz, cell_size, t_ambient,
                                   min_level_size)
        if not self.g:
            raise ValueError("oracle: invalid grid description")
        if self.L.orc_set_storage(self.g, self.STORAGE[storage]) != 0:
            raise ValueError(storage)

    def close(self):
        if self.g:
            self.L.orc_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration -------------------------------------------------
    @property
    def levels(self):
        return self.L.orc_levels(self.g)

    def level_size(self, level):
        out = (ctypes.c_int * 3)()
        self.L.orc_level_size(self.g, level, out)
        return tuple(out[: self.dim])

    def set_physics(self, settings12):
        a = np.ascontiguousarray(settings12, dtype=np.float32)
        assert a.size == 12
        self.L.orc_set_physics(self.g, _ptr(a, ctypes.c_float))

    def set_emitters(self, em):
        a = np.ascontiguousarray(em, dtype=np.int32).reshape(-1, 12)
        self.L.orc_set_emitters(self.g, _ptr(a, ctypes.c_int), a.shape[0])

    def set_emitter_inflow(self, inflow):
        a = np.ascontiguousarray(inflow, dtype=np.float32).ravel()
        self.L.orc_set_emitter_inflow(self.g, _ptr(a, ctypes.c_float), a.size)

    def set_obstacles(self, ob):
        a = np.ascontiguousarray(ob, dtype=np.float32).reshape(-1, 8)
        self.L.orc_set_obstacles(self.g, _ptr(a, ctypes.c_float), a.shape[0])

    def set_boxes(self, boxes):
        a = np.ascontiguousarray(boxes, dtype=np.int32).reshape(-1, 6)
        self.L.orc_set_boxes(self.g, _ptr(a, ctypes.c_int), a.shape[0])

    def get_obstacles(self, n):
        a = np.zeros((n, 8), np.float32)
        self.L.orc_get_obstacles(self.g, _ptr(a, ctypes.c_float))
        return a

    def get_emitters(self, n):
        a = np.zeros((n, 12), np.int32)
        self.L.orc_get_emitters(self.g, _ptr(a, ctypes.c_int))
        return a

    def set_time(self, t):
        self.L.orc_set_time(self.g, float(t))

    def set_mg(self, pre, post, coarse, vcycles):
        self.L.orc_set_mg(self.g, pre, post, coarse, vcycles)

    def set_f16_rounding(self, mode):
        """Rounding of binary16 stores: "rtne" (default) or "rtz" (toward
        zero, as Mesa llvmpipe stores RGBA16F images)."""
        if self.L.orc_set_f16_rounding(self.g, {"rtne": 0, "rtz": 1}[mode]) != 0:
            raise ValueError(mode)

    def set_lerp(self, mode):
        """Linear-filter lerp: "split" (multiply, then add; default) or
        "fused" (one fused multiply-add, as llvmpipe's texture unit)."""
        if self.L.orc_set_lerp(self.g, {"split": 0, "fused": 1}[mode]) != 0:
            raise ValueError(mode)

    def set_refinement(self, outer, coarse_fp64=True):
        if self.L.orc_set_refinement(self.g, int(outer), int(bool(coarse_fp64))) != 0:
            raise ValueError("oracle: refinement setup failed")
        self.refine = int(outer)

    def refine_norms(self):
        """(|r| after each residual pass of the last solve, |b|)."""
        out = np.zeros(self.refine + 2, np.float64)
        self.L.orc_get_refine_norms(self.g, _ptr(out, ctypes.c_double))
        return np.sqrt(out[:-1]), float(np.sqrt(out[-1]))

    def rollout(self, steps, src_lo, src_size, src, v0, rho0, cycles=30, rho_bar=None):
        """Differentiable rollout (2D or 3D): the final density and, with
        rho_bar, (dL/dsrc, dL/dvx, dL/dvy[, dL/dvz]) of the FP64 C adjoint.
        v0 holds one array per velocity component."""
        d = len(v0)
        lo = (ctypes.c_int * 3)(*([int(x) for x in src_lo] + [0] * (3 - len(src_lo))))
        sz = (ctypes.c_int * 3)(*([int(x) for x in src_size] + [1] * (3 - len(src_size))))
        f = lambda a: np.ascontiguousarray(np.asarray(a, np.float64).ravel())
        src, r0 = f(src), f(rho0)
        vs = [f(v) for v in v0]
        out = np.empty(r0.size, np.float64)
        grads = None
        ptr = lambda a: a.ctypes.data if a is not None else None
        if rho_bar is not None:
            rb = f(rho_bar)
            grads = [np.empty(src.size, np.float64)] + [np.empty(v.size, np.float64) for v in vs]
        else:
            rb = None
        vp = [ptr(v) for v in vs] + [None] * (3 - d)
        gp = ([ptr(x) for x in grads] + [None] * (3 - d)) if grads else [None] * 4
        rc = self.L.orc_rollout(self.g, int(steps), lo, sz, src.ctypes.data, *vp, r0.ctypes.data,
                                int(cycles), ptr(rb), out.ctypes.data, *gp)
        if rc != 0:
            raise ValueError("oracle rollout failed")
        return (out, grads) if grads else out

    # -- fields ----------------------------------------------------------
    def field_shape(self, field, level=0):
        f = FIELD[field]
        n = self.L.orc_field_count(self.g, f, level)
        if n < 0:
            raise KeyError(field)
        return n

    def get(self, field, level=0):
        f = FIELD[field]
        n = self.L.orc_field_count(self.g, f, level)
        if n < 0:
            raise KeyError(field)
        out = np.empty(n, np.float64)
        self.L.orc_get_field(self.g, f, level, _ptr(out, ctypes.c_double))
        return out

    def set(self, field, data, level=0):
        f = FIELD[field]
        n = self.L.orc_field_count(self.g, f, level)
        a = np.ascontiguousarray(data, dtype=np.float64).ravel()
        if a.size != n:
            raise ValueError(f"{field}: expected {n} values, got {a.size}")
        self.L.orc_set_field(self.g, f, level, _ptr(a, ctypes.c_double))

    def sample(self, field, uvw, level=0):
        u = np.zeros(3, np.float32)
        u[: len(uvw)] = uvw
        out = np.zeros(4, np.float64)
        self.L.orc_sample(self.g, FIELD[field], level, _ptr(u, ctypes.c_float),
                          _ptr(out, ctypes.c_double))
        return out

    # -- stages ------------------------------------------------------------
    def __getattr__(self, name):
        # stage shortcuts: o.advect_velocities() -> orc_advect_velocities(g)
        fn = getattr(self.__dict__.get("L"), "orc_" + name, None) if "L" in self.__dict__ else None
        if fn is None:
            raise AttributeError(name)

        def call(*args):
            return fn(self.g, *args)
        return call

    def step(self, toggles=None):
        if toggles is None:
            self.L.orc_run_physics_step(self.g, None)
        else:
            t = np.ascontiguousarray(toggles, dtype=np.int32)
            self.L.orc_run_physics_step(self.g, _ptr(t, ctypes.c_int))
