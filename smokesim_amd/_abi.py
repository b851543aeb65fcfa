"""ctypes binding of libsmoke_hip.so (include/smoke_abi.h)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libsmoke_hip.so")

PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2}
FIELD = {"vx": 0, "vy": 1, "vz": 2, "smoke": 3, "vorticity": 4, "pressure": 5,
         "residual": 6, "solid": 7, "divergence": 8}
KERNEL = {"smooth_l0": 0, "advect_vel": 1, "advect_smoke": 2, "residual_l0": 3}
NOALLOC_MAX_VALUES = 4 << 20  # SMK_NOALLOC_MAX_VALUES

# Structured dtypes with the byte layouts of the reference's structs.
EMITTER_DTYPE = np.dtype([("pos_radius", "<i4", 4), ("density_temp", "<i4", 4),
                          ("emitter_settings", "<i4", 4)])
OBSTACLE_DTYPE = np.dtype([("pos_radius", "<f4", 4), ("velocities_anchored", "<f4", 4)])
BOX_DTYPE = np.dtype([("lo", "<i4", 3), ("hi", "<i4", 3)])


class SmokeError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"smk status {status}: {msg}")
        self.status = status


class _Desc(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int * 3), ("cell_size", ctypes.c_float),
                ("t_ambient", ctypes.c_float), ("storage", ctypes.c_int),
                ("min_level_size", ctypes.c_int), ("enclosed", ctypes.c_int),
                ("device", ctypes.c_int)]


class _Phys(ctypes.Structure):
    _fields_ = [("wind", ctypes.c_float * 4), ("time_step", ctypes.c_float),
                ("density", ctypes.c_float), ("t_ambient", ctypes.c_float),
                ("t_weight_coeff", ctypes.c_float), ("t_temp_coeff", ctypes.c_float),
                ("vorticity_confinement", ctypes.c_float), ("_pad0", ctypes.c_float),
                ("_pad1", ctypes.c_float)]


class _Toggles(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "advect_smoke", "advect_velocities", "apply_vorticity", "apply_buoyancy",
        "resolve_pressure", "update_velocities", "handle_emitters")]


class _MG(ctypes.Structure):
    _fields_ = [("pre_iterations", ctypes.c_int), ("post_iterations", ctypes.c_int),
                ("coarse_iterations", ctypes.c_int), ("v_cycles", ctypes.c_int)]


_LIB = None

_STAGES = ("update_emitters_status", "handle_emitters", "advect_velocities", "apply_buoyancy",
           "calculate_vorticity", "apply_vorticity", "pressure_solve", "update_velocities",
           "advect_smoke", "obstacle_update_step", "update_solid_map", "init_solid_map",
           "run_physics_step", "grid_clear", "sync")

EXPORTS = ("smk_grid_create", "smk_grid_destroy", "smk_grid_clear", "smk_grid_levels",
           "smk_grid_level_size", "smk_set_physics", "smk_set_step_toggles", "smk_set_mg_params",
           "smk_set_emitters", "smk_get_emitters", "smk_set_emitter_inflow", "smk_set_obstacles",
           "smk_get_obstacles", "smk_set_box_obstacles", "smk_set_time",
           "smk_update_emitters_status", "smk_handle_emitters", "smk_advect_velocities",
           "smk_apply_buoyancy", "smk_calculate_vorticity", "smk_apply_vorticity",
           "smk_pressure_solve", "smk_update_velocities", "smk_advect_smoke",
           "smk_obstacle_update_step", "smk_update_solid_map", "smk_init_solid_map",
           "smk_run_physics_step", "smk_run_steps", "smk_field_count", "smk_download_field",
           "smk_upload_field", "smk_sync", "smk_device_sync", "smk_calibrate_copy", "smk_residual_norm", "smk_set_stream",
           "smk_get_stream", "smk_set_graph_mode", "smk_profile_enable", "smk_profile_read",
           "smk_last_error", "smk_abi_version", "smk_set_refinement", "smk_refinement_norms",
           "smk_rollout_create", "smk_rollout_destroy", "smk_rollout_set_initial",
           "smk_rollout_set_source", "smk_rollout_forward", "smk_rollout_backward",
           "smk_rollout_residuals", "smk_rollout_set_target", "smk_rollout_fit", "smk_rollout_fit_result",
           "smk_set_tile", "smk_get_tile", "smk_set_sampler_path", "smk_autotune", "smk_batch_create", "smk_batch_destroy",
           "smk_batch_grid", "smk_batch_autotune", "smk_batch_step", "smk_batch_sync",
           "smk_batch_set_mode", "smk_batch_fused_widths", "smk_download_field_noalloc", "smk_set_arithmetic")


class _RolloutDesc(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("checkpoint_every", ctypes.c_int),
                ("src_lo", ctypes.c_int * 3), ("src_size", ctypes.c_int * 3),
                ("refine_cycles", ctypes.c_int), ("adjoint_cycles", ctypes.c_int)]


def lib():
    """Load libsmoke_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(lib_path):
        raise OSError(f"{lib_path} not built: run python -m smokesim_amd.build")
    L = ctypes.CDLL(lib_path)
    P = ctypes.c_void_p
    I = ctypes.c_int
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    L.smk_last_error.restype = ctypes.c_char_p
    L.smk_grid_create.argtypes = [ctypes.POINTER(_Desc), ctypes.POINTER(P)]
    L.smk_grid_destroy.argtypes = [P]
    L.smk_grid_levels.argtypes = [P, ctypes.POINTER(I)]
    L.smk_grid_level_size.argtypes = [P, I, ctypes.POINTER(I)]
    L.smk_set_physics.argtypes = [P, ctypes.POINTER(_Phys)]
    L.smk_set_step_toggles.argtypes = [P, ctypes.POINTER(_Toggles)]
    L.smk_set_mg_params.argtypes = [P, ctypes.POINTER(_MG)]
    for n in ("smk_set_emitters", "smk_get_emitters", "smk_set_obstacles", "smk_get_obstacles",
              "smk_set_box_obstacles", "smk_set_emitter_inflow"):
        getattr(L, n).argtypes = [P, P, I]
    L.smk_set_time.argtypes = [P, ctypes.c_float]
    for s in _STAGES:
        getattr(L, "smk_" + s).argtypes = [P]
    L.smk_run_steps.argtypes = [P, I, ctypes.c_float, ctypes.c_float]
    L.smk_field_count.argtypes = [P, I, I, ctypes.POINTER(ctypes.c_size_t)]
    L.smk_download_field.argtypes = [P, I, I, P, ctypes.c_size_t, I]
    L.smk_upload_field.argtypes = [P, I, I, P, ctypes.c_size_t, I]
    L.smk_download_field_noalloc.argtypes = [P, I, I, P, ctypes.c_size_t, I]
    L.smk_residual_norm.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    L.smk_device_sync.argtypes = [I]
    L.smk_calibrate_copy.argtypes = [I, ctypes.c_size_t, I, ctypes.POINTER(ctypes.c_double)]
    L.smk_set_stream.argtypes = [P, P]
    L.smk_get_stream.argtypes = [P, ctypes.POINTER(P)]
    L.smk_set_graph_mode.argtypes = [P, I]
    L.smk_profile_enable.argtypes = [P, I]
    L.smk_profile_read.argtypes = [P, I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I)]
    L.smk_set_refinement.argtypes = [P, I, I]
    L.smk_set_tile.argtypes = [P, I]
    L.smk_set_sampler_path.argtypes = [P, I]
    L.smk_set_arithmetic.argtypes = [P, I]
    L.smk_get_tile.argtypes = [P, ctypes.POINTER(I)]
    L.smk_autotune.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_double)]
    L.smk_batch_create.argtypes = [ctypes.POINTER(_Desc), I, I, ctypes.POINTER(P)]
    L.smk_batch_destroy.argtypes = [P]
    L.smk_batch_grid.argtypes = [P, I, ctypes.POINTER(P)]
    L.smk_batch_autotune.argtypes = [P, I, ctypes.POINTER(I)]
    L.smk_batch_step.argtypes = [P, I, ctypes.c_float, ctypes.c_float]
    L.smk_batch_sync.argtypes = [P]
    L.smk_batch_set_mode.argtypes = [P, I]
    L.smk_batch_fused_widths.argtypes = [P, ctypes.POINTER(I), I]
    L.smk_rollout_create.argtypes = [P, ctypes.POINTER(_RolloutDesc), ctypes.POINTER(P)]
    L.smk_rollout_destroy.argtypes = [P]
    L.smk_rollout_set_initial.argtypes = [P, P, P, P, P]
    L.smk_rollout_set_source.argtypes = [P, P]
    L.smk_rollout_forward.argtypes = [P, P]
    L.smk_rollout_backward.argtypes = [P, P, P, P, P, P]
    L.smk_rollout_residuals.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.smk_rollout_set_target.argtypes = [P, P]
    L.smk_rollout_fit.argtypes = [P, ctypes.c_int]
    L.smk_rollout_fit_result.argtypes = [P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), P]
    L.smk_refinement_norms.argtypes = [P, ctypes.POINTER(ctypes.c_double), I]
    _LIB = L
    return L


def device_sync(device=0):
    """hipDeviceSynchronize through the library's HIP runtime.  torch bundles
    a HIP runtime of its own; the two must not both drive the GPU in one
    process, so GPU code here never initialises torch.cuda."""
    _check(lib().smk_device_sync(int(device)))


def calibrate_copy(nbytes, reps=20, device=0):
    """Average milliseconds of one device-to-device copy of nbytes."""
    ms = ctypes.c_double()
    _check(lib().smk_calibrate_copy(int(device), int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def _check(status):
    if status != 0:
        raise SmokeError(status, lib().smk_last_error().decode(errors="replace"))


def desc(size, cell_size=0.1, t_ambient=293.0, storage="fp32", min_level_size=4, enclosed=True,
         device=0):
    size = tuple(int(s) for s in size)
    d = _Desc()
    d.size[0], d.size[1] = size[0], size[1]
    d.size[2] = size[2] if len(size) == 3 else 0
    d.cell_size = cell_size
    d.t_ambient = t_ambient
    d.storage = PRECISION[storage]
    d.min_level_size = min_level_size
    d.enclosed = 1 if enclosed else 0
    d.device = device
    return d


class Grid:
    """One simulation grid on a GPU.  ``size`` is (nx, ny) or (nx, ny, nz)."""

    def __init__(self, size, cell_size=0.1, t_ambient=293.0, storage="fp32",
                 min_level_size=4, enclosed=True, device=0, _borrow=None):
        L = lib()
        self.size = tuple(int(s) for s in size)
        self.dim = len(self.size)
        self._owned = _borrow is None
        if _borrow is not None:
            h = _borrow
        else:
            d = desc(size, cell_size, t_ambient, storage, min_level_size, enclosed, device)
            h = ctypes.c_void_p()
            _check(L.smk_grid_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h
        self._n_em = 0
        self._n_ob = 0
        # the configuration this object was given, for smokesim_amd.state
        self._cfg = {"size": list(self.size), "cell_size": float(cell_size),
                     "t_ambient": float(t_ambient), "storage": storage,
                     "min_level_size": int(min_level_size), "enclosed": bool(enclosed)}

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                lib().smk_grid_destroy(self._h)
            self._h = None

    def set_tile(self, tile_bx):
        _check(lib().smk_set_tile(self._h, int(tile_bx)))
        self._cfg["tile"] = int(tile_bx)

    def set_sampler_path(self, path):
        """0: range-checked buffer samplers where the fields fit (default);
        1: the clamped global-load samplers of large grids, on any grid."""
        _check(lib().smk_set_sampler_path(self._h, int(path)))

    def set_arithmetic(self, mode):
        """"default": binary16 stores rounded to nearest even, lerp as a
        multiply and an add; "llvmpipe": binary16 stores rounded toward zero
        and a fused-multiply-add lerp, as Mesa llvmpipe runs the reference
        shaders."""
        _check(lib().smk_set_arithmetic(self._h, {"default": 0, "llvmpipe": 1}[mode]))
        self._cfg["arithmetic"] = mode

    def tile(self):
        v = ctypes.c_int()
        _check(lib().smk_get_tile(self._h, ctypes.byref(v)))
        return v.value

    def autotune(self, reps=3):
        """Pick the launch tile width by timing a scratch grid; returns
        (tile width, ms per step)."""
        bx, ms = ctypes.c_int(), ctypes.c_double()
        _check(lib().smk_autotune(self._h, int(reps), ctypes.byref(bx), ctypes.byref(ms)))
        self._cfg["tile"] = bx.value
        return bx.value, ms.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- configuration -------------------------------------------------
    @property
    def levels(self):
        n = ctypes.c_int()
        _check(lib().smk_grid_levels(self._h, ctypes.byref(n)))
        return n.value

    def level_size(self, level):
        out = (ctypes.c_int * 3)()
        _check(lib().smk_grid_level_size(self._h, level, out))
        return tuple(out[: self.dim])

    def set_physics(self, time_step=1 / 120, density=1.0, t_ambient=293.0, t_weight_coeff=0.07,
                    t_temp_coeff=1.0, vorticity_confinement=1.0, wind=(0, 0, 0, 0)):
        p = _Phys()
        for i in range(4):
            p.wind[i] = wind[i] if i < len(wind) else 0.0
        p.time_step, p.density, p.t_ambient = time_step, density, t_ambient
        p.t_weight_coeff, p.t_temp_coeff = t_weight_coeff, t_temp_coeff
        p.vorticity_confinement = vorticity_confinement
        _check(lib().smk_set_physics(self._h, ctypes.byref(p)))
        w = [float(x) for x in wind] + [0.0] * (4 - len(wind))
        self._cfg["physics"] = w[:4] + [float(time_step), float(density), float(t_ambient),
                                       float(t_weight_coeff), float(t_temp_coeff),
                                       float(vorticity_confinement), 0.0, 0.0]

    def set_physics_array(self, settings12):
        a = [float(x) for x in settings12]
        self.set_physics(time_step=a[4], density=a[5], t_ambient=a[6], t_weight_coeff=a[7],
                         t_temp_coeff=a[8], vorticity_confinement=a[9], wind=a[0:4])

    def set_step_toggles(self, **kw):
        t = _Toggles()
        for n, _ in _Toggles._fields_:
            setattr(t, n, int(kw.get(n, 1)))
        _check(lib().smk_set_step_toggles(self._h, ctypes.byref(t)))
        self._cfg["toggles"] = {n: int(getattr(t, n)) for n, _ in _Toggles._fields_}

    def set_mg(self, pre, post, coarse, vcycles):
        m = _MG(pre, post, coarse, vcycles)
        _check(lib().smk_set_mg_params(self._h, ctypes.byref(m)))
        self._cfg["mg"] = [int(pre), int(post), int(coarse), int(vcycles)]

    def set_refinement(self, outer, coarse_fp64=True):
        """FP64 iterative refinement of the pressure solve (0: off)."""
        _check(lib().smk_set_refinement(self._h, int(outer), 1 if coarse_fp64 else 0))
        self._refine = int(outer)
        self._cfg["refinement"] = [int(outer), bool(coarse_fp64)]

    def refine_norms(self):
        """(|r| after each residual pass of the last solve, |b|)."""
        n = getattr(self, "_refine", 0) + 2
        out = (ctypes.c_double * n)()
        _check(lib().smk_refinement_norms(self._h, out, n))
        a = np.array(out[:], np.float64)
        return a[:-1], float(a[-1])

    def set_emitters(self, em):
        a = np.ascontiguousarray(np.asarray(em, dtype=np.int32).reshape(-1, 12))
        self._n_em = a.shape[0]
        _check(lib().smk_set_emitters(self._h, a.ctypes.data, a.shape[0]))

    def get_emitters(self):
        a = np.zeros((self._n_em, 12), np.int32)
        _check(lib().smk_get_emitters(self._h, a.ctypes.data, self._n_em))
        return a

    def set_emitter_inflow(self, inflow):
        a = np.ascontiguousarray(np.asarray(inflow, dtype=np.float32).ravel())
        _check(lib().smk_set_emitter_inflow(self._h, a.ctypes.data, a.size))
        self._cfg["inflow"] = a.tolist()

    def set_obstacles(self, ob):
        a = np.ascontiguousarray(np.asarray(ob, dtype=np.float32).reshape(-1, 8))
        self._n_ob = a.shape[0]
        _check(lib().smk_set_obstacles(self._h, a.ctypes.data, a.shape[0]))

    def get_obstacles(self):
        a = np.zeros((self._n_ob, 8), np.float32)
        _check(lib().smk_get_obstacles(self._h, a.ctypes.data, self._n_ob))
        return a

    def set_boxes(self, boxes):
        a = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        _check(lib().smk_set_box_obstacles(self._h, a.ctypes.data, a.shape[0]))
        self._cfg["boxes"] = a.tolist()

    def set_time(self, t):
        _check(lib().smk_set_time(self._h, float(t)))
        self._cfg["time"] = float(t)

    def set_graph_mode(self, on):
        _check(lib().smk_set_graph_mode(self._h, 1 if on else 0))

    def profile(self, on):
        _check(lib().smk_profile_enable(self._h, 1 if on else 0))

    def profile_read(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        _check(lib().smk_profile_read(self._h, KERNEL[kernel], ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def stream(self):
        s = ctypes.c_void_p()
        _check(lib().smk_get_stream(self._h, ctypes.byref(s)))
        return s.value

    # -- fields ---------------------------------------------------------
    def field_count(self, field, level=0):
        n = ctypes.c_size_t()
        _check(lib().smk_field_count(self._h, FIELD[field], level, ctypes.byref(n)))
        return n.value

    def get(self, field, level=0, dtype=np.float64):
        n = self.field_count(field, level)
        out = np.empty(n, dtype=dtype)
        code = 1 if out.dtype == np.float64 else 0
        _check(lib().smk_download_field(self._h, FIELD[field], level, out.ctypes.data,
                                        out.nbytes, code))
        return out

    def get_noalloc(self, field, level=0, dtype=np.float64):
        """get() through smk_download_field_noalloc: staged in the library's
        fixed per-GPU array, no allocation (fields up to
        NOALLOC_MAX_VALUES values)."""
        n = self.field_count(field, level)
        out = np.empty(n, dtype=dtype)
        code = 1 if out.dtype == np.float64 else 0
        _check(lib().smk_download_field_noalloc(self._h, FIELD[field], level, out.ctypes.data,
                                                out.nbytes, code))
        return out

    def set(self, field, data, level=0):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        _check(lib().smk_upload_field(self._h, FIELD[field], level, a.ctypes.data, a.nbytes, 1))

    def residual_norm(self):
        v = ctypes.c_double()
        _check(lib().smk_residual_norm(self._h, ctypes.byref(v)))
        return v.value

    # -- stages (grid.h:36-84) --------------------------------------------
    def _stage(self, name):
        _check(getattr(lib(), "smk_" + name)(self._h))

    def update_emitters_status(self): self._stage("update_emitters_status")
    def handle_emitters(self): self._stage("handle_emitters")
    def advect_velocities(self): self._stage("advect_velocities")
    def apply_buoyancy(self): self._stage("apply_buoyancy")
    def calculate_vorticity(self): self._stage("calculate_vorticity")
    def apply_vorticity(self): self._stage("apply_vorticity")
    def pressure_solve(self): self._stage("pressure_solve")
    def update_velocities(self): self._stage("update_velocities")
    def advect_smoke(self): self._stage("advect_smoke")
    def obstacle_update_step(self): self._stage("obstacle_update_step")
    def update_solid_map(self): self._stage("update_solid_map")
    def init_solid_map(self): self._stage("init_solid_map")
    def run_physics_step(self): self._stage("run_physics_step")
    def clear(self): self._stage("grid_clear")
    def sync(self): self._stage("sync")

    def step(self):
        self.run_physics_step()

    def run_steps(self, n, t0=0.0, dt_time=0.0):
        _check(lib().smk_run_steps(self._h, int(n), float(t0), float(dt_time)))
        if n > 0:
            self._cfg["time"] = float(t0) + (int(n) - 1) * float(dt_time)


class Rollout:
    """A differentiable rollout on a 2D or 3D FP32 grid (smk_rollout_*):
    `steps` differentiable steps with a density source on the box src_lo..
    src_lo + src_size, checkpoints every `checkpoint_every` steps and the
    hand-written adjoint.  In 2D the velocity tuples have two components.
    Each forward projection is refined with `refine_cycles` cycles from a
    cold start; each adjoint projection with `adjoint_cycles` (0: the same
    number) from the previous adjoint step's solution (the first adjoint
    solve, from zero, with at least `refine_cycles`)."""

    def __init__(self, grid, steps, checkpoint_every, src_lo, src_size, refine_cycles=12, adjoint_cycles=0):
        d = _RolloutDesc()
        d.steps, d.checkpoint_every, d.refine_cycles = int(steps), int(checkpoint_every), int(refine_cycles)
        d.adjoint_cycles = int(adjoint_cycles)
        lo = list(src_lo) + [0] * (3 - len(src_lo))
        sz = list(src_size) + [1] * (3 - len(src_size))
        for a in range(3):
            d.src_lo[a] = int(lo[a])
            d.src_size[a] = int(sz[a])
        self.grid = grid
        self.dim = len(grid.size)
        self.size = tuple(grid.size) + (1,) * (3 - self.dim)
        nx, ny, nz = self.size
        if self.dim == 3:
            self.nf = ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1))
        else:
            self.nf = ((nx + 1) * ny, nx * (ny + 1))
        self.nc = nx * ny * nz
        self.nsrc = int(np.prod(sz))
        h = ctypes.c_void_p()
        _check(lib().smk_rollout_create(grid._h, ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().smk_rollout_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f32(a, n):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())
        if a.size != n:
            raise ValueError(f"expected {n} values, got {a.size}")
        return a

    def set_initial(self, *args):
        """set_initial(vx, vy, [vz,] rho): one array per velocity component,
        then the density."""
        if len(args) != self.dim + 1:
            raise ValueError(f"expected {self.dim} velocity components and the density")
        a = [self._f32(v, n) for v, n in zip(args[:-1], self.nf)] + [self._f32(args[-1], self.nc)]
        ptrs = [x.ctypes.data for x in a[:-1]] + [None] * (3 - self.dim) + [a[-1].ctypes.data]
        _check(lib().smk_rollout_set_initial(self._h, *ptrs))

    def set_source(self, src):
        a = self._f32(src, self.nsrc)
        _check(lib().smk_rollout_set_source(self._h, a.ctypes.data))

    def forward(self):
        out = np.empty(self.nc, np.float32)
        _check(lib().smk_rollout_forward(self._h, out.ctypes.data))
        return out

    def backward(self, rho_bar):
        """(dL/dsrc, dL/dvx, dL/dvy[, dL/dvz]) for dL/d(final density)."""
        rb = self._f32(rho_bar, self.nc)
        sb = np.empty(max(self.nsrc, 1), np.float32)
        vb = [np.empty(n, np.float32) for n in self.nf]
        ptrs = [x.ctypes.data for x in vb] + [None] * (3 - self.dim)
        _check(lib().smk_rollout_backward(self._h, rb.ctypes.data, sb.ctypes.data, *ptrs))
        return (sb[: self.nsrc],) + tuple(vb)

    def set_target(self, target):
        """Target final density of the device-side source fit; starts a fit
        (the recorded losses restart)."""
        t = self._f32(target, self.nc)
        _check(lib().smk_rollout_set_target(self._h, t.ctypes.data))

    def fit(self, iterations):
        """Queue `iterations` more fit iterations (one captured graph each) on
        the grid's stream; returns at once."""
        _check(lib().smk_rollout_fit(self._h, int(iterations)))

    def fit_result(self, n):
        """Wait for the fit; (losses before each step since set_target, the
        current source)."""
        out = np.zeros(max(int(n), 1), np.float64)
        rec = ctypes.c_int()
        src = np.empty(max(self.nsrc, 1), np.float32)
        _check(lib().smk_rollout_fit_result(self._h, out.ctypes.data, int(n), ctypes.byref(rec), src.ctypes.data))
        return out[: min(int(n), rec.value)], src[: self.nsrc]

    def residuals(self):
        """(largest relative residual of the last forward's projection solves,
        largest of the last backward's adjoint projection solves)."""
        f, a = ctypes.c_double(), ctypes.c_double()
        _check(lib().smk_rollout_residuals(self._h, ctypes.byref(f), ctypes.byref(a)))
        return f.value, a.value


BATCH_MODE = {"per_scene": 0, "single_graph": 1, "fused": 2}
# the kinds of fused launches, in SMK_FUSED_* order
FUSED_KINDS = ("emit", "advect_velocity", "vorticity", "apply_vorticity", "divergence", "sweep", "restrict",
               "prolong", "update_velocity", "advect_smoke")


class Batch:
    """A batch of uneven scenes (smk_batch_*).  ``scenes`` is a list of dicts
    with the Grid keyword arguments (size, cell_size, t_ambient, storage,
    min_level_size).  ``mode``: "per_scene" (each scene replays its own
    captured step on one of `streams` streams), "single_graph" (all scenes'
    steps in one graph) or "fused" (one launch per stage for all scenes,
    one graph per batch step)."""

    def __init__(self, scenes, streams=4, single_graph=False, mode=None):
        arr = (_Desc * len(scenes))()
        for i, s in enumerate(scenes):
            arr[i] = desc(**s)
        h = ctypes.c_void_p()
        _check(lib().smk_batch_create(arr, len(scenes), int(streams), ctypes.byref(h)))
        self._h = h
        mode = mode or ("single_graph" if single_graph else "per_scene")
        if mode != "per_scene":
            _check(lib().smk_batch_set_mode(h, BATCH_MODE[mode]))
        self.grids = []
        for i, s in enumerate(scenes):
            gh = ctypes.c_void_p()
            _check(lib().smk_batch_grid(h, i, ctypes.byref(gh)))
            self.grids.append(Grid(s["size"], _borrow=gh))

    def autotune(self, reps=2):
        tiles = (ctypes.c_int * len(self.grids))()
        _check(lib().smk_batch_autotune(self._h, int(reps), tiles))
        return list(tiles)

    def fused_widths(self):
        """Block width of each kind of fused launch ({kind: width}; 0: the
        idle-lane rule), as set by autotune() in fused mode."""
        w = (ctypes.c_int * len(FUSED_KINDS))()
        _check(lib().smk_batch_fused_widths(self._h, w, len(FUSED_KINDS)))
        return dict(zip(FUSED_KINDS, list(w)))

    def step(self, steps=1, t0=0.0, dt_time=0.1):
        _check(lib().smk_batch_step(self._h, int(steps), float(t0), float(dt_time)))

    def sync(self):
        _check(lib().smk_batch_sync(self._h))

    def close(self):
        if getattr(self, "_h", None):
            for g in self.grids:
                g._h = None
            lib().smk_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
