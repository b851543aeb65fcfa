"""Seeded synthetic scenes of the benchmark workloads.

A scene is a plain dict of numpy arrays and scalars: physics settings (the
12 floats of physics_settings, physics.h:22-32), emitters (12 ints each,
emitters.h:10-14), per-emitter inflow, sphere obstacles (8 floats each,
obstacles.h:10-13), box obstacles (6 ints each), the initial velocity
faces and the multigrid schedule.  ``apply_scene`` uploads one to any
object with the Grid interface (the HIP grid, or the oracle in tests).

Defaults the workloads leave open: cell size 0.1 and ambient temperature
293 (the reference's placeholder grid, grid.c:357-366, and physics.c:12),
emitter temperature 0, no falloff, anchored obstacles, 80 coarsest-level
iterations (grid.c:179).
"""
import numpy as np

SEEDS = {"scene": 1, "velocity": 2, "source": 3, "target": 4, "probes": 5, "batch": 6}

BASE_PHYSICS = dict(time_step=0.1, density=1.0, t_ambient=293.0, t_weight_coeff=0.05,
                    t_temp_coeff=0.1, vorticity_confinement=0.3)


def physics_array(p):
    return np.array([0, 0, 0, 0, p["time_step"], p["density"], p["t_ambient"],
                     p["t_weight_coeff"], p["t_temp_coeff"], p["vorticity_confinement"], 0, 0],
                    np.float32)


def emitter(pos, radius, density=(1, 0, 0), temperature=0, falloff=0, delay=0):
    """emitters.c:52-58 get_emitter layout."""
    p = list(pos) + [0] * (3 - len(pos))
    return [int(p[0]), int(p[1]), int(p[2]), int(radius), int(density[0]), int(density[1]),
            int(density[2]), int(temperature), int(delay), 0, 0, int(falloff)]


def sphere(pos, radius, vel=(0, 0, 0), anchored=1):
    """obstacles.c:60-65 get_obstacle layout."""
    p = list(pos) + [0.0] * (3 - len(pos))
    v = list(vel) + [0.0] * (3 - len(vel))
    return [float(p[0]), float(p[1]), float(p[2]), float(radius), float(v[0]), float(v[1]),
            float(v[2]), float(anchored)]


def velocity_noise(size, amplitude, seed):
    """Uniform +-amplitude on every face of the MAC grid."""
    rng = np.random.default_rng(seed)
    dim = len(size)
    out = {}
    for c, name in enumerate(("vx", "vy", "vz")[:dim]):
        shape = [s + (1 if d == c else 0) for d, s in enumerate(size)]
        out[name] = rng.uniform(-amplitude, amplitude, int(np.prod(shape))).astype(np.float32)
    return out


def _overlap_sphere_box(c, r, lo, hi):
    d2 = 0.0
    for x, a, b in zip(c, lo, hi):
        q = min(max(x, a), b)
        d2 += (x - q) ** 2
    return d2 < r * r


def make_scene(size, n_emitters=1, emitter_radius=12, inflow=2.0, n_spheres=0, sphere_radius=16,
               n_boxes=0, box_sides=(16, 32), noise=0.01, seed=SEEDS["scene"],
               vel_seed=SEEDS["velocity"], mg=(2, 2, 80, 4), min_level_size=8, physics=None):
    """Emitters in the bottom quarter, obstacles in the upper three quarters,
    rejecting obstacles that overlap an emitter."""
    size = tuple(int(s) for s in size)
    dim = len(size)
    rng = np.random.default_rng(seed)
    phys = dict(BASE_PHYSICS, **(physics or {}))
    em = []
    centres = []
    for _ in range(n_emitters):
        r = emitter_radius
        c = [int(rng.integers(r + 2, size[0] - r - 2))]
        c.append(int(rng.integers(r + 2, max(r + 3, size[1] // 4))))
        if dim == 3:
            c.append(int(rng.integers(r + 2, size[2] - r - 2)))
        centres.append((c, r))
        em.append(emitter(c, r))
    spheres = []
    tries = 0
    while len(spheres) < n_spheres and tries < 1000:
        tries += 1
        r = sphere_radius
        c = [float(rng.integers(r + 2, size[0] - r - 2)),
             float(rng.integers(max(size[1] // 4 + r, r + 2), size[1] - r - 2))]
        if dim == 3:
            c.append(float(rng.integers(r + 2, size[2] - r - 2)))
        if any(np.linalg.norm(np.subtract(c, ec)) < r + er + 2 for ec, er in centres):
            continue
        spheres.append(sphere(c, r))
    boxes = []
    tries = 0
    while len(boxes) < n_boxes and tries < 1000:
        tries += 1
        side = [int(rng.integers(box_sides[0], box_sides[1] + 1)) for _ in range(dim)]
        lo = [int(rng.integers(2, max(3, size[d] - side[d] - 2))) for d in range(dim)]
        lo[1] = int(rng.integers(max(2, size[1] // 4), max(size[1] // 4 + 1, size[1] - side[1] - 2)))
        hi = [lo[d] + side[d] for d in range(dim)]
        if any(_overlap_sphere_box(ec, er + 2, lo, hi) for ec, er in centres):
            continue
        if dim == 2:
            lo.append(0)
            hi.append(1)
        boxes.append(lo + hi)
    vel = velocity_noise(size, noise, vel_seed) if noise else {}
    return dict(size=size, physics=physics_array(phys), emitters=np.array(em, np.int32).reshape(-1, 12),
                inflow=np.full(len(em), inflow, np.float32),
                obstacles=np.array(spheres, np.float32).reshape(-1, 8),
                boxes=np.array(boxes, np.int32).reshape(-1, 6), velocity=vel, mg=tuple(mg),
                min_level_size=min_level_size, cell_size=0.1, t_ambient=phys["t_ambient"])


def config_2d_256(seed=SEEDS["scene"]):
    """2D 256x256: one emitter r=12 (inflow 2.0), one disc obstacle r=16."""
    return make_scene((256, 256), n_emitters=1, emitter_radius=12, n_spheres=1, sphere_radius=16,
                      seed=seed)


def config_3d_128(seed=SEEDS["scene"]):
    """3D 128^3: two spherical emitters r=8 (inflow 2.0), one box 16-32 cells."""
    return make_scene((128, 128, 128), n_emitters=2, emitter_radius=8, n_boxes=1, seed=seed)


def config_3d_256(seed=SEEDS["scene"]):
    """3D 256^3: four emitters r=12, six boxes 16-32 cells."""
    return make_scene((256, 256, 256), n_emitters=4, emitter_radius=12, n_boxes=6, seed=seed)


def batch_uneven(n=64, seed=SEEDS["batch"]):
    """The batch of uneven scenes: half 2D with sides 128-512, half 3D with
    sides 64-160 (non-cubic), 1-4 emitters and 0-5 obstacles each, seeded."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if i < n // 2:
            size = tuple(int(x) for x in rng.integers(128, 513, 2))
            kw = dict(n_emitters=int(rng.integers(1, 5)), emitter_radius=8, n_spheres=int(rng.integers(0, 6)),
                      sphere_radius=10)
        else:
            size = tuple(int(x) for x in rng.integers(64, 161, 3))
            kw = dict(n_emitters=int(rng.integers(1, 5)), emitter_radius=5, n_boxes=int(rng.integers(0, 6)),
                      box_sides=(6, 14))
        out.append(make_scene(size, seed=seed * 1000 + i, min_level_size=8, **kw))
    return out


def apply_scene(target, scene, upload_velocity=True, enclosed=True):
    """Upload a scene to a Grid (or the oracle) and rasterise the obstacles:
    init_solid_map when the domain is enclosed (done by the HIP grid at
    creation, gui.c:121-123) then update_solid_map, so obstacles exist from
    the first step."""
    if hasattr(target, "set_physics_array"):
        target.set_physics_array(scene["physics"])
    else:
        target.set_physics(scene["physics"])
    target.set_emitters(scene["emitters"])
    if len(scene["emitters"]):
        target.set_emitter_inflow(scene["inflow"])
    target.set_obstacles(scene["obstacles"])
    target.set_boxes(scene["boxes"])
    target.set_mg(*scene["mg"])
    if enclosed and not hasattr(target, "set_physics_array"):
        target.init_solid_map()
    target.update_solid_map()
    if upload_velocity:
        for name, arr in scene["velocity"].items():
            target.set(name, arr)
