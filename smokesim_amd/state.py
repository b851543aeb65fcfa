"""Scene serialisation and field export for a Grid.

The reference has only stubs for saving scenes (grid_parser.c:1-5,
window.c:177-178), and its renderer samples the live textures
(shaders3d/graphics/ray_march_smoke.frag:6, velocity_arrows.vert:5-7).
Here:

- ``save(grid, path)`` writes one ``.npz`` holding the grid description,
  every setting the Grid object was given, the emitter and obstacle state
  (status flags and positions included), the simulation time and the
  fields that carry state from one step to the next: the three velocity
  components, smoke, the level-0 pressure (the warm start of the next
  solve; FP64 in the refined mode) and the level-0 solid map.
  ``load(path)`` builds a new Grid from it; stepping the loaded grid
  continues the saved run bit for bit.  (Not saved: the stale contents of
  the ping-pong write buffers, which the reference exposes only when
  buoyancy runs without velocity advection, apply_buoyancy.comp:42.)
- ``export_fields(grid, directory)`` writes one ``.npy`` per field, shaped
  (nz, ny, nx') [+ (4,)] in GL texel order (x fastest), for inspection
  with any NumPy-aware tool.
"""
import json
import os

import numpy as np

from ._abi import Grid

STATE_FIELDS = ("vx", "vy", "vz", "smoke", "pressure", "solid")
EXPORT_FIELDS = ("vx", "vy", "vz", "smoke", "vorticity", "pressure", "residual", "solid",
                 "divergence")
FORMAT = 1


def _shape(grid, field):
    """Array shape of `field` (level 0) in texel order: (nz, ny, nx) with the
    staggered axis one longer, and a trailing 4 for 4-channel fields."""
    nx, ny = grid.size[0], grid.size[1]
    nz = grid.size[2] if grid.dim == 3 else 1
    ext = {"vx": (0, 0, 1), "vy": (0, 1, 0), "vz": (1, 0, 0)}.get(field, (0, 0, 0))
    shp = (nz + ext[0], ny + ext[1], nx + ext[2])
    if grid.dim == 2 and field == "vz":
        return None
    count = grid.field_count(field)
    cells = int(np.prod(shp))
    if count == 4 * cells:
        return shp + (4,)
    if count != cells:
        raise ValueError(f"{field}: {count} values do not fit {shp}")
    return shp


def save(grid, path):
    """Write the grid's restartable state to `path` (.npz)."""
    cfg = dict(grid._cfg)
    arrays = {"meta": np.frombuffer(json.dumps({"format": FORMAT, "config": cfg}).encode(), np.uint8),
              "emitters": grid.get_emitters(), "obstacles": grid.get_obstacles()}
    for f in STATE_FIELDS:
        if grid.dim == 2 and f == "vz":
            continue
        # every storage format is exact in FP32, but BF16 smoke holds
        # T - t_ambient, and T itself (t_ambient + a small offset) needs
        # FP64; the refined mode keeps the level-0 pressure in FP64
        fp64 = ((f == "pressure" and cfg.get("refinement", [0])[0] > 0)
                or (f == "smoke" and cfg.get("storage") == "bf16"))
        arrays[f] = grid.get(f, dtype=np.float64 if fp64 else np.float32)
    np.savez_compressed(path, **arrays)


def load(path, device=0):
    """A new Grid holding the state saved in `path`."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        if meta.get("format") != FORMAT:
            raise ValueError(f"{path}: unknown state format {meta.get('format')}")
        c = meta["config"]
        g = Grid(tuple(c["size"]), cell_size=c["cell_size"], t_ambient=c["t_ambient"],
                 storage=c["storage"], min_level_size=c["min_level_size"], enclosed=c["enclosed"],
                 device=device)
        if "physics" in c:
            g.set_physics_array(c["physics"])
        if "mg" in c:
            g.set_mg(*c["mg"])
        if "toggles" in c:
            g.set_step_toggles(**c["toggles"])
        if len(z["emitters"]):
            g.set_emitters(z["emitters"])
        if "inflow" in c:
            g.set_emitter_inflow(c["inflow"])
        if len(z["obstacles"]):
            g.set_obstacles(z["obstacles"])
        if "boxes" in c:
            g.set_boxes(c["boxes"])
        if "refinement" in c:
            g.set_refinement(*c["refinement"])
        if "arithmetic" in c:
            g.set_arithmetic(c["arithmetic"])
        for f in STATE_FIELDS:
            if f in z.files:
                g.set(f, z[f])
        if "time" in c:
            g.set_time(c["time"])
        if c.get("tile"):
            g.set_tile(c["tile"])
    return g


def export_fields(grid, directory, fields=EXPORT_FIELDS):
    """One float32 .npy per field (level 0) under `directory`; returns the
    paths written."""
    os.makedirs(directory, exist_ok=True)
    out = []
    for f in fields:
        shp = _shape(grid, f)
        if shp is None:
            continue
        a = grid.get(f, dtype=np.float32).reshape(shp)
        p = os.path.join(directory, f + ".npy")
        np.save(p, a)
        out.append(p)
    return out
