"""GPU parity of the step's less common paths, each against the float oracle.

- Moving obstacles: non-anchored spheres that fall, bounce off the walls and
  floor and push each other apart (obstacle_update.comp 3D:51-134), with the
  solid map re-rasterised every step (update_solid_map.comp 3D:46-82).
- Open boundary and wind: no border ring (gui.c:121-123) and wind inflow at
  face 0 (advect_vx.comp 3D:58-61).
- clear_grid (grid.c:442-448) between runs.
- Step toggles (physics_step_settings, physics.h:10-18; physics.c:76-100).
- Field export and save/load of a running scene.
- The C host driver (smk_driver), its checksums against the oracle's run of
  the same scene.
Every comparison is bit for bit.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pair(size, sc, storage="fp32", enclosed=True):
    from oracle import Oracle
    from smokesim_amd import Grid
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], storage=storage,
             min_level_size=sc["min_level_size"], enclosed=enclosed)
    o = Oracle(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
               min_level_size=sc["min_level_size"], storage=storage)
    scenes.apply_scene(g, sc, enclosed=enclosed)
    scenes.apply_scene(o, sc, enclosed=enclosed)
    return g, o


def names(dim):
    f = ["vx", "vy", "smoke", "pressure", "residual", "vorticity", "solid"]
    return f + (["vz"] if dim == 3 else [])


def assert_same(g, o, fields, where):
    for f in fields:
        a, b = g.get(f), o.get(f)
        assert a.shape == b.shape, (where, f, a.shape, b.shape)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, (f"{where} {f}: {bad.size} of {a.size} differ, first at {bad[0]}: "
                               f"hip {a[bad[0]]!r} oracle {b[bad[0]]!r}")


def step_both(g, o, t, toggles=None):
    g.set_time(t)
    o.set_time(t)
    g.step()
    o.step(toggles)


# three non-anchored spheres: two thrown at each other (they collide at step
# 3), one rising and drifting in z; walls, floor and the z faces are hit
# within 24 steps (3D: x walls at step 13, floor at 14, z wall at 5)
MOVING = {
    "3d": dict(size=(32, 28, 24), obstacles=[((9, 16, 12), 3, (12, 0, 1)), ((23, 16, 12), 3, (-12, 0, -1)),
                                             ((16, 22, 6), 3, (0, 4, -6))]),
    "2d": dict(size=(64, 48), obstacles=[((20, 30), 4, (12, 0)), ((44, 30), 4, (-12, 0)),
                                         ((32, 40), 4, (3, 6))]),
}


@pytest.mark.parametrize("case", ["3d", "2d"])
def test_moving_obstacles_bit_exact(case, oracle_built):
    c = MOVING[case]
    size = c["size"]
    sc = scenes.make_scene(size, n_emitters=1, emitter_radius=4, min_level_size=4)
    sc["obstacles"] = np.array([scenes.sphere(p, r, v, anchored=0) for p, r, v in c["obstacles"]],
                               np.float32)
    g, o = pair(size, sc)
    n = len(c["obstacles"])
    print(f"moving obstacles {case} {size}", flush=True)
    moved = bounced = 0
    prev = g.get_obstacles()
    for s in range(24):
        step_both(g, o, 0.1 * s)
        ob_g, ob_o = g.get_obstacles(), o.get_obstacles(n)
        assert np.array_equal(ob_g, ob_o), (s, ob_g, ob_o)
        moved += int(np.any(ob_g[:, :3] != prev[:, :3]))
        bounced += int(np.any(np.sign(ob_g[:, 4:7]) * np.sign(prev[:, 4:7]) < 0))
        prev = ob_g
        assert_same(g, o, ["solid"] if s % 4 != 3 else names(len(size)), f"step {s}")
    assert moved == 24 and bounced >= 2, (moved, bounced)
    g.close()
    o.close()


@pytest.mark.parametrize("case", ["3d", "2d"])
def test_open_boundary_with_wind_bit_exact(case, oracle_built):
    size = (24, 20, 16) if case == "3d" else (64, 48)
    wind = (0.4, 0.15, 0.1) if case == "3d" else (0.5, 0.2)
    sc = scenes.make_scene(size, n_emitters=1, emitter_radius=4, n_spheres=1, sphere_radius=3,
                           min_level_size=4)
    sc["physics"][: len(wind)] = wind
    g, o = pair(size, sc, enclosed=False)
    print(f"open boundary {case} wind {wind}", flush=True)
    # no border ring: the domain edge is fluid
    a = g.get("solid").reshape(-1, 4)[:, 3]
    assert a[0] == 0.0
    for s in range(5):
        step_both(g, o, 0.1 * s)
        assert_same(g, o, names(len(size)), f"step {s}")
    g.close()
    o.close()


@pytest.mark.parametrize("case", ["3d", "2d"])
def test_clear_then_steps_bit_exact(case, oracle_built):
    size = (24, 20, 16) if case == "3d" else (64, 48)
    sc = scenes.make_scene(size, n_emitters=1, emitter_radius=4, n_spheres=1, sphere_radius=3,
                           min_level_size=4)
    g, o = pair(size, sc)
    for s in range(3):
        step_both(g, o, 0.1 * s)
    g.clear()
    o.clear_grid()
    # 3D clears the pressure pyramid, 2D keeps it (3d/grid3d.c:181-183)
    assert_same(g, o, names(len(size)), "after clear")
    for s in range(3, 6):
        step_both(g, o, 0.1 * s)
        assert_same(g, o, names(len(size)), f"step {s}")
    g.close()
    o.close()


TOGGLES = ("advect_smoke", "advect_velocities", "apply_vorticity", "apply_buoyancy",
           "resolve_pressure", "update_velocities", "handle_emitters")


@pytest.mark.parametrize("off", list(TOGGLES) + ["all", "advect_velocities+apply_vorticity"])
def test_step_toggles_bit_exact(off, oracle_built):
    size = (24, 20, 16)
    sc = scenes.make_scene(size, n_emitters=2, emitter_radius=4, n_boxes=1, box_sides=(4, 6),
                           min_level_size=4)
    g, o = pair(size, sc)
    offs = set(TOGGLES) if off == "all" else set(off.split("+"))
    kw = {n: 0 if n in offs else 1 for n in TOGGLES}
    g.set_step_toggles(**kw)
    tog = [kw[n] for n in TOGGLES]
    print("toggles", kw, flush=True)
    for s in range(3):
        step_both(g, o, 0.1 * s, tog)
        assert_same(g, o, names(3), f"{off} step {s}")
    assert np.array_equal(g.get_emitters(), o.get_emitters(len(sc["emitters"])))
    g.close()
    o.close()


def test_export_matches_oracle(tmp_path, oracle_built):
    from smokesim_amd import state
    size = (24, 20, 16)
    sc = scenes.make_scene(size, n_emitters=2, emitter_radius=4, n_boxes=1, box_sides=(4, 6),
                           min_level_size=4)
    g, o = pair(size, sc)
    for s in range(3):
        step_both(g, o, 0.1 * s)
    paths = state.export_fields(g, os.path.join(tmp_path, "fields"))
    assert len(paths) == len(state.EXPORT_FIELDS)
    for p in paths:
        f = os.path.basename(p)[:-4]
        a = np.load(p)
        b = o.get(f).astype(np.float32)
        assert a.size == b.size, f
        assert np.array_equal(a.ravel(), b), f"{f}: {np.sum(a.ravel() != b)} values differ"
    g.close()
    o.close()


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_save_load_continuation_matches_oracle(storage, tmp_path, oracle_built):
    from smokesim_amd import state
    size = (24, 20, 16)
    sc = scenes.make_scene(size, n_emitters=2, emitter_radius=4, n_boxes=1, box_sides=(4, 6),
                           min_level_size=4)
    sc["obstacles"] = np.array([scenes.sphere((12, 14, 8), 3, (2, 0, 0), anchored=0)], np.float32)
    g, o = pair(size, sc, storage=storage)
    for s in range(3):
        step_both(g, o, 0.1 * s)
    path = os.path.join(tmp_path, "state.npz")
    state.save(g, path)
    g.close()
    h = state.load(path)
    for s in range(3, 6):
        step_both(h, o, 0.1 * s)
    assert_same(h, o, names(3), f"{storage} continued")
    assert np.array_equal(h.get_obstacles(), o.get_obstacles(1))
    h.close()
    o.close()


def test_bf16_smoke_offset_round_trip(tmp_path):
    """BF16 smoke stores T - t_ambient; a temperature offset with low bits
    survives download, save and load exactly."""
    from smokesim_amd import Grid, state
    size = (12, 10, 8)
    g = Grid(size, storage="bf16", min_level_size=4)
    sm = g.get("smoke").reshape(-1, 4)
    t = 293.0 + 2.0 ** -9 + 2.0 ** -16
    sm[5, 3] = t
    sm[6, 0] = 2.0 ** -9 + 2.0 ** -16
    g.set("smoke", sm.ravel())
    got = g.get("smoke").reshape(-1, 4)
    assert got[5, 3] == t and got[6, 0] == 2.0 ** -9 + 2.0 ** -16
    path = os.path.join(tmp_path, "bf16.npz")
    state.save(g, path)
    h = state.load(path)
    back = h.get("smoke").reshape(-1, 4)
    assert back[5, 3] == t and np.array_equal(back, got)
    g.close()
    h.close()


def test_c_driver_checksums_match_oracle(oracle_built):
    """smk_driver (the main.c loop over the C-ABI, graph replay) against the
    oracle running the scene the driver builds."""
    from oracle import Oracle
    drv = os.path.join(ROOT, "smokesim_amd", "_build", "smk_driver")
    assert os.path.exists(drv), "smk_driver not built"
    nx, ny, nz, steps = 24, 20, 16, 5
    print("running", drv, flush=True)
    out = subprocess.run([drv, str(nx), str(ny), str(nz), str(steps), "fp32"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"total density (\S+), max \|vy\| (\S+)", out.stdout)
    assert m, out.stdout
    dens_hip, vmax_hip = float(m.group(1)), np.float32(m.group(2))
    # the driver's scene (smk_driver.c)
    r = nx // 10 if nx // 10 > 3 else 3
    f32 = np.float32
    o = Oracle((nx, ny, nz), cell_size=0.1, t_ambient=293.0, min_level_size=8)
    o.set_physics(np.array([0, 0, 0, 0, 0.1, 1.0, 293.0, 0.05, 0.1, 0.3, 0, 0], np.float32))
    o.set_mg(2, 2, 80, 4)
    o.set_emitters(np.array([[nx // 2, ny // 8 + r, nz // 2, r, 1, 0, 0, 0, 0, 0, 1, 0]], np.int32))
    o.set_emitter_inflow(np.array([2.0], np.float32))
    o.set_obstacles(np.array([[f32(nx) / f32(2.0), f32(ny) * f32(0.6), f32(nz) / f32(2.0), r,
                               0, 0, 0, 1]], np.float32))
    o.init_solid_map()
    o.update_solid_map()
    for s in range(3 + steps):
        o.set_time(f32(0.0) + f32(s) * f32(0.1) if s < 3 else f32(0.3) + f32(s - 3) * f32(0.1))
        o.step()
    sm = o.get("smoke").astype(np.float32).reshape(-1, 4)[:, 0].astype(np.float64)
    dens = float(np.cumsum(sm)[-1])  # sequential sum in texel order, like the driver
    vmax = np.float32(np.max(np.abs(o.get("vy").astype(np.float32))))
    print("driver", dens_hip, vmax_hip, "oracle", dens, vmax, flush=True)
    assert dens_hip == dens and vmax_hip == vmax
    o.close()
