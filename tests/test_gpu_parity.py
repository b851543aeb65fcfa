"""GPU parity: the HIP step (through the C-ABI) against the CPU oracle.

Both sides receive identical seeded inputs.  Bars: solid masks bit-exact;
velocity, pressure, smoke and vorticity relative L2 <= 1e-5 per step (the
floating-point tolerance of the parity contract), and whole steps bit-exact.
"""
import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return d / n if n > 0 else d


def fields(dim):
    f = ["vx", "vy", "smoke", "pressure", "vorticity", "solid"]
    if dim == 3:
        f.insert(2, "vz")
    return f


def compare(grid, orc, names, tol=TOL, where=""):
    worst = {}
    for f in names:
        g = grid.get(f)
        o = orc.get(f)
        assert g.shape == o.shape, f"{f}: {g.shape} vs {o.shape}"
        if f == "solid":
            assert np.array_equal(g, o), f"{where} solid mask differs"
            continue
        e = rel_l2(g, o)
        worst[f] = e
        assert e <= tol, f"{where} {f}: rel L2 {e:.3e} > {tol:.0e} (max abs {np.abs(g - o).max():.3e})"
    return worst


def make_pair(size, scene, storage="fp32"):
    from oracle import Oracle
    from smokesim_amd import Grid
    g = Grid(size, cell_size=scene["cell_size"], t_ambient=scene["t_ambient"],
             storage=storage, min_level_size=scene["min_level_size"])
    o = Oracle(size, cell_size=scene["cell_size"], t_ambient=scene["t_ambient"],
               min_level_size=scene["min_level_size"], storage=storage)
    scenes.apply_scene(g, scene)
    scenes.apply_scene(o, scene)
    return g, o


SMALL = {
    "2d": dict(size=(64, 48), kw=dict(n_emitters=1, emitter_radius=6, n_spheres=1,
                                         sphere_radius=6, min_level_size=4)),
    "3d": dict(size=(24, 20, 16), kw=dict(n_emitters=2, emitter_radius=4, n_boxes=1,
                                             box_sides=(4, 6), min_level_size=4)),
    # level 0 large enough for the multi-launch sweep kernel in 2D
    "2d_big": dict(size=(160, 120), kw=dict(n_emitters=2, emitter_radius=8, n_spheres=2,
                                               sphere_radius=10, min_level_size=8)),
    # odd sizes: padded colour-split rows, trailing fine cells in restriction
    "3d_odd": dict(size=(27, 21, 19), kw=dict(n_emitters=2, emitter_radius=4, n_boxes=2,
                                                 n_spheres=1, sphere_radius=4, box_sides=(3, 6),
                                                 min_level_size=4)),
    "2d_odd": dict(size=(75, 53), kw=dict(n_emitters=1, emitter_radius=5, n_spheres=1,
                                             sphere_radius=7, min_level_size=4)),
    # level 1 above the one-workgroup bottom, odd sizes, several obstacles
    "3d_mid": dict(size=(72, 64, 48), kw=dict(n_emitters=2, emitter_radius=6, n_boxes=3,
                                                box_sides=(6, 12), n_spheres=1, sphere_radius=6,
                                                min_level_size=8)),
    "3d_mid_odd": dict(size=(70, 58, 46), kw=dict(n_emitters=2, emitter_radius=6, n_boxes=3,
                                                    box_sides=(5, 11), min_level_size=8)),
    # coarsest level 20x13x18 (4680 cells): no LDS bottom, the coarsest
    # level in the single-workgroup smoother with more than 64 KiB of LDS
    "3d_bigcoarse": dict(size=(40, 26, 36), kw=dict(n_emitters=2, emitter_radius=4, n_boxes=2,
                                                      box_sides=(4, 8), min_level_size=8)),
    # level 1 above 64^3 cells
    "3d_big": dict(size=(136, 136, 128), kw=dict(n_emitters=2, emitter_radius=10, n_boxes=4,
                                                   box_sides=(8, 20), min_level_size=8)),
}


@pytest.mark.parametrize("case", ["2d", "3d", "2d_big", "3d_odd", "2d_odd"])
def test_step_parity(case, oracle_built):
    c = SMALL[case]
    sc = scenes.make_scene(c["size"], **c["kw"])
    g, o = make_pair(c["size"], sc)
    print(f"case {case} size {c['size']} levels {g.levels}", flush=True)
    assert g.levels == o.levels
    compare(g, o, ["solid"], where="init")
    for s in range(6):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        worst = compare(g, o, fields(len(c["size"])), where=f"step {s}")
        print(case, s, {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    assert np.array_equal(g.get_emitters(), o.get_emitters(len(sc["emitters"])))


@pytest.mark.parametrize("case", ["3d", "2d", "3d_odd", "2d_big", "3d_mid", "3d_mid_odd", "3d_big",
                                  "3d_bigcoarse"])
def test_step_bit_exact(case, oracle_built):
    """Whole steps (4 V-cycles per solve) agree with the float oracle bit for
    bit, the level-0 residual image of the last V-cycle included."""
    c = SMALL[case]
    sc = scenes.make_scene(c["size"], **c["kw"])
    g, o = make_pair(c["size"], sc)
    print(f"case {case} size {c['size']} levels {g.levels}", flush=True)
    names = fields(len(c["size"])) + ["residual"]
    for s in range(2 if case == "3d_big" else 3):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        for f in names:
            a, b = g.get(f), o.get(f)
            bad = np.nonzero(a != b)[0]
            assert bad.size == 0, (f"step {s} {f}: {bad.size} of {a.size} differ, first at {bad[0]}: "
                                   f"hip {a[bad[0]]!r} oracle {b[bad[0]]!r}")


@pytest.mark.parametrize("case,storage", [("3d", "fp32"), ("2d", "fp32"), ("3d_odd", "fp32"), ("3d", "bf16")])
def test_step_bit_exact_global_samplers(case, storage, oracle_built):
    """The clamped global-load samplers, which grids with a field of 2 GiB
    or more use instead of the range-checked buffer loads, agree with the
    oracle bit for bit too."""
    c = SMALL[case]
    sc = scenes.make_scene(c["size"], **c["kw"])
    g, o = make_pair(c["size"], sc, storage)
    g.set_sampler_path(1)
    print(f"case {case} {storage} size {c['size']} global samplers", flush=True)
    for s in range(3):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        for f in fields(len(c["size"])):
            a, b = g.get(f), o.get(f)
            bad = np.nonzero(a != b)[0]
            assert bad.size == 0, (f"step {s} {f}: {bad.size} of {a.size} differ, first at {bad[0]}: "
                                   f"hip {a[bad[0]]!r} oracle {b[bad[0]]!r}")


@pytest.mark.parametrize("case", ["2d", "3d", "2d_big", "3d_odd"])
def test_stage_parity(case, oracle_built):
    """Each stage separately from the same random state."""
    c = SMALL[case]
    size = c["size"]
    sc = scenes.make_scene(size, **c["kw"])
    g, o = make_pair(size, sc)
    rng = np.random.default_rng(11)
    for f in ["vx", "vy", "vz"][: len(size)]:
        v = rng.uniform(-1.5, 1.5, g.field_count(f))
        g.set(f, v)
        o.set(f, v)
    n = g.field_count("smoke") // 4
    sm = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n),
                   293 + rng.uniform(-5, 5, n)], 1).ravel()
    g.set("smoke", sm)
    o.set("smoke", sm)
    p = rng.uniform(-1, 1, g.field_count("pressure"))
    g.set("pressure", p)
    o.set("pressure", p)
    names = fields(len(size))
    for stage in ["update_emitters_status", "handle_emitters", "advect_velocities",
                  "apply_buoyancy", "calculate_vorticity", "apply_vorticity", "pressure_solve",
                  "update_velocities", "advect_smoke", "obstacle_update_step",
                  "update_solid_map"]:
        getattr(g, stage)()
        getattr(o, stage)()
        if stage == "handle_emitters":
            o.apply_inflow()
        compare(g, o, names, tol=2e-6, where=stage)
        print(case, stage, "ok", flush=True)


@pytest.mark.parametrize("case", ["3d", "2d"])
def test_100_steps_bit_exact(case, oracle_built):
    """No drift at all against the float oracle over 100 steps: every field
    bit-identical at every tenth step and at the end."""
    c = SMALL[case]
    sc = scenes.make_scene(c["size"], **c["kw"])
    g, o = make_pair(c["size"], sc)
    names = fields(len(c["size"])) + ["residual"]
    for s in range(100):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        if (s + 1) % 10:
            continue
        for f in names:
            a, b = g.get(f), o.get(f)
            assert np.array_equal(a, b), f"step {s + 1} {f}: {np.sum(a != b)} of {a.size} differ"
        print(case, "step", s + 1, "bit-exact", flush=True)
    assert np.array_equal(g.get_emitters(), o.get_emitters(len(sc["emitters"])))


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_large_level0_pressure_solve_bit_exact(storage, oracle_built):
    """One V-cycle (2+2 sweeps, 80 coarsest iterations) on 256x256x128, where
    level 0 takes the four-cells-per-thread divergence and sweep and the
    two-coarse-cells-per-thread residual/restriction and prolongation
    kernels: pressure and residual bit-identical to the float oracle (FP32
    or BF16 velocity storage)."""
    size = (256, 256, 128)
    sc = scenes.make_scene(size, n_emitters=2, emitter_radius=10, n_boxes=4, box_sides=(8, 20),
                           min_level_size=8)
    g, o = make_pair(size, sc, storage)
    g.set_mg(2, 2, 80, 1)
    o.set_mg(2, 2, 80, 1)
    rng = np.random.default_rng(21)
    for f in ["vx", "vy", "vz"]:
        v = rng.uniform(-1.0, 1.0, g.field_count(f))
        g.set(f, v)
        o.set(f, v)
    print("large level-0 solve", storage, size, "levels", g.levels, flush=True)
    g.pressure_solve()
    o.pressure_solve()
    for f in ["pressure", "residual"]:
        a, b = g.get(f), o.get(f)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{f}: {bad.size} of {a.size} differ, first at {bad[0]}"


def _config_pair(sc):
    from oracle import Oracle
    from smokesim_amd import Grid
    g = Grid(sc["size"], cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
             min_level_size=sc["min_level_size"])
    o = Oracle(sc["size"], cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
               min_level_size=sc["min_level_size"])
    scenes.apply_scene(g, sc)
    scenes.apply_scene(o, sc)
    return g, o


def test_2d_256_config_100_steps_bit_exact(oracle_built):
    """The 2D 256x256 configuration at its full size (one emitter r=12 with
    inflow, one disc obstacle, 2+2 sweeps x 4 V-cycles): 100 steps, every
    field bit-identical to the float oracle at steps 25, 50, 75 and 100."""
    sc = scenes.config_2d_256()
    g, o = _config_pair(sc)
    names = fields(2) + ["residual"]
    for s in range(100):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        if (s + 1) % 25:
            continue
        for f in names:
            a, b = g.get(f), o.get(f)
            assert np.array_equal(a, b), f"step {s + 1} {f}: {np.sum(a != b)} of {a.size} differ"
        print("2d 256 step", s + 1, "bit-exact", flush=True)
    assert np.array_equal(g.get_emitters(), o.get_emitters(len(sc["emitters"])))


def test_3d_128_config_steps_bit_exact(oracle_built):
    """The 3D 128^3 configuration at its full size (two emitters r=8, one box,
    2+2 sweeps x 4 V-cycles): 3 steps, every field bit-identical to the float
    oracle after each step."""
    sc = scenes.config_3d_128()
    g, o = _config_pair(sc)
    names = fields(3) + ["residual"]
    for s in range(3):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        for f in names:
            a, b = g.get(f), o.get(f)
            assert np.array_equal(a, b), f"step {s + 1} {f}: {np.sum(a != b)} of {a.size} differ"
        print("3d 128 step", s + 1, "bit-exact", flush=True)


def test_3d_128_config_200_steps_match_oracle_trajectory():
    """The 3D 128^3 configuration for its full 200 steps: every 10 steps the
    digest of each field (velocities, smoke, level-0 pressure and residual)
    equals the CPU oracle's digest at that step, from
    tests/golden/trajectory_c2.json (written by make_trajectory_hashes.py):
    bit-identical over the whole run."""
    import hashlib
    import json
    import os
    from smokesim_amd import Grid
    path = os.path.join(os.path.dirname(__file__), "golden", "trajectory_c2.json")
    with open(path) as f:
        ref = json.load(f)
    sc = scenes.config_3d_128()
    assert list(sc["size"]) == ref["size"]
    g = Grid(sc["size"], cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
             min_level_size=sc["min_level_size"])
    scenes.apply_scene(g, sc)
    recs = {r["step"]: r for r in ref["records"]}
    for s in range(ref["steps"]):
        g.set_time(s * 0.1)
        g.step()
        r = recs.get(s + 1)
        if r is None:
            continue
        for f in ("vx", "vy", "vz", "smoke", "pressure", "residual"):
            a = np.ascontiguousarray(g.get(f), np.float64)
            assert hashlib.sha256(a.tobytes()).hexdigest() == r[f], \
                f"step {s + 1} {f}: L2 {np.linalg.norm(a):.9e} vs oracle {r[f + '_l2']:.9e}"
        print("3d 128 step", s + 1, "matches the oracle's digests", flush=True)
    g.close()


def test_3d_256_bench_config_step_bit_exact(oracle_built):
    """The benchmark's own 256^3 scene (four emitters r=12, six boxes, 2+2
    sweeps x 4 V-cycles, 80 coarsest iterations): one whole step bit-identical
    to the float oracle."""
    sc = scenes.config_3d_256()
    g, o = _config_pair(sc)
    g.set_time(0.0)
    o.set_time(0.0)
    print("3d 256 bench scene: one step on the GPU and the oracle", flush=True)
    g.step()
    o.step()
    for f in fields(3) + ["residual"]:
        a, b = g.get(f), o.get(f)
        assert np.array_equal(a, b), f"{f}: {np.sum(a != b)} of {a.size} differ"
