"""The HIP step against the reference's own GLSL shaders (GPU).

tests/golden/gl_*.npz hold seeded scenes and the fields that SmokeSim's
compute shaders produced for them on Mesa llvmpipe (stage by stage, three
whole steps, and 100-step runs with emitters and a falling obstacle).  With
smk_set_arithmetic(SMK_ARITH_LLVMPIPE) the HIP kernels make llvmpipe's two
implementation choices (binary16 stores rounded toward zero, fused-lerp
linear filter); every case is replayed through the C-ABI and checked

1. bit for bit against the float oracle in the same mode, and
2. against the shaders' outputs with the bars of test_gl_fixtures.py: bit
   for bit for the stages without divisions (emitters, advection, buoyancy,
   smoke advection, solid maps, obstacles), within the written tolerances
   for the stages whose per-cell divisions llvmpipe does not round as IEEE
   division (vorticity, pressure).
"""
import numpy as np
import pytest

import test_gl_fixtures as T

pytestmark = pytest.mark.gpu

TOGGLES = ("advect_smoke", "advect_velocities", "apply_vorticity", "apply_buoyancy",
           "resolve_pressure", "update_velocities", "handle_emitters")


def hip_grid(fx, size, enclosed, mode="llvmpipe"):
    from smokesim_amd import Grid
    g = Grid(size, cell_size=float(fx["in_cell_size"]), t_ambient=float(fx["in_t_ambient"]),
             min_level_size=4, enclosed=enclosed)
    g.set_arithmetic(mode)  # before the uploads and the first rasterisation
    g.set_physics_array(fx["in_settings"])
    g.set_emitters(fx["in_emitters"])
    g.set_obstacles(fx["in_obstacles"])
    for f in ("vx", "vy", "vz")[: len(size)]:
        g.set(f, fx["in_" + f].astype(np.float32).ravel())
    g.set("smoke", fx["in_smoke"].astype(np.float64))
    g.set("pressure", fx["in_pressure0"], 0)
    # the grid wrote the border ring at creation when enclosed; then one
    # update_solid_map, as the fixture runner does
    g.update_solid_map()
    return g


def size_of(fx):
    n = [int(v) for v in fx["in_size"]]
    return tuple(n[:2]) if n[2] == 0 else tuple(n)


def hip_replay(fx, stage):
    size = size_of(fx)
    g = hip_grid(fx, size, bool(int(fx[f"{stage}__enclosed"])))
    tog = [int(x) for x in fx[f"{stage}__toggles"]]
    g.set_step_toggles(**dict(zip(TOGGLES, tog)))
    out = []
    for t in fx[f"{stage}__times"]:
        g.set_time(float(t))
        g.step()
        st = {"vx": g.get("vx"), "vy": g.get("vy"), "smoke": g.get("smoke"),
              "vorticity": g.get("vorticity"), "emitters": g.get_emitters(),
              "obstacles": g.get_obstacles()}
        if len(size) == 3:
            st["vz"] = g.get("vz")
        for lv in range(g.levels):
            st[f"pressure{lv}"] = g.get("pressure", lv)
            st[f"residual{lv}"] = g.get("residual", lv)
            st[f"solid{lv}"] = g.get("solid", lv)
        out.append(st)
    g.close()
    return out


@pytest.mark.parametrize("scene", T.SCENES)
def test_hip_matches_reference_shaders(scene, oracle_built):
    fx = T.load(scene)
    for stage in T.cases(fx):
        print(f"{scene}/{stage}", flush=True)
        hip = hip_replay(fx, stage)
        orc = T.replay(fx, stage)
        assert len(hip) == len(orc)
        for s, (h, o) in enumerate(zip(hip, orc)):
            for name, arr in o.items():
                a, b = np.asarray(h[name], np.float64).ravel(), np.asarray(arr, np.float64).ravel()
                bad = np.nonzero(a != b)[0]
                assert bad.size == 0, (f"{scene}/{stage} step {s} {name}: {bad.size} of {a.size} differ "
                                       f"from the oracle, first at {bad[0]}: hip {a[bad[0]]!r} oracle {b[bad[0]]!r}")
            ref = T.shader_fields(fx, stage, s)
            assert ref, (stage, s)
            for name, arr in ref.items():
                T.check_field(stage, name, arr, h[name])


@pytest.mark.parametrize("scene", ["3d", "2d"])
def test_hip_100_step_drift_against_reference_shaders(scene, oracle_built):
    """100 whole steps on the GPU against the reference shaders, with the
    drift bars of test_gl_fixtures.py (1e-3 relative L2 while the flow is
    insensitive, 10x the oracle's own one-ulp sensitivity where it is not);
    the HIP fields equal the oracle's bit for bit at every compared step."""
    fx = T.load(f"drift_{scene}")
    size = size_of(fx)
    g = hip_grid(fx, size, True)
    o = T.drift_oracle(fx, size)
    q = T.drift_oracle(fx, size, perturb=True)
    keep = set(int(k) for k in fx["drift__keep"])
    comps = ("vx", "vy", "vz")[: len(size)]
    for s, t in enumerate(fx["drift__times"]):
        for x in (g, o, q):
            x.set_time(float(t))
            x.step()
        if s not in keep:
            continue
        got_v = np.concatenate([g.get(c) for c in comps])
        assert np.array_equal(got_v, np.concatenate([o.get(c) for c in comps])), f"step {s + 1}"
        assert np.array_equal(g.get("smoke"), o.get("smoke")), f"step {s + 1}"
        ref_v = np.concatenate([fx[f"drift__{s}__{c}"].ravel() for c in comps])
        sens = T.rel_l2(np.concatenate([q.get(c) for c in comps]), np.concatenate([o.get(c) for c in comps]))
        ref_sm = fx[f"drift__{s}__smoke"].astype(np.float64).reshape(-1, 4)
        got_sm = g.get("smoke").reshape(-1, 4)
        e = {"velocity": T.rel_l2(got_v, ref_v),
             "pressure": T.rel_l2(g.get("pressure", 0), fx[f"drift__{s}__pressure0"]),
             "density": T.rel_l2(got_sm[:, :3], ref_sm[:, :3])}
        bar = max(1e-3, 10.0 * sens)
        print(scene, "step", s + 1, {k: f"{v:.2e}" for k, v in e.items()},
              f"1-ulp self-sensitivity {sens:.2e}", flush=True)
        for k, v in e.items():
            assert v <= bar, f"step {s + 1}: {k} drift {v:.3e} > {bar:.1e}"
        if sens < 1e-4:
            assert max(e.values()) <= 1e-3
    assert np.array_equal(g.get_emitters(), fx["drift__final__emitters"])
    np.testing.assert_array_equal(g.get_obstacles(), fx["drift__final__obstacles"])
    g.close()
    o.close()
    q.close()


def test_default_arithmetic_differs_from_llvmpipe_only_in_rounding(oracle_built):
    """The default mode (binary16 to nearest even, split lerp) stays bit-exact
    with the default oracle on a fixture scene, and differs from the
    llvmpipe mode: the switch really changes the kernels."""
    fx = T.load("3d")
    size = size_of(fx)
    out = {}
    for mode in ("default", "llvmpipe"):
        g = hip_grid(fx, size, True, mode)
        for t in fx["full__times"]:
            g.set_time(float(t))
            g.step()
        out[mode] = g.get("smoke")
        g.close()
    from oracle import Oracle
    o = Oracle(size, cell_size=float(fx["in_cell_size"]), t_ambient=float(fx["in_t_ambient"]),
               min_level_size=4)
    o.set_physics(fx["in_settings"])
    o.set_emitters(fx["in_emitters"])
    o.set_obstacles(fx["in_obstacles"])
    for f in ("vx", "vy", "vz"):
        o.set(f, fx["in_" + f].astype(np.float32).ravel())
    o.set("smoke", fx["in_smoke"].astype(np.float64))
    o.set("pressure", fx["in_pressure0"], 0)
    o.init_solid_map()
    o.update_solid_map()
    for t in fx["full__times"]:
        o.set_time(float(t))
        o.step()
    assert np.array_equal(out["default"], o.get("smoke"))
    assert not np.array_equal(out["default"], out["llvmpipe"])
    o.close()
