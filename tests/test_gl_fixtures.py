"""The oracle against the reference's own GLSL compute shaders.

tests/golden/gl_*.npz hold seeded scenes and the fields that SmokeSim's
shaders produced for them, stage by stage and over three whole steps, when
run unchanged on Mesa llvmpipe by oracle/glref (tests/golden/make_gl_fixtures.py
wrote them).  Each case is replayed here through the float oracle with the
same scene, toggles and time values.

llvmpipe makes two choices that the GL specification leaves open; the
oracle is switched to them for this comparison only:
- binary16 stores (RGBA16F smoke, 3D vorticity, solid maps) round toward
  zero; the oracle's default, and the HIP kernels', is round to nearest even;
- the linear filter's lerp is one fused multiply-add.
With both, the stages below are compared bit for bit.  The vorticity and
pressure stages divide per cell, and GLSL 4.60 (4.7.1) only bounds division
to 2.5 ULP; llvmpipe's quotients differ from IEEE division in the last one or
two bits, so those fields are compared within the tolerances written below.
"""
import os

import numpy as np
import pytest

from oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

SCENES = ("3d", "2d", "3d_wind")

# stage -> fields that must match bit for bit
EXACT = {
    "emitters": ("smoke", "emitters"),
    "advect_velocities": ("vx", "vy", "vz"),
    "buoyancy": ("vy",),
    "advect_smoke": ("smoke",),
}

# tolerance classes for the stages with divisions: |oracle - shader| <=
# REL * max|shader field|; FP16 fields may differ in at most FRAC16 of their
# elements, by at most one binary16 step at the field's largest magnitude.
REL = {"v": 2e-6, "pressure": 1e-5, "residual": 1e-4}
REL_FULL = {"v": 2e-5, "pressure": 2e-5, "residual": 2e-4}
FRAC16 = 0.01


def load(scene):
    path = os.path.join(GOLDEN, f"gl_{scene}.npz")
    return dict(np.load(path, allow_pickle=False))


def cases(fx):
    return sorted({k.split("__")[0] for k in fx if "__" in k})


def replay(fx, stage):
    """Run `stage` of the fixture through the oracle; per-step field dicts."""
    n = [int(v) for v in fx["in_size"]]
    size = tuple(n[:2]) if n[2] == 0 else tuple(n)
    o = Oracle(size, cell_size=float(fx["in_cell_size"]), t_ambient=float(fx["in_t_ambient"]),
               min_level_size=4)
    o.set_f16_rounding("rtz")
    o.set_lerp("fused")
    o.set_physics(fx["in_settings"])
    o.set_emitters(fx["in_emitters"])
    o.set_obstacles(fx["in_obstacles"])
    o.set("vx", fx["in_vx"])
    o.set("vy", fx["in_vy"])
    if len(size) == 3:
        o.set("vz", fx["in_vz"])
    o.set("smoke", fx["in_smoke"].astype(np.float64))
    o.set("pressure", fx["in_pressure0"], 0)
    # the runner's set-up: init_solid_map when enclosed (main.c:90), then one
    # update_solid_map so that obstacles are solid from the first step
    if int(fx[f"{stage}__enclosed"]):
        o.init_solid_map()
    o.update_solid_map()
    tog = fx[f"{stage}__toggles"]
    out = []
    for t in fx[f"{stage}__times"]:
        o.set_time(float(t))
        o.step(tog)
        st = {"vx": o.get("vx"), "vy": o.get("vy"), "smoke": o.get("smoke"),
              "vorticity": o.get("vorticity"),
              "emitters": o.get_emitters(len(fx["in_emitters"])),
              "obstacles": o.get_obstacles(len(fx["in_obstacles"]))}
        if len(size) == 3:
            st["vz"] = o.get("vz")
        for lv in range(o.levels):
            st[f"pressure{lv}"] = o.get("pressure", lv)
            st[f"residual{lv}"] = o.get("residual", lv)
            st[f"solid{lv}"] = o.get("solid", lv)
        out.append(st)
    o.close()
    return out


def shader_fields(fx, stage, step):
    pre = f"{stage}__{step}__"
    return {k[len(pre):]: fx[k] for k in fx if k.startswith(pre)}


def check_field(stage, name, ref, got):
    ref = np.asarray(ref)
    ref64 = ref.astype(np.float64).ravel()
    got = np.asarray(got, np.float64).ravel()
    assert ref64.shape == got.shape, (stage, name, ref64.shape, got.shape)
    exact = (name in EXACT.get(stage, ()) or name.startswith("solid") or name in ("obstacles", "emitters"))
    if exact:
        bad = np.nonzero(ref64 != got)[0]
        assert bad.size == 0, (f"{stage}/{name}: {bad.size} of {ref64.size} differ, first at "
                               f"{bad[0]}: shader {ref64[bad[0]]!r} oracle {got[bad[0]]!r}")
        return
    if ref.dtype == np.float16:
        assert np.all(got.astype(np.float16).astype(np.float64) == got), f"{stage}/{name} not binary16"
        frac = float(np.mean(ref64 != got))
        step = 2.0 ** (np.floor(np.log2(np.max(np.abs(ref64)) + 1e-30)) - 10)
        err = float(np.max(np.abs(ref64 - got)))
        assert err <= step and frac <= FRAC16, (stage, name, err, step, frac)
        return
    cls = "v" if name in ("vx", "vy", "vz", "vorticity") else ("pressure" if name.startswith("pressure") else "residual")
    rel = (REL_FULL if stage == "full" else REL)[cls]
    scale = np.max(np.abs(ref64)) + 1e-30
    err = np.max(np.abs(ref64 - got)) / scale
    assert err <= rel, f"{stage}/{name}: max error {err:.3e} of max|field| > {rel:.1e}"


@pytest.mark.parametrize("scene", SCENES)
def test_oracle_matches_reference_shaders(scene):
    fx = load(scene)
    for stage in cases(fx):
        got = replay(fx, stage)
        for s, st in enumerate(got):
            ref = shader_fields(fx, stage, s)
            assert ref, (stage, s)
            for name, arr in ref.items():
                check_field(stage, name, arr, st[name])


def test_fixture_stages_cover_every_toggle():
    fx = load("3d")
    seen = np.zeros(7, int)
    for stage in cases(fx):
        seen += fx[f"{stage}__toggles"]
    assert np.all(seen > 0)
    # bit-exact stages really are compared bit for bit and are not trivial
    for stage, names in EXACT.items():
        a = shader_fields(fx, stage, 0)
        for n in names:
            if n in a and n != "emitters":
                base = fx["in_" + n] if "in_" + n in fx else None
                if base is not None:
                    assert not np.array_equal(np.asarray(a[n], np.float64), np.asarray(base, np.float64)), (stage, n)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def drift_oracle(fx, size, perturb=False):
    o = Oracle(size, cell_size=float(fx["in_cell_size"]), t_ambient=float(fx["in_t_ambient"]),
               min_level_size=4)
    o.set_f16_rounding("rtz")
    o.set_lerp("fused")
    o.set_physics(fx["in_settings"])
    o.set_emitters(fx["in_emitters"])
    o.set_obstacles(fx["in_obstacles"])
    for f in ("vx", "vy", "vz")[: len(size)]:
        v = fx["in_" + f].astype(np.float32).ravel().copy()
        if perturb and f == "vx":
            k = v.size // 2 + 7  # one interior face, one ulp up
            v[k] = np.nextafter(v[k], np.float32(1))
        o.set(f, v)
    o.set("smoke", fx["in_smoke"].astype(np.float64))
    o.set("pressure", fx["in_pressure0"], 0)
    o.init_solid_map()
    o.update_solid_map()
    return o


@pytest.mark.parametrize("scene", ["3d", "2d"])
def test_100_step_drift_against_reference_shaders(scene):
    """100 whole steps (every stage, emitters, a falling obstacle) against
    the reference shaders.  Velocity, pressure and density stay within 1e-3
    relative L2 while the flow itself is insensitive; where it is not, a
    one-ulp change of a single initial face (the oracle against itself)
    grows just as fast, so the bar there is 10x that self-sensitivity.  The
    emitter and obstacle state agree exactly throughout."""
    fx = load(f"drift_{scene}")
    n = [int(v) for v in fx["in_size"]]
    size = tuple(n[:2]) if n[2] == 0 else tuple(n)
    o = drift_oracle(fx, size)
    q = drift_oracle(fx, size, perturb=True)
    keep = set(int(k) for k in fx["drift__keep"])
    comps = ("vx", "vy", "vz")[: len(size)]
    for s, t in enumerate(fx["drift__times"]):
        for x in (o, q):
            x.set_time(float(t))
            x.step()
        if s not in keep:
            continue
        ref_v = np.concatenate([fx[f"drift__{s}__{c}"].ravel() for c in comps])
        got_v = np.concatenate([o.get(c) for c in comps])
        sens = rel_l2(np.concatenate([q.get(c) for c in comps]), got_v)
        ref_sm = fx[f"drift__{s}__smoke"].astype(np.float64).reshape(-1, 4)
        got_sm = o.get("smoke").reshape(-1, 4)
        e = {"velocity": rel_l2(got_v, ref_v),
             "pressure": rel_l2(o.get("pressure", 0), fx[f"drift__{s}__pressure0"]),
             "density": rel_l2(got_sm[:, :3], ref_sm[:, :3])}
        bar = max(1e-3, 10.0 * sens)
        print(scene, "step", s + 1, {k: f"{v:.2e}" for k, v in e.items()},
              f"1-ulp self-sensitivity {sens:.2e}", flush=True)
        for k, v in e.items():
            assert v <= bar, f"step {s + 1}: {k} drift {v:.3e} > {bar:.1e}"
        if sens < 1e-4:
            assert max(e.values()) <= 1e-3
    assert np.array_equal(o.get_emitters(len(fx["in_emitters"])), fx["drift__final__emitters"])
    np.testing.assert_array_equal(o.get_obstacles(len(fx["in_obstacles"])), fx["drift__final__obstacles"])
    o.close()
    q.close()
