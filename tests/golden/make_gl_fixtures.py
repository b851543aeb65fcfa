"""Writes tests/golden/gl_*.npz: SmokeSim's own GLSL compute shaders run on
Mesa llvmpipe (oracle/_ref/glref, built by oracle/glref/Makefile) on seeded
small scenes, stage by stage and for whole steps, plus 100-step runs
(gl_drift_*.npz) kept at a few snapshot steps.

Run in a container that holds the reference checkout:

    make -C oracle/glref
    python tests/golden/make_gl_fixtures.py [--ref /root/reference]

Each fixture holds the scene (sizes, settings, emitters, obstacles, initial
fields) and the fields the shaders produced after every step.  The fixtures
are data only; tests/test_gl_fixtures.py replays the same scenes through the
oracle and compares.
"""
import argparse
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
GLREF = os.path.join(REPO, "oracle", "_ref", "glref")

# physics_step_settings order (physics.h:10-18)
TOGGLE_NAMES = ("advect_smoke", "advect_velocities", "apply_vorticity", "apply_buoyancy",
                "resolve_pressure", "update_velocities", "handle_emitters")


def toggles(*on):
    return [1 if n in on else 0 for n in TOGGLE_NAMES]


ALL = [1] * 7


def pyramid(n, is_2d):
    """Level sizes: halve while every extent is >= 4 (3d/grid3d.c:44-52)."""
    dims = list(n)
    out = []
    cur = list(n)
    while dims[0] >= 4 and dims[1] >= 4 and (is_2d or dims[2] >= 4):
        dims = [d >> 1 for d in dims]
        out.append(tuple(cur))
        cur = [c // 2 for c in cur]
    return out


def make_scene(size, seed, wind=(0.0, 0.0, 0.0), delay=2):
    """A seeded scene: random face velocities, a smoke blob with heat, two
    emitters (falloff on/off, the second with `delay`), one anchored and one
    free obstacle."""
    rng = np.random.default_rng(seed)
    is_2d = len(size) == 2
    nx, ny = size[0], size[1]
    nz = 1 if is_2d else size[2]
    h = 0.1
    tamb = 293.0
    # physics_settings (physics.h:22-32): wind, dt, density, t_ambient,
    # t_weight_coeff, t_temp_coeff, vorticity_confinement, pad, pad
    settings = np.array([wind[0], wind[1], wind[2], 0.0, 0.05, 1.0, tamb, 0.07, 0.1, 0.6, 0, 0],
                        np.float32)
    vx = rng.uniform(-0.6, 0.6, (nz, ny, nx + 1)).astype(np.float32)
    vy = rng.uniform(-0.6, 0.6, (nz, ny + 1, nx)).astype(np.float32)
    vz = None if is_2d else rng.uniform(-0.6, 0.6, (nz + 1, ny, nx)).astype(np.float32)
    sm = np.zeros((nz, ny, nx, 4), np.float32)
    sm[..., 3] = tamb
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cz = 0 if is_2d else nz * 0.4
    blob = np.exp(-(((xx - nx * 0.45) ** 2 + (yy - ny * 0.35) ** 2 + (zz - cz) ** 2) / (0.08 * nx * nx)))
    for c in range(3):
        sm[..., c] = blob * rng.uniform(0.3, 1.0) + rng.uniform(0, 0.05, blob.shape)
    sm[..., 3] = tamb + 40.0 * blob + rng.uniform(-1, 1, blob.shape)
    sm = sm.astype(np.float16)
    p0 = rng.uniform(-0.05, 0.05, (nz, ny, nx)).astype(np.float32)
    zc = 0 if is_2d else nz // 2
    # emitters.h: pos_radius, density_temp, settings (delay, last_emit, emit, falloff)
    em = np.array([[nx // 3, 3, zc, 3, 2, 1, 0, 150, 0, 0, 0, 1],
                   [2 * nx // 3, 4, zc, 2, 0, 3, 1, 80, delay, 0, 0, 0]], np.int32)
    # obstacles.h: pos_radius, velocities_anchored
    zo = 0.0 if is_2d else nz * 0.5
    ob = np.array([[nx * 0.5, ny * 0.62, zo, 2.6, 0, 0, 0, 1],
                   [nx * 0.3, ny * 0.8, zo, 1.8, 0.5, -0.3, 0.0, 0]], np.float32)
    return dict(size=np.array([nx, ny, 0 if is_2d else nz], np.int32), cell_size=np.float32(h),
                t_ambient=np.float32(tamb), settings=settings, emitters=em, obstacles=ob,
                vx=vx, vy=vy, vz=vz, smoke=sm, pressure0=p0)


def write_case(path, sc, tog, steps, enclosed=1, pre_raster=1, times=None):
    nx, ny, nz = (int(v) for v in sc["size"])
    dt = float(sc["settings"][4])
    if times is None:
        times = np.arange(steps, dtype=np.float32) * dt
    with open(path, "wb") as f:
        f.write(struct.pack("<2i", 0x524C4753, 1))
        f.write(struct.pack("<3i", nx, ny, nz))
        f.write(struct.pack("<5i", steps, len(sc["emitters"]), len(sc["obstacles"]), enclosed,
                            pre_raster))
        f.write(np.asarray(tog, np.int32).tobytes())
        f.write(struct.pack("<2f", float(sc["cell_size"]), float(sc["t_ambient"])))
        f.write(sc["settings"].astype("<f4").tobytes())
        f.write(np.asarray(times, "<f4").tobytes())
        f.write(sc["emitters"].astype("<i4").tobytes())
        f.write(sc["obstacles"].astype("<f4").tobytes())
        f.write(sc["vx"].astype("<f4").tobytes())
        f.write(sc["vy"].astype("<f4").tobytes())
        if nz:
            f.write(sc["vz"].astype("<f4").tobytes())
        f.write(sc["smoke"].astype("<f2").view("<u2").tobytes())
        f.write(sc["pressure0"].astype("<f4").tobytes())
    return np.asarray(times, np.float32)


def read_out(path, sc, steps):
    nx, ny, nz3 = (int(v) for v in sc["size"])
    is_2d = nz3 == 0
    nz = 1 if is_2d else nz3
    data = open(path, "rb").read()
    off = 0

    def take(dtype, shape):
        nonlocal off
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        a = np.frombuffer(data, dtype, count=int(np.prod(shape)), offset=off).reshape(shape)
        off += n
        return a.copy()

    levels = pyramid((nx, ny) if is_2d else (nx, ny, nz), is_2d)
    out = []
    for _ in range(steps):
        st = {}
        st["vx"] = take("<f4", (nz, ny, nx + 1))
        st["vy"] = take("<f4", (nz, ny + 1, nx))
        if not is_2d:
            st["vz"] = take("<f4", (nz + 1, ny, nx))
        st["smoke"] = take("<u2", (nz, ny, nx, 4)).view("<f2")
        st["vorticity"] = take("<f4", (nz, ny, nx)) if is_2d else take("<u2", (nz, ny, nx, 4)).view("<f2")
        for l, ls in enumerate(levels):
            lz = 1 if is_2d else ls[2]
            st[f"pressure{l}"] = take("<f4", (lz, ls[1], ls[0]))
            st[f"residual{l}"] = take("<f4", (lz, ls[1], ls[0]))
            st[f"solid{l}"] = take("<u2", (lz, ls[1], ls[0], 4)).view("<f2")
        st["emitters"] = take("<i4", (len(sc["emitters"]), 12))
        st["obstacles"] = take("<f4", (len(sc["obstacles"]), 8))
        out.append(st)
    if off != len(data):
        raise RuntimeError(f"output size mismatch: read {off} of {len(data)} bytes")
    return out


def run(ref, sc, tog, steps, enclosed=1, pre_raster=1):
    with tempfile.TemporaryDirectory() as td:
        cp, op = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
        times = write_case(cp, sc, tog, steps, enclosed, pre_raster)
        env = dict(os.environ)
        env.setdefault("LP_NUM_THREADS", "4")
        subprocess.run([GLREF, ref, cp, op], check=True, env=env)
        return times, read_out(op, sc, steps)


# (name, toggles, steps, enclosed, fields kept from the shader output)
STAGES = [
    ("emitters", toggles("handle_emitters"), 2, 1, ("smoke", "emitters")),
    ("advect_velocities", toggles("advect_velocities"), 1, 1, ("vx", "vy", "vz")),
    ("buoyancy", toggles("apply_buoyancy"), 1, 1, ("vy",)),
    ("vorticity", toggles("apply_vorticity"), 1, 1, ("vx", "vy", "vz", "vorticity")),
    ("pressure", toggles("resolve_pressure", "update_velocities"), 1, 1,
     ("vx", "vy", "vz", "pressure*", "residual*", "solid*")),
    ("pressure_open", toggles("resolve_pressure", "update_velocities"), 1, 0,
     ("vx", "vy", "vz", "pressure*", "residual*", "solid*")),
    ("advect_smoke", toggles("advect_smoke"), 1, 1, ("smoke",)),
    ("full", ALL, 3, 1, ("*",)),
]


def keep(st, pats):
    out = {}
    for k, v in st.items():
        for p in pats:
            if p == "*" or k == p or (p.endswith("*") and k.startswith(p[:-1])):
                out[k] = v
    return out


# 100-step drift runs: whole steps, snapshots of these steps (0-based)
DRIFT_STEPS = 100
DRIFT_KEEP = (9, 24, 49, 74, 99)
DRIFT_FIELDS = ("vx", "vy", "vz", "smoke", "pressure0")


def drift_fixture(ref, sc):
    times, steps_out = run(ref, sc, ALL, DRIFT_STEPS, 1)
    rec = {f"in_{k}": v for k, v in sc.items() if v is not None}
    rec["drift__times"] = times
    rec["drift__keep"] = np.asarray(DRIFT_KEEP, np.int32)
    for s in DRIFT_KEEP:
        for k, v in steps_out[s].items():
            if k in DRIFT_FIELDS:
                rec[f"drift__{s}__{k}"] = v
    rec["drift__final__emitters"] = steps_out[-1]["emitters"]
    rec["drift__final__obstacles"] = steps_out[-1]["obstacles"]
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--drift-only", action="store_true", help="write only gl_drift_*.npz")
    args = ap.parse_args()
    if not os.path.exists(GLREF):
        sys.exit("build oracle/_ref/glref first: make -C oracle/glref")
    scenes = {"3d": make_scene((16, 20, 12), seed=11), "2d": make_scene((40, 28), seed=12),
              "3d_wind": make_scene((16, 20, 12), seed=13, wind=(0.4, 0.0, 0.0))}
    for sname, sc in scenes.items():
        if args.drift_only:
            break
        rec = {f"in_{k}": v for k, v in sc.items() if v is not None}
        stages = STAGES if sname != "3d_wind" else [s for s in STAGES if s[0] in ("advect_velocities", "full")]
        for name, tog, steps, enclosed, pats in stages:
            times, steps_out = run(args.ref, sc, tog, steps, enclosed)
            rec[f"{name}__toggles"] = np.asarray(tog, np.int32)
            rec[f"{name}__enclosed"] = np.int32(enclosed)
            rec[f"{name}__times"] = times
            for s, st in enumerate(steps_out):
                for k, v in keep(st, pats).items():
                    rec[f"{name}__{s}__{k}"] = v
            print(f"{sname}: {name} ({steps} step(s)) done", flush=True)
        path = os.path.join(args.out, f"gl_{sname}.npz")
        np.savez_compressed(path, **rec)
        print(path, os.path.getsize(path), "bytes")
    # 2D: emitters without delay.  The 2D update_emitters_status.comp runs
    # 8x8 invocations per group but indexes emitters by gl_GlobalInvocationID.x
    # alone, so eight invocations update each emitter concurrently; with a
    # delay the outcome depends on their order.  With delay 0 every
    # invocation writes the same state.
    drift_scenes = {"3d": scenes["3d"], "2d": make_scene((40, 28), seed=12, delay=0)}
    for sname in ("3d", "2d"):
        rec = drift_fixture(args.ref, drift_scenes[sname])
        path = os.path.join(args.out, f"gl_drift_{sname}.npz")
        np.savez_compressed(path, **rec)
        print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
