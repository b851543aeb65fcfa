"""Times SmokeSim's own GLSL compute shaders on Mesa llvmpipe (oracle/_ref/glref)
on the benchmark's scenes and multigrid schedule: the reference CPU path.

TEST INFRASTRUCTURE ONLY (the shaders never leave the container holding the
reference checkout).  Run there:

    make -C oracle/glref
    python oracle/glref/time_reference.py [--ref /root/reference] [--size 256] [--steps 3]

The scene is the benchmark's seeded one (emitters of radius 12 in the bottom
quarter, velocity noise +-0.01, dt = 0.1, the physics coefficients of the
configurations) without the box obstacles and the emitter inflow, which the
shaders do not have; every shader dispatch covers the whole grid whatever
the obstacles, so the step time does not depend on them.  Schedule: 2 + 2
red-black iterations, 4 V-cycles, levels down to 8 cells, 80 coarsest
iterations (glref's GLREF_* knobs).  llvmpipe runs with LP_NUM_THREADS
threads (default: every core).  Prints one line per step and a summary.
"""
import argparse
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))

from make_gl_fixtures import write_case  # noqa: E402
from smokesim_amd import scenes  # noqa: E402

GLREF = os.path.join(REPO, "oracle", "_ref", "glref")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    a = ap.parse_args()
    n = a.size
    sc = scenes.make_scene((n, n, n), n_emitters=4, emitter_radius=12, n_boxes=0)
    vel = sc["velocity"]
    case = {"size": (n, n, n), "settings": sc["physics"], "emitters": sc["emitters"],
            "obstacles": np.zeros((0, 8), np.float32), "vx": vel["vx"], "vy": vel["vy"], "vz": vel["vz"],
            "smoke": np.tile(np.array([0, 0, 0, sc["t_ambient"]], np.float16), n ** 3),
            "pressure0": np.zeros(n ** 3, np.float32), "cell_size": sc["cell_size"],
            "t_ambient": sc["t_ambient"]}
    env = dict(os.environ, GLREF_TIME="1", LP_NUM_THREADS=str(a.threads), GLREF_PRE="2",
               GLREF_POST="2", GLREF_COARSE="80", GLREF_VCYCLES="4", GLREF_MIN_LEVEL="8")
    with tempfile.TemporaryDirectory() as td:
        cp = os.path.join(td, "case.bin")
        # the first step also compiles llvmpipe's shader variants: one extra
        write_case(cp, case, [1] * 7, a.steps + 1)
        out = subprocess.run([GLREF, a.ref, cp, os.path.join(td, "unused.bin")], env=env,
                             capture_output=True, text=True, check=True)
    ms = [float(m) for m in re.findall(r"step \d+ ([0-9.]+) ms", out.stderr)]
    for i, t in enumerate(ms):
        print(f"step {i}: {t:.1f} ms" + (" (includes shader compilation)" if i == 0 else ""))
    timed = ms[1:]
    med = float(np.median(timed))
    print(f"{n}^3, {a.threads} llvmpipe threads, 2+2 sweeps x 4 V-cycles to 8^3: median {med:.1f} ms/step "
          f"over {len(timed)} steps = {n ** 3 / (med * 1e-3) / 1e6:.3f} Mcell-steps/s")


if __name__ == "__main__":
    main()
