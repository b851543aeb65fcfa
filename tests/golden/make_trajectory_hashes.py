"""Write per-step digests of the CPU oracle's trajectory for a full-length
configuration run, so that a GPU test can compare its own run bit for bit
without the oracle's minutes of CPU time.

  python tests/golden/make_trajectory_hashes.py c2   # 3D 128^3, 200 steps

Every `every` steps (and after step 1) the SHA-256 of each field's float64
download (vx, vy, vz, smoke, level-0 pressure and residual) is recorded, with
the field's L2 norm for diagnostics.  Time is step * 0.1, as in the GPU
tests.  Output: tests/golden/trajectory_<name>.json.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

FIELDS = ("vx", "vy", "vz", "smoke", "pressure", "residual")


def digest(a):
    a = np.ascontiguousarray(np.asarray(a, np.float64))
    return hashlib.sha256(a.tobytes()).hexdigest()


def configs():
    from smokesim_amd import scenes
    return {"c2": (scenes.config_3d_128, 200, 10)}


def main(name):
    import oracle
    from smokesim_amd import scenes
    make, steps, every = configs()[name]
    sc = make()
    o = oracle.Oracle(sc["size"], cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
                      min_level_size=sc["min_level_size"])
    scenes.apply_scene(o, sc)
    out = {"config": name, "size": list(sc["size"]), "steps": steps, "every": every, "records": []}
    t0 = time.time()
    for s in range(steps):
        o.set_time(s * 0.1)
        o.step()
        if s == 0 or (s + 1) % every == 0:
            rec = {"step": s + 1}
            for f in FIELDS:
                a = o.get(f)
                rec[f] = digest(a)
                rec[f + "_l2"] = float(np.linalg.norm(a))
            out["records"].append(rec)
            print(name, "step", s + 1, f"{time.time() - t0:.0f} s", flush=True)
    o.close()
    with open(os.path.join(HERE, f"trajectory_{name}.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "c2")
