"""Scene serialisation and field export (smokesim_amd.state) on the GPU:
a grid saved mid-run and loaded into a new grid continues the run bit for
bit; exported fields have the texel-order shapes."""
import os

import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu

CASES = {
    "3d": dict(size=(24, 20, 16), kw=dict(n_emitters=2, emitter_radius=4, n_boxes=1, n_spheres=1,
                                           sphere_radius=3, box_sides=(4, 6), min_level_size=4)),
    "2d": dict(size=(64, 48), kw=dict(n_emitters=1, emitter_radius=6, n_spheres=1, sphere_radius=6,
                                       min_level_size=4)),
}


@pytest.mark.parametrize("case", ["3d", "2d"])
@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_save_load_continues_bit_for_bit(case, storage, tmp_path):
    from smokesim_amd import Grid, state
    c = CASES[case]
    sc = scenes.make_scene(c["size"], **c["kw"])
    g = Grid(c["size"], cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], storage=storage,
             min_level_size=sc["min_level_size"])
    scenes.apply_scene(g, sc)
    g.run_steps(3, 0.0, 0.1)
    path = os.path.join(tmp_path, "state.npz")
    state.save(g, path)
    g.run_steps(3, 0.3, 0.1)
    h = state.load(path)
    h.run_steps(3, 0.3, 0.1)
    names = ["vx", "vy", "smoke", "pressure", "vorticity", "solid", "residual"]
    if len(c["size"]) == 3:
        names.append("vz")
    for f in names:
        a, b = g.get(f), h.get(f)
        assert np.array_equal(a, b), f"{f}: {np.sum(a != b)} of {a.size} differ"
    assert np.array_equal(g.get_emitters(), h.get_emitters())
    assert np.array_equal(g.get_obstacles(), h.get_obstacles())
    g.close()
    h.close()


def test_export_shapes(tmp_path):
    from smokesim_amd import Grid, state
    c = CASES["3d"]
    sc = scenes.make_scene(c["size"], **c["kw"])
    g = Grid(c["size"], cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
             min_level_size=sc["min_level_size"])
    scenes.apply_scene(g, sc)
    g.run_steps(2, 0.0, 0.1)
    paths = state.export_fields(g, os.path.join(tmp_path, "fields"))
    nx, ny, nz = c["size"]
    want = {"vx": (nz, ny, nx + 1), "vy": (nz, ny + 1, nx), "vz": (nz + 1, ny, nx),
            "smoke": (nz, ny, nx, 4), "vorticity": (nz, ny, nx, 4), "pressure": (nz, ny, nx),
            "solid": (nz, ny, nx, 4)}
    got = {os.path.basename(p)[:-4]: np.load(p).shape for p in paths}
    for f, shp in want.items():
        assert got[f] == shp, (f, got[f], shp)
    v = np.load(os.path.join(tmp_path, "fields", "vy.npy"))
    assert np.array_equal(v.ravel(), g.get("vy").astype(np.float32))
    g.close()
