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


def test_residual_norm_is_the_fp64_residual_and_deterministic():
    """smk_residual_norm: |b - A p| on level 0 with compute_residual.comp's
    operator (out-of-domain neighbours solid) and b = rho*div/(dt*h), in
    FP64, against a scipy matrix built from the downloaded fields; and the
    same value on every call (per-block partial sums added in a fixed
    order)."""
    pytest.importorskip("scipy.sparse")
    from smokesim_amd import Grid
    import test_oracle as T
    size = (24, 20, 16)
    sc = scenes.make_scene(size, n_emitters=2, emitter_radius=4, n_boxes=1, box_sides=(4, 6),
                           min_level_size=4)
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=sc["min_level_size"])
    scenes.apply_scene(g, sc)
    for s in range(3):
        g.set_time(0.1 * s)
        g.step()
    norms = [g.residual_norm() for _ in range(3)]
    assert norms[0] == norms[1] == norms[2]
    # the face divergence in float, in the kernels' order
    # ((((vr - vl) + vtop) - vbot) + vfront) - vback, solid neighbours'
    # faces taking the solid map's velocity, out-of-domain neighbours solid
    sh = T.cshape(size)
    sol = g.get("solid").reshape(*sh, 4)
    a = sol[..., 3]
    dv = None
    for c, k in enumerate(("vx", "vy", "vz")):
        ax = 2 - c
        v = g.get(k).reshape(T.vshape(size, c)).astype(np.float32)
        lo, hi = np.take(v, range(sh[ax]), axis=ax), np.take(v, range(1, sh[ax] + 1), axis=ax)
        one = np.ones_like(np.take(a, [0], axis=ax))
        zero = np.zeros_like(one)
        a_hi = np.concatenate([np.take(a, range(1, sh[ax]), axis=ax), one], ax)
        a_lo = np.concatenate([one, np.take(a, range(sh[ax] - 1), axis=ax)], ax)
        s_hi = np.concatenate([np.take(sol[..., c], range(1, sh[ax]), axis=ax), zero], ax).astype(np.float32)
        s_lo = np.concatenate([zero, np.take(sol[..., c], range(sh[ax] - 1), axis=ax)], ax).astype(np.float32)
        hi = np.where(a_hi >= 1, s_hi, hi).astype(np.float32)
        lo = np.where(a_lo >= 1, s_lo, lo).astype(np.float32)
        dv = (hi - lo) if dv is None else (dv + hi) - lo
    dv = dv.astype(np.float64)
    p = g.get("pressure").reshape(sh)
    h = float(np.float32(sc["cell_size"]))  # the grid keeps the cell size in float
    rho, dt = float(sc["physics"][5]), float(sc["physics"][4])
    A = T.residual_operator(a, len(size))
    r = rho * dv.ravel() / (dt * h) - (A @ p.ravel()) / (h * h)
    r = np.where(a.ravel() >= 1, 0.0, r)
    want = float(np.linalg.norm(r))
    print("residual norm", norms[0], "numpy", want, flush=True)
    assert norms[0] == pytest.approx(want, rel=1e-9)
    g.close()
