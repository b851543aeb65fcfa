"""smk_download_field_noalloc against smk_download_field (GPU).

The no-allocation download stages through the library's fixed per-GPU
array; it must return exactly what the allocating download returns, for
every field kind (natural-order FP32/BF16/FP16 storage, colour-split levels,
the FP64 refined pressure, the derived divergence, the BF16 smoke
temperature offset) and both output types -- also on the second of two
calls in a row whose input is smaller than the first's, when the array
still holds the first call's values -- and reject fields over its limit.
"""
import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = [("vx", 0), ("vy", 0), ("vz", 0), ("smoke", 0), ("vorticity", 0), ("pressure", 0),
          ("pressure", 1), ("residual", 0), ("residual", 2), ("solid", 0), ("solid", 1),
          ("divergence", 0)]


def make_grid(size, storage="fp32", refine=0):
    from smokesim_amd import Grid
    sc = scenes.make_scene(size, n_emitters=2, emitter_radius=4, n_boxes=1, box_sides=(4, 6),
                           min_level_size=4)
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], storage=storage,
             min_level_size=4)
    scenes.apply_scene(g, sc)
    # temperatures off ambient: BF16 smoke stores them as offsets
    sm = g.get("smoke").reshape(-1, 4)
    sm[:, 3] = 293.0 + np.random.default_rng(7).uniform(-3, 3, sm.shape[0])
    g.set("smoke", sm.ravel())
    if refine:
        g.set_refinement(refine)
    g.run_steps(2, 0.0, 0.1)
    return g


def check(g, field, level, dtype):
    a = g.get(field, level, dtype=dtype)
    b = g.get_noalloc(field, level, dtype=dtype)
    assert a.dtype == b.dtype and a.shape == b.shape
    assert np.array_equal(a, b, equal_nan=True), (field, level, dtype, np.sum(a != b))
    return a


@pytest.mark.parametrize("storage,refine", [("fp32", 0), ("bf16", 0), ("fp16", 0), ("fp32", 3)])
def test_noalloc_download_equals_download(storage, refine):
    big = make_grid((40, 36, 32), storage, refine)
    small = make_grid((20, 18, 16), storage, refine)
    print(f"noalloc {storage} refine {refine}", flush=True)
    for field, level in FIELDS:
        for dtype in (np.float32, np.float64):
            # two calls in a row, the second on a smaller input
            check(big, field, level, dtype)
            check(small, field, level, dtype)
    if storage == "bf16":
        sm = small.get_noalloc("smoke").reshape(-1, 4)
        assert np.any(sm[:, 3] != 293.0) and np.all(np.abs(sm[:, 3] - 293.0) < 50)
    big.close()
    small.close()


def test_noalloc_download_2d_and_after_steps():
    from smokesim_amd import Grid
    size = (64, 48)
    sc = scenes.make_scene(size, n_emitters=1, emitter_radius=6, n_spheres=1, sphere_radius=6,
                           min_level_size=4)
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=4)
    scenes.apply_scene(g, sc)
    for s in range(3):
        g.set_time(0.1 * s)
        g.step()
        for f in ("vx", "vy", "smoke", "pressure", "vorticity"):
            check(g, f, 0, np.float32)
    g.close()


def test_noalloc_download_rejects_fields_over_the_limit():
    from smokesim_amd import Grid, SmokeError
    from smokesim_amd._abi import NOALLOC_MAX_VALUES
    g = Grid((128, 128, 72), min_level_size=8)  # smoke: 4 * 128*128*72 values
    assert g.field_count("smoke") > NOALLOC_MAX_VALUES >= g.field_count("vx")
    with pytest.raises(SmokeError, match="limit"):
        g.get_noalloc("smoke")
    check(g, "vx", 0, np.float32)  # within the limit
    g.close()
