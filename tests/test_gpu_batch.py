"""Batches of uneven scenes and launch tiles (GPU).

- Stepping scenes through one batch graph gives bit-identical fields to
  stepping each scene alone.
- The launch tile width changes only the launch shape, never a result.
- The autotuner picks one of its candidate widths.
"""
import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu

SCENES = [
    ((64, 48), dict(n_emitters=1, emitter_radius=6, n_spheres=1, sphere_radius=6, min_level_size=4)),
    ((24, 20, 16), dict(n_emitters=2, emitter_radius=4, n_boxes=1, box_sides=(4, 6), min_level_size=4)),
    ((75, 53), dict(n_emitters=1, emitter_radius=5, n_spheres=1, sphere_radius=7, min_level_size=4)),
    ((27, 21, 19), dict(n_emitters=2, emitter_radius=4, n_boxes=2, box_sides=(3, 6), min_level_size=4)),
]


def fields(dim):
    return ["vx", "vy", "smoke", "pressure", "solid"] + (["vz"] if dim == 3 else [])


@pytest.mark.parametrize("mode", ["per_scene", "single_graph", "fused"])
def test_batch_matches_individual_grids(mode, oracle_built):
    from smokesim_amd import Batch, Grid
    descs, scs = [], []
    for i, (size, kw) in enumerate(SCENES):
        sc = scenes.make_scene(size, seed=10 + i, **kw)
        scs.append(sc)
        descs.append(dict(size=size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
                          storage="bf16", min_level_size=sc["min_level_size"]))
    b = Batch(descs, streams=3, mode=mode)
    singles = []
    for (size, _), sc, g, d in zip(SCENES, scs, b.grids, descs):
        scenes.apply_scene(g, sc)
        s = Grid(**d)
        scenes.apply_scene(s, sc)
        singles.append(s)
    print("batch of", len(descs), "scenes", flush=True)
    b.step(4, 0.0, 0.1)
    b.sync()
    for s in singles:
        for n in range(4):
            s.set_time(n * 0.1)
            s.step()
    for (size, _), g, s in zip(SCENES, b.grids, singles):
        for f in fields(len(size)):
            assert np.array_equal(g.get(f), s.get(f)), (size, f)
    b.close()


def test_fused_autotune_picks_widths_and_keeps_results(oracle_built):
    """The fused mode's autotuner sets one candidate width per launch kind
    that the batch has, and the batch then still equals single grids bit for
    bit (two steps before the autotune, two after)."""
    from smokesim_amd import Batch, Grid
    descs, scs = [], []
    for i, (size, kw) in enumerate(SCENES):
        sc = scenes.make_scene(size, seed=20 + i, **kw)
        scs.append(sc)
        descs.append(dict(size=size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
                          storage="bf16", min_level_size=sc["min_level_size"]))
    b = Batch(descs, streams=1, mode="fused")
    singles = []
    for sc, g, d in zip(scs, b.grids, descs):
        scenes.apply_scene(g, sc)
        s = Grid(**d)
        scenes.apply_scene(s, sc)
        singles.append(s)
    assert all(w == 0 for w in b.fused_widths().values())
    b.step(2, 0.0, 0.1)
    b.sync()
    b.autotune(reps=1)
    widths = b.fused_widths()
    print("fused widths", widths, flush=True)
    for k in ("advect_velocity", "restrict", "prolong", "advect_smoke", "emit"):
        assert widths[k] in (8, 16, 32, 64, 128), (k, widths[k])
    # these take 256-thread blocks over linear thread ranges
    for k in ("sweep", "vorticity", "apply_vorticity", "divergence", "update_velocity"):
        assert widths[k] == 256, (k, widths[k])
    b.step(2, 0.2, 0.1)
    b.sync()
    for s in singles:
        for n in range(4):
            s.set_time(n * 0.1)
            s.step()
    for (size, _), g, s in zip(SCENES, b.grids, singles):
        for f in fields(len(size)):
            assert np.array_equal(g.get(f), s.get(f)), (size, f)
    b.close()


def test_tile_width_does_not_change_results(oracle_built):
    from smokesim_amd import Grid
    size, kw = SCENES[1]
    sc = scenes.make_scene(size, **kw)
    out = []
    for tile in (0, 16, 128):
        g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
                 min_level_size=sc["min_level_size"])
        scenes.apply_scene(g, sc)
        g.set_tile(tile)
        for n in range(3):
            g.set_time(n * 0.1)
            g.step()
        out.append({f: g.get(f) for f in fields(3)})
    for f in fields(3):
        assert np.array_equal(out[0][f], out[1][f]) and np.array_equal(out[0][f], out[2][f]), f


def test_autotune_picks_a_candidate(oracle_built):
    from smokesim_amd import Grid
    g = Grid((48, 40, 32), min_level_size=4)
    bx, ms = g.autotune(reps=2)
    print("autotuned tile", bx, f"{ms:.3f} ms/step", flush=True)
    assert bx in (16, 32, 64, 128, 256) and ms > 0
    assert g.tile() == bx


def test_single_graph_batch_orders_against_scene_streams(oracle_built):
    """A single-graph batch step runs on the batch's first stream; reading a
    scene that lives on another stream right after it (no smk_batch_sync)
    sees the finished step, and work queued on that stream before the step
    (a time update) is ordered before it."""
    from smokesim_amd import Batch, Grid
    descs, scs = [], []
    for i, (size, kw) in enumerate(SCENES):
        sc = scenes.make_scene(size, seed=20 + i, **kw)
        scs.append(sc)
        descs.append(dict(size=size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
                          storage="fp32", min_level_size=sc["min_level_size"]))
    b = Batch(descs, streams=3, single_graph=True)
    singles = []
    for sc, g, d in zip(scs, b.grids, descs):
        scenes.apply_scene(g, sc)
        s = Grid(**d)
        scenes.apply_scene(s, sc)
        singles.append(s)
    for n in range(3):
        b.step(1, 0.1 * n, 0.1)
        for s in singles:
            s.set_time(0.1 * n)
            s.step()
        # scenes 1 and 2 live on streams 1 and 2: read without a batch sync
        for i in (1, 2):
            size = SCENES[i][0]
            for f in fields(len(size)):
                assert np.array_equal(b.grids[i].get(f), singles[i].get(f)), (n, i, f)
    b.close()


def test_fused_batch_refuses_what_it_does_not_fuse(oracle_built):
    """A fused batch needs static obstacles and no refinement; otherwise the
    step says which scene stops it instead of running."""
    from smokesim_amd import Batch, SmokeError
    descs = [dict(size=s, min_level_size=4) for s, _ in SCENES[:2]]
    b = Batch(descs, streams=2, mode="fused")
    b.grids[1].set_refinement(2, True)
    with pytest.raises(SmokeError, match="refinement"):
        b.step(1, 0.0, 0.1)
    b.grids[1].set_refinement(0, False)
    b.grids[0].set_obstacles([scenes.sphere((20, 20), 4, (1, 0), anchored=0)])
    with pytest.raises(SmokeError, match="moves"):
        b.step(1, 0.0, 0.1)
    b.close()
