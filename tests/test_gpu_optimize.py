"""GPU check of the batched source optimisation (differentiable rollouts of
uneven 2D and 3D scenes inside a gradient-descent loop): the loss of every
scene falls, and the fitted sources do not depend on how many worker
threads share the GPU."""
import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu


def small_batch():
    return [scenes.make_scene((48, 40), n_emitters=1, emitter_radius=5, n_spheres=1, sphere_radius=5,
                              min_level_size=4, seed=11),
            scenes.make_scene((64, 44), n_emitters=2, emitter_radius=5, min_level_size=4, seed=12),
            scenes.make_scene((24, 20, 16), n_emitters=1, emitter_radius=3, n_boxes=1, box_sides=(3, 5),
                              min_level_size=4, seed=13),
            scenes.make_scene((20, 24, 18), n_emitters=1, emitter_radius=3, min_level_size=4, seed=14)]


def test_source_fit_reduces_every_loss_and_is_thread_invariant(oracle_built):
    from smokesim_amd.optimize import fit_batch
    b = small_batch()
    print("source fit, 4 scenes, 2 workers", flush=True)
    two = fit_batch(b, steps=4, iterations=4, workers=2)
    for r in two["results"]:
        print(r["size"], ["%.3e" % x for x in r["losses"]], flush=True)
        assert r["losses"][0] > 0
        assert r["losses"][-1] < 0.5 * r["losses"][0]
    print("source fit, 1 worker", flush=True)
    one = fit_batch(b, steps=4, iterations=4, workers=1)
    for a, c in zip(one["results"], two["results"]):
        assert a["losses"] == c["losses"]
        assert np.array_equal(a["source"], c["source"])


def test_device_fit_equals_host_fit(oracle_built):
    """The fit with every iteration on the device (forward, loss, backward,
    projected step captured as one graph per scene, the scenes sharing a
    batch's streams) against the host-driven loop on the same seeds: the
    same losses (sums in another order: 1e-12 relative) and sources (1e-6),
    bit-identical between two device fits, every loss falls."""
    from smokesim_amd.optimize import fit_batch, fit_batch_device
    b = small_batch()
    host = fit_batch(b, steps=4, iterations=4, workers=1)
    dev = fit_batch_device(b, steps=4, iterations=4, streams=2)
    again = fit_batch_device(b, steps=4, iterations=4, streams=3)
    for h, d, a in zip(host["results"], dev["results"], again["results"]):
        print(d["size"], ["%.3e" % x for x in d["losses"]], flush=True)
        lh, ld = np.array(h["losses"]), np.array(d["losses"])
        assert lh.shape == ld.shape
        assert np.max(np.abs(lh - ld) / lh) < 1e-12
        np.testing.assert_allclose(d["source"], h["source"], rtol=1e-6, atol=1e-7)
        assert d["losses"] == a["losses"] and np.array_equal(d["source"], a["source"])
        assert ld[-1] < 0.5 * ld[0]


def test_device_fit_recaptures_after_a_settings_change(oracle_built):
    """A rollout's captured fit iteration holds the grid's settings: after
    set_physics the next fit captures again and equals the fit of a rollout
    made after the change."""
    from smokesim_amd import Grid, Rollout
    from smokesim_amd.optimize import source_box
    sc = small_batch()[2]
    size = tuple(sc["size"])
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=4)
    scenes.apply_scene(g, sc)
    lo, sz = source_box(size)
    n = int(np.prod(sz))
    rng = np.random.default_rng(21)
    s0, s1 = (rng.uniform(0, 0.5, n).astype(np.float32) for _ in range(2))
    v0 = [g.get(c).astype(np.float32) for c in ("vx", "vy", "vz")]
    rho0 = np.zeros(int(np.prod(size)), np.float32)

    def fit(r, it=2):
        r.set_initial(*v0, rho0)
        r.set_source(s0)
        r.set_target(r.forward())
        r.set_source(s1)
        r.fit(it)
        return r.fit_result(it)

    r = Rollout(g, 4, 2, lo, sz, refine_cycles=8)
    before, _ = fit(r)
    phys = np.array(sc["physics"], np.float64)
    phys[9] = 2.5 * phys[9] + 0.1  # vorticity confinement
    g.set_physics_array(phys)
    after, src_after = fit(r)
    fresh = Rollout(g, 4, 2, lo, sz, refine_cycles=8)
    ref, src_ref = fit(fresh)
    fresh.close()
    r.close()
    g.close()
    assert not np.array_equal(before, after)
    assert np.array_equal(after, ref) and np.array_equal(src_after, src_ref)


def test_bf16_device_fit_reduces_every_loss(oracle_built):
    """The device-side fit on BF16 grids (rollout state in BF16, FP32
    arithmetic): reproducible, and every loss falls."""
    from smokesim_amd.optimize import fit_batch_device
    b = small_batch()
    one = fit_batch_device(b, steps=4, iterations=4, storage="bf16")
    two = fit_batch_device(b, steps=4, iterations=4, storage="bf16", streams=1)
    for r, q in zip(one["results"], two["results"]):
        print(r["size"], ["%.3e" % x for x in r["losses"]], flush=True)
        assert r["losses"] == q["losses"] and np.array_equal(r["source"], q["source"])
        assert r["losses"][-1] < 0.5 * r["losses"][0]
