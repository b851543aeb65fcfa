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
