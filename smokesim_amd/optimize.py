"""Gradient-descent source optimisation over batches of uneven scenes.

Each scene fits a per-cell density source near its floor so that the density
after `steps` differentiable steps matches a target density (made by a
forward run with a different seeded source).  One iteration is a forward
rollout, the hand-written adjoint (smk_rollout_*) and a projected gradient
step on the source.  The reference has no gradients; the scenes' grids,
emitters and obstacles follow its scene data (emitters.c, obstacles.c), with
the emitters replaced by the source being fitted.

Two drivers:
- fit_batch: scenes run in worker threads, each on its own grid and HIP
  stream (the library's handles are independent; its calls release the
  GIL); every iteration goes through the host (forward, loss, backward,
  step in numpy).
- fit_batch_device: the grids of one batch share its streams; each scene's
  iteration (forward, loss, backward, projected step) runs on the device and
  is captured once as a HIP graph (smk_rollout_fit), so the whole fit is
  queued at once and the host only waits for the result.
Rollouts compute in FP32 and store their state in the grid's format.
"""
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from ._abi import Batch, Grid, Rollout
from . import scenes as _scenes


def source_box(size):
    """The fitted source region: a centred box near the floor, a quarter of
    the domain wide on each horizontal axis and a few cells high."""
    lo, sz = [], []
    for a, n in enumerate(size):
        if a == 1:
            s = max(2, n // 16)
            lo.append(2)
        else:
            s = max(2, n // 4)
            lo.append(n // 2 - s // 2)
        sz.append(s)
    return tuple(lo), tuple(sz)


def fit_scene(scene, steps=16, iterations=10, seed=0, refine_cycles=12, device=0):
    """Fit one scene's source.  Returns the loss before every iteration and
    after the last, the cells, and the wall time of the iterations."""
    size = tuple(scene["size"])
    g = Grid(size, cell_size=scene["cell_size"], t_ambient=scene["t_ambient"],
             min_level_size=scene["min_level_size"], device=device)
    try:
        _scenes.apply_scene(g, scene)
        lo, sz = source_box(size)
        r = Rollout(g, steps, steps, lo, sz, refine_cycles=refine_cycles)
        try:
            fields = ("vx", "vy", "vz")[: len(size)]
            v0 = [g.get(f).astype(np.float32) for f in fields]
            rho0 = np.zeros(int(np.prod(size)), np.float32)
            rng = np.random.default_rng(seed)
            n = int(np.prod(sz))
            r.set_initial(*v0, rho0)
            r.set_source(rng.uniform(0.0, 0.5, n).astype(np.float32))
            target = r.forward().astype(np.float64)
            src = rng.uniform(0.0, 0.5, n).astype(np.float32)
            losses = []
            t0 = time.perf_counter()
            for _ in range(iterations):
                r.set_initial(*v0, rho0)
                r.set_source(src)
                rho = r.forward().astype(np.float64)
                diff = rho - target
                loss = float(np.dot(diff, diff))
                losses.append(loss)
                sb = r.backward(2.0 * diff)[0].astype(np.float64)
                gg = float(np.dot(sb, sb))
                if gg == 0.0:
                    break
                # Polyak step for a loss whose minimum is 0, kept >= 0
                src = np.maximum(src - (0.5 * loss / gg) * sb, 0.0).astype(np.float32)
            r.set_initial(*v0, rho0)
            r.set_source(src)
            rho = r.forward().astype(np.float64)
            losses.append(float(np.dot(rho - target, rho - target)))
            secs = time.perf_counter() - t0
        finally:
            r.close()
    finally:
        g.close()
    return {"size": size, "cells": int(np.prod(size)), "losses": losses, "seconds": secs,
            "source": src}


def fit_batch(scene_list, steps=16, iterations=10, workers=4, seed=0, refine_cycles=12):
    """Fit every scene of the batch (workers threads).  Returns the per-scene
    results and the aggregate rate: cells x steps x iterations over the
    wall time, the forward steps of the fitted rollouts per second."""
    results = [None] * len(scene_list)
    lock = threading.Lock()
    t0 = time.perf_counter()

    def run(i):
        res = fit_scene(scene_list[i], steps, iterations, seed + i, refine_cycles)
        with lock:
            results[i] = res

    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        for f in [ex.submit(run, i) for i in range(len(scene_list))]:
            f.result()
    wall = time.perf_counter() - t0
    cell_steps = sum(r["cells"] for r in results) * steps * iterations
    return {"results": results, "wall_seconds": wall,
            "Gcell-steps/s": cell_steps / wall / 1e9}


def fit_batch_device(scene_list, steps=16, iterations=10, streams=4, seed=0, refine_cycles=12,
                     adjoint_cycles=0, storage="fp32"):
    """The same fit as fit_batch (same seeds, targets and initial sources),
    every iteration on the device: each scene's iteration is one captured
    graph replayed `iterations` times on the batch's streams; the rollouts
    store their state in `storage` (fp32, bf16 or fp16).  Returns the
    per-scene results (losses before every step and after the last, the
    fitted source) and the aggregate rate over the queued fit (set-up and
    graph capture timed separately)."""
    descs = [dict(size=s["size"], cell_size=s["cell_size"], t_ambient=s["t_ambient"], storage=storage,
                  min_level_size=s["min_level_size"]) for s in scene_list]
    t_setup = time.perf_counter()
    b = Batch(descs, streams=streams)
    rolls, fields = [], []
    try:
        for i, (g, sc) in enumerate(zip(b.grids, scene_list)):
            size = tuple(sc["size"])
            _scenes.apply_scene(g, sc)
            lo, sz = source_box(size)
            r = Rollout(g, steps, steps, lo, sz, refine_cycles=refine_cycles, adjoint_cycles=adjoint_cycles)
            rolls.append(r)
            f = ("vx", "vy", "vz")[: len(size)]
            v0 = [g.get(c).astype(np.float32) for c in f]
            rho0 = np.zeros(int(np.prod(size)), np.float32)
            rng = np.random.default_rng(seed + i)
            n = int(np.prod(sz))
            r.set_initial(*v0, rho0)
            r.set_source(rng.uniform(0.0, 0.5, n).astype(np.float32))
            target = r.forward()
            r.set_target(target)
            r.set_source(rng.uniform(0.0, 0.5, n).astype(np.float32))
            fields.append(target.astype(np.float64))
        t_capture = time.perf_counter()
        for r in rolls:
            r.fit(0)  # capture only
        b.sync()
        t0 = time.perf_counter()
        # one iteration per scene in turn, so that no stream's queue fills
        # while the others wait
        for _ in range(iterations):
            for r in rolls:
                r.fit(1)
        out = [r.fit_result(iterations) for r in rolls]
        wall = time.perf_counter() - t0
        results = []
        for r, (losses, src), target, sc in zip(rolls, out, fields, scene_list):
            d = r.forward().astype(np.float64) - target
            results.append({"size": tuple(sc["size"]), "cells": int(np.prod(sc["size"])),
                            "losses": list(losses) + [float(np.dot(d, d))], "source": src})
    finally:
        for r in rolls:
            r.close()
        b.close()
    cell_steps = sum(r["cells"] for r in results) * steps * iterations
    return {"results": results, "wall_seconds": wall, "setup_seconds": t_capture - t_setup,
            "capture_seconds": t0 - t_capture, "Gcell-steps/s": cell_steps / wall / 1e9}
