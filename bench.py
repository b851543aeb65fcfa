#!/usr/bin/env python3
"""Benchmark: the smoke simulation step at 256^3 on one MI355X.

One step = one run_physics_step (physics.c:72-101) over the whole grid:
emission, MAC advection with buoyancy, vorticity confinement, a 4-V-cycle
red-black multigrid pressure solve (2+2 sweeps, coarsest 8^3), gradient
subtraction, smoke advection and the obstacle update.  Inputs are synthetic
(seeded emitters, boxes and velocity noise) and resident in HBM before the
timed region.  Prints ONE JSON line.

Metric: Gcell-steps/s = cells * steps / seconds, timed over the production
path (the step captured once and replayed as a HIP graph, launch tile from
the project's autotuner).  The roofline object is for the dominant kernel
(the level-0 red-black sweep), timed with HIP events on the grid's stream in
a second timed region of the same length with eager launches (the events
would cost ~5 % of the headline).  The cpu_baseline object times the CPU
oracle (a scalar C port of the reference shaders) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell of one red-black half sweep on level 0: read
# the other colour's pressure (P/2), the own colour's right-hand side (R/2),
# the solid mask of every cell (1 B) and write the own colour's pressure
# (P/2), with P = R = 4 B.
SWEEP_BYTES_PER_CELL = 7.0
# Algorithmic bytes per cell of one whole 3D step (element-granular
# compulsory traffic, every array element read once and written once per
# stage): advect velocity 25, buoyancy 17, vorticity 21, confinement 32,
# divergence 17, 4 V-cycles 377.1, coarse masks 1.3, gradient 29, smoke 29.
STEP_BYTES_PER_CELL = {"fp32": 548.4, "bf16": 490.4, "fp16": 490.4}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--storage", default="fp32", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches for the headline instead of HIP-graph replay")
    ap.add_argument("--graph", action="store_true", help="(default) HIP-graph replay for the headline")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch tile instead of the autotuned one")
    ap.add_argument("--extra", action="store_true", help="also time the 128^3 and 2D 256^2 configs")
    ap.add_argument("--calibrate", action="store_true",
                    help="time a plain 256 MiB device copy first (HBM counter calibration)")
    ap.add_argument("--no-fit", action="store_true",
                    help="skip the batched source-fitting mode (about 45 s)")
    ap.add_argument("--only", choices=["rollout128", "rollout128ckpt8", "batch", "batch_fused", "fit"],
                    help="run just one mode measurement and print its JSON (for profiling)")
    ap.add_argument("--no-modes", action="store_true",
                    help="skip the BF16 and FP64-refined mode lines")
    ap.add_argument("--refine", type=int, default=11,
                    help="correction V-cycles per solve of the refined modes")
    ap.add_argument("--refine-main", action="store_true",
                    help="run the main measurement with FP64 refinement (--refine cycles)")
    return ap.parse_args()


def make_grid(scene, storage, graph=True):
    from smokesim_amd import Grid, scenes
    g = Grid(scene["size"], cell_size=scene["cell_size"], t_ambient=scene["t_ambient"],
             storage=storage, min_level_size=scene["min_level_size"])
    scenes.apply_scene(g, scene)
    g.set_graph_mode(graph)
    return g


def time_steps(g, steps, warmup, dt):
    """Warm up, then time exactly `steps` steps bracketed by device-wide
    synchronisation (hipDeviceSynchronize of the library's HIP runtime; one
    GPU, so no barrier across ranks is needed)."""
    from smokesim_amd import device_sync
    g.run_steps(warmup, 0.0, dt)
    g.sync()
    device_sync(0)
    t0 = time.perf_counter()
    g.run_steps(steps, warmup * dt, dt)
    g.sync()
    device_sync(0)
    return time.perf_counter() - t0


def measure_mode(sc, storage, refine, steps, warmup, dt, autotune=True):
    """Steps/s of one storage / solver mode on the same scene (graph replay,
    the mode's own autotuned tile like the headline); with refinement also
    the relative FP64 residual of the last solve."""
    g = make_grid(sc, storage, graph=True)
    if refine:
        # correction cycles reach the contraction of 80 coarsest iterations
        # with 20
        mg = sc["mg"]
        g.set_mg(mg[0], mg[1], 20, mg[3])
        g.set_refinement(refine, True)
    tile = g.autotune(3)[0] if autotune else g.tile()
    secs = time_steps(g, steps, warmup, dt)
    cells = int(np.prod(sc["size"]))
    out = {"Gcell-steps/s": round(cells * steps / secs / 1e9, 4),
           "ms_per_step": round(secs / steps * 1e3, 4), "storage": storage, "tile_bx": tile}
    if refine:
        rn, bn = g.refine_norms()
        out["refine_cycles"] = refine
        out["rel_residual_last_solve"] = float(rn[-1] / bn)
    g.close()
    return out


def measure_rollout(n, steps, ce, refine, reps=2, adjoint=8):
    """Differentiable rollout on n^3: forward alone, then forward + adjoint
    (checkpoints every ce steps, recomputation of all but the last segment).
    Velocity noise +-0.05, a 32^3 (n/4 per axis) source near the floor.
    Forward projections: `refine` correction cycles from zero; adjoint
    projections: `adjoint` cycles warm-started from the previous adjoint
    step's solution (the first from zero with `refine`)."""
    from smokesim_amd import Grid, Rollout, scenes
    sc = scenes.make_scene((n, n, n), n_emitters=0, n_boxes=1, box_sides=(n // 8, n // 4),
                           noise=0.05, min_level_size=8)
    g = Grid((n, n, n), cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=8)
    scenes.apply_scene(g, sc)
    g.set_mg(2, 2, 20, 1)
    s = n // 4
    r = Rollout(g, steps, ce, (n // 2 - s // 2, 2, n // 2 - s // 2), (s, s, s), refine_cycles=refine,
                adjoint_cycles=adjoint)
    rng = np.random.default_rng(3)
    v0 = [g.get(f).astype(np.float32) for f in ("vx", "vy", "vz")]
    rho0 = np.zeros(n ** 3, np.float32)
    r.set_initial(*v0, rho0)
    r.set_source(rng.uniform(0, 0.5, s ** 3).astype(np.float32))
    r.forward()
    rb = rng.uniform(-1, 1, n ** 3).astype(np.float32)
    r.backward(rb)
    tf = tb = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        r.forward()
        t1 = time.perf_counter()
        r.backward(rb)
        t2 = time.perf_counter()
        tf, tb = min(tf, t1 - t0), min(tb, t2 - t1)
    res_f, res_a = r.residuals()
    r.close()
    g.close()
    cells = n ** 3
    return {"Gcell-steps/s": round(cells * steps / (tf + tb) / 1e9, 4),
            "forward_only_Gcell-steps/s": round(cells * steps / tf / 1e9, 4),
            "ms_forward": round(tf * 1e3, 2), "ms_backward": round(tb * 1e3, 2),
            "fwd_plus_adjoint_over_forward": round((tf + tb) / tf, 3),
            "steps": steps, "checkpoint_every": ce, "refine_cycles": refine, "adjoint_cycles": adjoint,
            "max_rel_residual_forward_solves": float(f"{res_f:.3g}"),
            "max_rel_residual_adjoint_solves": float(f"{res_a:.3g}"),
            "recomputed_steps": steps - min(ce, steps),
            "note": "host<->device copies of the initial state, source and gradients included"}


def make_batch_scenes(n=64, seed=6):
    from smokesim_amd import scenes
    return scenes.batch_uneven(n, seed)


def measure_batch(n_scenes, steps, storage="bf16", streams=4, modes=("per_scene", "one_stream", "fused")):
    """Aggregate Gcell-steps/s of a batch of uneven scenes (each scene
    replays its captured step; scenes spread over `streams` streams), the
    same batch on one stream, the fused mode (one multi-scene launch per
    stage, level and half sweep, the batch step captured as one graph), and
    the tiles the autotuner picked."""
    from smokesim_amd import Batch, device_sync, scenes
    scs = make_batch_scenes(n_scenes)
    descs = [dict(size=s["size"], cell_size=s["cell_size"], t_ambient=s["t_ambient"], storage=storage,
                  min_level_size=s["min_level_size"]) for s in scs]
    cells = sum(int(np.prod(s["size"])) for s in scs)
    out = {"scenes": n_scenes, "cells": cells, "steps": steps, "storage": storage, "streams": streams}
    runs = {"per_scene": (streams, "", "per_scene"), "one_stream": (1, "one_stream_", "per_scene"),
            "fused": (1, "fused_", "fused")}
    tiles = []
    for k, label, mode in (runs[m] for m in modes):
        b = Batch(descs, streams=k, mode=mode)
        for g, s in zip(b.grids, scs):
            scenes.apply_scene(g, s)
        if mode == "fused":
            # the fused launches' block widths, timed per launch kind
            b.autotune(reps=2)
            out["fused_block_widths"] = b.fused_widths()
        else:
            tiles = b.autotune(reps=2)
        b.step(2, 0.0, 0.1)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps, 0.2, 0.1)
        b.sync()
        device_sync(0)
        tb = time.perf_counter() - t0
        b.close()
        out[label + "Gcell-steps/s"] = round(cells * steps / tb / 1e9, 4)
        out[label + "ms_per_batch_step"] = round(tb / steps * 1e3, 3)
    if tiles:
        hist = {}
        for t in tiles:
            hist[str(t)] = hist.get(str(t), 0) + 1
        out["autotuned_tiles"] = hist
    return out


def measure_source_fit(n_scenes=64, steps=16, iterations=10, streams=4, storage="bf16"):
    """The batch of uneven scenes (half 2D, half 3D) each fitting its floor
    source by gradient descent through the differentiable rollout, every
    iteration on the device (one captured graph per scene, replayed on the
    batch's streams): forward steps of the fitted rollouts per second over
    the whole batch (adjoints included in the time), and the loss reduction."""
    from smokesim_amd.optimize import fit_batch_device
    scs = make_batch_scenes(n_scenes)
    out = fit_batch_device(scs, steps=steps, iterations=iterations, streams=streams, storage=storage)
    ratios = sorted(r["losses"][-1] / r["losses"][0] for r in out["results"])
    return {"scenes": n_scenes, "cells": sum(r["cells"] for r in out["results"]), "steps": steps,
            "iterations": iterations, "streams": streams, "storage": storage,
            "Gcell-steps/s": round(out["Gcell-steps/s"], 4), "wall_seconds": round(out["wall_seconds"], 2),
            "setup_seconds": round(out["setup_seconds"], 2), "capture_seconds": round(out["capture_seconds"], 2),
            "final_over_initial_loss_median": float(ratios[len(ratios) // 2]),
            "final_over_initial_loss_max": float(ratios[-1]),
            "note": "each iteration on the device: forward rollout + loss + adjoint + projected Polyak step, "
                    "captured as one graph per scene"}


def copy_ceiling_gbs(nbytes=1 << 30):
    """Read + write GB/s of the library's plain 16-byte copy kernel over
    nbytes (best of 5)."""
    from smokesim_amd import calibrate_copy
    ms = calibrate_copy(nbytes, reps=5)
    return 2 * nbytes / (ms * 1e-3) / 1e9


def traffic_from_profiles(size):
    """HBM bytes per launch of the level-0 sweep measured with rocprofv3
    counters (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if a summary
    for this size is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", f"k_sweep4_l0_{size}_hbm.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def reference_llvmpipe_256():
    """The reference shaders' own step on Mesa llvmpipe at 256^3 (the
    benchmark's schedule), as measured in the build container, which holds
    the reference checkout: profiles/reference_llvmpipe_step.md."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "reference_llvmpipe_step.md")
    try:
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if cells and cells[0] == "256^3":
                return {"value": round(float(cells[4]) / 1e3, 6), "unit": "Gcell-steps/s",
                        "cores": int(cells[1]), "kind": "reference",
                        "sample": f"the reference's GLSL shaders on Mesa llvmpipe (oracle/_ref/glref), "
                                  f"256^3, 2+2 sweeps x 4 V-cycles to 8^3, median of 3 steps "
                                  f"({cells[3]} ms/step), build container, {cells[1]} cores; "
                                  f"from profiles/reference_llvmpipe_step.md"}
    except OSError:
        pass
    return None


def cpu_baseline():
    """The oracle (the C restatement of the reference shaders) timed on this
    host for one step of the benchmark scene at 256^3 on its OpenMP threads;
    beside it the reference shaders' own figure on llvmpipe."""
    import oracle
    from smokesim_amd import scenes
    sc = scenes.config_3d_256()
    n = sc["size"][0]
    o = oracle.Oracle(sc["size"], cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
                      min_level_size=sc["min_level_size"])
    scenes.apply_scene(o, sc)
    steps = 1
    t0 = time.perf_counter()
    for s in range(steps):
        o.set_time(s * 0.1)
        o.step()
    dt = time.perf_counter() - t0
    o.close()
    cores = oracle.threads()
    out = {"value": round(n ** 3 * steps / dt / 1e9, 6), "unit": "Gcell-steps/s", "cores": cores,
           "kind": "port",
           "sample": f"CPU oracle (C restatement of the reference shaders, -O2, {cores} OpenMP threads), "
                     f"{steps} step of the benchmark scene at {n}^3 (4 emitters, 6 boxes, 2+2 sweeps x 4 "
                     f"V-cycles to 8^3), {dt:.1f} s"}
    ref = reference_llvmpipe_256()
    if ref:
        out["reference_llvmpipe"] = ref
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.only:
        run = {"rollout128": lambda: measure_rollout(128, 32, 32, 12, reps=1),
               "rollout128ckpt8": lambda: measure_rollout(128, 32, 8, 12),
               "batch": lambda: measure_batch(64, 8),
               "batch_fused": lambda: measure_batch(64, 8, modes=("fused",)),
               "fit": lambda: measure_source_fit(iterations=2)}[a.only]
        print(json.dumps({"mode": a.only, **run()}), flush=True)
        return
    from smokesim_amd import calibrate_copy, scenes
    copy_gbs = None
    if a.calibrate:
        nbytes = 256 << 20
        ms = calibrate_copy(nbytes, reps=5)
        copy_gbs = 2 * nbytes / (ms * 1e-3) / 1e9  # read + write
    n = a.size
    sc = scenes.make_scene((n, n, n), n_emitters=4, emitter_radius=12, n_boxes=6)
    graph = not a.eager
    g = make_grid(sc, a.storage, graph=graph)
    if a.refine_main:
        g.set_mg(2, 2, 20, 4)
        g.set_refinement(a.refine, True)
    tile = None
    if not a.no_autotune:
        # the project's tile autotuner, on a scratch grid (not timed)
        tile, _ = g.autotune(3)
    dt = float(sc["physics"][4])
    # headline: the production path (captured step replayed), no events
    g.run_steps(a.warmup, 0.0, dt)
    secs = time_steps(g, a.steps, 0, dt)
    cells = n ** 3
    value = cells * a.steps / secs / 1e9
    # roofline: a timed region of its own with eager launches; every level-0
    # sweep is dispatched with HIP start/stop events attached to it
    # (hipExtLaunchKernelGGL), stamped at the kernel's own start and end on
    # the grid's stream (eager launches, so this region stays out of the
    # headline)
    from smokesim_amd import device_sync
    g.set_graph_mode(False)
    t_ev = (a.warmup + a.steps) * dt
    g.run_steps(1, t_ev, dt)
    g.profile(True)
    g.sync()
    device_sync(0)
    t0 = time.perf_counter()
    g.run_steps(a.steps, t_ev + dt, dt)
    g.sync()
    device_sync(0)
    secs_ev = time.perf_counter() - t0
    ms_l0, launches = g.profile_read("smooth_l0")
    g.profile(False)
    avg_ms = ms_l0 / max(launches, 1)
    step_gbs = STEP_BYTES_PER_CELL[a.storage] * cells * a.steps / secs / 1e9
    achieved = SWEEP_BYTES_PER_CELL * cells / (avg_ms * 1e-3) / 1e9 if launches else None
    res_norm = g.residual_norm()
    out = {
        "metric": "Gcell-steps/s at 256^3 (plain step) and % of 8 TB/s HBM roofline",
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(secs / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.storage == "fp32" else a.storage,
        "data": "synthetic",
        "config": {
            "workload": f"3d-{n}-plain-{a.storage}: {n}^3 MAC grid, 4 emitters r=12 inflow 2.0, "
                        f"6 boxes 16-32, velocity noise +-0.01, dt=0.1, 2+2 red-black sweeps, "
                        f"4 V-cycles to 8^3 (80 coarsest iterations)",
            "size": [n, n, n], "storage": a.storage, "graph": graph, "tile_bx": tile,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "kernel": "k_sweep4<3>: level-0 red-black half sweep (4 cells per thread)",
            "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic_from_profiles(n),
            "algorithmic_bytes_per_launch": SWEEP_BYTES_PER_CELL * cells,
            "avg_launch_ms": round(avg_ms, 5),
            "launches_per_step": launches // max(a.steps, 1),
            "events_region_ms_per_step": round(secs_ev / a.steps * 1e3, 4),
            # what a plain streaming copy reaches on this GPU (1 GiB, not
            # cache resident): the practical ceiling beside the 8 TB/s peak
            "copy_1gib_gbs": round(copy_ceiling_gbs(), 1),
        },
        "step_roofline": {
            "algorithmic_bytes_per_cell_step": STEP_BYTES_PER_CELL[a.storage],
            "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(step_gbs / HBM_PEAK_GBS, 4),
        },
        "residual_l2_after_run": res_norm,
    }
    if copy_gbs is not None:
        out["copy_kernel_gbs"] = round(copy_gbs, 1)
    g.close()
    if not a.no_modes:
        at = not a.no_autotune
        modes = {"bf16": measure_mode(sc, "bf16", 0, a.steps, 2, dt, at),
                 "fp32_refined": measure_mode(sc, "fp32", a.refine, 10, 2, dt, at),
                 "bf16_refined": measure_mode(sc, "bf16", a.refine, 10, 2, dt, at)}
        modes["bf16_over_fp32_refined"] = round(
            modes["bf16_refined"]["Gcell-steps/s"] / modes["fp32_refined"]["Gcell-steps/s"], 3)
        # every step state kept (about 2.7 GB at 128^3: no recomputation in
        # the backward pass), and the checkpointed variant beside it
        modes["rollout_128_fwd_adjoint"] = measure_rollout(128, 32, 32, 12)
        modes["rollout_128_fwd_adjoint_ckpt8"] = measure_rollout(128, 32, 8, 12)
        modes["rollout_256_fwd_adjoint"] = measure_rollout(256, 8, 8, 12, reps=1)
        modes["batch_64_uneven"] = measure_batch(64, 8)
        if not a.no_fit:
            modes["source_fit_64_uneven"] = measure_source_fit()
        out["modes"] = modes
    if a.extra:
        ext = {}
        gg = make_grid(sc, a.storage, graph=True)
        s = time_steps(gg, a.steps, a.warmup, dt)
        ext[f"3d-{n}-plain-{a.storage}-graph"] = {"Gcell-steps/s": round(cells * a.steps / s / 1e9, 4),
                                                  "ms_per_step": round(s / a.steps * 1e3, 4)}
        gg.close()
        for name, scn in (("3d-128-plain-fp32", scenes.config_3d_128()),
                          ("2d-256-plain-fp32", scenes.config_2d_256())):
            gg = make_grid(scn, "fp32", graph=True)
            s = time_steps(gg, 50, 5, float(scn["physics"][4]))
            c = int(np.prod(scn["size"]))
            ext[name] = {"Gcell-steps/s": round(c * 50 / s / 1e9, 4), "ms_per_step": round(s / 50 * 1e3, 4)}
            gg.close()
        out["extra"] = ext
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
