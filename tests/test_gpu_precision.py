"""GPU parity of the mixed-precision modes through the C-ABI.

- 16-bit storage: BF16 faces + BF16 smoke holding T - t_ambient, or FP16
  faces + FP16 smoke.  The float oracle rounds the same stores the same way,
  so the bar stays rel-L2 <= 1e-5 per step.
- FP64 iterative refinement of the pressure solve.  The per-cell FP64
  residual passes, the correction V-cycles and the FP64 coarsest level follow
  the oracle operation for operation, so the bar is the same.  Only the norms
  use atomics, whose summation order differs, so they are compared to 1e-9.
- Convergence: refinement reaches a relative residual <= 1e-6.
- Error of 16-bit storage against the FP64 oracle: reported and bounded.
"""
import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / n if n > 0 else d


def fields(dim):
    f = ["vx", "vy", "smoke", "pressure", "vorticity", "solid"]
    if dim == 3:
        f.insert(2, "vz")
    return f


def compare(g, o, names, tol=TOL, where=""):
    for f in names:
        a, b = g.get(f), o.get(f)
        assert a.shape == b.shape, f
        if f == "solid":
            assert np.array_equal(a, b), f"{where} solid mask differs"
            continue
        e = rel_l2(a, b)
        assert e <= tol, f"{where} {f}: rel L2 {e:.3e} (max abs {np.abs(a - b).max():.3e})"


def make_pair(size, sc, storage, refine=0, fp64_oracle=False):
    from oracle import Oracle
    from smokesim_amd import Grid
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], storage=storage,
             min_level_size=sc["min_level_size"])
    o = Oracle(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"],
               min_level_size=sc["min_level_size"], storage=storage, fp64=fp64_oracle)
    scenes.apply_scene(g, sc)
    scenes.apply_scene(o, sc)
    if refine:
        g.set_refinement(refine, True)
        o.set_refinement(refine, True)
    return g, o


CASES = {
    "3d": ((24, 20, 16), dict(n_emitters=2, emitter_radius=4, n_boxes=1, box_sides=(4, 6),
                              min_level_size=4)),
    "2d": ((64, 48), dict(n_emitters=1, emitter_radius=6, n_spheres=1, sphere_radius=6,
                          min_level_size=4)),
    "3d_odd": ((27, 21, 19), dict(n_emitters=2, emitter_radius=4, n_boxes=2, n_spheres=1,
                                  sphere_radius=4, box_sides=(3, 6), min_level_size=4)),
}


@pytest.mark.parametrize("storage", ["bf16", "fp16"])
@pytest.mark.parametrize("case", ["3d", "2d"])
def test_16bit_storage_step_parity(case, storage, oracle_built):
    size, kw = CASES[case]
    sc = scenes.make_scene(size, **kw)
    g, o = make_pair(size, sc, storage)
    print(f"16-bit storage {storage} case {case} size {size}", flush=True)
    for s in range(5):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        compare(g, o, fields(len(size)), where=f"{storage} step {s}")
    # the temperature channel comes back in kelvin
    t = g.get("smoke")[3::4]
    assert abs(np.median(t) - sc["t_ambient"]) < 1.0


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
@pytest.mark.parametrize("case", ["3d", "2d", "3d_odd"])
def test_refined_step_parity(case, storage, oracle_built):
    size, kw = CASES[case]
    sc = scenes.make_scene(size, **kw)
    g, o = make_pair(size, sc, storage, refine=4)
    print(f"refined {storage} case {case} size {size}", flush=True)
    for s in range(4):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
        compare(g, o, fields(len(size)), where=f"refined {storage} step {s}")
        rg, bg = g.refine_norms()
        ro, bo = o.refine_norms()
        np.testing.assert_allclose(rg, ro, rtol=1e-9, atol=1e-300)
        assert abs(bg - bo) <= 1e-9 * bo


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_refinement_reaches_1e6(storage, oracle_built):
    n = 48
    sc = scenes.make_scene((n, n, n), n_emitters=2, emitter_radius=5, n_boxes=1,
                           box_sides=(8, 14), min_level_size=4)
    from smokesim_amd import Grid
    g = Grid((n, n, n), cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], storage=storage,
             min_level_size=4)
    scenes.apply_scene(g, sc)
    g.set_refinement(12, True)
    for s in range(3):
        g.set_time(s * 0.1)
        g.step()
        rn, bn = g.refine_norms()
        rel = rn / bn
        print(storage, s, np.array2string(rel, precision=2), flush=True)
        assert rel[-1] <= 1e-6
        # every correction cycle contracts the residual
        assert np.all(rel[1:] < rel[:-1])


def test_bf16_error_against_fp64_oracle(oracle_built):
    """BF16 storage + refinement on the GPU against the FP64 oracle solving the
    same steps in double (refined to 1e-12).  Reports the error of pressure,
    post-projection divergence and density; bounds are loose sanity limits."""
    from oracle import Oracle
    from smokesim_amd import Grid
    size = (32, 32, 32)
    sc = scenes.make_scene(size, n_emitters=2, emitter_radius=5, n_boxes=1, box_sides=(6, 10),
                           min_level_size=4)
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], storage="bf16",
             min_level_size=4)
    o = Oracle(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=4,
               fp64=True)
    scenes.apply_scene(g, sc)
    scenes.apply_scene(o, sc)
    g.set_refinement(12, True)
    o.set_refinement(30, True)
    for s in range(4):
        g.set_time(s * 0.1)
        o.set_time(s * 0.1)
        g.step()
        o.step()
    ep = rel_l2(g.get("pressure"), o.get("pressure"))
    dg, do = g.get("divergence"), o.get("divergence")
    v_scale = np.sqrt(np.mean(o.get("vy") ** 2)) / sc["cell_size"]
    ed = np.sqrt(np.mean((dg - do) ** 2)) / v_scale
    sg, so = g.get("smoke")[0::4], o.get("smoke")[0::4]
    es = rel_l2(sg, so)
    print(f"bf16 vs fp64 oracle after 4 steps: pressure {ep:.2e}, divergence {ed:.2e} "
          f"(of |v|/h), density {es:.2e}", flush=True)
    assert ep < 5e-2 and es < 5e-2 and ed < 5e-2
