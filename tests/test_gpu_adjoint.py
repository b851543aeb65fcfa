"""GPU checks of the differentiable rollout and its hand-written adjoint.

- The adjoint gradient w.r.t. the source and the initial velocity matches
  central finite differences of the GPU forward.  The forward runs in FP32,
  so the bar is the finite differences' own error, measured as the
  difference between steps h and 2h (at least 1e-3).
- Gradients are bit-reproducible (fixed-point scatter).
- Checkpointing does not change the gradient.
- The torch.autograd.Function wrapper returns the same gradients.
"""
import numpy as np
import pytest

from smokesim_amd import scenes

pytestmark = pytest.mark.gpu


def make_case(n=20, steps=4, ce=2, refine=12, seed=3):
    from smokesim_amd import Grid, Rollout
    sc = scenes.make_scene((n, n, n), n_emitters=0, n_boxes=1, box_sides=(4, 6), min_level_size=4,
                           seed=seed)
    g = Grid((n, n, n), cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=4)
    scenes.apply_scene(g, sc)
    lo, size = (n // 2 - 2, 2, n // 2 - 2), (4, 4, 4)
    r = Rollout(g, steps, ce, lo, size, refine_cycles=refine)
    rng = np.random.default_rng(seed)
    nx = ny = nz = n
    v0 = [rng.uniform(-0.05, 0.05, m).astype(np.float32)
          for m in ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1))]
    rho0 = np.zeros(n ** 3, np.float32)
    src = rng.uniform(0, 0.5, int(np.prod(size))).astype(np.float32)
    tgt_src = rng.uniform(0, 0.5, int(np.prod(size))).astype(np.float32)
    r.set_initial(*v0, rho0)
    r.set_source(tgt_src)
    target = r.forward().astype(np.float64)
    return g, r, v0, rho0, src, target


def loss_and_grad(r, v0, rho0, src, target):
    r.set_initial(*v0, rho0)
    r.set_source(src)
    rho = r.forward().astype(np.float64)
    L = float(np.sum((rho - target) ** 2))
    sb, vxb, vyb, vzb = r.backward(2.0 * (rho - target))
    return L, sb, (vxb, vyb, vzb)


def loss(r, v0, rho0, src, target):
    r.set_initial(*v0, rho0)
    r.set_source(src)
    rho = r.forward().astype(np.float64)
    return float(np.sum((rho - target) ** 2))


@pytest.mark.parametrize("n", [20, 23])
def test_rollout_gradient_matches_finite_differences(n, oracle_built):
    g, r, v0, rho0, src, target = make_case(n=n)
    L, sb, vb = loss_and_grad(r, v0, rho0, src, target)
    print(f"loss {L:.6e}", flush=True)
    rng = np.random.default_rng(5)
    h = 1e-2
    # source probes
    def fd_src(p, step):
        sp, sm = src.copy(), src.copy()
        sp[p] += step
        sm[p] -= step
        return (loss(r, v0, rho0, sp, target) - loss(r, v0, rho0, sm, target)) / (2 * step)

    probes = rng.choice(src.size, 12, replace=False)
    fd = np.array([fd_src(p, h) for p in probes])
    fd2 = np.array([fd_src(p, 2 * h) for p in probes])
    ad = np.array([float(sb[p]) for p in probes])
    es = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    ns = np.linalg.norm(fd2 - fd) / np.linalg.norm(fd)
    print("source fd", np.array2string(fd, precision=3), "\nsource ad", np.array2string(ad, precision=3),
          f"\nrel err {es:.2e} (fd h vs 2h {ns:.2e})", flush=True)
    # velocity probes on faces with a large adjoint
    hv = 1e-3

    def fd_vel(c, p, step):
        vp = [x.copy() for x in v0]
        vm = [x.copy() for x in v0]
        vp[c][p] += step
        vm[c][p] -= step
        return (loss(r, vp, rho0, src, target) - loss(r, vm, rho0, src, target)) / (2 * step)

    picks = [(c, p) for c in range(3) for p in np.argsort(-np.abs(vb[c]))[:6]]
    fdv = np.array([fd_vel(c, p, hv) for c, p in picks])
    fdv2 = np.array([fd_vel(c, p, 2 * hv) for c, p in picks])
    adv = np.array([float(vb[c][p]) for c, p in picks])
    ev = np.linalg.norm(fdv - adv) / np.linalg.norm(fdv)
    nv = np.linalg.norm(fdv2 - fdv) / np.linalg.norm(fdv)
    print("velocity fd", np.array2string(fdv, precision=3), "\nvelocity ad",
          np.array2string(adv, precision=3), f"\nrel err {ev:.2e} (fd h vs 2h {nv:.2e})", flush=True)
    # the target is 1e-3; where the FP32 forward's finite differences are
    # noisier than that (h vs 2h), 3x their own noise
    assert es <= max(1e-3, 3 * ns)
    assert ev <= max(1e-3, 3 * nv)


def test_gradient_is_reproducible_and_checkpoint_invariant(oracle_built):
    from smokesim_amd import Rollout
    g, r, v0, rho0, src, target = make_case(steps=4, ce=1)
    L1, sb1, vb1 = loss_and_grad(r, v0, rho0, src, target)
    L2, sb2, vb2 = loss_and_grad(r, v0, rho0, src, target)
    assert L1 == L2
    assert np.array_equal(sb1, sb2)
    for a, b in zip(vb1, vb2):
        assert np.array_equal(a, b)
    # the same rollout with a single segment (no recomputation)
    r.close()
    n = 20
    r2 = Rollout(g, 4, 4, (n // 2 - 2, 2, n // 2 - 2), (4, 4, 4), refine_cycles=12)
    L3, sb3, vb3 = loss_and_grad(r2, v0, rho0, src, target)
    assert L3 == L1
    assert np.array_equal(sb3, sb1)


@pytest.mark.parametrize("speed", [0.05, 3.0])
def test_binned_scatter_equals_direct_scatter(speed, oracle_built):
    """The LDS-binned adjoint advection scatter (z-marching 16x16 tiles, runs
    of 16 planes) against one global atomic per contribution (sampler path
    1): bit-identical gradients on 40x36x42 (partial tiles and z runs), with
    departure points within one cell (speed 0.05) and up to 3 cells away
    (speed 3: corners outside the bin windows take the direct atomics)."""
    from smokesim_amd import Grid, Rollout
    size = (40, 36, 42)
    sc = scenes.make_scene(size, n_emitters=0, n_boxes=1, box_sides=(4, 6), min_level_size=4, seed=9)
    g = Grid(size, cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=4)
    scenes.apply_scene(g, sc)
    nx, ny, nz = size
    rng = np.random.default_rng(9)
    v0 = [rng.uniform(-speed, speed, m).astype(np.float32)
          for m in ((nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1))]
    rho0 = rng.uniform(0, 1, nx * ny * nz).astype(np.float32)
    src = rng.uniform(0, 0.5, 64).astype(np.float32)
    rb = rng.uniform(-1, 1, nx * ny * nz).astype(np.float32)
    out = []
    for path in (1, 0):
        g.set_sampler_path(path)
        r = Rollout(g, 3, 3, (18, 2, 19), (4, 4, 4), refine_cycles=6)
        r.set_initial(*v0, rho0)
        r.set_source(src)
        r.forward()
        out.append(r.backward(rb))
        r.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert np.abs(out[0][1]).max() > 0
    g.close()


def test_autograd_function_matches(oracle_built):
    import torch
    from smokesim_amd.autograd import rollout
    g, r, v0, rho0, src, target = make_case(steps=2, ce=1)
    L, sb, vb = loss_and_grad(r, v0, rho0, src, target)
    ts = torch.tensor(src, dtype=torch.float64, requires_grad=True)
    tv = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in v0]
    rho = rollout(r, ts, *tv, torch.tensor(rho0, dtype=torch.float64))
    tl = torch.sum((rho - torch.tensor(target)) ** 2)
    tl.backward()
    assert abs(tl.item() - L) <= 1e-12 * max(1.0, L)
    np.testing.assert_allclose(ts.grad.numpy(), sb, rtol=1e-6, atol=1e-12)
    for t, a in zip(tv, vb):
        np.testing.assert_allclose(t.grad.numpy(), a, rtol=1e-6, atol=1e-12)


def test_gradient_matches_fp64_c_adjoint(oracle_built):
    """The GPU rollout (FP32 fields, 12 correction cycles) against the FP64
    C adjoint (30 cycles) on the same inputs: final density and both
    gradients, norm-wise."""
    from oracle import Oracle
    n = 20
    g, r, v0, rho0, src, target = make_case(n=n)
    L, sb, vb = loss_and_grad(r, v0, rho0, src, target)
    sc = scenes.make_scene((n, n, n), n_emitters=0, n_boxes=1, box_sides=(4, 6), min_level_size=4, seed=3)
    o = Oracle((n, n, n), cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=4, fp64=True)
    scenes.apply_scene(o, sc)
    lo, size = (n // 2 - 2, 2, n // 2 - 2), (4, 4, 4)
    v64 = [x.astype(np.float64) for x in v0]
    rho_o = o.rollout(4, lo, size, src.astype(np.float64), v64, rho0, cycles=30)
    r.set_initial(*v0, rho0)
    r.set_source(src)
    rho_g = r.forward().astype(np.float64)
    er = np.linalg.norm(rho_g - rho_o) / np.linalg.norm(rho_o)
    _, grads = o.rollout(4, lo, size, src.astype(np.float64), v64, rho0, cycles=30,
                         rho_bar=2.0 * (rho_g - target))
    es = np.linalg.norm(sb - grads[0]) / np.linalg.norm(grads[0])
    ev = [np.linalg.norm(vb[c] - grads[1 + c]) / np.linalg.norm(grads[1 + c]) for c in range(3)]
    print(f"final density {er:.2e}, source grad {es:.2e}, velocity grads "
          + ", ".join(f"{e:.2e}" for e in ev), flush=True)
    assert er < 1e-5
    assert es < 1e-4
    assert max(ev) < 1e-3


def test_torch_gradcheck(oracle_built):
    """torch.autograd.gradcheck of the autograd wrapper: the hand-written
    adjoint against torch's own central differences, for the source and one
    initial-velocity component (the others held fixed), through a fixed
    random projection of the final density.  The forward is FP32, so the
    finite-difference step is 1e-2 and the tolerances are FP32 ones."""
    import torch
    from smokesim_amd.autograd import rollout
    g, r, v0, rho0, src, target = make_case(steps=2, ce=1)
    w = torch.tensor(np.random.default_rng(9).uniform(-1, 1, rho0.size))
    vx = torch.tensor(v0[0], dtype=torch.float64)
    vz = torch.tensor(v0[2], dtype=torch.float64)
    r0 = torch.tensor(rho0, dtype=torch.float64)

    def f(s, vy):
        return (rollout(r, s, vx, vy, vz, r0) * w).sum()

    s = torch.tensor(src, dtype=torch.float64, requires_grad=True)
    vy = torch.tensor(v0[1], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(f, (s, vy), eps=1e-2, atol=1e-4, rtol=2e-2, fast_mode=True)


@pytest.mark.parametrize("storage,bars", [("bf16", (1e-2, 2e-2, 6e-2)), ("fp16", (1e-3, 3e-3, 5e-2))])
def test_16bit_storage_rollout_tracks_fp32(storage, bars, oracle_built):
    """Rollouts on a BF16 / FP16 grid keep velocities, density, checkpoints
    and saved step states in 16 bits (arithmetic, adjoint fields and the
    source in FP32).  From the same initial state (BF16-representable) the
    final density and the gradients stay within the storage's rounding of the
    FP32 rollout's (measured at 20^3, 4 steps: BF16 3.4e-3 / 6.8e-3 / 6.4e-3,
    FP16 4.1e-4 / 7.9e-4 / 1.1e-3 for density / source / velocity), and the
    gradients are bit-reproducible and checkpoint-invariant."""
    from smokesim_amd import Grid, Rollout
    n = 20
    sc = scenes.make_scene((n, n, n), n_emitters=0, n_boxes=1, box_sides=(4, 6), min_level_size=4, seed=3)
    lo, size = (n // 2 - 2, 2, n // 2 - 2), (4, 4, 4)
    rng = np.random.default_rng(3)
    nf = ((n + 1) * n * n, n * (n + 1) * n, n * n * (n + 1))
    # values exact in BF16 (and FP16): 8-bit significands
    v0 = [np.round(rng.uniform(-0.05, 0.05, m) * 2 ** 12).astype(np.float32) / 2 ** 12 for m in nf]
    rho0 = np.zeros(n ** 3, np.float32)
    src = rng.uniform(0, 0.5, 64).astype(np.float32)
    tgt = rng.uniform(0, 0.5, 64).astype(np.float32)
    out = {}
    for st, ce in (("fp32", 2), (storage, 2), (storage, 1), (storage, 4)):
        g = Grid((n, n, n), cell_size=sc["cell_size"], t_ambient=sc["t_ambient"], min_level_size=4, storage=st)
        scenes.apply_scene(g, sc)
        r = Rollout(g, 4, ce, lo, size, refine_cycles=12)
        r.set_initial(*v0, rho0)
        r.set_source(tgt)
        target = r.forward().astype(np.float64)
        r.set_initial(*v0, rho0)
        r.set_source(src)
        rho = r.forward().astype(np.float64)
        sb, *vb = r.backward((2.0 * (rho - target)).astype(np.float32))
        out[(st, ce)] = (rho, sb, vb)
        r.close()
        g.close()

    def rel(a, b):
        return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))
    a, f = out[(storage, 2)], out[("fp32", 2)]
    e = (rel(a[0], f[0]), rel(a[1], f[1]), max(rel(x, y) for x, y in zip(a[2], f[2])))
    print(f"{storage} vs fp32: final density {e[0]:.2e}, source grad {e[1]:.2e}, velocity grads {e[2]:.2e}",
          flush=True)
    assert all(x < b for x, b in zip(e, bars)) and e[0] > 0
    for ce in (1, 4):
        b = out[(storage, ce)]
        assert np.array_equal(b[0], a[0]) and np.array_equal(b[1], a[1])
        assert all(np.array_equal(x, y) for x, y in zip(b[2], a[2]))
