"""torch.autograd.Function over a differentiable rollout (smk_rollout_*).

The rollout runs on the GPU through libsmoke_hip.so; tensors cross the
boundary as host (CPU) arrays, because torch's own HIP runtime must not be
initialised in the same process as the library's.  The forward returns the
final density; the backward re-runs the forward from the saved inputs (so it
is correct even when other forwards ran in between, as in gradcheck) and
returns the hand-written adjoint's gradients w.r.t. the source and the
initial velocity (two components in 2D, three in 3D).
"""
import numpy as np
import torch


class RolloutFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rollout, src, rho0, *vel):
        ctx.rollout = rollout
        ctx.save_for_backward(src, rho0, *vel)
        _load(rollout, src, vel, rho0)
        out = rollout.forward()
        return torch.from_numpy(out.astype(np.float64)).to(src.dtype)

    @staticmethod
    def backward(ctx, grad):
        src, rho0, *vel = ctx.saved_tensors
        r = ctx.rollout
        _load(r, src, vel, rho0)
        r.forward()
        sb, *vb = r.backward(grad.detach().cpu().numpy())
        t = lambda a, like: torch.from_numpy(a.astype(np.float64)).to(like.dtype)
        return (None, t(sb, src), None) + tuple(t(a, v) for a, v in zip(vb, vel))


def _load(r, src, vel, rho0):
    f = lambda x: x.detach().cpu().numpy().astype(np.float32)
    r.set_initial(*[f(v) for v in vel], f(rho0))
    r.set_source(f(src))


def rollout(r, src, *args):
    """rollout(r, src, vx, vy, [vz,] rho0): final density of rollout `r` as a
    differentiable function of the source and the initial velocity (1-D
    tensors in GL texel order; the initial density is not differentiated)."""
    *vel, rho0 = args
    if len(vel) != r.dim:
        raise ValueError(f"expected {r.dim} velocity components")
    return RolloutFunction.apply(r, src, rho0, *vel)
