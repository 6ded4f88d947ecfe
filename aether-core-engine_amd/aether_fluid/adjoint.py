"""Tape API of the differentiable smoke rollout, in the shape of a
torch.autograd.Function: forward(controls, beta) -> loss records what
backward needs; backward(dloss) returns the control and beta gradients
computed by the library's hand-written adjoint kernels."""
from __future__ import annotations

import numpy as np


class SmokeRollout:
    """A differentiable rollout of `solver` (its state is the initial
    condition and is not modified)."""

    def __init__(self, solver, nsteps: int, dt: float, patch_lo, patch_hi, target,
                 checkpoint_every: int = 8, adjoint_cg_iters=None, adjoint_cg_rel_tol=None):
        self.solver = solver
        self.kw = dict(nsteps=nsteps, dt=dt, patch_lo=tuple(patch_lo), patch_hi=tuple(patch_hi),
                       target=np.asarray(target, dtype=np.float64),
                       checkpoint_every=checkpoint_every, adjoint_cg_iters=adjoint_cg_iters,
                       adjoint_cg_rel_tol=adjoint_cg_rel_tol)
        self._saved = None

    def forward(self, controls, beta: float, need_grad: bool = True) -> float:
        loss, gc, gb = self.solver.rollout(controls=controls, beta=beta, grad=need_grad,
                                           **self.kw)
        self._saved = (gc, gb)
        return loss

    def backward(self, dloss: float = 1.0):
        if self._saved is None or self._saved[0] is None:
            raise RuntimeError("forward(..., need_grad=True) must run before backward()")
        gc, gb = self._saved
        return dloss * gc, dloss * gb


def torch_function():
    """The same rollout as a torch.autograd.Function (when torch imports):
    loss = Fn.apply(controls_tensor, beta_tensor, rollout)."""
    import torch

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, controls, beta, rollout):
            loss = rollout.forward(controls.detach().cpu().double().numpy(),
                                   float(beta.detach().cpu()), need_grad=True)
            ctx.rollout = rollout
            ctx.dev = (controls.device, controls.dtype, beta.device, beta.dtype)
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            gc, gb = ctx.rollout.backward(float(g))
            cd, ct, bd, bt = ctx.dev
            return (torch.as_tensor(gc, device=cd, dtype=ct),
                    torch.as_tensor(gb, device=bd, dtype=bt), None)

    return _Fn
