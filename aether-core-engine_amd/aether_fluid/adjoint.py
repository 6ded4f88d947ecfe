"""Reverse-mode node of the differentiable smoke rollout.

SmokeRollout is a tape over the library's aether_fluid_rollout_forward /
aether_fluid_rollout_backward: forward(controls, beta[, initial_state])
runs the checkpointed forward sweep on the device and returns the final
state; backward(adjoints of the final state) runs the hand-written adjoint
kernels and returns the vector-Jacobian products for the controls, beta and
the initial state.  Any loss of the final state can therefore be
differentiated; torch_function() wraps the same node as a
torch.autograd.Function whose output is the final state.
"""
from __future__ import annotations

import numpy as np

FIELDS = ("u", "v", "w", "density", "temperature")


class SmokeRollout:
    """A differentiable rollout of `solver` over `nsteps` steps of `dt`,
    with density and temperature set from the controls on the inflow patch
    [patch_lo, patch_hi) and buoyancy coefficient beta.  The solver's
    current state is the initial condition; forward leaves it unchanged
    unless an initial state is passed, which is uploaded first."""

    def __init__(self, solver, nsteps: int, dt: float, patch_lo, patch_hi,
                 checkpoint_every: int = 8, adjoint_cg_iters=None, adjoint_cg_rel_tol=None):
        self.solver = solver
        self.kw = dict(nsteps=nsteps, dt=dt, patch_lo=tuple(patch_lo), patch_hi=tuple(patch_hi),
                       checkpoint_every=checkpoint_every, adjoint_cg_iters=adjoint_cg_iters,
                       adjoint_cg_rel_tol=adjoint_cg_rel_tol)
        self._recorded = False

    def forward(self, controls, beta: float, initial_state: dict | None = None) -> dict:
        """Runs the forward sweep and records the tape; returns the final
        {u, v, w, density, temperature} (flat f64 arrays)."""
        if initial_state is not None:
            for name in FIELDS:
                if name in initial_state:
                    self.solver.upload(name, initial_state[name])
        final = self.solver.rollout_forward(controls=controls, beta=beta, **self.kw)
        self._recorded = True
        return final

    def backward(self, lam_final: dict):
        """Adjoints of the final fields (missing names are zero) ->
        (d/dcontrols, d/dbeta, {field: d/d initial field}).  Consumes the
        tape of the last forward."""
        if not self._recorded:
            raise RuntimeError("forward() must run before backward()")
        self._recorded = False
        return self.solver.rollout_backward(lam_final)

    def loss_and_grad(self, controls, beta: float, target, grad: bool = True):
        """The fused form for the loss sum (density_T - target)^2:
        (loss, d/dcontrols, d/dbeta) from one library call."""
        kw = dict(self.kw)
        return self.solver.rollout(controls=controls, beta=beta, target=target, grad=grad, **kw)


def torch_function():
    """The rollout as a torch.autograd.Function:
    final = Fn.apply(controls, beta, initial_state_or_None, rollout), with
    final a (5, N) float64 tensor (u, v, w, density, temperature) and
    initial_state a (5, N) tensor or None (then the solver's state).  The
    backward pass is the library's adjoint."""
    import torch

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, controls, beta, initial_state, rollout):
            init = None
            if initial_state is not None:
                a = initial_state.detach().cpu().double().numpy()
                init = {name: a[q] for q, name in enumerate(FIELDS)}
            final = rollout.forward(controls.detach().cpu().double().numpy(),
                                    float(beta.detach().cpu()), init)
            ctx.rollout = rollout
            ctx.meta = (controls.device, controls.dtype, beta.device, beta.dtype,
                        None if initial_state is None else (initial_state.device,
                                                            initial_state.dtype))
            return torch.from_numpy(np.stack([final[name] for name in FIELDS]))

        @staticmethod
        def backward(ctx, g):
            gn = g.detach().cpu().double().numpy()
            gc, gb, gi = ctx.rollout.backward({name: gn[q] for q, name in enumerate(FIELDS)})
            cd, ct, bd, bt, im = ctx.meta
            g_init = None
            if im is not None:
                g_init = torch.as_tensor(np.stack([gi[name] for name in FIELDS]), device=im[0],
                                         dtype=im[1])
            return (torch.as_tensor(gc, device=cd, dtype=ct),
                    torch.as_tensor(gb, device=bd, dtype=bt), g_init, None)

    return _Fn
