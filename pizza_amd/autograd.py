"""torch.autograd binding of the step's discrete adjoint.

``window_objective(temp0, run, objective, nsteps, ckpt_every)`` returns the
objective of ``nsteps`` steps started from the temperature field ``temp0``
(a complex128 tensor (n_m, n_r), the layout of ``Run.set_field``) as a
0-d float64 tensor whose backward runs the adjoint sweep of libpizza_amd
(pz_adjoint_backward): the gradient with respect to ``temp0`` comes from the
transposed step kernels, replayed segment by segment from the snapshots
the forward sweep kept.  The other fields of ``run`` (velocity, xi) are
taken as they were set before the call.
"""
import numpy as np
import torch


class _WindowObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, temp0, run, objective, nsteps, ckpt_every, weights):
        t0 = temp0.detach().to("cpu").numpy()
        run.set_field("temp", t0)
        run.start()
        w = None if weights is None else weights.detach().to("cpu").numpy()
        J = run.adjoint_forward(objective, nsteps, ckpt_every, w)
        ctx.run = run
        ctx.device = temp0.device
        return torch.tensor(J, dtype=torch.float64, device=temp0.device)

    @staticmethod
    def backward(ctx, gJ):
        g, _ = ctx.run.adjoint_backward(float(gJ))
        return torch.from_numpy(np.ascontiguousarray(g)).to(ctx.device), None, None, None, None, None


def window_objective(temp0, run, objective="nutop", nsteps=400, ckpt_every=20, weights=None):
    """Differentiable objective of a window of steps (see module docstring)."""
    return _WindowObjective.apply(temp0, run, objective, nsteps, ckpt_every, weights)
