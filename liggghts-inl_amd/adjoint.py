"""Gradients of batched-pour observables with respect to each instance's
Young's modulus, friction coefficient and restitution coefficient.

``BatchObservables`` is a torch.autograd.Function over the engine's adjoint
window: its forward runs the batch from the given parameters for ``nsteps``
(``dem_adjoint_forward``, checkpoints every ``checkpoint_every`` steps) and
returns, per instance, the mean height of the spheres and the observed
plane's force on them at the last step; its backward runs the reverse sweep
(``dem_adjoint_backward``) and maps the gradients with respect to the
material tables (Yeff, Geff, betaeff, mu) onto (Y, mu, e) by the chain rule
of the table formulas (global_properties.cpp:428-562, like materials):
Yeff = Y / (2 (1 - nu^2)), Geff = Y / (4 (2 - nu)(1 + nu)),
betaeff = ln e / sqrt(ln^2 e + pi^2).

torch carries the per-instance parameters and gradients (CPU tensors); the
simulation itself lives in the engine's device memory.
"""
import numpy as np
import torch

from .engine import Simulation


def table_jacobian(Y, e, nu):
    """d Yeff/dY, d Geff/dY, d betaeff/de for like materials."""
    Y = np.asarray(Y, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    dYeff = np.full_like(Y, 1.0 / (2.0 * (1.0 - nu * nu)))
    dGeff = np.full_like(Y, 1.0 / (4.0 * (2.0 - nu) * (1.0 + nu)))
    lg = np.log(e)
    dbeta = (1.0 / e) * np.pi ** 2 / (lg * lg + np.pi ** 2) ** 1.5
    return dYeff, dGeff, dbeta


class BatchObservables(torch.autograd.Function):
    """(height, wall) = BatchObservables.apply(Y, mu, e, spec, atoms, inst,
    nsteps, checkpoint_every); Y, mu, e: float64 tensors, one per instance."""

    @staticmethod
    def forward(ctx, Y, mu, e, spec, atoms, inst, nsteps, checkpoint_every):
        params = dict(Y=Y.detach().cpu().numpy().astype(np.float64), e=e.detach().cpu().numpy().astype(np.float64),
                      mu=mu.detach().cpu().numpy().astype(np.float64))
        sp = dict(spec, instances=dict(inst=inst, **params))
        sp.setdefault("observe_plane", 4)
        sim = Simulation(sp, atoms)
        sim.setup()
        sim.adjoint_forward(int(nsteps), int(checkpoint_every))
        obs = sim.observables()
        ctx.sim = sim
        ctx.params = params
        ctx.nu = float(np.asarray(spec["materials"]["nu"])[1])
        height = torch.from_numpy(obs[:, 1] / obs[:, 0])
        wall = torch.from_numpy(obs[:, 2].copy())
        return height, wall

    @staticmethod
    def backward(ctx, g_height, g_wall):
        wh = g_height.detach().cpu().numpy().astype(np.float64) if g_height is not None else None
        ww = g_wall.detach().cpu().numpy().astype(np.float64) if g_wall is not None else None
        g = ctx.sim.adjoint_backward(wh, ww)
        dYeff, dGeff, dbeta = table_jacobian(ctx.params["Y"], ctx.params["e"], ctx.nu)
        dY = g[:, 0] * dYeff + g[:, 1] * dGeff
        de = g[:, 2] * dbeta
        dmu = g[:, 3]
        return (torch.from_numpy(dY), torch.from_numpy(dmu), torch.from_numpy(de), None, None, None, None, None)
