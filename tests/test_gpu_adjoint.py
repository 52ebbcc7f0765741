"""Adjoint gradients on the MI355X (FP64), checked against central finite
differences of the forward run.

A batch of small pours, each with its own Y, e, mu, stepped over a window
with collisions between spheres and with the walls.  Observables per
instance at the last step: mean height of the spheres and the floor's force
on them.  Their gradients with respect to (Y, mu, e) of every instance come
from one reverse sweep (torch.autograd through BatchObservables); the
reference gradient is the central difference of two forward runs with every
instance's parameter moved by a relative 1e-6 (instances are independent,
so one pair of runs gives every instance's derivative).  Also: the window's
forward (fixed list, history mask where rebuilds fall) reaches the state of
normal stepping, and the gradient does not depend on the checkpoint
interval (recomputation is deterministic).
"""
import importlib

import numpy as np
import pytest
import torch

from conftest import load_engine

pytestmark = pytest.mark.gpu

STEPS = 1200


def batch(ninst=3, nmax=700, seed=31):
    W = load_engine().workloads
    spec, atoms, inst, params = W.batch_pours(ninst=ninst, seed=seed, nmin=300, nmax=nmax)
    # the window keeps one list: half the skin (0.4 mm) covers 1200 steps at
    # the pours' speeds (|v| <= 0.17 m/s, plus gravity)
    spec = dict(spec, skin=8e-4)
    return spec, atoms, inst, params


def observe(spec, atoms, inst, params, steps):
    eng = load_engine()
    s = eng.Simulation(dict(spec, instances=dict(inst=inst, **params), observe_plane=4), atoms)
    s.setup()
    s.step(steps)
    o = s.observables()
    return o[:, 1] / o[:, 0], o[:, 2], s


def test_window_forward_matches_stepping():
    spec, atoms, inst, params = batch()
    eng = load_engine()
    h0, w0, s0 = observe(spec, atoms, inst, params, STEPS)
    s = eng.Simulation(dict(spec, instances=dict(inst=inst, **params), observe_plane=4), atoms)
    s.setup()
    s.adjoint_forward(STEPS, 100)
    o = s.observables()
    assert np.abs(o[:, 1] / o[:, 0] - h0).max() <= 1e-12 * np.abs(h0).max()
    assert np.abs(o[:, 2] - w0).max() <= 1e-9 * np.abs(w0).max()
    a, b = s.state(), s0.state()
    assert np.abs(a["x"] - b["x"]).max() <= 1e-12 * 0.1
    assert np.abs(a["v"] - b["v"]).max() <= 1e-9 * np.abs(b["v"]).max()
    assert (w0 > 0).all()           # every instance is on the floor


def test_gradients_against_finite_differences():
    eng = load_engine()
    adj = importlib.import_module("liggghts_inl_amd.adjoint")
    spec, atoms, inst, params = batch()
    B = int(inst.max()) + 1
    Y = torch.tensor(params["Y"], dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(params["mu"], dtype=torch.float64, requires_grad=True)
    e = torch.tensor(params["e"], dtype=torch.float64, requires_grad=True)
    grads = {}
    for name, wsel in (("height", (1.0, 0.0)), ("wall", (0.0, 1.0))):
        h, w = adj.BatchObservables.apply(Y, mu, e, spec, atoms, inst, STEPS, 100)
        (wsel[0] * h.sum() + wsel[1] * w.sum()).backward()
        grads[name] = {k: t.grad.numpy().copy() for k, t in (("Y", Y), ("mu", mu), ("e", e))}
        for t in (Y, mu, e):
            t.grad = None
    h0, w0, _ = observe(spec, atoms, inst, params, STEPS)
    J = {"height": np.abs(h0), "wall": np.abs(w0)}
    for key in ("Y", "mu", "e"):
        p = params[key]
        hstep = 1e-6 * p
        pp = dict(params)
        pm = dict(params)
        pp[key] = p + hstep
        pm[key] = p - hstep
        hp, wp, _ = observe(spec, atoms, inst, pp, STEPS)
        hm, wm, _ = observe(spec, atoms, inst, pm, STEPS)
        fd = {"height": (hp - hm) / (2 * hstep), "wall": (wp - wm) / (2 * hstep)}
        for name in ("height", "wall"):
            g = grads[name][key]
            print(key, name, "adjoint", g, "fd", fd[name], flush=True)
            # 1e-4 of the finite difference; where the observable does not
            # depend on the parameter (the floor force on mu at this state),
            # a floor of 1e-8 of the observable per unit relative change
            assert np.all(np.abs(g - fd[name]) <= 1e-4 * np.abs(fd[name]) + 1e-8 * J[name] / np.abs(p)), \
                (key, name, g, fd[name])
    assert B == 3


def test_gradient_independent_of_checkpoint_interval():
    spec, atoms, inst, params = batch(ninst=2, nmax=500, seed=5)
    eng = load_engine()
    out = []
    for every in (100, 350):
        s = eng.Simulation(dict(spec, instances=dict(inst=inst, **params), observe_plane=4), atoms)
        s.setup()
        s.adjoint_forward(700, every)
        out.append(s.adjoint_backward(np.ones(2), np.ones(2)))
    assert np.array_equal(out[0], out[1])
    assert np.all(np.isfinite(out[0])) and np.abs(out[0]).max() > 0
