"""Oracle vs the reference's own gold outputs.

Replays LIGGGHTS-INL's examples/LIGGGHTS/INL/normal_contact_hysteretic_
nonlinear{1,2}/input.lmp on the oracle: two 3 mm spheres, the lower one
frozen, the upper one pushed down by `fix aveforce` through 20 loading /
holding / unloading stages (2,000,000 steps of 1e-6 s), five hertz primitive
planes, pair model hysteretic/nonlinear{1,2} + tangential history.  The
printed columns (time, displacement, velocity, force, group force) are
compared row by row with the gold CSVs shipped with the reference (data
copied under tests/golden/reference_gold/).

Script semantics reproduced: equal-style variables substituted with
"%.15g" (Input::substitute), `unfix loading` + `fix loading ...` moving the
aveforce to the end of the fix list from the second run on, a setup pass
(shearupdate = 0) at the start of every run, and the time bookkeeping of
update.cpp.
"""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import Oracle

GOLD = os.path.join(ROOT, "tests", "golden", "reference_gold")


def sub(x):
    """Value of ${var} for an equal-style variable (printed with %.15g)."""
    return float("%.15g" % x)


def stages():
    """(c0, t0, c1) of `my_force = c0 - (time - t0)*c1` for the 20 runs."""
    my_area = sub(0.003 * 0.003)
    flt = sub(-1000.0 * my_area)                 # force_loading_tar
    slope = sub(flt / 0.10)
    fut = sub(-10.0 * my_area)                   # force_unloading_tar
    us = sub((flt - fut) / 0.10)                 # unloading_slope
    out = [(0.0, 0.0, -slope), (flt, 0.0, 0.0), (flt, 0.2, us), (fut, 0.0, 0.0)]
    ft2 = sub(-2500.0 * my_area)
    s2 = sub((fut - ft2) / 0.1)
    us2 = sub((ft2 - fut) / 0.10)
    out += [(fut, 0.4, s2), (ft2, 0.0, 0.0), (ft2, 0.6, us2), (fut, 0.0, 0.0)]
    t = 0.8
    for stress in (-5000.0, -7500.0, -10000.0):
        ft3 = sub(stress * my_area)
        s3 = sub((fut - ft3) / 0.1)
        us3 = sub((ft3 - fut) / 0.10)
        out += [(fut, t, s3), (ft3, 0.0, 0.0), (ft3, t + 0.2, us3), (fut, 0.0, 0.0)]
        t += 0.4
    return out


def build(model):
    w = 3
    E, nu, e = 1e7, 0.3, 0.3
    full = lambda v: np.full((w, w), v)  # noqa: E731
    mats = dict(ntypes=2, Y=np.full(w, E), nu=np.full(w, nu), e=full(e), mu=full(0.0),
                mur=full(0.0), etar=full(0.0), krs=0.0,
                hysteretic=dict(LoadingStiffness=full(5e4), UnloadingStiffness=full(2.5),
                                coefficientAdhesionStiffness=full(0.0),
                                coefficientPlasticityDepth=full(0.1), pullOffForce=full(0.0),
                                alphaCustom=full(20.0), cinCustom=full(1e-7), aoneCustom=full(10e8),
                                atwoCustom=full(4e4), athreeCustom=full(8.0), kcinCustom=full(1e-4)))
    spec = dict(box_lo=[-0.1, -0.1, -0.1], box_hi=[0.1, 0.1, 0.2], periodic=[0, 0, 0], dt=1e-6,
                skin=0.01, neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off",
                normal=model, wall_normal="hertz", materials=mats, fix_order=[])
    r = 0.0015
    rho = 290.0
    m = 4.0 / 3.0 * np.pi * r ** 3 * rho
    atoms = dict(x=np.array([[0.0015, 0.0015, 0.0015], [0.0015, 0.0015, 0.0045]]), v=np.zeros((2, 3)),
                 omega=np.zeros((2, 3)), radius=np.array([r, r]), rmass=np.array([m, m]),
                 type=np.array([1, 1]), tag=np.array([1, 2]), freeze=np.array([1, 0]))
    o = Oracle(spec, atoms)
    o.set_group(1, [0, 1])                 # top_lid
    o.add_freeze()                          # fix fz bulk freeze
    o.add_setforce(1, [0.0, 0.0, None])     # fix freez top_lid setforce 0 0 NULL
    return o


def replay(model, nrows=None):
    o = build(model)
    st = stages()
    rows = []
    loading = None
    walls_added = False
    for k, (c0, t0, c1) in enumerate(st):
        if loading is not None:
            o.unfix(loading)
        loading = o.add_aveforce(1, 0.0, 0.0, c0, t0, c1)
        if not walls_added:
            for dim, pos in ((2, 0.0), (0, 0.0), (0, 0.003), (1, 0.0), (1, 0.003)):
                o.add_plane(dim, pos, 2)
            walls_added = True
        o.setup()
        for _ in range(100):
            o.run(1000)
            s = o.state()
            rows.append([o.time(), 0.003 - s["x"][1, 2], s["v"][1, 2], s["f"][1, 2],
                         o.aveforce_vector(loading)[2]])
            if nrows and len(rows) >= nrows:
                return np.array(rows)
        o.cleanup()
    return np.array(rows)


def gold(name):
    return np.loadtxt(os.path.join(GOLD, name), delimiter=",", skiprows=1)[:, :5]


@pytest.mark.parametrize("model,fname", [
    ("hysteretic/nonlinear1", "hysteretic_nonlinear1_out_loading.csv"),
    ("hysteretic/nonlinear2", "hysteretic_nonlinear2_out_loading.csv"),
])
def test_hysteretic_gold(model, fname):
    g = gold(fname)
    got = replay(model)
    assert got.shape == g.shape
    scale = np.abs(g).max(0)
    err = (np.abs(got - g) / scale).max(0)
    # column-scaled: time, displacement, velocity, net force on the top
    # sphere (a small difference of large terms), group force.  The gold is
    # printed with 15 significant digits.
    assert err[0] < 1e-15, err
    assert err[1] < 1e-13, err
    assert err[2] < 1e-10, err
    assert err[3] < 1e-9, err
    assert err[4] < 1e-12, err
