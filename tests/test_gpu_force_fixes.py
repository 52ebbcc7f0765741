"""Groups, fix setforce and fix aveforce on the MI355X.

* The contact pass's fused integrator against the separate kernel: bitwise.
* Semantics from an identical state: the same pour with and without the
  fixes, one set-up force pass each.  A setforce group gets the set values
  in its styled components; an aveforce group gets the group's mean force
  plus the added value; everyone else keeps its force; each fix's vector is
  the group's total force before it acted (summed in a fixed order here,
  atom order in the reference: 1e-12 relative).
* The reference's own gold for the INL hysteretic examples
  (examples/LIGGGHTS/INL/normal_contact_hysteretic_nonlinear{1,2}): the
  20 load stages replayed on the device through the same script
  semantics the oracle test follows (a set-up pass per run, `unfix
  loading` + a new aveforce per run, the time bookkeeping of update.cpp),
  every printed row compared with the gold CSV.  The engine applies the
  fixes after the walls; in this script the top sphere touches the side
  walls at exactly zero overlap and the frozen sphere never moves, so every
  wall term the order could move is exactly zero.
"""
import numpy as np
import pytest

from conftest import load_engine
from test_oracle_gold import gold, stages

pytestmark = pytest.mark.gpu


def test_setforce_aveforce_semantics():
    eng = load_engine()
    spec, atoms = eng.workloads.pour_box()
    n = len(atoms["radius"])
    rng = np.random.default_rng(2)
    ga = rng.random(n) < 0.3
    gb = rng.random(n) < 0.2
    ref = eng.Simulation(spec, atoms)
    ref.setup()
    f0 = ref.state()["f"]
    s = eng.Simulation(spec, atoms)
    s.set_group(2, ga)
    s.set_group(5, gb & ~ga)
    sid = s.add_setforce(2, [1.5, None, -0.25])
    aid = s.add_aveforce(5, None, 0.125, -3.0)
    s.setup()
    f = s.state()["f"]
    b = gb & ~ga
    rest = ~(ga | b)
    assert np.array_equal(f[rest], f0[rest])
    assert np.all(f[ga, 0] == 1.5) and np.all(f[ga, 2] == -0.25)
    assert np.array_equal(f[ga, 1], f0[ga, 1])
    scale = np.abs(f0).sum(0)
    for g, fid in ((ga, sid), (b, aid)):
        want = f0[g].sum(0)
        assert np.all(np.abs(s.force_op_vector(fid) - want) <= 1e-12 * scale), (s.force_op_vector(fid), want)
    mean = f0[b].sum(0) / b.sum()
    assert np.array_equal(f[b, 0], f0[b, 0])   # NULL x
    assert np.all(np.abs(f[b, 1] - (mean[1] + 0.125)) <= 1e-12 * scale[1])
    assert np.all(np.abs(f[b, 2] - (mean[2] - 3.0)) <= 1e-12 * scale[2])
    # unfix: the next pass leaves the group alone
    s.unfix(sid)
    s.unfix(aid)
    s.step(1)
    ref.step(1)
    st, rt = s.state(), ref.state()
    assert not np.array_equal(st["x"], rt["x"])     # the fixes moved the groups
    with pytest.raises(Exception):
        s.force_op_vector(sid)


def test_fused_integration_bitwise():
    """The contact pass integrates when nothing follows it; a setforce on an
    empty group changes no force but keeps the integrator a kernel of its
    own.  Both runs, over rebuilds, give the same bits."""
    eng = load_engine()
    spec, atoms = eng.workloads.pour_box()
    a = eng.Simulation(spec, atoms)
    b = eng.Simulation(spec, atoms)
    b.set_group(3, np.zeros(len(atoms["radius"]), dtype=np.int32))
    b.add_setforce(3, [0.0, 0.0, 0.0])
    for s in (a, b):
        s.setup()
        s.step(37)
        s.step(25)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sb[k]), k


def gold_sim(model):
    """The example's two spheres, freeze, setforce and five hertz planes."""
    eng = load_engine()
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
    planes = [(2, 0.0, 2), (0, 0.0, 2), (0, 0.003, 2), (1, 0.0, 2), (1, 0.003, 2)]
    spec = dict(box_lo=[-0.1, -0.1, -0.1], box_hi=[0.1, 0.1, 0.2], periodic=[0, 0, 0], dt=1e-6,
                skin=0.01, neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off",
                normal=model, wall_normal="hertz", materials=mats, planes=planes, freeze=True)
    r = 0.0015
    m = 4.0 / 3.0 * np.pi * r ** 3 * 290.0
    atoms = dict(x=np.array([[0.0015, 0.0015, 0.0015], [0.0015, 0.0015, 0.0045]]), v=np.zeros((2, 3)),
                 omega=np.zeros((2, 3)), radius=np.array([r, r]), rmass=np.array([m, m]),
                 type=np.array([1, 1]), tag=np.array([1, 2]), freeze=np.array([1, 0]))
    s = eng.Simulation(spec, atoms)
    s.set_group(1, [0, 1])                  # top_lid
    s.add_setforce(1, [0.0, 0.0, None])     # fix freez top_lid setforce 0 0 NULL
    return s


def replay_gpu(model, nstages):
    s = gold_sim(model)
    rows = []
    loading = None
    for c0, t0, c1 in stages()[:nstages]:
        if loading is not None:
            s.unfix(loading)
        loading = s.add_aveforce(1, 0.0, 0.0, c0, t0, c1)
        s.setup()
        for _ in range(100):
            s.step(1000)
            st = s.state()
            rows.append([s.time(), 0.003 - st["x"][1, 2], st["v"][1, 2], st["f"][1, 2],
                         s.force_op_vector(loading)[2]])
        s.cleanup()
    return np.array(rows)


@pytest.mark.parametrize("model,fname,nstages", [
    ("hysteretic/nonlinear1", "hysteretic_nonlinear1_out_loading.csv", 20),
    ("hysteretic/nonlinear2", "hysteretic_nonlinear2_out_loading.csv", 20),
])
def test_hysteretic_gold_on_device(model, fname, nstages):
    g = gold(fname)[: 100 * nstages]
    got = replay_gpu(model, nstages)
    assert got.shape == g.shape
    scale = np.abs(g).max(0)
    err = (np.abs(got - g) / scale).max(0)
    print(model, "column-scaled errors", err, flush=True)
    # all 20 load stages (2,000,000 steps), the bars of the oracle's own
    # replay of these golds: the device evaluates the same FP64 operations
    assert err[0] < 1e-15, err
    assert err[1] < 1e-13, err
    assert err[2] < 1e-10, err
    assert err[3] < 1e-9, err
    assert err[4] < 1e-12, err
