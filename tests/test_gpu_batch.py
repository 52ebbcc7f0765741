"""Batched runs (many independent pours in one context), graph stepping and
the autotuner, on the MI355X.

* a batch of uneven pours, each with its own Y, e, mu: every instance of
  the batched run is bitwise the same as that instance run alone on the
  device (same kernels, same per-instance cell order), with and without
  graph stepping; and within the short-horizon tolerance of the oracle
  running the instance by itself;
* per-instance observables equal numpy sums over the downloaded state, and
  the observed plane's force equals, for spheres that touch only the floor,
  f_z - m g_z of each sphere;
* the autotuner leaves the state and the contact history unchanged;
* stream ordering: graph-stepped chunks interleaved with downloads of the
  state and the per-instance observables, and with a gravity change, each
  read right after the step call returns (no synchronisation by the
  caller), equal bitwise to the same chunks stepped eagerly.  The context's
  stream does not order against the null stream the copies use; a download
  that did not wait for the graph would read a state the graph is still
  writing.
"""
import numpy as np
import pytest

from conftest import load_engine
from oracle import Oracle

pytestmark = pytest.mark.gpu


def small_batch(ninst=6, nmax=1500, seed=21):
    W = load_engine().workloads
    return W.batch_pours(ninst=ninst, seed=seed, nmin=400, nmax=nmax)


def run_batch(spec, atoms, inst, params, steps, graph=False, precision="fp64"):
    eng = load_engine()
    sp = dict(spec, instances=dict(inst=inst, **params), observe_plane=4, graph=graph)
    s = eng.Simulation(sp, atoms, precision=precision)
    s.setup()
    s.step(steps)
    return s


def by_tag(st, tags):
    o = np.argsort(tags)
    return {k: v[o] for k, v in st.items() if k in ("x", "v", "omega", "f", "torque")}


def test_batch_equals_solo_bitwise_and_graph():
    eng = load_engine()
    W = eng.workloads
    spec, atoms, inst, params = small_batch()
    steps = 95
    sb = run_batch(spec, atoms, inst, params, steps)
    gb = by_tag(sb.state(), atoms["tag"])
    sg = run_batch(spec, atoms, inst, params, steps, graph=True)
    gg = by_tag(sg.state(), atoms["tag"])
    for k in gb:
        assert np.array_equal(gb[k], gg[k]), k
    order = np.argsort(atoms["tag"])

    def cells(maxrad):   # the cell grid (cells of cutneighmax) follows the largest radius
        cut = 2 * maxrad + spec["skin"]
        return [int(np.floor((h - l) / (cut * 1.0001))) for l, h in zip(spec["box_lo"], spec["box_hi"])]

    for b in range(int(inst.max()) + 1):
        s1, a1 = W.instance(spec, atoms, inst, params, b)
        # same grid, hence the same cell order in every row
        assert cells(a1["radius"].max()) == cells(atoms["radius"].max())
        s = eng.Simulation(s1, a1)
        s.setup()
        s.step(steps)
        g1 = by_tag(s.state(), a1["tag"])
        sel = np.isin(atoms["tag"][order], a1["tag"])
        for k in g1:
            assert np.array_equal(g1[k], gb[k][sel]), (b, k)


def test_batch_against_oracle_short_horizon():
    eng = load_engine()
    W = eng.workloads
    spec, atoms, inst, params = small_batch(ninst=3, nmax=1200, seed=4)
    steps = 60
    sb = run_batch(spec, atoms, inst, params, steps)
    g = by_tag(sb.state(), atoms["tag"])
    order = np.argsort(atoms["tag"])
    for b in range(3):
        s1, a1 = W.instance(spec, atoms, inst, params, b)
        o = Oracle(s1, a1)
        o.setup()
        o.run(steps)
        r = by_tag(o.state(), a1["tag"])
        sel = np.isin(atoms["tag"][order], a1["tag"])
        assert np.abs(g["x"][sel] - r["x"]).max() <= 1e-12 * 0.1
        assert np.abs(g["v"][sel] - r["v"]).max() <= 1e-9 * np.abs(r["v"]).max()
        fs = np.abs(r["f"]).max()
        assert np.abs(g["f"][sel] - r["f"]).max() <= 1e-9 * fs


@pytest.mark.parametrize("lanes,block", [(1, 64), (4, 256), (16, 128), (-32, 256)])
def test_batch_mixed_against_oracle(lanes, block):
    """The batch in its stated mixed precision, with the contact shapes the
    autotuner chooses between: each instance is advanced on the oracle until
    it has contacts, the batch starts from those states and histories, and
    its set-up forces and a 40-step horizon match each instance's FP64
    oracle (forces within 1e-5 of each atom's force scale)."""
    eng = load_engine()
    W = eng.workloads
    spec, atoms, inst, params = small_batch(ninst=3, nmax=1200, seed=4)
    a = dict(atoms)
    a["x"], a["v"], a["omega"] = atoms["x"].copy(), atoms["v"].copy(), atoms["omega"].copy()
    oracles, hist, wh = [], [], None
    for b in range(3):
        s1, a1 = W.instance(spec, atoms, inst, params, b)
        o = Oracle(s1, a1)
        o.setup()
        o.run(1500)
        st = o.state()
        k = inst == b
        a["x"][k], a["v"][k], a["omega"][k] = st["x"], st["v"], st["omega"]
        ti, tj, sh, hi, fl = o.pairs()
        hist.append((ti[fl != 0], tj[fl != 0], hi[fl != 0]))
        w = o.wall_history()
        if wh is None:
            wh = np.zeros((len(atoms["radius"]),) + w.shape[1:])
        wh[k] = w
        oracles.append((o, a1, k))
    sp = dict(spec, instances=dict(inst=inst, **params), observe_plane=4)
    s = eng.Simulation(sp, a, precision="mixed")
    s.set_contact_shape(lanes, block)
    s.set_history(np.concatenate([h[0] for h in hist]), np.concatenate([h[1] for h in hist]),
                  np.concatenate([h[2] for h in hist]))
    s.set_wall_history(wh)
    s.setup()
    for o, _, _ in oracles:
        o.rebuild()
        o.compute_forces(0)

    def compare(tol_x):
        g = s.state()
        ncontact = 0
        for o, a1, k in oracles:
            r = o.state()
            fs, ts = o.force_scale()
            ncontact += (fs > 0).sum()
            fs = np.maximum(fs, 1e-300)
            ts = np.maximum(ts + fs * a1["radius"], 1e-300)
            assert np.abs(g["x"][k] - r["x"]).max() <= tol_x
            assert (np.abs(g["f"][k] - r["f"]).max(1) / fs).max() <= 1e-5
            assert (np.abs(g["torque"][k] - r["torque"]).max(1) / ts).max() <= 1e-5
        return ncontact

    assert compare(0.0) > 300
    s.step(40)
    for o, _, _ in oracles:
        o.run(40)
    compare(1e-11)


def test_observables():
    spec, atoms, inst, params = small_batch(ninst=4, nmax=1000, seed=8)
    s = run_batch(spec, atoms, inst, params, 40, graph=True)
    obs = s.observables()
    st = s.state()
    for b in range(4):
        k = inst == b
        assert obs[b, 0] == k.sum()
        m = atoms["rmass"][k]
        assert abs(obs[b, 1] - st["x"][k, 2].sum()) <= 1e-12 * abs(obs[b, 1])
        assert abs(obs[b, 3] - m.sum()) <= 1e-12 * m.sum()
        ke = 0.5 * (m * (st["v"][k] ** 2).sum(1)).sum()
        assert abs(obs[b, 4] - ke) <= 1e-12 * ke


def test_observed_plane_force_closed_form():
    """Spheres pressed into the floor, apart from each other: the floor's
    force on each is its f_z minus its weight."""
    spec, _, _, _ = small_batch(ninst=1)
    n = 50
    rng = np.random.default_rng(3)
    r = rng.uniform(0.8e-3, 1.2e-3, n)
    x = np.stack([0.003 + 0.0035 * (np.arange(n) % 10), 0.003 + 0.0035 * (np.arange(n) // 10),
                  r - rng.uniform(1e-6, 2e-5, n)], 1)
    atoms = dict(x=x, v=np.zeros((n, 3)), omega=np.zeros((n, 3)), radius=r, rmass=4 / 3 * np.pi * r ** 3 * 2500,
                 type=np.ones(n, np.int32), tag=np.arange(1, n + 1, dtype=np.int32))
    inst = (np.arange(n) % 2).astype(np.int32)
    params = dict(Y=np.array([2e6, 8e6]), e=np.array([0.4, 0.8]), mu=np.array([0.3, 0.6]))
    s = run_batch(spec, atoms, inst, params, 0)
    st = s.state()
    obs = s.observables()
    wall = st["f"][:, 2] - atoms["rmass"] * -9.81
    for b in range(2):
        want = wall[inst == b].sum()
        assert want > 0
        assert abs(obs[b, 2] - want) <= 1e-12 * want


def test_autotune_keeps_state():
    spec, atoms, inst, params = small_batch(ninst=4, seed=9)
    s = run_batch(spec, atoms, inst, params, 30)
    before = s.state()
    pb = s.pairs()
    g, b, ms = s.autotune(reps=2)
    assert g in (1, 2, 4, 8, 16, -8, -16, -32, -64) and b in (64, 128, 256) and ms > 0
    after = s.state()
    pa = s.pairs()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(before[k], after[k]), k
    for u, v in zip(pb, pa):
        assert np.array_equal(u, v)
    s.step(20)
    s2 = run_batch(spec, atoms, inst, params, 50)
    a, c = s.state(), s2.state()
    assert np.abs(a["x"] - c["x"]).max() <= 1e-12 * 0.1


def test_graph_follows_setters():
    """A recorded graph holds the values its kernels were given: after
    set_gravity, or a new upload with larger spheres (another cell grid),
    graph stepping must record again and stay bitwise equal to stepping
    without the graph."""
    eng = load_engine()
    spec, atoms = eng.workloads.packed_column(n=6000, slab_n=3000)
    big = dict(atoms, radius=atoms["radius"] * 1.15)

    def run(graph):
        s = eng.Simulation(dict(spec, graph=graph), atoms)
        s.setup()
        out = []
        s.step(47)
        out.append(s.state())
        s.set_gravity(0.0, 3.0, -25.0)
        s.step(46)
        out.append(s.state())
        s.upload(big)
        s.setup()
        s.step(45)
        out.append(s.state())
        s.close()
        return out

    plain, graph = run(False), run(True)
    for a, b in zip(plain, graph):
        for k in ("x", "v", "omega", "f", "torque"):
            assert np.array_equal(a[k], b[k]), k
    # the gravity change and the larger spheres do show in the state
    assert np.abs(plain[1]["v"][:, 1]).max() > 0
    assert not np.array_equal(plain[2]["f"], plain[1]["f"])


def test_graph_chunks_interleaved_with_host_copies():
    spec, atoms, inst, params = small_batch(ninst=8, nmax=3000, seed=11)
    eng = load_engine()
    runs = []
    for graph in (False, True):
        sp = dict(spec, instances=dict(inst=inst, **params), observe_plane=4, graph=graph)
        s = eng.Simulation(sp, atoms, precision="fp64")
        s.setup()
        seen = []
        for k, chunk in enumerate([20, 7, 33, 20, 1, 40, 19]):
            if k == 3:
                s.set_gravity(0.0, 0.0, -4.0)
            s.step(chunk)
            st = s.state()
            seen.append((by_tag(st, st["tag"]), s.observables()))
        runs.append(seen)
        s.close()
    for (ge, oe), (gg, og) in zip(*runs):
        for key in ge:
            assert np.array_equal(ge[key], gg[key]), key
        assert np.array_equal(oe, og)
