"""Config-4 fibres on the MI355X against the oracle: bonded chains
(bond/nonlinear along each chain, created at step 1), strain-hardening
hysteretic/nonlinear1 contact between spheres, a floor plane and a plate
mesh moving down onto the top layer, gravity.

Two layers of the fibre mat (50 chains of 40 spheres, 1,950 bonds), run
from step 0 on both: the bonds the device creates are the oracle's, and
after 300 steps positions (1e-12 of the box), forces (1e-9), pair
histories (bond state included, 1e-8 of the largest slot) and the plate's
contact slots agree.

The same mat over the configuration's full 20,000 steps (2 ms, the plate
0.1 mm down): the hysteretic contacts make it chaotic, so the oracle itself,
with every coordinate changed by 1e-15 relative, ends 2e-7 m away per sphere
with single forces differing by their own size, while its bulk figures move
little (kinetic energy 6e-7 relative, 2e-6 for a 1e-12 change; mean height
3e-8; every bond intact; 1,000 plate contact slots).  Bars: the same bond
count and plate contact count, kinetic energy within 1e-4 relative, mean
height within 1e-6, positions within 1e-5 m (2 % of the radius).

The configuration's full size (500 chains, 20,000 spheres, 19,500 bonds)
with bonds breaking: maxDistance cut to 1.0005 d and Gaussian velocities of
0.2 m/s per component, so about a third of the bonds break within 300
steps.  The device breaks the same bonds as the oracle (the bonded pair
sets by tag are equal), and positions, forces and pair histories agree to
the bars of the two-layer test.
"""
import numpy as np
import pytest

from conftest import load_engine
from oracle import Oracle
from test_gpu_parity import check_forces, check_pairs

pytestmark = pytest.mark.gpu


def nbonded(hist):
    # reference slot order: 12 hysteretic slots, then the bond's (flag first)
    return int((hist[:, 12] > 0.5).sum()) if len(hist) else 0


def test_fibres_against_oracle():
    eng = load_engine()
    spec, atoms = eng.workloads.fibres(nchains=50, nsph=40)
    o = Oracle(spec, atoms)
    s = eng.Simulation(spec, atoms)
    o.setup()
    s.setup()
    o.run(1)
    s.step(1)
    assert nbonded(o.pairs()[3]) == 2 * 50 * 39
    assert nbonded(s.pairs()[3]) == 2 * 50 * 39
    o.run(299)
    s.step(299)
    g, r = s.state(), o.state()
    assert np.abs(g["x"] - r["x"]).max() <= 1e-12 * 0.04
    assert np.abs(g["v"] - r["v"]).max() <= 1e-9 * np.abs(r["v"]).max()
    check_forces(s, o, 1e-9)
    # the hysteretic slots hold values up to ~4e4 (pow/exp of the device
    # math library, as in the hysteretic horizon test)
    check_pairs(s, o, hist_tol=1e-8)
    gt, gtri, gsl, gh = s.mesh_contacts(0)
    ot, otri, osl, oh = o.mesh_contacts(0, slots=True)
    G = {(a, b): h for a, b, h in zip(gt.tolist(), gtri.tolist(), gh)}
    O = {(a, b): h for a, b, h in zip(ot.tolist(), otri.tolist(), oh)}
    assert len(G) > 0 and set(G) == set(O)
    sc = max(np.abs(oh).max(), 1e-300)
    assert max(np.abs(G[k] - O[k]).max() for k in O) <= 1e-9 * sc


def test_fibres_full_run_bulk():
    eng = load_engine()
    spec, atoms = eng.workloads.fibres(nchains=50, nsph=40)
    o = Oracle(spec, atoms)
    s = eng.Simulation(spec, atoms)
    o.setup()
    s.setup()
    o.run(20000)
    s.step(20000)
    g, r = s.state(), o.state()
    m = atoms["rmass"]
    ke = lambda st: 0.5 * (m * (st["v"] ** 2).sum(1)).sum()  # noqa: E731
    bonds_g, bonds_o = nbonded(s.pairs()[3]) // 2, nbonded(o.pairs()[3]) // 2
    mc_g, mc_o = len(s.mesh_contacts(0)[0]), len(o.mesh_contacts(0, slots=True)[0])
    dke = abs(ke(g) - ke(r)) / ke(r)
    dmz = abs(g["x"][:, 2].mean() - r["x"][:, 2].mean()) / abs(r["x"][:, 2].mean())
    dx = np.abs(g["x"] - r["x"]).max()
    print("bonds", bonds_g, bonds_o, "plate contacts", mc_g, mc_o, "ke", dke, "mean z", dmz, "dx", dx, flush=True)
    assert bonds_g == bonds_o == 1950
    assert mc_g == mc_o
    assert dke <= 1e-4
    assert dmz <= 1e-6
    assert dx <= 1e-5


def test_fibres_full_size_bonds_break():
    eng = load_engine()
    spec, atoms = eng.workloads.fibres(nchains=500, nsph=40)
    d = 2 * atoms["radius"][0]
    bond = dict(spec["bond"])
    tables = dict(bond["tables"])
    tables["maxDistance"] = np.full((3, 3), 1.0005 * d)
    bond["tables"] = tables
    spec = dict(spec, bond=bond)
    atoms = dict(atoms, v=np.random.default_rng(7).normal(0.0, 0.2, atoms["x"].shape))
    o = Oracle(spec, atoms)
    s = eng.Simulation(spec, atoms)
    o.setup()
    s.setup()
    o.run(1)
    s.step(1)
    assert nbonded(o.pairs()[3]) == nbonded(s.pairs()[3]) == 2 * 500 * 39
    o.run(299)
    s.step(299)

    def bonded_set(sim):
        ti, tj, _, h = sim.pairs()[:4]
        on = h[:, 12] > 0.5 if len(h) else np.zeros(0, bool)
        return {(min(a, b), max(a, b)) for a, b in zip(np.asarray(ti)[on].tolist(), np.asarray(tj)[on].tolist())}

    bg, bo = bonded_set(s), bonded_set(o)
    print("intact bonds", len(bg), len(bo), "of", 500 * 39, flush=True)
    assert bg == bo
    assert 0.3 * 500 * 39 < len(bo) < 0.9 * 500 * 39   # a sizeable share broke
    g, r = s.state(), o.state()
    assert np.abs(g["x"] - r["x"]).max() <= 1e-12 * 0.04
    assert np.abs(g["v"] - r["v"]).max() <= 1e-9 * np.abs(r["v"]).max()
    check_forces(s, o, 1e-9)
    check_pairs(s, o, hist_tol=1e-8)
