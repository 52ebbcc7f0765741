"""Config-4 fibres on the MI355X against the oracle: bonded chains
(bond/nonlinear along each chain, created at step 1), strain-hardening
hysteretic/nonlinear1 contact between spheres, a floor plane and a plate
mesh moving down onto the top layer, gravity.

Two layers of the fibre mat (50 chains of 40 spheres, 1,950 bonds), run
from step 0 on both: the bonds the device creates are the oracle's, and
after 300 steps positions (1e-12 of the box), forces (1e-9), pair
histories (bond state included, 1e-8 of the largest slot) and the plate's
contact slots agree.
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
