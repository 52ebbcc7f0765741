"""The 4,000-sphere pour (6 planes, Hertz + history, rebuild every 10 steps)
run to its full 20,000 steps on the MI355X, in FP64 and in mixed precision,
against the oracle running the same input.

Over 20,000 steps (0.02 s) the pour is dilute (about 1,500 directed contacts
at the end), so rounding differences grow slowly: on the oracle itself, a
relative change of 1e-15 in every x-velocity moves positions by 1.4e-12 m and
velocities by 4.9e-9 m/s at step 20,000; a change of 1e-12 moves them by
9.6e-11 m and 1.3e-7 m/s.  Measured on the MI355X: FP64 ends 4.4e-15 m
from the oracle in position, 8e-13 of the largest speed in velocity,
1e-15 relative in the bulk figures, with the same 1,504 touching pairs; mixed precision ends 2.7e-6 m
away per sphere (FP32 force errors of ~1e-7 relative enter at every
collision and grow like the perturbations above), 4e-8 in mean height,
1.4e-5 / 1.1e-4 in translational / rotational kinetic energy, with 1,500
touching pairs.  Bars:
* FP64: every sphere's position within 1e-12 m, its velocity within 1e-10
  of the largest speed; mean height, height spread and translational /
  rotational kinetic energy within 1e-12 relative; the same number of
  touching directed pairs;
* mixed precision: positions within 1e-5 m (1% of the radius), mean
  height and height spread within 1e-6 relative, kinetic energies within
  1e-3 relative, touching pairs within 1%.
"""
import numpy as np
import pytest

from conftest import load_engine
from oracle import Oracle

pytestmark = pytest.mark.gpu

STEPS = 20000


def bulk(st, atoms, flags):
    m = atoms["rmass"]
    z = st["x"][:, 2]
    ke = 0.5 * (m * (st["v"] ** 2).sum(1)).sum()
    kr = 0.5 * (0.4 * m * atoms["radius"] ** 2 * (st["omega"] ** 2).sum(1)).sum()
    return dict(mean_z=z.mean(), std_z=z.std(), ke=ke, kr=kr, touching=int((flags != 0).sum()))


def by_tag(st, tags):
    o = np.argsort(tags)
    return {k: v[o] for k, v in st.items() if k in ("x", "v", "omega")}


@pytest.fixture(scope="module")
def oracle_run():
    spec, atoms = load_engine().workloads.pour_box()
    o = Oracle(spec, atoms)
    o.setup()
    o.run(STEPS)
    st = o.state()
    return spec, atoms, by_tag(st, atoms["tag"]), bulk(st, atoms, o.pairs()[4])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pour_full_run(oracle_run, precision):
    spec, atoms, ref, rb = oracle_run
    s = load_engine().Simulation(dict(spec, graph=True), atoms, precision=precision)
    s.setup()
    s.step(STEPS)
    raw = s.state()
    g = by_tag(raw, atoms["tag"])
    order = np.argsort(atoms["tag"])
    a = dict(rmass=atoms["rmass"][order], radius=atoms["radius"][order])
    gb = bulk(g, a, s.pairs()[4])
    dx = np.abs(g["x"] - ref["x"]).max()
    dv = np.abs(g["v"] - ref["v"]).max() / np.abs(ref["v"]).max()
    rel = {k: abs(gb[k] - rb[k]) / abs(rb[k]) for k in ("mean_z", "std_z", "ke", "kr")}
    print(precision, "dx", dx, "dv", dv, rel, "touching", gb["touching"], rb["touching"], flush=True)
    if precision == "fp64":
        assert dx <= 1e-12
        assert dv <= 1e-10
        assert max(rel.values()) <= 1e-12, rel
        assert gb["touching"] == rb["touching"]
    else:
        assert dx <= 1e-5
        assert max(rel["mean_z"], rel["std_z"]) <= 1e-6, rel
        assert max(rel["ke"], rel["kr"]) <= 1e-3, rel
        assert abs(gb["touching"] - rb["touching"]) <= 0.01 * rb["touching"]
