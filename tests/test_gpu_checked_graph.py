"""Graph stepping with neigh_modify check yes (every 1, delay 0) on the
MI355X: the distance check runs inside the recorded steps and halts the
period on the device when a rebuild is due; the host goes on eagerly from
that step.  The runs must be bit for bit the eager ones:

* the config-4 fibre mat (bonds, hysteretic contact, a plate mesh moving
  down, the floor plane), with a skin of 1 um so that the plate's trigger
  fires every 100 steps and particle checks halt periods;
* a pour with gravity and walls (the integrator fused into the contact pass,
  so each step swaps the state buffers and a halt after an odd number of
  steps leaves them swapped), rebuilt every few steps by its checks.
"""
import numpy as np
import pytest

from conftest import load_engine

pytestmark = pytest.mark.gpu


def run(spec, atoms, steps, chunks, graph):
    eng = load_engine()
    s = eng.Simulation(dict(spec, graph=graph), atoms)
    s.setup()
    for c in chunks:
        s.step(c)
    assert s.timestep() == steps
    st = s.state()
    out = dict(x=st["x"], v=st["v"], omega=st["omega"], f=st["f"])
    ti, tj, _, h, fl = s.pairs()
    o = np.lexsort((tj, ti))
    out.update(ti=ti[o], tj=tj[o], h=h[o], flag=fl[o])
    if spec.get("meshes"):
        mt, mtri, _, mh = s.mesh_contacts(0)
        o = np.lexsort((mtri, mt))
        out.update(mt=mt[o], mtri=mtri[o], mh=mh[o])
    s.close()
    return out


def same(a, b):
    assert set(a) == set(b)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_fibres_checked_graph_equals_eager():
    eng = load_engine()
    spec, atoms = eng.workloads.fibres(nchains=50, nsph=40)
    spec = dict(spec, skin=1e-6)
    chunks = [1, 999, 1500, 500]
    eager = run(spec, atoms, 3000, chunks, graph=False)
    graph = run(spec, atoms, 3000, chunks, graph=True)
    print("pairs %d, plate contacts %d" % (len(eager["ti"]), len(eager.get("mt", []))), flush=True)
    same(graph, eager)


def test_pour_checked_graph_equals_eager():
    eng = load_engine()
    spec, atoms = eng.workloads.pour_box(n=2000)
    spec = dict(spec, neigh_every=1, neigh_delay=0, neigh_check=True)
    chunks = [1, 1200, 1799]
    eager = run(spec, atoms, 3000, chunks, graph=False)
    graph = run(spec, atoms, 3000, chunks, graph=True)
    same(graph, eager)
