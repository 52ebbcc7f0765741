"""Binary restart of the device state (dem_write_restart / dem_read_restart)
on the MI355X.

* the file's contact-history section is FixContactHistory::pack_restart's
  per-atom record (fix_contact_history.cpp:588-602) of exactly the flagged
  pairs the device holds, with their history;
* a context read from the file holds the same atoms, wall history, groups
  and mesh contact slots; after its set-up rebuild it carries the history of
  every pair that touches;
* both continue over 100 steps to the same trajectory.  A run after
  read_restart starts with Verlet::setup (verlet.cpp:88-140): a rebuild,
  which carries the history of the pairs that touch and drops the stale
  rest, and a set-up force pass.  The continued context gets the same
  set-up, as a second run command gives it, so only the order of force
  sums differs (the restarted context sorts from upload order).
"""
import numpy as np
import pytest

from conftest import load_engine
from test_gpu_mesh import mesh_dict, wall_spheres

pytestmark = pytest.mark.gpu


def parse_restart(path):
    with open(path, "rb") as fp:
        buf = fp.read()
    head = np.frombuffer(buf, dtype=np.int32, count=6, offset=8)
    n, dnum, wdn, nmesh, dw = (int(v) for v in head[1:])
    off = 56
    ai = np.frombuffer(buf, dtype=np.int32, count=4 * n, offset=off).reshape(n, 4)
    off += 16 * n
    ad = np.frombuffer(buf, dtype=np.float64, count=11 * n, offset=off).reshape(n, 11)
    off += 88 * n
    nrec = int(np.frombuffer(buf, dtype=np.int64, count=1, offset=off)[0])
    off += 8
    rec = np.frombuffer(buf, dtype=np.float64, count=nrec, offset=off)
    recs = {}
    q = 0
    for o in range(n):
        size = int(rec[q])
        npart = int(rec[q + 1:q + 2].view(np.int64)[0])
        assert size == (dnum + 1) * npart + 2
        p = q + 2
        for _ in range(npart):
            tj = int(rec[p:p + 1].view(np.int64)[0])
            recs[(int(ai[o, 0]), tj)] = rec[p + 1:p + 1 + dnum].copy()
            p += 1 + dnum
        q += size
    return ai, ad, recs


def flagged(s):
    ti, tj, sh, hi, fl = s.pairs()
    k = fl != 0
    return {(a, b): h for a, b, h in zip(ti[k].tolist(), tj[k].tolist(), hi[k])}


def test_restart_pour_round_trip(tmp_path):
    eng = load_engine()
    spec, atoms = eng.workloads.pour_box()
    s1 = eng.Simulation(spec, atoms)
    s1.set_group(4, (np.arange(len(atoms["radius"])) % 3 == 0).astype(np.int32))
    s1.setup()
    s1.step(3003)
    path = str(tmp_path / "pour.restart")
    s1.write_restart(path)
    ai, ad, recs = parse_restart(path)
    st1 = s1.state()
    F1 = flagged(s1)
    assert len(F1) > 500
    assert set(recs) == set(F1)
    assert all(np.array_equal(recs[k], F1[k]) for k in F1)
    o = np.argsort(ai[:, 0])
    assert np.array_equal(ad[o, 0:3], st1["x"][np.argsort(st1["tag"])])
    assert np.array_equal(ai[o, 3] >> 4 & 1, (np.arange(len(o)) % 3 == 0).astype(np.int32))

    s2 = eng.Simulation(spec, atoms)
    s2.read_restart(path)
    assert s2.timestep() == s1.timestep()
    st2 = s2.state()
    for k in ("x", "v", "omega"):
        assert np.array_equal(st2[k][np.argsort(st2["tag"])], st1[k][np.argsort(st1["tag"])])
    w1, w2 = s1.wall_history(), s2.wall_history()
    assert np.array_equal(w2[np.argsort(st2["tag"])], w1[np.argsort(st1["tag"])])
    s2.setup()
    F2 = flagged(s2)
    # the set-up rebuild keeps the history of the pairs that touch now
    x = dict(zip(st1["tag"].tolist(), st1["x"]))
    r = dict(zip(atoms["tag"].tolist(), atoms["radius"]))
    touching = {k for k in F1 if np.sum((x[k[0]] - x[k[1]]) ** 2) < (r[k[0]] + r[k[1]]) ** 2}
    assert touching and touching == set(F2)
    assert all(np.array_equal(F2[k], F1[k]) for k in touching)
    s1.setup()
    s1.step(100)
    s2.step(100)
    a, b = s1.state(), s2.state()
    oa, ob = np.argsort(a["tag"]), np.argsort(b["tag"])
    assert np.abs(a["x"][oa] - b["x"][ob]).max() <= 1e-13
    assert np.abs(a["f"][oa] - b["f"][ob]).max() <= 1e-8 * np.abs(a["f"]).max()


def test_restart_mesh_slots(tmp_path):
    eng = load_engine()
    spec, atoms = eng.workloads.hopper(n=2000)
    atoms = wall_spheres(spec, atoms)
    s1 = eng.Simulation(spec, atoms)
    s1.setup()
    s1.step(150)
    path = str(tmp_path / "hopper.restart")
    s1.write_restart(path)
    M1 = mesh_dict(*s1.mesh_contacts(0))
    assert len(M1) > 100
    s2 = eng.Simulation(spec, atoms)
    s2.read_restart(path)
    s2.setup()
    M2 = mesh_dict(*s2.mesh_contacts(0))
    assert set(M1) == set(M2)
    assert all(M1[k][0] == M2[k][0] and np.array_equal(M1[k][1], M2[k][1]) for k in M1)
