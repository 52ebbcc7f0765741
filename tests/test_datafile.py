"""read_data / write_data (liggghts-inl_amd/datafile.py).

CPU: the reference's own data files (examples/LIGGGHTS/Tutorials_public:
hydrogel_multicontact/data/input.data, 514 spheres in a fixed box, and
ParScale/fluidizedBedTempParScale/particleFew.data, 10 spheres with tags out
of order, boundary p p f) read with the header, Atoms and Velocities rules
of the reference; mass from diameter and density as AtomVecSphere does;
write_data -> read_data reproduces every double; periodic remap with image
flags and the dropping of atoms outside a fixed boundary.
GPU: the engine and the oracle started from the same reference data file
give the same pair set and forces (1e-12 of the force scale) after the
set-up pass.
"""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, load_engine

DATA = os.path.join(ROOT, "tests", "golden", "reference_data")


def DF():
    return load_engine().datafile


def test_hydrogel_file():
    d = DF().read_data(os.path.join(DATA, "hydrogel_multicontact_input.data"))
    a = d["atoms"]
    assert d["natoms_header"] == 514 and len(a["tag"]) == 514
    assert d["ntypes"] == 2
    assert d["box_lo"] == [-0.0101, -0.0101, -0.0101] and d["box_hi"] == [0.1701, 0.1701, 0.4101]
    # first line: 1 1 0.0192777 11.5 0.0281955 0.15621 0.0747432
    assert a["tag"][0] == 1 and a["type"][0] == 1
    assert a["radius"][0] == 0.5 * 0.0192777
    assert np.array_equal(a["x"][0], [0.0281955, 0.15621, 0.0747432])
    r = a["radius"]
    assert np.array_equal(a["rmass"], 4.0 * math.pi / 3.0 * r * r * r * 11.5)
    # last Velocities line: 514 -0.000004905 -9.81703E-005 -0.000239189 0.0 0.0 0.0
    k = int(np.nonzero(a["tag"] == 514)[0][0])
    assert np.array_equal(a["v"][k], [-0.000004905, -9.81703e-5, -0.000239189])
    assert not a["omega"].any()
    assert not a["image"].any()


def test_particle_few_tags_out_of_order():
    d = DF().read_data(os.path.join(DATA, "fluidizedBedTempParScale_particleFew.data"), periodic=(1, 1, 0))
    a = d["atoms"]
    assert list(a["tag"]) == [1, 7, 8, 4, 5, 6, 2, 3, 9, 10]
    assert d["ntypes"] == 3
    vz = dict(zip(a["tag"].tolist(), a["v"][:, 2].tolist()))
    assert vz[1] == vz[7] == vz[8] == 1.0
    assert all(vz[t] == -1.0 for t in (4, 5, 6, 2, 3, 9, 10))
    assert np.all(a["radius"] == 1e-3)
    assert np.allclose(a["rmass"], 4 / 3 * math.pi * 1e-9 * 2500, rtol=1e-15, atol=0)


def test_write_read_round_trip(tmp_path):
    d = DF().read_data(os.path.join(DATA, "hydrogel_multicontact_input.data"))
    a = d["atoms"]
    rng = np.random.default_rng(3)
    a["omega"] = rng.normal(size=a["omega"].shape)
    a["image"] = rng.integers(-3, 4, size=a["image"].shape).astype(np.int32)
    p = str(tmp_path / "out.data")
    DF().write_data(p, d["box_lo"], d["box_hi"], a, d["ntypes"], timestep=1234)
    e = DF().read_data(p)
    b = e["atoms"]
    assert e["box_lo"] == d["box_lo"] and e["box_hi"] == d["box_hi"] and e["ntypes"] == 2
    for k in ("tag", "type", "radius", "x", "v", "omega", "image"):
        assert np.array_equal(a[k], b[k]), k
    # density = m / (4/3 pi r^3) and back: two roundings, as in the reference
    assert np.all(np.abs(b["rmass"] - a["rmass"]) <= 4e-16 * a["rmass"])


def test_remap_and_fixed_boundary(tmp_path):
    p = str(tmp_path / "box.data")
    with open(p, "w") as fh:
        fh.write("title\n\n5 atoms\n1 atom types\n\n0 1 xlo xhi\n0 1 ylo yhi\n0 1 zlo zhi\n\nAtoms # sphere\n\n")
        fh.write("1 1 0.1 1000 1.25 0.5 0.5\n")          # x past hi: one period back, image +1
        fh.write("2 1 0.1 1000 -1.75 0.5 0.5 0 2 0\n")   # two periods forward, image -2 + given
        fh.write("3 1 0.1 1000 1.0 0.5 0.5\n")           # exactly at hi: wraps to lo
        fh.write("4 1 0.1 1000 0.5 0.5 1.5\n")           # outside the fixed z boundary: not created
        fh.write("5 1 0.1 1000 0.5 0.5 0.25\n")
    d = DF().read_data(p, periodic=(1, 0, 0))
    a = d["atoms"]
    assert list(a["tag"]) == [1, 2, 3, 5]
    assert a["x"][0, 0] == 1.25 - 1.0 and list(a["image"][0]) == [1, 0, 0]
    assert a["x"][1, 0] == -1.75 + 1.0 + 1.0 and list(a["image"][1]) == [-2, 2, 0]
    assert a["x"][2, 0] == 0.0 and list(a["image"][2]) == [1, 0, 0]


@pytest.mark.gpu
def test_engine_from_reference_data_file():
    """Identical start from the file: the set-up pass of the engine and of
    the oracle (2,786 touching directed pairs; the packing overlaps under
    these materials) give the same pair set and forces within 1e-12 of the
    force scale."""
    from oracle import Oracle
    eng = load_engine()
    d = eng.datafile.read_data(os.path.join(DATA, "hydrogel_multicontact_input.data"))
    a = d["atoms"]
    atoms = {k: a[k] for k in ("x", "v", "omega", "radius", "rmass", "type", "tag")}
    spec = dict(box_lo=d["box_lo"], box_hi=d["box_hi"], periodic=[0, 0, 0], dt=1e-5, skin=2e-3,
                materials=eng.workloads._materials(ntypes=2), planes=[(2, -0.0101, 1)],
                gravity=(0.0, 0.0, -9.81))
    o = Oracle(spec, atoms)
    o.setup()
    s = eng.Simulation(spec, atoms)
    s.setup()
    g, r = s.state(), o.state()
    gi, gj, gs, _, _ = s.pairs()
    oi, oj, osh, _, _ = o.pairs()
    assert set(zip(gi.tolist(), gj.tolist(), gs.tolist())) == set(zip(oi.tolist(), oj.tolist(), osh.tolist()))
    fs = np.abs(r["f"]).max()
    assert fs > 1.0
    assert np.abs(g["f"] - r["f"]).max() <= 1e-12 * fs
    assert np.abs(g["torque"] - r["torque"]).max() <= 1e-12 * max(np.abs(r["torque"]).max(), fs * 0.01)
