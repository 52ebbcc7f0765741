"""fix insert/pack on the MI355X against the oracle (oracle/insert_oracle.c).

* the reference's packing tutorial (examples/LIGGGHTS/Tutorials_public/
  packing/in.packing: two templates, mass fractions 0.3 / 0.7, volume
  fraction 0.25, all_in, maxattempt 200, inserted once into an empty box) on
  the device: the same particles, bit for bit, tags 1..N, at rest after the
  insertion step (no two particles and no particle and wall touch);
* insertion into a box holding resident particles, periodic in x and y: the
  device's overlap test (grid over the residents, periodic images) gives the
  oracle's brute-force result bit for bit;
* a pour with gravity and insert_every: at every insertion step the new
  particles are the ones the oracle places against the device's resident
  particles at that step, and their velocity after the step is the inserted
  one plus the final half kick;
* appending atoms carries every existing contact history: a bed with a far
  particle appended mid-run steps exactly as the bed without it;
* particletemplate/multiplespheres and create_particles single on the
  device (microstructure_compression's 114,704-sphere template, the Monte
  Carlo inside tests on the GPU): the oracle's template and placement bit
  for bit, sphere 1 where the reference's gold prints it; an oriented,
  bonded two-sphere clump;
* a pour of 50,000 spheres into the top of the 1M-sphere packed column,
  overlapping its surface: no overlaps (checked on the host with a k-d tree),
  the aimed count placed, times reported.
"""
import time

import numpy as np
import pytest

from conftest import load_engine
from test_insert import BOX, PACKING, oracle_fix, resident_bed

pytestmark = pytest.mark.gpu


def empty_atoms():
    return dict(x=np.zeros((0, 3)), v=np.zeros((0, 3)), omega=np.zeros((0, 3)), radius=np.zeros(0),
                rmass=np.zeros(0), type=np.zeros(0, np.int32), tag=np.zeros(0, np.int32))


def materials(Y=5e6, nu=0.45, e=0.3, mu=0.5, ntypes=1):
    w = ntypes + 1
    return dict(ntypes=ntypes, Y=np.full(w, Y), nu=np.full(w, nu), e=np.full((w, w), e), mu=np.full((w, w), mu))


def packing_spec():
    lo, hi = BOX
    return dict(box_lo=list(lo), box_hi=list(hi), periodic=[0, 0, 0], dt=1e-5, skin=0.002, cdf=1.0,
                neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off", materials=materials(),
                planes=[(0, lo[0], 1), (0, hi[0], 1), (1, lo[1], 1), (1, hi[1], 1), (2, lo[2], 1), (2, hi[2], 1)])


def by_tag(st, props):
    o = np.argsort(props["tag"])
    return {k: v[o] for k, v in dict(st, **props).items() if k != "tag"}, props["tag"][o]


def test_packing_tutorial_insertion_on_device():
    eng = load_engine()
    s = eng.Simulation(packing_spec(), empty_atoms())
    fid = s.add_insert_pack(dict(PACKING, group_mask=0))
    s.setup()
    assert s.n == 0
    s.step(1)
    o = oracle_fix(PACKING)
    ref = o.insert(1, np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    n = len(ref["radius"])
    print("packing tutorial: %d spheres inserted, device %d" % (n, s.n))
    assert s.n == n > 300
    st, props = s.state(), s.props()
    np.testing.assert_array_equal(props["tag"], np.arange(1, n + 1))
    np.testing.assert_array_equal(st["x"], ref["x"])
    np.testing.assert_array_equal(props["radius"], ref["radius"])
    np.testing.assert_array_equal(props["rmass"], ref["rmass"])
    np.testing.assert_array_equal(props["type"], ref["type"])
    assert (st["v"] == 0).all() and (st["f"] == 0).all()
    np.testing.assert_array_equal(s.insert_vector(fid), o.vector())
    # once: no more insertions
    s.step(50)
    assert s.n == n
    s.close()


def test_insertion_against_residents_periodic():
    eng = load_engine()
    # residents clear of each other's periodic images (0.1 periodic, r <= 6 mm)
    x, r, m = resident_bed(n=300, seed=5, half=0.036)
    n0 = len(r)
    atoms = dict(x=x, v=np.zeros((n0, 3)), omega=np.zeros((n0, 3)), radius=r, rmass=m,
                 type=np.ones(n0, np.int32), tag=np.arange(101, 101 + n0, dtype=np.int32))
    spec = dict(packing_spec(), periodic=[1, 1, 0], planes=[])
    ins = dict(PACKING, volumefraction_region=0.35, all_in=False, vel=("uniform", (0.0, 0.0, 0.0), (0.1, 0.1, 0.1)))
    s = eng.Simulation(spec, atoms)
    fid = s.add_insert_pack(dict(ins, group_mask=0))
    s.setup()
    s.step(1)
    o = oracle_fix(ins, periodic=(1, 1, 0))
    ref = o.insert(1, x, r, m)
    k = len(ref["radius"])
    print("residents %d, inserted %d (oracle %d)" % (n0, s.n - n0, k))
    assert s.n == n0 + k and k > 50
    st, props = s.state(), s.props()
    np.testing.assert_array_equal(st["x"][:n0], x)          # the residents did not move
    np.testing.assert_array_equal(st["x"][n0:], ref["x"])
    np.testing.assert_array_equal(props["tag"][n0:], np.arange(101 + n0, 101 + n0 + k))
    np.testing.assert_array_equal(props["radius"][n0:], ref["radius"])
    # the inserted velocities are set (no force acts: nothing touches)
    np.testing.assert_array_equal(st["v"][n0:], ref["v"])
    np.testing.assert_array_equal(s.insert_vector(fid), o.vector())
    s.close()


def pour_spec():
    spec = dict(box_lo=[0.0, 0.0, 0.0], box_hi=[0.04, 0.04, 0.08], periodic=[0, 0, 0], dt=1e-5, skin=2.5e-4,
                cdf=1.0, neigh_every=10, neigh_delay=0, neigh_check=False, rolling="off",
                materials=materials(Y=5e6, nu=0.3, e=0.5, mu=0.5),
                planes=[(0, 0.0, 1), (0, 0.04, 1), (1, 0.0, 1), (1, 0.04, 1), (2, 0.0, 1), (2, 0.08, 1)],
                gravity=(0.0, 0.0, -9.81))
    ins = dict(seed=15485863, dist_seed=15485867,
               templates=[dict(type=1, radius=1.0e-3, density=2500.0), dict(type=1, radius=1.3e-3, density=2500.0)],
               weights=[0.5, 0.5], region=((0.0, 0.0, 0.05), (0.04, 0.04, 0.08)), volumefraction_region=0.12,
               maxattempt=50, insert_every=1500, all_in=True,
               vel=("uniform", (0.0, 0.0, -0.6), (0.05, 0.05, 0.05)))
    return spec, ins


def test_pour_with_insert_every_matches_oracle():
    eng = load_engine()
    spec, ins = pour_spec()
    s = eng.Simulation(spec, empty_atoms())
    fid = s.add_insert_pack(dict(ins, group_mask=0))
    o = oracle_fix(ins, box=(spec["box_lo"], spec["box_hi"]))
    s.setup()
    dtf = 0.5 * spec["dt"]
    n_before = 0
    for event in range(5):
        t_ins = int(s.insert_vector(fid)[5])
        assert t_ins == o.next_step()
        s.step(t_ins - s.timestep())
        assert s.timestep() == t_ins
        st, props = s.state(), s.props()
        ref = o.insert(t_ins, st["x"][:n_before], props["radius"][:n_before], props["rmass"][:n_before])
        k = len(ref["radius"])
        print("event %d at step %d: %d resident, %d inserted" % (event, t_ins, n_before, k))
        assert s.n == n_before + k
        new = slice(n_before, n_before + k)
        np.testing.assert_array_equal(st["x"][new], ref["x"])
        np.testing.assert_array_equal(props["radius"][new], ref["radius"])
        np.testing.assert_array_equal(props["rmass"][new], ref["rmass"])
        # after the step's final half kick: v = v_ins + dtf / m * f
        np.testing.assert_array_equal(st["v"][new], ref["v"] + (dtf / ref["rmass"])[:, None] * st["f"][new])
        np.testing.assert_array_equal(s.insert_vector(fid)[:5], o.vector()[:5])
        n_before += k
    assert n_before > 1000
    eager = s.state()
    s.close()
    # graph stepping ends its periods before each insertion step, which runs
    # eagerly: the same pour
    g = eng.Simulation(dict(spec, graph=True), empty_atoms())
    g.add_insert_pack(dict(ins, group_mask=0))
    g.setup()
    g.step(eager_steps := t_ins)
    assert g.timestep() == eager_steps and g.n == n_before
    gs = g.state()
    for key in ("x", "v", "omega"):
        np.testing.assert_array_equal(gs[key], eager[key], err_msg=key)
    g.close()


def test_append_keeps_contact_history():
    """A bed with contacts; one far particle appended after 2300 steps, then
    300 more.  Rebuilding every step in both runs, the bed steps exactly as
    without it: its rows and their history are carried unchanged (the far
    particle sits in the last cell of the sort order, so the bed keeps its
    order too)."""
    eng = load_engine()
    spec, atoms = eng.workloads.pour_box(n=2000)
    spec = dict(spec, neigh_every=1)
    far = dict(x=np.array([[0.0385, 0.0385, 0.0785]]), v=np.zeros((1, 3)), omega=np.zeros((1, 3)),
               radius=np.array([1e-3]), rmass=np.array([1e-5]), type=np.ones(1, np.int32),
               tag=np.array([9999], np.int32))
    runs = []
    for append in (False, True):
        s = eng.Simulation(spec, atoms)
        s.setup()
        s.step(2300)
        if append:
            s.append_atoms(far)
        s.step(300)
        st = s.state()
        ti, tj, sh, hi, fl = s.pairs()
        keep = (ti < 9999) & (tj < 9999) & (fl != 0)
        pairs = {(a, b): h for a, b, h in zip(ti[keep].tolist(), tj[keep].tolist(), hi[keep])}
        runs.append((st, pairs, s.n))
        s.close()
    (a, pa, na), (b, pb, nb) = runs
    assert nb == na + 1
    assert len(pa) > 200
    for k in ("x", "v", "omega", "f", "torque"):
        np.testing.assert_array_equal(a[k], b[k][:na], err_msg=k)
    assert set(pa) == set(pb)
    assert all(np.array_equal(pa[k], pb[k]) for k in pa)


def test_pour_into_packed_column_1M():
    from scipy.spatial import cKDTree

    eng = load_engine()
    spec, atoms = eng.workloads.packed_column()
    top = float((atoms["x"][:, 2] + atoms["radius"]).max())
    n0 = len(atoms["radius"])
    s = eng.Simulation(spec, atoms, precision="mixed")
    ins = dict(seed=32452843, dist_seed=32452867, templates=[dict(type=1, radius=1.0e-3, density=2500.0)],
               weights=[1.0], region=((0.0, 0.0, top - 0.01), (0.06, 0.06, top + 0.3)),
               particles_in_region=50000, maxattempt=50, insert_every="once", all_in=False,
               vel=("constant", (0.0, 0.0, -0.5)))
    # the residents already in the region count toward the target
    lo, hi = np.array(ins["region"][0]), np.array(ins["region"][1])
    present = int(((atoms["x"] >= lo) & (atoms["x"] <= hi)).all(1).sum())
    ins["particles_in_region"] = present + 50000
    fid = s.add_insert_pack(dict(ins, group_mask=0))
    s.setup()
    s.sync()
    t = time.perf_counter()
    s.step(1)
    s.sync()
    dt = time.perf_counter() - t
    vec = s.insert_vector(fid)
    print("1M column: %d residents in the region, %d inserted of %d aimed, %d tries, step with insertion %.1f ms"
          % (present, vec[2], vec[3], vec[4], 1e3 * dt))
    assert vec[3] == 50000 and vec[2] == 50000
    st, props = s.state(), s.props()
    assert s.n == n0 + 50000
    x, r = st["x"], props["radius"]
    # the residents' positions here are the ones the insertion tested
    # against (a step's final half kick moves no particle); periodic in x, y
    bs = np.array([0.06, 0.06, 100.0])
    xp = np.mod(x + np.array([0.0, 0.0, 1.0]), bs)
    tree = cKDTree(xp[:n0], boxsize=bs)
    cand = tree.query_ball_point(xp[n0:], r=1.0e-3 + float(atoms["radius"].max()))
    bad = 0
    for i, js in enumerate(cand):
        if js:
            d = xp[js] - xp[n0 + i]
            d -= bs * np.round(d / bs)
            bad += int((np.sqrt((d * d).sum(1)) <= 1.0e-3 + r[js] + 1e-12).any())
    assert bad == 0
    # nor each other (they keep their radius from the box faces, so no
    # periodic image comes closer)
    assert len(cKDTree(x[n0:]).query_pairs(2.0e-3 - 1e-12)) == 0
    s.close()


def test_template_and_create_particles_on_device():
    from test_template import GOLD_SPHERE1, microstructure_template, tmpl_oracle

    eng = load_engine()
    x = microstructure_template()
    n = len(x)
    spec = dict(box_lo=[-10.0, -10.0, 0.0], box_hi=[110.0, 110.0, 200.0], periodic=[0, 0, 0], dt=2e-3, skin=2.0,
                cdf=1.0, neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off",
                materials=materials(Y=2.5e5, nu=0.3, e=0.1, mu=0.5, ntypes=2))
    s = eng.Simulation(spec, empty_atoms())
    t0 = time.perf_counter()
    tid = s.add_template_multisphere(900019, 1, 1089.0, x, np.ones(n), types=np.full(n, 2), ntry=100000)
    t_setup = time.perf_counter() - t0
    o = tmpl_oracle(900019, x, np.ones(n), ntry=100000)
    info, oi = s.template_info(tid), o.info()
    print("template of %d spheres set up in %.2f s; r_bound %.6f, volume %.1f" % (n, t_setup, info[0], info[4]))
    np.testing.assert_array_equal(info[[0, 1, 2, 3, 4, 6, 7, 8]], oi[[0, 1, 2, 3, 4, 6, 7, 8]])
    assert info[5] == info[4] * 1089.0
    assert s.create_particles_single(tid, (500.0, 0.0, 0.0)) == 0      # outside the box: none
    k = s.create_particles_single(tid, (49.0, 49.0, 99.0))
    assert k == n == s.n
    st, props = s.state(), s.props()
    np.testing.assert_array_equal(st["x"], o.place([49.0, 49.0, 99.0]))
    assert np.abs(st["x"][0] - np.array(GOLD_SPHERE1)).max() < 1e-15
    np.testing.assert_array_equal(props["tag"], np.arange(1, n + 1))
    assert (props["type"] == 2).all() and (props["radius"] == 1.0).all()
    assert (props["rmass"] == 4.18879020479 * 1.0 * 1.0 * 1.0 * 1089.0).all()
    s.close()

    # an oriented two-sphere clump, velocity and spin, after resident atoms
    s = eng.Simulation(spec, empty_atoms())
    x2, r2 = [[0.0, 0.0, 0.0], [0.0004, 0.0, 0.0]], [0.0002, 0.0002]
    tid = s.add_template_multisphere(900007, 1, 2500.0, x2, r2, ntry=100000, bonded=True)
    o2 = tmpl_oracle(900007, x2, r2, ntry=100000)
    q = [0.3, -0.2, 0.9, 0.1]
    assert s.create_particles_single(tid, (1.0, 2.0, 3.0)) == 2
    assert s.create_particles_single(tid, (4.0, 5.0, 6.0), v=(0.1, 0.2, 0.3), omega=(1.0, 0.0, 0.0), quat=q) == 2
    st, props = s.state(), s.props()
    np.testing.assert_array_equal(st["x"][:2], o2.place([1.0, 2.0, 3.0]))
    np.testing.assert_array_equal(st["x"][2:], o2.place([4.0, 5.0, 6.0], q))
    np.testing.assert_array_equal(props["tag"], [1, 2, 3, 4])
    np.testing.assert_array_equal(st["v"][2:], [[0.1, 0.2, 0.3]] * 2)
    np.testing.assert_array_equal(st["omega"][2:], [[1.0, 0.0, 0.0]] * 2)
    s.close()


def bond_tables(create=0.007):
    w = 2
    full = lambda v: np.full((w, w), v)  # noqa: E731
    return dict(radiusMultiplier=full(1.0), normalBondStiffnessPerUnitArea=full(1e7),
                tangentialBondStiffnessPerUnitArea=full(1e6), maxDistance=full(0.009), maxSigma=full(1e12),
                maxTau=full(1e12), createDistance=full(create), dampingNormalForce=full(0.0),
                dampingTangentialForce=full(0.0), dampingNormalTorque=full(0.0), dampingTangentialTorque=full(0.0))


def test_bonded_clumps_poured_and_bonded():
    """insert/pack pouring bonded 3-sphere clumps (particletemplate/
    multiplespheres ... bonded yes) and single spheres with random
    orientations under gravity: every insertion equals the oracle's bit for
    bit, and at each insertion step the bond model's rule B bonds exactly
    the spheres of one clump closer than createDistance (the neighbours in
    the chain, not its ends) and nothing else, ever"""
    from oracle import TemplateOracle
    from test_insert import CLUMP3, clump_spec

    eng = load_engine()
    spec = dict(packing_spec(), gravity=(0.0, 0.0, -9.81), cohesion="bond")
    s = eng.Simulation(spec, empty_atoms())
    x3, r3 = CLUMP3
    tid = s.add_template_multisphere(999671, 1, 2500.0, x3, r3, ntry=100000, bonded=True)
    to = TemplateOracle(999671, x3, r3, ntry=100000, atom_type=1, density=2500.0, bonded=True)
    ins = dict(clump_spec(), volumefraction_region=0.06, insert_every=400,
               region=((-0.05, -0.05, 0.06), (0.05, 0.05, 0.15)))
    fid = s.add_insert_pack(dict(ins, multisphere=[tid, 0], group_mask=0))
    # the bond model sees the templates' smallest radius (an empty box)
    s.set_bond(bond_tables(), ts_create=-1.0)
    o = oracle_fix(dict(ins, multisphere_oracle=[to, None]))
    s.setup()
    n_before = 0
    clumps = []
    for event in range(3):
        t_ins = int(s.insert_vector(fid)[5])
        assert t_ins == o.next_step()
        s.step(t_ins - s.timestep())
        st, props = s.state(), s.props()
        ref = o.insert(t_ins, st["x"][:n_before], props["radius"][:n_before], props["rmass"][:n_before])
        k = len(ref["radius"])
        new = slice(n_before, n_before + k)
        assert s.n == n_before + k
        np.testing.assert_array_equal(st["x"][new], ref["x"])
        np.testing.assert_array_equal(props["radius"][new], ref["radius"])
        np.testing.assert_array_equal(props["rmass"][new], ref["rmass"])
        nc = int(np.sum(ref["brid"] >= 0)) // 3
        tags = props["tag"][new]
        clumps += [tuple(tags[3 * c:3 * c + 3]) for c in range(nc)]
        print("event %d at step %d: %d clumps and %d spheres inserted" % (event, t_ins, nc, k - 3 * nc))
        assert nc > 50 or event > 0
        # bonds: within each clump, neighbours only
        ti, tj, sh, hi, fl = s.pairs()
        # the reference's slot order puts the bond's flag first (Hertz keeps
        # no normal history)
        bonded = {(a, b) for a, b, h in zip(ti.tolist(), tj.tolist(), hi[:, 0]) if h == 1.0}
        want = set()
        for c in clumps:
            for a, b in ((c[0], c[1]), (c[1], c[2])):
                want |= {(a, b), (b, a)}
        assert bonded == want, (len(bonded), len(want))
        n_before += k
    np.testing.assert_array_equal(s.insert_vector(fid)[:5], o.vector()[:5])
    s.step(300)
    ti, tj, sh, hi, fl = s.pairs()
    bonded = {(a, b) for a, b, h in zip(ti.tolist(), tj.tolist(), hi[:, 0]) if h == 1.0}
    assert bonded == want
    s.close()
