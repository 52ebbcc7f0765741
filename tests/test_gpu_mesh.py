"""Mesh walls on the MI355X against the oracle and the reference's gold
(FP64 engine, MI355X only).

* the reference's rolling_resistance_cdtnonlinear2 example replayed on the
  device: one sphere on the 886-facet STL floor, hysteretic/nonlinear2 wall
  contact + tangential history + cdtnonlinear2 rolling resistance, gravity
  tilted by 10 degrees after 100,000 steps, 200,000 steps in all;
* a hopper (cone + cylinder, ~13,000 triangles, 3,000 spheres touching the
  wall) from an identical state: forces within 1e-12 of each atom's force
  scale, the same (atom, triangle) contact slots with the same history;
  then a short horizon;
* a moving plate (fix move/mesh linear) pressing on a sphere, step by step
  against the oracle.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, load_engine
from oracle import Oracle
from test_gpu_parity import check_forces, check_pairs, snapshot_atoms
from test_oracle_mesh import plate_run, rolling_resistance_replay  # noqa: F401

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "reference_gold")


def sub(x):
    return float("%.15g" % x)


def test_rolling_resistance_gold_on_device():
    eng = load_engine()
    tris = eng.meshio.read_stl(os.path.join(GOLD, "rolling_resistance_bottom.stl"))
    w = 4
    full = lambda v: np.full((w, w), v)  # noqa: E731
    mu = np.zeros((w, w))
    mu[1:, 1:] = [[0.9, 0.9, 0.0], [0.9, 0.9, 0.9], [0.0, 0.9, 0.9]]
    mur = np.zeros((w, w))
    mur[1:, 1:] = [[0.9, 0.2, 0.9], [0.2, 0.9, 0.9], [0.9, 0.9, 0.9]]
    mats = dict(ntypes=3, Y=np.full(w, 1e7), nu=np.full(w, 0.3), e=full(0.3), mu=mu, mur=mur,
                etar=full(0.0), krs=0.0,
                hysteretic=dict(LoadingStiffness=full(5e4), UnloadingStiffness=full(2.5),
                                coefficientAdhesionStiffness=full(0.0),
                                coefficientPlasticityDepth=full(0.1), pullOffForce=full(0.0),
                                alphaCustom=full(20.0), cinCustom=full(1e-7), aoneCustom=full(16e8),
                                atwoCustom=full(4e4), athreeCustom=full(11.0), kcinCustom=full(1e-4)))
    spec = dict(box_lo=[-0.015, -0.005, -0.001], box_hi=[0.035, 0.005, 0.01], periodic=[0, 0, 0],
                dt=1e-7, skin=0.001, cdf=1.1 * 1.1, neigh_every=1, neigh_delay=0, neigh_check=True,
                rolling="cdtnonlinear2", normal="hysteretic/nonlinear2", wall_normal="hysteretic/nonlinear2",
                materials=mats, meshes=[dict(nodes=tris, type=2)],
                planes=[(1, -0.005, 3), (1, 0.005, 3)], gravity=(0.0, 0.0, -9.81))
    r = 0.5 * 0.003
    m = 4.0 * np.pi / 3.0 * r * r * r * 320.0
    atoms = dict(x=np.array([[0.0, 0.0015, 0.0015]]), v=np.zeros((1, 3)), omega=np.zeros((1, 3)),
                 radius=np.array([r]), rmass=np.array([m]), type=np.array([1]), tag=np.array([1]))
    s = eng.Simulation(spec, atoms)
    rows = []

    def record():
        st = s.state()
        rows.append([1, s.timestep() * 1e-7, *st["x"][0], *st["v"][0], *st["omega"][0], *st["torque"][0]])

    s.setup()
    for _ in range(100):
        s.step(1000)
        record()
    gx = sub(1.0 * np.sin(10.0 / 180.0 * 3.1415926))
    gz = sub(-1.0 * np.cos(10.0 / 180.0 * 3.1415926))
    ln = np.sqrt(gx * gx + 0.0 + gz * gz)
    s.set_gravity(9.81 * (gx / ln), 9.81 * (0.0 / ln), 9.81 * (gz / ln))
    s.setup()
    for _ in range(100):
        s.step(1000)
        record()
    got = np.array(rows)
    g = np.loadtxt(os.path.join(GOLD, "rolling_resistance_cdtnonlinear2_out_loading.csv"),
                   delimiter=",", skiprows=1)
    scale = np.abs(g).max(0)
    scale[scale == 0] = 1.0
    err = (np.abs(got - g) / scale).max(0)
    print("column-scaled error vs gold:", err)
    # measured on MI355X: <= 2e-14 positions, <= 8.7e-14 velocity, spin and
    # torque (pow/exp of the device math library in the hysteretic law may
    # differ from glibc by an ulp); the gold prints 15 digits
    assert err[1] < 1e-15, err
    assert err[2:5].max() < 1e-12, err
    assert err[5:].max() < 1e-12, err
    tg, tr, sl, hi = s.mesh_contacts(0)
    assert len(tg) == 1 and tr[0] == 845


def wall_spheres(spec, atoms, n_wall=800, seed=3):
    """Add spheres resting on random points of the hopper's triangles (faces,
    edges and corners alike), overlapping the wall by 0-30 um, not each
    other."""
    rng = np.random.default_rng(seed)
    tris = spec["meshes"][0]["nodes"]
    placed = np.array(atoms["x"])
    rad = np.array(atoms["radius"])
    out_x, out_r = [], []
    tries = 0
    while len(out_x) < n_wall and tries < 50 * n_wall:
        tries += 1
        t = tris[rng.integers(len(tris))]
        if t[:, 2].max() > 0.05:
            continue
        b = rng.dirichlet([1.0, 1.0, 1.0])
        kind = rng.integers(3)
        if kind == 1:        # on an edge
            b[rng.integers(3)] = 0.0
            b /= b.sum()
        elif kind == 2:      # at a node
            b = np.eye(3)[rng.integers(3)]
        p = b @ t
        nrm = np.cross(t[1] - t[0], t[2] - t[0])
        nrm /= np.linalg.norm(nrm)
        if nrm[:2] @ p[:2] > 0:   # point into the hopper
            nrm = -nrm
        ri = rng.uniform(0.8e-3, 1.2e-3)
        c = p + nrm * (ri - rng.uniform(0.0, 3e-5))
        allx = np.concatenate([placed] + ([np.array(out_x)] if out_x else []))
        allr = np.concatenate([rad] + ([np.array(out_r)] if out_r else []))
        if (np.linalg.norm(allx - c, axis=1) < allr + ri + 1e-5).any():
            continue
        out_x.append(c)
        out_r.append(ri)
    a = dict(atoms)
    a["x"] = np.concatenate([atoms["x"], np.array(out_x)])
    a["radius"] = np.concatenate([atoms["radius"], np.array(out_r)])
    a["rmass"] = 4.0 / 3.0 * np.pi * a["radius"] ** 3 * 2500.0
    m = len(a["radius"])
    a["v"] = np.zeros((m, 3))
    a["omega"] = np.zeros((m, 3))
    a["type"] = np.ones(m, np.int32)
    a["tag"] = np.arange(1, m + 1, dtype=np.int32)
    return a


def hopper_state(steps=200):
    eng = load_engine()
    spec, atoms = eng.workloads.hopper(n=2000)
    atoms = wall_spheres(spec, atoms)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(steps)
    return spec, atoms, o


def sim_from_oracle_mesh(spec, atoms, o, precision="fp64"):
    eng = load_engine()
    a = snapshot_atoms(atoms, o)
    s = eng.Simulation(spec, a, precision=precision)
    ti, tj, sh, hi, fl = o.pairs()
    k = fl != 0
    s.set_history(ti[k], tj[k], hi[k])
    s.set_wall_history(o.wall_history())
    for m in range(o.nmesh):
        tg, tr, sl, mh = o.mesh_contacts(m, slots=True)
        s.set_mesh_contacts(m, tg, tr, sl, mh)
    s.setup()
    o.rebuild()
    o.compute_forces(0)
    return s


def mesh_dict(tg, tr, sl, hi):
    return {(a, b): (c, h) for a, b, c, h in zip(tg.tolist(), tr.tolist(), sl.tolist(), hi)}


def check_mesh_contacts(s, o, tol):
    G = mesh_dict(*s.mesh_contacts(0))
    O = mesh_dict(*o.mesh_contacts(0, slots=True))
    assert set(G) == set(O)
    assert all(G[k][0] == O[k][0] for k in O), "slot layout differs"
    scale = max(max((np.abs(v[1]).max() for v in O.values()), default=0.0), 1e-300)
    err = max((np.abs(G[k][1] - O[k][1]).max() for k in O), default=0.0)
    assert err <= tol * scale, err / scale
    return len(O)


def test_hopper_identical_state():
    spec, atoms, o = hopper_state()
    s = sim_from_oracle_mesh(spec, atoms, o)
    nc = check_mesh_contacts(s, o, 1e-12)
    assert nc > 100
    check_pairs(s, o)
    check_forces(s, o, 1e-12)


def test_hopper_short_horizon():
    spec, atoms, o = hopper_state()
    s = sim_from_oracle_mesh(spec, atoms, o)
    s.step(40)
    o.run(40)
    g, r = s.state(), o.state()
    assert np.abs(g["x"] - r["x"]).max() <= 1e-12 * 0.5
    check_forces(s, o, 1e-9)
    check_mesh_contacts(s, o, 1e-9)


def test_hopper_mixed_precision():
    """The hopper in its stated mixed precision (FP32 pair law in the
    contact queue, FP64 mesh walls and state): forces within 1e-5 of the
    FP64 oracle from an identical state, and after 40 steps."""
    spec, atoms, o = hopper_state()
    s = sim_from_oracle_mesh(spec, atoms, o, precision="mixed")
    check_mesh_contacts(s, o, 1e-12)
    check_forces(s, o, 1e-5)
    s.step(40)
    o.run(40)
    g, r = s.state(), o.state()
    assert np.abs(g["x"] - r["x"]).max() <= 1e-11
    check_forces(s, o, 1e-5)
    check_mesh_contacts(s, o, 1e-5)


def test_moving_plate_device():
    """fix move/mesh linear: a plate moving down onto a sphere, device vs
    oracle over 3,000 steps (mesh mover, node velocities in the wall contact,
    half-skin re-neighbour trigger)."""
    eng = load_engine()
    from test_oracle_mesh import base_spec, flat_grid, mats, one_atom

    r = 1e-3
    sq = flat_grid(1, 0.01, z=r + 3e-5, x0=-0.005, y0=-0.005)[:, ::-1]
    spec = base_spec(dt=1e-7, skin=1e-4, materials=mats(mu=0.3),
                     meshes=[dict(nodes=sq, type=2, vel=(0.0, 0.0, -0.4))])
    atoms = one_atom([0.0007, -0.0004, 0.0], w=(5.0, 0, 0), r=r)
    o = Oracle(spec, atoms)
    s = eng.Simulation(spec, atoms)
    o.setup()
    s.setup()
    for _ in range(30):
        o.run(100)
        s.step(100)
        g, q = s.state(), o.state()
        assert np.abs(g["x"] - q["x"]).max() <= 1e-15
        sc = np.abs(q["f"]).max() + 1e-300
        assert np.abs(g["f"] - q["f"]).max() <= 1e-12 * sc
    assert np.abs(q["f"][0, 2]) > 0


def test_hopper_graph_stepping_bitwise():
    """A static mesh steps through recorded graphs (two rebuild periods per
    graph, the mesh candidate rows rebuilt inside it): state and mesh
    contact slots equal, bit for bit, to plain stepping over 95 steps."""
    eng = load_engine()
    spec, atoms = eng.workloads.hopper(n=2000)
    atoms = wall_spheres(spec, atoms)
    out = []
    for graph in (False, True):
        s = eng.Simulation(dict(spec, graph=graph), atoms)
        s.setup()
        s.step(95)
        st = s.state()
        out.append((st, mesh_dict(*s.mesh_contacts(0))))
    (a, ma), (b, mb) = out
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(a[k], b[k]), k
    assert set(ma) == set(mb) and len(ma) > 0
    assert all(ma[k][0] == mb[k][0] and np.array_equal(ma[k][1], mb[k][1]) for k in ma)


def test_hopper_fused_integration_bitwise():
    """With mesh walls the last mesh pass integrates; a setforce on an empty
    group changes no force but keeps the integrator a kernel of its own.
    Both runs, over rebuilds, give the same state and mesh slots bit for bit."""
    eng = load_engine()
    spec, atoms = eng.workloads.hopper(n=2000)
    atoms = wall_spheres(spec, atoms)
    out = []
    for split in (False, True):
        s = eng.Simulation(spec, atoms)
        if split:
            s.set_group(3, np.zeros(len(atoms["radius"]), dtype=np.int32))
            s.add_setforce(3, [0.0, 0.0, 0.0])
        s.setup()
        s.step(37)
        s.step(26)
        out.append((s.state(), mesh_dict(*s.mesh_contacts(0))))
    (a, ma), (b, mb) = out
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(a[k], b[k]), k
    assert set(ma) == set(mb) and len(ma) > 0
    assert all(ma[k][0] == mb[k][0] and np.array_equal(ma[k][1], mb[k][1]) for k in ma)


def test_autotune_leaves_mesh_state():
    """The autotuner's trial passes create and clean mesh contact slots; it
    restores them with the pair history and forces, so the steps after it
    are bitwise those of a run given the same contact shape by hand."""
    eng = load_engine()
    spec, atoms = eng.workloads.hopper(n=2000)
    atoms = wall_spheres(spec, atoms)
    out = []
    shape = None
    for tune in (True, False):
        s = eng.Simulation(spec, atoms)
        s.setup()
        if tune:
            lanes, block, _ = s.autotune(reps=2)
            shape = (lanes, block)
        else:
            s.set_contact_shape(*shape)
        s.step(30)
        out.append((s.state(), mesh_dict(*s.mesh_contacts(0))))
        s.close()
    (a, ma), (b, mb) = out
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(a[k], b[k]), k
    assert set(ma) == set(mb) and len(ma) > 0
    assert all(ma[k][0] == mb[k][0] and np.array_equal(ma[k][1], mb[k][1]) for k in ma)


def test_hopper_mesh_stress():
    """fix mesh/surface/stress on the hopper: the total force the spheres
    put on the mesh and its torque about a reference point, from an
    identical state and after 40 steps, against the oracle (the sums run in
    another order: 1e-10 of their magnitude)."""
    eng = load_engine()
    spec, atoms, o = hopper_state()
    pref = np.array([0.001, -0.002, 0.03])
    o.mesh_stress(0, pref)
    a = snapshot_atoms(atoms, o)
    s = eng.Simulation(spec, a)
    ti, tj, sh, hi, fl = o.pairs()
    k = fl != 0
    s.set_history(ti[k], tj[k], hi[k])
    s.set_wall_history(o.wall_history())
    tg, tr, sl, mh = o.mesh_contacts(0, slots=True)
    s.set_mesh_contacts(0, tg, tr, sl, mh)
    s.mesh_stress(0, pref)
    s.setup()
    o.rebuild()
    o.compute_forces(0)
    for _ in range(2):
        g, r = s.mesh_stress_vector(0), o.mesh_stress_vector(0)
        assert np.linalg.norm(r[:3]) > 0
        assert np.abs(g[:3] - r[:3]).max() <= 1e-10 * np.linalg.norm(r[:3])
        assert np.abs(g[3:6] - r[3:6]).max() <= 1e-10 * np.linalg.norm(r[3:6])
        assert np.array_equal(g[6:], r[6:])
        s.step(40)
        o.run(40)


def freefall_massflow(vec_side, steps, dt=1e-5):
    """Spheres falling freely (no contacts, no walls) onto a horizontal
    square face mesh at z0 = 50 mm (two triangles split along x = y), counted
    by fix massflow/mesh."""
    eng = load_engine()
    from test_oracle_mesh import base_spec, flat_grid, mats

    rng = np.random.default_rng(5)
    g = np.arange(7) * 0.004 + 0.0075
    xy = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    xy = xy[np.abs(xy[:, 0] - xy[:, 1]) > 1.5e-3]          # off the shared diagonal edge
    n = len(xy)
    r = 0.5e-3
    z = 0.0505 + rng.permutation(n) * (0.035 / n) + rng.uniform(0, 1e-4, n)
    x = np.column_stack([xy, z])
    m = 4.0 / 3.0 * np.pi * r ** 3 * 2500.0
    atoms = dict(x=x, v=np.zeros((n, 3)), omega=np.zeros((n, 3)), radius=np.full(n, r), rmass=np.full(n, m),
                 type=np.ones(n, int), tag=np.arange(1, n + 1))
    spec = base_spec(box_lo=[0, 0, 0], box_hi=[0.04, 0.04, 0.1], dt=dt, skin=2.5e-4, neigh_every=1,
                     neigh_check=False, gravity=(0.0, 0.0, -9.81), materials=mats(ntypes=1))
    s = eng.Simulation(spec, atoms)
    s.add_massflow(flat_grid(1, 0.03, z=0.05, x0=0.005, y0=0.005), vec_side, once=True)
    s.setup()
    return s, atoms, m


def test_massflow_free_fall_closed_form():
    """fix massflow/mesh counts exactly the spheres whose free-fall height
    z0 - g t^2 / 2 has passed the face (velocity Verlet is exact for a
    constant force), and its rates are the count differences over the time
    between two reads; with vec_side pointing up nothing is counted."""
    dt = 1e-5
    s, atoms, m = freefall_massflow((0.0, 0.0, -1.0), 0, dt)
    z0 = atoms["x"][:, 2]
    out = []
    # reads after 3000 and 7000 steps: 5 and 29 of the 42 spheres crossed
    for nsteps in (3000, 4000):
        s.step(nsteps)
        out.append(s.massflow_vector())
    t1, t2 = 3000 * dt, 7000 * dt
    for t, v in ((t1, out[0]), (t2, out[1])):
        zt = z0 - 0.5 * 9.81 * t * t
        assert np.abs(zt - 0.05).min() > 1e-6        # no sphere at the face at the read
        crossed = (zt < 0.05).sum()
        assert v[1] == crossed
        assert v[0] == pytest.approx(crossed * m, rel=1e-14)
    st = s.state()
    assert np.allclose(st["x"][:, 2], z0 - 0.5 * 9.81 * t2 * t2, rtol=0, atol=1e-12)
    assert 0 < out[0][1] < out[1][1] < len(z0)
    # rates: the second read over the 4000 steps since the first
    assert out[1][3] == pytest.approx((out[1][1] - out[0][1]) / (4000 * dt), rel=1e-12)
    assert out[1][2] == pytest.approx((out[1][0] - out[0][0]) / (4000 * dt), rel=1e-12)
    s.close()
    s, atoms, m = freefall_massflow((0.0, 0.0, 1.0), 0, dt)
    s.step(10000)
    assert s.massflow_vector()[1] == 0
    s.close()
