"""Pins the oracle's triangle-mesh wall (CPU only).

* sphere-triangle gap, contact vector and region against an independent
  closest-point-on-triangle construction (Voronoi regions of the vertices,
  edges and face) on random triangles;
* the obtuse-angle bookkeeping of SurfaceMesh::calcObtuseAngleIndex, which
  records an obtuse angle only at node 2 (surface_mesh_I.h:326-336,455-467);
* topology of a flat triangulated square: interior edges and nodes inactive,
  boundary edges active, exactly one active copy of each square corner;
* a sphere sliding and rolling across a flat triangulated floor follows the
  same trajectory as on a primitive plane at the same height (history handed
  between coplanar triangles, no double contact on shared edges or nodes);
* Galilean invariance of a moving mesh: a plate approaching a resting
  sphere at -v gives the same relative motion as the sphere approaching a
  static plate at +v (wall velocity from the node velocities, the mover and
  the half-skin re-neighbour trigger);
* LIGGGHTS-INL's own gold output examples/LIGGGHTS/INL/rolling_resistance_
  cdtnonlinear2 (one sphere on an 886-facet STL floor, hysteretic/nonlinear2
  wall contact with tangential history and cdtnonlinear2 rolling
  resistance, 200,000 steps).
"""
import importlib.util
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import Oracle

GOLD = os.path.join(ROOT, "tests", "golden", "reference_gold")


def meshio():
    spec = importlib.util.spec_from_file_location(
        "meshio_mod", os.path.join(ROOT, "liggghts-inl_amd", "meshio.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def mats(ntypes=2, Y=5e6, nu=0.3, e=0.9, mu=0.5, mur=0.0, etar=0.0):
    w = ntypes + 1
    return dict(ntypes=ntypes, Y=np.full(w, Y), nu=np.full(w, nu), e=np.full((w, w), e),
                mu=np.full((w, w), mu), mur=np.full((w, w), mur), etar=np.full((w, w), etar), krs=0.0)


def one_atom(x, v=(0, 0, 0), w=(0, 0, 0), r=1e-3, rho=2500.0, typ=1):
    m = 4.0 * np.pi / 3.0 * r * r * r * rho
    return dict(x=np.array([x], float), v=np.array([v], float), omega=np.array([w], float),
                radius=np.array([r]), rmass=np.array([m]), type=np.array([typ]), tag=np.array([1]))


def base_spec(**kw):
    s = dict(box_lo=[-0.05] * 3, box_hi=[0.05] * 3, periodic=[0, 0, 0], dt=1e-7, skin=2.5e-4,
             neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off", materials=mats())
    s.update(kw)
    return s


# ---------------------------------------------------------------- geometry

def closest_point(p, a, b, c):
    """Closest point of triangle abc to p by explicit Voronoi regions
    (vertex, edge, face), written from the geometry alone."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a, "v0"
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b, "v1"
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab, "e01"
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c, "v2"
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac, "e02"
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b), "e12"
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den), "face"


def angle_ok(t):
    """No obtuse angle at nodes 0 and 1 (the only case the reference's
    single obtuse index describes completely)."""
    for k in (0, 1):
        e1 = t[(k + 1) % 3] - t[k]
        e2 = t[(k + 2) % 3] - t[k]
        if e1 @ e2 <= 0:
            return False
    return True


def single_triangle_oracle(tri, box=0.05):
    s = base_spec(box_lo=[-box] * 3, box_hi=[box] * 3)
    o = Oracle(s, one_atom([0.04, 0.04, 0.04]))
    o.add_mesh(np.asarray(tri)[None], type=2)
    return o


def test_distance_random_triangles():
    rng = np.random.default_rng(5)
    nt = 0
    checked = {"face": 0, "edge": 0, "corner": 0}
    while nt < 40:
        t = rng.uniform(-0.01, 0.01, size=(3, 3))
        if not angle_ok(t):
            continue
        nt += 1
        o = single_triangle_oracle(t)
        r = 1e-3
        for _ in range(60):
            p = t.mean(0) + rng.normal(scale=0.008, size=3)
            gap, delta, bary, region = o.mesh_distance(0, 0, r, p)
            q, kind = closest_point(p, *t)
            d = np.linalg.norm(q - p)
            assert abs(gap - (d - r)) <= 1e-12 * max(d, 1e-3), (kind, region, gap, d - r)
            assert np.abs(delta - (q - p)).max() <= 1e-12, (kind, region)
            # the barycentric weights reproduce the closest point
            assert np.abs(bary @ t - q).max() <= 1e-11
            if region == 7:
                checked["face"] += 1
                assert kind == "face"
            elif region in (3, 5, 6):
                checked["edge"] += 1
            else:
                checked["corner"] += 1
    assert min(checked.values()) > 50, checked


def test_obtuse_index_only_node2():
    # obtuse at node 2: recorded; obtuse at node 0: not recorded
    t2 = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.005, 0.001, 0.0]])
    t0 = t2[[2, 0, 1]]
    for t, want in ((t2, 2), (t0, -1)):
        o = single_triangle_oracle(t)
        assert o.mesh_topology()["obtuse"][0] == want
    # with the angle recorded at node 2 the corner region near node 2 still
    # resolves to the closest edge point
    o = single_triangle_oracle(t2)
    p = np.array([0.004, 0.0025, 0.0005])
    gap, delta, bary, region = o.mesh_distance(0, 0, 1e-3, p)
    q, kind = closest_point(p, *t2)
    assert kind.startswith("e") and region == 4
    assert abs(gap - (np.linalg.norm(q - p) - 1e-3)) < 1e-15


def flat_grid(n=2, L=0.01, z=0.0, x0=0.0, y0=0.0, flip=False):
    """n x n squares, each split into two triangles along a diagonal."""
    tris = []
    h = L / n
    for i in range(n):
        for j in range(n):
            a = np.array([x0 + i * h, y0 + j * h, z])
            b = a + [h, 0, 0]
            c = a + [h, h, 0]
            d = a + [0, h, 0]
            if flip and (i + j) % 2:
                tris += [[a, b, d], [b, c, d]]
            else:
                tris += [[a, b, c], [a, c, d]]
    return np.asarray(tris)


def test_flat_grid_topology():
    n, L = 3, 0.01
    tris = flat_grid(n, L, flip=True)
    o = Oracle(base_spec(), one_atom([0.04, 0.04, 0.04]))
    o.add_mesh(tris, type=2)
    top = o.mesh_topology()
    ea, ca = top["edge_active"], top["corner_active"]
    for t in range(len(tris)):
        for k in range(3):
            a, b = tris[t][k], tris[t][(k + 1) % 3]
            boundary = any(np.all(np.isclose([a[d], b[d]], v)) for d in (0, 1) for v in (0.0, L))
            assert ea[t, k] == (1 if boundary else 0), (t, k, a, b)
    # active corners: exactly the four square corners, each once
    act = {}
    for t in range(len(tris)):
        for k in range(3):
            if ca[t, k]:
                key = tuple(np.round(tris[t][k] / L, 6))
                act[key] = act.get(key, 0) + 1
    assert act == {(0.0, 0.0, 0.0): 1, (1.0, 0.0, 0.0): 1, (0.0, 1.0, 0.0): 1, (1.0, 1.0, 0.0): 1}
    assert (top["nneigh"] >= 1).all() and (top["nneigh"] <= 3).all()


def folded_pair():
    """Two triangles sharing the edge (0,0,0)-(0.01,0,0) at a right angle."""
    a, b = np.array([0.0, 0, 0]), np.array([0.01, 0, 0])
    return np.array([[a, b, [0.005, 0.01, 0.0]], [b, a, [0.005, 0.0, 0.01]]])


def test_folded_edge_active_once():
    o = Oracle(base_spec(), one_atom([0.04, 0.04, 0.04]))
    o.add_mesh(folded_pair(), type=2)
    ea = o.mesh_topology()["edge_active"]
    # the shared edge is edge 0 of both; the higher id keeps it active
    assert ea[0, 0] == 0 and ea[1, 0] == 1


# ------------------------------------------------------------ dynamics

def floor_run(use_mesh, v, w, steps, rolling="off", mur=0.0, etar=0.0):
    s = base_spec(rolling=rolling, materials=mats(mur=mur, etar=etar, mu=0.3), gravity=(0.0, 0.0, -9.81),
                  dt=2e-7)
    r = 1e-3
    # start slightly overlapping the floor at z = 0
    a = one_atom([0.0031, 0.0043, r - 2e-6], v=v, w=w, r=r)
    if use_mesh:
        s["meshes"] = [dict(nodes=flat_grid(12, 0.018, flip=True, x0=-0.002, y0=-0.002), type=2)]
    else:
        s["planes"] = [(2, 0.0, 2)]
    o = Oracle(s, a)
    o.setup()
    out = []
    for _ in range(steps // 500):
        o.run(500)
        st = o.state()
        out.append(np.concatenate([st["x"][0], st["v"][0], st["omega"][0], st["f"][0], st["torque"][0]]))
    return np.array(out), o


@pytest.mark.parametrize("rolling", ["off", "epsd"])
def test_flat_mesh_equals_plane(rolling):
    kw = dict(v=(0.5, 0.3, 0.0), w=(0.0, 0.0, 3.0), steps=40000, rolling=rolling,
              mur=0.2 if rolling != "off" else 0.0, etar=0.3)
    pm, om = floor_run(True, **kw)
    pp, _ = floor_run(False, **kw)
    # the sphere crossed several triangles and shared edges
    assert np.linalg.norm(pm[-1, :2] - np.array([0.0031, 0.0043])) > 0.004
    scale = np.abs(pp).max(0) + 1e-300
    err = (np.abs(pm - pp) / scale).max()
    assert err <= 1e-13, err
    tg, tr, hi = om.mesh_contacts()
    assert len(tg) == 1


def plate_run(moving, steps=3000):
    vz = 0.4
    r = 1e-3
    s = base_spec(dt=1e-7, skin=1e-4, materials=mats(mu=0.3))
    gap0 = 3e-5
    # a plate above the sphere, normal pointing down
    sq = flat_grid(1, 0.01, z=r + gap0, x0=-0.005, y0=-0.005)[:, ::-1]
    if moving:
        s["meshes"] = [dict(nodes=sq, type=2, vel=(0.0, 0.0, -vz))]
        a = one_atom([0.0007, -0.0004, 0.0], v=(0.0, 0.0, 0.0), w=(5.0, 0, 0), r=r)
    else:
        s["meshes"] = [dict(nodes=sq, type=2)]
        a = one_atom([0.0007, -0.0004, 0.0], v=(0.0, 0.0, vz), w=(5.0, 0, 0), r=r)
    o = Oracle(s, a)
    o.setup()
    rel = []
    for k in range(steps // 100):
        o.run(100)
        st = o.state()
        t = o.time()
        plate_z = r + gap0 - (vz * t if moving else 0.0)
        vplate = -vz if moving else 0.0
        rel.append([plate_z - st["x"][0, 2], vplate - st["v"][0, 2], st["v"][0, 0], st["v"][0, 1],
                    st["omega"][0, 0], st["f"][0, 2]])
    return np.array(rel)


def test_moving_plate_galilean():
    a = plate_run(True)
    b = plate_run(False)
    # the plate reached the sphere and pushed it (contact force present)
    assert np.abs(b[:, 5]).max() > 0
    scale = np.abs(b).max(0) + 1e-300
    err = (np.abs(a - b) / scale).max(0)
    assert err.max() < 1e-9, err


# ------------------------------------------------------------ gold replay

def sub(x):
    return float("%.15g" % x)


def rolling_resistance_replay():
    tris = meshio().read_stl(os.path.join(GOLD, "rolling_resistance_bottom.stl"))
    w = 4
    E, nu, e = 1e7, 0.3, 0.3
    full = lambda v: np.full((w, w), v)  # noqa: E731
    mu = np.zeros((w, w))
    mu[1:, 1:] = [[0.9, 0.9, 0.0], [0.9, 0.9, 0.9], [0.0, 0.9, 0.9]]
    mur = np.zeros((w, w))
    mur[1:, 1:] = [[0.9, 0.2, 0.9], [0.2, 0.9, 0.9], [0.9, 0.9, 0.9]]
    mats3 = dict(ntypes=3, Y=np.full(w, E), nu=np.full(w, nu), e=full(e), mu=mu, mur=mur,
                 etar=full(0.0), krs=0.0,
                 hysteretic=dict(LoadingStiffness=full(5e4), UnloadingStiffness=full(2.5),
                                 coefficientAdhesionStiffness=full(0.0),
                                 coefficientPlasticityDepth=full(0.1), pullOffForce=full(0.0),
                                 alphaCustom=full(20.0), cinCustom=full(1e-7), aoneCustom=full(16e8),
                                 atwoCustom=full(4e4), athreeCustom=full(11.0), kcinCustom=full(1e-4)))
    # cohesion easo/capillary/viscous registers maxSeparationDistanceRatio*1.1
    # as the contact distance factor (cohesion_model_easo_capillary_viscous.h:181)
    spec = dict(box_lo=[-0.015, -0.005, -0.001], box_hi=[0.035, 0.005, 0.01], periodic=[0, 0, 0],
                dt=1e-7, skin=0.001, cdf=1.1 * 1.1, neigh_every=1, neigh_delay=0, neigh_check=True,
                rolling="cdtnonlinear2", normal="hysteretic/nonlinear2",
                wall_normal="hysteretic/nonlinear2", materials=mats3, fix_order=[])
    r = 0.5 * 0.003
    m = 4.0 * np.pi / 3.0 * r * r * r * 320.0
    atoms = dict(x=np.array([[0.0, 0.0015, 0.0015]]), v=np.zeros((1, 3)), omega=np.zeros((1, 3)),
                 radius=np.array([r]), rmass=np.array([m]), type=np.array([1]), tag=np.array([1]))
    o = Oracle(spec, atoms)
    o.set_group(1, [1])
    sf1 = o.add_setforce(1, [None, 0.0, None])
    o.add_mesh(tris, type=2)
    o.add_plane(1, -0.005, 3)
    o.add_plane(1, 0.005, 3)
    grav = o.add_gravity(0.0, 0.0, -9.81)   # 9.81 * (0,0,-1)/|(0,0,-1)|
    rows = []

    def record():
        st = o.state()
        rows.append([1, o.time(), *st["x"][0], *st["v"][0], *st["omega"][0], *st["torque"][0]])

    o.setup()
    for _ in range(100):
        o.run(1000)
        record()
    o.cleanup()
    # unfix grav; fix grav ... vector ${i_grav_x} ${i_grav_y} ${i_grav_z}
    ang = 10.0
    gx = sub(1.0 * np.sin(ang / 180.0 * 3.1415926))
    gy = sub(0.0)
    gz = sub(-1.0 * np.cos(ang / 180.0 * 3.1415926))
    ln = np.sqrt(gx * gx + gy * gy + gz * gz)
    o.unfix(grav)
    o.add_gravity(9.81 * (gx / ln), 9.81 * (gy / ln), 9.81 * (gz / ln))
    o.unfix(sf1)
    o.add_setforce(1, [None, 0.0, None])
    o.setup()
    for _ in range(100):
        o.run(1000)
        record()
    o.cleanup()
    return np.array(rows), o


def test_stl_reader():
    tris = meshio().read_stl(os.path.join(GOLD, "rolling_resistance_bottom.stl"))
    assert tris.shape == (886, 3, 3)
    assert tris[0, 0].tolist() == [-4.500013e-03, -4.133915e-03, 0.0]
    assert tris[0, 2].tolist() == [-5.0e-03, -5.0e-03, 0.0]


def test_rolling_resistance_gold():
    g = np.loadtxt(os.path.join(GOLD, "rolling_resistance_cdtnonlinear2_out_loading.csv"),
                   delimiter=",", skiprows=1)
    got, o = rolling_resistance_replay()
    assert got.shape == g.shape
    # every column, scaled by its largest gold magnitude; the gold prints 15
    # significant digits
    scale = np.abs(g).max(0)
    scale[scale == 0] = 1.0
    err = (np.abs(got - g) / scale).max(0)
    # measured: <= 3.4e-15 in every column
    assert err[1] < 1e-15, err                       # time
    assert err[2:5].max() < 1e-13, err               # position
    assert err[5:].max() < 1e-13, err                # velocity, spin, torque
    # the sphere stayed on the mesh: one mesh contact with history
    tg, tr, hi = o.mesh_contacts()
    assert len(tg) == 1 and np.abs(hi).max() > 0


def test_mesh_stress_closed_form():
    """fix mesh/surface/stress: spheres resting at their static Hertz
    overlap on a flat triangulated floor (one of them on an edge shared by
    two triangles, one on a node) load the mesh with their weights.  The
    total force is -sum m g along z and the torque about p_ref is
    sum (x_c - p_ref) x (0, 0, -m g), x_c the contact point under each
    sphere (mesh_module_stress.cpp:318-344)."""
    g = 9.81
    r = 1e-3
    L = 0.02
    tris = flat_grid(4, L, x0=-0.01, y0=-0.01)
    # Hertz static overlap against the floor: m g = 4/3 E* sqrt(r) d^1.5,
    # E* of two type pairs (1, 2) with equal Y and nu
    Y, nu = 5e6, 0.3
    Es = 1.0 / (2 * (1 - nu * nu) / Y)
    pos = [(0.0031, -0.0042), (0.0, 0.0025), (-0.005, -0.005), (0.004, 0.0066)]   # face, edge, node, face
    xs, rs = [], []
    for (a, b), rr in zip(pos, (r, 0.8 * r, 1.2 * r, 0.9 * r)):
        m = 4.0 / 3.0 * np.pi * rr ** 3 * 2500.0
        d = (m * g / (4.0 / 3.0 * Es * np.sqrt(rr))) ** (2.0 / 3.0)
        xs.append([a, b, rr - d])
        rs.append(rr)
    n = len(xs)
    rs = np.array(rs)
    ms = 4.0 / 3.0 * np.pi * rs ** 3 * 2500.0
    atoms = dict(x=np.array(xs), v=np.zeros((n, 3)), omega=np.zeros((n, 3)), radius=rs, rmass=ms,
                 type=np.ones(n, int), tag=np.arange(1, n + 1))
    spec = base_spec(gravity=(0.0, 0.0, -g), dt=1e-7, meshes=[dict(nodes=tris, type=2)])
    o = Oracle(spec, atoms)
    pref = np.array([0.002, -0.003, 0.01])
    o.mesh_stress(0, pref)
    o.setup()
    o.run(20)
    sv = o.mesh_stress_vector(0)
    st = o.state()
    W = (ms * g).sum()
    assert sv[2] == pytest.approx(-W, rel=1e-9)
    assert abs(sv[0]) + abs(sv[1]) < 1e-12 * W
    cp = np.stack([st["x"][:, 0], st["x"][:, 1], np.zeros(n)], 1) - pref
    torque = np.cross(cp, np.stack([np.zeros(n), np.zeros(n), -ms * g], 1)).sum(0)
    assert np.allclose(sv[3:6], torque, rtol=1e-8, atol=1e-12 * W * L)
    assert np.array_equal(sv[6:9], pref)
    # the force on the mesh is minus the force the mesh puts on the spheres
    assert np.allclose(sv[:3], -o.mesh_force(0), rtol=1e-14, atol=0)
