"""fix insert/stream: the oracle pinned by closed forms and by the geometry
the reference fixes, then the engine's host loop (insert_pack.cpp through
tests/insert_host_driver.cpp, brute-force device hooks) against the oracle
bit for bit.

The workload is the insert_stream tutorial (examples/LIGGGHTS/
Tutorials_public/insert_stream/in.insert_stream): its face.stl (86
triangles in z = 443, scale 0.005), two templates r = 0.015 / 0.025 at
2500 kg/m^3 mixed 0.3 / 0.7 by mass, vel constant 0 -0.5 -2, particlerate
1000, nparticles 5000, extrude_length 0.6, dt 1e-5, box (-1, 1)^2 x
(-0.05, 2.85).  Pins:

* insert_every = (int)((0.6 + 1e-14) / (dt |v_n|)) = 30000 with v_n = -2 z
  (the face normal +z turned toward the velocity), 300 bodies per insertion;
* every inserted sphere lies on the face's footprint (barycentric test
  against the triangles, written here with numpy), at a depth r..0.6 + r
  behind the face, moving with v_n and no spin; its release step is the
  steps it needs to reach the face at |v_n| (floor of depth / (|v_n| dt));
* the first insertion's count is (int)(300 + u) with u the fix stream's
  first draw (Park-Miller, pinned in test_insert.py);
* the face is sampled by area: over many points without overlap check each
  triangle gets its area's share within binomial noise, and the points'
  mean is the face's area centroid;
* the insertions stop at nparticles exactly, the last one short;
* residents whose projection misses the face (beside it, or over its
  central hole) are not tested, one over the face is (FixInsertStream::
  is_nearby).
"""
import os

import numpy as np
import pytest

from conftest import ROOT, load_engine
from test_insert import HostLoop, same

DT = 1e-5
BOX = ((-1.0, -1.0, -0.05), (1.0, 1.0, 2.85))
FACE_Z = 443 * 0.005


def face_tris():
    eng = load_engine()
    t = eng.meshio.read_stl(os.path.join(ROOT, "tests", "golden", "reference_data", "insert_stream_face.stl"))
    return np.asarray(t, dtype=np.float64).reshape(-1, 3, 3) * 0.005


def tutorial(**kw):
    d = dict(seed=32452867, dist_seed=32452843,
             templates=[dict(type=1, radius=0.015, density=2500.0), dict(type=1, radius=0.025, density=2500.0)],
             weights=[0.3, 0.7], nparticles=5000, vel=("constant", (0.0, -0.5, -2.0)), particlerate=1000.0,
             overlapcheck=True, insertion_face=face_tris().reshape(-1, 9), extrude_length=0.6)
    d.update(kw)
    return d


def oracle(spec, box=BOX, periodic=(0, 0, 0), ntimestep=0):
    from oracle import InsertStreamOracle

    return InsertStreamOracle(spec, box[0], box[1], periodic, DT, ntimestep)


def on_footprint(p, tris):
    """the (x, y) of each point inside some triangle of the face (z
    ignored: the face is horizontal)"""
    a, b, c = tris[:, 0, :2], tris[:, 1, :2], tris[:, 2, :2]
    q = p[:, None, :2]
    v0, v1, v2 = c - a, b - a, q - a
    d00 = (v0 * v0).sum(-1)
    d01 = (v0 * v1).sum(-1)
    d11 = (v1 * v1).sum(-1)
    d02 = (v0 * v2).sum(-1)
    d12 = (v1 * v2).sum(-1)
    den = d00 * d11 - d01 * d01
    u = (d11 * d02 - d01 * d12) / den
    v = (d00 * d12 - d01 * d02) / den
    return ((u > -1e-9) & (v > -1e-9) & (u + v < 1 + 1e-9)).any(-1)


def advance(atoms, step):
    """residents for the next insertion: each atom the stream made rides
    its release data (x_ins + v_int (t - i_step) dt) up to its release step,
    then moves on with the release velocity; no forces"""
    rel = atoms["release"]
    x = np.empty_like(atoms["x"])
    for k in range(len(x)):
        i, r = int(rel[k, 3]), int(rel[k, 4])
        if i == 0:                       # not made by the stream: stays put
            x[k] = atoms["x"][k]
        elif step <= r:
            x[k] = rel[k, :3] + rel[k, 5:8] * ((step - i) * DT)
        else:
            x[k] = rel[k, :3] + rel[k, 5:8] * ((r - i) * DT) + rel[k, 8:11] * ((step - r) * DT)
    return x


def run_events(fix, nevents, extra=None):
    """nevents insertions, the residents advanced between them; returns the
    per-event outputs and the atoms"""
    atoms = dict(x=np.zeros((0, 3)), radius=np.zeros(0), rmass=np.zeros(0), release=np.zeros((0, 14)))
    if extra is not None:
        atoms = {k: np.concatenate([extra[k], atoms[k]]) for k in atoms}
    outs = []
    for _ in range(nevents):
        t = fix.next_step()
        if t <= 0:
            break
        x = advance(atoms, t)
        o = fix.insert(t, x, atoms["radius"], atoms["rmass"])
        outs.append((t, x, o))
        for k in atoms:
            atoms[k] = np.concatenate([atoms[k], o[k]])
    return outs, atoms


def test_stream_properties_closed_forms():
    o = oracle(tutorial())
    info = o.info()
    assert info[0] == 30000
    assert abs(info[1] - 300.0) < 1e-9
    # the scaled nodes leave the normal off z by rounding only
    np.testing.assert_allclose(info[3:6], [0.0, 0.0, -1.0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(info[6:9], [0.0, 0.0, -2.0], rtol=0, atol=1e-15)
    assert o.next_step() == 1
    # the first insertion: (int)(300 + u), u the fix stream's first draw
    from oracle import park_states

    u1 = park_states(32452867, 1)[1][0]
    out = o.insert(1, np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    assert len(out["radius"]) == int(300.0 + u1)
    # the insertion fraction: the extruded face lies inside the box
    assert o.info()[2] == 1.0
    assert o.next_step() == 30001


def test_stream_geometry_and_release():
    o = oracle(tutorial())
    tris = face_tris()
    outs, atoms = run_events(o, 3)
    for t, _, out in outs:
        x, r, rel = out["x"], out["radius"], out["release"]
        assert len(r) > 250
        assert on_footprint(x, tris).all()
        depth = x[:, 2] - FACE_Z
        assert (depth >= r - 1e-12).all() and (depth <= 0.6 + r + 1e-12).all()
        np.testing.assert_allclose(out["v"], np.tile([0.0, 0.0, -2.0], (len(r), 1)), rtol=0, atol=1e-15)
        assert not out["omega"].any()
        # release data: x_ins, insertion step, release step after the
        # crossing time, v_integrate = v_n, release velocity and spin
        np.testing.assert_array_equal(rel[:, :3], x)
        assert (rel[:, 3] == t).all()
        n_steps = np.floor((depth + 1e-14) / (2.0 * DT))
        assert np.abs(rel[:, 4] - t - n_steps).max() <= 1   # the normal's rounding may move a floor
        assert (rel[:, 4] - t == n_steps).mean() > 0.99
        np.testing.assert_array_equal(rel[:, 8:11], np.tile([0.0, -0.5, -2.0], (len(r), 1)))
        assert not rel[:, 11:].any()
        # the bodies of one insertion do not overlap
        d = np.linalg.norm(x[:, None] - x[None], axis=-1) - (r[:, None] + r[None])
        np.fill_diagonal(d, 1.0)
        assert d.min() > 0.0
    # big spheres are 0.7 of the mass: n_big / n_small = 0.7 / 0.3 * (0.015 / 0.025)^3
    nb = (atoms["radius"] == 0.025).sum()
    ns = (atoms["radius"] == 0.015).sum()
    assert abs(nb / ns - 0.7 / 0.3 * 0.216) < 0.06


def test_stream_stops_at_nparticles():
    spec = tutorial(nparticles=700)
    o = oracle(spec)
    outs, atoms = run_events(o, 10)
    assert len(outs) == 3
    assert len(atoms["radius"]) == 700
    assert o.vector()[0] == 700 and o.next_step() == 0


def test_face_sampled_by_area():
    """overlapcheck no, one small template: 20 insertions of 3000 points"""
    tris = face_tris()
    spec = tutorial(templates=[dict(type=1, radius=0.001, density=2500.0)], weights=[1.0], nparticles="INF",
                    particlerate=10000.0, overlapcheck=False)
    o = oracle(spec)
    pts = []
    for _ in range(20):
        t = o.next_step()
        pts.append(o.insert(t, np.zeros((0, 3)), np.zeros(0), np.zeros(0))["x"])
    p = np.concatenate(pts)
    assert len(p) > 50000
    assert on_footprint(p, tris).all()
    e0, e1 = tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    area = 0.5 * np.abs(e0[:, 0] * e1[:, 1] - e0[:, 1] * e1[:, 0])
    frac = area / area.sum()
    # which triangle each point falls in (first match)
    a, b, c = tris[:, 0, :2], tris[:, 1, :2], tris[:, 2, :2]
    hits = np.zeros(len(tris))
    for chunk in np.array_split(p, 20):
        q = chunk[:, None, :2]
        v0, v1, v2 = c - a, b - a, q - a
        d00, d01, d11 = (v0 * v0).sum(-1), (v0 * v1).sum(-1), (v1 * v1).sum(-1)
        d02, d12 = (v0 * v2).sum(-1), (v1 * v2).sum(-1)
        den = d00 * d11 - d01 * d01
        u = (d11 * d02 - d01 * d12) / den
        v = (d00 * d12 - d01 * d02) / den
        inside = (u > -1e-9) & (v > -1e-9) & (u + v < 1 + 1e-9)
        hits += np.bincount(inside.argmax(1), minlength=len(tris))
    n = len(p)
    sig = np.sqrt(n * frac * (1 - frac))
    assert (np.abs(hits - n * frac) < 5 * sig + 2).all()
    centroid = ((tris.mean(1)) * area[:, None]).sum(0) / area.sum()
    spread = np.sqrt(((p[:, :2] - centroid[:2]) ** 2).mean(0) / n)
    assert (np.abs(p[:, :2].mean(0) - centroid[:2]) < 5 * spread).all()
    # depth uniform in [r, 0.6 - r] (overlapcheck no shortens the extrusion)
    depth = p[:, 2] - FACE_Z
    assert depth.min() >= 0.001 - 1e-12 and depth.max() <= 0.6 - 0.001 + 1e-12
    assert abs(depth.mean() - 0.3) < 5 * 0.6 / np.sqrt(12 * n)


def test_nearby_filter():
    """residents beside the face's footprint and above its central hole (the
    face is a ring) are not tested: inserted spheres overlap them; one over
    the ring is"""
    spec = tutorial(nparticles="INF", particlerate=20000.0)
    o = oracle(spec)
    tris = face_tris()
    x = np.array([[0.24, 0.32, FACE_Z + 0.3],      # beside: footprint x in [-0.15, 0.15]
                  [0.0, 0.325, FACE_Z + 0.3],      # above the hole
                  [0.11, 0.3, FACE_Z + 0.3]])      # above the ring
    np.testing.assert_array_equal(on_footprint(x, tris), [False, False, True])
    r = np.array([0.1, 0.1, 0.03])
    m = 2500.0 * 4.0 / 3.0 * np.pi * r ** 3
    out = o.insert(1, x, r, m)
    assert len(out["radius"]) > 250      # the extrusion jams near 310 large spheres
    gap = np.linalg.norm(out["x"][:, None] - x[None], axis=-1) - (out["radius"][:, None] + r[None])
    assert (gap[:, 0] < 0).any() and (gap[:, 1] < 0).any()
    assert (gap[:, 2] > 0).all()


# ---- the engine's host loop against the oracle

def clump_template_pair():
    from oracle import TemplateOracle
    from test_template import host_template

    x = np.array([[0.0, 0.0, 0.0], [0.012, 0.0, 0.0], [0.0, 0.01, 0.004]])
    r = np.array([0.01, 0.008, 0.006])
    L, h, _ = host_template(15485863, x, r, ntry=20000)
    return TemplateOracle(15485863, x, r, ntry=20000), (L, h)


CASES = {
    "tutorial": dict(),
    "no_overlapcheck": dict(overlapcheck=False),
    "uniform_velocity_uncorrelated": dict(vel=("uniform", (0.0, -0.5, -2.0), (0.1, 0.2, 0.3)),
                                          random_distribute="uncorrelated", omega=(1.0, 2.0, 3.0)),
    "gaussian_insert_every_duration": dict(vel=("gaussian", (0.0, 0.0, -1.5), (0.05, 0.05, 0.1)),
                                           extrude_length=0.0, insert_every=40000, duration=30000,
                                           nparticles=1000),
    "mass_massrate_parallel": dict(nparticles=0, mass=60.0, massrate=40.0, particlerate=0.0, parallel=True),
    "random_orientation_clumps": "clumps",
}


@pytest.mark.parametrize("case", list(CASES))
def test_host_stream_equals_oracle(case):
    kw = CASES[case]
    host_ms = None
    if kw == "clumps":
        t_o, (L, h) = clump_template_pair()
        kw = dict(templates=[dict(type=1, radius=0.015, density=2500.0), dict(type=1, radius=0.02, density=2500.0)],
                  multisphere=[0, 1], orientation="random", nparticles=900)
        spec_o = tutorial(**kw, multisphere_oracle=[None, t_o])
        host_ms = [None, h]
    else:
        spec_o = tutorial(**kw)
    spec_h = dict(spec_o, multisphere_host=host_ms) if host_ms else spec_o
    o = oracle(spec_o)
    hl = HostLoop(spec_h, box=BOX, stream_dt=DT)
    # a resident beside the footprint and one inside the extrusion
    extra = dict(x=np.array([[0.2, 0.3, FACE_Z + 0.2], [0.0, 0.3, FACE_Z + 0.25]]), radius=np.array([0.06, 0.03]),
                 rmass=np.array([2.0, 0.3]), release=np.zeros((2, 14)))
    atoms = dict(extra)
    for event in range(4):
        t = o.next_step()
        assert hl.next_step() == t
        if t <= 0:
            break
        x = advance(atoms, t)
        ro = o.insert(t, x, atoms["radius"], atoms["rmass"])
        rh = hl.insert(t, x, atoms["radius"], atoms["rmass"])
        print("%s event %d at %d: %d atoms" % (case, event, t, len(ro["radius"])))
        rh.pop("brid")
        ro.pop("brid")
        same(rh, ro)
        np.testing.assert_array_equal(hl.vector(), o.vector())
        for k in atoms:
            atoms[k] = np.concatenate([atoms[k], ro[k]])
    hl.close()
    if host_ms:
        L.it_destroy(h)


def test_host_stream_equals_oracle_periodic():
    """x and y periodic, the box's x span that of the face: residents'
    images are tested through is_nearby on the image position"""
    box = ((-0.15, 0.1, -0.05), (0.15, 0.55, 2.85))
    spec = tutorial(nparticles=900)
    o = oracle(spec, box=box, periodic=(1, 1, 0))
    hl = HostLoop(spec, box=box, periodic=(1, 1, 0), stream_dt=DT)
    # residents near the periodic faces, inside the extrusion
    extra = dict(x=np.array([[0.148, 0.3, FACE_Z + 0.2], [-0.149, 0.25, FACE_Z + 0.3], [0.0, 0.545, FACE_Z + 0.1]]),
                 radius=np.array([0.02, 0.03, 0.025]), rmass=np.array([0.08, 0.3, 0.16]), release=np.zeros((3, 14)))
    atoms = dict(extra)
    for event in range(3):
        t = o.next_step()
        assert hl.next_step() == t
        x = advance(atoms, t)
        ro = o.insert(t, x, atoms["radius"], atoms["rmass"])
        rh = hl.insert(t, x, atoms["radius"], atoms["rmass"])
        ro.pop("brid")
        rh.pop("brid")
        same(rh, ro)
        assert len(ro["radius"]) > 100
        for k in atoms:
            atoms[k] = np.concatenate([atoms[k], ro[k]])
    hl.close()

