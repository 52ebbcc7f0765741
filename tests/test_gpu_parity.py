"""HIP engine vs the CPU oracle on the same seeded inputs (MI355X only).

Bars:
* pair sets (directed full list with periodic image codes): bit-exact,
  also against an O(N^2) brute force;
* FP64 forces and torques from an identical state: |gpu - oracle| <=
  1e-12 x the atom's force scale (sum of |contact contributions| and
  |gravity|), since only the order of summation differs;
* contact history per directed pair after a pass: <= 1e-12 relative;
* short horizon (20-40 steps, two rebuilds): positions within 1e-12 of the
  box size, forces within 1e-9 of the force scale (chaos amplifies the
  summation-order rounding step by step);
* mixed precision vs FP64 oracle: forces within 1e-5 of the force scale;
* two runs of the same input give bit-identical results (no atomics).
"""
import numpy as np
import pytest

from conftest import load_engine
from oracle import Oracle
from test_oracle_pins import brute_pairs

pytestmark = pytest.mark.gpu


def W():
    return load_engine().workloads


def pour_state(steps=3000):
    """config-1 pour advanced on the oracle until contacts exist."""
    spec, atoms = W().pour_box()
    o = Oracle(spec, atoms)
    o.setup()
    o.run(steps)
    return spec, atoms, o


def column_state(n=20000, steps=40):
    spec, atoms = W().packed_column(n=n, slab_n=5000)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(steps)
    return spec, atoms, o


def snapshot_atoms(atoms, o):
    st = o.state()
    a = dict(atoms)
    a["x"], a["v"], a["omega"] = st["x"], st["v"], st["omega"]
    return a


def sim_from_oracle(spec, atoms, o, precision="fp64"):
    """GPU simulation holding exactly the oracle's state and history; both
    then rebuild and take a force pass with shearupdate = 0."""
    eng = load_engine()
    a = snapshot_atoms(atoms, o)
    s = eng.Simulation(spec, a, precision=precision)
    ti, tj, sh, hi, fl = o.pairs()
    k = fl != 0
    s.set_history(ti[k], tj[k], hi[k])
    s.set_wall_history(o.wall_history())
    s.setup()
    o.rebuild()
    o.compute_forces(0)
    return s


def pair_dict(ti, tj, sh, hi):
    return {(a, b, c): h for a, b, c, h in zip(ti.tolist(), tj.tolist(), sh.tolist(), hi)}


def check_forces(s, o, tol):
    g = s.state()
    r = o.state()
    fs, ts = o.force_scale()
    # a contact torque is bounded by its force times the lever arm (< radius)
    ts = ts + fs * r["radius"] if "radius" in r else ts
    fs = np.maximum(fs, 1e-300)
    ts = np.maximum(ts, 1e-300)
    ef = np.abs(g["f"] - r["f"]).max(1) / fs
    et = np.abs(g["torque"] - r["torque"]).max(1) / ts
    assert ef.max() <= tol, ef.max()
    assert et.max() <= tol, et.max()
    return ef.max(), et.max()


def check_pairs(s, o, hist_tol=1e-12):
    gi, gj, gs, gh, gf = s.pairs()
    oi, oj, osh, oh, of = o.pairs()
    G = pair_dict(gi, gj, gs, gh)
    O = pair_dict(oi, oj, osh, oh)
    assert len(G) == len(gi)
    assert set(G) == set(O)
    scale = max(np.abs(oh).max(), 1e-300)
    err = max(np.abs(G[k] - O[k]).max() for k in O) if O else 0.0
    assert err <= hist_tol * scale, err / scale


@pytest.mark.parametrize("case", ["pour", "column"])
def test_forces_identical_state(case):
    spec, atoms, o = pour_state() if case == "pour" else column_state()
    s = sim_from_oracle(spec, atoms, o)
    check_pairs(s, o)
    check_forces(s, o, 1e-12)
    fs, _ = o.force_scale()
    assert (fs > 0).sum() > 100


def test_pair_set_brute_force_periodic():
    eng = load_engine()
    rng = np.random.default_rng(11)
    n = 600
    lo, hi = [0, 0, 0], [0.02, 0.02, 0.02]
    x = rng.uniform(0, 0.02, size=(n, 3))
    r = rng.uniform(0.4e-3, 0.9e-3, size=n)
    spec = dict(box_lo=lo, box_hi=hi, periodic=[1, 1, 1], dt=1e-7, skin=3e-4,
                materials=W()._materials())
    atoms = dict(x=x, v=np.zeros((n, 3)), omega=np.zeros((n, 3)), radius=r,
                 rmass=4 / 3 * np.pi * r ** 3 * 2500, type=np.ones(n, int), tag=np.arange(1, n + 1))
    s = eng.Simulation(spec, atoms)
    s.setup()
    gi, gj, gs, _, _ = s.pairs()
    got = set(zip(gi.tolist(), gj.tolist(), gs.tolist()))
    assert len(got) == len(gi)
    assert got == brute_pairs(x, r, lo, hi, [1, 1, 1], 3e-4)


@pytest.mark.parametrize("case", ["pour", "column"])
def test_short_horizon_trajectory(case):
    spec, atoms, o = pour_state() if case == "pour" else column_state()
    s = sim_from_oracle(spec, atoms, o)
    nsteps = 25
    s.step(nsteps)
    o.run(nsteps)
    g, r = s.state(), o.state()
    L = np.max(np.asarray(spec["box_hi"]) - np.asarray(spec["box_lo"]))
    assert np.abs(g["x"] - r["x"]).max() <= 1e-12 * L
    vs = np.abs(r["v"]).max()
    assert np.abs(g["v"] - r["v"]).max() <= 1e-9 * vs
    check_forces(s, o, 1e-9)
    check_pairs(s, o, hist_tol=1e-9)


def test_mixed_precision_force_error():
    spec, atoms, o = column_state()
    s = sim_from_oracle(spec, atoms, o, precision="mixed")
    ef, et = check_forces(s, o, 1e-5)


HYST = {k: np.full((2, 2), v) for k, v in dict(
    LoadingStiffness=5e4, UnloadingStiffness=2.5, coefficientAdhesionStiffness=0.0,
    coefficientPlasticityDepth=0.1, pullOffForce=0.0, alphaCustom=20.0, cinCustom=1e-7,
    aoneCustom=10e8, atwoCustom=4e4, athreeCustom=8.0, kcinCustom=1e-4).items()}


@pytest.mark.parametrize("model,wall", [("hysteretic/nonlinear1", "hertz"),
                                        ("hysteretic/nonlinear2", "hertz"),
                                        ("hysteretic/nonlinear1", "hysteretic/nonlinear1")])
def test_hysteretic_identical_state_and_horizon(model, wall):
    """Strain-hardening contact (the INL gold examples' parameters) on the
    pour: forces from an identical state, then a short horizon.  pow/exp/log
    come from the device math library, hence 1e-11 instead of 1e-12."""
    spec, atoms = W().pour_box()
    spec = dict(spec, normal=model, wall_normal=wall)
    spec["materials"] = dict(spec["materials"], hysteretic=HYST)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(3000)
    s = sim_from_oracle(spec, atoms, o)
    check_pairs(s, o, hist_tol=1e-11)
    check_forces(s, o, 1e-11)
    fs, _ = o.force_scale()
    assert (fs > 0).sum() > 100
    s.step(25)
    o.run(25)
    g, r = s.state(), o.state()
    assert np.abs(g["x"] - r["x"]).max() <= 1e-12 * 0.08
    check_forces(s, o, 1e-8)


def test_deterministic_bitwise():
    spec, atoms = W().packed_column(n=20000, slab_n=5000)
    eng = load_engine()
    out = []
    for _ in range(2):
        s = eng.Simulation(spec, atoms)
        s.setup()
        s.step(30)
        out.append(s.state())
        s.close()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(out[0][k], out[1][k])


def test_neighbor_overflow_error():
    eng = load_engine()
    spec, atoms = W().packed_column(n=5000, slab_n=5000)
    spec = dict(spec, neigh_one=3)
    s = eng.Simulation(spec, atoms)
    with pytest.raises(eng.DemError, match="Neighbor list overflow"):
        s.setup()


SHAPES = ((4, 256), (1, 64), (2, 128), (8, 64), (16, 128), (-8, 256), (-16, 128), (-32, 256), (-64, 256))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["pour", "column"])
def test_contact_shapes_agree(case, precision):
    """Every contact-pass shape (lane groups 1..16, contact-queue tiles
    8..64, blocks 64..256) gives the oracle's forces on an identical state,
    in FP64 (1e-12: the shapes differ only in the order each particle's
    contacts are summed) and in mixed precision (1e-5, the FP32 law)."""
    spec, atoms, o = pour_state() if case == "pour" else column_state(n=6000, steps=40)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for lanes, block in SHAPES:
        eng = load_engine()
        s = eng.Simulation(spec, snapshot_atoms(atoms, o), precision=precision)
        s.set_contact_shape(lanes, block)
        ti, tj, sh, hi, fl = o.pairs()
        k = fl != 0
        s.set_history(ti[k], tj[k], hi[k])
        s.set_wall_history(o.wall_history())
        s.setup()
        o2 = o   # the oracle's set-up pass on the same state
        o2.rebuild()
        o2.compute_forces(0)
        check_forces(s, o2, tol)
        check_pairs(s, o2, hist_tol=1e-12 if precision == "fp64" else 1e-5)
        s.close()


@pytest.mark.parametrize("rolling", ["epsd2", "epsd3"])
def test_rolling_models_identical_state_and_horizon(rolling):
    """EPSD2 (k_r = k_t R*^2, no viscous term) and EPSD3 (k_r = k_rs k_n
    mu_r^2 R*^2) on the dense column: forces, torques and rolling history
    from an identical state, then a short horizon over a rebuild."""
    spec, atoms = W().packed_column(n=20000, slab_n=5000, rolling=rolling)
    spec["materials"] = dict(spec["materials"], krs=1.7)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(40)
    s = sim_from_oracle(spec, atoms, o)
    check_pairs(s, o)
    check_forces(s, o, 1e-12)
    _, ts = o.force_scale()
    assert (ts > 0).sum() > 1000
    s.step(25)
    o.run(25)
    g, r = s.state(), o.state()
    assert np.abs(g["x"] - r["x"]).max() <= 1e-12 * 0.06
    check_forces(s, o, 1e-9)
    check_pairs(s, o, hist_tol=1e-9)


@pytest.mark.parametrize("big", [False, True])
def test_rows_touching_first(big):
    """Every row lists its touching partners first, then the others, and
    the pair set is the brute-force one: on random overlapping spheres whose
    rows stay within the 32 staged slots (single-search rebuild) and, with a
    3 mm sphere among them whose row is longer, on the two-search rebuild
    that such a row selects."""
    eng = load_engine()
    rng = np.random.default_rng(17)
    n = 1500
    lo, hi = [0, 0, 0], [0.02, 0.02, 0.02]
    x = rng.uniform(0, 0.02, size=(n, 3))
    r = rng.uniform(0.4e-3, 0.9e-3, size=n)
    if big:
        x[0] = [0.01, 0.01, 0.01]
        r[0] = 3e-3
    skin = 3e-4
    spec = dict(box_lo=lo, box_hi=hi, periodic=[1, 1, 1], dt=1e-7, skin=skin, materials=W()._materials())
    atoms = dict(x=x, v=np.zeros((n, 3)), omega=np.zeros((n, 3)), radius=r,
                 rmass=4 / 3 * np.pi * r ** 3 * 2500, type=np.ones(n, int), tag=np.arange(1, n + 1))
    s = eng.Simulation(spec, atoms)
    s.setup()
    gi, gj, gs, _, _ = s.pairs()
    got = set(zip(gi.tolist(), gj.tolist(), gs.tolist()))
    assert len(got) == len(gi)
    assert got == brute_pairs(x, r, lo, hi, [1, 1, 1], skin)
    rowlen = np.bincount(gi, minlength=n + 1)
    assert (rowlen.max() > 32) == big
    # touching at the build: rsq < (ri + rj)^2 for the entry's image
    L = np.asarray(hi) - np.asarray(lo)
    sh = np.stack([gs % 3 - 1, (gs // 3) % 3 - 1, gs // 9 - 1], 1)
    d = x[gi - 1] - (x[gj - 1] + sh * L)
    touch = (d * d).sum(1) < (r[gi - 1] + r[gj - 1]) ** 2
    assert touch.sum() > n
    starts = np.flatnonzero(np.r_[True, gi[1:] != gi[:-1]])
    assert len(starts) == len(np.unique(gi))   # rows are contiguous
    for a, b in zip(starts, np.r_[starts[1:], len(gi)]):
        t = touch[a:b]
        assert not np.any(~t[:-1] & t[1:]), gi[a]   # no touching entry after one that does not touch
