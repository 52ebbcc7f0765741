"""Pins the CPU oracle (oracle/dem_oracle.c) against what the mathematics
fixes, independently of the oracle's own formulas:

* Hertz elastic force of a static overlap vs the textbook law
  F = 4/3 E* sqrt(R*) delta^(3/2) with E*, R* from the textbook definitions;
* effective shear modulus vs G = E / (2(1+nu)) (Mindlin), beta vs the
  restitution relation e = exp(pi beta / sqrt(1 - beta^2));
* restitution of a head-on impact (velocity independent, equal to e);
* pair membership vs an O(N^2) brute force over all periodic images;
* Newton's third law and angular momentum: linear and total angular
  momentum (orbital + spin) of a colliding cluster with friction and
  rolling resistance are conserved to rounding, which fails for a wrong
  lever arm (cri/crj), a sign error or a missing j-side term;
* Coulomb cap |Ft| = mu |Fn| when sliding, rolling cap |M| = |Fn| reff mu_r;
* velocity Verlet is exact for constant force (free fall);
* history carry-over rules at a rebuild (kept iff touching, dropped else,
  stale between rebuilds).
These run on the CPU only.
"""
import numpy as np
import pytest

from oracle import Oracle

R1, R2 = 1.0e-3, 0.7e-3
RHO = 2500.0


def mass(r):
    return 4.0 / 3.0 * np.pi * r ** 3 * RHO


def mats(Y=5e6, nu=0.3, e=0.9, mu=0.5, mur=0.0, etar=0.0, ntypes=1, Y2=None, nu2=None):
    w = ntypes + 1
    Yv = np.full(w, Y)
    nuv = np.full(w, nu)
    if Y2 is not None:
        Yv[2] = Y2
        nuv[2] = nu2
    return dict(ntypes=ntypes, Y=Yv, nu=nuv, e=np.full((w, w), e), mu=np.full((w, w), mu),
                mur=np.full((w, w), mur), etar=np.full((w, w), etar), krs=0.0)


def spec(**kw):
    s = dict(box_lo=[-0.05] * 3, box_hi=[0.05] * 3, periodic=[0, 0, 0], dt=1e-7, skin=2.5e-4,
             neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off", materials=mats())
    s.update(kw)
    return s


def atoms(x, v, r, w=None, types=None):
    x = np.asarray(x, float)
    n = len(x)
    r = np.asarray(r, float)
    return dict(x=x, v=np.asarray(v, float), omega=np.zeros((n, 3)) if w is None else np.asarray(w, float),
                radius=r, rmass=mass(r), type=np.ones(n, int) if types is None else np.asarray(types),
                tag=np.arange(1, n + 1))


def test_hertz_static_textbook():
    Y1, Y2, n1, n2 = 5e6, 2e7, 0.3, 0.25
    delta = 2.0e-5
    d = R1 + R2 - delta
    o = Oracle(spec(materials=mats(Y=Y1, nu=n1, ntypes=2, Y2=Y2, nu2=n2)),
               atoms([[0, 0, 0], [d, 0, 0]], np.zeros((2, 3)), [R1, R2], types=[1, 2]))
    o.setup()
    f = o.state()["f"]
    Estar = 1.0 / ((1 - n1 ** 2) / Y1 + (1 - n2 ** 2) / Y2)
    Rstar = R1 * R2 / (R1 + R2)
    F = 4.0 / 3.0 * Estar * np.sqrt(Rstar) * delta ** 1.5
    assert f[0, 0] == pytest.approx(-F, rel=1e-12)
    assert f[1, 0] == pytest.approx(F, rel=1e-12)
    assert abs(f[0, 1]) + abs(f[0, 2]) == 0.0


def test_material_tables_closed_forms():
    e = 0.37
    Y = [0, 5e6, 3e7]
    nu = [0, 0.3, 0.2]
    w = 3
    m = dict(ntypes=2, Y=np.array(Y), nu=np.array(nu), e=np.full((w, w), e), mu=np.full((w, w), 0.5),
             mur=np.zeros((w, w)), etar=np.zeros((w, w)), krs=0.0)
    o = Oracle(spec(materials=m), atoms([[0, 0, 0]], [[0, 0, 0]], [R1]))
    import ctypes as C

    Ye = np.zeros((w, w))
    Ge = np.zeros((w, w))
    be = np.zeros((w, w))
    o.L.or_get_material_tables(o.h, Ye.ctypes.data_as(C.POINTER(C.c_double)),
                               Ge.ctypes.data_as(C.POINTER(C.c_double)),
                               be.ctypes.data_as(C.POINTER(C.c_double)))
    G = [0] + [Y[i] / (2 * (1 + nu[i])) for i in (1, 2)]
    assert Ge[1, 2] == pytest.approx(1.0 / ((2 - nu[1]) / G[1] + (2 - nu[2]) / G[2]), rel=1e-14)
    assert Ye[2, 1] == pytest.approx(1.0 / ((1 - nu[1] ** 2) / Y[1] + (1 - nu[2] ** 2) / Y[2]), rel=1e-14)
    b = be[1, 1]
    assert b < 0
    assert np.exp(np.pi * b / np.sqrt(1 - b * b)) == pytest.approx(e, rel=1e-13)


def restitution(e, v0, r=R1, dt=2e-8):
    gap = 1e-6
    o = Oracle(spec(dt=dt, materials=mats(e=e, mu=0.0)),
               atoms([[-r - gap, 0, 0], [r + gap, 0, 0]], [[v0, 0, 0], [-v0, 0, 0]], [r, r]))
    o.setup()
    # contact time of a Hertz impact ~ 2.87 (m*^2/(R* E*^2 v))^(1/5)
    mstar = mass(r) / 2
    Estar = 1.0 / (2 * (1 - 0.09) / 5e6)
    tc = 2.87 * (mstar ** 2 / (r / 2 * Estar ** 2 * v0)) ** 0.2
    o.run(int(3 * tc / dt) + int(gap / v0 / dt) + 10)
    s = o.state()
    assert s["x"][1, 0] - s["x"][0, 0] > 2 * r  # separated
    return (s["v"][1, 0] - s["v"][0, 0]) / (2 * v0)


@pytest.mark.parametrize("e", [0.3, 0.6, 0.9])
def test_restitution_head_on(e):
    e1 = restitution(e, 0.05)
    e2 = restitution(e, 0.4)
    assert e1 == pytest.approx(e, abs=2e-3)
    assert e2 == pytest.approx(e1, abs=2e-4)  # velocity independent


def brute_pairs(x, r, box_lo, box_hi, periodic, skin, cdf=1.0):
    """All directed (i, j, shift) with rsq <= ((ri+rj)*cdf+skin)^2, written
    out without the oracle: every periodic image of every atom."""
    L = np.asarray(box_hi) - np.asarray(box_lo)
    out = set()
    n = len(r)
    rng = [range(-1, 2) if periodic[d] else range(0, 1) for d in range(3)]
    for sx in rng[0]:
        for sy in rng[1]:
            for sz in rng[2]:
                img = x + np.array([sx * L[0], sy * L[1], sz * L[2]])
                code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1)
                dx = x[:, None, 0] - img[None, :, 0]
                dy = x[:, None, 1] - img[None, :, 1]
                dz = x[:, None, 2] - img[None, :, 2]
                rsq = dx * dx + dy * dy + dz * dz
                rs = (r[:, None] + r[None, :]) * cdf
                ok = rsq <= (rs + skin) * (rs + skin)
                if code == 13:
                    np.fill_diagonal(ok, False)
                ii, jj = np.nonzero(ok)
                for a, b in zip(ii, jj):
                    out.add((a + 1, b + 1, code))
    return out


@pytest.mark.parametrize("periodic", [[0, 0, 0], [1, 1, 0], [1, 1, 1]])
def test_pair_set_matches_brute_force(periodic):
    rng = np.random.default_rng(7)
    n = 400
    lo, hi = [0, 0, 0], [0.02, 0.02, 0.02]
    x = rng.uniform(0, 0.02, size=(n, 3))
    r = rng.uniform(0.4e-3, 0.9e-3, size=n)
    s = spec(box_lo=lo, box_hi=hi, periodic=periodic, skin=3e-4)
    o = Oracle(s, atoms(x, np.zeros((n, 3)), r))
    o.setup()
    ti, tj, sh, _, _ = o.pairs()
    got = set(zip(ti.tolist(), tj.tolist(), sh.tolist()))
    assert len(got) == len(ti)  # no duplicates
    assert got == brute_pairs(x, r, lo, hi, periodic, 3e-4)


def cluster(n=64, seed=3, rolling="epsd", mur=0.2, etar=0.3, e=0.5, mu=0.4):
    """4x4x4 lattice of slightly overlapping spheres, random spin, inward
    velocities: many sticking and sliding contacts for a few thousand steps."""
    rng = np.random.default_rng(seed)
    g = np.arange(4) * 1.6e-3 - 2.4e-3
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)[:n]
    x = x + rng.uniform(-2e-5, 2e-5, size=x.shape)
    r = rng.uniform(0.78e-3, 0.82e-3, size=n)
    v = -x * 40.0 + rng.uniform(-0.02, 0.02, size=(n, 3))
    w = rng.uniform(-50, 50, size=(n, 3))
    s = spec(rolling=rolling, materials=mats(e=e, mu=mu, mur=mur, etar=etar), dt=2e-7)
    return s, atoms(x, v, r, w)


def momenta(st, r):
    m = mass(r)
    P = (m[:, None] * st["v"]).sum(0)
    Lorb = (m[:, None] * np.cross(st["x"], st["v"])).sum(0)
    Lspin = (0.4 * m[:, None] * r[:, None] ** 2 * st["omega"]).sum(0)
    return P, Lorb + Lspin


@pytest.mark.parametrize("rolling", ["off", "epsd", "epsd2", "epsd3"])
def test_momentum_and_angular_momentum_conserved(rolling):
    s, a = cluster(rolling=rolling)
    if rolling == "epsd3":
        s["materials"]["krs"] = 1.5
    o = Oracle(s, a)
    o.setup()
    P0, L0 = momenta(o.state(), a["radius"])
    o.run(1500)
    fs, ts = o.force_scale()
    assert (fs > 0).sum() > 30 and (ts > 0).sum() > 30  # contacts with torques
    o.run(1500)
    st = o.state()
    P1, L1 = momenta(st, a["radius"])
    pscale = (mass(a["radius"]) * np.abs(a["v"]).max()).sum()
    lscale = pscale * 4e-3
    assert np.abs(P1 - P0).max() < 1e-12 * pscale
    assert np.abs(L1 - L0).max() < 1e-12 * lscale


def test_sliding_on_plane_coulomb_cap():
    r = R1
    mu = 0.35
    s = spec(materials=mats(mu=mu), planes=[(2, 0.0, 1)], gravity=(0, 0, -9.81), dt=1e-7)
    o = Oracle(s, atoms([[0, 0, r - 1e-6]], [[0.5, 0, 0]], [r]))
    o.setup()
    o.run(2000)
    st = o.state()
    g = mass(r) * 9.81
    Fn = st["f"][0, 2] + g
    Ft = np.hypot(st["f"][0, 0], st["f"][0, 1])
    assert Fn > 0
    assert Ft == pytest.approx(mu * Fn, rel=1e-12)
    assert st["f"][0, 0] < 0  # opposes sliding


def test_rolling_cap_epsd():
    r = R1
    mur = 0.1
    s = spec(rolling="epsd", materials=mats(mu=0.0, mur=mur, etar=0.3), planes=[(2, 0.0, 1)],
             gravity=(0, 0, -9.81), dt=1e-7)
    o = Oracle(s, atoms([[0, 0, r - 1e-6]], [[0, 0, 0]], [r], w=[[0, 400.0, 0]]))
    o.setup()
    o.run(3000)
    st = o.state()
    Fn = st["f"][0, 2] + mass(r) * 9.81
    M = np.linalg.norm(st["torque"][0])
    assert M == pytest.approx(Fn * r * mur, rel=1e-9)
    assert st["torque"][0, 1] < 0  # resists the spin


def test_free_fall_exact():
    g = -9.81
    s = spec(gravity=(0, 0, g), dt=1e-5)
    o = Oracle(s, atoms([[0, 0, 0.01]], [[0.1, 0, 0.2]], [R1]))
    o.setup()
    n = 1000
    o.run(n)
    t = n * 1e-5
    st = o.state()
    assert st["x"][0, 2] == pytest.approx(0.01 + 0.2 * t + 0.5 * g * t * t, rel=1e-12)
    assert st["v"][0, 2] == pytest.approx(0.2 + g * t, rel=1e-12)
    assert st["x"][0, 0] == pytest.approx(0.1 * t, rel=1e-12)


def pair_hist(o, a, b):
    ti, tj, sh, hi, fl = o.pairs()
    k = np.nonzero((ti == a) & (tj == b))[0]
    return (hi[k[0]], fl[k[0]]) if len(k) else (None, None)


def test_history_carry_rules():
    r = R1
    s = spec(mu=0.5, neigh_every=1, neigh_check=False, dt=1e-7, skin=1e-3)
    # two spheres overlapping and sliding past each other in y
    o = Oracle(s, atoms([[0, 0, 0], [2 * r - 2e-5, 0, 0]], [[0, 0.02, 0], [0, -0.02, 0]], [r, r]))
    o.setup()
    o.run(50)
    h, fl = pair_hist(o, 1, 2)
    assert fl != 0 and abs(h[1]) > 0          # shear accumulated, carried across rebuilds
    h21, _ = pair_hist(o, 2, 1)
    assert np.array_equal(h21, -h)            # partner copy is negated (newtonflag 1)


def test_history_dropped_when_apart_at_rebuild():
    r = R1
    s = spec(mu=0.5, neigh_every=1000, neigh_check=False, dt=1e-7, skin=2e-3)
    o = Oracle(s, atoms([[0, 0, 0], [2 * r - 1e-6, 0, 0]], [[-0.05, 0.02, 0], [0.05, -0.02, 0]], [r, r]))
    o.setup()
    o.run(200)                                # separate; no rebuild yet
    st = o.state()
    assert st["x"][1, 0] - st["x"][0, 0] > 2 * r
    h, fl = pair_hist(o, 1, 2)
    assert fl != 0 and abs(h[1]) > 0          # stale shear kept (cdf = 1: no surfacesClose)
    o.rebuild()
    h, fl = pair_hist(o, 1, 2)
    assert fl == 0 and not h.any()            # not touching at rebuild: dropped
