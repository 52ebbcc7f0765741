"""Pins the oracle's sticking-regime laws by the oscillations they produce.

A sphere resting under gravity on a plane, or on a frozen sphere, sits at
the static Hertz overlap.  Below the friction caps its tangential contact
displacement s and its rolling-spring torque M then obey linear damped
oscillators whose coefficients follow from textbook contact mechanics:

* tangential:  m_t s'' + gamma_t s' + k_t s = 0, with the Mindlin stiffness
  k_t = 8 G* sqrt(R* delta) (G* = 1 / ((2-nu_1)/G_1 + (2-nu_2)/G_2)), the
  dashpot gamma_t = 2 sqrt(5/6) |beta| sqrt(k_t m*) of the Hertz model
  (beta the restitution coefficient's damping ratio) and the mass of the
  contact point 1/m_t = 1/m + c^2/I, where c = r - delta/2 is the lever
  from the centre to the contact point and I = 2/5 m r^2;
* rolling (no sliding friction, so the sphere only spins):  the spring
  torque M = k_r theta of the rolling angle, with k_r = 2.25 k_n mu_r^2 R*^2
  (EPSD, k_n = 4/3 E* sqrt(R* delta) the secant normal stiffness),
  k_t R*^2 (EPSD2) or k_rs k_n mu_r^2 R*^2 (EPSD3), and the viscous torque
  eta_r 2 sqrt(I_r k_r) theta' (EPSD, EPSD3; none for EPSD2) with I_r the
  rolling inertia about the contact point, 7/5 m r^2 for one sphere and
  7/5 I_i I_j / (I_i + I_j) (I = m r^2) for a pair.

LIGGGHTS steps these with velocity Verlet whose velocity-dependent forces
see the half-kicked velocity (fix_nve_sphere.cpp:134-243 around
PairGran::compute).  For x'' = -k x - c x' that scheme is exactly the
linear recurrence

    x[n+2] - 2 x[n+1] + x[n] = -c dt (x[n+1] - x[n]) - k dt^2 x[n+1]

on the values after each force pass (the history slots hold x[n+1]).  The
tests read s or M after every step, fit (c dt, k dt^2) of that recurrence by
least squares and compare them with the closed forms to 1e-6: a factor
error in k_t, gamma_t, k_r or the viscous term, a missing sqrt(5/6), a wrong
lever arm or inertia, or a lagged velocity all fail.  CPU only.
"""
import numpy as np
import pytest

from oracle import Oracle

RHO = 2500.0
G = 9.81
Y, NU = 5e6, 0.3


def mass(r):
    return 4.0 / 3.0 * np.pi * r ** 3 * RHO


def beta(e):
    le = np.log(e)
    return le / np.sqrt(le * le + np.pi * np.pi)


def estar():
    return 1.0 / (2 * (1 - NU * NU) / Y)


def gstar():
    Gm = Y / (2 * (1 + NU))
    return 1.0 / (2 * (2 - NU) / Gm)


def static_overlap(m, reff):
    """delta with 4/3 E* sqrt(R*) delta^(3/2) = m g."""
    return (m * G / (4.0 / 3.0 * estar() * np.sqrt(reff))) ** (2.0 / 3.0)


def materials(e, mu, mur=0.0, etar=0.0, krs=0.0):
    w = 2
    return dict(ntypes=1, Y=np.full(w, Y), nu=np.full(w, NU), e=np.full((w, w), e),
                mu=np.full((w, w), mu), mur=np.full((w, w), mur), etar=np.full((w, w), etar), krs=krs)


def spec(dt, mats, rolling="off", planes=(), freeze=False, half=0.05):
    s = dict(box_lo=[-half] * 3, box_hi=[half] * 3, periodic=[0, 0, 0], dt=dt, skin=2.5e-4,
             neigh_every=1000000, neigh_delay=0, neigh_check=False, rolling=rolling,
             materials=mats, planes=list(planes), gravity=(0.0, 0.0, -G))
    order = ["wall:%d" % k for k in range(len(planes))] + ["gravity"]
    if freeze:
        order.append("freeze")
    s["fix_order"] = order
    return s


def fit_recurrence(x, dt):
    """(c, k) of x[n+2]-2x[n+1]+x[n] = -c dt (x[n+1]-x[n]) - k dt^2 x[n+1]."""
    x = np.asarray(x, float)
    d2 = x[2:] - 2 * x[1:-1] + x[:-2]
    A = np.stack([-(x[1:-1] - x[:-2]), -x[1:-1]], 1)
    sc = np.abs(A).max(0)
    sol, *_ = np.linalg.lstsq(A / sc, d2, rcond=None)
    cdt, kdt2 = sol / sc
    resid = d2 - A @ (sol / sc)
    return cdt / dt, kdt2 / dt ** 2, np.abs(resid).max() / np.abs(d2).max()


def series(o, nsteps, read):
    out = np.empty(nsteps)
    for n in range(nsteps):
        o.run(1)
        out[n] = read(o)
    return out


def sphere_on_plane(r, dt, mats, rolling, v0=(0, 0, 0), w0=(0, 0, 0)):
    m = mass(r)
    delta = static_overlap(m, r)
    z = r - delta
    s = spec(dt, mats, rolling=rolling, planes=[(2, 0.0, 1)])
    a = dict(x=np.array([[0.0, 0.0, z]]), v=np.array([v0], float), omega=np.array([w0], float),
             radius=np.array([r]), rmass=np.array([m]), type=np.ones(1, int), tag=np.array([1]))
    o = Oracle(s, a)
    o.setup()
    return o, m, r - z  # the overlap the oracle sees


def sphere_on_frozen(ri, rj, dt, mats, rolling, v0=(0, 0, 0), w0=(0, 0, 0)):
    mi, mj = mass(ri), mass(rj)
    reff = ri * rj / (ri + rj)
    delta = static_overlap(mi, reff)
    zi = rj + ri - delta
    s = spec(dt, mats, rolling=rolling, freeze=True, half=1.5 * (ri + rj))
    a = dict(x=np.array([[0.0, 0.0, 0.0], [0.0, 0.0, zi]]), v=np.array([[0, 0, 0], v0], float),
             omega=np.array([[0, 0, 0], w0], float), radius=np.array([rj, ri]), rmass=np.array([mj, mi]),
             type=np.ones(2, int), tag=np.array([1, 2]), freeze=np.array([1, 0]))
    o = Oracle(s, a)
    o.setup()
    return o, mi, mj, reff, (ri + rj) - zi


def pair_slot(o, slot):
    ti, tj, sh, hi, fl = o.pairs()
    k = np.nonzero((ti == 2) & (tj == 1))[0][0]
    return hi[k, slot]


@pytest.mark.parametrize("e", [0.9, 0.4])
def test_tangential_stick_on_plane(e):
    r, dt, mu = 1.0e-3, 1.0e-7, 0.5
    o, m, delta = sphere_on_plane(r, dt, materials(e, mu), "off", v0=(1e-4, 0, 0))
    s = series(o, 30000, lambda o: o.wall_history()[0, 0, 0])
    kt = 8 * gstar() * np.sqrt(r * delta)
    gam = 2 * np.sqrt(5.0 / 6.0) * abs(beta(e)) * np.sqrt(kt * m)
    lever = r - 0.5 * delta
    inv_mt = 1.0 / m + lever * lever / (0.4 * m * r * r)
    c, k, res = fit_recurrence(s, dt)
    assert res < 1e-6                      # a linear oscillator: still sticking
    assert np.abs(s).max() * kt < 0.2 * mu * m * G
    assert k == pytest.approx(kt * inv_mt, rel=1e-6)
    assert c == pytest.approx(gam * inv_mt, rel=1e-6)
    # oscillating, not overdamped
    assert np.count_nonzero(np.diff(np.sign(s)) != 0) >= 4


def test_tangential_stick_on_frozen_sphere():
    # Sliding moves the top sphere over the curved partner, where gravity
    # adds a tangential force ~ m g / (r_i + r_j) per unit displacement: a
    # 2 m partner keeps that far below k_t, and R* still differs from r_i
    # by 4e-4.
    # The spin is chosen so that the angular momentum about the contact
    # point is zero: the sphere then oscillates in place, with no steady
    # rolling that would carry it over the curved partner.
    ri, rj, dt, mu, e = 0.8e-3, 2.0, 1.0e-7, 0.5, 0.7
    reff = ri * rj / (ri + rj)
    d0 = static_overlap(mass(ri), reff)
    v0 = 8e-5
    w0 = (ri - 0.5 * d0) * v0 / (0.4 * ri * ri)
    o, mi, mj, reff, delta = sphere_on_frozen(ri, rj, dt, materials(e, mu), "off", v0=(0, v0, 0),
                                              w0=(w0, 0, 0))
    s = series(o, 20000, lambda o: pair_slot(o, 1))
    kt = 8 * gstar() * np.sqrt(reff * delta)
    # the frozen partner leaves m* = m_i (pair_gran_base.h:389-393)
    gam = 2 * np.sqrt(5.0 / 6.0) * abs(beta(e)) * np.sqrt(kt * mi)
    lever = ri - 0.5 * delta
    inv_mt = 1.0 / mi + lever * lever / (0.4 * mi * ri * ri)
    c, k, res = fit_recurrence(s, dt)
    assert res < 1e-8
    assert k == pytest.approx(kt * inv_mt, rel=1e-6)
    assert c == pytest.approx(gam * inv_mt, rel=1e-6)
    assert np.count_nonzero(np.diff(np.sign(s)) != 0) >= 4


def rolling_coeffs(model, kn, kt, reff, mur, etar, krs, Ir):
    if model == "epsd":
        kr = 2.25 * kn * mur * mur * reff * reff
        return kr, etar * 2 * np.sqrt(Ir * kr)
    if model == "epsd2":
        return kt * reff * reff, 0.0
    kr = krs * kn * mur * mur * reff * reff
    return kr, etar * 2 * np.sqrt(Ir * kr)


@pytest.mark.parametrize("model", ["epsd", "epsd2", "epsd3"])
def test_rolling_stick_on_plane(model):
    r, dt, mur, etar, krs = 1.0e-3, 5.0e-7, 0.1, 0.3, 1.7
    mats = materials(0.9, 0.0, mur=mur, etar=etar, krs=krs)
    o, m, delta = sphere_on_plane(r, dt, mats, model, w0=(0, 0.05, 0))
    M = series(o, 60000, lambda o: o.wall_history()[0, 0, 4])
    kn = 4.0 / 3.0 * estar() * np.sqrt(r * delta)
    kt = 8 * gstar() * np.sqrt(r * delta)
    kr, cr = rolling_coeffs(model, kn, kt, r, mur, etar, krs, 1.4 * m * r * r)
    # on a wall the rolling velocity is omega scaled by (r - delta/2)/(r - delta)
    # (surface_model_default.h:181-186)
    alpha = (r - 0.5 * delta) / (r - delta)
    inertia = 0.4 * m * r * r
    c, k, res = fit_recurrence(M, dt)
    assert res < 1e-6
    assert np.abs(M).max() < 0.2 * m * G * r * mur     # below the rolling cap
    assert k == pytest.approx(alpha * kr / inertia, rel=1e-6)
    if cr == 0.0:
        assert abs(c) * dt < 1e-12
    else:
        assert c == pytest.approx(alpha * cr / inertia, rel=1e-6)
    assert np.count_nonzero(np.diff(np.sign(M)) != 0) >= 4


@pytest.mark.parametrize("model", ["epsd", "epsd2", "epsd3"])
def test_rolling_stick_on_frozen_sphere(model):
    ri, rj, dt, mur, etar, krs = 0.9e-3, 1.2e-3, 5.0e-7, 0.12, 0.25, 2.3
    mats = materials(0.9, 0.0, mur=mur, etar=etar, krs=krs)
    o, mi, mj, reff, delta = sphere_on_frozen(ri, rj, dt, mats, model, w0=(0.05, 0, 0))
    M = series(o, 40000, lambda o: pair_slot(o, 3))
    kn = 4.0 / 3.0 * estar() * np.sqrt(reff * delta)
    kt = 8 * gstar() * np.sqrt(reff * delta)
    Ii, Ij = mi * ri * ri, mj * rj * rj
    kr, cr = rolling_coeffs(model, kn, kt, reff, mur, etar, krs, 1.4 * Ii * Ij / (Ii + Ij))
    inertia = 0.4 * mi * ri * ri
    c, k, res = fit_recurrence(M, dt)
    assert res < 1e-6
    assert k == pytest.approx(kr / inertia, rel=1e-6)
    if cr == 0.0:
        assert abs(c) * dt < 1e-12
    else:
        assert c == pytest.approx(cr / inertia, rel=1e-6)
