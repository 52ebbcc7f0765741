"""The adjoint's contact law and its reverse sweep, CPU only.

The law of csrc/contact_adj.h is built for the host (tests/adjoint/
law_harness.cpp) and checked:
* law_fwd against the oracle's contact pass: a sphere pair and a sphere on
  a plane, with tangential history built over earlier steps, in the sliding
  (Coulomb-capped) and the sticking regime, normal steps and set-up passes:
  force, torque and new shear to 1e-12;
* law_vjp against central finite differences of law_fwd, every input and
  parameter, random states in both regimes, for pairs and planes.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle import Oracle

CSRC = os.path.join(ROOT, "liggghts-inl_amd", "csrc")
DP = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def law(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("adj") / "liblaw.so")
    src = os.path.join(ROOT, "tests", "adjoint", "law_harness.cpp")
    subprocess.check_call(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-std=c++17", "-I", CSRC,
                           src, "-o", out])
    L = C.CDLL(out)
    L.law_fwd_c.argtypes = [DP, DP, C.c_double, C.c_int, DP]
    L.law_vjp_c.argtypes = [DP, DP, C.c_double, C.c_int, DP, DP, DP]
    return L


def p_(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(DP)


def fwd(L, x, p, dt, update):
    xa, xp = p_(x)
    pa, pp = p_(p)
    o = np.zeros(9)
    L.law_fwd_c(xp, pp, dt, update, o.ctypes.data_as(DP))
    return o


def vjp(L, x, p, dt, update, seed):
    xa, xp = p_(x)
    pa, pp = p_(p)
    sa, sp = p_(seed)
    gin = np.zeros(18)
    gp = np.zeros(4)
    L.law_vjp_c(xp, pp, dt, update, sp, gin.ctypes.data_as(DP), gp.ctypes.data_as(DP))
    return gin, gp


def pack(d, vi, vj, wi, wj, sh, radi, radj, meff, wall):
    reff = radi if wall else radi * radj / (radi + radj)
    radsum = radi if wall else radi + radj
    return np.concatenate([d, vi, vj, wi, wj, sh, [radi, radj, radsum, reff, meff, 1.0 if wall else 0.0]])


def mats(mu, e=0.5, Y=5e6, nu=0.3):
    return dict(ntypes=1, Y=np.full(2, Y), nu=np.full(2, nu), e=np.full((2, 2), e), mu=np.full((2, 2), mu),
                mur=np.zeros((2, 2)), etar=np.zeros((2, 2)), krs=0.0)


def law_params(o):
    Yo = np.zeros((2, 2))
    Go = np.zeros((2, 2))
    bo = np.zeros((2, 2))
    o.L.or_get_material_tables(o.h, Yo.ctypes.data_as(DP), Go.ctypes.data_as(DP), bo.ctypes.data_as(DP))
    return np.array([Yo[1, 1], Go[1, 1], bo[1, 1]])


@pytest.mark.parametrize("mu", [0.05, 2.0])
def test_law_matches_oracle_pair(law, mu):
    r1, r2 = 1.0e-3, 0.8e-3
    m1, m2 = 4 / 3 * np.pi * r1 ** 3 * 2500, 4 / 3 * np.pi * r2 ** 3 * 2500
    x = np.array([[0.0, 0.0, 0.0], [1.79e-3, 0.1e-3, -0.05e-3]])
    atoms = dict(x=x, v=np.array([[0.02, 0.3, -0.1], [-0.01, -0.2, 0.05]]),
                 omega=np.array([[30.0, -20.0, 5.0], [-10.0, 40.0, 25.0]]), radius=np.array([r1, r2]),
                 rmass=np.array([m1, m2]), type=np.array([1, 1]), tag=np.array([1, 2]))
    spec = dict(box_lo=[-0.01] * 3, box_hi=[0.01] * 3, periodic=[0, 0, 0], dt=1e-7, skin=2e-4,
                materials=mats(mu), neigh_every=1000, neigh_check=False)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(40)
    P = np.append(law_params(o), mu)
    ti, tj, sh, hist, fl = o.pairs()
    k = int(np.nonzero((ti == 1) & (tj == 2))[0][0])
    st = o.state()
    for update in (1, 0):
        h0 = hist[k].copy()
        o.compute_forces(update)
        s2 = o.state()
        h1 = o.pairs()[3][k]
        inp = pack(st["x"][0] - st["x"][1], st["v"][0], st["v"][1], st["omega"][0], st["omega"][1], h0,
                   r1, r2, m1 * m2 / (m1 + m2), False)
        out = fwd(law, inp, P, 1e-7, update)
        fs = np.abs(s2["f"][0]).max()
        assert np.abs(out[:3] - s2["f"][0]).max() <= 1e-12 * fs
        assert np.abs(out[3:6] - s2["torque"][0]).max() <= 1e-12 * (np.abs(s2["torque"][0]).max() + fs * r1)
        assert np.abs(out[6:] - h1).max() <= 1e-12 * np.abs(h1).max()
        hist = o.pairs()[3]
    if mu == 0.05:   # sliding: the tangential force sits on the Coulomb cap
        ft = out[:3] - np.dot(out[:3], inp[:3]) * inp[:3] / np.dot(inp[:3], inp[:3])
        fn = np.dot(out[:3], inp[:3]) / np.linalg.norm(inp[:3])
        assert abs(np.linalg.norm(ft) - mu * abs(fn)) <= 1e-12 * abs(fn)


def test_law_matches_oracle_plane(law):
    r = 1.0e-3
    m = 4 / 3 * np.pi * r ** 3 * 2500
    atoms = dict(x=np.array([[0.0, 0.0, r - 2e-5]]), v=np.array([[0.1, -0.05, -0.02]]),
                 omega=np.array([[10.0, 20.0, -5.0]]), radius=np.array([r]), rmass=np.array([m]),
                 type=np.array([1]), tag=np.array([1]))
    spec = dict(box_lo=[-0.01, -0.01, 0.0], box_hi=[0.01, 0.01, 0.02], periodic=[0, 0, 0], dt=1e-7,
                skin=2e-4, materials=mats(0.4), planes=[(2, 0.0, 1)], neigh_every=1000, neigh_check=False)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(30)
    P = np.append(law_params(o), 0.4)
    st = o.state()
    h0 = o.wall_history()[0, 0].copy()
    o.compute_forces(1)
    s2 = o.state()
    h1 = o.wall_history()[0, 0]
    z = st["x"][0, 2]
    inp = pack(np.array([0.0, 0.0, z]), st["v"][0], np.zeros(3), st["omega"][0], np.zeros(3), h0, r, 0.0, m, True)
    out = fwd(law, inp, P, 1e-7, 1)
    fs = np.abs(s2["f"][0]).max()
    assert np.abs(out[:3] - s2["f"][0]).max() <= 1e-12 * fs
    assert np.abs(out[3:6] - s2["torque"][0]).max() <= 1e-12 * fs * r
    assert np.abs(out[6:] - h1).max() <= 1e-12 * np.abs(h1).max()


def random_state(rng, wall, capped):
    radi = rng.uniform(0.8e-3, 1.2e-3)
    radj = 0.0 if wall else rng.uniform(0.8e-3, 1.2e-3)
    radsum = radi if wall else radi + radj
    ov = rng.uniform(1e-6, 3e-5)
    if wall:
        d = np.array([0.0, 0.0, radi - ov])
    else:
        u = rng.normal(size=3)
        d = u / np.linalg.norm(u) * (radsum - ov)
    vi, vj = rng.normal(scale=0.2, size=3), (np.zeros(3) if wall else rng.normal(scale=0.2, size=3))
    wi, wj = rng.normal(scale=50, size=3), (np.zeros(3) if wall else rng.normal(scale=50, size=3))
    sh = rng.normal(scale=(5e-5 if capped else 1e-8), size=3)
    meff = rng.uniform(2e-6, 1e-5)
    x = pack(d, vi, vj, wi, wj, sh, radi, radj, meff, wall)
    p = np.array([rng.uniform(5e5, 5e6), rng.uniform(2e5, 2e6), -rng.uniform(0.05, 0.9),
                  rng.uniform(0.2, 0.7) if not capped else 0.05])
    return x, p


@pytest.mark.parametrize("wall", [False, True])
@pytest.mark.parametrize("capped", [False, True])
@pytest.mark.parametrize("update", [1, 0])
def test_law_vjp_finite_differences(law, wall, capped, update):
    rng = np.random.default_rng(100 + 4 * wall + 2 * capped + update)
    dt = 1e-6
    # inputs that vary: the plane's separation is along its axis only, and
    # a plane has no partner velocity
    groups = [(0, 3), (3, 6), (6, 9), (9, 12), (12, 15), (15, 18)]
    if wall:
        groups = [(2, 3), (3, 6), (9, 12), (15, 18)]
    for trial in range(6):
        x, p = random_state(rng, wall, capped)
        o = fwd(law, x, p, dt, update)
        if capped and update:   # on the cap (a set-up pass does not project the shear)
            ft = o[:3] - np.dot(o[:3], x[:3]) * x[:3] / np.dot(x[:3], x[:3])
            fn = np.dot(o[:3], x[:3]) / np.linalg.norm(x[:3])
            assert abs(np.linalg.norm(ft) - p[3] * abs(fn)) <= 1e-10 * abs(fn)
        scale_out = np.concatenate([np.full(3, np.abs(o[:3]).max()), np.full(3, np.abs(o[3:6]).max()),
                                    np.full(3, max(np.abs(o[6:]).max(), 1e-30))])
        seed = rng.normal(size=9) / scale_out
        gin, gp = vjp(law, x, p, dt, update, seed)
        for a, b in groups:
            xs = np.abs(x[a:b]).max()
            gs = np.abs(gin[a:b]).max()
            for k in range(a, b):
                h = 1e-6 * xs
                xp, xm = x.copy(), x.copy()
                xp[k] += h
                xm[k] -= h
                fd = np.dot(seed, fwd(law, xp, p, dt, update) - fwd(law, xm, p, dt, update)) / (2 * h)
                assert abs(fd - gin[k]) <= 1e-5 * (abs(gin[k]) + 1e-3 * gs) + 1e-300, (trial, k, fd, gin[k])
        for k in range(4):
            h = 1e-6 * abs(p[k])
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = np.dot(seed, fwd(law, x, pp, dt, update) - fwd(law, x, pm, dt, update)) / (2 * h)
            # in units of the relative parameter change; a seed of unit size
            # per output scale makes the finite-difference noise ~1e-10 here
            assert abs(fd - gp[k]) * abs(p[k]) <= 1e-5 * abs(gp[k] * p[k]) + 1e-8, (trial, "param", k, fd, gp[k])
