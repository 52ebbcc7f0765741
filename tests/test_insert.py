"""fix insert/pack: the oracle pinned by the mathematics of its random
streams, and the engine's host placement loop checked against the oracle on
the CPU.

The reference ships no output of an insertion (its packing and pour
examples write only VTK dumps, not kept in the repository), so the oracle is
pinned by what the algorithm fixes in closed form:

* Park-Miller's minimal standard generator (random_park.cpp): the published
  check value (seed 1, the 10,000th state is 1043618065) and the exact
  modular recurrence seed' = 16807 seed mod (2^31 - 1) on many seeds;
* the region volume of a block inside the box is the block's volume exactly
  (every Monte Carlo sample falls inside), and of a block half outside the
  box, the inside half within the binomial error;
* the number to insert: int((f V - V_present) / E[V] + u) with u the fix
  stream's first draw and E[V] the distribution's number-weighted mean
  volume (mass weights divided by the template masses);
* how many of each template: the exact-number split (truncation, then the
  gap drawn from the distribution's stream);
* the first particle: the region stream reset to seed + 12 and advanced by
  the 3 x ntry_mc draws of the volume estimate gives its position; the fix
  stream after the count draw gives its velocity;
* no two particles overlap and every particle keeps its radius from the
  region's faces (all_in) and from the box (the sub-domain border check).

The engine's host loop (insert_pack.cpp) draws attempts in batches from copies of
the streams and walks them serially; tests/insert_host_driver.cpp links it
with brute-force overlap hooks, and it must give the oracle's particles bit
for bit.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_engine

IM = 2147483647
BOX = ((-0.05, -0.05, 0.0), (0.05, 0.05, 0.15))
PACKING = dict(seed=32452867, dist_seed=32452843,
               templates=[dict(type=1, radius=0.005, density=2500.0), dict(type=1, radius=0.008, density=2500.0)],
               weights=[0.3, 0.7], region=BOX, volumefraction_region=0.25, maxattempt=200, insert_every="once",
               overlapcheck=True, all_in=True, vel=("constant", (0.0, 0.0, 0.0)))


def park(seed, n):
    from oracle import park_states

    return park_states(seed, n)


def test_park_miller_check_value():
    st, va = park(1, 10000)
    assert st[-1] == 1043618065
    assert st[0] == 16807
    np.testing.assert_array_equal(va, (1.0 / IM) * st.astype(np.float64))


def test_park_miller_is_the_modular_recurrence():
    rng = np.random.default_rng(7)
    for s0 in [1, 2, IM - 1, 16807, 127773, 127774] + list(rng.integers(1, IM, 40)):
        st, _ = park(int(s0), 50)
        s = int(s0)
        for k in range(50):
            s = (16807 * s) % IM
            assert st[k] == s


def oracle_fix(spec, box=BOX, periodic=(0, 0, 0), ntimestep=0):
    from oracle import InsertPackOracle

    return InsertPackOracle(spec, box[0], box[1], periodic, ntimestep)


def test_region_volume_inside_box_is_exact():
    o = oracle_fix(PACKING)
    lo, hi = np.array(BOX[0]), np.array(BOX[1])
    assert o.region_volume() == np.prod(hi - lo)


def test_region_volume_half_outside_box():
    spec = dict(PACKING, region=((-0.05, -0.05, -0.15), (0.05, 0.05, 0.15)), all_in=False)
    o = oracle_fix(spec)
    p = 0.5
    n = 100000
    v_bbox = 0.1 * 0.1 * 0.3
    est = o.region_volume() / v_bbox
    assert abs(est - p) < 4 * np.sqrt(p * (1 - p) / n)


def expected_numbers(spec, n_present_vol=0.0):
    """count to insert and per-template split, from the streams' draws"""
    r = np.array([t["radius"] for t in spec["templates"]])
    rho = np.array([t["density"] for t in spec["templates"]])
    vol = r * r * r * 4.0 * np.pi / 3.0
    w = np.array(spec["weights"], dtype=np.float64)
    w = w / w.sum()
    w = w / (rho * vol)
    w = w / w.sum()
    ev = float((vol * w).sum())
    lo, hi = np.array(spec["region"][0]), np.array(spec["region"][1])
    V = float(np.prod(hi - lo))
    u = park(spec["seed"], 1)[1][0]
    n = int((spec["volumefraction_region"] * V - n_present_vol) / ev + u)
    part = np.floor(n * w).astype(int)
    rem = n * w - part
    gap = n - part.sum()
    du = park(spec["dist_seed"], max(gap, 1))[1]
    for i in range(gap):
        rr = du[i] * gap
        j = 0
        rs = rem[0]
        while rs < rr and j < len(w) - 1:
            j += 1
            rs += rem[j]
        part[j] += 1
    return n, part


def test_packing_counts_and_split():
    o = oracle_fix(PACKING)
    out = o.insert(1, np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    n, part = expected_numbers(PACKING)
    assert n > 300
    # enough room: every particle placed, large ones first
    assert len(out["radius"]) == n
    r = out["radius"]
    assert (np.sum(r == 0.005), np.sum(r == 0.008)) == (part[0], part[1])
    assert (r[:part[1]] == 0.008).all() and (r[part[1]:] == 0.005).all()
    np.testing.assert_array_equal(out["rmass"], 2500.0 * (r * r * r * 4.0 * np.pi / 3.0))
    vec = o.vector()
    assert vec[0] == n and vec[5] == 0.0   # once: no further insertion


def test_first_particle_from_the_streams():
    spec = dict(PACKING, all_in=False, vel=("uniform", (0.1, -0.2, 0.3), (0.01, 0.02, 0.03)))
    o = oracle_fix(spec)
    out = o.insert(1, np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    lo, hi = np.array(spec["region"][0]), np.array(spec["region"][1])
    rb = 0.008                      # the first particle is of the larger template
    _, u = park(spec["seed"] + 12, 3 * 100000 + 300)
    k = 3 * 100000
    while True:                     # redrawn while closer than rb to the box
        p = lo + u[k:k + 3] * (hi - lo)
        k += 3
        if (np.abs(lo - p) >= rb).all() and (np.abs(hi - p) >= rb).all():
            break
    np.testing.assert_array_equal(out["x"][0], p)
    fu = park(spec["seed"], 4)[1]
    v = np.array([0.1, -0.2, 0.3]) + np.array([0.01, 0.02, 0.03]) * 2.0 * (fu[1:4] - 0.5)
    np.testing.assert_array_equal(out["v"][0], v)


def no_overlap(x, r, xr=None, rr=None, prd=None):
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    rs = r[:, None] + r[None, :]
    np.fill_diagonal(d2, np.inf)
    ok = (d2 > rs * rs).all()
    if xr is not None and len(xr):
        shifts = [np.zeros(3)]
        if prd is not None:
            shifts = [np.array([a, b, 0.0]) * prd for a in (-1, 0, 1) for b in (-1, 0, 1)]
        for sh in shifts:
            d2 = (((x[:, None, :] - (xr[None, :, :] + sh))) ** 2).sum(-1)
            rs = r[:, None] + rr[None, :]
            ok = ok and (d2 > rs * rs).all()
    return ok


def test_packing_geometry():
    o = oracle_fix(PACKING)
    out = o.insert(1, np.zeros((0, 3)), np.zeros(0), np.zeros(0))
    x, r = out["x"], out["radius"]
    lo, hi = np.array(BOX[0]), np.array(BOX[1])
    assert no_overlap(x, r)
    assert ((x - r[:, None] >= lo) & (x + r[:, None] <= hi)).all()


def resident_bed(n=300, seed=3, half=0.045):
    """non-overlapping spheres in the lower part of the packing box"""
    rng = np.random.default_rng(seed)
    x, r = [], []
    while len(r) < n:
        p = rng.uniform([-half, -half, 0.004], [half, half, 0.06])
        q = rng.uniform(0.003, 0.006)
        if all(np.sum((p - y) ** 2) > (q + s) ** 2 for y, s in zip(x, r)):
            x.append(p)
            r.append(q)
    x = np.array(x)
    r = np.array(r)
    return x, r, 2500.0 * 4.0 / 3.0 * np.pi * r ** 3


def test_particles_in_region_count_and_no_overlap():
    x, r, m = resident_bed()
    spec = dict(PACKING, particles_in_region=450, maxattempt=100)
    del spec["volumefraction_region"]
    o = oracle_fix(spec)
    out = o.insert(1, x, r, m)
    n_present = np.sum((x >= np.array(BOX[0])).all(1) & (x <= np.array(BOX[1])).all(1))
    assert o.vector()[3] == 450 - n_present
    assert len(out["radius"]) == 450 - n_present
    assert no_overlap(out["x"], out["radius"], x, r)


# ---- the engine's host loop against the oracle (CPU)

_DRV = None


def driver():
    global _DRV
    if _DRV is not None:
        return _DRV
    out = os.path.join(ROOT, "_build", "tests", "libinsert_host_driver.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "include"), "-o", out,
                    os.path.join(ROOT, "tests", "insert_host_driver.cpp"),
                    os.path.join(ROOT, "liggghts-inl_amd", "csrc", "insert_pack.cpp"),
                    os.path.join(ROOT, "liggghts-inl_amd", "csrc", "mesh_host.cpp")], check=True)
    L = C.CDLL(out)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    eng = load_engine()
    L.ih_create.restype = C.c_void_p
    L.ih_create.argtypes = [C.POINTER(eng.engine.InsertPackSpec), dp, dp, ip, C.c_longlong,
                            C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
    L.ih_destroy.argtypes = [C.c_void_p]
    L.ih_next_step.restype = C.c_longlong
    L.ih_next_step.argtypes = [C.c_void_p]
    L.ih_vector.argtypes = [C.c_void_p, dp]
    L.ih_insert.argtypes = [C.c_void_p, C.c_longlong, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp, dp, dp, ip, dp,
                            ip, dp, C.c_double]
    L.ih_stream_create.restype = C.c_void_p
    L.ih_stream_create.argtypes = [C.POINTER(eng.engine.InsertStreamSpec), dp, dp, ip, C.c_longlong, C.c_double,
                                   C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
    _DRV = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class HostLoop:
    """spec["multisphere_host"]: per template an it_create handle or None"""

    def __init__(self, spec, box=BOX, periodic=(0, 0, 0), ntimestep=0, stream_dt=None):
        eng = load_engine()
        self.L = driver()
        self.stream = stream_dt is not None
        if self.stream:
            self.s, self.keep = eng.engine.insert_stream_spec(spec)
        else:
            self.s, self.keep = eng.engine.insert_pack_spec(spec)
        hs = spec.get("multisphere_host")
        self.ms = None if hs is None else (C.c_void_p * len(hs))(*hs)
        lo = np.ascontiguousarray(box[0], dtype=np.float64)
        hi = np.ascontiguousarray(box[1], dtype=np.float64)
        pe = np.ascontiguousarray(periodic, dtype=np.int32)
        err = C.create_string_buffer(256)
        if self.stream:
            self.h = self.L.ih_stream_create(C.byref(self.s), _dp(lo), _dp(hi), pe.ctypes.data_as(C.POINTER(C.c_int)),
                                             int(ntimestep), float(stream_dt), self.ms, err, 256)
        else:
            self.h = self.L.ih_create(C.byref(self.s), _dp(lo), _dp(hi), pe.ctypes.data_as(C.POINTER(C.c_int)),
                                      int(ntimestep), self.ms, err, 256)
        if not self.h:
            raise RuntimeError(err.value.decode())

    def insert(self, step, x, r, m, cap=1 << 18, maxrad=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        r = np.ascontiguousarray(r, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.float64)
        o = dict(x=np.zeros((cap, 3)), v=np.zeros((cap, 3)), omega=np.zeros((cap, 3)), radius=np.zeros(cap),
                 rmass=np.zeros(cap), type=np.zeros(cap, np.int32), brid=np.zeros(cap))
        if self.stream:
            o["release"] = np.zeros((cap, 14))
        nb = C.c_int(0)
        k = self.L.ih_insert(self.h, int(step), len(r), _dp(x) if len(r) else None, _dp(r) if len(r) else None,
                             _dp(m) if len(r) else None, cap, _dp(o["x"]), _dp(o["v"]), _dp(o["omega"]),
                             _dp(o["radius"]), _dp(o["rmass"]), o["type"].ctypes.data_as(C.POINTER(C.c_int)),
                             _dp(o["brid"]), C.byref(nb), _dp(o["release"]) if self.stream else None, float(maxrad))
        assert k >= 0
        self.batches = nb.value
        return {key: val[:k].copy() for key, val in o.items()}

    def vector(self):
        out = np.zeros(6)
        self.L.ih_vector(self.h, _dp(out))
        return out

    def next_step(self):
        return self.L.ih_next_step(self.h)

    def close(self):
        self.L.ih_destroy(self.h)


def same(a, b):
    assert set(a) == set(b)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


CASES = {
    "packing": (PACKING, None, (0, 0, 0)),
    "packing_no_all_in_uniform_vel": (dict(PACKING, all_in=False, vel=("uniform", (0.0, 0.0, -1.0), (0.1, 0.1, 0.2)),
                                           omega=(1.0, 2.0, 3.0)), None, (0, 0, 0)),
    "gaussian_vel_numberbased_three_templates": (dict(
        PACKING, templates=PACKING["templates"] + [dict(type=2, radius=0.0065, density=1500.0)],
        weights=[0.5, 0.2, 0.3], number_based=True, vel=("gaussian", (0.0, 0.0, -0.5), (0.05, 0.05, 0.05))),
        None, (0, 0, 0)),
    "dense_hits_maxtry": (dict(PACKING, volumefraction_region=0.55, maxattempt=3), None, (0, 0, 0)),
    "residents_periodic": (dict(PACKING, volumefraction_region=0.35, all_in=False), "bed", (1, 1, 0)),
    "mass_target_uncorrelated": (dict({k: v for k, v in PACKING.items() if k != "volumefraction_region"},
                                      mass_in_region=1.2, random_distribute="uncorrelated"), "bed", (0, 0, 0)),
    "no_overlapcheck": (dict(PACKING, overlapcheck=False, volumefraction_region=0.1), None, (0, 0, 0)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_loop_equals_oracle(name):
    spec, res, periodic = CASES[name]
    if res == "bed":
        x, r, m = resident_bed()
    else:
        x, r, m = np.zeros((0, 3)), np.zeros(0), np.zeros(0)
    o = oracle_fix(spec, periodic=periodic)
    h = HostLoop(spec, periodic=periodic)
    a = o.insert(1, x, r, m)
    b = h.insert(1, x, r, m)
    same(a, b)
    np.testing.assert_array_equal(o.vector(), h.vector())
    assert len(a["radius"]) > 0
    h.close()


def test_host_loop_repeated_insertions():
    """insert_every: each insertion sees the particles of the earlier ones"""
    spec = dict(PACKING, volumefraction_region=0.12, insert_every=100, all_in=False,
                vel=("uniform", (0.0, 0.0, -0.5), (0.1, 0.1, 0.1)))
    o = oracle_fix(spec)
    h = HostLoop(spec)
    x, r, m = np.zeros((0, 3)), np.zeros(0), np.zeros(0)
    step = 1
    for _ in range(4):
        assert o.next_step() == h.next_step() == step
        a = o.insert(step, x, r, m)
        b = h.insert(step, x, r, m)
        same(a, b)
        # the placed particles fall by 2 cm before the next insertion
        x = np.concatenate([x, a["x"] - [0.0, 0.0, 0.02]])
        r = np.concatenate([r, a["radius"]])
        m = np.concatenate([m, a["rmass"]])
        step += 100
    np.testing.assert_array_equal(o.vector(), h.vector())
    h.close()


# ---- clumps: particletemplate/multiplespheres in the distribution

CLUMP3 = (np.array([[-0.006, 0.0, 0.0], [0.0, 0.0, 0.0], [0.006, 0.0, 0.0]]), np.full(3, 0.003))


def clump_spec(periodic_res=False):
    spec = dict(PACKING, volumefraction_region=0.2, all_in=False, orientation="random",
                templates=[dict(type=1, radius=0.0, density=0.0), dict(type=1, radius=0.004, density=2500.0)],
                weights=[0.6, 0.4], vel=("uniform", (0.0, 0.0, -0.5), (0.1, 0.1, 0.1)))
    return spec


def test_clump_insertion_host_loop_equals_oracle():
    """a bonded 3-sphere clump template (the particle3.txt of the reference's
    cohesive_bond examples) with a sphere template, random orientations,
    residents: the host loop equals the oracle bit for bit, bond ids
    included; every clump keeps the template's shape; the first clump's
    bond id is the insertion step plus the template stream's first draw"""
    from oracle import TemplateOracle
    from test_template import host_template

    x3, r3 = CLUMP3
    seed = 999671
    to = TemplateOracle(seed, x3, r3, ntry=100000, atom_type=1, density=2500.0, bonded=True)
    L, th, _ = host_template(seed, x3, r3, ntry=100000, bonded=True)
    xr, rr, mr = resident_bed(n=150, seed=9)
    spec = clump_spec()
    o = oracle_fix(dict(spec, multisphere_oracle=[to, None]))
    h = HostLoop(dict(spec, multisphere=[1, 0], multisphere_host=[th, None]))
    a = o.insert(1, xr, rr, mr)
    b = h.insert(1, xr, rr, mr)
    same(a, b)
    np.testing.assert_array_equal(o.vector(), h.vector())
    nclump = int(np.sum(a["brid"] >= 0)) // 3
    assert nclump > 20 and np.sum(a["brid"] < 0) > 20
    # clumps come first (larger volume), three spheres each, one id per clump
    ids = a["brid"][:3 * nclump].reshape(-1, 3)
    assert (ids == ids[:, :1]).all() and len(np.unique(ids[:, 0])) == nclump
    from oracle import park_states
    assert ids[0, 0] == 1.0 + park_states(seed, 1)[1][0]
    xs = a["x"][:3 * nclump].reshape(-1, 3, 3)
    d01 = np.linalg.norm(xs[:, 0] - xs[:, 1], axis=1)
    d02 = np.linalg.norm(xs[:, 0] - xs[:, 2], axis=1)
    np.testing.assert_allclose(d01, 0.006, rtol=1e-12)
    np.testing.assert_allclose(d02, 0.012, rtol=1e-12)
    np.testing.assert_array_equal(a["rmass"][:3], 4.18879020479 * 0.003 ** 3 * 2500.0)
    assert no_overlap(a["x"][3 * nclump:], a["radius"][3 * nclump:], xr, rr)
    L.it_destroy(th)
    h.close()
