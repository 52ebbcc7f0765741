"""fix particletemplate/multiplespheres and create_particles single: the
oracle pinned by closed forms and by the reference's microstructure gold, and
the engine's host template code checked against the oracle on the CPU.

* a clump of two touching spheres: the bounding sphere is the one through
  both (radius 2r, centre at the contact point) whatever order the
  randomised construction takes; the clump is placed with its bounding
  sphere's centre at the given point, so the spheres sit at -r and +r (the
  2-sphere tension example's gold starts from there), both to the rounding
  of the construction; the
  Monte Carlo centre of mass and volume lie within their sampling error of
  the contact point and of two sphere volumes; an orientation quaternion
  rotates the clump as the rotation it stands for;
* microstructure_compression's 114,704-sphere template (seed 900019, the
  second line of its prime_numbers.txt after `next gy_prime`; ntry 100,000)
  placed with create_particles at (49, 49, 99): sphere 1 lands where the
  reference's gold prints it, to all 15 printed digits.  That position is
  the file's (2, 2, 2) minus the bounding sphere's centre, both shifted by
  the Monte Carlo centre of mass: it pins the 200 randomised bounding-sphere
  passes, the centre-of-mass sampling and the order of the stream's draws;
* the engine's host code (insert_pack.cpp, through
  tests/insert_host_driver.cpp with a brute-force inside test) gives the
  oracle's template and placement bit for bit.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT, load_engine
from test_insert import driver

GOLD_SPHERE1 = (0.842052335551081, 0.412012706847342, 0.620945518938754)


def tmpl_oracle(seed, x, r, scale=1.0, ntry=100000):
    from oracle import TemplateOracle

    return TemplateOracle(seed, x, r, scale, ntry)


def microstructure_template():
    d = np.load(os.path.join(ROOT, "tests", "golden", "reference_data", "microstructure_vox_1_occupancy.npz"))
    shape = tuple(int(v) for v in d["shape"])
    g = np.unpackbits(d["occ"])[:np.prod(shape)].reshape(shape).astype(bool)
    kji = np.argwhere(g)
    return np.stack([2.0 * kji[:, 2] + 2.0, 2.0 * kji[:, 1] + 2.0, 2.0 * kji[:, 0] + 2.0], 1)


def test_two_sphere_clump_closed_forms():
    r0 = 0.0002
    x = [[0.0, 0.0, 0.0], [0.0004, 0.0, 0.0]]
    ntry = 1000000
    t = tmpl_oracle(900007, x, [r0, r0], ntry=ntry)
    info = t.info()
    assert abs(info[0] - 0.0004) < 1e-19            # to the rounding of the construction
    np.testing.assert_allclose(info[1:4] + info[6:9], [0.0002, 0.0, 0.0], rtol=0, atol=1e-19)
    pos = t.place([0.0, 0.0, 0.0])
    np.testing.assert_allclose(pos, [[-0.0002, 0.0, 0.0], [0.0002, 0.0, 0.0]], rtol=0, atol=1e-19)
    vol = 2 * 4.0 / 3.0 * np.pi * r0 ** 3
    box = 0.0008 * 0.0004 * 0.0004
    p = vol / box
    assert abs(info[4] - vol) < 4 * np.sqrt(p * (1 - p) / ntry) * box
    n_in = p * ntry
    sig = r0 / np.sqrt(5.0 * n_in)
    assert abs(info[6] - 0.0002) < 4 * np.sqrt(2.0) * sig and np.abs(info[7:9]).max() < 4 * sig
    # orientation (0.7071, 0, 0.7071, 0), normalised: 90 degrees about y,
    # x -> -z
    q = t.place([0.01, 0.02, 0.03], [0.7071, 0.0, 0.7071, 0.0])
    np.testing.assert_allclose(q, [[0.01, 0.02, 0.03 + 0.0002], [0.01, 0.02, 0.03 - 0.0002]], rtol=0, atol=1e-18)


def test_microstructure_template_reproduces_the_gold():
    x = microstructure_template()
    t = tmpl_oracle(900019, x, np.ones(len(x)), ntry=100000)
    pos = t.place([49.0, 49.0, 99.0])
    print("sphere 1 at %r, gold %r" % (tuple(pos[0]), GOLD_SPHERE1))
    assert np.abs(pos[0] - np.array(GOLD_SPHERE1)).max() < 1e-15
    # the clump keeps the lattice: every sphere moved by the same offset
    # (within the rounding of the two shifts)
    off = pos - x
    assert np.abs(off - off[0]).max() < 1e-13


# ---- the engine's host template code against the oracle (CPU)

def host_template(seed, x, r, types=None, scale=1.0, ntry=100000, bonded=False):
    L = driver()
    eng = load_engine()
    L.it_create.restype = C.c_void_p
    L.it_create.argtypes = [C.POINTER(eng.engine.TemplateMultisphereSpec), C.c_char_p, C.c_int]
    L.it_destroy.argtypes = [C.c_void_p]
    L.it_info.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.it_place.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    s = eng.engine.TemplateMultisphereSpec()
    s.seed, s.atom_type, s.density, s.nspheres = seed, 1, 2500.0, len(rr)
    s.x = xx.ctypes.data_as(C.POINTER(C.c_double))
    s.r = rr.ctypes.data_as(C.POINTER(C.c_double))
    s.types = None
    s.scale, s.ntry, s.bonded = scale, ntry, int(bonded)
    err = C.create_string_buffer(256)
    h = L.it_create(C.byref(s), err, 256)
    assert h, err.value
    return L, h, len(rr)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize("case", ["two", "cluster"])
def test_host_template_equals_oracle(case):
    if case == "two":
        x, r, scale = np.array([[0.0, 0.0, 0.0], [0.0004, 0.0, 0.0]]), np.array([0.0002, 0.0002]), 1.0
    else:
        rng = np.random.default_rng(11)
        x = rng.normal(0.0, 1.0, size=(60, 3))
        r = rng.uniform(0.4, 0.9, size=60)
        scale = 0.37
    o = tmpl_oracle(15485863, x, r, scale=scale, ntry=200000)
    L, h, n = host_template(15485863, x, r, scale=scale, ntry=200000)
    info = np.zeros(9)
    L.it_info(h, _dp(info))
    oi = o.info()
    np.testing.assert_array_equal(info[[0, 1, 2, 3, 4, 6, 7, 8]], oi[[0, 1, 2, 3, 4, 6, 7, 8]])
    for quat in (None, np.array([0.3, -0.2, 0.9, 0.1])):
        pos = np.zeros((n, 3))
        xc = np.array([0.5, -1.5, 2.25])
        L.it_place(h, _dp(xc), None if quat is None else _dp(quat), _dp(pos))
        np.testing.assert_array_equal(pos, o.place(xc, quat))
    L.it_destroy(h)
