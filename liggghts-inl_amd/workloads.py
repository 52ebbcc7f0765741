"""Seeded synthetic inputs for the benchmark and parity cases.

Every generator is deterministic in its seed and returns (spec, atoms) in
the form ``engine.Simulation`` and the oracle both accept, so the GPU run and
the CPU checker start from bit-identical state (the role `read_data` files
play for the reference).  Units are SI.

* ``pour_box``   -- 4,000 monodisperse r = 1 mm spheres in a 40x40x80 mm box
  with 6 planes, jittered lattice, v ~ U[-0.1, 0.1] m/s, Hertz + history,
  Y = 5e6, nu = 0.3, e = 0.9, mu = 0.5, dt = 1e-6, skin 0.25 mm, rebuild
  every 10 steps.
* ``packed_column`` -- N polydisperse spheres, r ~ U[0.5, 1.5] mm, periodic
  60x60 mm in x/y, bottom plane, packing fraction 0.60, Hertz + history +
  EPSD rolling (mu_r = 0.1), dt = 5e-7.  The packing is a periodic slab
  relaxed by overlap removal at 0.999 of the radii and stacked in z, so at
  full radii neighbouring spheres overlap by ~1e-3 r: a dense, contact-rich
  state, not a gravity-equilibrated one.
"""
import numpy as np

RHO = 2500.0


def _materials(Y=5e6, nu=0.3, e=0.9, mu=0.5, mur=0.0, etar=0.0, krs=0.0, ntypes=1):
    w = ntypes + 1
    return dict(ntypes=ntypes, Y=np.full(w, Y), nu=np.full(w, nu), e=np.full((w, w), e),
                mu=np.full((w, w), mu), mur=np.full((w, w), mur), etar=np.full((w, w), etar), krs=krs)


def _atoms(x, r, rng, vmax, rho=RHO):
    n = len(r)
    return dict(
        x=np.ascontiguousarray(x, dtype=np.float64),
        v=rng.uniform(-vmax, vmax, size=(n, 3)) if vmax > 0 else np.zeros((n, 3)),
        omega=np.zeros((n, 3)),
        radius=np.ascontiguousarray(r, dtype=np.float64),
        rmass=4.0 / 3.0 * np.pi * r ** 3 * rho,
        type=np.ones(n, np.int32),
        tag=np.arange(1, n + 1, dtype=np.int32),
    )


def pour_box(n=4000, seed=12345, spacing=2.2e-3, jitter=5e-5):
    """Config of the reference's CPU-runnable pour (4,000 spheres, 6 planes)."""
    rng = np.random.default_rng(seed)
    r0 = 1e-3
    L = np.array([0.04, 0.04, 0.08])
    nx = int((L[0] - 2 * r0 - 2 * jitter) // spacing) + 1
    ny = int((L[1] - 2 * r0 - 2 * jitter) // spacing) + 1
    nz = int(np.ceil(n / (nx * ny)))
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sites = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1)
    sites = sites[np.lexsort((sites[:, 0], sites[:, 1], sites[:, 2]))][:n]
    x = (r0 + jitter) + sites * spacing + rng.uniform(-jitter, jitter, size=(n, 3))
    r = np.full(n, r0)
    spec = dict(
        box_lo=[0.0, 0.0, 0.0], box_hi=list(L), periodic=[0, 0, 0], dt=1e-6, skin=2.5e-4, cdf=1.0,
        neigh_every=10, neigh_delay=0, neigh_check=False, rolling="off",
        materials=_materials(),
        planes=[(0, 0.0, 1), (0, L[0], 1), (1, 0.0, 1), (1, L[1], 1), (2, 0.0, 1), (2, L[2], 1)],
        gravity=(0.0, 0.0, -9.81),
    )
    return spec, _atoms(x, r, rng, 0.1)


def _relax_slab(r, L, H, rng, shrink=0.999, iters=3000, tol=1e-3):
    """Random periodic (x, y, z) packing of spheres with radii r*shrink and
    no overlap: grow from 0.55x size while removing overlaps pairwise
    (csrc/packing.c, host code)."""
    import ctypes as C
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdem_workloads.so")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    lib.dem_relax_packing.restype = C.c_int
    lib.dem_relax_packing.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_int, C.c_int, C.c_double]
    n = len(r)
    box = np.array([L, L, H], dtype=np.float64)
    pos = np.ascontiguousarray(rng.uniform(0, 1, size=(n, 3)) * box)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    lib.dem_relax_packing(n, pos.ctypes.data_as(dp), rr.ctypes.data_as(dp), box.ctypes.data_as(dp),
                          shrink, iters, 250, tol)
    return pos


def packed_column(n=1_000_000, seed=2024, slab_n=12500, phi=0.60, rmin=0.5e-3, rmax=1.5e-3,
                  Lxy=0.06, skin=2.5e-4, mur=0.1, etar=0.3, rolling="epsd"):
    """Dense polydisperse column, periodic x/y, bottom plane."""
    rng = np.random.default_rng(seed)
    slab_n = min(slab_n, n)
    r_slab = rng.uniform(rmin, rmax, size=slab_n)
    vol = (4.0 / 3.0 * np.pi * r_slab ** 3).sum()
    H = vol / (phi * Lxy * Lxy)
    pos = _relax_slab(r_slab, Lxy, H, rng)
    ntile = int(np.ceil(n / slab_n))
    xs = np.concatenate([pos + np.array([0, 0, k * H]) for k in range(ntile)])
    rs = np.tile(r_slab, ntile)
    order = np.argsort(xs[:, 2], kind="stable")[:n]
    xs, rs = xs[order], rs[order]
    xs[:, 2] += -(xs[:, 2] - rs).min()
    top = (xs[:, 2] + rs).max()
    spec = dict(
        box_lo=[0.0, 0.0, -0.001], box_hi=[Lxy, Lxy, top * 1.25 + 0.01], periodic=[1, 1, 0], dt=5e-7,
        skin=skin, cdf=1.0, neigh_every=10, neigh_delay=0, neigh_check=False, rolling=rolling,
        materials=_materials(mur=mur, etar=etar),
        planes=[(2, 0.0, 1)],
        gravity=(0.0, 0.0, -9.81),
    )
    return spec, _atoms(xs, rs, rng, 0.0)
