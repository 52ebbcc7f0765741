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
* ``hopper`` -- N spheres r ~ U[0.8, 1.2] mm in a 30-degree conical hopper
  (100 mm top, 12 mm outlet) with a cylinder on top, ~20,000-30,000
  triangles, mu = 0.4, e = 0.5, dt = 5e-7, floor plane below the outlet.
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


def hopper_mesh(r_out=6e-3, r_top=50e-3, half_angle_deg=30.0, cyl_height=0.1, ntheta=160,
                ring_len=2.2e-3, cyl_ring_len=12e-3):
    """Conical hopper (apex down, outlet open) with a cylinder on top, as a
    triangle list (ntri, 3, 3).  Rings of ntheta quads, two triangles each;
    node order gives normals pointing into the hopper."""
    t = np.tan(np.radians(half_angle_deg))
    hc = (r_top - r_out) / t
    ncone = max(1, int(round(hc / np.cos(np.radians(half_angle_deg)) / ring_len)))
    ncyl = max(1, int(round(cyl_height / cyl_ring_len))) if cyl_height > 0 else 0
    zs = list(np.linspace(0.0, hc, ncone + 1))
    rs = [r_out + z * t for z in zs]
    for k in range(1, ncyl + 1):
        zs.append(hc + cyl_height * k / ncyl)
        rs.append(r_top)
    th = np.linspace(0.0, 2 * np.pi, ntheta + 1)[:-1]
    ring = [np.stack([r * np.cos(th), r * np.sin(th), np.full(ntheta, z)], 1) for z, r in zip(zs, rs)]
    tris = []
    for k in range(len(zs) - 1):
        a, b = ring[k], ring[k + 1]
        for j in range(ntheta):
            j1 = (j + 1) % ntheta
            tris.append([a[j], b[j], a[j1]])
            tris.append([a[j1], b[j], b[j1]])
    return np.asarray(tris), hc


def hopper(n=200_000, seed=7, rmin=0.8e-3, rmax=1.2e-3, spacing=2.5e-3, jitter=5e-5, ntheta=160,
           mu=0.4, e=0.5, Y=5e6, dt=5e-7, skin=2.5e-4, wall_overlap=0.0):
    """Hopper discharge: n spheres r ~ U[rmin, rmax] on a jittered lattice
    inside a 30-degree cone (100 mm top, 12 mm outlet) topped by a cylinder
    tall enough to hold them, ~20,000 triangles, a floor plane 40 mm below
    the outlet, gravity.  wall_overlap > 0 moves the outermost lattice
    sites that far into the wall (contacts from the first step)."""
    rng = np.random.default_rng(seed)
    r_out, r_top, ang = 6e-3, 50e-3, 30.0
    t = np.tan(np.radians(ang))
    hc = (r_top - r_out) / t
    # lattice sites in the interior, bottom up, until n fit
    lim = lambda z: np.minimum(r_out + np.maximum(z, 0.0) * t, r_top)  # noqa: E731
    c = np.cos(np.radians(ang))
    nxy = int(2 * r_top / spacing) + 2
    g = (np.arange(nxy) - nxy / 2 + 0.5) * spacing
    gx, gy = np.meshgrid(g, g, indexing="ij")
    gx, gy = gx.ravel(), gy.ravel()
    rho_xy = np.hypot(gx, gy)
    sites = []
    z = rmax + 2e-3
    count = 0
    while count < n:
        room = (lim(z) - rho_xy) * c if z < hc else (r_top - rho_xy)
        ok = room >= rmax + jitter - wall_overlap
        layer = np.stack([gx[ok], gy[ok], np.full(ok.sum(), z)], 1)
        sites.append(layer)
        count += len(layer)
        z += spacing
    x = np.concatenate(sites)[:n]
    x = x + rng.uniform(-jitter, jitter, size=x.shape)
    r = rng.uniform(rmin, rmax, size=n)
    top = x[:, 2].max() + rmax
    cyl = max(top - hc + 5e-3, 5e-3)
    tris, _ = hopper_mesh(r_out, r_top, ang, cyl, ntheta)
    mats = _materials(Y=Y, e=e, mu=mu, ntypes=2)
    ztop = hc + cyl
    spec = dict(
        box_lo=[-0.06, -0.06, -0.045], box_hi=[0.06, 0.06, ztop + 0.01], periodic=[0, 0, 0], dt=dt, skin=skin,
        cdf=1.0, neigh_every=10, neigh_delay=0, neigh_check=False, rolling="off", materials=mats,
        meshes=[dict(nodes=tris, type=2)], planes=[(2, -0.04, 2)], gravity=(0.0, 0.0, -9.81),
    )
    return spec, _atoms(x, r, rng, 0.0)


def batch_pours(ninst=256, seed=5, nmin=500, nmax=8000, rmin=0.8e-3, rmax=1.2e-3, spacing=2.5e-3,
                jitter=5e-5):
    """Independent small pours for a batched run: instance b holds N_b ~
    U[nmin, nmax] spheres, r ~ U[rmin, rmax], on a jittered lattice in the
    same 40x40x100 mm box with 6 planes, v ~ U[-0.1, 0.1] m/s, Hertz +
    history, dt = 1e-6, skin 0.25 mm, rebuild every 10 steps, gravity; its
    own Y ~ U[1e6, 1e7] Pa, mu ~ U[0.2, 0.7], e ~ U[0.3, 0.9] (nu = 0.3).

    Returns (spec, atoms, inst, params): atoms of all instances concatenated
    (tags unique over the batch), inst = instance of each atom, params =
    dict(Y, e, mu) per instance.  ``instance(spec, atoms, inst, params, b)``
    gives one instance as a stand-alone (spec, atoms)."""
    rng = np.random.default_rng(seed)
    L = np.array([0.04, 0.04, 0.10])
    counts = rng.integers(nmin, nmax + 1, size=ninst)
    Y = rng.uniform(1e6, 1e7, size=ninst)
    mu = rng.uniform(0.2, 0.7, size=ninst)
    e = rng.uniform(0.3, 0.9, size=ninst)
    nx = int((L[0] - 2 * rmax - 2 * jitter) // spacing) + 1
    ny = int((L[1] - 2 * rmax - 2 * jitter) // spacing) + 1
    xs, rs, vs, ins = [], [], [], []
    for b in range(ninst):
        n = int(counts[b])
        nz = int(np.ceil(n / (nx * ny)))
        ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        sites = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1)
        sites = sites[np.lexsort((sites[:, 0], sites[:, 1], sites[:, 2]))][:n]
        xs.append((rmax + jitter) + sites * spacing + rng.uniform(-jitter, jitter, size=(n, 3)))
        rs.append(rng.uniform(rmin, rmax, size=n))
        vs.append(rng.uniform(-0.1, 0.1, size=(n, 3)))
        ins.append(np.full(n, b, np.int32))
    x = np.concatenate(xs)
    r = np.concatenate(rs)
    if x[:, 2].max() + rmax > L[2]:
        raise ValueError("batch_pours: lattice taller than the box")
    atoms = _atoms(x, r, rng, 0.0)
    atoms["v"] = np.concatenate(vs)
    spec = dict(
        box_lo=[0.0, 0.0, 0.0], box_hi=list(L), periodic=[0, 0, 0], dt=1e-6, skin=2.5e-4, cdf=1.0,
        neigh_every=10, neigh_delay=0, neigh_check=False, rolling="off",
        materials=_materials(Y=5e6, e=0.6, mu=0.45),
        planes=[(0, 0.0, 1), (0, L[0], 1), (1, 0.0, 1), (1, L[1], 1), (2, 0.0, 1), (2, L[2], 1)],
        gravity=(0.0, 0.0, -9.81),
    )
    return spec, atoms, np.concatenate(ins), dict(Y=Y, e=e, mu=mu)


def instance(spec, atoms, inst, params, b):
    """Instance b of a batch as a stand-alone (spec, atoms): its atoms (tags
    kept) and its own materials."""
    k = inst == b
    a = {key: np.ascontiguousarray(v[k]) for key, v in atoms.items()}
    m = dict(spec["materials"])
    w = m["ntypes"] + 1
    m["Y"] = np.full(w, float(params["Y"][b]))
    m["e"] = np.full((w, w), float(params["e"][b]))
    m["mu"] = np.full((w, w), float(params["mu"][b]))
    s = dict(spec, materials=m)
    s.pop("instances", None)
    return s, a
