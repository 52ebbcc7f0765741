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


def fibres(nchains=500, nsph=40, seed=4, L=0.04, r=0.5e-3, plate_gap=-2e-6):
    """Bonded-sphere fibres under a moving plate: nchains straight chains of
    nsph touching spheres (r = 0.5 mm, rho 2500) in an L-sized cube, laid as
    a fibre mat: layers 1.4 d apart in z, each layer's chains parallel to x
    or to y (chosen per layer from the seed) at 1.5 d pitch, each chain with
    its own small random tilt (up to 3 mrad in the layer, 1 mrad out of it)
    and shift along its axis.  No two chains come closer than 1.15 d, so bonds
    form only along chains.  Pair chain hysteretic/nonlinear1 + tangential
    history + bond/nonlinear, bonds created at step 1 within 1.1 d; tension
    stiffness K_fn2 = 1e10 N/m^3 (k A d), unloading 2x and reloading 0.5x
    (the yield hysteresis), break beyond 1.05 d0 (5 % tensile strain); the
    compression branch is K_fn1 |d|^(1/2) with K_fn1 = 8 N/m^(1/2).  Hertz
    walls: a fixed floor plane and a plate (two triangles) moving down at
    0.05 m/s, starting plate_gap above the highest sphere's top (negative:
    pressing).  dt = 1e-7, gravity."""
    rng = np.random.default_rng(seed)
    d = 2 * r
    span = (nsph - 1) * d
    pitch = 1.5 * d
    per_layer = int((L - 2 * r - 0.001) // pitch)
    pts = []
    layer = 0
    while len(pts) < nchains:
        axis = int(rng.integers(0, 2))
        z = r + 0.0005 + layer * 1.4 * d
        if z + r > L:
            raise RuntimeError("fibres: the mat does not fit the box")
        for c in range(per_layer):
            if len(pts) == nchains:
                break
            u = np.zeros(3)
            u[axis] = 1.0
            u[1 - axis] = rng.uniform(-3e-3, 3e-3)
            u[2] = rng.uniform(-1e-3, 1e-3)
            u /= np.linalg.norm(u)
            centre = np.zeros(3)
            centre[axis] = L / 2 + rng.uniform(-0.2, 0.2) * (L - span - 2 * r) / 2
            centre[1 - axis] = r + 0.0005 + c * pitch + 0.5 * pitch
            centre[2] = z + span / 2 * abs(u[2]) + 1e-5
            pts.append(centre + np.outer(np.arange(nsph) - (nsph - 1) / 2.0, u) * d)
        layer += 1
    x = np.concatenate(pts)
    n = len(x)
    rad = np.full(n, r)
    atoms = _atoms(x, rad, rng, 0.0)
    w = 3
    full = lambda v: np.full((w, w), v)  # noqa: E731
    mats = _materials(Y=5e6, nu=0.3, e=0.5, mu=0.5, ntypes=2)
    mats["hysteretic"] = {k: full(v) for k, v in dict(
        LoadingStiffness=5e4, UnloadingStiffness=2.5, coefficientAdhesionStiffness=0.0,
        coefficientPlasticityDepth=0.1, pullOffForce=0.0, alphaCustom=20.0, cinCustom=1e-7,
        aoneCustom=10e8, atwoCustom=4e4, athreeCustom=8.0, kcinCustom=1e-4).items()}
    kt = 1e10
    bond = dict(radiusMultiplier=full(1.0), createDistance=full(1.1 * d), maxDistance=full(1.05 * d),
                maxSigma=full(1e12), maxTau=full(1e12), ratioTensionCompression=full(1.0),
                dampingNormalForce=full(0.0), dampingTangentialForce=full(0.0), dampingNormalTorque=full(0.0),
                dampingTangentialTorque=full(0.0),
                K_fn1=full(8.0), Ku_fn1=full(16.0), Kc_fn1=full(4.0),
                K_fn2=full(kt), Ku_fn2=full(2 * kt), Kc_fn2=full(0.5 * kt),
                K_ft=full(0.4 * kt), K_tn=full(0.4 * kt), Ku_tn=full(0.8 * kt), Kc_tn=full(0.2 * kt),
                K_tt=full(kt), Ku_tt=full(2 * kt), Kc_tt=full(0.5 * kt))
    ztop = x[:, 2].max() + r + plate_gap
    plate = np.array([[[0.0, 0.0, ztop], [L, 0.0, ztop], [L, L, ztop]],
                      [[0.0, 0.0, ztop], [L, L, ztop], [0.0, L, ztop]]])
    spec = dict(
        box_lo=[0.0, 0.0, 0.0], box_hi=[L, L, L + 0.005], periodic=[0, 0, 0], dt=1e-7, skin=2e-4, cdf=1.0,
        neigh_every=1, neigh_delay=0, neigh_check=True, rolling="off",
        normal="hysteretic/nonlinear1", wall_normal="hertz", cohesion="bond/nonlinear",
        materials=mats, planes=[(2, 0.0, 2)],
        meshes=[dict(nodes=plate, type=2, vel=(0.0, 0.0, -0.05))],
        bond=dict(tables=bond, ts_create=1.0),
        gravity=(0.0, 0.0, -9.81),
    )
    return spec, atoms
