"""ctypes binding of the engine's C-ABI (include/liggghts_dem.h).

This mirrors what the LIGGGHTS style plug-in does from C++: set up a
context from the script's pair/fix/neighbor settings, upload the Atom
arrays, run Verlet steps on the device and read results back.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBNAME = "libliggghts_dem.so"
_LIB = None

DEM_FP64 = 0
DEM_MIXED = 1
ROLL = {"off": 0, "epsd": 1, "epsd2": 2, "epsd3": 3, "cdtnonlinear2": 4}


class DemError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("boxlo", C.c_double * 3),
        ("boxhi", C.c_double * 3),
        ("periodic", C.c_int * 3),
        ("skin", C.c_double),
        ("cdf", C.c_double),
        ("neigh_every", C.c_int),
        ("neigh_delay", C.c_int),
        ("neigh_check", C.c_int),
        ("neigh_one", C.c_int),
        ("rolling_model", C.c_int),
        ("tangential_damping", C.c_int),
        ("limit_force", C.c_int),
        ("torsion_torque", C.c_int),
        ("precision", C.c_int),
        ("device", C.c_int),
        ("normal_model", C.c_int),
        ("wall_normal_model", C.c_int),
        ("wall_limit_force", C.c_int),
        ("cohesion_model", C.c_int),
    ]


class BondSettings(C.Structure):
    _fields_ = [
        ("ts_create", C.c_double),
        ("stress_break", C.c_int),
        ("create_always", C.c_int),
        ("damping", C.c_int),
        ("damping_smooth", C.c_int),
        ("ratio_tc", C.c_int),
        ("tension_variant", C.c_int),
    ]


class TemplateSphere(C.Structure):
    _fields_ = [("atom_type", C.c_int), ("radius", C.c_double), ("density", C.c_double)]


class InsertPackSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_int),
        ("dist_seed", C.c_int),
        ("ntemplates", C.c_int),
        ("templates", C.POINTER(TemplateSphere)),
        ("weights", C.POINTER(C.c_double)),
        ("number_based", C.c_int),
        ("region_lo", C.c_double * 3),
        ("region_hi", C.c_double * 3),
        ("target", C.c_int),
        ("target_value", C.c_double),
        ("maxattempt", C.c_int),
        ("insert_every", C.c_int),
        ("start", C.c_longlong),
        ("overlapcheck", C.c_int),
        ("all_in", C.c_int),
        ("check_dist_from_subdomain_border", C.c_int),
        ("ntry_mc", C.c_int),
        ("exact_number", C.c_int),
        ("vel_style", C.c_int),
        ("vel", C.c_double * 3),
        ("vel_fluct", C.c_double * 3),
        ("omega", C.c_double * 3),
        ("group_mask", C.c_int),
        ("multisphere", C.POINTER(C.c_int)),
        ("orientation", C.c_int),
        ("quat", C.c_double * 4),
    ]


class InsertStreamSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_int),
        ("dist_seed", C.c_int),
        ("ntemplates", C.c_int),
        ("templates", C.POINTER(TemplateSphere)),
        ("weights", C.POINTER(C.c_double)),
        ("number_based", C.c_int),
        ("multisphere", C.POINTER(C.c_int)),
        ("face_ntri", C.c_int),
        ("face_nodes", C.POINTER(C.c_double)),
        ("extrude_length", C.c_double),
        ("insert_every", C.c_int),
        ("duration", C.c_int),
        ("start", C.c_longlong),
        ("nparticles", C.c_int),
        ("mass", C.c_double),
        ("particlerate", C.c_double),
        ("massrate", C.c_double),
        ("maxattempt", C.c_int),
        ("overlapcheck", C.c_int),
        ("all_in", C.c_int),
        ("parallel", C.c_int),
        ("ntry_mc", C.c_int),
        ("exact_number", C.c_int),
        ("vel_style", C.c_int),
        ("vel", C.c_double * 3),
        ("vel_fluct", C.c_double * 3),
        ("omega", C.c_double * 3),
        ("group_mask", C.c_int),
        ("orientation", C.c_int),
        ("quat", C.c_double * 4),
    ]


class TemplateMultisphereSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_int),
        ("atom_type", C.c_int),
        ("density", C.c_double),
        ("nspheres", C.c_int),
        ("x", C.POINTER(C.c_double)),
        ("r", C.POINTER(C.c_double)),
        ("types", C.POINTER(C.c_int)),
        ("scale", C.c_double),
        ("ntry", C.c_int),
        ("bonded", C.c_int),
    ]


INSERT_TARGET = {"volumefraction_region": 0, "particles_in_region": 1, "mass_in_region": 2}
VEL_STYLE = {"constant": 0, "uniform": 1, "gaussian": 2}


def insert_pack_spec(d):
    """dem_insert_pack_spec from a dict of the script's keywords:

    seed, dist_seed, templates [{type, radius, density}], weights,
    number_based, region (lo, hi), one of volumefraction_region /
    particles_in_region / mass_in_region, maxattempt, insert_every ("once"
    or steps), start, overlapcheck, all_in, check_dist_from_subdomain_border,
    ntry_mc, random_distribute ("exact" / "uncorrelated"), vel ("constant",
    v) / ("uniform" | "gaussian", v, fluct), omega, group_mask, multisphere
    (per template an add_template_multisphere id, 0 for a sphere),
    orientation ("template", "random" or ("constant", q)).

    Returns (spec, keep): keep holds the arrays the spec points to."""
    L = library()
    s = InsertPackSpec()
    L.dem_insert_pack_defaults(C.byref(s))
    tl = d["templates"]
    tarr = (TemplateSphere * len(tl))()
    for k, t in enumerate(tl):
        tarr[k].atom_type = int(t["type"])
        tarr[k].radius = float(t["radius"])
        tarr[k].density = float(t["density"])
    w = np.ascontiguousarray(d.get("weights", [1.0] * len(tl)), dtype=np.float64)
    s.seed = int(d["seed"])
    s.dist_seed = int(d["dist_seed"])
    s.ntemplates = len(tl)
    s.templates = tarr
    s.weights = w.ctypes.data_as(C.POINTER(C.c_double))
    s.number_based = int(d.get("number_based", False))
    lo, hi = d["region"]
    for k in range(3):
        s.region_lo[k] = float(lo[k])
        s.region_hi[k] = float(hi[k])
    targets = [key for key in INSERT_TARGET if key in d]
    if len(targets) != 1:
        raise DemError("must define exactly one keyword out of 'volumefraction_region', 'particles_in_region', "
                       "and 'mass_in_region'")
    s.target = INSERT_TARGET[targets[0]]
    s.target_value = float(d[targets[0]])
    s.maxattempt = int(d.get("maxattempt", 50))
    ev = d.get("insert_every", -1)
    s.insert_every = 0 if ev == "once" else int(ev)
    s.start = int(d.get("start", 0))
    s.overlapcheck = int(d.get("overlapcheck", True))
    s.all_in = int(d.get("all_in", False))
    s.check_dist_from_subdomain_border = int(d.get("check_dist_from_subdomain_border", True))
    s.ntry_mc = int(d.get("ntry_mc", 100000))
    s.exact_number = int(d.get("random_distribute", "exact") == "exact")
    vel = d.get("vel", ("constant", (0.0, 0.0, 0.0)))
    s.vel_style = VEL_STYLE[vel[0]]
    for k in range(3):
        s.vel[k] = float(vel[1][k])
        s.vel_fluct[k] = float(vel[2][k]) if len(vel) > 2 else 0.0
        s.omega[k] = float(d.get("omega", (0.0, 0.0, 0.0))[k])
    s.group_mask = int(d.get("group_mask", 0))
    # multisphere: per template an add_template_multisphere id, 0 for a sphere
    ms = np.ascontiguousarray(d.get("multisphere", [0] * len(tl)), dtype=np.int32)
    s.multisphere = ms.ctypes.data_as(C.POINTER(C.c_int))
    o = d.get("orientation", "template")
    s.orientation = {"template": 0, "random": 1}.get(o, 2) if isinstance(o, str) else 2
    q = o[1] if not isinstance(o, str) else (1.0, 0.0, 0.0, 0.0)
    for k in range(4):
        s.quat[k] = float(q[k])
    return s, (tarr, w, ms)


def insert_stream_spec(d):
    """dem_insert_stream_spec from a dict of the script's keywords:

    seed, dist_seed, templates, weights, number_based, multisphere,
    orientation, vel, omega, group_mask, maxattempt, overlapcheck, all_in,
    ntry_mc, random_distribute, start (as insert_pack_spec);
    insertion_face (triangles, (ntri, 9) or (ntri, 3, 3), scaled and moved
    as fix mesh/surface leaves them), extrude_length or insert_every
    [duration], nparticles (or "INF") or mass (or "INF"), particlerate or
    massrate, parallel.

    Returns (spec, keep): keep holds the arrays the spec points to."""
    L = library()
    s = InsertStreamSpec()
    L.dem_insert_stream_defaults(C.byref(s))
    tl = d["templates"]
    tarr = (TemplateSphere * len(tl))()
    for k, t in enumerate(tl):
        tarr[k].atom_type = int(t["type"])
        tarr[k].radius = float(t["radius"])
        tarr[k].density = float(t["density"])
    w = np.ascontiguousarray(d.get("weights", [1.0] * len(tl)), dtype=np.float64)
    s.seed = int(d["seed"])
    s.dist_seed = int(d["dist_seed"])
    s.ntemplates = len(tl)
    s.templates = tarr
    s.weights = w.ctypes.data_as(C.POINTER(C.c_double))
    s.number_based = int(d.get("number_based", False))
    ms = np.ascontiguousarray(d.get("multisphere", [0] * len(tl)), dtype=np.int32)
    s.multisphere = ms.ctypes.data_as(C.POINTER(C.c_int))
    face = np.ascontiguousarray(d["insertion_face"], dtype=np.float64).reshape(-1, 9)
    s.face_ntri = len(face)
    s.face_nodes = face.ctypes.data_as(C.POINTER(C.c_double))
    s.extrude_length = float(d.get("extrude_length", 0.0))
    s.insert_every = int(d.get("insert_every", -1))
    s.duration = int(d.get("duration", 0))
    s.start = int(d.get("start", 0))
    npart = d.get("nparticles", 0)
    s.nparticles = -1 if npart == "INF" else int(npart)
    mass = d.get("mass", 0.0)
    s.mass = -1.0 if mass == "INF" else float(mass)
    s.particlerate = float(d.get("particlerate", 0.0))
    s.massrate = float(d.get("massrate", 0.0))
    s.maxattempt = int(d.get("maxattempt", 50))
    s.overlapcheck = int(d.get("overlapcheck", True))
    s.all_in = int(d.get("all_in", False))
    s.parallel = int(d.get("parallel", False))
    s.ntry_mc = int(d.get("ntry_mc", 100000))
    s.exact_number = int(d.get("random_distribute", "exact") == "exact")
    vel = d.get("vel", ("constant", (0.0, 0.0, 0.0)))
    s.vel_style = VEL_STYLE[vel[0]]
    for k in range(3):
        s.vel[k] = float(vel[1][k])
        s.vel_fluct[k] = float(vel[2][k]) if len(vel) > 2 else 0.0
        s.omega[k] = float(d.get("omega", (0.0, 0.0, 0.0))[k])
    s.group_mask = int(d.get("group_mask", 0))
    o = d.get("orientation", "template")
    s.orientation = {"template": 0, "random": 1}.get(o, 2) if isinstance(o, str) else 2
    q = o[1] if not isinstance(o, str) else (1.0, 0.0, 0.0, 0.0)
    for k in range(4):
        s.quat[k] = float(q[k])
    return s, (tarr, w, ms, face)


COHESION = {"off": 0, "bond": 1, "bond/nonlinear": 2}
# bond tables in the order of DEM_NBOND (include/liggghts_dem.h); names of the
# script properties, the bond/nonlinear "...Bondnonlinear" / "stiffnessPerUnitArea..."
# spellings mapped onto the same slots
BOND_KEYS = ("radiusMultiplier", "normalBondStiffnessPerUnitArea", "tangentialBondStiffnessPerUnitArea",
             "dampingNormalForce", "dampingTangentialForce", "dampingNormalTorque", "dampingTangentialTorque",
             "maxDistance", "maxSigma", "maxTau", "ratioTensionCompression", "createDistance",
             "K_fn1", "Ku_fn1", "Kc_fn1", "K_fn2", "Ku_fn2", "Kc_fn2", "K_ft", "K_tn", "Ku_tn", "Kc_tn",
             "K_tt", "Ku_tt", "Kc_tt")


NORMAL = {"hertz": 0, "hysteretic/nonlinear1": 1, "hysteretic/nonlinear2": 2}
HYST_KEYS = ("LoadingStiffness", "UnloadingStiffness", "coefficientAdhesionStiffness",
             "coefficientPlasticityDepth", "pullOffForce", "alphaCustom", "cinCustom",
             "aoneCustom", "atwoCustom", "athreeCustom", "kcinCustom")


def library_path():
    return os.path.join(_HERE, _LIBNAME)


def build():
    """Compile csrc/ for gfx950 with its Makefile (hipcc cross-compiles)."""
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc")], check=True)


def library():
    """Load the HIP library; raise if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise DemError("%s is not built; run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    i64p = C.POINTER(C.c_int64)
    vp = C.c_void_p
    sig = {
        "dem_ctx_create": (vp, [C.POINTER(Params)]),
        "dem_ctx_destroy": (None, [vp]),
        "dem_last_error": (C.c_char_p, [vp]),
        "dem_material_tables": (C.c_int, [C.c_int, dp, dp, dp, dp, dp, dp]),
        "dem_set_materials": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, dp, dp, C.c_double]),
        "dem_upload_state": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, dp, ip, ip, ip]),
        "dem_download": (C.c_int, [vp, dp, dp, dp, dp, dp]),
        "dem_add_plane": (C.c_int, [vp, C.c_int, C.c_double, C.c_int]),
        "dem_set_gravity": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "dem_set_freeze": (C.c_int, [vp, C.c_int]),
        "dem_setup": (C.c_int, [vp]),
        "dem_step": (C.c_int, [vp, C.c_int]),
        "dem_compute_forces": (C.c_int, [vp, C.c_int]),
        "dem_rebuild": (C.c_int, [vp]),
        "dem_get_pairs": (C.c_int, [vp, i64p, ip, ip, ip, dp, ip]),
        "dem_history_size": (C.c_int, [vp]),
        "dem_wall_history_size": (C.c_int, [vp]),
        "dem_set_hysteretic": (C.c_int, [vp, dp, dp]),
        "dem_set_history": (C.c_int, [vp, C.c_int64, ip, ip, dp]),
        "dem_get_wall_history": (C.c_int, [vp, dp]),
        "dem_set_wall_history": (C.c_int, [vp, dp]),
        "dem_timing_enable": (C.c_int, [vp, C.c_int]),
        "dem_timing_get": (C.c_int, [vp, C.c_char_p, dp, i64p]),
        "dem_list_stats": (C.c_int, [vp, dp, dp, i64p]),
        "dem_sync": (C.c_int, [vp]),
        "dem_timestep": (C.c_int64, [vp]),
        "dem_set_instances": (C.c_int, [vp, C.c_int, ip]),
        "dem_set_instance_materials": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, dp, dp]),
        "dem_observe_plane": (C.c_int, [vp, C.c_int]),
        "dem_instance_observables": (C.c_int, [vp, dp]),
        "dem_instance_observables_noalloc": (C.c_int, [vp, dp]),
        "dem_use_graph": (C.c_int, [vp, C.c_int]),
        "dem_autotune": (C.c_int, [vp, C.c_int, ip, ip, dp]),
        "dem_adjoint_forward": (C.c_int, [vp, C.c_int, C.c_int]),
        "dem_set_contact_shape": (C.c_int, [vp, C.c_int, C.c_int]),
        "dem_adjoint_backward": (C.c_int, [vp, dp, dp, dp]),
        "dem_add_mesh": (C.c_int, [vp, C.c_int, dp, C.c_int, C.c_double, dp]),
        "dem_mesh_size": (C.c_int, [vp, C.c_int]),
        "dem_mesh_topology": (C.c_int, [C.c_int, dp, C.c_double, dp, dp, ip, ip, ip, ip, ip]),
        "dem_get_mesh_contacts": (C.c_int, [vp, C.c_int, i64p, ip, ip, ip, dp]),
        "dem_set_mesh_contacts": (C.c_int, [vp, C.c_int, C.c_int64, ip, ip, ip, dp]),
        "dem_set_bond": (C.c_int, [vp, dp, C.POINTER(BondSettings)]),
        "dem_set_group": (C.c_int, [vp, C.c_int, ip]),
        "dem_add_setforce": (C.c_int, [vp, C.c_int, ip, dp, ip]),
        "dem_add_aveforce": (C.c_int, [vp, C.c_int, ip, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, ip]),
        "dem_unfix": (C.c_int, [vp, C.c_int]),
        "dem_mesh_stress": (C.c_int, [vp, C.c_int, dp]),
        "dem_write_restart": (C.c_int, [vp, C.c_char_p]),
        "dem_read_restart": (C.c_int, [vp, C.c_char_p]),
        "dem_add_massflow": (C.c_int, [vp, C.c_int, dp, dp, C.c_int, C.c_int]),
        "dem_massflow_vector": (C.c_int, [vp, dp]),
        "dem_mesh_stress_vector": (C.c_int, [vp, C.c_int, dp]),
        "dem_set_integrate_group": (C.c_int, [vp, C.c_int]),
        "dem_add_move_linear": (C.c_int, [vp, C.c_int, ip, dp, ip]),
        "dem_force_op_vector": (C.c_int, [vp, C.c_int, dp]),
        "dem_cleanup": (C.c_int, [vp]),
        "dem_time": (C.c_double, [vp]),
        "dem_insert_pack_defaults": (None, [C.POINTER(InsertPackSpec)]),
        "dem_add_insert_pack": (C.c_int, [vp, C.POINTER(InsertPackSpec), ip]),
        "dem_insert_stream_defaults": (None, [C.POINTER(InsertStreamSpec)]),
        "dem_add_insert_stream": (C.c_int, [vp, C.POINTER(InsertStreamSpec), ip]),
        "dem_stream_release_data": (C.c_int, [vp, C.c_int, dp]),
        "dem_insert_vector": (C.c_int, [vp, C.c_int, dp]),
        "dem_append_atoms": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, dp, ip, ip, ip]),
        "dem_natoms": (C.c_int, [vp]),
        "dem_add_template_multisphere": (C.c_int, [vp, C.POINTER(TemplateMultisphereSpec), ip]),
        "dem_template_info": (C.c_int, [vp, C.c_int, dp]),
        "dem_create_particles_single": (C.c_int, [vp, C.c_int, dp, dp, dp, dp, C.c_int]),
        "dem_download_props": (C.c_int, [vp, ip, ip, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def material_tables(ntypes, Y, nu, e):
    """Yeff, Geff, betaeff as global_properties.cpp:428-562 builds them."""
    L = library()
    w = ntypes + 1
    Y = np.ascontiguousarray(Y, dtype=np.float64).reshape(w)
    nu = np.ascontiguousarray(nu, dtype=np.float64).reshape(w)
    e = np.ascontiguousarray(e, dtype=np.float64).reshape(w, w)
    out = [np.zeros((w, w)) for _ in range(3)]
    rc = L.dem_material_tables(ntypes, _dp(Y), _dp(nu), _dp(e), _dp(out[0]), _dp(out[1]), _dp(out[2]))
    if rc:
        raise DemError(L.dem_last_error(None).decode())
    return out


def mesh_topology(nodes, box_lo, box_hi, curvature=None):
    """Active edges/corners, obtuse index and edge-neighbour counts the
    engine computes for a triangle mesh (host only, no GPU needed)."""
    L = library()
    nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).reshape(-1, 9))
    nt = len(nd)
    lo = np.ascontiguousarray(box_lo, dtype=np.float64)
    hi = np.ascontiguousarray(box_hi, dtype=np.float64)
    kept = np.zeros(1, np.int32)
    ea = np.zeros((nt, 3), np.int32)
    ca = np.zeros((nt, 3), np.int32)
    ob = np.zeros(nt, np.int32)
    nn = np.zeros(nt, np.int32)
    rc = L.dem_mesh_topology(nt, _dp(nd), 0.0 if curvature is None else float(curvature), _dp(lo), _dp(hi),
                             _ip(kept), _ip(ea), _ip(ca), _ip(ob), _ip(nn))
    if rc:
        raise DemError(L.dem_last_error(None).decode())
    k = int(kept[0])
    return dict(edge_active=ea[:k], corner_active=ca[:k], obtuse=ob[:k], nneigh=nn[:k])


class Simulation:
    """One device simulation.

    ``spec``: dict with box_lo, box_hi, periodic, dt, skin, cdf,
    neigh_every, neigh_delay, neigh_check, rolling, tangential_damping,
    limit_force, materials {ntypes, Y, nu, e, mu, mur, etar, krs},
    planes [(dim, pos, wtype)], meshes [dict(nodes (ntri,3,3), type,
    curvature=None, vel=None)], gravity (ax, ay, az) or None, freeze (bool).
    ``atoms``: dict of numpy arrays x, v, omega (n,3), radius, rmass, type,
    tag, optional freeze.
    """

    def __init__(self, spec, atoms, precision="fp64", device=0):
        self.L = library()
        p = Params()
        p.dt = spec["dt"]
        for d in range(3):
            p.boxlo[d] = spec["box_lo"][d]
            p.boxhi[d] = spec["box_hi"][d]
            p.periodic[d] = int(spec["periodic"][d])
        p.skin = spec["skin"]
        p.cdf = spec.get("cdf", 1.0)
        p.neigh_every = spec.get("neigh_every", 1)
        p.neigh_delay = spec.get("neigh_delay", 0)
        p.neigh_check = int(spec.get("neigh_check", False))
        p.neigh_one = spec.get("neigh_one", 2000)
        p.rolling_model = ROLL[spec.get("rolling", "off")]
        p.tangential_damping = int(spec.get("tangential_damping", True))
        p.normal_model = NORMAL[spec.get("normal", "hertz")]
        p.wall_normal_model = NORMAL[spec.get("wall_normal", "hertz")]
        # limitForce defaults: off for hertz, on for the hysteretic models
        p.limit_force = int(spec.get("limit_force", p.normal_model != 0))
        p.wall_limit_force = int(spec.get("wall_limit_force", p.wall_normal_model != 0))
        p.torsion_torque = int(spec.get("torsion_torque", False))
        p.cohesion_model = COHESION[spec.get("cohesion", "off")]
        p.precision = DEM_MIXED if precision == "mixed" else DEM_FP64
        p.device = device
        self.h = self.L.dem_ctx_create(C.byref(p))
        if not self.h:
            raise DemError(self.L.dem_last_error(None).decode())
        for dim, pos, wt in spec.get("planes", []):
            self._chk(self.L.dem_add_plane(self.h, int(dim), float(pos), int(wt)))
        self.nplanes = len(spec.get("planes", []))
        self.nmesh = 0
        for mspec in spec.get("meshes", []):
            self.add_mesh(**mspec)
        g = spec.get("gravity")
        if g is not None:
            self._chk(self.L.dem_set_gravity(self.h, float(g[0]), float(g[1]), float(g[2])))
        if spec.get("freeze", False):
            self._chk(self.L.dem_set_freeze(self.h, 1))
        m = spec["materials"]
        nt = m["ntypes"]
        w = nt + 1
        Yeff, Geff, beta = material_tables(nt, m["Y"], m["nu"], m["e"])
        mu = np.ascontiguousarray(m["mu"], dtype=np.float64).reshape(w, w)
        mur = np.ascontiguousarray(m.get("mur", np.zeros((w, w))), dtype=np.float64).reshape(w, w)
        etar = np.ascontiguousarray(m.get("etar", np.zeros((w, w))), dtype=np.float64).reshape(w, w)
        self._chk(self.L.dem_set_materials(self.h, nt, _dp(Yeff), _dp(Geff), _dp(beta), _dp(mu),
                                           _dp(mur), _dp(etar), float(m.get("krs", 0.0))))
        if "hysteretic" in m:
            crl = np.ascontiguousarray(np.log(np.asarray(m["e"], dtype=np.float64).reshape(w, w)))
            tabs = np.ascontiguousarray(np.stack([np.asarray(m["hysteretic"][k], dtype=np.float64).reshape(w, w)
                                                  for k in HYST_KEYS]))
            self._chk(self.L.dem_set_hysteretic(self.h, _dp(crl), _dp(tabs)))
        self.ntypes = nt
        self.nu = np.asarray(m["nu"], dtype=np.float64)
        self.upload(atoms)
        if "bond" in spec:
            self.set_bond(**spec["bond"])
        self.ninst = 1
        if "instances" in spec:
            self.set_instances(**spec["instances"])
        if spec.get("observe_plane") is not None:
            self.observe_plane(spec["observe_plane"])
        if spec.get("graph", False):
            self.use_graph(True)

    def add_mesh(self, nodes, type, curvature=None, vel=None):
        nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).reshape(-1, 9))
        vv = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        k = self.L.dem_add_mesh(self.h, len(nd), _dp(nd), int(type),
                                0.0 if curvature is None else float(curvature), None if vv is None else _dp(vv))
        if k < 0:
            raise DemError("dem error %d: %s" % (k, self.L.dem_last_error(self.h).decode()))
        self.nmesh += 1
        return k

    def mesh_contacts(self, k=0):
        """(tag, triangle id, slot, history) of every filled mesh contact slot."""
        cnt = C.c_int64(0)
        self._chk(self.L.dem_get_mesh_contacts(self.h, int(k), C.byref(cnt), None, None, None, None))
        c = cnt.value
        dw = self.wall_dnum
        tg = np.zeros(max(c, 1), np.int32)
        tr = np.zeros(max(c, 1), np.int32)
        sl = np.zeros(max(c, 1), np.int32)
        hi = np.zeros((max(c, 1), dw))
        self._chk(self.L.dem_get_mesh_contacts(self.h, int(k), C.byref(cnt), _ip(tg), _ip(tr), _ip(sl), _dp(hi)))
        return tg[:c], tr[:c], sl[:c], hi[:c]

    def write_restart(self, path):
        """Binary restart of the device state (dem_write_restart)."""
        self._chk(self.L.dem_write_restart(self.h, os.fsencode(path)))

    def read_restart(self, path):
        """Restores a restart written by write_restart; call setup() next."""
        self._chk(self.L.dem_read_restart(self.h, os.fsencode(path)))
        # the atoms are now in the file's order
        self._tag = self.props()["tag"]

    def add_massflow(self, nodes, vec_side, once=True, group_bit=-1):
        """fix massflow/mesh on a planar face mesh (ntri, 3, 3)."""
        nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).reshape(-1, 9))
        vs = np.ascontiguousarray(vec_side, dtype=np.float64)
        self._chk(self.L.dem_add_massflow(self.h, len(nd), _dp(nd), _dp(vs), int(bool(once)), int(group_bit)))

    def massflow_vector(self):
        """(mass, count, mass rate, count rate) as compute_vector gives them."""
        out = np.zeros(4)
        self._chk(self.L.dem_massflow_vector(self.h, _dp(out)))
        return out

    def mesh_stress(self, k=0, pref=(0.0, 0.0, 0.0)):
        """fix mesh/surface/stress on mesh k with reference point pref."""
        p = np.ascontiguousarray(pref, dtype=np.float64)
        self._chk(self.L.dem_mesh_stress(self.h, int(k), _dp(p)))

    def mesh_stress_vector(self, k=0):
        """(force[3], torque[3], p_ref[3]) of the last force pass."""
        out = np.zeros(9)
        self._chk(self.L.dem_mesh_stress_vector(self.h, int(k), _dp(out)))
        return out

    def set_mesh_contacts(self, k, tag, tri, slot, hist):
        tg = np.ascontiguousarray(tag, dtype=np.int32)
        tr = np.ascontiguousarray(tri, dtype=np.int32)
        sl = np.ascontiguousarray(slot, dtype=np.int32)
        h = np.ascontiguousarray(hist, dtype=np.float64).reshape(len(tg), self.wall_dnum)
        self._chk(self.L.dem_set_mesh_contacts(self.h, int(k), len(tg), _ip(tg), _ip(tr), _ip(sl), _dp(h)))

    def set_bond(self, tables, ts_create=0.0, stress_break=False, create_always=False, damping=True,
                 damping_smooth=False, ratio_tc=False, tension_variant=False):
        """Bond tables {BOND_KEYS name: (ntypes+1)^2} and settings."""
        w = self.ntypes + 1
        tabs = np.zeros((len(BOND_KEYS), w, w))
        for k, v in tables.items():
            tabs[BOND_KEYS.index(k)] = np.asarray(v, dtype=np.float64).reshape(w, w)
        tabs = np.ascontiguousarray(tabs)
        b = BondSettings(float(ts_create), int(stress_break), int(create_always), int(damping),
                         int(damping_smooth), int(ratio_tc), int(tension_variant))
        self._chk(self.L.dem_set_bond(self.h, _dp(tabs), C.byref(b)))

    def set_gravity(self, ax, ay, az):
        self._chk(self.L.dem_set_gravity(self.h, float(ax), float(ay), float(az)))

    def _chk(self, rc):
        if rc != 0:
            raise DemError("dem error %d: %s" % (rc, self.L.dem_last_error(self.h).decode()))

    @property
    def n(self):
        """atom count (it grows when an insertion fix places particles)"""
        return self.L.dem_natoms(self.h)

    @property
    def tag(self):
        if len(self._tag) != self.n:
            self._tag = self.props()["tag"]
        return self._tag

    def props(self):
        """tag, type, radius, rmass in download order"""
        n = self.n
        out = dict(tag=np.zeros(n, np.int32), type=np.zeros(n, np.int32), radius=np.zeros(n), rmass=np.zeros(n))
        if n:
            self._chk(self.L.dem_download_props(self.h, _ip(out["tag"]), _ip(out["type"]), _dp(out["radius"]),
                                                _dp(out["rmass"])))
        return out

    def add_insert_pack(self, spec):
        """fix insert/pack (+ its particledistribution/discrete and
        particletemplate/sphere templates, region block); see
        insert_pack_spec.  Returns the fix id."""
        s, keep = insert_pack_spec(spec)
        fid = C.c_int(0)
        self._chk(self.L.dem_add_insert_pack(self.h, C.byref(s), C.byref(fid)))
        del keep
        return fid.value

    def add_insert_stream(self, spec):
        """fix insert/stream (+ its distribution and templates, insertion
        face); see insert_stream_spec.  Returns the fix id."""
        s, keep = insert_stream_spec(spec)
        fid = C.c_int(0)
        self._chk(self.L.dem_add_insert_stream(self.h, C.byref(s), C.byref(fid)))
        del keep
        return fid.value

    def stream_release_data(self, ntags=None):
        """(ntags, 14) release data by tag - 1: x_ins, i_step, r_step,
        v_integrate, v_release, omega_release (zero for atoms no stream made)"""
        nt = (int(self.tag.max()) if self.n else 0) if ntags is None else int(ntags)
        out = np.zeros((max(nt, 0), 14))
        self._chk(self.L.dem_stream_release_data(self.h, nt, _dp(out)))
        return out

    def insert_vector(self, fid):
        """ninserted, massinserted, last inserted, last aimed, last tries, next step"""
        out = np.zeros(6)
        self._chk(self.L.dem_insert_vector(self.h, int(fid), _dp(out)))
        return out

    def add_template_multisphere(self, seed, atom_type, density, x, r, types=None, scale=1.0, ntry=100000,
                                 bonded=False):
        """fix particletemplate/multiplespheres: spheres x (n, 3) and r (n,) as
        the file lists them.  Runs the template's set-up; returns its id."""
        xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        rr = np.ascontiguousarray(r, dtype=np.float64)
        tt = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
        s = TemplateMultisphereSpec()
        s.seed = int(seed)
        s.atom_type = int(atom_type)
        s.density = float(density)
        s.nspheres = len(rr)
        s.x = _dp(xx)
        s.r = _dp(rr)
        s.types = None if tt is None else _ip(tt)
        s.scale = float(scale)
        s.ntry = int(ntry)
        s.bonded = int(bool(bonded))
        tid = C.c_int(0)
        self._chk(self.L.dem_add_template_multisphere(self.h, C.byref(s), C.byref(tid)))
        return tid.value

    def template_info(self, tid):
        """r_bound, x_bound (centre-of-mass frame), volume, mass, centre of mass (file frame)"""
        out = np.zeros(9)
        self._chk(self.L.dem_template_info(self.h, int(tid), _dp(out)))
        return out

    def create_particles_single(self, tid, x, v=None, omega=None, quat=None, gmask=0):
        """create_particles <template> single x y z [velocity] [omega] [orientation]"""
        xs = np.ascontiguousarray(x, dtype=np.float64)
        vs = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        ws = None if omega is None else np.ascontiguousarray(omega, dtype=np.float64)
        qs = None if quat is None else np.ascontiguousarray(quat, dtype=np.float64)
        k = self.L.dem_create_particles_single(self.h, int(tid), _dp(xs), None if vs is None else _dp(vs),
                                               None if ws is None else _dp(ws), None if qs is None else _dp(qs),
                                               int(gmask))
        if k < 0:
            raise DemError("dem error %d: %s" % (k, self.L.dem_last_error(self.h).decode()))
        return k

    def append_atoms(self, atoms, gmask=None):
        m = len(atoms["radius"])
        a = {k: np.ascontiguousarray(atoms[k], dtype=np.float64) for k in ("x", "v", "omega", "radius", "rmass")}
        ty = np.ascontiguousarray(atoms["type"], dtype=np.int32)
        tag = np.ascontiguousarray(atoms["tag"], dtype=np.int32)
        gm = None if gmask is None else np.ascontiguousarray(gmask, dtype=np.int32)
        self._chk(self.L.dem_append_atoms(self.h, m, _dp(a["x"]), _dp(a["v"]), _dp(a["omega"]), _dp(a["radius"]),
                                          _dp(a["rmass"]), _ip(ty), _ip(tag), None if gm is None else _ip(gm)))

    def upload(self, atoms):
        n = len(atoms["radius"])
        a = {k: np.ascontiguousarray(atoms[k], dtype=np.float64) for k in ("x", "v", "omega", "radius", "rmass")}
        ty = np.ascontiguousarray(atoms["type"], dtype=np.int32)
        tag = np.ascontiguousarray(atoms["tag"], dtype=np.int32)
        fr = np.ascontiguousarray(atoms.get("freeze", np.zeros(n)), dtype=np.int32)
        self._tag = tag.copy()
        self._chk(self.L.dem_upload_state(self.h, n, _dp(a["x"]), _dp(a["v"]), _dp(a["omega"]),
                                          _dp(a["radius"]), _dp(a["rmass"]), _ip(ty), _ip(tag), _ip(fr)))

    def close(self):
        if getattr(self, "h", None):
            self.L.dem_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def setup(self):
        self._chk(self.L.dem_setup(self.h))

    def step(self, n):
        self._chk(self.L.dem_step(self.h, int(n)))

    def compute_forces(self, shearupdate=1):
        self._chk(self.L.dem_compute_forces(self.h, int(shearupdate)))

    def rebuild(self):
        self._chk(self.L.dem_rebuild(self.h))

    def sync(self):
        self._chk(self.L.dem_sync(self.h))

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "omega", "f", "torque")}
        self._chk(self.L.dem_download(self.h, _dp(out["x"]), _dp(out["v"]), _dp(out["omega"]),
                                      _dp(out["f"]), _dp(out["torque"])))
        out["tag"] = self.tag.copy()
        return out

    @property
    def dnum(self):
        return self.L.dem_history_size(self.h)

    def pairs(self):
        """Directed full list: (tag_i, tag_j, shift, hist[dnum], flag)."""
        cnt = C.c_int64(0)
        self._chk(self.L.dem_get_pairs(self.h, C.byref(cnt), None, None, None, None, None))
        c = cnt.value
        ti = np.zeros(max(c, 1), np.int32)
        tj = np.zeros(max(c, 1), np.int32)
        sh = np.zeros(max(c, 1), np.int32)
        fl = np.zeros(max(c, 1), np.int32)
        hi = np.zeros((max(c, 1), self.dnum))
        self._chk(self.L.dem_get_pairs(self.h, C.byref(cnt), _ip(ti), _ip(tj), _ip(sh), _dp(hi), _ip(fl)))
        return ti[:c], tj[:c], sh[:c], hi[:c], fl[:c]

    def set_history(self, tag_i, tag_j, hist):
        ti = np.ascontiguousarray(tag_i, dtype=np.int32)
        tj = np.ascontiguousarray(tag_j, dtype=np.int32)
        h = np.ascontiguousarray(hist, dtype=np.float64).reshape(len(ti), self.dnum)
        self._chk(self.L.dem_set_history(self.h, len(ti), _ip(ti), _ip(tj), _dp(h)))

    @property
    def wall_dnum(self):
        return self.L.dem_wall_history_size(self.h)

    def wall_history(self):
        h = np.zeros((self.n, max(self.nplanes, 1), self.wall_dnum))
        if self.nplanes:
            self._chk(self.L.dem_get_wall_history(self.h, _dp(h)))
        return h

    def set_wall_history(self, hist):
        if self.nplanes:
            h = np.ascontiguousarray(hist, dtype=np.float64).reshape(self.n, self.nplanes, self.wall_dnum)
            self._chk(self.L.dem_set_wall_history(self.h, _dp(h)))

    def timing(self, on=True):
        self._chk(self.L.dem_timing_enable(self.h, int(on)))

    def timer(self, name):
        ms = C.c_double(0)
        cnt = C.c_int64(0)
        self._chk(self.L.dem_timing_get(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def list_stats(self):
        a = C.c_double(0)
        b = C.c_double(0)
        e = C.c_int64(0)
        self._chk(self.L.dem_list_stats(self.h, C.byref(a), C.byref(b), C.byref(e)))
        return a.value, b.value, e.value

    def timestep(self):
        return self.L.dem_timestep(self.h)

    # ---- groups and per-group force fixes (applied after the other
    # post_force contributions, in the order added)
    def set_group(self, bit, members):
        """group assign: members[i] != 0 puts atom i (upload order) in group bit."""
        m = np.ascontiguousarray(members, dtype=np.int32)
        self._chk(self.L.dem_set_group(self.h, int(bit), _ip(m)))

    def add_setforce(self, bit, values):
        """fix setforce with constant values; None = NULL.  Returns the fix id."""
        st = np.array([v is not None for v in values], dtype=np.int32)
        va = np.array([0.0 if v is None else float(v) for v in values])
        fid = C.c_int(0)
        self._chk(self.L.dem_add_setforce(self.h, int(bit), _ip(st), _dp(va), C.byref(fid)))
        return fid.value

    def add_aveforce(self, bit, xv, yv, c0, t0=0.0, c1=0.0):
        """fix aveforce with z = c0 - (time - t0)*c1 (None x/y = NULL).
        Returns the fix id."""
        st = np.array([xv is not None, yv is not None, True], dtype=np.int32)
        fid = C.c_int(0)
        self._chk(self.L.dem_add_aveforce(self.h, int(bit), _ip(st), 0.0 if xv is None else float(xv),
                                          0.0 if yv is None else float(yv), float(c0), float(t0), float(c1),
                                          C.byref(fid)))
        return fid.value

    def set_integrate_group(self, bit):
        """fix nve/sphere on group bit `bit` (-1: every atom)."""
        self._chk(self.L.dem_set_integrate_group(self.h, int(bit)))

    def add_move_linear(self, bit, v):
        """fix move linear on group bit; None components are integrated."""
        st = np.array([c is not None for c in v], dtype=np.int32)
        va = np.array([0.0 if c is None else c for c in v])
        fid = C.c_int(0)
        self._chk(self.L.dem_add_move_linear(self.h, int(bit), _ip(st), _dp(va), C.byref(fid)))
        return fid.value

    def unfix(self, fid):
        self._chk(self.L.dem_unfix(self.h, int(fid)))

    def force_op_vector(self, fid):
        """The group's force before the fix acted, at the last force pass."""
        out = np.zeros(3)
        self._chk(self.L.dem_force_op_vector(self.h, int(fid), _dp(out)))
        return out

    aveforce_vector = force_op_vector

    def cleanup(self):
        """End of a `run` command (time bookkeeping)."""
        self._chk(self.L.dem_cleanup(self.h))

    def time(self):
        return self.L.dem_time(self.h)

    # ---- batched runs
    def set_instances(self, inst, Y=None, e=None, mu=None):
        """Atom -> instance (upload order).  Y, e, mu: per-instance values
        (every type of instance b gets Y[b], every type pair e[b], mu[b])."""
        inst = np.ascontiguousarray(inst, dtype=np.int32)
        ninst = int(inst.max()) + 1 if len(inst) else 1
        self._chk(self.L.dem_set_instances(self.h, ninst, _ip(inst)))
        self.ninst = ninst
        if Y is None:
            return
        w = self.ntypes + 1
        for b in range(ninst):
            Yb = np.full(w, float(Y[b]))
            eb = np.full((w, w), float(e[b]))
            Yeff, Geff, beta = material_tables(self.ntypes, Yb, self.nu, eb)
            mub = np.ascontiguousarray(np.full((w, w), float(mu[b])))
            self._chk(self.L.dem_set_instance_materials(self.h, b, _dp(Yeff), _dp(Geff), _dp(beta), _dp(mub),
                                                        None, None))

    def observe_plane(self, k):
        self._chk(self.L.dem_observe_plane(self.h, int(k)))

    def observables(self):
        """(ninst, 5): count, sum z, observed plane force, mass, kinetic energy."""
        out = np.zeros((self.ninst, 5))
        self._chk(self.L.dem_instance_observables(self.h, _dp(out)))
        return out

    def observables_noalloc(self):
        """observables() through the entry point that allocates no device
        memory (a fixed per-GPU scratch array of 4096 instances)."""
        out = np.zeros((self.ninst, 5))
        self._chk(self.L.dem_instance_observables_noalloc(self.h, _dp(out)))
        return out

    def use_graph(self, on=True):
        self._chk(self.L.dem_use_graph(self.h, int(bool(on))))

    def adjoint_forward(self, nsteps, checkpoint_every=100):
        self._chk(self.L.dem_adjoint_forward(self.h, int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, w_height=None, w_wall=None):
        """(ninst, 4): gradient of sum_b w_height[b] mean_z_b + w_wall[b] Fwall_b
        with respect to each instance's Yeff, Geff, betaeff, mu."""
        g = np.zeros((self.ninst, 4))
        wh = None if w_height is None else np.ascontiguousarray(w_height, dtype=np.float64)
        ww = None if w_wall is None else np.ascontiguousarray(w_wall, dtype=np.float64)
        self._chk(self.L.dem_adjoint_backward(self.h, None if wh is None else _dp(wh),
                                              None if ww is None else _dp(ww), _dp(g)))
        return g

    def set_contact_shape(self, lanes=0, block=0):
        self._chk(self.L.dem_set_contact_shape(self.h, int(lanes), int(block)))

    def autotune(self, reps=5):
        g = C.c_int(0)
        b = C.c_int(0)
        ms = C.c_double(0)
        self._chk(self.L.dem_autotune(self.h, int(reps), C.byref(g), C.byref(b), C.byref(ms)))
        return g.value, b.value, ms.value
