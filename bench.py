"""Benchmark: particle-steps/s of the DEM timestep on one MI355X.

Workload (the reference metric's configuration): 1,000,000 polydisperse
spheres, r ~ U[0.5, 1.5] mm, rho = 2500, periodic 60 x 60 mm in x/y, bottom
plane, packing fraction 0.60, Hertz + history tangential + EPSD rolling
friction (mu_r = 0.1, eta_r = 0.3), Y = 5e6, nu = 0.3, e = 0.9, mu = 0.5,
gravity, dt = 5e-7, skin 0.25 mm, rebuild every 10 steps.  Seeded synthetic
packing (see liggghts-inl_amd/workloads.py), all state resident in HBM.

A step is one full timestep (integrate, rebuild every 10th step, contact
pass with walls and gravity).  The timed steps run through recorded HIP
graphs; a second run from the same start, stepped eagerly with HIP events
around every launch, gives the per-kernel durations the roofline uses.  `value` is the mixed-precision form (FP32
contact law, FP64 positions, history and accumulation); the plain FP64
form is timed beside it, and the mixed form's force error against the FP64
oracle is measured on the full initial state.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
# HBM bytes per particle of one mixed contact pass (integrator fused, no force
# output) on this workload, from the FETCH_SIZE and WRITE_SIZE counter passes
# of rocprofv3 on k_force_q<true,32> (FETCH_SIZE scaled x2.0, WRITE_SIZE x1.0
# against a 256 MiB device copy), averaged over the stepping dispatches:
# profiles/pmc_traffic_k_force_column1M.md.  Re-measured when the kernel changes.
MEASURED_TRAFFIC_BPP = 836.2


def package():
    name = "liggghts_inl_amd"
    pdir = os.path.join(ROOT, "liggghts-inl_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(pdir, "__init__.py"),
                                                  submodule_search_locations=[pdir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def contact_bytes_per_particle(K, C, H, wall_contacts_per_particle, fused=True):
    """Algorithmic HBM bytes of one contact pass per particle, FP64 state:
    own x,v,omega (72) + r,m (16) + type (4) = 92, neighbour indices 4K,
    history read+write 16 H per contact (pairs and walls); then either the
    force and torque written (48), or, with the integrator fused into a pass
    that also starts the next step, the new x/r, v/m and omega/meta records
    (96) written to the other state buffers and no force output (nothing
    reads it before the next pass recomputes it)."""
    return 92.0 + 4.0 * K + 16.0 * H * (C + wall_contacts_per_particle) + (96.0 if fused else 48.0)


def run_gpu(pkg, spec, atoms, precision, steps, warmup, kernel_steps):
    """The timed region steps through recorded HIP graphs (rebuilds every 10
    steps on a fixed cadence); per-kernel durations come from a second run
    of kernel_steps steps from the same start, stepped eagerly with HIP
    events around every launch on the context's stream."""
    s = pkg.Simulation(dict(spec, graph=True), atoms, precision=precision)
    s.setup()
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    t1 = time.perf_counter()
    res = dict(seconds=t1 - t0)
    res["K"], res["C"], res["E"] = s.list_stats()
    st = s.state()
    res["ke"] = float(0.5 * (atoms["rmass"] * (st["v"] ** 2).sum(1)).sum())
    res["mean_z"] = float(st["x"][:, 2].mean())
    s.close()
    s = pkg.Simulation(spec, atoms, precision=precision)
    s.setup()
    s.step(warmup)
    s.sync()
    s.timing(True)
    t0 = time.perf_counter()
    s.step(kernel_steps)
    s.sync()
    res["eager_seconds"] = time.perf_counter() - t0
    res["kernel_steps"] = kernel_steps
    for k in ("contact", "integrate", "rebuild"):
        ms, cnt = s.timer(k)
        res[k] = (ms, cnt)
    res["K_eager"], res["C_eager"], _ = s.list_stats()
    s.close()
    return res


CPU_SAMPLE_N = 50000


def cpu_worker(start_at, seconds):
    """One host core's share of the CPU baseline: the oracle steps its own
    50,000-sphere sample of the column workload from start_at (a shared
    wall-clock time, so every worker's window overlaps) for about `seconds`;
    prints its sphere count, steps and elapsed time."""
    pkg = package()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle

    spec, atoms = pkg.workloads.packed_column(n=CPU_SAMPLE_N, slab_n=12500)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(10)
    while time.time() < start_at:
        time.sleep(0.01)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds:
        o.run(10)
        steps += 10
    print(json.dumps({"n": CPU_SAMPLE_N, "steps": steps, "seconds": time.perf_counter() - t0}))


def cpu_baseline(seconds=12.0):
    """The oracle (the C restatement of the reference's path) on the host's
    cores, one process per core, each stepping its own copy of a
    50,000-sphere sample of the workload over the same wall-clock window:
    the weak-scaled throughput of the CPU path on this host, like one MPI
    rank per core without the ghost exchange.  At most 16 workers (this
    host's share of the machine)."""
    import subprocess

    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    workers = max(1, min(16, ncpu))
    start_at = time.time() + 15.0
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", repr(start_at),
                               repr(seconds)], stdout=subprocess.PIPE, stderr=subprocess.PIPE)
             for _ in range(workers)]
    outs = []
    for p in procs:
        out, err = p.communicate(timeout=seconds + 120)
        if p.returncode != 0:
            raise RuntimeError("cpu baseline worker failed: %s" % err.decode()[-500:])
        outs.append(json.loads(out.decode().strip().splitlines()[-1]))
    rate = sum(o["n"] * o["steps"] / o["seconds"] for o in outs)
    steps = sum(o["steps"] for o in outs)
    return dict(value=rate, unit="particle-steps/s", cores=workers, kind="port",
                sample="oracle (C restatement, 1 thread per process) on %d host cores, each stepping its own "
                       "%d-sphere sample of the same workload (periodic 60x60 mm column, Hertz+history+EPSD, "
                       "bottom plane) for %.0f s over one shared window; %d steps in all"
                % (workers, CPU_SAMPLE_N, seconds, steps))


def mixed_force_error(pkg, n=50000, steps=300):
    """max over atoms of |f_mixed - f_oracle| / (sum of |contact forces|),
    and the same for torques, on an n-sphere sample of the workload that the
    oracle has stepped, so shear and rolling history are live: the mixed GPU
    context gets the oracle's state and history, both take a set-up force
    pass."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle

    spec, atoms = pkg.workloads.packed_column(n=n, slab_n=12500)
    o = Oracle(spec, atoms)
    o.setup()
    o.run(steps)
    st = o.state()
    a2 = dict(atoms, x=st["x"], v=st["v"], omega=st["omega"])
    s = pkg.Simulation(spec, a2, precision="mixed")
    ti, tj, sh, hi, fl = o.pairs()
    k = fl != 0
    s.set_history(ti[k], tj[k], hi[k])
    s.set_wall_history(o.wall_history())
    s.setup()
    o.rebuild()
    o.compute_forces(0)
    fs, ts = o.force_scale()
    r = o.state()
    g = s.state()
    s.close()
    fs = np.maximum(fs, 1e-300)
    ts = np.maximum(ts + fs * atoms["radius"], 1e-300)
    ef = (np.abs(g["f"] - r["f"]).max(1) / fs).max()
    et = (np.abs(g["torque"] - r["torque"]).max(1) / ts).max()
    return float(ef), float(et), int(k.sum())


def bench_batch(pkg, args):
    """Batched pours: 256 independent instances of uneven size (N_b ~
    U[500, 8000], 1.09M spheres in all) with their own Y, e, mu, stepped
    in one context (every kernel launch covers every instance), recorded
    HIP graphs, contact-pass shape from the autotuner.  Reported beside the
    headline metric, not instead of it."""
    spec, atoms, inst, params = pkg.workloads.batch_pours(ninst=args.instances)
    n = len(atoms["radius"])
    out = {}
    for prec in (["mixed", "fp64"] if args.only is None else [args.only]):
        sp = dict(spec, instances=dict(inst=inst, **params), observe_plane=4, graph=True)
        s = pkg.Simulation(sp, atoms, precision=prec)
        s.setup()
        tune = s.autotune(reps=3)
        s.step(args.warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(args.steps)
        s.sync()
        dt = time.perf_counter() - t0
        obs = s.observables()
        s.close()
        out[prec] = dict(value=n * args.steps / dt, ms_per_step=dt * 1e3 / args.steps,
                         autotune={"lanes": tune[0], "block": tune[1], "contact_ms": tune[2]},
                         mean_bed_height_mm=float((obs[:, 1] / obs[:, 0]).mean() * 1e3))
        print("%s: %.3f s for %d steps" % (prec, dt, args.steps), file=sys.stderr, flush=True)
    adj = None
    if args.adjoint:
        # gradients of every instance's mean height and floor force with
        # respect to its Y, mu, e (FP64), checkpoints every 100 steps; the
        # window keeps one neighbour list, so the skin covers the motion
        sp = dict(spec, skin=8e-4, instances=dict(inst=inst, **params), observe_plane=4)
        s = pkg.Simulation(sp, atoms, precision="fp64")
        s.setup()
        s.sync()
        t0 = time.perf_counter()
        s.adjoint_forward(args.adjoint_steps, 100)
        s.sync()
        t1 = time.perf_counter()
        g = s.adjoint_backward(np.ones(args.instances), np.ones(args.instances))
        t2 = time.perf_counter()
        s.close()
        adj = {"steps": args.adjoint_steps, "checkpoint_every": 100, "forward_s": t1 - t0,
               "backward_s": t2 - t1, "backward_particle_steps_per_s": n * args.adjoint_steps / (t2 - t1),
               "grad_finite": bool(np.isfinite(g).all())}
        print("adjoint: forward %.2f s, backward %.2f s" % (t1 - t0, t2 - t1), file=sys.stderr, flush=True)
    main_prec = "mixed" if "mixed" in out else "fp64"
    R = out[main_prec]
    line = {
        "metric": "particle-steps/s, %d batched pours (uneven sizes) in one context" % args.instances,
        "value": R["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": R["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 law / f64 state" if main_prec == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": "%d pours, N_b~U[500,8000], r~U[0.8,1.2] mm, 40x40x100 mm box, 6 planes, "
                               "Hertz+history, Y~U[1e6,1e7], mu~U[0.2,0.7], e~U[0.3,0.9], dt 1e-6, "
                               "rebuild every 10, graph stepping" % args.instances,
                   "n_particles": n, "precision": main_prec},
        "forms": out,
    }
    if adj is not None:
        line["adjoint"] = adj
    print(json.dumps(line))


def bench_fibres(pkg, args):
    """Bonded fibres (500 chains x 40 spheres, bond/nonlinear along the
    chains, hysteretic/nonlinear1 contact, floor plane, plate mesh moving
    down at 0.05 m/s, dt 1e-7), FP64 (bonds and hysteretic chains run in
    FP64).  Reported beside the headline metric."""
    spec, atoms = pkg.workloads.fibres()
    n = len(atoms["radius"])
    # graph stepping with the distance check on the device (neigh_modify
    # check yes): periods halt on the device when a rebuild is due
    s = pkg.Simulation(dict(spec, graph=True), atoms, precision="fp64")
    s.setup()
    s.step(args.warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    dt = time.perf_counter() - t0
    hist = s.pairs()[3]
    bonds = int((hist[:, 12] > 0.5).sum()) // 2
    s.close()
    print("fibres: %.3f s for %d steps, %d bonds intact" % (dt, args.steps, bonds), file=sys.stderr, flush=True)
    line = {
        "metric": "particle-steps/s, bonded fibres under a moving plate", "value": n * args.steps / dt,
        "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "500 chains x 40 spheres r 0.5 mm in a 40 mm cube (fibre mat), bond/nonlinear "
                               "(K_fn2 1e10 N/m3, break at 5 % strain), hysteretic/nonlinear1, floor plane, plate "
                               "mesh at -0.05 m/s, dt 1e-7, gravity", "n_particles": n,
                   "bonds_created": 19500, "bonds_intact": bonds},
    }
    print(json.dumps(line))


def bench_hopper(pkg, args):
    """Hopper discharge on a triangle mesh (200,000 spheres r ~ U[0.8, 1.2] mm
    filled on a jittered lattice inside a 30-degree cone with a cylinder on
    top, 23,360 triangles, Hertz + history, mu 0.4, e 0.5, dt 5e-7), mixed
    and FP64.  Reported beside the headline metric."""
    spec, atoms = pkg.workloads.hopper()
    n = len(atoms["radius"])
    ntri = len(spec["meshes"][0]["nodes"])
    out = {}
    for prec in (["mixed", "fp64"] if args.only is None else [args.only]):
        s = pkg.Simulation(dict(spec, graph=True), atoms, precision=prec)   # a static mesh: graph stepping
        s.setup()
        s.step(args.warmup)
        s.sync()
        t0 = time.perf_counter()
        s.step(args.steps)
        s.sync()
        dt = time.perf_counter() - t0
        st = s.state()
        s.close()
        out[prec] = {"value": n * args.steps / dt, "ms_per_step": dt * 1e3 / args.steps,
                     "mean_height_mm": float(st["x"][:, 2].mean() * 1e3)}
        print("hopper %s: %.3f s for %d steps" % (prec, dt, args.steps), file=sys.stderr, flush=True)
    main_prec = "mixed" if "mixed" in out else "fp64"
    R = out[main_prec]
    line = {
        "metric": "particle-steps/s, hopper discharge on a triangle-mesh wall", "value": R["value"],
        "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 law / f64 state" if main_prec == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": "%d spheres r~U[0.8,1.2] mm in a 30-degree conical hopper (100 mm top, 12 mm "
                               "outlet) with a cylinder on top, %d triangles, Hertz+history, mu 0.4, e 0.5, "
                               "floor plane, gravity, dt 5e-7, rebuild every 10, graph stepping" % (n, ntri),
                   "n_particles": n, "triangles": ntri, "precision": main_prec},
        "forms": out,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the mixed-vs-oracle force check")
    ap.add_argument("--only", choices=["mixed", "fp64"], default=None)
    ap.add_argument("--workload", choices=["column", "batch", "fibres", "hopper"], default="column",
                    help="column: the headline 1M-sphere metric; batch: 256 batched pours; fibres: bonded "
                         "fibres under a moving plate; hopper: 200,000 spheres in a triangle-mesh hopper")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--adjoint", action="store_true", help="batch: also time the adjoint window (FP64)")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="batch: steps of the adjoint window (one neighbour list: motion stays within half the skin)")
    ap.add_argument("--kernel-steps", type=int, default=1000,
                    help="steps of the eager run that times each kernel with HIP events")
    ap.add_argument("--cpu-worker", nargs=2, type=float, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run a device-to-device copy of MB MiB, 10 times (for calibrating "
                         "the FETCH_SIZE/WRITE_SIZE counters against known traffic)")
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(*args.cpu_worker)
        return
    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": n * 4, "launches": 10}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is a design note, "
              "not a measured path (asked for --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    pkg = package()
    if args.workload == "batch":
        bench_batch(pkg, args)
        return
    if args.workload == "fibres":
        bench_fibres(pkg, args)
        return
    if args.workload == "hopper":
        bench_hopper(pkg, args)
        return
    t0 = time.perf_counter()
    spec, atoms = pkg.workloads.packed_column(n=args.n)
    print("workload generated in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        # before any GPU work: the workers are started as child processes
        cpu = cpu_baseline()
        print("cpu baseline: %.3e particle-steps/s on %d cores" % (cpu["value"], cpu["cores"]),
              file=sys.stderr, flush=True)
    results = {}
    for prec in (["mixed", "fp64"] if args.only is None else [args.only]):
        results[prec] = run_gpu(pkg, spec, atoms, prec, args.steps, args.warmup, args.kernel_steps)
        r = results[prec]
        print("%s: %.3f s for %d steps (graph), contact %.3f ms/launch over %d eager steps" %
              (prec, r["seconds"], args.steps, r["contact"][0] / max(r["contact"][1], 1), r["kernel_steps"]),
              file=sys.stderr, flush=True)

    main_prec = "mixed" if "mixed" in results else "fp64"
    R = results[main_prec]
    n = args.n
    value = n * args.steps / R["seconds"]
    H = 6
    wallc = 0.0
    # the timed contact launches are those of the eager run: its list figures
    bpp = contact_bytes_per_particle(R["K_eager"], R["C_eager"], H, wallc)
    avg_ms = R["contact"][0] / max(R["contact"][1], 1)
    achieved = bpp * n / (avg_ms * 1e-3) / 1e9

    out = {
        "metric": "particle-steps/s, 1M Hertz-history spheres, mixed vs FP64 at <=1e-5 force err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["seconds"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 law / f64 state" if main_prec == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "1M polydisperse spheres r~U[0.5,1.5] mm, periodic 60x60 mm column, phi 0.60, "
                        "Hertz+history+EPSD (mu_r 0.1), bottom plane, gravity, dt 5e-7, rebuild every 10",
            "n_particles": n,
            "precision": main_prec,
            "mean_full_neighbours": R["K_eager"],
            "mean_contacts": R["C_eager"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC_BPP * n if main_prec == "mixed" else None,
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE passes, copy-calibrated "
                              "(profiles/pmc_traffic_k_force_column1M.md)",
            "kernel": "k_force_q<mixed, 32> (contact pass, velocity-Verlet integrator fused)",
            "bytes_per_particle": bpp,
            "avg_launch_ms": avg_ms,
        },
    }
    if "fp64" in results and "mixed" in results:
        F = results["fp64"]
        out["fp64"] = {
            "value": n * args.steps / F["seconds"],
            "ms_per_step": F["seconds"] * 1e3 / args.steps,
            "contact_avg_ms": F["contact"][0] / max(F["contact"][1], 1),
            "speedup_mixed_over_fp64": F["seconds"] / R["seconds"],
        }
    out["kernels_ms_per_step"] = {k: R[k][0] / R["kernel_steps"] for k in ("contact", "integrate", "rebuild")}
    out["kernels_ms_per_step"]["eager_ms_per_step"] = R["eager_seconds"] * 1e3 / R["kernel_steps"]
    out["kernels_ms_per_step"]["steps"] = R["kernel_steps"]
    if not args.no_check:
        ef, et, npairs = mixed_force_error(pkg)
        out["mixed_force_error"] = {"force": ef, "torque": et, "bar": 1e-5,
                                    "vs": "FP64 oracle, 50,000-sphere sample of the workload after 300 oracle "
                                          "steps (%d directed pairs with live shear and rolling history), all "
                                          "atoms, relative to the atom's sum of |contact forces| (torques: plus "
                                          "that times the radius)" % npairs}
    if "fp64" in results and "mixed" in results:
        F, R2 = results["fp64"], results["mixed"]
        out["mixed_observables"] = {
            "steps": args.warmup + args.steps,
            "kinetic_energy_rel": abs(R2["ke"] - F["ke"]) / max(abs(F["ke"]), 1e-300),
            "mean_height_rel": abs(R2["mean_z"] - F["mean_z"]) / max(abs(F["mean_z"]), 1e-300),
            "vs": "the FP64 GPU run from the same initial state, at the end of the timed run"}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
