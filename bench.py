"""Benchmark: MD steps/s of the device-resident fused LJ path on 1,048,576 argon atoms.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-dsf] [--no-fp64] [--no-learned]
                    [--no-graph] [--no-drift] [--no-richardson] [--no-gravity] [--copy-probe MIB]

Workload (one step = one velocity-Verlet step of the whole system, state resident in
HBM, energies recorded every step): 64^3 x 4 = 1,048,576-atom FCC argon at
1374 kg/m^3, positions jittered by N(0, 0.01 sigma), 94.4 K, LJ cutoff 2.5 sigma,
Verlet skin 0.3 sigma (rebuild at skin/2), dt = 10 fs, brick-tiled LDS engine.
`value` is the mixed-precision path (float32 pair arithmetic, float64 sums and
integration); the plain float64 path on the same GPU is reported beside it, with
the 65,536-ion NaCl LJ + DSF workload and the learned-correction throughput.

roofline: the fused force + integrate kernel. Algorithmic bytes per launch
  = N * 148 B (read x, v, F and the type; write x, v, F) + 4 B per half-list pair
  within rc + skin, divided by its mean launch time from HIP events on the engine's
  stream over the timed region; peak 8 TB/s (HBM3E). traffic: the HBM bytes per
  launch that rocprofv3's FETCH_SIZE / WRITE_SIZE counters measure (calibrated on a
  plain copy).
cpu_baseline: the NumPy float64 oracle on the same 1,048,576-atom workload, one core
  (a child process pinned to one CPU, with OMP/OpenBLAS/MKL threads set to 1 before
  numpy loads), started first and run beside the GPU sections: one list build with the
  first force evaluation (timed on its own), then 2 timed VV steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the mixed fused kernel on this workload (the tuned tile, 3x3x3
# bricks at 512 threads, energies recorded): rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# per dispatch, the factors being what the two counters read for a 1 GiB device copy
# (bench.py --copy-probe 1024); the summary with its commands is
# profiles/k_b_fused_lj_argon_1048576.md
FUSED_TRAFFIC_BYTES = 3.749e8
# VALU wave instructions per launch of the same kernel (rocprofv3 SQ_INSTS_VALU, summed over
# the device, same profile) and the issue capacity they are set against: 256 CUs x 4 SIMDs,
# one wave64 VALU instruction per 4 cycles each, at 2.4 GHz
FUSED_VALU_INSTS = 4.86e7
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4
DT = 10e-15
# brick tiles: the autotuner's choice on MI355X for each workload and precision
# (python -m devinproj_amd.autotune writes devinproj_amd/data/tiles_mi355x.json); the
# fallback is the best tile of a manual sweep on the 1M-atom argon workload
FALLBACK_TILE = {"bricks": (3, 3, 3), "threads": 512}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tile_for(system, forces, precision):
    from devinproj_amd import autotune
    return autotune.cached_tile(system, forces, precision, autotune.DEFAULT_CACHE) or FALLBACK_TILE


def brick_lj(s, precision, steps, warmup, graph_steps=0):
    import torch

    import devinproj_amd as dp
    from devinproj_amd import batch
    from devinproj_amd import workloads as W
    lj = dp.LennardJones(W.argon_params(), cutoff=2.5, skin=0.3, precision=precision)
    tile = tile_for(s, [lj], precision)
    eng = batch.BrickEngine([s], [[lj]], dp.VelocityVerlet(), DT, precision=precision, bricks=[tile["bricks"]],
                            threads=tile["threads"], log_cap=steps + warmup + graph_steps + 64)
    eng.run(warmup)
    eng.set_timing(True)
    torch.cuda.synchronize()
    eng.sync()
    t0 = time.perf_counter()
    eng.run(steps, sync=False)
    eng.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms, launches = eng.timing()
    eng.set_timing(False)
    eng.fetch(state=False)
    out = {"tile": {"bricks": list(tile["bricks"]), "threads": tile["threads"]},
           "steps_per_s": steps / (t1 - t0), "ms_per_step": 1e3 * (t1 - t0) / steps,
           "kernel_ms": ms / max(launches, 1), "launches": launches, "rebuilds": int(eng.rebuilds[0]),
           "list_entries": eng.info["list_entries"], "energy_drift": eng.energy_drift(),
           "region_capacity": eng.info["region_capacity"], "rebuild_phase_us": eng.info["rebuild_phase_us"]}
    if graph_steps:
        blk = 10
        eng.run(blk, use_graph=True)   # capture + first replay
        t0 = time.perf_counter()
        for _ in range(graph_steps // blk):
            eng.run(blk, use_graph=True, sync=False)
        eng.sync()
        t1 = time.perf_counter()
        out["graph_steps_per_s"] = (graph_steps // blk) * blk / (t1 - t0)
    eng.close()
    return out


def brick_nacl(precision, steps, warmup):
    import devinproj_amd as dp
    from devinproj_amd import batch
    from devinproj_amd import workloads as W
    s = W.rocksalt(32, 32, 64, seed=3)
    forces = [dp.LennardJones(W.NACL_LJ, cutoff=2.5, precision=precision),
              dp.DampedShiftedCoulomb(rcut=12e-10, alpha=0.2e10, precision=precision)]
    tile = tile_for(s, forces, precision)
    if tile is FALLBACK_TILE:
        tile = {"bricks": (1, 1, 1), "threads": 256}
    eng = batch.BrickEngine([s], [forces], dp.VelocityVerlet(), 1e-15, precision=precision, bricks=[tile["bricks"]],
                            threads=tile["threads"], log_cap=steps + warmup + 16)
    eng.run(warmup)
    t0 = time.perf_counter()
    eng.run(steps, sync=False)
    eng.sync()
    t1 = time.perf_counter()
    eng.fetch(state=False)
    out = {"steps_per_s": steps / (t1 - t0), "rebuilds": int(eng.rebuilds[0]), "energy_drift": eng.energy_drift(),
           "tile": {"bricks": list(tile["bricks"]), "threads": tile["threads"]}}
    eng.close()
    return out


def drift_long(s, precision, steps=10_000, transient=1_000):
    """Energy drift (accuracy.py:72-76 formula) over `steps` recorded VV steps of the 1M-atom
    workload, state resident in HBM; also the drift after the start-up transient (from
    step `transient` on, same normalisation) and the energy series for the mixed-minus-fp64
    comparison."""
    import devinproj_amd as dp
    from devinproj_amd import batch
    from devinproj_amd import workloads as W
    lj = dp.LennardJones(W.argon_params(), cutoff=2.5, skin=0.3, precision=precision)
    tile = tile_for(s, [lj], precision)
    eng = batch.BrickEngine([s], [[lj]], dp.VelocityVerlet(), DT, precision=precision, bricks=[tile["bricks"]],
                            threads=tile["threads"], log_cap=steps + 64)
    t0 = time.perf_counter()
    eng.run(steps, record_every=1)
    t1 = time.perf_counter()
    eng.fetch(state=False)
    lg = eng.logs[0]
    E = lg[:, 0] + lg[:, 1]
    scale = max(abs(E[0]), np.max(np.abs(lg[:, 1])), np.max(np.abs(lg[:, 0])), 1e-300)
    late = E[transient:]
    slope = np.polyfit(np.arange(len(late), dtype=float), late, 1)[0]
    out = {"steps": steps, "energy_drift": eng.energy_drift(), "rebuilds": int(eng.rebuilds[0]),
           "steps_per_s": steps / (t1 - t0),
           "energy_drift_after_step_%d" % transient: float(np.max(np.abs(late - late[0])) / scale),
           "secular_drift_per_10k_steps": float(abs(slope) * 10_000 / scale)}
    eng.close()
    return out, E, scale


def richardson_batch(coarse_steps=2000, short_steps=100, precision="mixed"):
    """48 uneven reduced-unit LJ systems (N log-uniform in [4000, 262144], rho* in
    [0.7, 0.9], T* = 1), each at dt, dt/2, dt/4 (144 trajectories) in one batched engine,
    blocks of 10 steps captured as one hipGraph, brick shapes and workgroup size from the
    autotuner's cache (one tile per size bucket); timestep_convergence (accuracy.py:191-221)
    per system, over `coarse_steps` coarse steps of dt = 0.005 tau (10 tau: past the
    Lyapunov time of the liquid) and over `short_steps` (0.5 tau: within it)."""
    import re

    from devinproj_amd import autotune, batch
    systems, make, dt = autotune.richardson_workload()
    bricks, threads = autotune.batch_tiles(systems, make, precision, autotune.DEFAULT_CACHE)
    n_atoms = sum(sy.n for sy in systems)
    out = {"systems": len(systems), "trajectories": 3 * len(systems), "atoms_per_replica_set": n_atoms,
           "dt": dt, "tuned_systems": sum(b is not None for b in bricks), "threads": threads}
    for tag, steps in (("long", coarse_steps), ("short", short_steps)):
        t0 = time.perf_counter()
        checks, eng, _ = batch.richardson(systems, make, dt, steps, precision=precision, bricks=bricks,
                                          use_graph=True, block=10, threads=threads)
        t1 = time.perf_counter()
        status = {}
        for c in checks:
            status[c.status] = status.get(c.status, 0) + 1
        est = sorted(c.value for c in checks if c.value is not None)
        orders = sorted(float(m.group(1)) for c in checks for m in [re.search(r"observed order (-?[\d.]+)", c.detail)]
                        if m)
        out[tag] = {"coarse_steps": steps, "run_seconds": eng.run_seconds, "total_seconds": t1 - t0,
                    "coarse_steps_per_s": steps / eng.run_seconds,
                    "atom_steps_per_s": n_atoms * steps * (1 + 2 + 4) / eng.run_seconds, "status_counts": status,
                    "median_estimate": est[len(est) // 2] if est else None,
                    "median_observed_order": orders[len(orders) // 2] if orders else None}
        eng.close()
    return out


def gravity_plummer(n=262144, seed=11):
    """Plummer sphere of n equal-mass stars (1e5 Msun, a = 1 pc), softening 0.01 pc: direct
    O(N^2) forces (float64 and mixed), Barnes-Hut theta = 0.5 with quadrupoles and its RMS
    force error against direct sum on 200 sampled stars (accuracy.py:135-146), and one
    Yoshida4 step with direct forces."""
    import torch

    import devinproj_amd as dp
    from devinproj_amd import workloads as W
    from devinproj_amd.accuracy import rms_rel_error
    from devinproj_amd.units import Msun, parsec
    s, a, M = W.plummer(n, a=parsec, m=1e5 * Msun / n, seed=seed)
    soft = 0.01 * parsec

    def timed(fn, reps=2):
        fn(s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn(s)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    out = {"stars": n}
    for prec in ("mixed", "fp64"):
        g = dp.Gravity(softening=soft, precision=prec)
        sec, _ = timed(g.compute)
        out[f"direct_{prec}_ms"] = 1e3 * sec
        out[f"direct_{prec}_interactions_per_s"] = n * (n - 1) / sec
    bh = dp.BarnesHut(softening=soft, theta=0.5, quadrupole=True)
    sec, (Fb, _) = timed(bh.compute)
    out["bh_theta05_quad_ms"] = 1e3 * sec
    # the same tree build and walk on device-resident positions, timed with HIP events
    pd = torch.from_numpy(s.pos).cuda()
    md = torch.from_numpy(s.masses).cuda()
    F = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    U = torch.empty(1, dtype=torch.float64, device="cuda")
    bh.forces_device(pd, md, n, F, U)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    build_ms = walk_ms = 0.0
    reps = 3
    for _ in range(reps):
        ev[0].record()
        bh.tree.build(pd, md, n, bh.leaf_size, bh.quadrupole)
        ev[1].record()
        bh.tree.walk(bh.G, bh.eps2, bh.theta, F, U)
        ev[2].record()
        torch.cuda.synchronize()
        build_ms += ev[0].elapsed_time(ev[1]) / reps
        walk_ms += ev[1].elapsed_time(ev[2]) / reps
    out["bh_device_build_ms"] = build_ms
    out["bh_device_walk_ms"] = walk_ms
    out["bh_tree"] = bh.tree.info()
    rng = np.random.default_rng(0)
    sub = rng.choice(n, 200, replace=False)
    Fd = dp.direct_subset(dp.Gravity(softening=soft), s, sub)
    out["bh_rms_rel_force_error"] = rms_rel_error(Fb[sub], Fd)
    t_cross = float(np.sqrt(a ** 3 / (dp.G * M)))
    # device-resident Yoshida4 (3 fused force + kick + drift launches per step, one hipGraph
    # per block of steps), direct mixed and fp64, and with the Barnes-Hut tree rebuilt per substep
    for name, force, graph, steps in (("direct_mixed", dp.Gravity(softening=soft, precision="mixed"), True, 4),
                                      ("direct_fp64", dp.Gravity(softening=soft), True, 2),
                                      ("bh_theta05_quad", dp.BarnesHut(softening=soft, theta=0.5, quadrupole=True),
                                       False, 4)):
        eng = dp.NBodyEngine(s.copy(), force, dp.Yoshida4(), 1e-3 * t_cross, log_cap=64)
        eng.run(1, use_graph=graph)
        t0 = time.perf_counter()
        eng.run(steps, use_graph=graph)
        out[f"yoshida4_{name}_s_per_step"] = (time.perf_counter() - t0) / steps
        eng.fetch(state=False)
        out[f"yoshida4_{name}_energy_drift"] = eng.energy_drift()
        eng.close()
    return out


def learned_throughput(reps=5):
    """The learned correction on the 65,536-ion NaCl workload, on the GPU: evaluations/s of
    DSF + the shipped network against Ewald (real-space error 1e-6), each DSF and
    correction call rebuilding its neighbour list (a rebuild every evaluation, into the
    instance's kept device buffers), and
    their RMS relative force error against Ewald on a snapshot of an independent melt
    (Langevin 3000 K then 1200 K, seed 77) that the network never saw, beside DSF alone."""
    import torch

    import devinproj_amd as dp
    from devinproj_amd import learned, train_nacl
    rc = train_nacl.RC
    s, rms = train_nacl.melt((32, 32, 64), seed=77)
    ewald = train_nacl.ewald_1e6()
    corr = learned.load_correction()

    dsf = dp.DampedShiftedCoulomb(rcut=rc, alpha=train_nacl.ALPHA_DSF, precision="mixed")
    ldsf = learned.LearnedDSF(rcut=rc, alpha=train_nacl.ALPHA_DSF, precision="mixed", correction=corr)

    def dsf_fresh(sys_):  # the list is rebuilt inside every call
        dsf.invalidate_list()
        return dsf.compute(sys_)

    def learned_fresh(sys_):  # DSF + correction over one list, rebuilt inside every call
        ldsf.invalidate_list()
        return ldsf.compute(sys_)

    def timed(fn):
        fn(s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(s)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    t_e = timed(ewald.compute)
    t_d = timed(dsf_fresh)
    t_l = timed(learned_fresh)
    Fe, Fd, Fl = ewald.compute(s)[0], dsf_fresh(s)[0], learned_fresh(s)[0]
    from devinproj_amd.accuracy import rms_rel_error
    return {"ewald_ms": 1e3 * t_e, "dsf_ms": 1e3 * t_d, "dsf_plus_correction_ms": 1e3 * t_l,
            "dsf_plus_correction_over_ewald": t_e / t_l,
            "heldout_rms_rel_error_dsf_vs_ewald": rms_rel_error(Fd, Fe),
            "heldout_rms_rel_error_dsf_plus_correction_vs_ewald": rms_rel_error(Fl, Fe),
            "heldout_melt_rms_displacement_A": rms * 1e10,
            "training_record": {k: corr.meta.get(k) for k in ("rms_rel_error_dsf_vs_ewald",
                                                              "rms_rel_error_dsf_plus_ls_pair_bound_vs_ewald",
                                                              "rms_rel_error_dsf_plus_correction_vs_ewald")},
            "note": "drop-in compute() wall times including host<->device copies and the list builds"}


def rattle_dimers(n_mol=131072, steps=200, block=20):
    """Rigid bonds on the device: 131,072 rigid dimers (262,144 atoms, reduced LJ units,
    eps 1, sigma 0.8, rho 0.2, fp64, dt 1e-3) with Rattle (SHAKE and RATTLE per bond
    cluster, tol 1e-10) against velocity Verlet on the same fluid, hipGraph replays of
    `block` steps, no records."""
    import devinproj_amd as dp
    from devinproj_amd import batch
    from devinproj_amd import workloads as W
    s, cons = W.rigid_dimers(n_mol, rho=0.2, seed=3)
    out = {"atoms": s.n, "bonds": len(cons)}
    for name, integ in (("vv", dp.VelocityVerlet()), ("rattle", dp.Rattle(cons, tol=1e-10))):
        eng = batch.BrickEngine([s.copy()], [[dp.LennardJones({"A": (1.0, 0.8)}, cutoff=2.5, skin=0.3)]], integ,
                                1e-3, log_cap=64)
        eng.run(block, record_every=0, use_graph=True)
        t0 = time.perf_counter()
        for _ in range(steps // block):
            eng.run(block, record_every=0, use_graph=True, sync=False)
        eng.sync()
        sec = time.perf_counter() - t0
        eng.fetch(state=False)
        out[f"{name}_ms_per_step"] = 1e3 * sec / (steps // block * block)
        if name == "rattle":
            out["max_violation"] = integ.max_violation
        eng.close()
    return out


def copy_probe(mib=1024, reps=3):
    """Plain device-to-device copies of known size: the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters (each copy reads and writes `mib` MiB)."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"copy_mib": mib, "reps": reps}


def pairs_within(s, rc):
    import devinproj_amd as dp
    i, _ = dp.cell_list_pairs(s, rc)
    return len(i)


def cpu_child(steps=2):
    """The oracle on the 1,048,576-atom workload in this (child) process: list build plus
    first forces, then `steps` timed VV steps. Prints one JSON line."""
    from oracle import accusim_np as A
    from oracle import builders as B
    s = B.argon(64, seed=2, jitter=0.01)
    t0 = time.perf_counter()
    sim = A.Simulation(s, [A.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3)], integrator=A.VelocityVerlet(),
                       dt=DT)
    t1 = time.perf_counter()
    sim.run(steps=steps, positions=False)
    t2 = time.perf_counter()
    print(json.dumps({"atoms": s.n, "init_s": t1 - t0, "steps": steps, "steps_s": t2 - t1}), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_start():
    """Start the oracle child on one CPU of this process's set, thread pools at 1."""
    import subprocess
    env = dict(os.environ)
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS", "NUMEXPR_NUM_THREADS"):
        env[k] = "1"
    cpus = sorted(os.sched_getaffinity(0))
    cpu = cpus[-1]
    p = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-child", str(cpu)], env=env,
                         stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    return p, len(cpus)


def cpu_baseline_finish(p, host_cpus, timeout=900):
    out, err = p.communicate(timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({p.returncode}): {err[-2000:]}")
    r = json.loads(out.strip().splitlines()[-1])
    return {"value": r["steps"] / r["steps_s"], "unit": "steps/s", "cores": 1, "kind": "port",
            "host_cpus": host_cpus, "cpu_model": cpu_model(),
            "list_build_and_first_forces_s": r["init_s"],
            "sample": f"the {r['atoms']}-atom argon workload itself: one cell-list build with the first force "
                      f"evaluation ({r['init_s']:.1f} s), then {r['steps']} VV steps timed "
                      f"({r['steps_s']:.1f} s); one core, thread pools at 1"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dsf", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the captured-graph timing")
    ap.add_argument("--no-drift", action="store_true", help="skip the 10^4-step energy-drift runs")
    ap.add_argument("--no-richardson", action="store_true", help="skip the 48-system Richardson batch")
    ap.add_argument("--no-gravity", action="store_true", help="skip the Plummer gravity section")
    ap.add_argument("--no-rattle", action="store_true", help="skip the rigid-dimer Rattle section")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run plain copies of MIB MiB (counter calibration under rocprofv3 --pmc)")
    ap.add_argument("--cpu-child", type=int, default=None, metavar="CPU", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child is not None:
        os.sched_setaffinity(0, {args.cpu_child})
        cpu_child()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    cpu_proc = None
    if not args.no_cpu:
        cpu_proc, host_cpus = cpu_baseline_start()
        log("cpu baseline (oracle) started in a child process")
    import torch

    from devinproj_amd import _lib
    from devinproj_amd import workloads as W
    _lib.lib()
    name = torch.cuda.get_device_name(0)
    log(f"device: {name}")
    if args.copy_probe:
        log(f"copy probe: {copy_probe(args.copy_probe)}")
    s = W.argon(64, seed=2, jitter=0.01)
    n = s.n
    npairs = pairs_within(s, 2.5 * W.AR_SIGMA)
    log(f"workload: {n} atoms, box {s.box[0] * 1e10:.2f} A, {npairs} pairs within rc")
    gsteps = 0 if args.no_graph else args.steps
    mixed = brick_lj(s.copy(), "mixed", args.steps, args.warmup, graph_steps=gsteps)
    log(f"mixed: {mixed}")
    fp64 = None if args.no_fp64 else brick_lj(s.copy(), "fp64", args.steps, args.warmup, graph_steps=gsteps)
    log(f"fp64: {fp64}")
    nacl = None
    if not args.no_dsf:
        nacl = {p: brick_nacl(p, max(args.steps // 2, 20), 10) for p in ("mixed", "fp64")}
        log(f"nacl: {nacl}")
    learned = None if args.no_learned else learned_throughput()
    log(f"learned: {learned}")
    drift = None
    if not args.no_drift:
        runs = {p: drift_long(s.copy(), p) for p in ("mixed", "fp64")}
        drift = {p: r[0] for p, r in runs.items()}
        (Em, scale), (Ed, _) = (runs["mixed"][1:], runs["fp64"][1:])
        k = min(len(Em), len(Ed))
        drift["mixed_minus_fp64_energy_max_rel"] = float(np.max(np.abs(Em[:k] - Ed[:k])) / scale)
        log(f"drift: {drift}")
    rich = None if args.no_richardson else richardson_batch()
    log(f"richardson: {rich}")
    grav = None if args.no_gravity else gravity_plummer()
    log(f"gravity: {grav}")
    rattle = None if args.no_rattle else rattle_dimers()
    log(f"rattle: {rattle}")
    alg_bytes = n * 148 + 4 * (mixed["list_entries"] / 2)
    achieved = alg_bytes / (mixed["kernel_ms"] * 1e-3) / 1e9
    cpu = None
    if cpu_proc is not None:
        log("waiting for the cpu baseline (oracle)")
        cpu = cpu_baseline_finish(cpu_proc, host_cpus)
        log(f"cpu: {cpu}")
    line = {
        "metric": "MD steps/s at 1M-atom LJ & 65k-ion DSF: mixed/learned vs plain FP64, same GPU",
        "value": mixed["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair / f64 accumulate",
        "data": "synthetic (seeded FCC argon, jitter 0.01 sigma, 94.4 K)",
        "config": {"workload": "lj-argon-fcc-1048576-vv-dt10fs", "atoms": n, "cutoff_sigma": 2.5,
                   "skin_sigma": 0.3, "precision": "mixed", "record_every": 1, "engine": "brick",
                   "bricks": mixed["tile"]["bricks"], "threads": mixed["tile"]["threads"], "device": name},
        "pair_interactions_per_s": npairs * mixed["steps_per_s"],
        "fp64_steps_per_s": fp64["steps_per_s"] if fp64 else None,
        "mixed_over_fp64": mixed["steps_per_s"] / fp64["steps_per_s"] if fp64 else None,
        "graph_steps_per_s": {"mixed": mixed.get("graph_steps_per_s"),
                              "fp64": fp64.get("graph_steps_per_s") if fp64 else None},
        "rebuilds": {"mixed": mixed["rebuilds"], "fp64": fp64["rebuilds"] if fp64 else None},
        "energy_drift": {"mixed": mixed["energy_drift"], "fp64": fp64["energy_drift"] if fp64 else None},
        "nacl_65536_lj_dsf_steps_per_s": {k: v["steps_per_s"] for k, v in nacl.items()} if nacl else None,
        "learned_dsf_vs_ewald": learned,
        "energy_drift_10k_steps": drift,
        "richardson_48_uneven_lj": rich,
        "plummer_262144_gravity": grav,
        "rattle_131072_rigid_dimers": rattle,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": FUSED_TRAFFIC_BYTES,
                     "kernel": "k_b_fused<float,LJ>", "kernel_ms": mixed["kernel_ms"],
                     "alg_bytes_per_launch": alg_bytes,
                     # the kernel's other limit: its measured VALU instructions over the issue
                     # capacity of its measured launch time
                     "valu_issue_frac": FUSED_VALU_INSTS / (1e-3 * mixed["kernel_ms"] * VALU_ISSUE_PER_S)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
