"""Benchmark: MD steps/s of the device-resident fused LJ path on 1,048,576 argon atoms.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-dsf]

Workload (one step = one velocity-Verlet step of the whole system, state resident in
HBM, energies recorded every step): 64^3 x 4 = 1,048,576-atom FCC argon at
1374 kg/m^3, positions jittered by N(0, 0.01 sigma), 94.4 K, LJ cutoff 2.5 sigma,
Verlet skin 0.3 sigma, dt = 10 fs. `value` is the mixed-precision path (float32 pair
arithmetic, float64 sums and integration); the plain float64 path on the same GPU is
reported beside it, plus the 65,536-ion NaCl LJ + DSF workload.

roofline: the fused force+integrate kernel, algorithmic bytes per launch
  = N * 148 B (read x, v, F and the type; write x, v, F) + 4 B per half-list pair
  within rc + skin, divided by its mean launch time from HIP events on the engine's
  stream over the timed region; peak 8 TB/s (HBM3E).
cpu_baseline: the NumPy float64 oracle (single thread) on a 32,000-atom sample of the
  same density, one list build + 5 timed steps, scaled to 1,048,576 atoms by N.
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
DT = 10e-15


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gpu_lj(s, precision, steps, warmup, record_every=1, graph_steps=0):
    import torch

    import devinproj_amd as dp
    from devinproj_amd import workloads as W
    eng = dp.MDEngine(s, [dp.LennardJones(W.argon_params(), cutoff=2.5, skin=0.3, precision=precision)],
                      dp.VelocityVerlet(), dt=DT, precision=precision, log_cap=steps + warmup + graph_steps + 64)
    eng.run(warmup, record_every=record_every)
    eng.set_timing(True)
    torch.cuda.synchronize()
    eng.sync()
    t0 = time.perf_counter()
    eng.run(steps, record_every=record_every, sync=False)
    eng.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms, launches = eng.timing()
    eng.set_timing(False)
    eng.fetch(state=False)
    out = {
        "steps_per_s": steps / (t1 - t0),
        "ms_per_step": 1e3 * (t1 - t0) / steps,
        "kernel_ms": ms / max(launches, 1),
        "launches": launches,
        "rebuilds": eng.rebuilds,
        "list_entries": eng.stats["list_entries"],
        "energy_drift": eng.energy_drift(),
    }
    # the same run continued as replays of one captured hipGraph (blocks of 10 steps)
    if graph_steps:
        eng.use_graph = True
        blk = 10
        eng.run(blk, record_every=record_every)     # capture + first replay
        t0 = time.perf_counter()
        for _ in range(graph_steps // blk):
            eng.run(blk, record_every=record_every, sync=False)
        eng.sync()
        t1 = time.perf_counter()
        out["graph_steps_per_s"] = (graph_steps // blk) * blk / (t1 - t0)
    eng.close()
    return out


def gpu_nacl(precision, steps, warmup):
    import torch

    import devinproj_amd as dp
    from devinproj_amd import workloads as W
    s = W.rocksalt(32, 32, 64, seed=3)
    forces = [dp.LennardJones(W.NACL_LJ, cutoff=2.5, precision=precision),
              dp.DampedShiftedCoulomb(rcut=12e-10, alpha=0.2e10, precision=precision)]
    eng = dp.MDEngine(s, forces, dp.VelocityVerlet(), dt=1e-15, precision=precision, log_cap=steps + warmup + 16)
    eng.run(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(steps, sync=False)
    eng.sync()
    t1 = time.perf_counter()
    eng.fetch(state=False)
    out = {"steps_per_s": steps / (t1 - t0), "rebuilds": eng.rebuilds, "energy_drift": eng.energy_drift()}
    eng.close()
    return out


def pairs_within(s, rc):
    import devinproj_amd as dp
    i, _ = dp.cell_list_pairs(s, rc)
    return len(i)


def cpu_baseline(n_cells=20, steps=5):
    """Oracle (NumPy float64, one thread) on a 4*n_cells^3-atom sample at the same density."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import accusim_np as A
    from oracle import builders as B
    s = B.argon(n_cells, seed=2, jitter=0.01)
    sim = A.Simulation(s, [A.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3)], integrator=A.VelocityVerlet(),
                       dt=DT)
    t0 = time.perf_counter()
    sim.run(steps=steps, positions=False)
    t1 = time.perf_counter()
    sps = steps / (t1 - t0)
    return {"value": sps * s.n / 1048576, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{s.n}-atom FCC argon at the same density, {steps} VV steps after one list build, "
                      f"{sps:.3f} steps/s measured, scaled by N to 1,048,576 atoms"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dsf", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see the sharding notes)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from devinproj_amd import workloads as W
    from devinproj_amd import _lib
    _lib.lib()
    name = torch.cuda.get_device_name(0)
    log(f"device: {name}")
    s = W.argon(64, seed=2, jitter=0.01)
    n = s.n
    rc = 2.5 * W.AR_SIGMA
    log(f"workload: {n} atoms, box {s.box[0] * 1e10:.2f} A")
    npairs = pairs_within(s, rc)
    log(f"pairs within rc: {npairs}")
    log("mixed precision run")
    mixed = gpu_lj(s.copy(), "mixed", args.steps, args.warmup, graph_steps=args.steps)
    log(f"mixed: {mixed}")
    fp64 = None
    if not args.no_fp64:
        log("fp64 run")
        fp64 = gpu_lj(s.copy(), "fp64", args.steps, args.warmup, graph_steps=args.steps)
        log(f"fp64: {fp64}")
    nacl = None
    if not args.no_dsf:
        log("NaCl LJ+DSF runs")
        nacl = {p: gpu_nacl(p, max(args.steps // 3, 10), 10) for p in ("mixed", "fp64")}
        log(f"nacl: {nacl}")
    half_list = mixed["list_entries"] / 2
    alg_bytes = n * 148 + 4 * half_list
    achieved = alg_bytes / (mixed["kernel_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cpu = cpu_baseline()
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
                   "skin_sigma": 0.3, "precision": "mixed", "record_every": 1, "device": name},
        "pair_interactions_per_s": npairs * mixed["steps_per_s"],
        "graph_steps_per_s": {"mixed": mixed.get("graph_steps_per_s"),
                              "fp64": fp64.get("graph_steps_per_s") if fp64 else None},
        "fp64_steps_per_s": fp64["steps_per_s"] if fp64 else None,
        "mixed_over_fp64": mixed["steps_per_s"] / fp64["steps_per_s"] if fp64 else None,
        "rebuilds": {"mixed": mixed["rebuilds"], "fp64": fp64["rebuilds"] if fp64 else None},
        "energy_drift": {"mixed": mixed["energy_drift"], "fp64": fp64["energy_drift"] if fp64 else None},
        "nacl_65536_lj_dsf_steps_per_s": {k: v["steps_per_s"] for k, v in nacl.items()} if nacl else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_md_fused<float,LJ>", "kernel_ms": mixed["kernel_ms"],
                     "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
