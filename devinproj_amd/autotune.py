"""Autotuner for the brick engine's tile: the brick shape (cells per axis) and the
workgroup size of the per-brick kernels.

`choose_tile(system, make_forces, precision, dt)` builds a BrickEngine for every
candidate (each brick shape whose staged region fits the LDS, at 256 and 512 threads),
times whole MD steps on the wall clock (fused steps, records and the rebuilds that
happen in the run), and returns the fastest as {"bricks": (bx, by, bz), "threads": t}.
Results are cached per size bucket (atoms per cell, N, precision, interactions) in
memory and optionally in a JSON file, so a batch of many systems tunes one
representative per bucket (`tune_batch`). The list layout (16-byte groups of ten
12-bit entries, or eight 16-bit ones when a brick region exceeds 4,095 slots; one
group row per `stride` atoms) follows from the tile, so the list stride is not a
tuning parameter.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from .batch import BrickEngine
from .integrators import VelocityVerlet

CANDIDATES = [(1, 1, 1), (1, 1, 2), (1, 2, 2), (2, 2, 2), (2, 2, 3), (2, 3, 3), (3, 3, 3), (3, 3, 4), (3, 4, 4),
              (4, 4, 4)]
THREADS = (256, 512)
LDS_BYTES = 160 * 1024
_CACHE: dict = {}


def _list_radius(system, forces):
    radii, skins = [], []
    for f in forces:
        if hasattr(f, "list_radius"):
            rc, sk = f.list_radius(system)
        else:
            rc, sk = f.rcut, f.skin
        radii.append(rc)
        skins.append(sk)
    skin = min(skins)
    return max(r + skin for r in radii)


def bucket_key(system, forces, precision):
    """Size bucket of a system: precision, interactions, atoms per list cell in bands of 4
    (what decides which brick regions fit the LDS) and N in factors of 4 (what decides how
    many bricks fill the 256 CUs)."""
    rc = _list_radius(system, forces)
    nc = np.floor(np.asarray(system.box, float) / rc).astype(int)
    per_cell = system.n / float(np.prod(nc))
    kinds = "+".join(sorted(type(f).__name__ for f in forces))
    band = 4 * int(round(per_cell / 4))
    size = 2 * int(math.log2(max(system.n, 2)) // 2)
    return f"{precision}|{kinds}|cell{band}|n2^{size}"


def candidates_for(system, forces, precision, lds_bytes=LDS_BYTES, slack=0.25):
    """Brick shapes that fit the cell grid and whose staged region (with the engine's
    slack and room for density fluctuations) fits the LDS in the fused and list kernels."""
    rc = _list_radius(system, forces)
    nc = np.floor(np.asarray(system.box, float) / rc).astype(int)
    per_cell = system.n / float(np.prod(nc))
    elem = 32 if precision == "fp64" else 16
    out = []
    for b in CANDIDATES:
        if any(b[d] > nc[d] for d in range(3)):
            continue
        region = np.prod(np.asarray(b) + 2) * per_cell * (1.0 + slack) * 1.1 + 64
        if region * max(elem, 16) + 4096 <= lds_bytes:
            out.append(b)
    return out or [(1, 1, 1)]


def time_tile(system, make_forces, precision, bricks, threads, dt, steps=40):
    """Wall-clock seconds per MD step (VelocityVerlet) of one engine with this tile."""
    s = system.copy()
    eng = BrickEngine([s], [make_forces(s)], VelocityVerlet(), dt, precision=precision, bricks=[bricks],
                      threads=threads, log_cap=steps + 16)
    try:
        eng.run(5)
        t0 = time.perf_counter()
        eng.run(steps)
        t1 = time.perf_counter()
        eng.fetch(state=False)  # raises if a region or list overflowed
        return (t1 - t0) / steps
    finally:
        eng.close()


def _load(cache_file):
    if cache_file and not _CACHE and os.path.exists(cache_file):
        with open(cache_file) as fh:
            for k, v in json.load(fh).items():
                _CACHE[k] = {"bricks": tuple(v["bricks"]), "threads": int(v["threads"])}


def _save(cache_file):
    if cache_file:
        with open(cache_file, "w") as fh:
            json.dump({k: {"bricks": list(v["bricks"]), "threads": v["threads"]} for k, v in _CACHE.items()}, fh,
                      indent=1)


def cached_tile(system, forces, precision, cache_file):
    """The cached tile for this system's bucket, or None."""
    _load(cache_file)
    return _CACHE.get(bucket_key(system, forces, precision))


def choose_tile(system, make_forces, precision, dt, cache_file=None, steps=40, threads=THREADS, log=None,
                key_suffix=""):
    key = bucket_key(system, make_forces(system), precision) + key_suffix
    _load(cache_file)
    if key in _CACHE:
        return _CACHE[key]
    best, best_t = None, float("inf")
    for b in candidates_for(system, make_forces(system), precision):
        for t in threads:
            try:
                sec = time_tile(system, make_forces, precision, b, t, dt, steps)
            except (RuntimeError, ValueError) as e:  # region or list does not fit this tile
                if log:
                    log(f"  bricks {b} threads {t}: skipped ({e})")
                continue
            if log:
                log(f"  bricks {b} threads {t}: {1e3 * sec:.4f} ms per step")
            if sec < best_t:
                best, best_t = {"bricks": b, "threads": t}, sec
    if best is None:
        raise RuntimeError(f"no brick tile fits system bucket {key}")
    _CACHE[key] = best
    _save(cache_file)
    return best


def choose_bricks(system, make_forces, precision, dt, cache_file=None, steps=40, log=None):
    """Brick shape of the fastest tile (see choose_tile)."""
    return choose_tile(system, make_forces, precision, dt, cache_file=cache_file, steps=steps, log=log)["bricks"]


def _batch_key(keys, precision):
    """Cache key of a batch: its multiset of size buckets."""
    sig = ",".join(f"{k.split('|', 2)[2]}x{keys.count(k)}" for k in sorted(set(keys)))
    return f"batch|{precision}|{sig}"


def time_batch(systems, make_forces, precision, bricks, threads, dt, steps=20):
    """Seconds per coarse step of the batched timestep_convergence run (batch.richardson: every
    system at dt, dt/2 and dt/4 in one engine, blocks of 10 steps replayed as a hipGraph)."""
    from .batch import richardson
    _, eng, _ = richardson(systems, make_forces, dt, steps, precision=precision, bricks=bricks, use_graph=True,
                           block=10, threads=threads)
    try:
        return eng.run_seconds / steps
    finally:
        eng.close()


def tune_batch(systems, make_forces, precision, dt, cache_file=None, log=None):
    """Tiles for a batch of systems that share one launch (and so one workgroup size): for
    each workgroup size, the fastest brick shape of each size bucket (tuned on the bucket's
    largest system at that size); then, timed on the batched dt, dt/2, dt/4 run itself, the
    workgroup size and the choice between those shapes and the engine's own shape rule.
    Returns (bricks per system, threads) and caches the choice."""
    keys = [bucket_key(s, make_forces(s), precision) for s in systems]
    bkey = _batch_key(keys, precision)
    _load(cache_file)
    rep = {}
    for s, k in zip(systems, keys):
        if k not in rep or s.n > rep[k].n:
            rep[k] = s
    if bkey not in _CACHE:
        best, best_t = None, float("inf")
        for t in THREADS:
            for k, s in rep.items():
                if log:
                    log(f"tuning bucket {k} at {t} threads on a {s.n}-atom system")
                choose_tile(s, make_forces, precision, dt, cache_file=cache_file, threads=(t,), log=log,
                            key_suffix=f"|t{t}")
            # the per-bucket shapes, and the engine's own shape rule, on the whole batch
            for tuned in (True, False):
                bricks = [_CACHE[k + f"|t{t}"]["bricks"] for k in keys] if tuned else None
                sec = time_batch(systems, make_forces, precision, bricks, t, dt)
                if log:
                    log(f"batch of {len(systems)} at {t} threads, {'tuned' if tuned else 'engine default'} "
                        f"bricks: {1e3 * sec:.3f} ms per step")
                if sec < best_t:
                    best, best_t = (t, tuned), sec
        # "bricks" (1, 1, 1) marks per-bucket shapes, (0, 0, 0) the engine's default rule
        _CACHE[bkey] = {"bricks": (1, 1, 1) if best[1] else (0, 0, 0), "threads": best[0]}
        _save(cache_file)
    return batch_tiles(systems, make_forces, precision, cache_file)


def batch_tiles(systems, make_forces, precision, cache_file):
    """Cached tiles of a batch without tuning: (bricks per system, threads) as tune_batch
    chose them (None for a system: the engine's default shape), or (None per system, 256)
    when the batch was never tuned."""
    _load(cache_file)
    keys = [bucket_key(s, make_forces(s), precision) for s in systems]
    bkey = _batch_key(keys, precision)
    if bkey not in _CACHE:
        return [None] * len(systems), 256
    t = _CACHE[bkey]["threads"]
    if tuple(_CACHE[bkey]["bricks"]) == (0, 0, 0):
        return [None] * len(systems), t
    return [_CACHE[k + f"|t{t}"]["bricks"] if k + f"|t{t}" in _CACHE else None for k in keys], t


DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tiles_mi355x.json")


def _bench_workloads():
    """The LJ argon and molten NaCl systems that bench.py measures, with their forces."""
    from . import forces as F
    from . import workloads as W
    argon = W.argon(64, seed=2, jitter=0.01)
    nacl = W.rocksalt(32, 32, 64, seed=3)

    def lj_forces(prec):
        return lambda s: [F.LennardJones(W.argon_params(), cutoff=2.5, skin=0.3, precision=prec)]

    def nacl_forces(prec):
        return lambda s: [F.LennardJones(W.NACL_LJ, cutoff=2.5, precision=prec),
                          F.DampedShiftedCoulomb(rcut=12e-10, alpha=0.2e10, precision=prec)]

    return [(argon, lj_forces, 10e-15), (nacl, nacl_forces, 1e-15)]


def richardson_workload():
    """The 48 uneven reduced-unit LJ systems of the batched timestep_convergence benchmark,
    their force factory and coarsest dt."""
    from . import forces as F
    from . import workloads as W
    systems = [W.lj_reduced(n, rho, T=1.0, seed=sd) for n, rho, sd in W.uneven_batch()]
    return systems, (lambda s: [F.LennardJones({"A": (1.0, 1.0)}, cutoff=2.5, skin=0.3, precision="mixed")]), 0.005


def main(argv=None):
    """python -m devinproj_amd.autotune [--cache FILE] [--richardson]: tune the benchmark
    workloads (and, with --richardson, every size bucket of the 48-system batch)."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    ap.add_argument("--richardson", action="store_true")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.richardson:
        systems, make, dt = richardson_workload()
        bricks, threads = tune_batch(systems, make, "mixed", dt, cache_file=a.cache, log=log)
        print(json.dumps({"richardson_systems": len(systems), "threads": threads,
                          "bricks": [None if b is None else list(b) for b in bricks]}), flush=True)
    for system, make, dt in _bench_workloads():
        for prec in ("mixed", "fp64"):
            log(f"{system.n} atoms, {prec}")
            tile = choose_tile(system, make(prec), prec, dt, cache_file=a.cache, log=log)
            print(json.dumps({"atoms": system.n, "precision": prec, "bricks": list(tile["bricks"]),
                              "threads": tile["threads"]}), flush=True)


if __name__ == "__main__":
    main()
