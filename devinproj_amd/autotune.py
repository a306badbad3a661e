"""Autotuner for the brick engine's tile: the brick shape (cells per axis) of each system.

`choose_bricks(system, forces, precision)` builds a BrickEngine for every candidate
shape whose staged region fits the LDS, times a few fused steps with HIP events, and
returns the fastest. Results are cached per size bucket (atoms per cell, log2 of N,
precision, interactions) in memory and optionally in a JSON file, so a batch of many
systems tunes one representative per bucket (`tune_batch`).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from .batch import BrickEngine
from .integrators import VelocityVerlet

CANDIDATES = [(1, 1, 1), (1, 1, 2), (1, 2, 2), (2, 2, 2), (2, 2, 3), (2, 3, 3), (3, 3, 3)]
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
    rc = _list_radius(system, forces)
    nc = np.floor(np.asarray(system.box, float) / rc).astype(int)
    per_cell = system.n / float(np.prod(nc))
    kinds = "+".join(sorted(type(f).__name__ for f in forces))
    return f"{precision}|{kinds}|cell{int(round(per_cell))}|n2^{int(math.log2(max(system.n, 2)))}"


def candidates_for(system, forces, precision, lds_bytes=160 * 1024):
    rc = _list_radius(system, forces)
    nc = np.floor(np.asarray(system.box, float) / rc).astype(int)
    per_cell = system.n / float(np.prod(nc))
    elem = 32 if precision == "fp64" else 16
    out = []
    for b in CANDIDATES:
        if any(b[d] > nc[d] for d in range(3)):
            continue
        region = np.prod(np.asarray(b) + 2) * per_cell * 1.5 + 64
        if region * 32 + 4096 <= lds_bytes and region * elem + 4096 <= lds_bytes:
            out.append(b)
    return out or [(1, 1, 1)]


def time_bricks(system, make_forces, precision, bricks, dt, steps=20):
    s = system.copy()
    eng = BrickEngine([s], [make_forces(s)], VelocityVerlet(), dt, precision=precision, bricks=[bricks])
    try:
        eng.run(3)
        eng.set_timing(True)
        eng.run(steps)
        ms, k = eng.timing()
        return ms / max(k, 1)
    finally:
        eng.close()


def choose_bricks(system, make_forces, precision, dt, cache_file=None, steps=20, log=None):
    key = bucket_key(system, make_forces(system), precision)
    if cache_file and not _CACHE and os.path.exists(cache_file):
        with open(cache_file) as fh:
            _CACHE.update({k: tuple(v) for k, v in json.load(fh).items()})
    if key in _CACHE:
        return _CACHE[key]
    best, best_t = None, float("inf")
    for b in candidates_for(system, make_forces(system), precision):
        try:
            t = time_bricks(system, make_forces, precision, b, dt, steps)
        except (RuntimeError, ValueError) as e:  # region does not fit this shape
            if log:
                log(f"  bricks {b}: skipped ({e})")
            continue
        if log:
            log(f"  bricks {b}: {t:.4f} ms per fused step")
        if t < best_t:
            best, best_t = b, t
    _CACHE[key] = best
    if cache_file:
        with open(cache_file, "w") as fh:
            json.dump({k: list(v) for k, v in _CACHE.items()}, fh, indent=1)
    return best


def tune_batch(systems, make_forces, precision, dt, cache_file=None, log=None):
    """One brick shape per system, tuning one representative (the largest) per bucket."""
    keys = [bucket_key(s, make_forces(s), precision) for s in systems]
    rep = {}
    for s, k in zip(systems, keys):
        if k not in rep or s.n > rep[k].n:
            rep[k] = s
    for k, s in rep.items():
        if log:
            log(f"tuning bucket {k} on a {s.n}-atom system")
        choose_bricks(s, make_forces, precision, dt, cache_file=cache_file, log=log)
    return [_CACHE[k] for k in keys]
