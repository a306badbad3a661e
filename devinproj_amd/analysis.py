"""Structural and transport observables of accusim/analysis.py, with the pair and frame
loops on the GPU (csrc/analysis.hip):

  rdf(system, rmax, bins, species)   analysis.py:7-24; the pair counts come from the
      device (every pair i < j, minimum image, numpy.histogram's binning), the
      normalisation is the reference's;
  msd(trajectory, box) / vacf(trajectory, max_lag) / green_kubo_diffusion
      analysis.py:27-43, 77-94 over a Trajectory's frames, or over a BrickEngine's
      device snapshot ring without copying it back (msd_engine, vacf_engine);
  pressure(sim, h)   analysis.py:46-64: the reference's box-rescaling central
      difference, here through the device forces' compute();
  wall_pressure(system, duration)   analysis.py:67-74 (host arithmetic).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L


def _box3(s):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(s.box, float), (s.dim,)), float)


def rdf(system, rmax, bins=100, species=None):
    """Radial distribution function g(r) in a periodic box (minimum image). Returns (r, g)."""
    s = system
    idx = np.arange(s.n) if species is None else np.flatnonzero(s.sp == s.species_index(species))
    if not s.periodic:
        raise ValueError("rdf needs a periodic box (for normalisation)")
    if rmax > 0.5 * np.min(np.broadcast_to(s.box, (s.dim,))):
        raise ValueError("rmax must be at most half the box")
    if s.dim != 3:
        raise ValueError("the device rdf is three-dimensional")
    th = L.torch()
    p = th.from_numpy(np.ascontiguousarray(s.pos[idx], float)).to("cuda")
    edges = np.linspace(0.0, float(rmax), int(bins) + 1)
    counts = np.zeros(int(bins), np.uint64)
    L.call("accu_rdf_counts", L.ptr(p), len(idx), L.ptr(_box3(s)), 1, float(rmax), L.ptr(edges), int(bins),
           L.ptr(counts), L.stream())
    hist = counts.astype(np.int64)
    V = float(np.prod(np.broadcast_to(s.box, (s.dim,))))
    n = len(idx)
    shell = 4 / 3 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    ideal = shell * n * (n - 1) / 2 / V
    return 0.5 * (edges[1:] + edges[:-1]), hist / ideal


def _frames_dev(frames):
    th = L.torch()
    arr = np.ascontiguousarray(np.asarray(frames, float))
    if arr.ndim != 3 or arr.shape[2] != 3:
        raise ValueError("frames must be [T][N][3] (a constant particle set)")
    return th.from_numpy(arr).to("cuda"), arr.shape[0], arr.shape[1]


def _msd_dev(ptr, nf, stride, n, box):
    out = np.zeros(nf)
    b = None if box is None else np.ascontiguousarray(np.broadcast_to(np.asarray(box, float), (3,)), float)
    L.call("accu_msd", ptr, nf, int(stride), n, L.ptr(b), int(b is not None), L.ptr(out), L.stream())
    return out


def _vacf_dev(ptr, nf, stride, n, max_lag):
    out = np.zeros(max_lag)
    L.call("accu_vacf", ptr, nf, int(stride), n, int(max_lag), L.ptr(out), L.stream())
    return out


def msd(trajectory, box=None):
    """Mean squared displacement vs time from recorded positions (fixed particle set).

    With a periodic `box`, frames are unwrapped by minimum-image steps, so frames must be
    recorded often enough that no particle moves half a box between them."""
    t = np.asarray(trajectory.t)
    d, nf, n = _frames_dev(trajectory.positions)
    return t - t[0], _msd_dev(L.ptr(d), nf, 3 * n, n, box)


def vacf(trajectory, max_lag=None):
    """Velocity autocorrelation <v(0).v(t)> averaged over particles and time origins (fixed
    particle set). Returns (lag times, C(t))."""
    V = np.asarray(trajectory.velocities)
    if V.ndim != 3:
        raise ValueError("vacf needs recorded velocities of a constant particle set")
    n = len(V)
    max_lag = n // 2 if max_lag is None else max_lag
    d, nf, na = _frames_dev(V)
    C = _vacf_dev(L.ptr(d), nf, 3 * na, na, max_lag)
    t = np.asarray(trajectory.t)
    return t[:max_lag] - t[0], C


def green_kubo_diffusion(trajectory, max_lag=None):
    """Self-diffusion D = (1/d) * integral of the VACF (trapezoid), an independent route to
    the MSD slope."""
    t, C = vacf(trajectory, max_lag)
    dim = np.asarray(trajectory.velocities).shape[2]
    return float(np.sum(0.5 * (C[1:] + C[:-1]) * np.diff(t))) / dim


def _ring(engine, k):
    """(device pointer of system k's first position, frames stored, N of the arena, n_k,
    record indices of the frames) of a BrickEngine's snapshot ring."""
    if not engine.snapshots:
        raise ValueError("the engine keeps no snapshot frames (snapshots=0)")
    frames, stored = L.C.c_void_p(), L.C.c_int(0)
    L.call("accu_bmd_snapshot_frames", engine._h, L.C.byref(frames), L.C.byref(stored))
    N = sum(s.n for s in engine.systems)
    off = sum(s.n for s in engine.systems[:k])
    _, _, rec = engine.snapshot_frames() if stored.value else (None, None, np.zeros(0, np.int32))
    return frames.value + 8 * 3 * off, stored.value, N, engine.systems[k].n, rec


def msd_engine(engine, k=0, box=None):
    """msd of system k over the frames in a BrickEngine's device ring (no copy to the host;
    call after fetch()). Frames are unwrapped by minimum-image steps in `box` (the system's
    box when None and the system is periodic)."""
    ptr, nf, N, n, rec = _ring(engine, k)
    s = engine.systems[k]
    if box is None and s.periodic:
        box = s.box
    t = np.array([engine._rec_times[r][k] for r in rec[:nf]])
    return t - t[0], _msd_dev(L.C.c_void_p(ptr), nf, 6 * N, n, box)


def vacf_engine(engine, k=0, max_lag=None):
    """vacf of system k over the velocity frames of a BrickEngine's device ring."""
    ptr, nf, N, n, rec = _ring(engine, k)
    max_lag = nf // 2 if max_lag is None else max_lag
    t = np.array([engine._rec_times[r][k] for r in rec[:max_lag]])
    return t - t[0], _vacf_dev(L.C.c_void_p(ptr + 8 * 3 * N), nf, 6 * N, n, max_lag)


def pressure(sim, h=1e-6):
    """Instantaneous pressure P = 2K/(dV) - dU/dV for any set of forces (analysis.py:46-64):
    dU/dV by uniformly rescaling the box and all positions (central difference)."""
    s = sim.system
    if s.box is None:
        raise ValueError("pressure needs a box")
    box0, pos0 = np.array(s.box, dtype=float), s.pos.copy()
    V0 = float(np.prod(np.broadcast_to(box0, (s.dim,))))
    U = []
    for sign in (+1, -1):
        lam = (1 + sign * h) ** (1 / s.dim)
        s.box, s.pos = box0 * lam, pos0 * lam
        U.append(sum(f.compute(s)[1] for f in sim.forces if not f.external))
    s.box, s.pos = (box0 if np.ndim(box0) else float(box0)), pos0
    dUdV = (U[0] - U[1]) / (2 * h * V0)
    return 2 * s.kinetic_energy() / (s.dim * V0) - dUdV


def wall_pressure(system, duration):
    """Mechanical pressure on reflecting walls (analysis.py:67-74): momentum delivered
    (System.wall_impulse, reset it to 0 before measuring) per unit wall area per unit time."""
    if system.box is None or system.periodic:
        raise ValueError("wall_pressure needs a reflecting (non-periodic) box")
    V = float(np.prod(system.box))
    area = 2 * V * float(np.sum(1 / system.box))
    return system.wall_impulse / (area * duration)
