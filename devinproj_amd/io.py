"""Trajectories on disk: the compressed .npz layout of accusim/io.py:13-43.

Keys: t, kinetic, potential, internal, released, mass, charge (one value per record),
momentum and angular_momentum ([records][3]), counts and species_meta (JSON text as
uint8 bytes) and, when frames were kept, frame_len, positions, species and velocities
(frames concatenated along the first axis, frame_len giving each frame's row count).
A file written here loads with the reference's `accusim.io.load` and vice versa.
`BrickEngine.trajectory(k)` produces a Trajectory from the device records and snapshot
frames for `save`.
"""
from __future__ import annotations

import json

import numpy as np

from .simulation import Trajectory

SCALARS = ("t", "kinetic", "potential", "internal", "released", "mass", "charge")


def save(trajectory: Trajectory, system, path):
    """io.py:13-26: every per-record list, the species table, and the frames if any."""
    data = {k: np.asarray(getattr(trajectory, k)) for k in SCALARS}
    data["momentum"] = np.asarray(trajectory.momentum)
    data["angular_momentum"] = np.asarray(trajectory.angular_momentum)
    data["counts"] = np.frombuffer(json.dumps(trajectory.counts).encode(), dtype=np.uint8)
    meta = [dict(vars(sp)) for sp in system.species]
    data["species_meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    if trajectory.positions:
        data["frame_len"] = np.array([len(p) for p in trajectory.positions])
        data["positions"] = np.concatenate(trajectory.positions)
        data["species"] = np.concatenate(trajectory.species)
        if len(trajectory.velocities) == len(trajectory.positions):
            data["velocities"] = np.concatenate(trajectory.velocities)
    np.savez_compressed(path, **data)


def load(path) -> tuple[Trajectory, list]:
    """io.py:29-43: (Trajectory, species metadata list). Nothing in the file is unpickled."""
    with np.load(path, allow_pickle=False) as z:
        tr = Trajectory()
        for k in SCALARS:
            setattr(tr, k, z[k].tolist())
        tr.momentum = list(z["momentum"])
        tr.angular_momentum = list(z["angular_momentum"])
        tr.counts = json.loads(bytes(z["counts"]).decode())
        if "frame_len" in z:
            cuts = np.cumsum(z["frame_len"])[:-1]
            tr.positions = np.split(z["positions"], cuts)
            tr.species = np.split(z["species"], cuts)
            if "velocities" in z:
                tr.velocities = np.split(z["velocities"], cuts)
        meta = json.loads(bytes(z["species_meta"]).decode())
    return tr, meta
