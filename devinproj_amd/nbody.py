"""Device-resident N-body runs over gravity (the C ABI `accu_nbody_*`).

`NBodyEngine(system, force, integrator, dt)` is `Simulation.run` (simulation.py:95-112)
for one open-space or periodic system under `Gravity` (forces.py:31-48) or `BarnesHut`
(tree.py:19-78) with `VelocityVerlet` or `Yoshida4` (integrators.py:19-47), the state
staying in device memory between calls. With direct gravity every substep is one
launch (force sum, closing kick, record sums, next opening kick and drift), so a
Yoshida4 step is three launches and a block of steps can be replayed as one hipGraph.
Each record holds the float64 device sums of potential and kinetic energy, momentum and
angular momentum (simulation.py:76-93); `trajectory()` returns them as a `Trajectory`
that `AccuracyReport.from_run` grades.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .forces import Gravity
from .integrators import VelocityVerlet, Yoshida4
from .simulation import Trajectory
from .tree import BarnesHut
from .units import si

LOG_W = 8  # PE, KE, Px, Py, Pz, Lx, Ly, Lz


class NBodyEngine:
    def __init__(self, system, force, integrator=None, dt=None, log_cap=4096):
        integrator = VelocityVerlet() if integrator is None else integrator
        if isinstance(integrator, Yoshida4):
            self.weights = np.ascontiguousarray(Yoshida4.weights(), float)
        elif type(integrator) is VelocityVerlet:
            self.weights = np.ones(1)
        else:
            raise ValueError(f"the N-body engine runs VelocityVerlet and Yoshida4, not {type(integrator).__name__}")
        if system.dim != 3:
            raise ValueError("the N-body engine needs dim == 3")
        if system.box is not None and not system.periodic:
            raise ValueError("walls (a box that is not periodic) are not on this path")
        self.system, self.force, self.integrator = system, force, integrator
        self.dt = si(dt, "s")
        self.steps = 0
        self.time = 0.0
        self._rec_t = [0.0]
        h = L.C.c_void_p()
        pos = L.host(system.pos)
        vel = L.host(system.vel)
        m = L.host(system.masses)
        box = None if system.box is None else L.host(system.box)
        L.call("accu_nbody_create", L.C.byref(h), system.n, L.ptr(pos), L.ptr(vel), L.ptr(m), L.ptr(box),
               int(bool(system.periodic)))
        self._h = h
        if isinstance(force, BarnesHut):
            L.call("accu_nbody_set_bh", h, float(force.G), float(force.eps2), float(force.theta),
                   int(force.leaf_size), int(bool(force.quadrupole)))
        elif isinstance(force, Gravity):
            L.call("accu_nbody_set_direct", h, float(force.G), float(force.eps2), int(force._prec))
        else:
            raise ValueError(f"the N-body engine takes Gravity or BarnesHut, not {type(force).__name__}")
        self.log_cap = int(log_cap)
        L.call("accu_nbody_init", h, self.log_cap)

    def run(self, steps, record_every=1, use_graph=False, sync=True):
        L.call("accu_nbody_run", self._h, int(steps), float(self.dt), L.ptr(self.weights), len(self.weights),
               int(record_every), int(bool(use_graph)))
        for k in range(1, steps + 1):
            if record_every > 0 and (k % record_every == 0 or k == steps):
                self._rec_t.append(self.time + k * self.dt)
        self.steps += steps
        self.time += steps * self.dt
        if sync:
            self.sync()
        return self

    def sync(self):
        L.call("accu_nbody_sync", self._h)

    def fetch(self, state=True):
        """Copy the records (and, with state, positions, velocities and accelerations into
        the system, `acc`) back; returns the log [records, 8]."""
        n = self.system.n
        pos = np.empty((n, 3)) if state else None
        vel = np.empty((n, 3)) if state else None
        acc = np.empty((n, 3)) if state else None
        log = np.zeros((self.log_cap, LOG_W))
        nrec = L.C.c_longlong(0)
        L.call("accu_nbody_get", self._h, L.ptr(pos), L.ptr(vel), L.ptr(acc), L.ptr(log), self.log_cap,
               L.C.byref(nrec))
        self.log = log[:nrec.value].copy()
        if state:
            self.system.pos[...] = pos
            self.system.vel[...] = vel
            self.acc = acc
        return self.log

    def trajectory(self):
        """The records as a Trajectory (after fetch): t, kinetic, potential, momentum and
        angular momentum from the device; internal, released and bath 0, mass, charge and
        counts constant (no reactions on this path)."""
        s, lg = self.system, self.log
        n = len(lg)
        tr = Trajectory()
        tr.t = list(self._rec_t[:n])
        tr.potential = lg[:, 0].tolist()
        tr.kinetic = lg[:, 1].tolist()
        tr.internal = [0.0] * n
        tr.released = [0.0] * n
        tr.momentum = [lg[r, 2:5].copy() for r in range(n)]
        tr.angular_momentum = [lg[r, 5:8].copy() for r in range(n)]
        tr.mass = [float(np.sum(s.masses))] * n
        tr.charge = [float(np.sum(s.charges))] * n
        tr.counts = [s.counts() for _ in range(n)]
        tr.bath = [0.0] * n
        return tr

    def energy_drift(self):
        """max |E - E0| / max(|E0|, max |KE|, max |PE|) over the records (accuracy.py:72-76)."""
        lg = self.log
        E = lg[:, 0] + lg[:, 1]
        scale = max(abs(E[0]), np.max(np.abs(lg[:, 1])), np.max(np.abs(lg[:, 0])), 1e-300)
        return float(np.max(np.abs(E - E[0])) / scale)

    def close(self):
        h = getattr(self, "_h", None)
        lib = getattr(L, "_lib", None) if L is not None else None
        if h is not None and lib is not None:
            lib.accu_nbody_destroy(h)
        self._h = None

    def __del__(self):
        self.close()
