"""Device-resident MD engine: the fused force + velocity-Verlet / Yoshida4 path.

`MDEngine(system, forces, integrator, dt, precision)` takes the same System, Force
objects (LennardJones and/or DampedShiftedCoulomb) and integrator class a
Simulation would, keeps all state in HBM, and runs steps with one fused kernel
per (sub)step plus a flag-gated Verlet rebuild chain (optionally captured into a
hipGraph). `run()` returns a Trajectory of the recorded PE, KE and momentum; the
system's pos/vel are refreshed from the device afterwards, in the user's order.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .forces import DampedShiftedCoulomb, LennardJones, _erfc
from .integrators import VelocityVerlet, Yoshida4
from .simulation import Trajectory

LOG_W = 8


class MDEngine:
    def __init__(self, system, forces, integrator=None, dt=None, precision="fp64", slack=0.35, log_cap=100_000,
                 use_graph=False):
        s = system
        if s.dim != 3 or not s.periodic:
            raise ValueError("the device MD engine needs a 3D periodic box")
        self.system = s
        self.forces = list(forces)
        self.integrator = integrator or VelocityVerlet()
        if isinstance(self.integrator, Yoshida4):
            self.weights = np.array(Yoshida4.weights(), float)
        elif isinstance(self.integrator, VelocityVerlet):
            self.weights = np.array([1.0])
        else:
            raise ValueError("the device engine runs VelocityVerlet or Yoshida4")
        self.dt = dt
        self.precision = precision
        self.use_graph = use_graph
        prec = L.PRECISIONS[precision]
        lj = [f for f in self.forces if isinstance(f, LennardJones)]
        dsf = [f for f in self.forces if isinstance(f, DampedShiftedCoulomb)]
        if len(lj) > 1 or len(dsf) > 1 or len(lj) + len(dsf) != len(self.forces) or not self.forces:
            raise ValueError("the device engine takes one LennardJones and/or one DampedShiftedCoulomb")
        if any(f.exclude for f in dsf):
            raise ValueError("exclusions are not supported by the device engine")
        L.lib()
        ns = len(s.species)
        mass = np.ascontiguousarray([sp.mass for sp in s.species], float)
        charge = np.ascontiguousarray([sp.charge for sp in s.species], float)
        pos = np.ascontiguousarray(s.pos, float)
        vel = np.ascontiguousarray(s.vel, float)
        spi = np.ascontiguousarray(s.sp, np.int32)
        box = np.ascontiguousarray(s.box, float)
        h = L.C.c_void_p()
        L.call("accu_md_create", L.C.byref(h), s.n, L.ptr(pos), L.ptr(vel), L.ptr(spi), ns, L.ptr(mass),
               L.ptr(charge), L.ptr(box), 1)
        self._h = h
        # one Verlet list for all interactions: radius = max(rc + skin), skin = the smallest skin
        radii, skins = [], []
        if lj:
            f = lj[0]
            eps, sig2, rc2, shift = f.device_tables(s)
            L.call("accu_md_set_lj", h, ns, L.ptr(eps), L.ptr(sig2), L.ptr(rc2), shift)
            rc, skin = f.list_radius(s)
            radii.append(rc)
            skins.append(skin)
        if dsf:
            f = dsf[0]
            erc = float(_erfc(np.array(f.alpha * f.rcut)))
            L.call("accu_md_set_dsf", h, float(f.alpha), float(f.rcut), erc, float(f.k))
            radii.append(f.rcut)
            skins.append(f.skin)
        skin = min(skins)
        rc_list = max(r + skin for r in radii)
        self.rc_list, self.skin = rc_list, skin
        L.call("accu_md_init", h, float(rc_list), float(skin), prec, float(slack), int(log_cap))
        self.log_cap = log_cap
        self.trajectory = Trajectory()
        self.time = 0.0
        self.steps = 0

    def run(self, steps, record_every=1, dt=None, sync=True):
        dt = self.dt if dt is None else dt
        if dt is None:
            raise ValueError("set dt")
        w = np.ascontiguousarray(self.weights)
        L.call("accu_md_run", self._h, int(steps), float(dt), L.ptr(w), len(w), int(record_every),
               int(bool(self.use_graph)))
        self.steps += steps
        self.time += steps * dt
        if sync:
            L.call("accu_md_sync", self._h)
        return self

    def sync(self):
        L.call("accu_md_sync", self._h)

    def fetch(self, state=True):
        """Copy positions, velocities, forces (user order) and the energy log to the host."""
        s = self.system
        n = s.n
        pos = np.empty((n, 3)) if state else None
        vel = np.empty((n, 3)) if state else None
        F = np.empty((n, 3)) if state else None
        log = np.empty((self.log_cap, LOG_W))
        stats = np.zeros(5, np.int64)
        L.call("accu_md_get", self._h, L.ptr(pos), L.ptr(vel), L.ptr(F), L.ptr(log), self.log_cap, L.ptr(stats))
        nrec = int(stats[1])
        self.stats = {"rebuilds": int(stats[0]), "records": nrec, "max_neighbours": int(stats[2]),
                      "capacity": int(stats[3]), "list_entries": int(stats[4])}
        if state:
            s.pos[...] = pos
            s.vel[...] = vel
            self.force = F
        self.log = log[:nrec].copy()
        return self.log

    def trajectory_arrays(self):
        """PE, KE, momentum [records, 3] from the device log (call fetch() first)."""
        lg = self.log
        return lg[:, 0], lg[:, 1], lg[:, 2:5]

    def energy_drift(self):
        """accuracy.py:72-76 on the recorded series: max |E - E0| / max(|E0|, max|KE|, max|PE|)."""
        pe, ke, _ = self.trajectory_arrays()
        E = ke + pe
        scale = max(abs(E[0]), np.max(np.abs(ke)), np.max(np.abs(pe)), 1e-300)
        return float(np.max(np.abs(E - E[0])) / scale)

    @property
    def rebuilds(self):
        return self.stats["rebuilds"]

    def set_timing(self, enable=True):
        L.call("accu_md_set_timing", self._h, int(bool(enable)))

    def timing(self):
        """(total ms, launches) of the fused kernels timed with HIP events since set_timing()."""
        ms, k = L.C.c_double(0.0), L.C.c_longlong(0)
        L.call("accu_md_timing", self._h, L.C.byref(ms), L.C.byref(k))
        return ms.value, k.value

    def stream_handle(self):
        p = L.C.c_void_p()
        L.call("accu_md_stream", self._h, L.C.byref(p))
        return p

    def close(self):
        if getattr(self, "_h", None) is not None and L._lib is not None:
            L._lib.accu_md_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
