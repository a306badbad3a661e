"""The host-side run loop over the GPU forces (accusim/simulation.py:14-127).

`Simulation` keeps the reference's constructor, attributes (`system`, `forces`,
`integrator`, `dt`, `rng`, `seed`, `time`, `steps`, `trajectory`, `force`, `pe`,
`acc`) and methods (`potential`, `update_forces`, `record`, `run`, `clone`,
`accuracy_report`), so that integrators, the accuracy report and user scripts written
against accusim drive the device forces unchanged. Each `compute` returns a fresh float64
array; `potential` adds them into one array of its own.

What one record holds is defined in one place, `_observables`: time, kinetic and
potential energy, internal energy, momentum, angular momentum, mass, charge, species
counts and the thermostat's bath energy. Reactions are not part of this package's path:
`reactions` must be empty, so nothing is released and the mass is the rest mass.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np

from .integrators import VelocityVerlet
from .units import si


@dataclass
class Trajectory:
    t: list = field(default_factory=list)
    kinetic: list = field(default_factory=list)
    potential: list = field(default_factory=list)
    internal: list = field(default_factory=list)
    released: list = field(default_factory=list)
    momentum: list = field(default_factory=list)
    angular_momentum: list = field(default_factory=list)
    mass: list = field(default_factory=list)
    charge: list = field(default_factory=list)
    counts: list = field(default_factory=list)
    positions: list = field(default_factory=list)
    velocities: list = field(default_factory=list)
    bath: list = field(default_factory=list)
    species: list = field(default_factory=list)

    @property
    def total_energy(self) -> np.ndarray:
        """Kinetic + potential + internal energy, less what reactions released."""
        parts = [np.asarray(getattr(self, k), float) for k in ("kinetic", "potential", "internal", "released")]
        return parts[0] + parts[1] + parts[2] - parts[3]

    def count(self, name) -> np.ndarray:
        return np.array([c.get(name, 0) for c in self.counts])


class Simulation:
    def __init__(self, system, forces=(), reactions=None, integrator=None, dt=None, seed=None):
        if reactions:
            raise ValueError("reactions are not on the GPU force path (run such models with accusim itself)")
        self.system = system
        self.forces = list(forces)
        self.integrator = VelocityVerlet() if integrator is None else integrator
        self.dt = si(dt, "s") if dt is not None else None
        self.seed = seed
        self.rng = np.random.default_rng(seed)
        self.reactions = None
        self.time = 0.0
        self.steps = 0
        self.trajectory = Trajectory()
        self._initial = None
        self.update_forces()
        self._m0 = self._mass_energy()

    # ---- forces -------------------------------------------------------------------

    def potential(self):
        """(total force [N, dim], total potential energy) of every force on the current state."""
        total = np.zeros_like(self.system.pos)
        energy = 0.0
        for f in self.forces:
            Fi, Ui = f.compute(self.system)
            total += Fi
            energy += Ui
        return total, energy

    def update_forces(self):
        self.force, self.pe = self.potential()
        m = self.system.masses
        self.acc = self.force / m[:, None] if m.size else self.force

    def _mass_energy(self):
        return float(np.sum(self.system.masses))

    # ---- records ------------------------------------------------------------------

    def _observables(self, with_state):
        s = self.system
        bath = getattr(self.integrator, "bath_energy", None)
        obs = {
            "t": self.time,
            "kinetic": s.kinetic_energy(),
            "potential": self.pe,
            "internal": float(np.sum(getattr(s, "eint", 0.0))),
            "released": 0.0,
            "momentum": s.momentum(),
            "angular_momentum": s.angular_momentum() if s.dim in (2, 3) else np.zeros(1),
            "mass": self._mass_energy(),
            "charge": float(np.sum(s.charges)),
            "counts": s.counts(),
            "bath": bath(s) if bath else 0.0,
        }
        if with_state:
            obs.update(positions=s.pos.copy(), velocities=s.vel.copy(), species=s.sp.copy())
        return obs

    def record(self, positions=True):
        for name, value in self._observables(positions).items():
            getattr(self.trajectory, name).append(value)

    # ---- stepping -----------------------------------------------------------------

    def run(self, duration=None, steps=None, record_every=1, positions=True):
        if self.dt is None:
            raise ValueError("set dt (or use accusim.advisor.suggest_dt)")
        n = int(round(si(duration, "s") / self.dt)) if steps is None else int(steps)
        if not self.trajectory.t:
            # the first record also keeps the starting point for timestep_convergence
            if self._initial is None:
                self._initial = self.clone()
            self.record(positions)
        for k in range(1, n + 1):
            self.integrator.step(self, self.dt)
            self.time += self.dt
            self.steps += 1
            if k == n or k % record_every == 0:
                self.record(positions)
        return self.trajectory

    def clone(self, dt=None):
        """A new simulation from the current state: its own copy of the system and of the
        integrator's state, the same force objects (so their neighbour lists are shared),
        a fresh generator from the seed and an empty trajectory; time and step count go on
        from here."""
        twin = Simulation(self.system.copy(), self.forces, integrator=copy.deepcopy(self.integrator),
                          dt=self.dt if dt is None else dt, seed=self.seed)
        twin.time, twin.steps = self.time, self.steps
        twin._initial, twin._m0 = self._initial, self._m0
        return twin

    def accuracy_report(self, convergence_duration=None):
        from .accuracy import AccuracyReport
        return AccuracyReport.build(self, convergence_duration)

