"""The integrators of accusim/integrators.py (velocity Verlet, Yoshida4, Langevin,
Nose-Hoover, Brownian) with the reference's step(sim, dt) contract: they call
sim.update_forces() once per (sub)step, so any Force in this package (or the
reference's) plugs in. The device-resident forms of the same schemes run in
devinproj_amd.batch.BrickEngine (and velocity Verlet / Yoshida4 in devinproj_amd.engine).
"""
from __future__ import annotations

import numpy as np

from .units import k_B, si


class Integrator:
    order = 1
    symplectic = False
    stochastic = False
    thermostat = False

    def step(self, sim, dt):
        raise NotImplementedError


class VelocityVerlet(Integrator):
    """Kick-drift-kick leapfrog: 2nd order, symplectic, time reversible."""

    order = 2
    symplectic = True

    def step(self, sim, dt):
        s = sim.system
        s.vel += 0.5 * dt * sim.acc
        s.pos += dt * s.vel
        s.apply_boundaries()
        sim.update_forces()
        s.vel += 0.5 * dt * sim.acc


class Yoshida4(Integrator):
    """4th-order symplectic composition of three Verlet steps with weights w1, w0, w1."""

    order = 4
    symplectic = True
    _w1 = 1 / (2 - 2 ** (1 / 3))
    _w0 = -(2 ** (1 / 3)) * _w1

    def __init__(self):
        self._vv = VelocityVerlet()

    @classmethod
    def weights(cls):
        return (cls._w1, cls._w0, cls._w1)

    def step(self, sim, dt):
        for w in self.weights():
            self._vv.step(sim, w * dt)


class Langevin(Integrator):
    """BAOAB Langevin dynamics (integrators.py:50-71): kick, half drift, Ornstein-Uhlenbeck
    velocity update with c1 = exp(-gamma dt), half drift, forces, kick. The device
    engine (BrickEngine) runs the same step with Philox normals instead of sim.rng."""

    order = 2
    stochastic = True
    thermostat = True

    def __init__(self, temperature, friction, seed=0):
        self.T = si(temperature, "K")
        self.gamma = si(friction, "1/s")
        self.seed = int(seed)

    def step(self, sim, dt):
        s = sim.system
        s.vel += 0.5 * dt * sim.acc
        s.pos += 0.5 * dt * s.vel
        c1 = np.exp(-self.gamma * dt)
        sig = np.sqrt((1 - c1 ** 2) * k_B * self.T / s.masses)[:, None]
        s.vel = c1 * s.vel + sig * sim.rng.normal(size=s.vel.shape)
        s.pos += 0.5 * dt * s.vel
        s.apply_boundaries()
        sim.update_forces()
        s.vel += 0.5 * dt * sim.acc


class NoseHoover(Integrator):
    """Nose-Hoover thermostat (integrators.py:74-116): Trotter splitting of Martyna,
    Tuckerman and Klein (1996) with thermostat mass Q = N_f kT period^2 set at the first
    half step; conserved extended energy E + Q xi^2 / 2 + N_f kT eta (bath_energy)."""

    order = 2
    thermostat = True

    def __init__(self, temperature, period):
        self.T = si(temperature, "K")
        self.tau = si(period, "s")
        self.xi = 0.0
        self.eta = 0.0
        self.Q = None

    @staticmethod
    def dof(s):
        return s.dim * max(s.n - 1, 1)

    def _half(self, s, dt):
        kT, nf = k_B * self.T, self.dof(s)
        if self.Q is None:
            self.Q = nf * kT * self.tau ** 2
        k2 = float(np.sum(s.masses[:, None] * s.vel ** 2))
        self.xi += 0.25 * dt * (k2 - nf * kT) / self.Q
        scale = np.exp(-0.5 * dt * self.xi)
        s.vel *= scale
        self.eta += 0.5 * dt * self.xi
        self.xi += 0.25 * dt * (k2 * scale ** 2 - nf * kT) / self.Q

    def step(self, sim, dt):
        s = sim.system
        self._half(s, dt)
        s.vel += 0.5 * dt * sim.acc
        s.pos += dt * s.vel
        s.apply_boundaries()
        sim.update_forces()
        s.vel += 0.5 * dt * sim.acc
        self._half(s, dt)

    def bath_energy(self, s):
        if self.Q is None:
            return 0.0
        return 0.5 * self.Q * self.xi ** 2 + self.dof(s) * k_B * self.T * self.eta


class Brownian(Integrator):
    """Overdamped Langevin, Euler-Maruyama (integrators.py:237-257): with each species'
    diffusion coefficient D, dx = D F / (kT) dt + sqrt(2 D dt) xi and v = dx / dt."""

    order = 1
    stochastic = True
    thermostat = True

    def __init__(self, temperature, seed=0):
        self.T = si(temperature, "K")
        self.seed = int(seed)

    def step(self, sim, dt):
        s = sim.system
        D = np.array([sp.diffusion for sp in s.species])[s.sp][:, None]
        drift = D * sim.force / (k_B * self.T) * dt
        noise = np.sqrt(2 * D * dt) * sim.rng.normal(size=s.pos.shape)
        s.pos += drift + noise
        s.vel = (drift + noise) / dt
        s.apply_boundaries()
        sim.update_forces()
