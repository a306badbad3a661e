"""Velocity Verlet and Yoshida4 (accusim/integrators.py:9-47) with the reference's
step(sim, dt) contract: they call sim.update_forces() once per (sub)step, so any
Force in this package (or the reference's) plugs in. The device-resident fused
forms of the same schemes live in devinproj_amd.engine.
"""
from __future__ import annotations


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
