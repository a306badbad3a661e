"""The integrators of accusim/integrators.py (velocity Verlet, Yoshida4, Langevin,
Nose-Hoover, Brownian, Rattle) with the reference's step(sim, dt) contract: they call
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



def constraint_clusters(ia, ib):
    """Connected clusters of a constraint graph (atoms are nodes, constraints edges): a list
    of constraint-index arrays, one per cluster, ordered by each cluster's first constraint,
    with the constraints of a cluster in their given order."""
    root = {}

    def find(x):
        while root.setdefault(x, x) != x:
            root[x] = root[root[x]]
            x = root[x]
        return x

    for a, b in zip(ia, ib):
        ra, rb = find(int(a)), find(int(b))
        if ra != rb:
            root[max(ra, rb)] = min(ra, rb)
    groups = {}
    for k, a in enumerate(ia):
        groups.setdefault(find(int(a)), []).append(k)
    return [np.asarray(g, dtype=np.int64) for g in groups.values()]


def coupling(IA, IB, inv):
    """C[n, k, l] for a stack of clusters (atom indices IA, IB [n, K]): the change of bond
    vector k per unit multiplier on bond l moving atom a_l by +inv_a and b_l by -inv_b
    along that bond (integrators.py:202-205)."""
    a, b = IA[:, :, None], IB[:, :, None]
    al, bl = IA[:, None, :], IB[:, None, :]
    return (inv[a] * ((a == al).astype(float) - (a == bl).astype(float))
            - inv[b] * ((b == al).astype(float) - (b == bl).astype(float)))


class Rattle(Integrator):
    """Velocity Verlet with rigid bonds (integrators.py:119-234): SHAKE on positions, RATTLE
    on velocities (Andersen 1983). constraints: [(id_a, id_b, length)] by stable particle id.

    Bonds couple only through shared atoms, so the Newton system of M-SHAKE
    (integrators.py:197-226) is block diagonal over the connected clusters of the constraint
    graph: a rigid dimer is a 1 x 1 block, a rigid triangle 3 x 3. Clusters with the same
    number of bonds are solved together as one stack of small dense systems, so the cost
    grows with the number of bonds rather than its square (the reference solves one dense
    K x K system). The Newton iterations run until the worst relative residual of all bonds
    is at most tol, as the reference's; every `method` ("auto", "matrix", "gauss-seidel")
    takes this solve, the positions reaching the same constraint surface within tol.
    max_violation keeps the worst relative bond-length error of the run. BrickEngine runs
    the same scheme on the device (one cluster per thread)."""

    order = 2
    symplectic = True

    def __init__(self, constraints, tol=1e-12, max_iter=1000, method="auto"):
        self.constraints = [(int(a), int(b), si(d, "m")) for a, b, d in constraints]
        if method not in ("auto", "matrix", "gauss-seidel"):
            raise ValueError(f"unknown method {method!r}")
        self.method = method
        self.tol = tol
        self.max_iter = max_iter
        self.max_violation = 0.0
        self._key, self._groups = None, []

    @staticmethod
    def _mi(s, d):
        return d - s.box * np.round(d / s.box) if s.periodic else d

    def _setup(self, s):
        """Clusters of the constraints whose atoms are both present, grouped by size:
        [(IA, IB, lengths, C, 1 / masses)] with IA, IB [n, K] current atom indices."""
        key = (s.ids.tobytes(), s.masses.tobytes())
        if key == self._key:
            return self._groups
        where = {int(pid): n for n, pid in enumerate(s.ids)}
        cons = [(where[a], where[b], d) for a, b, d in self.constraints if a in where and b in where]
        s.constrained_dof = len(cons)
        ia = np.array([c[0] for c in cons], dtype=np.int64)
        ib = np.array([c[1] for c in cons], dtype=np.int64)
        lengths = np.array([c[2] for c in cons], dtype=float)
        by_size = {}
        for g in constraint_clusters(ia, ib):
            by_size.setdefault(len(g), []).append(g)
        inv = 1.0 / s.masses
        self._groups = []
        for K in sorted(by_size):
            idx = np.stack(by_size[K])
            IA, IB = ia[idx], ib[idx]
            self._groups.append((IA, IB, lengths[idx], coupling(IA, IB, inv), inv))
        self._key = key
        return self._groups

    def step(self, sim, dt):
        s = sim.system
        groups = self._setup(s)
        s.constrained_dof = sum(g[0].size for g in groups)
        s.vel += 0.5 * dt * sim.acc
        new = s.pos + dt * s.vel
        ro = [self._mi(s, s.pos[IA] - s.pos[IB]) for IA, IB, *_ in groups]
        best, stall = np.inf, 0
        for _ in range(self.max_iter):
            r = [self._mi(s, new[IA] - new[IB]) for IA, IB, *_ in groups]
            diff = [d * d - np.sum(rk * rk, axis=-1) for (_, _, d, _, _), rk in zip(groups, r)]
            res = max((float(np.max(np.abs(df) / (g[2] * g[2]))) for df, g in zip(diff, groups)), default=0.0)
            if res <= self.tol:
                break
            best, stall = (res, 0) if res < 0.5 * best else (best, stall + 1)
            if stall >= 10:
                raise RuntimeError(f"SHAKE stalled at relative residual {res:.1e} > tol {self.tol:.0e}: raise tol "
                                   "(float64 limit ~ eps * coordinate size / bond length) or reduce dt")
            for (IA, IB, _, C, inv), rk, rok, df in zip(groups, r, ro, diff):
                g = np.linalg.solve(2 * C * np.einsum("nkd,nld->nkl", rk, rok), df[..., None])[..., 0]
                for idx, sign in ((IA, 1.0), (IB, -1.0)):
                    dx = sign * (g * inv[idx])[..., None] * rok
                    np.add.at(new, idx, dx)
                    np.add.at(s.vel, idx, dx / dt)
        else:
            raise RuntimeError("SHAKE did not converge: reduce dt")
        s.pos = new
        s.apply_boundaries()
        sim.update_forces()
        s.vel += 0.5 * dt * sim.acc
        for IA, IB, d, C, inv in groups:
            r = self._mi(s, s.pos[IA] - s.pos[IB])
            rv = np.sum(r * (s.vel[IA] - s.vel[IB]), axis=-1)
            mu = np.linalg.solve(C * np.einsum("nkd,nld->nkl", r, r), rv[..., None])[..., 0]
            for idx, sign in ((IA, -1.0), (IB, 1.0)):
                np.add.at(s.vel, idx, sign * (mu * inv[idx])[..., None] * r)
            r = self._mi(s, s.pos[IA] - s.pos[IB])
            dev = np.abs(np.sqrt(np.sum(r * r, axis=-1)) - d) / d
            self.max_violation = max(self.max_violation, float(np.max(dev)) if dev.size else 0.0)
