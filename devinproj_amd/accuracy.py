"""Accuracy report for runs of this package (accusim/accuracy.py:18-221).

A report is a list of Check rows. Each row here is one function of a `RunFacts` record
(the trajectory, the final system, the forces, the integrator and dt) that returns a Check,
or None where the row does not apply. The same rows serve `Simulation` runs with the
drop-in forces and runs of the device engines (`AccuracyReport.from_run`), so both grade
alike. Which rows exist, their formulas, grade thresholds and messages, and their order
are the reference's, so that a report reads and grades the same:

  conservation rows (accuracy.py:71-123): energy, extended energy, momentum, angular
      momentum, mass-energy, charge, thermostat;
  state rows (accuracy.py:125-189): timestep vs physics (advisor.py:9-48), force
      approximation (200 sampled particles against exact summation, both on the GPU),
      DSF vs Ewald (GPU Ewald as the judge; `dsf_ewald_max_n` can lift the reference's
      n <= 2000 guard), electrostatics (plain Coulomb in a periodic box), constraints
      (Rattle's worst bond violation), relativity;
  timestep error (accuracy.py:191-221), when a convergence duration is given.
Reactions are not part of this package, so the reaction-timestep row never appears.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PASS, WARN, FAIL, NA = "PASS", "WARN", "FAIL", "n/a"


@dataclass
class Check:
    name: str
    status: str
    value: float | None
    detail: str

    def row(self):
        shown = "" if self.value is None else f"{self.value:.2e}"
        return f"  [{self.status:>4}] {self.name:<22} {shown:>10}  {self.detail}"


def grade(x, good, ok):
    """PASS up to `good`, WARN up to `ok`, FAIL beyond."""
    if x <= good:
        return PASS
    return WARN if x <= ok else FAIL


_grade = grade


def rms_rel_error(F_ap, F_ex):
    """sqrt(mean |F_ap - F_ex|^2 / mean |F_ex|^2) over particles."""
    num = np.mean(np.sum((F_ap - F_ex) ** 2, axis=1))
    den = np.mean(np.sum(F_ex ** 2, axis=1))
    return float(np.sqrt(num / den))


@dataclass
class RunFacts:
    """What the rows read: a finished run's records and its final state."""

    trajectory: object
    system: object
    forces: list
    integrator: object
    dt: float | None = None
    dsf_ewald_max_n: int = 2000
    _cache: dict = field(default_factory=dict, repr=False)

    @property
    def thermostatted(self):
        return bool(getattr(self.integrator, "thermostat", False))

    @property
    def conservative(self):
        return all(f.conservative for f in self.forces) and not self.thermostatted

    @property
    def external(self):
        return any(f.external for f in self.forces)

    @property
    def walls(self):
        return self.system.box is not None and not self.system.periodic

    @property
    def central(self):
        return all(f.central for f in self.forces)

    @property
    def energy(self):
        if "E" not in self._cache:
            self._cache["E"] = np.asarray(self.trajectory.total_energy, float)
        return self._cache["E"]

    @property
    def energy_scale(self):
        tr = self.trajectory
        E = self.energy
        return max(abs(E[0]), np.max(np.abs(tr.kinetic)), np.max(np.abs(tr.potential)), 1e-300)


def _max_change(series):
    """max_t |X(t) - X(0)| for a series of scalars or vectors."""
    a = np.asarray(series, float)
    if a.ndim == 1:
        return float(np.max(np.abs(a - a[0])))
    return float(np.max(np.linalg.norm(a - a[0], axis=1)))


# ---- conservation rows ---------------------------------------------------------------

def _row_energy(f: RunFacts):
    if not f.conservative:
        why = "thermostat" if f.thermostatted else "dissipative force"
        return Check("energy", NA, None, f"not conserved by design ({why})")
    drift = _max_change(f.energy) / f.energy_scale
    return Check("energy", grade(drift, 1e-5, 1e-3), drift, "max |E(t)-E0| / energy scale (incl. reaction ledger)")


def _row_extended_energy(f: RunFacts):
    bath = getattr(f.trajectory, "bath", [])
    if not (hasattr(f.integrator, "bath_energy") and all(x.conservative for x in f.forces)
            and len(bath) == len(f.trajectory.t)):
        return None
    dx = _max_change(f.energy + np.asarray(bath, float)) / f.energy_scale
    return Check("extended energy", grade(dx, 1e-5, 1e-3), dx, "Nose-Hoover conserved quantity: E + thermostat energy")


def _row_momentum(f: RunFacts):
    if f.external or f.walls or f.thermostatted:
        return Check("momentum", NA, None, "external field, walls or thermostat")
    s = f.system
    scale = max(float(np.sum(s.masses[:, None] * np.abs(s.vel))), 1e-300)
    dp = _max_change(f.trajectory.momentum) / scale
    return Check("momentum", grade(dp, 1e-9, 1e-6), dp, "max |P(t)-P0| / sum m|v|")


def _row_angular_momentum(f: RunFacts):
    s = f.system
    if not (f.central and not f.external and s.box is None and not f.thermostatted and s.dim in (2, 3)):
        return Check("angular momentum", NA, None, "needs central forces in open space")
    rxp = np.cross(s.pos, s.masses[:, None] * s.vel)
    per = np.linalg.norm(rxp, axis=-1) if s.dim == 3 else np.abs(rxp)
    dl = _max_change(f.trajectory.angular_momentum) / max(float(np.sum(per)), 1e-300)
    return Check("angular momentum", grade(dl, 1e-9, 1e-6), dl, "max |L(t)-L0| / sum |r x p|")


def _row_mass(f: RunFacts):
    M = np.asarray(f.trajectory.mass, float)
    dm = _max_change(M) / M[0] if M[0] else 0.0
    return Check("mass-energy", grade(dm, 1e-12, 1e-9), dm, "rest mass + released/c^2")


def _row_charge(f: RunFacts):
    dq = _max_change(f.trajectory.charge) / max(float(np.sum(np.abs(f.system.charges))), 1e-300)
    return Check("charge", grade(dq, 1e-12, 1e-9), dq, "total charge")


def _row_thermostat(f: RunFacts):
    from .integrators import Langevin, NoseHoover
    from .units import k_B
    ig, tr, s = f.integrator, f.trajectory, f.system
    if not isinstance(ig, (Langevin, NoseHoover)) or len(tr.t) < 4:
        return None
    # mean temperature over the second half of the records
    second = slice(len(tr.t) // 2, None)
    kin = np.asarray(tr.kinetic[second], float)
    n = np.array([sum(c.values()) for c in tr.counts[second]], float)
    dof = s.dim * (np.maximum(n - 1, 1) if isinstance(ig, NoseHoover) else np.maximum(n, 1))
    T_meas = float(np.mean(2 * kin / (dof * k_B)))
    dev = abs(T_meas - ig.T) / ig.T
    tol = max(0.02, 3 * np.sqrt(2 / (s.dim * max(s.n, 1))))
    status = PASS if dev <= tol else (WARN if dev <= 0.1 else FAIL)
    return Check("thermostat", status, dev, f"mean T over 2nd half {T_meas:.4g} K vs target {ig.T:.4g} K")


CONSERVATION_ROWS = (_row_energy, _row_extended_energy, _row_momentum, _row_angular_momentum, _row_mass,
                     _row_charge, _row_thermostat)


# ---- rows on the final state ----------------------------------------------------------

def timescales(system, forces, max_rows=1500, chunk=16):
    """Shortest characteristic times (s) per kind of physics on this path (advisor.py:9-48):
    orbital times of softened gravity, sigma sqrt(m / eps) of each LJ species, and closest
    approach over the fastest speed. Rows: every particle up to max_rows, else a sample of
    max_rows drawn by default_rng(0); distances are evaluated `chunk` rows at a time."""
    from .forces import Gravity, LennardJones
    from .tree import BarnesHut
    s = system
    out = {}
    if s.n < 2:
        return out
    rows = np.arange(s.n) if s.n <= max_rows else np.random.default_rng(0).choice(s.n, max_rows, replace=False)
    m = s.masses
    grav = [f for f in forces if isinstance(f, (Gravity, BarnesHut))]
    r_min, t_orb = np.inf, {}
    for a in range(0, len(rows), chunk):
        rk = rows[a:a + chunk]
        d = s.pos[None, :, :] - s.pos[rk][:, None, :]
        if s.periodic:
            d -= s.box * np.round(d / s.box)
        r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
        r[np.arange(len(rk)), rk] = np.inf
        r_min = min(r_min, float(np.min(r)))
        for f in grav:
            M = m[rk][:, None] + m[None, :]
            t = float(np.min(np.sqrt((r ** 2 + f.eps2) ** 1.5 / (f.G * M))))
            t_orb[id(f)] = min(t_orb.get(id(f), np.inf), t)
    for f in forces:
        if isinstance(f, (Gravity, BarnesHut)):
            out["gravity (orbital)"] = t_orb[id(f)]
        elif isinstance(f, LennardJones):
            eps_t, sig_t = f._tables(s)
            for k, sp in enumerate(s.species):
                if eps_t[k, k] > 0:
                    t = float(sig_t[k, k] * np.sqrt(sp.mass / eps_t[k, k]))
                    out["Lennard-Jones"] = min(out.get("Lennard-Jones", np.inf), t)
    speed = float(np.max(np.linalg.norm(s.vel, axis=1)))
    if speed > 0 and any(not f.external for f in forces):
        eps2 = max([getattr(f, "eps2", 0.0) for f in forces] or [0.0])
        out["closest approach / speed"] = float(np.sqrt(r_min ** 2 + eps2) / speed)
    return out


def _row_timestep_vs_physics(f: RunFacts):
    from .integrators import Brownian
    if isinstance(f.integrator, Brownian) or f.dt is None:
        return None
    ts = timescales(f.system, f.forces)
    if not ts:
        return None
    kind, tmin = min(ts.items(), key=lambda kv: kv[1])
    ratio = f.dt / tmin
    return Check("timestep vs physics", grade(ratio, 0.02, 0.1), ratio,
                 f"dt / shortest timescale ({kind}, {tmin:.3g} s) at end of run")


def _row_force_approximation(f: RunFacts):
    s = f.system
    approx = [x for x in f.forces if not getattr(x, "exact", True)]
    if not approx or s.n < 2:
        return None
    from .tree import direct_subset
    sub = np.random.default_rng(0).choice(s.n, size=min(s.n, 200), replace=False)
    e = max(rms_rel_error(x.compute(s)[0][sub], direct_subset(x, s, sub)) for x in approx)
    return Check("force approximation", grade(e, 1e-3, 1e-2), e,
                 f"RMS relative force error vs exact summation ({len(sub)} sampled particles)")


def _rows_dsf_vs_ewald(f: RunFacts):
    from .forces import DampedShiftedCoulomb, Ewald
    s = f.system
    out = []
    for x in f.forces:
        if not isinstance(x, DampedShiftedCoulomb):
            continue
        q = s.charges
        if not (s.periodic and s.dim == 3 and np.any(q)):
            continue
        neutral = abs(q.sum()) <= 1e-9 * np.sum(np.abs(q))
        if x.exclude or not neutral or s.n > f.dsf_ewald_max_n:
            out.append(Check("DSF vs Ewald", NA, None,
                             f"needs a neutral system, no exclusions, n <= {f.dsf_ewald_max_n}"))
            continue
        e = rms_rel_error(x.compute(s)[0], Ewald(k=x.k).compute(s)[0])
        out.append(Check("DSF vs Ewald", grade(e, 5e-2, 2e-1), e,
                         "RMS force error of the cutoff method vs Ewald, final snapshot "
                         "(inflated for near-crystalline states)"))
    return out


def _row_electrostatics(f: RunFacts):
    # plain Coulomb is not part of this package, but a caller may mix in the reference's
    if f.system.periodic and any(type(x).__name__ == "Coulomb" for x in f.forces):
        return Check("electrostatics", WARN, None,
                     "plain Coulomb in a periodic box uses minimum image (long range truncated): use Ewald")
    return None


def _row_constraints(f: RunFacts):
    if not hasattr(f.integrator, "max_violation"):
        return None
    v = float(f.integrator.max_violation)
    return Check("constraints", grade(v, 1e-9, 1e-6), v, "worst relative bond-length violation (SHAKE/RATTLE)")


def _row_relativity(f: RunFacts):
    from .units import c as C_LIGHT
    s = f.system
    beta2 = float(np.max(np.sum(s.vel ** 2, axis=1)) / C_LIGHT ** 2) if s.n else 0.0
    return Check("relativity", grade(beta2, 1e-4, 1e-2), beta2,
                 f"(v/c)^2 of fastest particle: Newtonian error ~ this size (v = {np.sqrt(beta2):.3g} c)")


STATE_ROWS = (_row_timestep_vs_physics, _row_force_approximation, _rows_dsf_vs_ewald, _row_electrostatics,
              _row_constraints, _row_relativity)


def _evaluate(rows, facts):
    out = []
    for row in rows:
        got = row(facts)
        if got is None:
            continue
        out.extend(got if isinstance(got, list) else [got])
    return out


def drift_rows(tr, s, forces, integ):
    """The conservation rows (accuracy.py:71-123) of a Trajectory-like record."""
    return _evaluate(CONSERVATION_ROWS, RunFacts(tr, s, list(forces), integ))


UNITS_ROW = ("units", PASS, None, "every input was dimension-checked on entry")


class AccuracyReport:
    def __init__(self, checks):
        self.checks = checks

    def __getitem__(self, name) -> Check:
        for c in self.checks:
            if c.name == name:
                return c
        raise KeyError(name)

    @property
    def ok(self) -> bool:
        return not any(c.status == FAIL for c in self.checks)

    def __str__(self):
        nf = sum(c.status == FAIL for c in self.checks)
        nw = sum(c.status == WARN for c in self.checks)
        if nf:
            head = f"Accuracy report (PROBLEMS FOUND: {nf} failed, {nw} warnings)"
        elif nw:
            head = f"Accuracy report (no failures, {nw} warning{'s' * (nw != 1)})"
        else:
            head = "Accuracy report (all applicable checks passed)"
        return "\n".join([head, *(c.row() for c in self.checks)])

    @classmethod
    def from_run(cls, trajectory, system, forces, integrator, dt=None, dsf_ewald_max_n=2000):
        """Every row but the timestep error, for any run that left a Trajectory and its final
        system (Simulation.run, or a device engine's trajectory() after fetch())."""
        if len(trajectory.t) < 2:
            raise ValueError("run the simulation before asking for a report")
        facts = RunFacts(trajectory, system, list(forces), integrator, dt, dsf_ewald_max_n)
        return cls([Check(*UNITS_ROW), *_evaluate(CONSERVATION_ROWS + STATE_ROWS, facts)])

    @classmethod
    def build(cls, sim, convergence_duration=None, dsf_ewald_max_n=2000):
        report = cls.from_run(sim.trajectory, sim.system, sim.forces, sim.integrator, sim.dt, dsf_ewald_max_n)
        if convergence_duration is not None:
            report.checks.append(timestep_convergence(sim, convergence_duration))
        return report


# ---- time-step error by Richardson extrapolation --------------------------------------

def richardson_check(finals, n, expected):
    """Grade the final positions of runs at dt, dt/2, dt/4 (accuracy.py:208-221): errors
    e1 = rms(x_dt - x_dt/2), e2 = rms(x_dt/2 - x_dt/4) relative to the spread of x_dt/4,
    observed order p = log2(e1 / e2), estimate e1 2^p / (2^p - 1) for p > 0.2."""
    coarse, mid, fine = finals
    size = max(float(np.ptp(fine)), 1e-300)
    e1 = float(np.sqrt(np.mean((coarse - mid) ** 2))) / size
    e2 = float(np.sqrt(np.mean((mid - fine) ** 2))) / size
    floor = 1e-13 * np.sqrt(4 * n)
    if e1 < 10 * floor:
        return Check("timestep error", PASS, e1, "converged to floating-point round-off")
    p = np.log2(e1 / e2)
    est = e1 * 2 ** p / (2 ** p - 1) if p > 0.2 else e1
    if p < 0.5 * expected:
        return Check("timestep error", FAIL if est > 1e-2 else WARN, est,
                     f"observed order {p:.1f} (expected {expected}): chaotic or dt too large; "
                     f"trust statistics, not trajectories")
    return Check("timestep error", grade(est, 1e-6, 1e-3), est,
                 f"est. position error at dt (relative to system size), observed order {p:.1f}/{expected}")


def timestep_convergence(sim, duration, start=None) -> Check:
    """Re-run from the initial state at dt, dt/2 and dt/4 for `duration` (accuracy.py:191-221)."""
    from .units import si
    if sim.integrator.stochastic:
        return Check("timestep error", NA, None, "stochastic run: compare ensemble statistics instead")
    origin = start or sim._initial
    n = max(int(round(si(duration, "s") / sim.dt)), 1)
    finals = []
    for k in (1, 2, 4):
        run = origin.clone(dt=sim.dt / k)
        run.run(steps=n * k, record_every=10 ** 9, positions=False)
        finals.append(run.system.pos.copy())
    return richardson_check(finals, n, sim.integrator.order)
