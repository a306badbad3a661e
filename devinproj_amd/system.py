"""Host-side particle state read by the GPU force classes (accusim/system.py:11-160).

Layout (what every Force.compute and the engines read): `pos`, `vel` float64 [N, dim]
C-contiguous, `sp` int species index, `ids` int, `eint` float64 [N], `box` float64 [dim]
or None, `periodic` bool, and the species table `species`. The force classes accept any
object with these attributes, the reference's own System included.

Per-species properties are kept as one float64 table that is rebuilt only when a species
is added, so `masses` and `charges` are single gathers (the reference rebuilds the table
on every call, system.py:83-89).

Boundaries: periodic boxes (positions wrapped into [0, L), system.py:127-129) and open
space (box None). A bounded box that is not periodic has reflecting walls in the
reference (system.py:130-140); those are not part of this package's force path, so
stepping such a system raises NotImplementedError.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .units import k_B, si


@dataclass(frozen=True)
class Species:
    name: str
    mass: float
    charge: float = 0.0
    radius: float = 0.0
    diffusion: float = 0.0


# columns of the cached species table
_COLUMNS = ("mass", "charge", "radius", "diffusion")


class System:
    """Particles in `dim` dimensions, SI units. box: edge lengths or None (open space)."""

    def __init__(self, dim: int = 3, box=None, periodic: bool = False):
        self.dim = int(dim)
        if box is None:
            self.box = None
        else:
            edges = np.array(np.broadcast_to(si(box, "m"), (self.dim,)), dtype=float)
            self.box = edges
        self.periodic = bool(periodic) and self.box is not None
        self.species: list[Species] = []
        self._table = np.zeros((0, len(_COLUMNS)))
        self.pos = np.zeros((0, self.dim))
        self.vel = np.zeros((0, self.dim))
        self.sp = np.zeros(0, dtype=int)
        self.ids = np.zeros(0, dtype=int)
        self.eint = np.zeros(0)
        self.wall_impulse = 0.0      # stays 0: no reflecting walls here
        self.constrained_dof = 0     # rigid-bond degrees of freedom (set by Rattle)
        self._next_id = 0

    # ---- species table -------------------------------------------------------------

    def add_species(self, name, mass, charge=0.0, radius=0.0, diffusion=0.0) -> Species:
        if any(x.name == name for x in self.species):
            raise ValueError(f"species '{name}' already defined")
        entry = Species(name, si(mass, "kg"), si(charge, "C"), si(radius, "m"), si(diffusion, "m^2/s"))
        if not entry.mass > 0:
            raise ValueError("mass must be positive")
        self.species.append(entry)
        row = np.array([[getattr(entry, c) for c in _COLUMNS]], dtype=float)
        self._table = np.vstack([self._table, row])
        return entry

    def species_index(self, name: str) -> int:
        for k, x in enumerate(self.species):
            if x.name == name:
                return k
        raise KeyError(f"unknown species '{name}'")

    def _per_particle(self, column: str) -> np.ndarray:
        if not len(self.sp):
            return np.zeros(0)
        return self._table[:, _COLUMNS.index(column)][self.sp]

    @property
    def masses(self) -> np.ndarray:
        return self._per_particle("mass")

    @property
    def charges(self) -> np.ndarray:
        return self._per_particle("charge")

    # ---- particles ------------------------------------------------------------------

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def add(self, species: str, pos, vel=None, eint=0.0) -> np.ndarray:
        """Append one particle (pos of shape (dim,)) or several ((k, dim)); returns their ids."""
        kind = self.species_index(species)
        new_pos = np.array(np.atleast_2d(si(pos, "m")), dtype=float)
        if new_pos.shape[1] != self.dim:
            raise ValueError(f"positions must have {self.dim} components")
        k = new_pos.shape[0]
        if vel is None:
            new_vel = np.zeros((k, self.dim))
        else:
            new_vel = np.array(np.broadcast_to(np.atleast_2d(si(vel, "m/s")), new_pos.shape), dtype=float)
        new_ids = self._next_id + np.arange(k)
        self._next_id += k
        self.pos = np.concatenate([self.pos, new_pos])
        self.vel = np.concatenate([self.vel, new_vel])
        self.sp = np.concatenate([self.sp, np.full(k, kind, dtype=int)])
        self.ids = np.concatenate([self.ids, new_ids])
        self.eint = np.concatenate([self.eint, np.full(k, float(eint)) if np.ndim(eint) == 0
                                    else np.asarray(eint, float).reshape(k)])
        return new_ids

    def remove(self, mask: np.ndarray):
        keep = np.logical_not(mask)
        for name in ("pos", "vel", "sp", "ids", "eint"):
            setattr(self, name, getattr(self, name)[keep])

    def count(self, name: str) -> int:
        return int(np.count_nonzero(self.sp == self.species_index(name)))

    def counts(self) -> dict[str, int]:
        per = np.bincount(self.sp, minlength=len(self.species)) if len(self.sp) else np.zeros(len(self.species))
        return {x.name: int(per[k]) for k, x in enumerate(self.species)}

    # ---- geometry -------------------------------------------------------------------

    def displacements(self, idx_a=None, idx_b=None) -> np.ndarray:
        """r_b - r_a for every (a, b), nearest image in a periodic box: shape (na, nb, dim)."""
        src = self.pos if idx_a is None else self.pos[idx_a]
        dst = self.pos if idx_b is None else self.pos[idx_b]
        d = dst[None, :, :] - src[:, None, :]
        if self.periodic:
            d -= self.box * np.round(d / self.box)
        return d

    def apply_boundaries(self):
        if self.box is None:
            return
        if not self.periodic:
            raise NotImplementedError("reflecting walls (a bounded box that is not periodic) are not part of "
                                      "this package; use periodic=True or box=None")
        self.pos %= self.box

    # ---- conserved quantities (summed as Simulation.record needs them) --------------

    def kinetic_energy(self) -> float:
        return 0.5 * float(np.sum(self.masses[:, None] * self.vel ** 2))

    def momentum(self) -> np.ndarray:
        return np.sum(self.masses[:, None] * self.vel, axis=0)

    def angular_momentum(self) -> np.ndarray:
        return np.atleast_1d(np.sum(np.cross(self.pos, self.masses[:, None] * self.vel), axis=0))

    def center_of_mass_velocity(self) -> np.ndarray:
        return self.momentum() / np.sum(self.masses)

    def temperature(self) -> float:
        dof = max(self.dim * (self.n - 1) - self.constrained_dof, 1)
        return 2.0 * self.kinetic_energy() / (dof * k_B)

    # ---- copies and initial velocities ----------------------------------------------

    def copy(self) -> "System":
        twin = System(self.dim, box=None if self.box is None else self.box.copy(), periodic=self.periodic)
        twin.species = list(self.species)
        twin._table = self._table.copy()
        for name in ("pos", "vel", "sp", "ids", "eint"):
            setattr(twin, name, getattr(self, name).copy())
        twin.wall_impulse = self.wall_impulse
        twin.constrained_dof = self.constrained_dof
        twin._next_id = self._next_id
        return twin

    def maxwell_boltzmann(self, temperature, rng=None, zero_momentum=True):
        """Velocities at `temperature` (K): one normal draw of shape (N, dim) scaled per
        particle by sqrt(kT/m), then the centre-of-mass velocity removed. The draw and the
        arithmetic are those of system.py:152-160, so a seed gives the reference's state."""
        gen = np.random.default_rng() if rng is None else rng
        spread = np.sqrt(k_B * si(temperature, "K") / self.masses)
        self.vel = gen.normal(size=self.pos.shape) * spread[:, None]
        if zero_momentum and self.n > 1:
            self.vel -= self.center_of_mass_velocity()
