"""Declarative JSON models (accusim/model.py:1-146) for the force path of this package.

The same file format as the reference, units as strings ("10 fs", "0.6502 kJ/mol"):

    {"dim": 3, "box": "34.68 A", "periodic": true, "seed": 64,
     "species": {"Ar": {"mass": "39.948 u"}},
     "particles": [{"species": "Ar", "count": 864, "random": true}],
     "temperature": "94.4 K",
     "forces": [{"type": "LennardJones", "params": {"Ar": ["0.996 kJ/mol", "3.405 A"]}, "cutoff": 2.5}],
     "integrator": {"type": "VelocityVerlet"},
     "dt": "10 fs", "duration": "10 ps"}

`build(model)` returns the package's Simulation (device forces through the drop-in
compute(), model.py:103-141) with the reference's random-number call sequence, so a model
gives the same initial state in both. The force types are those of this path
(LennardJones, DampedShiftedCoulomb, Ewald, Gravity, BarnesHut) and the integrators all
five of integrators.py; reactions, molecules and the bonded or external forces are not
on this path and are refused with a message naming them.
"""
from __future__ import annotations

import json
import re
from fractions import Fraction

import numpy as np

from . import forces as F
from . import integrators as I
from .simulation import Simulation
from .system import System
from .tree import BarnesHut

# unit -> (SI factor, exponents of (kg, m, s, C, K, mol)); CODATA 2018 and exact values
def _dim(kg=0, m=0, s=0, C=0, K=0, mol=0):
    return (Fraction(kg), Fraction(m), Fraction(s), Fraction(C), Fraction(K), Fraction(mol))


_AMU = 1.66053906660e-27
_EV = 1.602176634e-19
UNITS = {
    "1": (1.0, _dim()),
    "kg": (1.0, _dim(kg=1)), "g": (1e-3, _dim(kg=1)), "mg": (1e-6, _dim(kg=1)), "t": (1e3, _dim(kg=1)),
    "u": (_AMU, _dim(kg=1)), "amu": (_AMU, _dim(kg=1)), "Da": (_AMU, _dim(kg=1)),
    "kDa": (_AMU * 1e3, _dim(kg=1)), "MDa": (_AMU * 1e6, _dim(kg=1)),
    "Msun": (1.98847e30, _dim(kg=1)), "Mearth": (5.9722e24, _dim(kg=1)), "Mjup": (1.89813e27, _dim(kg=1)),
    "m": (1.0, _dim(m=1)), "km": (1e3, _dim(m=1)), "cm": (1e-2, _dim(m=1)), "mm": (1e-3, _dim(m=1)),
    "um": (1e-6, _dim(m=1)), "nm": (1e-9, _dim(m=1)), "pm": (1e-12, _dim(m=1)), "fm": (1e-15, _dim(m=1)),
    "A": (1e-10, _dim(m=1)), "angstrom": (1e-10, _dim(m=1)),
    "au": (1.495978707e11, _dim(m=1)), "ly": (9.4607304725808e15, _dim(m=1)),
    "pc": (3.0856775814913673e16, _dim(m=1)),
    "s": (1.0, _dim(s=1)), "ms": (1e-3, _dim(s=1)), "us": (1e-6, _dim(s=1)), "ns": (1e-9, _dim(s=1)),
    "ps": (1e-12, _dim(s=1)), "fs": (1e-15, _dim(s=1)), "min": (60.0, _dim(s=1)), "h": (3600.0, _dim(s=1)),
    "day": (86400.0, _dim(s=1)), "yr": (365.25 * 86400.0, _dim(s=1)),
    "C": (1.0, _dim(C=1)), "e": (_EV, _dim(C=1)),
    "K": (1.0, _dim(K=1)),
    "mol": (1.0, _dim(mol=1)),
    "N": (1.0, _dim(kg=1, m=1, s=-2)), "J": (1.0, _dim(kg=1, m=2, s=-2)), "kJ": (1e3, _dim(kg=1, m=2, s=-2)),
    "kcal": (4184.0, _dim(kg=1, m=2, s=-2)), "eV": (_EV, _dim(kg=1, m=2, s=-2)),
    "keV": (_EV * 1e3, _dim(kg=1, m=2, s=-2)), "MeV": (_EV * 1e6, _dim(kg=1, m=2, s=-2)),
    "W": (1.0, _dim(kg=1, m=2, s=-3)), "Pa": (1.0, _dim(kg=1, m=-1, s=-2)),
    "bar": (1e5, _dim(kg=1, m=-1, s=-2)), "atm": (101325.0, _dim(kg=1, m=-1, s=-2)),
    "Hz": (1.0, _dim(s=-1)), "L": (1e-3, _dim(m=3)), "V": (1.0, _dim(kg=1, m=2, s=-2, C=-1)),
    "M": (1e3, _dim(m=-3, mol=1)),
}
N_A = 6.02214076e23

_TOKEN = re.compile(r"\s*([A-Za-z0-9_.]+)(?:\^(-?[0-9]+(?:/[0-9]+)?))?\s*")


class UnitError(ValueError):
    pass


def parse_unit(text):
    """'km/s', 'kg*m^2/s^2', 'kJ/mol', 'm s^-1' -> (SI factor, dimension exponents).
    Factors after '/' are divided; a bare space multiplies (units.py:176-211)."""
    text = text.strip().replace("·", "*")
    value, dim = 1.0, [Fraction(0)] * 6
    if not text:
        return value, tuple(dim)
    sign, pos = 1, 0
    while pos < len(text):
        ch = text[pos]
        if ch in "*/":
            sign = -1 if ch == "/" else sign
            pos += 1
            continue
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise UnitError(f"cannot parse unit '{text}' at '{text[pos:]}'")
        name, exp = m.group(1), m.group(2)
        if name in UNITS:
            f, d = UNITS[name]
        else:
            try:
                f, d = float(name), _dim()
            except ValueError:
                raise UnitError(f"unknown unit '{name}' in '{text}'") from None
        p = (Fraction(exp) if exp else Fraction(1)) * sign
        value *= f ** float(p)
        dim = [a + b * p for a, b in zip(dim, d)]
        pos = m.end()
    return value, tuple(dim)


def quantity(x):
    """'3 km/s' -> (SI value, dims); numbers -> (value, dimensionless); lists element-wise
    (one common dimension)."""
    if isinstance(x, str):
        num, _, unit = x.strip().partition(" ")
        f, d = parse_unit(unit.strip() or "1")
        return float(num) * f, d
    if isinstance(x, (list, tuple)):
        items = [quantity(v) for v in x]
        dims = {d for _, d in items}
        if len(dims) > 1:
            raise UnitError(f"mixed units in {x}")
        return np.array([v for v, _ in items], float), (dims.pop() if dims else _dim())
    return x, _dim()


_DIMS = {
    "mass": _dim(kg=1), "length": _dim(m=1), "time": _dim(s=1), "temperature": _dim(K=1), "charge": _dim(C=1),
    "energy": _dim(kg=1, m=2, s=-2), "rate": _dim(s=-1), "diffusion": _dim(m=2, s=-1), "inv_length": _dim(m=-1),
    "gravity": _dim(kg=-1, m=3, s=-2), "none": _dim(), "coulomb": _dim(kg=1, m=3, s=-2, C=-2),
}


def si(x, kind, what):
    """SI value of x checked against dimension `kind` (plain numbers are taken as SI)."""
    v, d = quantity(x)
    if isinstance(x, (str, list, tuple)) and (isinstance(x, str) or any(isinstance(e, str) for e in x)):
        if kind == "energy" and d == _dim(kg=1, m=2, s=-2, mol=-1):
            return v / N_A  # per mole -> per particle (model.py:45-49)
        if d != _DIMS[kind]:
            raise UnitError(f"{what}: '{x}' is not a {kind}")
    return v


FORCE_TYPES = {"LennardJones": F.LennardJones, "DampedShiftedCoulomb": F.DampedShiftedCoulomb, "Ewald": F.Ewald,
               "Gravity": F.Gravity, "BarnesHut": BarnesHut}
INTEGRATORS = {"VelocityVerlet": I.VelocityVerlet, "Yoshida4": I.Yoshida4, "Langevin": I.Langevin,
               "Brownian": I.Brownian, "NoseHoover": I.NoseHoover}
_OFF_PATH = ("Coulomb", "HarmonicBond", "HarmonicAngle", "UniformField", "Drag")

# argument kinds of the forces and integrators (other keys pass through unconverted)
_ARG_KINDS = {
    "cutoff": "none", "skin": "none", "rcut": "length", "alpha": "inv_length", "softening": "length",
    "G": "gravity", "k": "coulomb", "theta": "none", "leaf_size": "none", "kmax": "none", "kcut": "inv_length",
    "temperature": "temperature", "friction": "rate", "period": "time", "seed": "none",
    "pairlist_threshold": "none", "quadrupole": None,
}


def _kwargs(d, what, kinds=None):
    """Constructor arguments in SI; `kinds` overrides the argument kinds of one type."""
    out = {}
    for k, v in d.items():
        kind = (kinds or {}).get(k, _ARG_KINDS.get(k))
        plain = kind is None or (kind == "none" and not isinstance(v, str))
        out[k] = v if plain else si(v, kind, f"{what}.{k}")
    return out


def build(model: dict, precision: str = "fp64"):
    """(Simulation, model) from a JSON model (model.py:103-141) with device forces of the
    given precision ("fp64" or "mixed")."""
    s, forces, integrator, dt = build_parts(model, precision)
    return Simulation(s, forces, integrator=integrator, dt=dt, seed=model.get("seed")), model


def build_parts(model: dict, precision: str = "fp64"):
    """(System, forces, integrator, dt) of a JSON model, before any force is evaluated."""
    for key in ("reactions", "molecules"):
        if model.get(key):
            raise ValueError(f"'{key}' are not on the GPU force path (run this model with accusim itself)")
    dim = model.get("dim", 3)
    box = si(model["box"], "length", "box") if "box" in model else None
    if box is not None and not model.get("periodic", False):
        raise ValueError("reflecting walls (a box without 'periodic') are not on the GPU force path "
                         "(run this model with accusim itself)")
    s = System(dim, box=box, periodic=model.get("periodic", False))
    kinds = {"mass": "mass", "charge": "charge", "radius": "length", "diffusion": "diffusion"}
    for name, sp in model["species"].items():
        s.add_species(name, **{k: si(v, kinds[k], f"species.{name}.{k}") for k, v in sp.items()})
    rng = np.random.default_rng(model.get("seed"))
    for p in model.get("particles", []):
        if p.get("random"):
            if s.box is None:
                raise ValueError("random placement needs a box")
            s.add(p["species"], rng.uniform(0, 1, (p["count"], s.dim)) * s.box)
        else:
            pos = si(p["pos"], "length", "particles.pos")
            vel = None
            if "vel" in p:
                vel, d = quantity(p["vel"])
                if d not in (_dim(m=1, s=-1), _dim()):
                    raise UnitError(f"particles.vel: '{p['vel']}' is not a velocity")
            s.add(p["species"], pos, vel)
    if "temperature" in model:
        s.maxwell_boltzmann(si(model["temperature"], "temperature", "temperature"), rng)
    forces = []
    for f in model.get("forces", []):
        f = dict(f)
        kind = f.pop("type")
        if kind in _OFF_PATH:
            raise ValueError(f"force '{kind}' is not on the GPU force path (run this model with accusim itself)")
        if kind not in FORCE_TYPES:
            raise ValueError(f"unknown force type '{kind}'")
        cls = FORCE_TYPES[kind]
        if cls is F.LennardJones:
            f["params"] = {k: (si(e, "energy", f"{kind}.{k}.epsilon"), si(sg, "length", f"{kind}.{k}.sigma"))
                           for k, (e, sg) in f["params"].items()}
            params = f.pop("params")
            forces.append(cls(params, **_kwargs(f, kind), precision=precision))
        elif cls is BarnesHut:
            forces.append(cls(**_kwargs(f, kind)))
        elif cls is F.DampedShiftedCoulomb:  # its skin is a length (forces.py:405), LJ's in sigma
            forces.append(cls(**_kwargs(f, kind, {"skin": "length"}), precision=precision))
        else:
            forces.append(cls(**_kwargs(f, kind), precision=precision))
    integ = dict(model.get("integrator", {"type": "VelocityVerlet"}))
    name = integ.pop("type")
    if name not in INTEGRATORS:
        raise ValueError(f"integrator '{name}' is not available (Rattle needs constraints, which JSON models "
                         f"do not describe)")
    integrator = INTEGRATORS[name](**_kwargs(integ, name))
    if "dt" not in model:
        raise ValueError("set dt (the reference's timestep advisor is not part of this package)")
    return s, forces, integrator, si(model["dt"], "time", "dt")


def load(path, precision="fp64"):
    with open(path) as fh:
        return build(json.load(fh), precision)
