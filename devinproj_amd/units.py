"""SI constants (CODATA 2018) and the plain-number pass-through of accusim's si().

accusim/units.py:214-233. Numbers and arrays are taken as SI; objects with an
`si(unit)` method (accusim Quantity) are converted by that method, so forces
built with the reference's Quantity inputs work unchanged.
"""
from __future__ import annotations

import math

import numpy as np

G = 6.67430e-11
c = 299792458.0
k_B = 1.380649e-23
N_A = 6.02214076e23
e_charge = 1.602176634e-19
epsilon_0 = 8.8541878128e-12
k_e = 1 / (4 * math.pi * epsilon_0)
amu = 1.66053906660e-27
Msun = 1.98847e30
parsec = 3.0856775814913673e16
angstrom = 1e-10
fs = 1e-15


def si(x, unit=None):
    """SI float (or float array) of x; Quantity-like objects are converted by x.si(unit)."""
    if hasattr(x, "si") and callable(x.si):
        return x.si(unit)
    if hasattr(x, "value") and hasattr(x, "dim"):
        return x.value
    return np.asarray(x, dtype=float) if not np.isscalar(x) else float(x)
