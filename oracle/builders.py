"""Seeded input builders for the force-path workloads (test infrastructure only).

Each builder returns an `oracle.accusim_np.System`; the product's Force classes
accept it unchanged (they read pos, vel, sp, ids, box, periodic, species,
masses, charges). RNG call sequences follow the reference's builders so that
seeded inputs are identical by construction.
"""
from __future__ import annotations

import numpy as np

from .accusim_np import AMU, ANGSTROM, G_SI, K_B, MSUN, N_A, PARSEC, System

# Argon, 12-6 LJ (workload parameters)
AR_MASS = 39.948 * AMU
AR_SIGMA = 3.405 * ANGSTROM
AR_EPS = 119.8 * K_B
AR_RHO = 1374.0          # kg/m^3
AR_T = 94.4              # K


def fcc_positions(n_cells, L):
    """FCC lattice of examples/rahman_argon.py:23-24 (base + 0.25 offset, i-j-k-b order)."""
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
    ijk = np.stack(np.meshgrid(np.arange(n_cells), np.arange(n_cells), np.arange(n_cells), indexing="ij"), -1).reshape(-1, 3)
    pts = (ijk[:, None, :] + base[None, :, :]).reshape(-1, 3)
    return pts * L / n_cells


def argon(n_cells=6, seed=64, jitter=0.0, T=AR_T, rho=AR_RHO, sigma=AR_SIGMA):
    """FCC argon at `rho`, Maxwell-Boltzmann velocities at T (examples/rahman_argon.py:19-29).

    n_cells=6 gives the 864-atom CPU-runnable case; n_cells=64 the 1,048,576-atom case
    (jitter = 0.01 sigma: positions += N(0, jitter) from the same rng, then wrapped,
    before the velocities are drawn)."""
    m = AR_MASS
    L = (4 * n_cells ** 3 * m / rho) ** (1 / 3)
    s = System(3, box=L, periodic=True)
    s.add_species("Ar", m)
    s.add("Ar", fcc_positions(n_cells, L))
    rng = np.random.default_rng(seed)
    if jitter:
        s.pos = (s.pos + rng.normal(0.0, jitter * sigma, s.pos.shape)) % s.box
    s.maxwell_boltzmann(T, rng)
    return s


def argon_params():
    return {"Ar": (AR_EPS, AR_SIGMA)}


# NaCl: Lorentz-Berthelot LJ parameters (kJ/mol -> J per ion) and masses
NA_MASS = 22.98977 * AMU
CL_MASS = 35.453 * AMU
NACL_LJ = {"Na": (0.544e3 / N_A, 2.35 * ANGSTROM), "Cl": (0.4186e3 / N_A, 4.40 * ANGSTROM)}


def rocksalt(nx=32, ny=32, nz=64, rho=1500.0, T=1200.0, seed=3, charge=1.602176634e-19):
    """Rocksalt ions on an nx*ny*nz site grid, (x+y+z) even -> Na+ (examples/molten_salt.py:24-33,
    orthorhombic box so 65,536 = 32*32*64 ions fit), MB velocities at T."""
    pair_mass = NA_MASS + CL_MASS
    a = (0.5 * pair_mass / rho) ** (1 / 3)
    box = np.array([nx, ny, nz], float) * a
    s = System(3, box=box, periodic=True)
    s.add_species("Na", NA_MASS, charge)
    s.add_species("Cl", CL_MASS, -charge)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    odd = g.sum(1) % 2 == 1
    s.add("Na", (g[~odd] + 0.5) * a)
    s.add("Cl", (g[odd] + 0.5) * a)
    s.maxwell_boltzmann(T, rng=np.random.default_rng(seed))
    return s


def rocksalt_unit(n):
    """tests/test_dsf.py:7-15: unit-spacing rocksalt, unit charges and masses, box n."""
    s = System(3, box=float(n), periodic=True)
    s.add_species("Na", 1.0, 1.0)
    s.add_species("Cl", 1.0, -1.0)
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    odd = g.sum(1) % 2 == 1
    s.add("Na", g[~odd].astype(float))
    s.add("Cl", g[odd].astype(float))
    return s


def nacl_cells(n_cells=1):
    """tests/test_forces.py:79-87: 2n x 2n x 2n NaCl lattice, spacing 1, box 2n."""
    s = System(3, box=2.0 * n_cells, periodic=True)
    s.add_species("Na", 1.0, 1.0)
    s.add_species("Cl", 1.0, -1.0)
    for x in range(2 * n_cells):
        for y in range(2 * n_cells):
            for z in range(2 * n_cells):
                s.add("Na" if (x + y + z) % 2 == 0 else "Cl", [[x, y, z]])
    return s


def plummer(n=400, a=PARSEC, m=MSUN, seed=0):
    """Aarseth-Henon-Wielen sampling, examples/star_cluster.py:14-37 (same rng call sequence)."""
    rng = np.random.default_rng(seed)
    M = n * m
    r = a / np.sqrt(rng.uniform(1e-3, 0.99, n) ** (-2 / 3) - 1)

    def iso(k):
        u = rng.normal(size=(k, 3))
        return u / np.linalg.norm(u, axis=1)[:, None]

    pos = r[:, None] * iso(n)
    q = np.empty(n)
    for i in range(n):
        while True:
            x, y = rng.uniform(0, 1), rng.uniform(0, 0.1)
            if y < x ** 2 * (1 - x ** 2) ** 3.5:
                q[i] = x
                break
    vesc = np.sqrt(2 * G_SI * M / np.sqrt(r ** 2 + a ** 2))
    vel = (q * vesc)[:, None] * iso(n)
    s = System(3)
    s.add_species("star", m)
    s.add("star", pos - pos.mean(0), vel - vel.mean(0))
    return s, a, M


def lj_reduced(n, rho, T=1.0, seed=0, jitter=0.01):
    """LJ system in reduced units (sigma = eps = m = 1): jittered FCC with
    n_c = ceil((n/4)^(1/3)) cells truncated to n atoms, box (n/rho)^(1/3), MB at T."""
    nc = int(np.ceil((n / 4) ** (1 / 3)))
    L = (n / rho) ** (1 / 3)
    pts = fcc_positions(nc, L)[:n]
    s = System(3, box=L, periodic=True)
    s.add_species("A", 1.0)
    rng = np.random.default_rng(seed)
    s.add("A", (pts + rng.normal(0, jitter, pts.shape)) % L)
    # K_B is not 1 in SI; draw MB directly in reduced units
    s.vel = rng.normal(size=s.pos.shape) * np.sqrt(T)
    s.vel -= s.vel.mean(0)
    return s


def uneven_batch(n_systems=48, seed=7, nmin=4000, nmax=262144):
    """48 LJ systems: N log-uniform in [nmin, nmax], rho uniform in [0.7, 0.9], T* = 1."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_systems))).astype(int)
    rhos = rng.uniform(0.7, 0.9, n_systems)
    return [(int(n), float(r), int(seed * 1000 + k)) for k, (n, r) in enumerate(zip(ns, rhos))]
