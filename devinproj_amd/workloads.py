"""Seeded workload builders (the reference's example input builders, same RNG
call sequences): FCC argon (examples/rahman_argon.py:19-29), rocksalt NaCl
(examples/molten_salt.py:24-33), Plummer sphere (examples/star_cluster.py:14-37)
and the uneven reduced-unit LJ batch. Returns devinproj_amd.System objects.
"""
from __future__ import annotations

import numpy as np

from .system import System
from .units import G, N_A, amu, angstrom, k_B, parsec, Msun

AR_MASS = 39.948 * amu
AR_SIGMA = 3.405 * angstrom
AR_EPS = 119.8 * k_B
AR_RHO = 1374.0
AR_T = 94.4
NA_MASS = 22.98977 * amu
CL_MASS = 35.453 * amu
NACL_LJ = {"Na": (0.544e3 / N_A, 2.35 * angstrom), "Cl": (0.4186e3 / N_A, 4.40 * angstrom)}
E_CHARGE = 1.602176634e-19


def argon_params():
    return {"Ar": (AR_EPS, AR_SIGMA)}


def fcc_positions(n_cells, L):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
    ijk = np.stack(np.meshgrid(np.arange(n_cells), np.arange(n_cells), np.arange(n_cells), indexing="ij"),
                   -1).reshape(-1, 3)
    return ((ijk[:, None, :] + base[None, :, :]).reshape(-1, 3)) * L / n_cells


def argon(n_cells=6, seed=64, jitter=0.0, T=AR_T, rho=AR_RHO, sigma=AR_SIGMA):
    """4 n_cells^3 argon atoms on an FCC lattice at rho; jitter (in sigma) from the same rng
    before the Maxwell-Boltzmann velocities (zero net momentum)."""
    L = (4 * n_cells ** 3 * AR_MASS / rho) ** (1 / 3)
    s = System(3, box=L, periodic=True)
    s.add_species("Ar", AR_MASS)
    s.add("Ar", fcc_positions(n_cells, L))
    rng = np.random.default_rng(seed)
    if jitter:
        s.pos = (s.pos + rng.normal(0.0, jitter * sigma, s.pos.shape)) % s.box
    s.maxwell_boltzmann(T, rng)
    return s


def rocksalt(nx=32, ny=32, nz=64, rho=1500.0, T=1200.0, seed=3):
    """nx*ny*nz ions on a rocksalt grid ((x+y+z) even -> Na+) in an orthorhombic box at rho."""
    a = (0.5 * (NA_MASS + CL_MASS) / rho) ** (1 / 3)
    s = System(3, box=np.array([nx, ny, nz], float) * a, periodic=True)
    s.add_species("Na", NA_MASS, E_CHARGE)
    s.add_species("Cl", CL_MASS, -E_CHARGE)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    odd = g.sum(1) % 2 == 1
    s.add("Na", (g[~odd] + 0.5) * a)
    s.add("Cl", (g[odd] + 0.5) * a)
    s.maxwell_boltzmann(T, rng=np.random.default_rng(seed))
    return s


def plummer(n=400, a=parsec, m=Msun, seed=0):
    """Aarseth-Henon-Wielen sampling of a Plummer sphere in virial equilibrium."""
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
    vesc = np.sqrt(2 * G * M / np.sqrt(r ** 2 + a ** 2))
    vel = (q * vesc)[:, None] * iso(n)
    s = System(3)
    s.add_species("star", m)
    s.add("star", pos - pos.mean(0), vel - vel.mean(0))
    return s, a, M


def lj_reduced(n, rho, T=1.0, seed=0, jitter=0.01):
    """Reduced-unit LJ (sigma = eps = m = 1): jittered FCC truncated to n atoms, MB velocities at T."""
    nc = int(np.ceil((n / 4) ** (1 / 3)))
    L = (n / rho) ** (1 / 3)
    pts = fcc_positions(nc, L)[:n]
    s = System(3, box=L, periodic=True)
    s.add_species("A", 1.0)
    rng = np.random.default_rng(seed)
    s.add("A", (pts + rng.normal(0, jitter, pts.shape)) % L)
    s.vel = rng.normal(size=s.pos.shape) * np.sqrt(T)
    s.vel -= s.vel.mean(0)
    return s


def uneven_batch(n_systems=48, seed=7, nmin=4000, nmax=262144):
    """(N, rho, seed) of n_systems LJ systems: N log-uniform in [nmin, nmax], rho uniform in [0.7, 0.9]."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_systems))).astype(int)
    rhos = rng.uniform(0.7, 0.9, n_systems)
    return [(int(n), float(r), int(seed * 1000 + k)) for k, (n, r) in enumerate(zip(ns, rhos))]


def rigid_dimers(n_mol, rho=0.2, bond=1.0, v_scale=0.5, seed=0):
    """n_mol rigid dimers in reduced units (mass 1, bond `bond`) at number density rho: centres
    on a cubic grid, random orientations, each molecule a random translation plus a rotation
    about its centre (velocities that satisfy the bonds), zero net momentum. Returns
    (System, constraints [(id_a, id_b, bond)]) for Rattle."""
    rng = np.random.default_rng(seed)
    L = (2 * n_mol / rho) ** (1 / 3)
    m = int(np.ceil(n_mol ** (1 / 3)))
    g = (np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n_mol] + 0.5) * (L / m)
    u = rng.normal(size=(n_mol, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    w = rng.normal(0, v_scale, (n_mol, 3))
    vc = rng.normal(0, v_scale, (n_mol, 3))
    half = 0.5 * bond * u
    s = System(3, box=L, periodic=True)
    s.add_species("A", 1.0)
    ia = s.add("A", (g - half) % L, vel=vc - np.cross(w, half))
    ib = s.add("A", (g + half) % L, vel=vc + np.cross(w, half))
    s.vel -= s.vel.mean(0)
    return s, [(int(a), int(b), bond) for a, b in zip(ia, ib)]
