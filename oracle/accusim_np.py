"""NumPy float64 restatement of accusim's non-bonded force path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Every function follows one reference function operation for operation; the
reference file:line is cited in each docstring (paths relative to the
reference repository root). The reference is read as text, not imported.

How each part is pinned (tests in tests/test_oracle_*.py):
  cell_list_pairs   brute-force O(N^2) minimum-image filter (set equality),
                    ncell<3 / open-box special cases
  VerletList        skin-vs-fresh equality, rebuild-count bounds
                    (tests/test_nose_hoover.py:37-48, tests/test_dsf.py:119-139)
  LennardJones      F = 0 at 2^(1/6) sigma, F = -grad U, Newton 3, dense vs
                    cell (tests/test_forces.py:30-76), two-body closed form
  DampedShiftedCoulomb  Madelung within 2e-3, vanishing at rc with the exact
                    self term, F = -grad U, dense vs cell, exclusions
                    (tests/test_dsf.py:18-61, 94-102), within 3% of Ewald
  Gravity           two-body closed form, F = -grad U, Kepler closure
  Ewald             NaCl Madelung 1.747564594633 at rel 1e-7 with zero
                    lattice forces (tests/test_forces.py:90-97), F = -grad U,
                    independence of alpha at fixed accuracy
  BarnesHut         theta=0 equals direct to 1e-12, error < 2e-3 / 1e-2 at
                    theta 0.3 / 0.6 (tests/test_tree.py:21-29), golden vectors
  VelocityVerlet / Yoshida4  harmonic oscillator closed form, Kepler closure,
                    observed order 2 / 4 via timestep_convergence
  Rattle            the reference's known answers (tests/test_rattle.py:6-44):
                    rigid rotor keeps its length, energy and angular momentum
                    to 1e-9; an LJ collision with a rigid dumbbell keeps the
                    bond to 1e-10, momentum to 1e-12, spins it up, and its
                    energy error is second order in dt; a closed-form rigid
                    rotor (uniform rotation at the exact angular velocity)
  record / report   drift formulas on a trajectory with a known answer,
                    golden vectors of a short argon run; report_rows against the
                    reference's own report rows written for small argon, Plummer
                    Barnes-Hut, NaCl DSF and rigid-dimer runs (golden vectors)
  timescales        closed forms: two-body softened orbit, sigma sqrt(m / eps),
                    closest approach over speed; reference values (golden vectors)
  Langevin          the reference's thermostat test (tests/test_dynamics.py:80-95:
                    50 K -> 200 K within 8 %), the closed form at T = 0 (velocity
                    times c1 per step)
  NoseHoover        the reference's test (tests/test_nose_hoover.py:20-36): target
                    temperature within 5 %, extended energy kept to 1e-3, its error
                    shrinking like dt^2
  Brownian          free diffusion: positions equal the summed noise, mean square
                    displacement 6 D t
  philox4x32_10     the Random123 known-answer vectors; the normals' moments
  rdf               pair counts of a perfect FCC crystal (12 N / 2 nearest-neighbour
                    pairs, 6 N / 2 at the next shell), brute-force pair loop on a
                    random box, g -> 1 for an ideal gas
  msd / vacf        ballistic closed forms (msd = <v^2> t^2, C = <v^2>), unwrapping
                    across box faces; green_kubo_diffusion = kT / (m gamma) for
                    free Langevin particles (tests/test_analysis.py:22-35)
  pressure          ideal gas P = N kT / V and the virial of LJ pairs
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.special import erfc as _scipy_erfc

# Physical constants, CODATA 2018 (accusim/units.py:226-233, unit table 151-171).
G_SI = 6.67430e-11
C_LIGHT = 299792458.0
K_B = 1.380649e-23
N_A = 6.02214076e23
E_CHARGE = 1.602176634e-19
EPSILON_0 = 8.8541878128e-12
K_E = 1 / (4 * math.pi * EPSILON_0)
AMU = 1.66053906660e-27
MSUN = 1.98847e30
PARSEC = 3.0856775814913673e16
ANGSTROM = 1e-10
FS = 1e-15


@dataclass(frozen=True)
class Species:
    """accusim/system.py:11-17."""
    name: str
    mass: float
    charge: float = 0.0
    diffusion: float = 0.0


class System:
    """Particle state, accusim/system.py:20-160 (the parts on the force path)."""

    def __init__(self, dim=3, box=None, periodic=False):
        # system.py:27-40
        self.dim = dim
        self.box = None if box is None else np.broadcast_to(np.asarray(box, float), (dim,)).astype(float).copy()
        self.periodic = periodic and self.box is not None
        self.species = []
        self._index = {}
        self.pos = np.zeros((0, dim))
        self.vel = np.zeros((0, dim))
        self.sp = np.zeros(0, dtype=int)
        self.ids = np.zeros(0, dtype=int)
        self._next_id = 0

    def add_species(self, name, mass, charge=0.0, diffusion=0.0):
        # system.py:42-50
        if name in self._index:
            raise ValueError(f"species '{name}' already defined")
        s = Species(name, float(mass), float(charge), float(diffusion))
        if s.mass <= 0:
            raise ValueError("mass must be positive")
        self._index[name] = len(self.species)
        self.species.append(s)
        return s

    def add(self, species, pos, vel=None):
        # system.py:58-72
        p = np.atleast_2d(np.asarray(pos, float)).astype(float)
        v = np.zeros_like(p) if vel is None else np.broadcast_to(np.atleast_2d(np.asarray(vel, float)), p.shape).astype(float)
        n = len(p)
        new_ids = np.arange(self._next_id, self._next_id + n)
        self._next_id += n
        self.pos = np.vstack([self.pos, p])
        self.vel = np.vstack([self.vel, v])
        self.sp = np.concatenate([self.sp, np.full(n, self._index[species])])
        self.ids = np.concatenate([self.ids, new_ids])
        return new_ids

    def species_index(self, name):
        # system.py:52-53
        return self._index[name]

    @property
    def n(self):
        return len(self.pos)

    @property
    def masses(self):
        # system.py:83-85
        return np.array([s.mass for s in self.species])[self.sp] if self.n else np.zeros(0)

    @property
    def charges(self):
        # system.py:87-89
        return np.array([s.charge for s in self.species])[self.sp] if self.n else np.zeros(0)

    def counts(self):
        # system.py:94-95
        return {s.name: int(np.sum(self.sp == i)) for i, s in enumerate(self.species)}

    def displacements(self, idx_a=None, idx_b=None):
        # system.py:97-104: r_j - r_i, minimum image if periodic
        a = self.pos if idx_a is None else self.pos[idx_a]
        b = self.pos if idx_b is None else self.pos[idx_b]
        d = b[None, :, :] - a[:, None, :]
        if self.periodic:
            d -= self.box * np.round(d / self.box)
        return d

    def kinetic_energy(self):
        # system.py:106-107
        return 0.5 * float(np.sum(self.masses[:, None] * self.vel ** 2))

    def momentum(self):
        # system.py:109-110
        return np.sum(self.masses[:, None] * self.vel, axis=0)

    def angular_momentum(self):
        # system.py:112-114
        L = np.cross(self.pos, self.masses[:, None] * self.vel)
        return np.atleast_1d(np.sum(L, axis=0))

    def center_of_mass_velocity(self):
        # system.py:116-117
        return self.momentum() / np.sum(self.masses)

    def apply_boundaries(self):
        # system.py:124-129 (periodic wrap; reflecting walls are not on the path)
        if self.box is None:
            return
        if self.periodic:
            self.pos %= self.box
            return
        raise NotImplementedError("reflecting walls are outside the force path")

    def copy(self):
        # system.py:142-150
        other = System.__new__(System)
        other.__dict__.update(self.__dict__)
        for k in ("pos", "vel", "sp", "ids"):
            setattr(other, k, getattr(self, k).copy())
        other.species = list(self.species)
        other._index = dict(self._index)
        other.box = None if self.box is None else self.box.copy()
        return other

    def maxwell_boltzmann(self, temperature, rng=None, zero_momentum=True):
        # system.py:152-160
        rng = rng or np.random.default_rng()
        sigma = np.sqrt(K_B * float(temperature) / self.masses)[:, None]
        self.vel = rng.normal(size=self.pos.shape) * sigma
        if zero_momentum and self.n > 1:
            self.vel -= self.center_of_mass_velocity()


class Force:
    """forces.py:15-21."""
    conservative = True
    external = False
    central = True

    def compute(self, system):
        raise NotImplementedError


def _pairs(system):
    """forces.py:24-28."""
    d = system.displacements()
    r2 = np.einsum("ijk,ijk->ij", d, d)
    np.fill_diagonal(r2, np.inf)
    return d, r2


def _erfc(x):
    """forces.py:324-330 (scipy branch)."""
    return _scipy_erfc(x)


def cell_grid(s, rc):
    """Cell geometry of forces.py:281-292: (lo, span, ncell, cell[N,3], flat[N])."""
    pos = s.pos
    if s.box is not None:
        lo, span = np.zeros(s.dim), s.box
    else:
        lo = pos.min(axis=0)
        span = np.maximum(pos.max(axis=0) - lo, rc)
    ncell = np.maximum((span // rc).astype(int), 1)
    if s.periodic:
        ncell = np.where(ncell < 3, 1, ncell)
    cell = np.minimum(((pos - lo) / span * ncell).astype(int), ncell - 1)
    cell = np.maximum(cell, 0)
    flat = np.ravel_multi_index(cell.T, ncell)
    return lo, span, ncell, cell, flat


def cell_list_pairs(s, rc):
    """All pairs (i < j) with minimum-image |d|^2 < rc^2: forces.py:279-321."""
    pos = s.pos
    lo, span, ncell, cell, flat = cell_grid(s, rc)
    order = np.argsort(flat, kind="stable")
    sorted_flat = flat[order]
    starts = np.searchsorted(sorted_flat, np.arange(np.prod(ncell)))
    ends = np.searchsorted(sorted_flat, np.arange(np.prod(ncell)), side="right")
    offsets = np.array(np.meshgrid(*[[-1, 0, 1]] * s.dim)).reshape(s.dim, -1).T
    I, J = [], []
    for c in range(int(np.prod(ncell))):
        a = order[starts[c]:ends[c]]
        if not len(a):
            continue
        cidx = np.array(np.unravel_index(c, ncell))
        neigh = cidx + offsets
        if s.periodic:
            neigh %= ncell
        else:
            neigh = neigh[np.all((neigh >= 0) & (neigh < ncell), axis=1)]
        nflat = np.unique(np.ravel_multi_index(neigh.T, ncell))
        b = np.concatenate([order[starts[n]:ends[n]] for n in nflat])
        ii, jj = np.meshgrid(a, b, indexing="ij")
        keep = ii < jj
        I.append(ii[keep])
        J.append(jj[keep])
    if not I:
        return np.zeros(0, int), np.zeros(0, int)
    i, j = np.concatenate(I), np.concatenate(J)
    d = pos[j] - pos[i]
    if s.periodic:
        d -= s.box * np.round(d / s.box)
    close = np.einsum("ij,ij->i", d, d) < rc ** 2
    return i[close], j[close]


class VerletList:
    """forces.py:78-98: reuse until max minimum-image displacement >= skin/2."""

    def __init__(self):
        self._state = None
        self.rebuilds = 0

    def pairs(self, s, rc, skin):
        st = self._state
        if st is not None and st[3].shape == s.ids.shape and np.array_equal(st[3], s.ids) \
                and np.array_equal(st[4], s.box) and st[5] == rc and skin > 0:
            disp = s.pos - st[2]
            if s.periodic:
                disp -= s.box * np.round(disp / s.box)
            if float(np.max(np.einsum("ij,ij->i", disp, disp))) < (0.5 * skin) ** 2:
                return st[0], st[1]
        i, j = cell_list_pairs(s, rc + skin)
        self._state = (i, j, s.pos.copy(), s.ids.copy(), None if s.box is None else s.box.copy(), rc)
        self.rebuilds += 1
        return i, j


class LennardJones(Force):
    """forces.py:101-175: truncated-and-shifted 12-6 LJ, Lorentz-Berthelot mixing."""

    def __init__(self, params, cutoff=2.5, pairlist_threshold=400, skin=0.3):
        self.params = {k: (float(e), float(sg)) for k, (e, sg) in params.items()}
        self.cutoff = cutoff
        self.pairlist_threshold = pairlist_threshold
        self.skin = skin
        self._vl = VerletList()

    @property
    def rebuilds(self):
        return self._vl.rebuilds

    def _tables(self, s):
        # forces.py:144-155
        n = len(s.species)
        eps = np.zeros((n, n))
        sig = np.ones((n, n))
        for i, a in enumerate(s.species):
            for j, b in enumerate(s.species):
                if a.name in self.params and b.name in self.params:
                    ea, sa = self.params[a.name]
                    eb, sb = self.params[b.name]
                    eps[i, j] = np.sqrt(ea * eb)
                    sig[i, j] = 0.5 * (sa + sb)
        return eps, sig

    def _compute_pairlist(self, s):
        # forces.py:119-142
        eps_t, sig_t = self._tables(s)
        smax = float(np.max(np.where(eps_t > 0, sig_t, 0)))
        i, j = self._vl.pairs(s, self.cutoff * smax, self.skin * smax)
        f = np.zeros_like(s.pos)
        if not len(i):
            return f, 0.0
        d = s.pos[j] - s.pos[i]
        if s.periodic:
            d -= s.box * np.round(d / s.box)
        r2 = np.einsum("ij,ij->i", d, d)
        eps = eps_t[s.sp[i], s.sp[j]]
        sig = sig_t[s.sp[i], s.sp[j]]
        mask = (r2 < (self.cutoff * sig) ** 2) & (eps > 0)
        d, r2, eps, sig, i, j = d[mask], r2[mask], eps[mask], sig[mask], i[mask], j[mask]
        sr6 = (sig ** 2 / r2) ** 3
        src6 = (1 / self.cutoff ** 2) ** 3
        U = float(np.sum(4 * eps * (sr6 ** 2 - sr6 - (src6 ** 2 - src6))))
        coef = -24 * eps * (2 * sr6 ** 2 - sr6) / r2
        fij = coef[:, None] * d
        np.add.at(f, i, fij)
        np.add.at(f, j, -fij)
        return f, U

    def compute(self, s):
        # forces.py:157-175
        if s.n < 2:
            return np.zeros_like(s.pos), 0.0
        if s.n >= self.pairlist_threshold:
            return self._compute_pairlist(s)
        eps_t, sig_t = self._tables(s)
        eps = eps_t[s.sp[:, None], s.sp[None, :]]
        sig = sig_t[s.sp[:, None], s.sp[None, :]]
        d, r2 = _pairs(s)
        rc2 = (self.cutoff * sig) ** 2
        mask = (r2 < rc2) & (eps > 0)
        sr2 = np.where(mask, sig ** 2 / np.where(mask, r2, 1.0), 0.0)
        sr6 = sr2 ** 3
        src6 = (1 / self.cutoff ** 2) ** 3
        u = np.where(mask, 4 * eps * (sr6 ** 2 - sr6 - (src6 ** 2 - src6)), 0.0)
        coef = np.where(mask, -24 * eps * (2 * sr6 ** 2 - sr6) / np.where(mask, r2, 1.0), 0.0)
        f = np.einsum("ij,ijk->ik", coef, d)
        return f, 0.5 * float(np.sum(u))


class DampedShiftedCoulomb(Force):
    """forces.py:391-452: Fennell-Gezelter damped shifted force electrostatics."""

    def __init__(self, rcut, alpha=None, k=K_E, pairlist_threshold=400, exclude=(), skin=None):
        self.pairlist_threshold = pairlist_threshold
        self.exclude = [tuple(p) for p in exclude]
        self.k = float(k)
        self.rcut = float(rcut)
        self.alpha = 2.0 / self.rcut if alpha is None else float(alpha)
        self.skin = 0.15 * self.rcut if skin is None else float(skin)
        self._vl = VerletList()

    def compute(self, s):
        q = s.charges
        f = np.zeros_like(s.pos)
        if not np.any(q):
            return f, 0.0
        a, rc = self.alpha, self.rcut
        c = 2 * a / np.sqrt(np.pi)
        erc = float(_erfc(np.array(a * rc)))
        fshift = erc / rc ** 2 + c * np.exp(-(a * rc) ** 2) / rc
        where = {pid: n for n, pid in enumerate(s.ids)} if self.exclude else {}
        ex = np.array([(where[x], where[y]) for x, y in self.exclude if x in where and y in where], dtype=int).reshape(-1, 2)
        if s.n >= self.pairlist_threshold:
            # forces.py:419-438
            i, j = self._vl.pairs(s, rc, self.skin)
            if len(ex):
                lo, hi = np.minimum(ex[:, 0], ex[:, 1]), np.maximum(ex[:, 0], ex[:, 1])
                keep = ~np.isin(i * s.n + j, lo * s.n + hi)
                i, j = i[keep], j[keep]
            d = s.pos[j] - s.pos[i]
            if s.periodic:
                d -= s.box * np.round(d / s.box)
            r = np.sqrt(np.einsum("ij,ij->i", d, d))
            inside = r < rc
            i, j, d, r = i[inside], j[inside], d[inside], r[inside]
            er = _erfc(a * r)
            qq = q[i] * q[j]
            U = float(np.sum(qq * (er / r - erc / rc + fshift * (r - rc))))
            U -= float(np.sum(q ** 2)) * (erc / (2 * rc) + a / np.sqrt(np.pi))
            fv = (qq * ((er / r ** 2 + c * np.exp(-(a * r) ** 2) / r) - fshift) / r)[:, None] * d
            np.add.at(f, i, -fv)
            np.add.at(f, j, fv)
            return self.k * f, self.k * U
        # forces.py:439-452
        d, r2 = _pairs(s)
        r = np.sqrt(r2)
        inside = r < rc
        rr = np.where(inside, r, 1.0)
        er = _erfc(a * rr)
        qq = np.where(inside, q[:, None] * q[None, :], 0.0)
        if len(ex):
            qq[ex[:, 0], ex[:, 1]] = 0.0
            qq[ex[:, 1], ex[:, 0]] = 0.0
        U = 0.5 * float(np.sum(qq * (er / rr - erc / rc + fshift * (rr - rc))))
        U -= float(np.sum(q ** 2)) * (erc / (2 * rc) + a / np.sqrt(np.pi))
        fmag = qq * ((er / rr ** 2 + c * np.exp(-(a * rr) ** 2) / rr) - fshift) / rr
        f = -self.k * np.einsum("ij,ijk->ik", fmag, d)
        return f, self.k * U


class Gravity(Force):
    """forces.py:31-48; `chunk` rows at a time so large N fits in memory."""

    def __init__(self, G=G_SI, softening=0.0, chunk=None):
        self.G = float(G)
        self.eps2 = float(softening) ** 2
        self.chunk = chunk

    def compute(self, s):
        if s.n < 2:
            return np.zeros_like(s.pos), 0.0
        if self.chunk is None:
            d, r2 = _pairs(s)
            r2 = r2 + self.eps2
            m = s.masses
            inv_r = 1 / np.sqrt(r2)
            mm = self.G * m[:, None] * m[None, :]
            f = np.einsum("ij,ijk->ik", mm * inv_r ** 3, d)
            U = -0.5 * float(np.sum(mm * inv_r))
            return f, U
        m = s.masses
        f = np.zeros_like(s.pos)
        U = 0.0
        for a in range(0, s.n, self.chunk):
            rows = np.arange(a, min(a + self.chunk, s.n))
            d = s.displacements(rows)
            r2 = np.einsum("ijk,ijk->ij", d, d)
            r2[np.arange(len(rows)), rows] = np.inf
            inv_r = 1 / np.sqrt(r2 + self.eps2)
            mm = self.G * m[rows, None] * m[None, :]
            f[rows] = np.einsum("ij,ijk->ik", mm * inv_r ** 3, d)
            U += -0.5 * float(np.sum(mm * inv_r))
        return f, U


def kvectors(box, alpha, kmax=12, kcut=None):
    """forces.py:347-357: k = 2 pi n / L for 0 < |n|^2 <= kmax^2, weight exp(-k^2/4a^2)/k^2.

    With `kcut` (an extension) the set is the ellipsoid |k| <= kcut instead,
    n_d up to floor(kcut L_d / 2 pi) per axis."""
    box = np.asarray(box, float)
    if kcut is None:
        n = np.arange(-kmax, kmax + 1)
        N = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3)
        N = N[np.sum(N ** 2, axis=1) <= kmax ** 2]
    else:
        nm = np.floor(kcut * box / (2 * np.pi)).astype(int)
        axes = [np.arange(-m, m + 1) for m in nm]
        N = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
        kk = 2 * np.pi * N / box
        N = N[np.sum(kk ** 2, axis=1) <= kcut ** 2]
    N = N[np.any(N != 0, axis=1)]
    kv = 2 * np.pi * N / box
    k2 = np.sum(kv ** 2, axis=1)
    return kv, np.exp(-k2 / (4 * alpha ** 2)) / k2


class Ewald(Force):
    """forces.py:333-388: real-space erfc sum + reciprocal structure factor + self term.

    `chunk` evaluates the same formulas in row / k-vector blocks (memory O(N*chunk));
    `kcut` selects k-vectors by |k| <= kcut (extension, see kvectors)."""

    def __init__(self, k=K_E, alpha=None, rcut=None, kmax=12, kcut=None, chunk=None):
        self.k = float(k)
        self.alpha = alpha
        self.rcut = rcut
        self.kmax = kmax
        self.kcut = kcut
        self.chunk = chunk

    def params(self, s):
        rcut = self.rcut or 0.5 * float(np.min(s.box))
        a = self.alpha or 4.0 / rcut
        return rcut, a

    def compute(self, s):
        if s.dim != 3 or not s.periodic:
            raise ValueError("Ewald needs a 3D periodic box")
        q = s.charges
        f = np.zeros_like(s.pos)
        if not np.any(q):
            return f, 0.0
        if abs(q.sum()) > 1e-9 * np.sum(np.abs(q)):
            raise ValueError("Ewald summation here requires a charge-neutral system")
        box = s.box
        rcut, a = self.params(s)
        kv, w = kvectors(box, a, self.kmax, self.kcut)
        V = float(np.prod(box))
        if self.chunk is None:
            d, r2 = _pairs(s)
            np.fill_diagonal(r2, np.inf)
            r = np.sqrt(r2)
            qq = q[:, None] * q[None, :]
            inside = r < rcut
            erfc_r = np.where(inside, _erfc(a * np.where(inside, r, 0)), 0.0)
            U_real = 0.5 * float(np.sum(np.where(inside, qq * erfc_r / r, 0)))
            coef = np.where(inside, qq * (erfc_r / r + 2 * a / np.sqrt(np.pi) * np.exp(-(a * r) ** 2)) / r2, 0)
            f -= np.einsum("ij,ijk->ik", coef, d)
            phase = s.pos @ kv.T
            S = (q[:, None] * np.exp(1j * phase)).sum(0)
            U_rec = 2 * np.pi / V * float(np.sum(w * np.abs(S) ** 2))
            im = np.imag(np.conj(S)[None, :] * np.exp(1j * phase))
            f += 4 * np.pi / V * q[:, None] * ((w[None, :] * im) @ kv)
        else:
            U_real = 0.0
            for a0 in range(0, s.n, self.chunk):
                rows = np.arange(a0, min(a0 + self.chunk, s.n))
                d = s.displacements(rows)
                r2 = np.einsum("ijk,ijk->ij", d, d)
                r2[np.arange(len(rows)), rows] = np.inf
                r = np.sqrt(r2)
                qq = q[rows, None] * q[None, :]
                inside = r < rcut
                erfc_r = np.where(inside, _erfc(a * np.where(inside, r, 0)), 0.0)
                U_real += 0.5 * float(np.sum(np.where(inside, qq * erfc_r / r, 0)))
                coef = np.where(inside, qq * (erfc_r / r + 2 * a / np.sqrt(np.pi) * np.exp(-(a * r) ** 2)) / r2, 0)
                f[rows] -= np.einsum("ij,ijk->ik", coef, d)
            S = np.zeros(len(kv), complex)
            for k0 in range(0, len(kv), self.chunk):
                ks = slice(k0, k0 + self.chunk)
                S[ks] = (q[:, None] * np.exp(1j * (s.pos @ kv[ks].T))).sum(0)
            U_rec = 2 * np.pi / V * float(np.sum(w * np.abs(S) ** 2))
            for k0 in range(0, len(kv), self.chunk):
                ks = slice(k0, k0 + self.chunk)
                im = np.imag(np.conj(S[ks])[None, :] * np.exp(1j * (s.pos @ kv[ks].T)))
                f += 4 * np.pi / V * q[:, None] * ((w[None, ks] * im) @ kv[ks])
        U_self = -a / np.sqrt(np.pi) * float(np.sum(q ** 2))
        return self.k * f, self.k * (U_real + U_rec + U_self)


class _Node:
    """tree.py:15-16."""
    __slots__ = ("center", "children", "com", "half", "idx", "mass")


class BarnesHut(Force):
    """tree.py:19-78: monopole octree, accept if outside the node box and (2 half)^2 < theta^2 r^2."""

    exact = False

    def __init__(self, G=G_SI, softening=0.0, theta=0.5, leaf_size=8):
        self.G = float(G)
        self.eps2 = float(softening) ** 2
        self.theta = theta
        self.leaf_size = leaf_size

    def _build(self, pos, m, idx, center, half, depth=0):
        # tree.py:30-43
        node = _Node()
        node.idx, node.center, node.half = idx, center, half
        node.mass = float(m[idx].sum())
        node.com = (m[idx, None] * pos[idx]).sum(0) / node.mass if node.mass > 0 else center
        node.children = []
        if len(idx) > self.leaf_size and depth < 48:
            octant = (pos[idx] > center).astype(int) @ (1 << np.arange(pos.shape[1]))
            for o in range(2 ** pos.shape[1]):
                sub = idx[octant == o]
                if len(sub):
                    sign = ((o >> np.arange(pos.shape[1])) & 1) * 2 - 1
                    node.children.append(self._build(pos, m, sub, center + sign * half / 2, half / 2, depth + 1))
        return node

    def compute(self, s):
        # tree.py:45-78
        pos, m = s.pos, s.masses
        acc = np.zeros_like(pos)
        phi = np.zeros(s.n)
        if s.n < 2:
            return acc, 0.0
        lo, hi = pos.min(0), pos.max(0)
        half = max(float(np.max(hi - lo)) / 2 * (1 + 1e-9), 1e-300)
        root = self._build(pos, m, np.arange(s.n), (lo + hi) / 2, half)
        stack = [(root, np.arange(s.n))]
        th2 = self.theta ** 2
        while stack:
            node, tgt = stack.pop()
            if not node.children:
                src = node.idx
                d = pos[src][None, :, :] - pos[tgt][:, None, :]
                r2 = np.einsum("ijk,ijk->ij", d, d) + self.eps2
                same = tgt[:, None] == src[None, :]
                inv = np.where(same, 0.0, 1 / np.sqrt(np.where(same, 1.0, r2)))
                acc[tgt] += self.G * np.einsum("ij,ijk->ik", m[src][None, :] * inv ** 3, d)
                phi[tgt] -= self.G * (inv * m[src][None, :]).sum(1)
                continue
            d = node.com - pos[tgt]
            r2 = np.einsum("ij,ij->i", d, d) + self.eps2
            outside = np.any(np.abs(pos[tgt] - node.center) > node.half, axis=1)
            accept = outside & ((2 * node.half) ** 2 < th2 * r2)
            if np.any(accept):
                a, ra = tgt[accept], r2[accept]
                acc[a] += self.G * node.mass * d[accept] / ra[:, None] ** 1.5
                phi[a] -= self.G * node.mass / np.sqrt(ra)
            rest = tgt[~accept]
            if len(rest):
                stack.extend((c, rest) for c in node.children)
        return m[:, None] * acc, 0.5 * float(np.sum(m * phi))


def bh_tree_preorder(pos, m, leaf_size=8):
    """The octree of tree.py:51-53 (root cube) and 30-43 (_build), flattened depth-first in
    pre-order with children in octant order 0..7, as flat arrays:
      center [nn, 3], half [nn], depth [nn], leaf [nn] (bool), count [nn], mass [nn],
      com [nn, 3], next [nn] (index of the first node after the subtree, nn at the end),
      members: list of sorted particle-index arrays, one per node.
    Pinned by BarnesHut itself: the walk over this tree is the one pinned by theta = 0 and
    the golden theta = 0.5 / 0.3 vectors."""
    pos = np.asarray(pos, float)
    m = np.asarray(m, float)
    lo, hi = pos.min(0), pos.max(0)
    half = max(float(np.max(hi - lo)) / 2 * (1 + 1e-9), 1e-300)
    root = BarnesHut(leaf_size=leaf_size)._build(pos, m, np.arange(len(pos)), (lo + hi) / 2, half)
    rows = []

    def visit(node, depth):
        k = len(rows)
        rows.append([node, depth, None])
        for c in node.children:
            visit(c, depth + 1)
        rows[k][2] = len(rows)

    visit(root, 0)
    return {
        "center": np.array([r[0].center for r in rows], float),
        "half": np.array([r[0].half for r in rows], float),
        "depth": np.array([r[1] for r in rows], np.int64),
        "leaf": np.array([not r[0].children for r in rows], bool),
        "count": np.array([len(r[0].idx) for r in rows], np.int64),
        "mass": np.array([r[0].mass for r in rows], float),
        "com": np.array([r[0].com for r in rows], float),
        "next": np.array([r[2] for r in rows], np.int64),
        "members": [np.sort(r[0].idx) for r in rows],
    }


def direct_subset(force, s, sub):
    """tree.py:81-87: exact gravitational force on particles `sub`."""
    d = s.pos[None, :, :] - s.pos[sub][:, None, :]
    r2 = np.einsum("ijk,ijk->ij", d, d) + force.eps2
    same = sub[:, None] == np.arange(s.n)[None, :]
    inv = np.where(same, 0.0, 1 / np.sqrt(np.where(same, 1.0, r2)))
    return force.G * s.masses[sub, None] * np.einsum("ij,ijk->ik", s.masses[None, :] * inv ** 3, d)


class VelocityVerlet:
    """integrators.py:19-31."""
    order = 2
    symplectic = True
    stochastic = False
    thermostat = False

    def step(self, sim, dt):
        s = sim.system
        s.vel += 0.5 * dt * sim.acc
        s.pos += dt * s.vel
        s.apply_boundaries()
        sim.update_forces()
        s.vel += 0.5 * dt * sim.acc


class Rattle:
    """integrators.py:119-234 (the "matrix" method, used by the reference for up to 1500
    constraints): velocity Verlet with rigid bonds [(id_a, id_b, length)].

    Positions (M-SHAKE, integrators.py:197-226): after the half kick and drift, Newton
    iterations on all constraints together, K g = L^2 - |r|^2 with
    K[k, l] = 2 (r_k . r0_l) C[k, l] and C[k, l] the change of bond vector k per unit
    multiplier on bond l (inverse masses, signs from shared atoms); each iteration moves
    positions and velocities along the old bond vectors r0. Velocities (RATTLE,
    integrators.py:227-234): after the forces and second half kick, one solve
    (C[k, l] r_k . r_l) mu = r_k . (v_a - v_b) removes the bond-stretching velocities.
    Iterations are float64; the worst relative bond-length error is kept in
    max_violation."""

    order = 2
    symplectic = True
    stochastic = False
    thermostat = False

    def __init__(self, constraints, tol=1e-12, max_iter=1000):
        self.constraints = [(int(a), int(b), float(d)) for a, b, d in constraints]
        self.tol = tol
        self.max_iter = max_iter
        self.max_violation = 0.0

    @staticmethod
    def _mi(s, d):
        return d - s.box * np.round(d / s.box) if s.periodic else d

    def step(self, sim, dt):
        s = sim.system
        slot = {int(pid): k for k, pid in enumerate(s.ids)}
        cons = [(slot[a], slot[b], d) for a, b, d in self.constraints if a in slot and b in slot]
        s.constrained_dof = len(cons)
        inv = 1.0 / s.masses
        if not cons:
            VelocityVerlet().step(sim, dt)
            return
        ia = np.array([c[0] for c in cons])
        ib = np.array([c[1] for c in cons])
        L2 = np.array([c[2] for c in cons]) ** 2
        same = lambda u, w: (u[:, None] == w[None, :]).astype(float)
        C = inv[ia][:, None] * (same(ia, ia) - same(ia, ib)) - inv[ib][:, None] * (same(ib, ia) - same(ib, ib))
        r0 = self._mi(s, s.pos[ia] - s.pos[ib])
        s.vel += 0.5 * dt * sim.acc
        new = s.pos + dt * s.vel
        best, stall = np.inf, 0
        for _ in range(self.max_iter):
            r = self._mi(s, new[ia] - new[ib])
            gap = L2 - np.einsum("ij,ij->i", r, r)
            res = float(np.max(np.abs(gap) / L2))
            if res <= self.tol:
                break
            if res < 0.5 * best:
                best, stall = res, 0
            else:
                stall += 1
            if stall >= 10:
                raise RuntimeError(f"SHAKE stalled at relative residual {res:.1e}")
            g = np.linalg.solve(2.0 * C * (r @ r0.T), gap)
            for idx, sign in ((ia, 1.0), (ib, -1.0)):
                step = sign * (g * inv[idx])[:, None] * r0
                np.add.at(new, idx, step)
                np.add.at(s.vel, idx, step / dt)
        else:
            raise RuntimeError("SHAKE did not converge")
        s.pos = new
        s.apply_boundaries()
        sim.update_forces()
        s.vel += 0.5 * dt * sim.acc
        r = self._mi(s, s.pos[ia] - s.pos[ib])
        rv = np.einsum("ij,ij->i", r, s.vel[ia] - s.vel[ib])
        mu = np.linalg.solve(C * (r @ r.T), rv)
        for idx, sign in ((ia, -1.0), (ib, 1.0)):
            np.add.at(s.vel, idx, sign * (mu * inv[idx])[:, None] * r)
        r = self._mi(s, s.pos[ia] - s.pos[ib])
        viol = np.abs(np.sqrt(np.einsum("ij,ij->i", r, r)) - np.sqrt(L2)) / np.sqrt(L2)
        self.max_violation = max(self.max_violation, float(np.max(viol)))


class Yoshida4:
    """integrators.py:34-47."""
    order = 4
    symplectic = True
    stochastic = False
    thermostat = False
    _w1 = 1 / (2 - 2 ** (1 / 3))
    _w0 = -(2 ** (1 / 3)) * _w1

    def __init__(self):
        self._vv = VelocityVerlet()

    def step(self, sim, dt):
        for w in (self._w1, self._w0, self._w1):
            self._vv.step(sim, w * dt)


class Langevin:
    """integrators.py:50-71: BAOAB Langevin dynamics (kick, half drift, Ornstein-Uhlenbeck
    velocity update, half drift, forces, kick)."""
    order = 2
    symplectic = False
    stochastic = True
    thermostat = True

    def __init__(self, temperature, friction):
        self.T = float(temperature)
        self.gamma = float(friction)

    def step(self, sim, dt):
        s = sim.system
        s.vel += 0.5 * dt * sim.acc
        s.pos += 0.5 * dt * s.vel
        c1 = np.exp(-self.gamma * dt)
        sig = np.sqrt((1 - c1 ** 2) * K_B * self.T / s.masses)[:, None]
        s.vel = c1 * s.vel + sig * sim.rng.normal(size=s.vel.shape)
        s.pos += 0.5 * dt * s.vel
        s.apply_boundaries()
        sim.update_forces()
        s.vel += 0.5 * dt * sim.acc


class NoseHoover:
    """integrators.py:74-116: Nose-Hoover chain of length one in the Trotter splitting of
    Martyna, Tuckerman and Klein (1996); thermostat mass Q = N_f kT tau^2 at the first
    half step, N_f = dim (n - 1); conserved extended energy E + Q xi^2 / 2 + N_f kT eta."""
    order = 2
    symplectic = False
    stochastic = False
    thermostat = True

    def __init__(self, temperature, period):
        self.T = float(temperature)
        self.tau = float(period)
        self.xi = 0.0
        self.eta = 0.0
        self.Q = None

    @staticmethod
    def dof(s):
        return s.dim * max(s.n - 1, 1)

    def _half(self, s, dt):
        kT, nf = K_B * self.T, self.dof(s)
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
        return 0.5 * self.Q * self.xi ** 2 + self.dof(s) * K_B * self.T * self.eta


class Brownian:
    """integrators.py:237-257: overdamped Langevin (Euler-Maruyama) with each species'
    diffusion coefficient D: dx = D F / (kT) dt + sqrt(2 D dt) xi, v = dx / dt."""
    order = 1
    symplectic = False
    stochastic = True
    thermostat = True

    def __init__(self, temperature):
        self.T = float(temperature)

    def step(self, sim, dt):
        s = sim.system
        D = np.array([sp.diffusion for sp in s.species])[s.sp][:, None]
        drift = D * sim.force / (K_B * self.T) * dt
        noise = np.sqrt(2 * D * dt) * sim.rng.normal(size=s.pos.shape)
        s.pos += drift + noise
        s.vel = (drift + noise) / dt
        s.apply_boundaries()
        sim.update_forces()


# ---------------------------------------------------------------- counter-based normals
# The device engines draw their Gaussian noise from Philox4x32-10 (Salmon, Moraes, Dror
# and Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC11) instead of numpy's
# PCG64 stream. This is a restatement of that generator and of the engines' mapping to
# normals, so that the reference's stochastic integrators above can be driven with the
# same noise (as `sim.rng`) and compared with the device step for step. Pinned by the
# known-answer vectors of the Random123 distribution (tests/test_oracle_thermostats.py).
_PH_M0, _PH_M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_PH_W0, _PH_W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)


def philox4x32_10(ctr, key):
    """ctr: uint32 [..., 4], key: uint32 [..., 2] -> uint32 [..., 4]."""
    c = [np.asarray(ctr[..., i], np.uint32).copy() for i in range(4)]
    k0 = np.asarray(key[..., 0], np.uint32).copy()
    k1 = np.asarray(key[..., 1], np.uint32).copy()
    mask = np.uint64(0xFFFFFFFF)
    for r in range(10):
        if r:
            k0 = (k0 + _PH_W0).astype(np.uint32)
            k1 = (k1 + _PH_W1).astype(np.uint32)
        p0 = _PH_M0 * c[0].astype(np.uint64)
        p1 = _PH_M1 * c[2].astype(np.uint64)
        hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & mask).astype(np.uint32)
        hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & mask).astype(np.uint32)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack(c, axis=-1)


def philox_normals(seed, counter, ids):
    """Three standard normals per id for drift number `counter` (the engines' noise):
    Philox4x32-10 with key (seed low, seed high) at counters (counter low, counter high,
    id, j) for j = 0, 1 gives 8 words; words (2i, 2i+1) make the uniform
    ((w0 >> 5) 2^26 + (w1 >> 6) + 0.5) 2^-53 in (0, 1); uniforms (u0, u1) and (u2, u3)
    go through Box-Muller, r = sqrt(-2 ln u0), n = (r cos 2 pi u1, r sin 2 pi u1, ...)."""
    ids = np.asarray(ids, np.uint64)
    n = len(ids)
    key = np.zeros((n, 2), np.uint32)
    key[:, 0] = np.uint32(seed & 0xFFFFFFFF)
    key[:, 1] = np.uint32((seed >> 32) & 0xFFFFFFFF)
    words = []
    for j in range(2):
        ctr = np.zeros((n, 4), np.uint32)
        ctr[:, 0] = np.uint32(counter & 0xFFFFFFFF)
        ctr[:, 1] = np.uint32((counter >> 32) & 0xFFFFFFFF)
        ctr[:, 2] = ids.astype(np.uint32)
        ctr[:, 3] = np.uint32(j)
        words.append(philox4x32_10(ctr, key))
    w = np.concatenate(words, axis=1).astype(np.float64)  # [n, 8]
    u = ((np.floor(w[:, 0::2] / 32.0)) * 67108864.0 + np.floor(w[:, 1::2] / 64.0) + 0.5) * 2.0 ** -53
    r0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    r1 = np.sqrt(-2.0 * np.log(u[:, 2]))
    t0 = 2.0 * np.pi * u[:, 1]
    t1 = 2.0 * np.pi * u[:, 3]
    return np.stack([r0 * np.cos(t0), r0 * np.sin(t0), r1 * np.cos(t1)], axis=1)


class PhiloxNormals:
    """Drop-in for `sim.rng` in the stochastic integrators: each normal() call is one
    drift, numbered from 0, over the atoms of the system in order (ids 0..n-1, offset by
    `id0` for a system placed after others in a device arena)."""

    def __init__(self, seed, id0=0):
        self.seed = int(seed)
        self.id0 = int(id0)
        self.counter = 0

    def normal(self, size):
        n = size[0]
        out = philox_normals(self.seed, self.counter, np.arange(self.id0, self.id0 + n))
        self.counter += 1
        return out[:, :size[1]] if len(size) > 1 else out[:, 0]


@dataclass
class Trajectory:
    """simulation.py:14-38 (fields on the force path)."""
    t: list = field(default_factory=list)
    kinetic: list = field(default_factory=list)
    potential: list = field(default_factory=list)
    momentum: list = field(default_factory=list)
    angular_momentum: list = field(default_factory=list)
    mass: list = field(default_factory=list)
    charge: list = field(default_factory=list)
    bath: list = field(default_factory=list)
    positions: list = field(default_factory=list)
    velocities: list = field(default_factory=list)

    @property
    def total_energy(self):
        # simulation.py:31-35 with no reactions: internal = released = 0
        return np.array(self.kinetic) + np.array(self.potential)


class Simulation:
    """simulation.py:41-127 without reactions."""

    def __init__(self, system, forces=(), integrator=None, dt=None, seed=None):
        self.system = system
        self.forces = list(forces)
        self.integrator = integrator or VelocityVerlet()
        self.dt = dt
        self.rng = np.random.default_rng(seed)
        self.seed = seed
        self.time = 0.0
        self.steps = 0
        self.trajectory = Trajectory()
        self.update_forces()
        self._initial = None

    def potential(self):
        # simulation.py:57-65
        s = self.system
        f = np.zeros_like(s.pos)
        u = 0.0
        for force in self.forces:
            fi, ui = force.compute(s)
            f += fi
            u += ui
        return f, u

    def update_forces(self):
        # simulation.py:67-70
        self.force, self.pe = self.potential()
        m = self.system.masses
        self.acc = self.force / m[:, None] if len(m) else self.force

    def record(self, positions=True):
        # simulation.py:76-93
        s, tr = self.system, self.trajectory
        tr.t.append(self.time)
        tr.kinetic.append(s.kinetic_energy())
        tr.potential.append(self.pe)
        tr.momentum.append(s.momentum())
        tr.angular_momentum.append(s.angular_momentum() if s.dim in (2, 3) else np.zeros(1))
        tr.mass.append(float(np.sum(s.masses)))
        tr.charge.append(float(np.sum(s.charges)))
        bath = getattr(self.integrator, "bath_energy", None)
        tr.bath.append(bath(s) if bath else 0.0)
        if positions:
            tr.positions.append(s.pos.copy())
            tr.velocities.append(s.vel.copy())

    def run(self, steps, record_every=1, positions=True):
        # simulation.py:95-112
        if self.dt is None:
            raise ValueError("set dt")
        if not self.trajectory.t:
            if self._initial is None:
                self._initial = self.clone()
            self.record(positions)
        for k in range(1, steps + 1):
            self.integrator.step(self, self.dt)
            self.time += self.dt
            self.steps += 1
            if k % record_every == 0 or k == steps:
                self.record(positions)
        return self.trajectory

    def clone(self, dt=None):
        # simulation.py:114-127
        import copy
        sim = Simulation.__new__(Simulation)
        sim.__dict__.update(self.__dict__)
        sim.system = self.system.copy()
        sim.forces = list(self.forces)
        sim.integrator = copy.deepcopy(self.integrator)
        sim.trajectory = Trajectory()
        sim.dt = self.dt if dt is None else dt
        sim.update_forces()
        return sim


def _grade(x, good, ok):
    """accuracy.py:30-31."""
    return "PASS" if x <= good else "WARN" if x <= ok else "FAIL"


def drift_checks(sim):
    """The drift rows of AccuracyReport.build, accuracy.py:71-110: {name: (status, value)}."""
    tr, s = sim.trajectory, sim.system
    out = {}
    conservative = all(f.conservative for f in sim.forces) and not sim.integrator.thermostat
    external = any(f.external for f in sim.forces)
    central = all(f.central for f in sim.forces)
    E = tr.total_energy
    scale = max(abs(E[0]), np.max(np.abs(tr.kinetic)), np.max(np.abs(tr.potential)), 1e-300)
    if conservative:
        drift = float(np.max(np.abs(E - E[0])) / scale)
        out["energy"] = (_grade(drift, 1e-5, 1e-3), drift)
    P = np.array(tr.momentum)
    pscale = max(float(np.sum(s.masses[:, None] * np.abs(s.vel))), 1e-300)
    if not external:
        dp = float(np.max(np.linalg.norm(P - P[0], axis=1)) / pscale)
        out["momentum"] = (_grade(dp, 1e-9, 1e-6), dp)
    if central and not external and s.box is None and s.dim in (2, 3):
        L = np.array(tr.angular_momentum)
        lscale = max(float(np.sum(np.linalg.norm(np.cross(s.pos, s.masses[:, None] * s.vel), axis=-1))), 1e-300)
        dl = float(np.max(np.linalg.norm(L - L[0], axis=1)) / lscale)
        out["angular momentum"] = (_grade(dl, 1e-9, 1e-6), dl)
    M = np.array(tr.mass)
    dm = float(np.max(np.abs(M - M[0])) / M[0]) if M[0] else 0.0
    out["mass-energy"] = (_grade(dm, 1e-12, 1e-9), dm)
    Qc = np.array(tr.charge)
    qscale = max(float(np.sum(np.abs(s.charges))), 1e-300)
    dq = float(np.max(np.abs(Qc - Qc[0])) / qscale)
    out["charge"] = (_grade(dq, 1e-12, 1e-9), dq)
    return out


def rms_rel_force_error(F_ap, F_ex):
    """accuracy.py:141-143 / 158-159: sqrt(mean |dF|^2 / mean |F_ex|^2)."""
    return float(np.sqrt(np.mean(np.sum((F_ap - F_ex) ** 2, 1)) / np.mean(np.sum(F_ex ** 2, 1))))


def force_approx_check(force, s):
    """accuracy.py:135-146: sampled RMS relative error of an approximate force vs direct_subset."""
    sub = np.random.default_rng(0).choice(s.n, size=min(s.n, 200), replace=False)
    F_ap = force.compute(s)[0][sub]
    F_ex = direct_subset(force, s, sub)
    e = rms_rel_force_error(F_ap, F_ex)
    return _grade(e, 1e-3, 1e-2), e



def timescales(sim, max_rows=1500):
    """advisor.py:9-48 for the forces of this path (gravity, Barnes-Hut, LJ; the plain
    Coulomb and bond rows need forces that are not on it): shortest times in seconds."""
    s = sim.system
    out = {}
    if s.n >= 2:
        sub = np.arange(s.n) if s.n <= max_rows else np.random.default_rng(0).choice(s.n, max_rows, replace=False)
        d = s.pos[None, :, :] - s.pos[sub][:, None, :]
        if s.periodic:
            d -= s.box * np.round(d / s.box)
        r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))
        r[np.arange(len(sub)), sub] = np.inf
        m_all = s.masses
        m = m_all[sub][:, None] * np.ones((1, s.n))
        m_col = np.ones((len(sub), 1)) * m_all[None, :]
        for f in sim.forces:
            kind = type(f).__name__
            if kind in ("Gravity", "BarnesHut"):
                M = m + m_col
                out["gravity (orbital)"] = float(np.min(np.sqrt((r ** 2 + f.eps2) ** 1.5 / (f.G * M))))
            elif kind == "LennardJones":
                eps_t, sig_t = f._tables(s)
                for i, sp in enumerate(s.species):
                    if eps_t[i, i] > 0:
                        t = sig_t[i, i] * np.sqrt(sp.mass / eps_t[i, i])
                        out["Lennard-Jones"] = min(out.get("Lennard-Jones", np.inf), float(t))
        speed = np.max(np.linalg.norm(s.vel, axis=1))
        if speed > 0 and any(not f.external for f in sim.forces):
            eps2 = max([getattr(f, "eps2", 0.0) for f in sim.forces] or [0.0])
            out["closest approach / speed"] = float(np.sqrt(np.min(r) ** 2 + eps2) / speed)
    return out


def report_rows(sim, dsf_ewald_max_n=2000):
    """AccuracyReport.build without the timestep error, accuracy.py:59-189: the rows in the
    reference's order as {name: (status, value or None)}. Reactions are absent, so the
    reaction-timestep row never appears; `dsf_ewald_max_n` is the reference's 2000 unless
    raised. The DSF row judges with this oracle's Ewald(k=f.k)."""
    tr, s, integ = sim.trajectory, sim.system, sim.integrator
    forces = sim.forces
    rows = {"units": ("PASS", None)}
    conservative = all(f.conservative for f in forces) and not integ.thermostat
    external = any(f.external for f in forces)
    walls = s.box is not None and not s.periodic
    central = all(f.central for f in forces)
    E = tr.total_energy
    scale = max(abs(E[0]), np.max(np.abs(tr.kinetic)), np.max(np.abs(tr.potential)), 1e-300)
    if conservative:
        drift = float(np.max(np.abs(E - E[0])) / scale)
        rows["energy"] = (_grade(drift, 1e-5, 1e-3), drift)
    else:
        rows["energy"] = ("n/a", None)
    if hasattr(integ, "bath_energy") and all(f.conservative for f in forces) and len(tr.bath) == len(tr.t):
        Ex = E + np.array(tr.bath)
        dx = float(np.max(np.abs(Ex - Ex[0])) / scale)
        rows["extended energy"] = (_grade(dx, 1e-5, 1e-3), dx)
    P = np.array(tr.momentum)
    pscale = max(float(np.sum(s.masses[:, None] * np.abs(s.vel))), 1e-300)
    if not external and not walls and not integ.thermostat:
        dp = float(np.max(np.linalg.norm(P - P[0], axis=1)) / pscale)
        rows["momentum"] = (_grade(dp, 1e-9, 1e-6), dp)
    else:
        rows["momentum"] = ("n/a", None)
    if central and not external and s.box is None and not integ.thermostat and s.dim in (2, 3):
        L = np.array(tr.angular_momentum)
        lscale = max(float(np.sum(np.linalg.norm(np.cross(s.pos, s.masses[:, None] * s.vel), axis=-1))), 1e-300)
        dl = float(np.max(np.linalg.norm(L - L[0], axis=1)) / lscale)
        rows["angular momentum"] = (_grade(dl, 1e-9, 1e-6), dl)
    else:
        rows["angular momentum"] = ("n/a", None)
    M = np.array(tr.mass)
    dm = float(np.max(np.abs(M - M[0])) / M[0]) if M[0] else 0.0
    rows["mass-energy"] = (_grade(dm, 1e-12, 1e-9), dm)
    Qc = np.array(tr.charge)
    qscale = max(float(np.sum(np.abs(s.charges))), 1e-300)
    dq = float(np.max(np.abs(Qc - Qc[0])) / qscale)
    rows["charge"] = (_grade(dq, 1e-12, 1e-9), dq)
    kind = type(integ).__name__
    if kind in ("Langevin", "NoseHoover") and len(tr.t) >= 4:
        half = len(tr.t) // 2
        kin = np.asarray(tr.kinetic[half:])
        counts = getattr(tr, "counts", None)
        n = (np.array([sum(c.values()) for c in counts[half:]], dtype=float) if counts
             else np.full(len(kin), float(s.n)))
        dof = s.dim * np.maximum(n - 1, 1) if kind == "NoseHoover" else s.dim * np.maximum(n, 1)
        T_meas = float(np.mean(2 * kin / (dof * K_B)))
        dev = abs(T_meas - integ.T) / integ.T
        tol = max(0.02, 3 * np.sqrt(2 / (s.dim * max(s.n, 1))))
        rows["thermostat"] = ("PASS" if dev <= tol else ("WARN" if dev <= 0.1 else "FAIL"), dev)
    if kind != "Brownian":
        ts = timescales(sim)
        if ts:
            _, tmin = min(ts.items(), key=lambda kv: kv[1])
            ratio = sim.dt / tmin
            rows["timestep vs physics"] = (_grade(ratio, 0.02, 0.1), ratio)
    approx = [f for f in forces if not getattr(f, "exact", True)]
    if approx and s.n >= 2:
        sub = np.random.default_rng(0).choice(s.n, size=min(s.n, 200), replace=False)
        e = max(rms_rel_force_error(f.compute(s)[0][sub], direct_subset(f, s, sub)) for f in approx)
        rows["force approximation"] = (_grade(e, 1e-3, 1e-2), e)
    for f in forces:
        if type(f).__name__ != "DampedShiftedCoulomb":
            continue
        q = s.charges
        if not (s.periodic and s.dim == 3 and np.any(q)):
            continue
        if f.exclude or abs(q.sum()) > 1e-9 * np.sum(np.abs(q)) or s.n > dsf_ewald_max_n:
            rows["DSF vs Ewald"] = ("n/a", None)
            continue
        e = rms_rel_force_error(f.compute(s)[0], Ewald(k=f.k).compute(s)[0])
        rows["DSF vs Ewald"] = (_grade(e, 5e-2, 2e-1), e)
    if s.periodic and any(type(f).__name__ == "Coulomb" for f in forces):
        rows["electrostatics"] = ("WARN", None)
    if hasattr(integ, "max_violation"):
        v = integ.max_violation
        rows["constraints"] = (_grade(v, 1e-9, 1e-6), v)
    beta2 = float(np.max(np.sum(s.vel ** 2, axis=1)) / C_LIGHT ** 2) if s.n else 0.0
    rows["relativity"] = (_grade(beta2, 1e-4, 1e-2), beta2)
    return rows

# ---------------------------------------------------------------- analysis.py:7-94

def rdf(system, rmax, bins=100, species=None):
    """analysis.py:7-24."""
    r, g, _ = rdf_counts(system, rmax, bins, species)
    return r, g


def rdf_counts(system, rmax, bins=100, species=None):
    """analysis.py:7-24, also returning the pair histogram."""
    s = system
    idx = np.arange(s.n) if species is None else np.flatnonzero(s.sp == s.species_index(species))
    if not s.periodic:
        raise ValueError("rdf needs a periodic box (for normalisation)")
    if rmax > 0.5 * np.min(np.broadcast_to(s.box, (s.dim,))):
        raise ValueError("rmax must be at most half the box")
    p = s.pos[idx]
    d = p[None, :, :] - p[:, None, :]
    d -= s.box * np.round(d / s.box)
    r = np.sqrt(np.einsum("ijk,ijk->ij", d, d))[np.triu_indices(len(idx), 1)]
    hist, edges = np.histogram(r, bins=bins, range=(0, rmax))
    V = float(np.prod(np.broadcast_to(s.box, (s.dim,))))
    n = len(idx)
    shell = (4 / 3 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)) if s.dim == 3 else (np.pi * (edges[1:] ** 2 - edges[:-1] ** 2))
    ideal = shell * n * (n - 1) / 2 / V
    return 0.5 * (edges[1:] + edges[:-1]), hist / ideal, hist


def msd(trajectory, box=None):
    """analysis.py:27-43."""
    t = np.asarray(trajectory.t)
    frames = trajectory.positions
    disp = np.zeros_like(frames[0])
    out = [0.0]
    for a, b in zip(frames[:-1], frames[1:]):
        step = b - a
        if box is not None:
            step -= box * np.round(step / box)
        disp = disp + step
        out.append(float(np.mean(np.sum(disp ** 2, axis=1))))
    return t - t[0], np.array(out)


def pressure(sim, h=1e-6):
    """analysis.py:46-64."""
    s = sim.system
    if s.box is None:
        raise ValueError("pressure needs a box")
    box0, pos0 = np.array(s.box, dtype=float), s.pos.copy()
    V0 = float(np.prod(np.broadcast_to(box0, (s.dim,))))
    U = []
    for sign in (+1, -1):
        lam = (1 + sign * h) ** (1 / s.dim)
        s.box, s.pos = box0 * lam, pos0 * lam
        U.append(sum(f.compute(s)[1] for f in sim.forces if not f.external))
    s.box, s.pos = (box0 if np.ndim(box0) else float(box0)), pos0
    dUdV = (U[0] - U[1]) / (2 * h * V0)
    return 2 * s.kinetic_energy() / (s.dim * V0) - dUdV


def vacf(trajectory, max_lag=None):
    """analysis.py:77-87."""
    V = np.asarray(trajectory.velocities)
    if V.ndim != 3:
        raise ValueError("vacf needs recorded velocities of a constant particle set")
    n = len(V)
    max_lag = n // 2 if max_lag is None else max_lag
    C = np.array([np.mean(np.einsum("tij,tij->ti", V[:n - k], V[k:])) for k in range(max_lag)])
    t = np.asarray(trajectory.t)
    return t[:max_lag] - t[0], C


def green_kubo_diffusion(trajectory, max_lag=None):
    """analysis.py:90-94."""
    t, C = vacf(trajectory, max_lag)
    dim = np.asarray(trajectory.velocities).shape[2]
    return float(np.sum(0.5 * (C[1:] + C[:-1]) * np.diff(t))) / dim


def timestep_convergence(sim, duration, start=None):
    """accuracy.py:191-221: Richardson estimate from dt, dt/2, dt/4 re-runs. Returns (status, est, p)."""
    base = start or sim._initial
    n = max(int(round(duration / sim.dt)), 1)
    finals = []
    for k in (1, 2, 4):
        c = base.clone(dt=sim.dt / k)
        c.run(steps=n * k, record_every=10 ** 9, positions=False)
        finals.append(c.system.pos.copy())
    return richardson(finals, n, sim.integrator.order)


def richardson(finals, n, expected):
    """The arithmetic of accuracy.py:208-221 on the three final position arrays."""
    size = max(float(np.ptp(finals[2])), 1e-300)
    roundoff = 1e-13 * np.sqrt(n * 4)
    e1 = float(np.sqrt(np.mean((finals[0] - finals[1]) ** 2))) / size
    e2 = float(np.sqrt(np.mean((finals[1] - finals[2]) ** 2))) / size
    if e1 < 10 * roundoff:
        return "PASS", e1, None
    p = float(np.log2(e1 / e2))
    est = e1 * 2 ** p / (2 ** p - 1) if p > 0.2 else e1
    if p < 0.5 * expected:
        return ("FAIL" if est > 1e-2 else "WARN"), est, p
    return _grade(est, 1e-6, 1e-3), est, p
