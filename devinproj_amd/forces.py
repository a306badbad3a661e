"""GPU force fields with accusim's Force API (accusim/forces.py).

Each class keeps the reference's constructor, attributes and
`compute(system) -> (F[N, 3] float64 SI, U float)` contract; the arithmetic runs
in the HIP library (include/accusim_hip.h) on the current torch stream.

Extensions beyond the reference (all default to the reference's behaviour):
  precision="fp64" | "mixed"   mixed = float32 pair arithmetic on the float64
                               minimum-image displacement, float64 sums
  Ewald(kcut=...)              k-vectors |k| <= kcut (ellipsoid) instead of |n| <= kmax
  Ewald(pairlist_threshold=...) real-space sum over a cell list at rcut for large N
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib as L
from .units import G as G_CONST
from .units import k_e as KE_CONST
from .units import si

try:  # forces.py:324-330 takes scipy's erfc when present
    from scipy.special import erfc as _erfc_impl
except ImportError:  # pragma: no cover
    _erfc_impl = np.vectorize(math.erfc)


def _erfc(x):
    return _erfc_impl(x)


class Force:
    conservative = True
    external = False
    central = True

    def compute(self, system):
        raise NotImplementedError


def _precision(p):
    try:
        return L.PRECISIONS[p]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(L.PRECISIONS)}") from None


def cell_grid(s, rc):
    """lo, span, ncell of forces.py:281-289 (host numpy, bit-identical to the reference)."""
    pos = s.pos
    if s.box is not None:
        lo, span = np.zeros(s.dim), np.asarray(s.box, float)
    else:
        lo = pos.min(axis=0)
        span = np.maximum(pos.max(axis=0) - lo, rc)
    ncell = np.maximum((span // rc).astype(int), 1)
    if s.periodic:
        ncell = np.where(ncell < 3, 1, ncell)
    return np.ascontiguousarray(lo, float), np.ascontiguousarray(span, float), np.ascontiguousarray(ncell, np.int32)


def _box_args(s):
    if s.box is None:
        return None, 0
    return np.ascontiguousarray(s.box, float), int(bool(s.periodic))


class NeighbourList:
    """Device cell + Verlet list (ELL, both directions) for one set of positions."""

    def __init__(self):
        self._h = None
        self.n = 0

    @property
    def handle(self):
        if self._h is None:
            h = L.C.c_void_p()
            L.call("accu_nlist_create", L.C.byref(h))
            self._h = h
        return self._h

    MAX_CELLS = 1 << 22

    def build(self, s, pos_d, rc):
        if s.dim != 3:
            raise ValueError("the GPU neighbour list supports 3D systems")
        lo, span, ncell = cell_grid(s, rc)
        if int(np.prod(ncell.astype(np.int64))) > self.MAX_CELLS:
            # a dilute system: coarser cells (still >= rc wide) give the same pair set
            f = int(np.ceil((np.prod(ncell.astype(float)) / self.MAX_CELLS) ** (1 / 3)))
            ncell = np.maximum(ncell // f, 1).astype(np.int32)
            if s.periodic:
                ncell = np.where(ncell < 3, 1, ncell).astype(np.int32)
        box, periodic = _box_args(s)
        L.call("accu_nlist_build", self.handle, L.ptr(pos_d), s.n, L.ptr(lo), L.ptr(span), L.ptr(ncell),
               L.ptr(box), periodic, float(rc), L.stream())
        self.n = s.n

    def info(self):
        out = np.zeros(8, np.int64)
        L.call("accu_nlist_info", self.handle, L.ptr(out))
        return {"n": int(out[0]), "ncells": int(out[1]), "stride": int(out[2]), "capacity": int(out[3]),
                "max_neighbours": int(out[4]), "ncell": tuple(int(x) for x in out[5:8])}

    def cells(self):
        th = L.torch()
        cell = th.empty(self.n, dtype=th.int32, device="cuda")
        order = th.empty(self.n, dtype=th.int32, device="cuda")
        nn = th.empty(self.n, dtype=th.int32, device="cuda")
        L.call("accu_nlist_cells", self.handle, L.ptr(cell), L.ptr(order), L.ptr(nn), L.stream())
        return cell.cpu().numpy(), order.cpu().numpy(), nn.cpu().numpy()

    def half_pairs(self):
        th = L.torch()
        npairs = L.C.c_longlong(0)
        L.call("accu_nlist_count_half", self.handle, L.C.byref(npairs), L.stream())
        n = npairs.value
        i = th.empty(max(n, 1), dtype=th.int64, device="cuda")
        j = th.empty(max(n, 1), dtype=th.int64, device="cuda")
        if n:
            L.call("accu_nlist_export_half", self.handle, L.ptr(i), L.ptr(j), L.stream())
        return i[:n].cpu().numpy(), j[:n].cpu().numpy()

    def __del__(self):
        if self._h is not None and L._lib is not None:
            L._lib.accu_nlist_destroy(self._h)
            self._h = None


def cell_list_pairs(s, rc):
    """All pairs (i < j) closer than rc (forces.py:279-321), built on the GPU."""
    if s.n < 2:
        return np.zeros(0, int), np.zeros(0, int)
    bufs = L.Buffers()
    pos_d = bufs.upload("pos", np.asarray(s.pos, float), L.torch().float64)
    nl = NeighbourList()
    nl.build(s, pos_d, rc)
    return nl.half_pairs()


class VerletList:
    """forces.py:78-98 on the device: the list is reused until the largest minimum-image
    displacement since the build reaches skin/2, and rebuilt whenever ids, box or rc change."""

    def __init__(self):
        self._state = None
        self.rebuilds = 0
        self.nl = NeighbourList()
        self._ref = L.Buffers()

    def update(self, s, pos_d, rc, skin):
        """Make the device list valid for pos_d; returns True if it was rebuilt."""
        st = self._state
        if st is not None and st[1].shape == s.ids.shape and np.array_equal(st[1], s.ids) \
                and np.array_equal(st[2], s.box) and st[3] == rc and skin > 0:
            box, periodic = _box_args(s)
            m = L.C.c_double(0.0)
            L.call("accu_max_disp2", L.ptr(pos_d), L.ptr(st[0]), s.n, L.ptr(box), periodic, L.C.byref(m),
                   L.stream())
            if m.value < (0.5 * skin) ** 2:
                return False
        self.nl.build(s, pos_d, rc + skin)
        ref = self._ref.get("ref", tuple(pos_d.shape), pos_d.dtype)
        ref.copy_(pos_d)
        self._state = (ref, s.ids.copy(), None if s.box is None else np.array(s.box, float), rc)
        self.rebuilds += 1
        return True

    def invalidate(self):
        """The next update rebuilds the list whatever the displacements."""
        self._state = None

    def pairs(self, s, rc, skin):
        pos_d = self._ref.upload("pos", np.asarray(s.pos, float), L.torch().float64)
        self.update(s, pos_d, rc, skin)
        return self.nl.half_pairs()


class _DeviceForce(Force):
    """Shared upload / download plumbing."""

    def _setup(self, precision):
        self.precision = precision
        self._prec = _precision(precision)
        self._buf = L.Buffers()

    def _pos(self, s):
        if s.dim != 3:
            raise ValueError("GPU forces need dim == 3")
        return self._buf.upload("pos", np.asarray(s.pos, float), L.torch().float64)

    def _out(self, n):
        th = L.torch()
        return self._buf.get("F", (n, 3), th.float64), self._buf.get("U", (1,), th.float64)

    @staticmethod
    def _fetch(F, U):
        return F.cpu().numpy(), float(U.cpu().numpy()[0])


class LennardJones(_DeviceForce):
    """12-6 Lennard-Jones with Lorentz-Berthelot mixing (forces.py:101-175)."""

    def __init__(self, params: dict, cutoff: float = 2.5, pairlist_threshold: int = 400, skin: float = 0.3,
                 precision: str = "fp64"):
        self.params = {k: (si(e, "J"), si(sg, "m")) for k, (e, sg) in params.items()}
        self.cutoff = cutoff
        self.pairlist_threshold = pairlist_threshold
        self.skin = skin
        self._vl = VerletList()
        self._setup(precision)

    @property
    def rebuilds(self):
        return self._vl.rebuilds

    def invalidate_list(self):
        """Rebuild the neighbour list at the next compute (a fresh list, with its device
        buffers kept)."""
        self._vl.invalidate()

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

    def device_tables(self, s):
        eps_t, sig_t = self._tables(s)
        src6 = (1 / self.cutoff ** 2) ** 3
        return (np.ascontiguousarray(eps_t), np.ascontiguousarray(sig_t ** 2),
                np.ascontiguousarray((self.cutoff * sig_t) ** 2), float(src6 ** 2 - src6))

    def list_radius(self, s):
        eps_t, sig_t = self._tables(s)
        smax = float(np.max(np.where(eps_t > 0, sig_t, 0)))
        return self.cutoff * smax, self.skin * smax

    def compute(self, s):
        if s.n < 2:
            return np.zeros_like(s.pos), 0.0
        th = L.torch()
        pos_d = self._pos(s)
        sp_d = self._buf.upload("sp", np.asarray(s.sp, np.int32), th.int32)
        h = None
        if s.n >= self.pairlist_threshold:
            rc, skin = self.list_radius(s)
            self._vl.update(s, pos_d, rc, skin)
            h = self._vl.nl.handle
        eps, sig2, rc2, shift = self.device_tables(s)
        box, periodic = _box_args(s)
        F, U = self._out(s.n)
        L.call("accu_lj_forces", h, L.ptr(pos_d), L.ptr(sp_d), s.n, eps.shape[0], L.ptr(eps), L.ptr(sig2),
               L.ptr(rc2), shift, L.ptr(box), periodic, self._prec, L.ptr(F), L.ptr(U), L.stream())
        return self._fetch(F, U)


def _exclusion_keys(exclude, s):
    """Sorted int64 keys lo*n+hi of excluded index pairs (forces.py:417-424)."""
    if not exclude:
        return np.zeros(0, np.int64)
    where = {pid: n for n, pid in enumerate(s.ids)}
    ex = np.array([(where[a], where[b]) for a, b in exclude if a in where and b in where], dtype=np.int64).reshape(-1, 2)
    if not len(ex):
        return np.zeros(0, np.int64)
    lo, hi = np.minimum(ex[:, 0], ex[:, 1]), np.maximum(ex[:, 0], ex[:, 1])
    return np.unique(lo * s.n + hi).astype(np.int64)


class DampedShiftedCoulomb(_DeviceForce):
    """Damped shifted force electrostatics (forces.py:391-452)."""

    def __init__(self, rcut, alpha=None, k=KE_CONST, pairlist_threshold=400, exclude=(), skin=None,
                 precision: str = "fp64"):
        self.pairlist_threshold = pairlist_threshold
        self.exclude = [tuple(p) for p in exclude]
        self.k = si(k, "J*m/C^2")
        self.rcut = si(rcut, "m")
        self.alpha = 2.0 / self.rcut if alpha is None else si(alpha, "1/m")
        self.skin = 0.15 * self.rcut if skin is None else si(skin, "m")
        self._vl = VerletList()
        self._setup(precision)

    @property
    def rebuilds(self):
        return self._vl.rebuilds

    def invalidate_list(self):
        """Rebuild the neighbour list at the next compute (a fresh list, with its device
        buffers kept)."""
        self._vl.invalidate()

    def device_forces(self, s):
        """The forces on the device: (F, U, positions) device buffers and the neighbour
        list they used (None for the all-pairs path); nothing is copied back."""
        th = L.torch()
        a, rc = self.alpha, self.rcut
        erc = float(_erfc(np.array(a * rc)))
        pos_d = self._pos(s)
        q_d = self._buf.upload("q", np.asarray(s.charges, float), th.float64)
        keys = _exclusion_keys(self.exclude, s)
        ex_d = self._buf.upload("ex", keys, th.int64) if len(keys) else None
        nl = None
        if s.n >= self.pairlist_threshold:
            self._vl.update(s, pos_d, rc, self.skin)
            nl = self._vl.nl
        box, periodic = _box_args(s)
        F, U = self._out(s.n)
        L.call("accu_coulomb_forces", None if nl is None else nl.handle, L.ptr(pos_d), L.ptr(q_d), s.n, 1, float(a),
               float(rc), erc, float(self.k), L.ptr(ex_d), len(keys), L.ptr(box), periodic, self._prec, L.ptr(F),
               L.ptr(U), L.stream())
        return F, U, pos_d, nl

    def compute(self, s):
        if not np.any(s.charges):
            return np.zeros_like(s.pos), 0.0
        F, U, _, _ = self.device_forces(s)
        return self._fetch(F, U)


class Gravity(_DeviceForce):
    """Newtonian gravity between all particles with Plummer softening (forces.py:31-48)."""

    def __init__(self, G=G_CONST, softening=0.0, precision: str = "fp64"):
        self.G = si(G, "m^3/kg/s^2")
        self.eps2 = si(softening, "m") ** 2
        self._setup(precision)

    def compute(self, s):
        if s.n < 2:
            return np.zeros_like(s.pos), 0.0
        th = L.torch()
        pos_d = self._pos(s)
        m_d = self._buf.upload("m", np.asarray(s.masses, float), th.float64)
        box, periodic = _box_args(s)
        origin = np.ascontiguousarray(np.mean(s.pos, axis=0)) if self._prec else None
        F, U = self._out(s.n)
        L.call("accu_gravity_direct", L.ptr(pos_d), L.ptr(m_d), s.n, None, 0, float(self.G), float(self.eps2),
               L.ptr(box), periodic, self._prec, L.ptr(origin), L.ptr(F), L.ptr(U), L.stream())
        return self._fetch(F, U)


def kvector_columns(box, alpha, kmax=12, kcut=None):
    """The k-vector set of forces.py:347-357 (|n|^2 <= kmax^2, n != 0; or |k| <= kcut),
    as columns (nx, ny, nz_lo, nz_hi) covering one half space (int32 [ncol, 4])."""
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
    half = (N[:, 0] > 0) | ((N[:, 0] == 0) & (N[:, 1] > 0)) | ((N[:, 0] == 0) & (N[:, 1] == 0) & (N[:, 2] > 0))
    N = N[half]
    if not len(N):
        return np.zeros((0, 4), np.int32)
    key = N[:, 0] * 1_000_003 + N[:, 1]
    order = np.lexsort((N[:, 2], key))
    N, key = N[order], key[order]
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    ends = np.r_[starts[1:], len(N)] - 1
    cols = np.stack([N[starts, 0], N[starts, 1], N[starts, 2], N[ends, 2]], 1).astype(np.int32)
    # each column must be a contiguous nz range
    assert np.all(cols[:, 3] - cols[:, 2] + 1 == ends - starts + 1)
    return np.ascontiguousarray(cols)


class Ewald(_DeviceForce):
    """Ewald summation for a neutral 3D orthorhombic periodic box (forces.py:333-388)."""

    def __init__(self, k=KE_CONST, alpha=None, rcut=None, kmax=12, kcut=None, pairlist_threshold=400,
                 precision: str = "fp64"):
        self.k = si(k, "J*m/C^2")
        self.alpha = None if alpha is None else si(alpha, "1/m")
        self.rcut = None if rcut is None else si(rcut, "m")
        self.kmax = kmax
        self.kcut = None if kcut is None else si(kcut, "1/m")
        self.pairlist_threshold = pairlist_threshold
        self._kcache = None
        self._nl = NeighbourList()
        self._setup(precision)

    def params(self, s):
        rcut = self.rcut or 0.5 * float(np.min(s.box))
        return rcut, self.alpha or 4.0 / rcut

    def columns(self, box, alpha):
        key = (tuple(np.asarray(box, float)), alpha, self.kmax, self.kcut)
        if self._kcache is None or self._kcache[0] != key:
            self._kcache = (key, kvector_columns(box, alpha, self.kmax, self.kcut))
        return self._kcache[1]

    def compute(self, s):
        if s.dim != 3 or not s.periodic:
            raise ValueError("Ewald needs a 3D periodic box")
        q = s.charges
        if not np.any(q):
            return np.zeros_like(s.pos), 0.0
        if abs(q.sum()) > 1e-9 * np.sum(np.abs(q)):
            raise ValueError("Ewald summation here requires a charge-neutral system")
        th = L.torch()
        rcut, a = self.params(s)
        pos_d = self._pos(s)
        q_d = self._buf.upload("q", np.asarray(q, float), th.float64)
        box = np.ascontiguousarray(s.box, float)
        h = None
        if s.n >= self.pairlist_threshold:
            # list radius a hair above rcut; the kernel applies the exact test r < rcut
            self._nl.build(s, pos_d, rcut * (1 + 1e-9))
            h = self._nl.handle
        F, U = self._out(s.n)
        U2 = self._buf.get("U2", (1,), th.float64)
        L.call("accu_coulomb_forces", h, L.ptr(pos_d), L.ptr(q_d), s.n, 0, float(a), float(rcut), 0.0, float(self.k),
               None, 0, L.ptr(box), 1, self._prec, L.ptr(F), L.ptr(U), L.stream())
        cols = self.columns(box, a)
        if len(cols):
            L.call("accu_ewald_recip", L.ptr(pos_d), L.ptr(q_d), s.n, L.ptr(box), float(a), L.ptr(cols), len(cols),
                   float(self.k), 1, L.ptr(F), L.ptr(U2), L.stream())
            Fh, Uh = self._fetch(F, U)
            return Fh, Uh + float(U2.cpu().numpy()[0])
        return self._fetch(F, U)
