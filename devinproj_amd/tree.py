"""Barnes-Hut octree gravity (accusim/tree.py) with the tree built and walked on the GPU.

The octree is the reference's (tree.py:30-53: root cube, octant order, leaf_size,
depth < 48), built by the device from sorted octant-path keys; the walk accepts the
same nodes (tree.py:67-70). `quadrupole=True` adds quadrupole moments to accepted
nodes (an extension; the reference expansion is monopole only).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .forces import Force
from .units import G as G_CONST
from .units import si


class Octree:
    """Device octree (accu_bh_* handle): build from device positions/masses, walk, export."""

    def __init__(self):
        self._h = None
        self.n = 0

    @property
    def handle(self):
        if self._h is None:
            h = L.C.c_void_p()
            L.call("accu_bh_create", L.C.byref(h))
            self._h = h
        return self._h

    def build(self, pos_d, m_d, n, leaf_size=8, quadrupole=False):
        L.call("accu_bh_build", self.handle, L.ptr(pos_d), L.ptr(m_d), int(n), int(leaf_size),
               int(bool(quadrupole)), L.stream())
        self.n = int(n)

    def walk(self, G, eps2, theta, F_d, U_d=None):
        L.call("accu_bh_walk", self.handle, float(G), float(eps2), float(theta), L.ptr(F_d), L.ptr(U_d), L.stream())

    def info(self):
        out = np.zeros(6, np.int64)
        L.call("accu_bh_info", self.handle, L.ptr(out))
        return {"n": int(out[0]), "nodes": int(out[1]), "leaves": int(out[2]), "max_depth": int(out[3]),
                "deep": bool(out[4]), "leaf_size": int(out[5])}

    def export(self, quadrupole=False):
        """The tree in pre-order as numpy arrays (cube, com, next, leaf, start, count, perm[, quad])."""
        th = L.torch()
        nn = self.info()["nodes"]
        f64 = dict(dtype=th.float64, device="cuda")
        i32 = dict(dtype=th.int32, device="cuda")
        t = {"cube": th.empty((nn, 4), **f64), "com": th.empty((nn, 4), **f64), "next": th.empty(nn, **i32),
             "leaf": th.empty(nn, **i32), "start": th.empty(nn, **i32), "count": th.empty(nn, **i32),
             "perm": th.empty(self.n, **i32)}
        q = th.empty((nn, 6), **f64) if quadrupole else None
        L.call("accu_bh_export", self.handle, L.ptr(t["cube"]), L.ptr(t["com"]), L.ptr(q), L.ptr(t["next"]),
               L.ptr(t["leaf"]), L.ptr(t["start"]), L.ptr(t["count"]), L.ptr(t["perm"]), L.stream())
        out = {k: v.cpu().numpy() for k, v in t.items()}
        if q is not None:
            out["quad"] = q.cpu().numpy()
        return out

    def __getstate__(self):
        return {"n": 0, "_h": None}

    def __del__(self):
        if self._h is not None and L._lib is not None:
            L._lib.accu_bh_destroy(self._h)
            self._h = None


class BarnesHut(Force):
    """Approximate Newtonian gravity. theta = 0 is exact direct summation."""

    exact = False

    def __init__(self, G=G_CONST, softening=0.0, theta=0.5, leaf_size=8, quadrupole=False):
        self.G = si(G, "m^3/kg/s^2")
        self.eps2 = si(softening, "m") ** 2
        self.theta = theta
        self.leaf_size = leaf_size
        self.quadrupole = quadrupole
        self.nodes = 0
        self.tree = Octree()
        self._buf = L.Buffers()

    def forces_device(self, pos_d, m_d, n, F_d, U_d=None):
        """Build the tree over device positions/masses and walk it into device F (and U)."""
        self.tree.build(pos_d, m_d, n, self.leaf_size, self.quadrupole)
        self.tree.walk(self.G, self.eps2, self.theta, F_d, U_d)

    def compute(self, s):
        if s.n < 2:
            return np.zeros_like(s.pos), 0.0
        if s.dim != 3:
            raise ValueError("GPU Barnes-Hut needs dim == 3")
        th = L.torch()
        pos = self._buf.upload("pos", np.asarray(s.pos, float), th.float64)
        m = self._buf.upload("m", np.asarray(s.masses, float), th.float64)
        F = self._buf.get("F", (s.n, 3), th.float64)
        U = self._buf.get("U", (1,), th.float64)
        self.forces_device(pos, m, s.n, F, U)
        self.nodes = self.tree.info()["nodes"]
        return F.cpu().numpy(), float(U.cpu().numpy()[0])


def direct_subset(force, s, sub):
    """Exact gravitational force on particles `sub` from all particles (tree.py:81-87), on the GPU."""
    th = L.torch()
    sub = np.ascontiguousarray(sub, np.int64).reshape(-1)
    if len(sub) and (sub.min() < -s.n or sub.max() >= s.n):
        bad = sub[(sub < -s.n) | (sub >= s.n)][0]
        raise IndexError(f"index {bad} is out of bounds for axis 0 with size {s.n}")
    sub = np.where(sub < 0, sub + s.n, sub)  # numpy's negative indices, as s.pos[sub] takes them
    buf = L.Buffers()
    pos_d = buf.upload("pos", np.asarray(s.pos, float), th.float64)
    m_d = buf.upload("m", np.asarray(s.masses, float), th.float64)
    t_d = buf.upload("t", sub, th.int64)
    F = buf.get("F", (len(sub), 3), th.float64)
    L.call("accu_gravity_direct", L.ptr(pos_d), L.ptr(m_d), s.n, L.ptr(t_d), len(sub), float(force.G),
           float(force.eps2), None, 0, L.ACCU_FP64, None, L.ptr(F), None, L.stream())
    return F.cpu().numpy()
