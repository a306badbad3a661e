"""Barnes-Hut octree gravity (accusim/tree.py) with the walk on the GPU.

The tree is built exactly as tree.py:30-53 builds it; the walk accepts the same
nodes (tree.py:67-70). `quadrupole=True` adds quadrupole moments to accepted
nodes (an extension; the reference expansion is monopole only).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .forces import Force
from .units import G as G_CONST
from .units import si


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
        self._buf = L.Buffers()

    def compute(self, s):
        if s.n < 2:
            return np.zeros_like(s.pos), 0.0
        if s.dim != 3:
            raise ValueError("GPU Barnes-Hut needs dim == 3")
        th = L.torch()
        pos = np.ascontiguousarray(s.pos, float)
        m = np.ascontiguousarray(s.masses, float)
        F = self._buf.get("F", (s.n, 3), th.float64)
        U = self._buf.get("U", (1,), th.float64)
        nn = L.C.c_longlong(0)
        L.call("accu_bh_forces", L.ptr(pos), L.ptr(m), s.n, float(self.G), float(self.eps2), float(self.theta),
               int(self.leaf_size), int(bool(self.quadrupole)), L.ptr(F), L.ptr(U), L.C.byref(nn), L.stream())
        self.nodes = nn.value
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
