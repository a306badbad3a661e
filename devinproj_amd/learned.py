"""Learned DSF -> Ewald pair correction (no counterpart in the reference).

`PairCorrection` is a small antisymmetric pair MLP, [r, q_i q_j, n_i, n_j, q_i + q_j] ->
64 -> 64 -> 1, whose scalar output acts along r_ij. The coordination numbers n_i enter
symmetrised as (n_i + n_j, |n_i - n_j|), and q_i + q_j tells like pairs of different
species apart (Na-Na from Cl-Cl, which share q_i q_j), so s_ij == s_ji and the correction
forces cancel in pairs (momentum is conserved). Because n_i depends on positions and is not
differentiated, the correction is not the gradient of an energy; it reports U = 0 and
`conservative = False`.

`LearnedDSF` is DampedShiftedCoulomb (forces.py:391-452) plus the correction, with the
same constructor and the trained network attached. `train_correction` fits the
network to Ewald-minus-DSF force residuals on the GPU (forward and backward kernels in
csrc/mlp.hip, Adam on device), keeping Ewald as the judge.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import _lib as L
from .forces import DampedShiftedCoulomb, Ewald, Force, NeighbourList, _box_args
from .units import e_charge

H = 64
N_IN = 5
N_PARAM = H * N_IN + H + H * H + H + H + 1
_W1, _B1 = 0, H * N_IN
_W2 = _B1 + H
_B2 = _W2 + H * H
_W3 = _B2 + H
_B3 = _W3 + H
OFFSETS = {"W1": (_W1, (H, N_IN)), "b1": (_B1, (H,)), "W2": (_W2, (H, H)), "b2": (_B2, (H,)), "w3": (_W3, (H,)),
           "b3": (_B3, (1,))}


def init_params(seed=0):
    """Flat float32 parameter vector: W1 ~ N(0, 1/5), W2 ~ N(0, 1/64), w3 ~ N(0, (0.1/8)^2), zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM, np.float32)
    p[_W1:_B1] = rng.normal(0, 1 / np.sqrt(N_IN), H * N_IN)
    p[_W2:_B2] = rng.normal(0, 1 / np.sqrt(H), H * H)
    p[_W3:_B3] = rng.normal(0, 0.1 / np.sqrt(H), H)
    return p


def unflatten(p):
    return {k: np.asarray(p[o:o + int(np.prod(sh))], np.float64).reshape(sh) for k, (o, sh) in OFFSETS.items()}


def flatten(d):
    p = np.zeros(N_PARAM, np.float32)
    for k, (o, sh) in OFFSETS.items():
        p[o:o + int(np.prod(sh))] = np.asarray(d[k], np.float32).reshape(-1)
    return p


class PairCorrection(Force):
    conservative = False

    def __init__(self, rcut, r_on, r_off, n_scale=8.0, F0=1.0, charge_unit=e_charge, params=None, seed=0):
        self.rcut, self.r_on, self.r_off = float(rcut), float(r_on), float(r_off)
        self.n_scale, self.F0, self.charge_unit = float(n_scale), float(F0), float(charge_unit)
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        if p.shape != (N_PARAM,):
            raise ValueError(f"expected {N_PARAM} parameters")
        h = L.C.c_void_p()
        L.call("accu_mlp_create", L.C.byref(h), L.ptr(p))
        self._h = h
        self._configure()
        self.nl = NeighbourList()
        self._buf = L.Buffers()
        self.steps = 0

    def _configure(self):
        L.call("accu_mlp_configure", self._h, self.rcut, self.r_on, self.r_off, self.n_scale, self.F0)

    def set_scale(self, F0):
        self.F0 = float(F0)
        self._configure()

    @property
    def params(self):
        p = np.empty(N_PARAM, np.float32)
        L.call("accu_mlp_params", self._h, L.ptr(p), None)
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        L.call("accu_mlp_params", self._h, None, L.ptr(p))

    def prepare(self, s, pos_d=None, nl=None):
        """Upload positions/charges and compute the coordination numbers over a neighbour
        list: `nl` when given (any list of these positions with radius >= max(rcut, r_off),
        e.g. the DSF force's own), else a list built here at that radius."""
        th = L.torch()
        if pos_d is None:
            pos_d = self._buf.upload("pos", np.asarray(s.pos, float), th.float64)
        q_d = self._buf.upload("q", np.asarray(s.charges, float), th.float64)
        if nl is None:
            self.nl.build(s, pos_d, max(self.rcut, self.r_off) * (1 + 1e-9))
            nl = self.nl
        self._cur = nl
        box, periodic = _box_args(s)
        L.call("accu_mlp_prepare", self._h, nl.handle, L.ptr(pos_d), L.ptr(q_d), 1.0 / self.charge_unit, s.n,
               L.ptr(box), periodic, L.stream())
        return pos_d

    def add_forces(self, s, pos_d, F_d, accumulate=True):
        """Correction forces over the list of the last prepare(), into F_d."""
        box, periodic = _box_args(s)
        L.call("accu_mlp_forces", self._h, self._cur.handle, L.ptr(pos_d), s.n, L.ptr(box), periodic,
               int(bool(accumulate)), L.ptr(F_d), L.stream())

    def compute(self, s):
        th = L.torch()
        pos_d = self.prepare(s)
        F = self._buf.get("F", (s.n, 3), th.float64)
        self.add_forces(s, pos_d, F, accumulate=False)
        return F.cpu().numpy(), 0.0

    def grad(self, s, target, pos_d=None, want_grad=True):
        """(loss, gradient) of mean(((F_corr - target)/F0)^2) for one snapshot; target in newtons."""
        th = L.torch()
        pos_d = self.prepare(s, pos_d)
        t_d = target if hasattr(target, "data_ptr") else self._buf.upload("target", np.asarray(target, float), th.float64)
        g = np.empty(N_PARAM, np.float32) if want_grad else None
        loss = L.C.c_double(0.0)
        box, periodic = _box_args(s)
        L.call("accu_mlp_grad", self._h, self._cur.handle, L.ptr(pos_d), s.n, L.ptr(box), periodic, L.ptr(t_d),
               L.ptr(g), L.C.byref(loss), L.stream())
        return loss.value, g

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        L.call("accu_mlp_adam", self._h, float(lr), float(b1), float(b2), float(eps), L.stream())
        self.steps += 1

    def close(self):
        if getattr(self, "_h", None) is not None and L._lib is not None:
            L._lib.accu_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class LearnedDSF(DampedShiftedCoulomb):
    """DampedShiftedCoulomb plus a trained PairCorrection (forces only; U is the DSF energy).
    The correction is added on the device into the DSF forces, over the DSF's own Verlet
    list when it has one (its radius rc + skin covers the correction's rcut and r_off), so
    one list serves both and one array is copied back."""

    conservative = False

    def __init__(self, rcut, alpha=None, k=None, correction=None, **kw):
        from .units import k_e
        super().__init__(rcut, alpha=alpha, k=k_e if k is None else k, **kw)
        self.correction = correction

    def compute(self, s):
        c = self.correction
        if c is None or not np.any(s.charges):
            return super().compute(s)
        F, U, pos_d, nl = self.device_forces(s)
        radius = self.rcut + self.skin
        c.prepare(s, pos_d, nl if nl is not None and radius >= max(c.rcut, c.r_off) else None)
        c.add_forces(s, pos_d, F, accumulate=True)
        return self._fetch(F, U)


NACL_CORRECTION = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "nacl_correction.npz")


def save_correction(corr, path, meta=None):
    """Parameters and geometry of a PairCorrection to an .npz (no pickled objects)."""
    np.savez(path, params=corr.params, F0=corr.F0, rc=corr.rcut, r_on=corr.r_on, r_off=corr.r_off,
             n_scale=corr.n_scale, meta=json.dumps(meta or {}))


def load_correction(path=NACL_CORRECTION):
    """PairCorrection from save_correction's file (allow_pickle=False). The default is the
    network trained by train_nacl on the 65,536-ion NaCl melt (DSF alpha 0.2/A, rc 12 A),
    whose training record is in `corr.meta`."""
    d = np.load(path, allow_pickle=False)
    corr = PairCorrection(float(d["rc"]), float(d["r_on"]), float(d["r_off"]), n_scale=float(d["n_scale"]),
                          F0=float(d["F0"]), params=np.asarray(d["params"], np.float32))
    corr.meta = json.loads(str(d["meta"]))
    return corr


def residual_targets(snapshots, dsf, ewald):
    """Ewald - DSF forces (newtons) for each snapshot system, computed on the GPU."""
    out = []
    for s in snapshots:
        Fe, _ = ewald.compute(s)
        Fd, _ = dsf.compute(s)
        out.append(Fe - Fd)
    return out


def train_correction(corr, snapshots, targets, epochs=10, lr=3e-3, seed=0, log=None, lr_final=0.05):
    """Adam over snapshots (one snapshot per step, shuffled each epoch), the learning rate
    decaying from lr to lr * lr_final on a cosine over all steps. Returns the loss history
    (mean per epoch)."""
    rng = np.random.default_rng(seed)
    th = L.torch()
    tgt = [th.from_numpy(np.ascontiguousarray(t, float)).cuda() for t in targets]
    hist = []
    total, step = epochs * len(snapshots), 0
    for ep in range(epochs):
        tot = 0.0
        for k in rng.permutation(len(snapshots)):
            loss, _ = corr.grad(snapshots[k], tgt[k], want_grad=False)
            frac = step / max(total - 1, 1)
            corr.adam(lr * (lr_final + (1.0 - lr_final) * 0.5 * (1.0 + np.cos(np.pi * frac))))
            step += 1
            tot += loss
        hist.append(tot / len(snapshots))
        if log:
            log(f"epoch {ep}: mean loss {hist[-1]:.4e}")
    return hist
