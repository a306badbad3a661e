"""Train the learned DSF -> Ewald correction on molten NaCl, entirely on the GPU.

    python -m devinproj_amd.train_nacl [--cells 32 32 64] [--snapshots 200] [--heldout 20]
                                       [--every 200] [--epochs 60] [--out path.npz]

The network shipped in devinproj_amd/data/nacl_correction.npz (learned.load_correction)
was trained by this script with its defaults.

1. Melt the rocksalt lattice the way examples/molten_salt.py:36-42 does: Langevin at
   3000 K for 2 ps (a perfect crystal superheats), then at the target 1200 K for 3 ps,
   on the brick engine (LJ + DSF); the rms displacement from the lattice (minimum image)
   must be far above 0.3 A. Then sample NVE snapshots `every` steps (0.2 ps) apart.
2. The held-out snapshots come from an independent melt (another seed), not from the
   training trajectory.
3. Targets: Ewald forces (real-space error 1e-6: alpha rc = 3.459, reciprocal cutoff at
   the same error) minus DSF forces, both from the GPU Force classes.
4. The least-squares bound: the best pairwise radial correction g_type(r) (piecewise
   linear in r per species pair) fitted to the training residuals, evaluated on the
   held-out snapshots. A pair correction that depends on r and the species alone cannot
   do better on average; the network (which also sees coordination numbers) is trained
   until it does at least as well.
5. Fit the pair MLP with the GPU forward/backward kernels and Adam, and report on the
   held-out snapshots the RMS relative force error against Ewald
   (accuracy.py's sqrt(mean|dF|^2 / mean|F|^2)) of DSF alone, of DSF + the least-squares
   pair correction and of DSF + the network.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib as L
from .accuracy import rms_rel_error
from .batch import BrickEngine
from .forces import DampedShiftedCoulomb, Ewald, LennardJones, cell_list_pairs
from .integrators import Langevin, VelocityVerlet
from .learned import PairCorrection, save_correction, train_correction
from .workloads import NACL_LJ, rocksalt

RC = 12e-10
ALPHA_DSF = 0.2e10


def ewald_1e6(rc=RC):
    a = 3.459 / rc
    return Ewald(alpha=a, rcut=rc, kcut=2 * a * np.sqrt(np.log(1e6)))


def _forces():
    return [LennardJones(NACL_LJ, cutoff=2.5, precision="mixed"),
            DampedShiftedCoulomb(rcut=RC, alpha=ALPHA_DSF, precision="mixed")]


def melt(cells, T=1200.0, T_hot=3000.0, hot_ps=2.0, cool_ps=3.0, friction=5e12, seed=3, log=None):
    """Rocksalt lattice melted under Langevin dynamics (hot, then at T), dt = 1 fs.
    Returns the system and its rms displacement from the lattice (metres)."""
    s = rocksalt(*cells, T=T_hot, seed=seed)
    lattice = s.pos.copy()
    for k, (temp, ps) in enumerate(((T_hot, hot_ps), (T, cool_ps))):
        eng = BrickEngine([s], [_forces()], Langevin(temp, friction, seed=seed * 7 + k), 1e-15, precision="mixed",
                          bricks=[(1, 1, 1)])
        eng.run(int(round(ps * 1000)), record_every=0)
        eng.fetch()
        eng.close()
    d = s.pos - lattice
    d -= s.box * np.round(d / s.box)
    rms = float(np.sqrt(np.mean(np.sum(d ** 2, 1))))
    if log:
        log(f"melt seed {seed}: Langevin {T_hot:.0f} K {hot_ps} ps then {T:.0f} K {cool_ps} ps, "
            f"rms displacement from the lattice {rms * 1e10:.2f} A")
    return s, rms


def sample(s, n_snap, every, log=None):
    """NVE snapshots `every` velocity-Verlet steps apart (dt = 1 fs)."""
    eng = BrickEngine([s], [_forces()], VelocityVerlet(), 1e-15, precision="mixed", bricks=[(1, 1, 1)])
    snaps = []
    for _ in range(n_snap):
        eng.run(every, record_every=0)
        eng.fetch()
        snaps.append(s.copy())
    eng.close()
    return snaps


def targets(snaps):
    ew = ewald_1e6()
    dsf = DampedShiftedCoulomb(rcut=RC, alpha=ALPHA_DSF)
    Fe, Fd = [], []
    for s in snaps:
        Fe.append(ew.compute(s)[0])
        Fd.append(dsf.compute(s)[0])
    return Fe, Fd


class PairLSFit:
    """Least-squares pairwise radial correction dF_i = sum_j g_t(r_ij) (r_j - r_i) / r_ij,
    t the species pair, g_t piecewise linear on `bins` nodes over [r0, rc]. The normal
    equations are accumulated over snapshots on the GPU (device buffers via torch)."""

    def __init__(self, nspecies=2, bins=128, r0=1.5e-10, rc=RC, ridge=1e-9):
        self.ns, self.B, self.r0, self.rc, self.ridge = nspecies, bins, r0, rc, ridge
        self.ntype = nspecies * (nspecies + 1) // 2
        self.P = self.ntype * bins
        th = L.torch()
        self.G = th.zeros((self.P, self.P), dtype=th.float64, device="cuda")
        self.h = th.zeros(self.P, dtype=th.float64, device="cuda")
        self.coef = None

    def _design(self, s):
        th = L.torch()
        i, j = cell_list_pairs(s, self.rc)
        i = th.from_numpy(i).cuda()
        j = th.from_numpy(j).cuda()
        pos = th.from_numpy(np.asarray(s.pos, float)).cuda()
        box = th.from_numpy(np.asarray(s.box, float)).cuda()
        sp = th.from_numpy(np.asarray(s.sp, np.int64)).cuda()
        d = pos[j] - pos[i]
        d -= box * th.round(d / box)
        r = th.sqrt((d * d).sum(1))
        u = ((r - self.r0) / (self.rc - self.r0) * (self.B - 1)).clamp(0, self.B - 1 - 1e-9)
        k = th.floor(u).long()
        f = u - k
        a, b = th.minimum(sp[i], sp[j]), th.maximum(sp[i], sp[j])
        t = a * self.ns - a * (a - 1) // 2 + (b - a)  # species pair index, 0 .. ntype-1
        col = t * self.B + k
        unit = d / r[:, None]
        A = th.zeros(3 * s.n * self.P, dtype=th.float64, device="cuda")
        for c in range(3):
            for w, cc in ((1 - f, col), (f, col + 1)):
                v = w * unit[:, c]
                A.index_add_(0, (3 * i + c) * self.P + cc, v)
                A.index_add_(0, (3 * j + c) * self.P + cc, -v)
        return A.view(3 * s.n, self.P)

    def add(self, s, dF):
        th = L.torch()
        A = self._design(s)
        b = th.from_numpy(np.asarray(dF, float).reshape(-1)).cuda()
        self.G += A.T @ A
        self.h += A.T @ b

    def solve(self):
        th = L.torch()
        lam = self.ridge * float(th.diagonal(self.G).mean())
        G = self.G + lam * th.eye(self.P, dtype=th.float64, device="cuda")
        self.coef = th.linalg.solve(G, self.h)
        return self

    def predict(self, s):
        return (self._design(s) @ self.coef).view(s.n, 3).cpu().numpy()


def heldout_errors(snaps, Fe, Fd, corr=None, ls=None):
    """Mean over snapshots of the RMS relative force error against Ewald: DSF alone, DSF +
    the least-squares pair correction, DSF + the network."""
    out = {"dsf": [], "ls": [], "net": []}
    for s, fe, fd in zip(snaps, Fe, Fd):
        out["dsf"].append(rms_rel_error(fd, fe))
        if ls is not None:
            out["ls"].append(rms_rel_error(fd + ls.predict(s), fe))
        if corr is not None:
            out["net"].append(rms_rel_error(fd + corr.compute(s)[0], fe))
    return {k: float(np.mean(v)) for k, v in out.items() if v}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, nargs=3, default=[32, 32, 64])
    ap.add_argument("--snapshots", type=int, default=200)
    ap.add_argument("--heldout", type=int, default=20)
    ap.add_argument("--every", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    L.lib()
    t0 = time.perf_counter()
    s_train, rms_train = melt(tuple(a.cells), seed=3, log=log)
    s_test, rms_test = melt(tuple(a.cells), seed=1003, log=log)
    train = sample(s_train, a.snapshots, a.every)
    test = sample(s_test, a.heldout, a.every)
    t1 = time.perf_counter()
    Fe_tr, Fd_tr = targets(train)
    Fe_te, Fd_te = targets(test)
    dF = [e - d for e, d in zip(Fe_tr, Fd_tr)]
    t2 = time.perf_counter()
    ls = PairLSFit()
    for s, d in zip(train, dF):
        ls.add(s, d)
    ls.solve()
    t3 = time.perf_counter()
    log(f"least-squares pair correction: {ls.P} coefficients")
    F0 = float(np.sqrt(np.mean([np.mean(np.sum(x ** 2, 1)) for x in dF])))
    corr = PairCorrection(RC, 2.8e-10, 3.8e-10, n_scale=8.0, F0=F0)
    hist = train_correction(corr, train, dF, epochs=a.epochs, lr=a.lr, log=log)
    t4 = time.perf_counter()
    err = heldout_errors(test, Fe_te, Fd_te, corr=corr, ls=ls)
    err_train = heldout_errors(train[:a.heldout], Fe_tr[:a.heldout], Fd_tr[:a.heldout], corr=corr, ls=ls)
    res = {"ions": train[0].n, "train_snapshots": len(train), "heldout": len(test), "every_steps": a.every,
           "heldout_source": "independent melt (seed 1003)",
           "rms_displacement_from_lattice_A": {"train": rms_train * 1e10, "heldout": rms_test * 1e10},
           "epochs": a.epochs, "F0_newton": F0, "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
           "rms_rel_error_dsf_vs_ewald": err["dsf"],
           "rms_rel_error_dsf_plus_ls_pair_bound_vs_ewald": err["ls"],
           "rms_rel_error_dsf_plus_correction_vs_ewald": err["net"],
           "training_set_errors": err_train,
           "seconds": {"melt_and_sample": t1 - t0, "ewald_targets": t2 - t1, "ls_fit": t3 - t2,
                       "training": t4 - t3}}
    print(json.dumps(res), flush=True)
    if a.out:
        save_correction(corr, a.out, meta=res)
    return res, corr


if __name__ == "__main__":
    main()
