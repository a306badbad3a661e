"""Pair counts of the 1,048,576-atom argon workload at its Verlet-list radius, from the
oracle's cell_list_pairs (restating forces.py:279-321), for the full-size GPU test.

The workload is 64^3 x 4 FCC argon at 1374 kg/m^3, positions jittered by N(0, 0.01 sigma)
from seed 2, then wrapped. Its list radius is (2.5 + 0.3) sigma. The brick engine's list
(every pair counted from both atoms, built from integer coordinates on a quantum of at most
h / 4096 of the cell edge h, with a threshold that covers rc_list plus the rounding of both
atoms) must hold between 2 * n_lo and 2 * n_hi entries, where
  n_lo = pairs i < j with minimum-image |d| < rc_list,
  n_hi = pairs i < j with |d| < rc_list + 2 sqrt(3) h / 4096.
Also stored: the pairs within the LJ cutoff 2.5 sigma, and a fingerprint of the positions
so that the test can tell that it rebuilt the same configuration.

Usage: python tests/golden/make_c2_pairs.py   (about 20 GB of memory, a few minutes)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import accusim_np as A  # noqa: E402
from oracle import builders as B  # noqa: E402


def main():
    s = B.argon(64, seed=2, jitter=0.01)
    sigma = B.AR_SIGMA
    rc_cut = 2.5 * sigma
    rc_list = rc_cut + 0.3 * sigma
    h = float(np.max(s.box / np.floor(s.box / rc_list)))  # the engine's cell edge
    rc_hi = rc_list + 2.0 * np.sqrt(3.0) * h / 4096.0
    # a radius just above the upper bracket, so one call returns every pair of both counts
    i, j = A.cell_list_pairs(s, rc_hi * (1 + 1e-9))
    d = s.pos[j] - s.pos[i]
    d -= s.box * np.round(d / s.box)
    r2 = np.einsum("ij,ij->i", d, d)
    out = {
        "n": s.n,
        "rc_cut": rc_cut,
        "rc_list": rc_list,
        "n_cut": int(np.sum(r2 < rc_cut ** 2)),
        "n_lo": int(np.sum(r2 < rc_list ** 2)),
        "rc_hi": rc_hi,
        "n_hi": int(np.sum(r2 < rc_hi ** 2)),
        "pos_sum": s.pos.sum(axis=0),
        "pos_head": s.pos[:4].copy(),
        "vel_head": s.vel[:4].copy(),
    }
    np.savez(os.path.join(HERE, "argon1m_pairs.npz"), **out)
    print({k: v for k, v in out.items() if np.ndim(v) == 0})


if __name__ == "__main__":
    main()
