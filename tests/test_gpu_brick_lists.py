"""Completeness of the brick engine's Verlet lists (csrc/brick.hip, k_b_build): right after
set-up the number of list entries (a full list: every pair counted from both atoms) lies
between twice the brute-force count of pairs within the list radius rc_list and twice the
count within rc_list + 2 sqrt(3) h / 4096, h the cell edge: the build tests integer
coordinates on a quantum of at most h / 4096 against a threshold that covers rc_list plus
the rounding of both atoms (at most sqrt(3) quanta each way). The build scans z-sorted
column windows, so a window that cut off a pair would show here as a count below the lower
bound."""
import numpy as np
import pytest

from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")
from devinproj_amd import batch  # noqa: E402


def _pair_counts(s, rc_list):
    L = np.asarray(s.box, float)
    h = float(np.max(L / np.floor(L / rc_list)))  # the engine's cell edge (cells of at least rc_list)
    rc_hi = rc_list + 2.0 * np.sqrt(3.0) * h / 4096.0
    n = s.n
    lo = hi = 0
    for i in range(0, n, 512):
        d = s.pos[i:i + 512, None, :] - s.pos[None, :, :]
        d -= L * np.round(d / L)
        r2 = np.einsum("ijk,ijk->ij", d, d)
        idx = np.arange(i, min(i + 512, n))
        r2[np.arange(len(idx)), idx] = np.inf
        lo += int(np.sum(r2 < rc_list ** 2))
        hi += int(np.sum(r2 < rc_hi ** 2 * (1 + 1e-9)))
    return lo, hi


def _uniform(n, L, seed):
    rng = np.random.default_rng(seed)
    s = B.argon(2, seed=seed)  # species table and masses of argon
    s2 = type(s)(3, box=L, periodic=True)
    s2.add_species("Ar", s.species[0].mass)
    s2.add("Ar", rng.uniform(0, L, (n, 3)), vel=np.zeros((n, 3)))
    return s2


@pytest.mark.parametrize("case", ["lattice", "uniform"])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_list_entries_bracket_the_brute_force_pair_count(case, precision):
    if case == "lattice":
        s = B.argon(12, seed=3, jitter=0.05)
    else:
        s = _uniform(6000, B.argon(12, seed=3).box[0], seed=5)
    lj = dp.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3)
    rc, sk = lj.list_radius(s)
    print(f"lists {case} {precision} n={s.n}", flush=True)
    for bricks in ([(1, 1, 1)], [(2, 2, 3)], None):
        eng = batch.BrickEngine([s.copy()], [[lj]], dp.VelocityVerlet(), 10e-15, precision=precision,
                                bricks=bricks)
        eng.fetch(state=False)
        lo, hi = _pair_counts(s, rc + sk)
        assert lo <= eng.info["list_entries"] <= hi, (bricks, lo, eng.info["list_entries"], hi)
        eng.close()
