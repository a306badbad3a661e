"""The oracle's restatement of accusim/analysis.py against closed forms and the
reference's own analysis test (tests/test_analysis.py:22-35)."""
import numpy as np
import pytest

from oracle import accusim_np as A
from oracle import builders as B


def test_rdf_counts_fcc_shells_exactly():
    # perfect FCC: every atom has 12 neighbours at a / sqrt(2) and 6 at a
    s = B.argon(6, seed=1)
    a = s.box[0] / 6
    # 37 bins of [0, 1.1 a]: both shells sit inside bins, away from the edges
    _, _, hist = A.rdf_counts(s, 1.1 * a, bins=37)
    first, second = int(37 / 1.1 / np.sqrt(2)), int(37 / 1.1)
    assert hist[first] == 12 * s.n // 2 and hist[second] == 6 * s.n // 2
    assert hist.sum() == (12 + 6) * s.n // 2


def test_rdf_counts_equal_a_brute_force_pair_loop():
    rng = np.random.default_rng(3)
    s = A.System(3, box=[7.0, 8.0, 9.0], periodic=True)
    s.add_species("X", 1.0)
    s.add("X", rng.uniform(0, 1, (150, 3)) * s.box)
    rmax, bins = 3.4, 17
    _, _, hist = A.rdf_counts(s, rmax, bins)
    ref = np.zeros(bins, int)
    w = rmax / bins
    for i in range(s.n):
        for j in range(i + 1, s.n):
            d = s.pos[j] - s.pos[i]
            d -= s.box * np.round(d / s.box)
            r = np.sqrt(d @ d)
            if r <= rmax:
                ref[min(int(r / w), bins - 1)] += 1
    assert np.array_equal(hist, ref)


def test_rdf_of_an_ideal_gas_is_one():
    rng = np.random.default_rng(0)
    s = A.System(3, box=10.0, periodic=True)
    s.add_species("X", 1.0)
    s.add("X", rng.uniform(0, 10, (2000, 3)))
    _, g = A.rdf(s, 5.0, bins=10)
    assert np.all(np.abs(g[2:] - 1) < 0.05)


def _traj(pos, vel, dt):
    tr = A.Trajectory()
    tr.t = [k * dt for k in range(len(pos))]
    tr.positions = [p.copy() for p in pos]
    tr.velocities = [v.copy() for v in vel]
    return tr


def test_msd_and_vacf_of_ballistic_motion_unwrap_across_faces():
    rng = np.random.default_rng(2)
    n, T, dt, L = 500, 40, 0.1, 5.0
    x0 = rng.uniform(0, L, (n, 3))
    v = rng.normal(size=(n, 3))
    pos = [(x0 + v * k * dt) % L for k in range(T)]
    tr = _traj(pos, [v] * T, dt)
    t, m = A.msd(tr, box=np.full(3, L))
    assert np.allclose(m, np.mean(np.sum(v ** 2, 1)) * t ** 2, rtol=1e-9, atol=1e-12)
    lag, C = A.vacf(tr, max_lag=10)
    assert np.allclose(C, np.mean(np.sum(v ** 2, 1)), rtol=1e-12)


def test_green_kubo_matches_langevin_diffusion():
    # tests/test_analysis.py:22-35: free Langevin particles, D = kT / (m gamma)
    s = A.System(3)
    m = 39.948 * 1.66053906660e-27
    s.add_species("Ar", m)
    s.add("Ar", np.zeros((400, 3)))
    s.maxwell_boltzmann(300, rng=np.random.default_rng(8))
    gamma = 1e12
    sim = A.Simulation(s, [], integrator=A.Langevin(300, gamma), dt=20e-15, seed=8)
    tr = sim.run(steps=4000, record_every=5)
    t, C = A.vacf(tr, max_lag=120)
    kT = A.K_B * 300
    assert abs(C[0] / (3 * kT / m) - 1) < 0.05
    assert abs(A.green_kubo_diffusion(tr, max_lag=120) / (kT / (m * gamma)) - 1) < 0.06


def test_pressure_of_an_ideal_gas_and_of_lj_pairs():
    rng = np.random.default_rng(5)
    s = A.System(3, box=20.0, periodic=True)
    s.add_species("X", 2.0)
    s.add("X", rng.uniform(0, 20, (300, 3)), vel=rng.normal(size=(300, 3)))
    sim = A.Simulation(s, [], dt=0.01)
    assert A.pressure(sim) == pytest.approx(2 * s.kinetic_energy() / (3 * 20.0 ** 3), rel=1e-12)
    # one LJ pair at rest: P = -dU/dV with U(r) scaled uniformly, dU/dV = r U'(r) / (3 V)
    s2 = A.System(3, box=10.0, periodic=True)
    s2.add_species("X", 1.0)
    s2.add("X", [[1.0, 1.0, 1.0], [2.2, 1.0, 1.0]])
    lj = A.LennardJones({"X": (1.0, 1.0)}, cutoff=2.5)
    sim2 = A.Simulation(s2, [lj], dt=0.01)
    r = 1.2
    dU = 4 * (-12 / r ** 13 + 6 / r ** 7)
    assert A.pressure(sim2, h=1e-7) == pytest.approx(-r * dU / (3 * 1000.0), rel=1e-6)
