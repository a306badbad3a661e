"""GPU parity of the brick engine's thermostatted integrators (integrators.py:50-116,
237-257) against the oracle's: Nose-Hoover step for step (state, records, bath energy
and thermostat variables), Langevin and Brownian step for step with the oracle driven by
the same Philox normals, the thermostats' physics, and graph replays."""
import dataclasses

import numpy as np
import pytest

from oracle import accusim_np as A
from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")
from devinproj_amd import batch  # noqa: E402

DT = 10e-15


def _lj(M):
    return [M.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3)]


def _close(a, b, rtol):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) <= rtol * np.max(np.abs(np.asarray(b)))


def _oracle(so, integ, steps, seed=None):
    sim = A.Simulation(so, _lj(A), integrator=integ, dt=DT)
    if seed is not None:
        sim.rng = A.PhiloxNormals(seed)
    tr = sim.run(steps=steps, positions=False)
    return sim, tr


def test_nose_hoover_matches_oracle_step_for_step():
    s = B.argon(6, seed=64, T=60.0)
    so = s.copy()
    eng = batch.BrickEngine([s], [_lj(dp)], dp.NoseHoover(120.0, 0.2e-12), DT)
    eng.run(40)
    eng.fetch()
    nh = A.NoseHoover(120.0, 0.2e-12)
    sim, tr = _oracle(so, nh, 40)
    lg = eng.logs[0]
    assert len(lg) == 41
    assert _close(lg[:, 0], tr.potential, 1e-9)
    assert _close(lg[:, 1], tr.kinetic, 1e-9)
    assert _close(lg[:, 9], tr.bath, 1e-9)
    assert _close(s.pos, so.pos, 1e-10)
    assert _close(s.vel, so.vel, 1e-9)
    xi, eta, Q = eng.thermostat_state()[0, :3]
    assert xi == pytest.approx(nh.xi, rel=1e-9) and eta == pytest.approx(nh.eta, rel=1e-9)
    assert Q == pytest.approx(nh.Q, rel=1e-12)
    tr_dev = eng.trajectory(0)
    assert np.allclose(tr_dev.bath, tr.bath, rtol=1e-9, atol=0)


def test_nose_hoover_runs_in_segments_equal_one_run():
    s1 = B.argon(6, seed=5, T=60.0)
    s2 = s1.copy()
    e1 = batch.BrickEngine([s1], [_lj(dp)], dp.NoseHoover(120.0, 0.2e-12), DT)
    e1.run(30)
    e1.fetch()
    e2 = batch.BrickEngine([s2], [_lj(dp)], dp.NoseHoover(120.0, 0.2e-12), DT)
    e2.run(12)
    e2.run(18)
    e2.fetch()
    assert np.array_equal(s1.pos, s2.pos) and np.array_equal(s1.vel, s2.vel)
    assert np.array_equal(e1.thermostat_state(), e2.thermostat_state())


def test_nose_hoover_reaches_target_and_conserves_extended_energy():
    # tests/test_nose_hoover.py:20-36 on the 864-atom argon box
    s = B.argon(6, seed=3, T=60.0)
    eng = batch.BrickEngine([s], [_lj(dp)], dp.NoseHoover(120.0, 0.2e-12), 5e-15)
    eng.run(4000, record_every=10)
    eng.fetch()
    lg = eng.logs[0]
    half = len(lg) // 2
    T = np.mean(2 * lg[half:, 1] / (3 * (s.n - 1) * A.K_B))
    assert abs(T / 120.0 - 1) < 0.05
    E = lg[:, 0] + lg[:, 1]
    scale = max(abs(E[0]), np.max(np.abs(lg[:, 1])), np.max(np.abs(lg[:, 0])))
    assert np.max(np.abs(E + lg[:, 9] - E[0] - lg[0, 9])) / scale < 1e-3
    rep = dp.accuracy.drift_rows(eng.trajectory(0), s, _lj(dp), dp.NoseHoover(120.0, 0.2e-12))
    assert {c.name: c for c in rep}["extended energy"].value < 1e-3


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 5e-5)])
def test_langevin_matches_oracle_with_the_same_noise(precision, tol):
    s = B.argon(6, seed=64)
    so = s.copy()
    eng = batch.BrickEngine([s], [_lj(dp)], dp.Langevin(94.4, 5e12, seed=21), DT, precision=precision)
    eng.run(30)
    eng.fetch()
    _, tr = _oracle(so, A.Langevin(94.4, 5e12), 30, seed=21)
    assert _close(eng.logs[0][:, 0], tr.potential, tol)
    assert _close(eng.logs[0][:, 1], tr.kinetic, tol)
    assert _close(s.pos, so.pos, tol)
    assert _close(s.vel, so.vel, 10 * tol)


def test_brownian_matches_oracle_with_the_same_noise():
    s = B.argon(6, seed=64)
    D = 2.4e-9
    s.species[0] = dataclasses.replace(s.species[0], diffusion=D)
    so = s.copy()
    eng = batch.BrickEngine([s], [_lj(dp)], dp.Brownian(94.4, seed=9), 1e-15)
    eng.run(20)
    eng.fetch()
    sim = A.Simulation(so, _lj(A), integrator=A.Brownian(94.4), dt=1e-15)
    sim.rng = A.PhiloxNormals(9)
    tr = sim.run(steps=20, positions=False)
    assert _close(s.pos, so.pos, 1e-10)
    assert _close(s.vel, so.vel, 1e-8)
    assert _close(eng.logs[0][:, 0], tr.potential, 1e-9)


def test_langevin_reaches_target_temperature_and_graph_replays_match():
    s = B.argon(6, seed=2, T=40.0)
    s2 = s.copy()
    eng = batch.BrickEngine([s], [_lj(dp)], dp.Langevin(120.0, 5e12, seed=1), 5e-15)
    eng.run(1500, record_every=10)
    eng.fetch()
    lg = eng.logs[0]
    T = np.mean(2 * lg[len(lg) // 2:, 1] / (3 * s.n * A.K_B))
    assert T == pytest.approx(120.0, rel=0.05)
    # the same 1500 steps as replays of a captured 50-step block
    e2 = batch.BrickEngine([s2], [_lj(dp)], dp.Langevin(120.0, 5e12, seed=1), 5e-15)
    for _ in range(30):
        e2.run(50, record_every=10, use_graph=True)
    e2.fetch()
    assert np.array_equal(s.pos, s2.pos) and np.array_equal(s.vel, s2.vel)
