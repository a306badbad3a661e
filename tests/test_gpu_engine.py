"""GPU parity of the device-resident fused MD engine against the oracle and the
reference's golden trajectory."""
import numpy as np
import pytest

from oracle import accusim_np as A
from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")

DT = 10e-15


def test_engine_vv_matches_reference_trajectory(golden):
    g = golden("argon864_vv20.npz")
    s = B.argon(6, seed=64)
    eng = dp.MDEngine(s, [dp.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3)], dp.VelocityVerlet(), dt=DT)
    eng.run(20)
    eng.fetch()
    pe, ke, mom = eng.trajectory_arrays()
    assert len(pe) == 21
    assert np.allclose(pe, g["pe"], rtol=1e-10, atol=0)
    assert np.allclose(ke, g["ke"], rtol=1e-10, atol=0)
    assert np.max(np.abs(s.pos - g["pos"])) <= 1e-10 * float(g["box"][0])
    assert np.max(np.abs(eng.force - g["F"])) <= 1e-9 * np.max(np.abs(g["F"]))
    assert eng.rebuilds == int(g["rebuilds"])


def test_engine_graph_replay_is_bitwise_identical():
    s1, s2 = B.argon(6, seed=64), B.argon(6, seed=64)
    e1 = dp.MDEngine(s1, [dp.LennardJones(B.argon_params())], dp.VelocityVerlet(), dt=DT)
    e2 = dp.MDEngine(s2, [dp.LennardJones(B.argon_params())], dp.VelocityVerlet(), dt=DT, use_graph=True)
    for _ in range(3):
        e1.run(10)
        e2.run(10)
    e1.fetch()
    e2.fetch()
    assert np.array_equal(e1.log, e2.log)
    assert np.array_equal(s1.pos, s2.pos) and np.array_equal(s1.vel, s2.vel)


def test_engine_yoshida_matches_oracle():
    s = B.argon(5, seed=8)
    so = s.copy()
    eng = dp.MDEngine(s, [dp.LennardJones(B.argon_params())], dp.Yoshida4(), dt=DT)
    eng.run(10)
    eng.fetch()
    sim = A.Simulation(so, [A.LennardJones(B.argon_params())], integrator=A.Yoshida4(), dt=DT)
    tr = sim.run(steps=10, positions=False)
    assert np.allclose(eng.log[:, 0], tr.potential, rtol=1e-10, atol=0)
    assert np.allclose(eng.log[:, 1], tr.kinetic, rtol=1e-10, atol=0)
    assert np.max(np.abs(s.pos - so.pos)) <= 1e-10 * so.box[0]


def test_engine_lj_plus_dsf_matches_oracle():
    s = B.rocksalt(8, 8, 8, seed=5)
    rng = np.random.default_rng(1)
    s.pos = (s.pos + rng.normal(0, 0.2e-10, s.pos.shape)) % s.box
    so = s.copy()
    rc = min(12e-10, 0.45 * s.box[0])
    mk = lambda M: [M.LennardJones(B.NACL_LJ, cutoff=2.5), M.DampedShiftedCoulomb(rcut=rc, alpha=0.2e10)]
    eng = dp.MDEngine(s, mk(dp), dp.VelocityVerlet(), dt=1e-15)
    eng.run(10)
    eng.fetch()
    sim = A.Simulation(so, mk(A), integrator=A.VelocityVerlet(), dt=1e-15)
    tr = sim.run(steps=10, positions=False)
    assert np.allclose(eng.log[:, 0], tr.potential, rtol=1e-9, atol=0)
    assert np.allclose(eng.log[:, 1], tr.kinetic, rtol=1e-9, atol=0)
    assert np.max(np.abs(s.pos - so.pos)) <= 1e-10 * so.box[0]


def test_engine_mixed_precision_tracks_fp64():
    s64 = B.argon(16, seed=2, jitter=0.01)
    s32 = s64.copy()
    e64 = dp.MDEngine(s64, [dp.LennardJones(B.argon_params())], dp.VelocityVerlet(), dt=DT)
    e32 = dp.MDEngine(s32, [dp.LennardJones(B.argon_params())], dp.VelocityVerlet(), dt=DT, precision="mixed")
    e64.run(50, record_every=1)
    e32.run(50, record_every=1)
    e64.fetch()
    e32.fetch()
    # forces at step 0 agree to the float32 pair precision
    assert A.rms_rel_force_error(e32.force, e64.force) < 1e-4
    assert abs(e32.energy_drift() - e64.energy_drift()) < 1e-5
    assert np.allclose(e32.log[:, 2:5], 0, atol=1e-6 * np.sum(np.abs(s64.vel) * B.AR_MASS))


def test_engine_momentum_conserved_and_rebuilds_counted():
    s = B.argon(8, seed=4)
    eng = dp.MDEngine(s, [dp.LennardJones(B.argon_params())], dp.VelocityVerlet(), dt=DT)
    eng.run(100, record_every=10)
    eng.fetch()
    assert len(eng.log) == 11
    P = eng.log[:, 2:5]
    pscale = np.sum(B.AR_MASS * np.abs(s.vel))
    assert np.max(np.linalg.norm(P - P[0], axis=1)) / pscale < 1e-12
    assert 1 < eng.rebuilds < 30
    assert eng.stats["max_neighbours"] <= eng.stats["capacity"]
