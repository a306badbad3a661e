"""GPU parity of the device-side trajectory: the brick engine's records (energies,
momentum, angular momentum summed in float64 on the device) and snapshot frames against
the oracle's Simulation.record (simulation.py:76-93), the frames against the records
themselves, graph replays, the ring's capacity, and the .npz round trip (io.py:13-43)."""
import numpy as np
import pytest

from oracle import accusim_np as A
from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")
from devinproj_amd import batch  # noqa: E402
from devinproj_amd import io as dio  # noqa: E402

DT = 10e-15


def _lj(precision="fp64"):
    return [dp.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3, precision=precision)]


def _oracle_run(s, steps, record_every):
    sim = A.Simulation(s, [A.LennardJones(B.argon_params())], integrator=A.VelocityVerlet(), dt=DT)
    return sim.run(steps=steps, record_every=record_every, positions=True)


def test_device_trajectory_matches_oracle_records_and_frames():
    s = B.argon(6, seed=64)
    so = s.copy()
    eng = batch.BrickEngine([s], [_lj()], dp.VelocityVerlet(), DT, snapshots=8)
    eng.run(20, record_every=5)
    eng.fetch()
    tr = eng.trajectory(0)
    ref = _oracle_run(so, 20, 5)
    assert len(tr.t) == len(ref.t) == 5
    assert np.allclose(tr.t, ref.t, rtol=1e-12, atol=0)
    assert np.allclose(tr.potential, ref.potential, rtol=1e-10, atol=0)
    assert np.allclose(tr.kinetic, ref.kinetic, rtol=1e-10, atol=0)
    pscale = float(np.sum(so.masses[:, None] * np.abs(so.vel)))
    assert np.max(np.abs(np.array(tr.momentum) - np.array(ref.momentum))) <= 1e-12 * pscale
    lscale = float(np.sum(np.linalg.norm(np.cross(so.pos, so.masses[:, None] * so.vel), axis=1)))
    assert np.max(np.abs(np.array(tr.angular_momentum) - np.array(ref.angular_momentum))) <= 1e-10 * lscale
    assert tr.frame_records == [0, 1, 2, 3, 4]
    assert len(tr.positions) == len(ref.positions) == 5
    L = float(so.box[0])
    for f in range(5):
        assert np.max(np.abs(tr.positions[f] - ref.positions[f])) <= 1e-10 * L
        vs = np.max(np.abs(ref.velocities[f]))
        assert np.max(np.abs(tr.velocities[f] - ref.velocities[f])) <= 1e-9 * vs
        assert np.array_equal(tr.species[f], so.sp)
    assert tr.mass == [float(np.sum(so.masses))] * 5
    assert tr.counts[0] == so.counts()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_reductions_equal_the_frames_they_describe(precision):
    """Kinetic energy, momentum and angular momentum of each record, recomputed on the
    host from that record's frame (system.py:106-114), equal the device's float64 sums
    up to summation order."""
    s = B.argon(8, seed=5, jitter=0.02)
    eng = batch.BrickEngine([s], [_lj(precision)], dp.Yoshida4(), DT, precision=precision, snapshots=4,
                            snapshot_every=2)
    eng.run(12, record_every=2)
    eng.fetch()
    tr = eng.trajectory(0)
    assert tr.frame_records == [0, 2, 4, 6]
    m = s.masses
    for f, r in enumerate(tr.frame_records):
        x, v = tr.positions[f], tr.velocities[f]
        assert tr.kinetic[r] == pytest.approx(0.5 * float(np.sum(m[:, None] * v ** 2)), rel=1e-12)
        P = np.sum(m[:, None] * v, axis=0)
        scale = float(np.sum(m[:, None] * np.abs(v)))
        assert np.max(np.abs(tr.momentum[r] - P)) <= 1e-13 * scale
        Lh = np.sum(np.cross(x, m[:, None] * v), axis=0)
        lscale = float(np.sum(np.linalg.norm(np.cross(x, m[:, None] * v), axis=1)))
        assert np.max(np.abs(tr.angular_momentum[r] - Lh)) <= 1e-13 * lscale
        assert np.all((x >= 0) & (x < s.box))


def test_graph_replays_take_the_same_frames():
    s1, s2 = B.argon(6, seed=9), B.argon(6, seed=9)
    e1 = batch.BrickEngine([s1], [_lj()], dp.VelocityVerlet(), DT, snapshots=6, snapshot_every=2)
    e2 = batch.BrickEngine([s2], [_lj()], dp.VelocityVerlet(), DT, snapshots=6, snapshot_every=2)
    for _ in range(4):
        e1.run(3, record_every=3)
        e2.run(3, record_every=3, use_graph=True)
    e1.fetch()
    e2.fetch()
    t1, t2 = e1.trajectory(0), e2.trajectory(0)
    assert t1.frame_records == t2.frame_records == [0, 2, 4]
    assert np.array_equal(e1.logs[0], e2.logs[0])
    for a, b in zip(t1.positions + t1.velocities, t2.positions + t2.velocities):
        assert np.array_equal(a, b)


def test_ring_keeps_the_first_frames_and_counts_the_rest():
    s = B.argon(5, seed=1)
    eng = batch.BrickEngine([s], [_lj()], dp.VelocityVerlet(), DT, snapshots=2)
    eng.run(4, record_every=1)
    eng.fetch()
    pos, vel, rec = eng.snapshot_frames()
    assert list(rec) == [0, 1]
    tr = eng.trajectory(0)
    assert len(tr.t) == 5 and len(tr.positions) == 2


def test_batched_systems_get_their_own_frames():
    systems = [B.lj_reduced(n, 0.8, seed=k) for k, n in enumerate([500, 864])]
    mk = lambda: [dp.LennardJones({"A": (1.0, 1.0)}, cutoff=2.5, skin=0.3)]
    eng = batch.BrickEngine([x.copy() for x in systems], [mk(), mk()], dp.VelocityVerlet(), [0.004, 0.002],
                            snapshots=3)
    eng.run(6, record_every=3)
    eng.fetch()
    for k, s in enumerate(systems):
        single = batch.BrickEngine([s.copy()], [mk()], dp.VelocityVerlet(), [0.004, 0.002][k], snapshots=3)
        single.run(6, record_every=3)
        single.fetch()
        a, b = eng.trajectory(k), single.trajectory(0)
        assert np.allclose(a.t, [0.0, 3 * [0.004, 0.002][k], 6 * [0.004, 0.002][k]], rtol=1e-12)
        assert a.frame_records == b.frame_records == [0, 1, 2]
        for x, y in zip(a.positions + a.velocities, b.positions + b.velocities):
            assert np.array_equal(x, y)


def test_device_trajectory_round_trips_through_npz(tmp_path):
    s = B.argon(5, seed=2)
    eng = batch.BrickEngine([s], [_lj()], dp.VelocityVerlet(), DT, snapshots=3)
    eng.run(10, record_every=5)
    eng.fetch()
    tr = eng.trajectory(0)
    dio.save(tr, s, tmp_path / "run.npz")
    tr2, meta = dio.load(tmp_path / "run.npz")
    assert meta[0]["name"] == s.species[0].name and tr2.counts == tr.counts
    assert np.array_equal(tr2.total_energy, tr.total_energy)
    assert np.array_equal(np.array(tr2.angular_momentum), np.array(tr.angular_momentum))
    assert all(np.array_equal(a, b) for a, b in zip(tr.positions, tr2.positions))
    assert all(np.array_equal(a, b) for a, b in zip(tr.velocities, tr2.velocities))
