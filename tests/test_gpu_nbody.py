"""The device-resident N-body engine (accu_nbody_*, devinproj_amd.NBodyEngine) against the
oracle's Simulation.run (simulation.py:95-112) with Yoshida4 / VelocityVerlet
(integrators.py:19-47) over Gravity (forces.py:31-48) and BarnesHut (tree.py:19-78).

Compared after a few steps: positions and velocities, and every record (potential and
kinetic energy, momentum, angular momentum) to 1e-10 of their scale in float64 (the force
sums differ from numpy's only in summation order); the Barnes-Hut run to 1e-9 (the tree is
the reference's, the walk's order is not). Also: graph replays equal eager runs bit for bit,
mixed precision stays within 1e-5 (RMS relative acceleration) of float64, and the
262,144-star workload conserves momentum over Yoshida4 steps."""
import numpy as np
import pytest

import report_cases as RC
from oracle import accusim_np as A
from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")


def _plummer(n, seed):
    s, a, M = B.plummer(n=n, seed=seed)
    st = {"pos": s.pos, "vel": s.vel, "sp": s.sp, "names": ["star"], "masses": s.masses[:1],
          "charges": np.zeros(1), "box": None}
    return st, a, float(np.sqrt(a ** 3 / (B.G_SI * M)))


def _oracle(st, force, integ, dt, steps):
    sim = A.Simulation(RC.system(A, st), [force], integrator=integ, dt=dt)
    sim.run(steps=steps, positions=False)
    return sim


def _close(eng, osim, tol, tag):
    eng.fetch()
    s, o = eng.system, osim.system
    assert np.max(np.abs(s.pos - o.pos)) <= tol * np.max(np.abs(o.pos)), tag
    assert np.max(np.abs(s.vel - o.vel)) <= tol * np.max(np.abs(o.vel)), tag
    tr, ot = eng.trajectory(), osim.trajectory
    assert len(tr.t) == len(ot.t), (tag, len(tr.t), len(ot.t))
    assert np.allclose(tr.t, ot.t, rtol=1e-14, atol=0), tag
    E = max(np.max(np.abs(ot.kinetic)), np.max(np.abs(ot.potential)))
    assert np.max(np.abs(np.subtract(tr.kinetic, ot.kinetic))) <= tol * E, tag
    assert np.max(np.abs(np.subtract(tr.potential, ot.potential))) <= tol * E, tag
    P = np.sum(o.masses[:, None] * np.abs(o.vel))
    assert np.max(np.abs(np.subtract(tr.momentum, ot.momentum))) <= tol * P, tag
    Lsc = np.sum(o.masses * np.linalg.norm(o.pos, axis=1) * np.linalg.norm(o.vel, axis=1))
    assert np.max(np.abs(np.subtract(tr.angular_momentum, ot.angular_momentum))) <= tol * Lsc, tag


def test_yoshida4_direct_fp64_equals_oracle():
    st, a, tc = _plummer(500, 1)
    print("nbody yoshida4 direct fp64 n=500", flush=True)
    osim = _oracle(st, A.Gravity(softening=0.01 * a), A.Yoshida4(), 1e-3 * tc, 6)
    eng = dp.NBodyEngine(RC.system(dp, st), dp.Gravity(softening=0.01 * a), dp.Yoshida4(), 1e-3 * tc)
    eng.run(4)
    eng.run(2)  # two runs: the second opens from the stored accelerations
    _close(eng, osim, 1e-10, "yoshida4 direct")
    eng.close()


def test_velocity_verlet_direct_periodic_fp64_equals_oracle():
    rng = np.random.default_rng(5)
    n, L = 400, 10.0
    st = {"pos": rng.random((n, 3)) * L, "vel": rng.normal(0, 0.05, (n, 3)), "sp": np.zeros(n, np.int64),
          "names": ["p"], "masses": np.array([1.0]), "charges": np.zeros(1), "box": np.array([L, L, L])}
    print("nbody vv direct periodic fp64 n=400", flush=True)
    eng = dp.NBodyEngine(RC.system(dp, st), dp.Gravity(G=1.0, softening=0.3), dp.VelocityVerlet(), 0.05)
    eng.run(12, record_every=3)
    osim2 = _oracle(st, A.Gravity(G=1.0, softening=0.3), A.VelocityVerlet(), 0.05, 0)
    osim2.run(steps=12, record_every=3, positions=False)
    _close(eng, osim2, 1e-10, "vv periodic")
    assert np.all(eng.system.pos >= 0) and np.all(eng.system.pos < L)
    eng.close()


def test_barnes_hut_yoshida4_equals_oracle():
    st, a, tc = _plummer(300, 2)
    print("nbody yoshida4 barnes-hut n=300", flush=True)
    osim = _oracle(st, A.BarnesHut(softening=0.01 * a, theta=0.5), A.Yoshida4(), 1e-3 * tc, 3)
    eng = dp.NBodyEngine(RC.system(dp, st), dp.BarnesHut(softening=0.01 * a, theta=0.5), dp.Yoshida4(), 1e-3 * tc)
    eng.run(3)
    _close(eng, osim, 1e-9, "yoshida4 barnes-hut")
    with pytest.raises(ValueError, match="not captured"):
        eng.run(1, use_graph=True)
    eng.close()


def test_graph_replays_equal_eager_runs():
    st, a, tc = _plummer(1000, 3)
    print("nbody graph vs eager n=1000", flush=True)
    runs = []
    for graph in (False, True):
        eng = dp.NBodyEngine(RC.system(dp, st), dp.Gravity(softening=0.01 * a), dp.Yoshida4(), 1e-3 * tc)
        for _ in range(4):
            eng.run(3, use_graph=graph)
        eng.fetch()
        runs.append((eng.system.pos.copy(), eng.system.vel.copy(), eng.log.copy()))
        eng.close()
    for x, y in zip(runs[0], runs[1]):
        assert np.array_equal(x, y)


def test_mixed_precision_close_to_fp64():
    st, a, tc = _plummer(4096, 4)
    print("nbody mixed vs fp64 n=4096", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        eng = dp.NBodyEngine(RC.system(dp, st), dp.Gravity(softening=0.01 * a, precision=prec), dp.Yoshida4(),
                             1e-3 * tc)
        eng.run(2)
        eng.fetch()
        out[prec] = (eng.acc.copy(), eng.log.copy())
        eng.close()
    a64, am = out["fp64"][0], out["mixed"][0]
    err = np.sqrt(np.mean(np.sum((am - a64) ** 2, 1)) / np.mean(np.sum(a64 ** 2, 1)))
    print(f"mixed vs fp64 rms relative acceleration {err:.2e}", flush=True)
    assert err < 1e-5
    E64 = out["fp64"][1][:, 0] + out["fp64"][1][:, 1]
    Em = out["mixed"][1][:, 0] + out["mixed"][1][:, 1]
    assert np.max(np.abs(Em - E64)) <= 1e-6 * np.max(np.abs(E64))


def test_plummer_262144_yoshida4_conserves_momentum():
    """The 262,144-star workload (Plummer sphere, softening 0.01 a, dt = 1e-3 t_cross),
    mixed-precision direct gravity: momentum stays at round-off over 3 Yoshida4 steps and
    the energy drift stays small."""
    s, a, M = B.plummer(n=262144, seed=11)
    st = {"pos": s.pos, "vel": s.vel, "sp": s.sp, "names": ["star"], "masses": s.masses[:1],
          "charges": np.zeros(1), "box": None}
    tc = float(np.sqrt(a ** 3 / (B.G_SI * M)))
    print("nbody plummer 262144 mixed yoshida4", flush=True)
    eng = dp.NBodyEngine(RC.system(dp, st), dp.Gravity(softening=0.01 * a, precision="mixed"), dp.Yoshida4(),
                         1e-3 * tc)
    eng.run(3, use_graph=True)
    eng.fetch()
    P = np.array(eng.log[:, 2:5])
    scale = np.sum(eng.system.masses[:, None] * np.abs(eng.system.vel))
    drift = eng.energy_drift()
    print(f"momentum drift {np.max(np.abs(P - P[0])) / scale:.2e}, energy drift {drift:.2e}", flush=True)
    assert np.max(np.abs(P - P[0])) / scale < 1e-9
    assert drift < 1e-6
    eng.close()
