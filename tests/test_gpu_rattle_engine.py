"""Rattle on the device (BrickEngine): SHAKE after each drift and RATTLE after each closing
kick in float64, one thread per constraint cluster, against the oracle's restatement of
the reference's dense M-SHAKE (integrators.py:119-234) on the same start."""
import numpy as np
import pytest

from oracle import accusim_np as A
from tests.rigid import fluid

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")
from devinproj_amd import batch  # noqa: E402
from devinproj_amd.integrators import constraint_clusters  # noqa: E402

LJ = {"A": (1.0, 0.8)}


def _oracle(s0, cons, steps, dt, tol):
    s = s0.copy()
    rat = A.Rattle(cons, tol=tol)
    sim = A.Simulation(s, [A.LennardJones(LJ, cutoff=2.5, skin=0.3)], integrator=rat, dt=dt)
    sim.run(steps=steps, positions=False)
    return s, rat, sim


def _engine(systems, cons, steps, dt, tol, record_every=1, use_graph=False):
    rat = dp.Rattle(cons, tol=tol)
    eng = batch.BrickEngine(systems, [[dp.LennardJones(LJ, cutoff=2.5, skin=0.3)] for _ in systems], rat, dt)
    eng.run(steps, record_every=record_every, use_graph=use_graph)
    eng.fetch()
    return eng, rat


def _close(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return float(np.max(np.abs(d)))


# dimers, triangles and 4-chains keep their cluster work in the LDS; 8-chains (7 bonds) in global memory
@pytest.mark.parametrize("shape,n_mol,rho", [("dimer", 256, 0.2), ("triangle", 125, 0.1), ("chain", 125, 0.06),
                                             ("chain8", 27, 0.025)])
def test_device_rattle_follows_the_oracle(shape, n_mol, rho):
    s0, cons = fluid(shape, n_mol, rho, seed=7)
    steps, dt, tol = 200, 1e-3, 1e-10
    print(f"rattle engine {shape} n_mol={n_mol} steps={steps}", flush=True)
    s_o, rat_o, sim_o = _oracle(s0, cons, steps, dt, tol)
    s = s0.copy()
    eng, rat = _engine([s], cons, steps, dt, tol)
    assert _close(s.pos, s_o.pos, s.box) < 1e-8
    assert np.max(np.abs(s.vel - s_o.vel)) < 1e-7
    assert rat.max_violation < 1e-9
    assert rat.max_violation == pytest.approx(rat_o.max_violation, rel=0.5, abs=1e-13)
    assert s.constrained_dof == len(cons)
    tr, tr_o = eng.trajectory(0), sim_o.trajectory
    ke, ke_o = np.array(tr.kinetic), np.array(tr_o.kinetic)
    pe, pe_o = np.array(tr.potential), np.array(tr_o.potential)
    assert len(ke) == len(ke_o) == steps + 1
    assert np.allclose(ke, ke_o, rtol=1e-8, atol=0) and np.allclose(pe, pe_o, rtol=1e-8, atol=1e-12)
    p, p_o = np.array(tr.momentum), np.array(tr_o.momentum)
    assert np.max(np.abs(p - p_o)) < 1e-9
    e = ke + pe
    assert np.max(np.abs(e - e[0])) / max(abs(e[0]), np.max(np.abs(ke))) < 1e-4


def test_device_rattle_two_systems_one_arena_and_graph_replays():
    """Two fluids (dimers, triangles) advance in one arena; clusters of each system map to
    its own atoms. The second half of the run replays a captured hipGraph."""
    s0, c0 = fluid("dimer", 216, 0.2, seed=11)
    s1, c1 = fluid("triangle", 125, 0.1, seed=12)
    # one constraint list for both: the ids of each system select its bonds (ids overlap, so
    # the dimer bonds are offset past the triangle fluid's ids and its ids shifted to match)
    shift = int(max(s1.ids.max(), max(max(a, b) for a, b, _ in c1))) + 1
    s0 = s0.copy()
    s0.ids = s0.ids + shift
    c0 = [(a + shift, b + shift, d) for a, b, d in c0]
    ref = []
    for s, c in ((s0, c0), (s1, c1)):
        ref.append(_oracle(s, c, 60, 1e-3, 1e-10)[0])
    systems = [s0.copy(), s1.copy()]
    rat = dp.Rattle(c0 + c1, tol=1e-10)
    eng = batch.BrickEngine(systems, [[dp.LennardJones(LJ, cutoff=2.5, skin=0.3)] for _ in systems], rat, 1e-3)
    eng.run(30, record_every=10)
    eng.run(10, record_every=10, use_graph=True)
    eng.run(10, record_every=10, use_graph=True)
    eng.run(10, record_every=10, use_graph=True)
    eng.fetch()
    for s, r in zip(systems, ref):
        assert _close(s.pos, r.pos, s.box) < 1e-9
    assert systems[0].constrained_dof == len(c0) and systems[1].constrained_dof == len(c1)
    assert rat.max_violation < 1e-9


def test_device_rattle_failure_is_reported():
    # a tolerance below float64's reach stalls the Newton iterations of every cluster: the
    # run that hit it raises (a synchronised run checks the constraints), and so does fetch
    s0, cons = fluid("dimer", 64, 0.05, seed=2)
    rat = dp.Rattle(cons, tol=1e-30)
    eng = batch.BrickEngine([s0.copy()], [[dp.LennardJones(LJ, cutoff=2.5, skin=0.3)]], rat, 1e-3)
    with pytest.raises(RuntimeError, match="stalled"):
        eng.run(2)
    with pytest.raises(RuntimeError, match="stalled"):
        eng.fetch()
    eng.run(1, sync=False)  # an unsynchronised run leaves the check to sync / fetch
    eng.sync()


def test_device_rattle_iteration_limit_is_reported():
    s0, cons = fluid("dimer", 64, 0.05, seed=2)
    rat = dp.Rattle(cons, tol=1e-12, max_iter=1)
    eng = batch.BrickEngine([s0.copy()], [[dp.LennardJones(LJ, cutoff=2.5, skin=0.3)]], rat, 1e-2)
    with pytest.raises(RuntimeError, match="did not converge"):
        eng.run(1)
    with pytest.raises(RuntimeError, match="did not converge"):
        eng.fetch()


def test_device_rattle_failure_names_the_callers_cluster():
    # one dimer's bond given twice makes that cluster's SHAKE system singular (two equal
    # rows), and only that cluster fails; the engine orders clusters by slot, the failure
    # is reported by the cluster's index in the caller's (constraint_clusters) order
    s0, cons = fluid("dimer", 64, 0.05, seed=2)
    bad = 37
    cons2 = list(cons) + [cons[bad]]
    eng = batch.BrickEngine([s0.copy()], [[dp.LennardJones(LJ, cutoff=2.5, skin=0.3)]],
                            dp.Rattle(cons2, tol=1e-10), 1e-3)
    with pytest.raises(RuntimeError, match="stalled"):
        eng.run(1)
    where = {int(p): n for n, p in enumerate(s0.ids)}
    ia = [where[a] for a, b, d in cons2]
    ib = [where[b] for a, b, d in cons2]
    groups = constraint_clusters(ia, ib)
    expect = next(k for k, g in enumerate(groups) if len(cons2) - 1 in g)
    assert eng.failed_cluster == expect
    assert bad in groups[expect]


def test_device_rattle_argument_checks():
    s0, cons = fluid("dimer", 64, 0.05, seed=2)
    with pytest.raises(ValueError, match="no constraint"):
        batch.BrickEngine([s0.copy()], [[dp.LennardJones(LJ, cutoff=2.5)]], dp.Rattle([(10**6, 10**6 + 1, 1.0)]),
                          1e-3)
    eng = batch.BrickEngine([s0.copy()], [[dp.LennardJones(LJ, cutoff=2.5)]], dp.Rattle(cons), 1e-3,
                            snapshots=4)
    with pytest.raises(ValueError, match="snapshot"):
        eng.run(2)


def test_device_rattle_with_mixed_precision_forces():
    """float32 pair forces (float64 sums), SHAKE / RATTLE still float64: bonds hold to the
    same tolerance and the trajectory stays within the force rounding of the oracle's."""
    s0, cons = fluid("dimer", 216, 0.2, seed=21)
    s_o, _, _ = _oracle(s0, cons, 200, 1e-3, 1e-10)
    s = s0.copy()
    rat = dp.Rattle(cons, tol=1e-10)
    eng = batch.BrickEngine([s], [[dp.LennardJones(LJ, cutoff=2.5, skin=0.3, precision="mixed")]], rat, 1e-3,
                            precision="mixed")
    eng.run(200, record_every=20)
    eng.fetch()
    assert rat.max_violation < 1e-9
    assert _close(s.pos, s_o.pos, s.box) < 1e-5
    e = np.array(eng.trajectory(0).kinetic) + np.array(eng.trajectory(0).potential)
    assert np.max(np.abs(e - e[0])) / abs(e[0]) < 1e-3
