"""Rattle as a caller of the GPU forces (the reference's integrators.py:119-234, restated
in the oracle, calls sim.update_forces(), which calls the drop-in Force.compute). The
constraint iterations stay float64 on the host; only the forces come from the GPU."""
import numpy as np
import pytest

from oracle import accusim_np as A

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")


def _collision(lj_cls, dt=1e-3):
    # tests/test_rattle.py:21-31 with the LJ force from lj_cls
    s = A.System(3)
    s.add_species("A", 1.0)
    s.add_species("C", 1.0)
    a, b = s.add("A", [[0.0, -0.5, 0.0], [0.0, 0.5, 0.0]])
    s.add("C", [[3.0, 0.3, 0.0]], vel=[[-1.0, 0.0, 0.0]])
    rat = A.Rattle([(a, b, 1.0)])
    sim = A.Simulation(s, [lj_cls({"A": (1.0, 1.0), "C": (1.0, 1.0)}, cutoff=2.5)], integrator=rat, dt=dt)
    sim.run(steps=int(round(6 / dt)), positions=False)
    return s, rat, A.drift_checks(sim)


def test_rattle_over_gpu_lj_follows_the_oracle_collision():
    s_gpu, rat_gpu, d_gpu = _collision(dp.LennardJones)
    s_ref, rat_ref, d_ref = _collision(A.LennardJones)
    assert np.max(np.abs(s_gpu.pos - s_ref.pos)) < 1e-10
    assert np.max(np.abs(s_gpu.vel - s_ref.vel)) < 1e-10
    # the reference's own assertions (tests/test_rattle.py:34-44) hold with GPU forces
    assert rat_gpu.max_violation < 1e-10
    assert d_gpu["energy"][1] < 1e-4
    assert d_gpu["momentum"][1] < 1e-12
    assert d_gpu["energy"][1] == pytest.approx(d_ref["energy"][1], rel=1e-6)


def test_rattle_over_gpu_lj_keeps_a_periodic_dimer_fluid_rigid():
    """256 rigid dimers (512 atoms, the GPU cell-list path) in a periodic box, reduced units;
    at rho = 0.2 the grid spacing (1.96) keeps atoms of neighbouring dimers apart (>= 0.96 > sigma)."""
    rng = np.random.default_rng(4)
    n_mol, rho, bond = 256, 0.2, 1.0
    L = (2 * n_mol / rho) ** (1 / 3)
    m = int(np.ceil(n_mol ** (1 / 3)))
    g = (np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n_mol] + 0.5) * (L / m)
    u = rng.normal(size=(n_mol, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    s = A.System(3, box=L, periodic=True)
    s.add_species("A", 1.0)
    va, vb = rng.normal(0, 0.5, (n_mol, 3)), rng.normal(0, 0.5, (n_mol, 3))
    along = np.sum((vb - va) * u, 1)[:, None] * u  # bond-stretching part of the relative velocity
    va, vb = va + 0.5 * along, vb - 0.5 * along     # removed (equal masses): velocities satisfy the constraints
    ia = s.add("A", (g - 0.5 * bond * u) % L, vel=va)
    ib = s.add("A", (g + 0.5 * bond * u) % L, vel=vb)
    s.vel -= s.vel.mean(0)  # the starting state; each run below copies it
    runs = []
    for lj_cls in (dp.LennardJones, A.LennardJones):
        r = A.System(3, box=L, periodic=True)
        r.add_species("A", 1.0)
        r.add("A", s.pos, vel=s.vel)
        rat = A.Rattle(list(zip(ia, ib, [bond] * n_mol)), tol=1e-10)
        sim = A.Simulation(r, [lj_cls({"A": (1.0, 0.8)}, cutoff=2.5, skin=0.3)], integrator=rat, dt=1e-3)
        sim.run(steps=300, positions=False)
        runs.append((r, rat, A.drift_checks(sim)))
    (r, rat, d), (r_ref, _, d_ref) = runs
    assert rat.max_violation < 1e-9
    assert d["momentum"][1] < 1e-10
    assert d["energy"][1] < 1e-4
    dx = r.pos - r_ref.pos
    dx -= L * np.round(dx / L)
    assert np.max(np.abs(dx)) < 1e-8
    assert d["energy"][1] == pytest.approx(d_ref["energy"][1], rel=1e-3)
