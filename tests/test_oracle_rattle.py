"""The oracle's Rattle (integrators.py:119-234) against the reference's own known answers
(tests/test_rattle.py:6-44) and a closed form."""
import numpy as np

from oracle import accusim_np as A


def _drift(sim):
    return A.drift_checks(sim)


def test_rigid_rotor_keeps_length_energy_and_angular_momentum():
    # tests/test_rattle.py:6-18
    s = A.System(3)
    s.add_species("A", 1.0)
    s.add_species("B", 3.0)
    a = s.add("A", [[0.0, 0.0, 0.0]], vel=[[0.0, 3.0, 0.0]])[0]
    b = s.add("B", [[1.0, 0.0, 0.0]], vel=[[0.0, -1.0, 0.0]])[0]
    rat = A.Rattle([(a, b, 1.0)])
    sim = A.Simulation(s, [], integrator=rat, dt=0.01)
    sim.run(steps=2000, positions=False)
    d = _drift(sim)
    assert rat.max_violation < 1e-10
    assert d["energy"][1] < 1e-9
    assert d["angular momentum"][1] < 1e-9


def test_rigid_rotor_rotates_at_the_exact_rate():
    """Closed form: masses 1 and 3 on a unit bond, relative speed 4 perpendicular to the
    bond, so the bond turns at omega = 4 about the fixed centre of mass. Each step drifts the
    bond by omega dt sideways and SHAKE restores its length along the old bond, so the bond
    turns by exactly asin(omega dt) per step (RATTLE keeps the speed)."""
    s = A.System(3)
    s.add_species("A", 1.0)
    s.add_species("B", 3.0)
    a = s.add("A", [[0.0, 0.0, 0.0]], vel=[[0.0, 3.0, 0.0]])[0]
    b = s.add("B", [[1.0, 0.0, 0.0]], vel=[[0.0, -1.0, 0.0]])[0]
    sim = A.Simulation(s, [], integrator=A.Rattle([(a, b, 1.0)]), dt=1e-3)
    sim.run(steps=1000, positions=False)
    r = s.pos[1] - s.pos[0]
    ang = np.arctan2(r[1], r[0])
    expected = -1000 * np.arcsin(4.0 * 1e-3)
    assert abs((ang - expected + np.pi) % (2 * np.pi) - np.pi) < 1e-12
    assert np.linalg.norm(r) == __import__("pytest").approx(1.0, abs=1e-12)


def _collision(dt):
    # tests/test_rattle.py:21-31
    s = A.System(3)
    s.add_species("A", 1.0)
    s.add_species("C", 1.0)
    a, b = s.add("A", [[0.0, -0.5, 0.0], [0.0, 0.5, 0.0]])
    s.add("C", [[3.0, 0.3, 0.0]], vel=[[-1.0, 0.0, 0.0]])
    lj = A.LennardJones({"A": (1.0, 1.0), "C": (1.0, 1.0)}, cutoff=2.5)
    rat = A.Rattle([(a, b, 1.0)])
    sim = A.Simulation(s, [lj], integrator=rat, dt=dt)
    sim.run(steps=int(round(6 / dt)), positions=False)
    return s, rat, _drift(sim)


def test_rigid_molecule_collision_conserves_energy_and_spins_up():
    # tests/test_rattle.py:34-44
    s, rat, d = _collision(1e-3)
    assert rat.max_violation < 1e-10
    assert d["energy"][1] < 1e-4
    assert d["momentum"][1] < 1e-12
    w = np.cross(s.pos[0] - s.pos[1], s.vel[0] - s.vel[1])
    assert np.linalg.norm(w) > 1e-3
    _, _, fine = _collision(5e-4)
    assert d["energy"][1] / fine["energy"][1] > 3
