"""Pins for the CPU oracle: golden vectors written from the reference, the
reference's own known-answer tests, closed forms and brute force."""
import numpy as np
import pytest

from oracle import accusim_np as A
from oracle import builders as B


# ---------------------------------------------------------------- golden vectors

def test_plummer_sampler_matches_reference(golden):
    g = golden("plummer_n64_seed1.npz")
    s, a, M = B.plummer(n=64, seed=1)
    assert np.array_equal(s.pos, g["pos"]) and np.array_equal(s.vel, g["vel"])
    assert float(g["a"]) == a and float(g["M"]) == M


@pytest.mark.parametrize("theta", [0.5, 0.3])
def test_barnes_hut_matches_reference(golden, theta):
    g = golden("bh_theta05_n300.npz")
    s, a, M = B.plummer(n=300, seed=1)
    assert np.array_equal(s.pos, g["pos"])
    F, U = A.BarnesHut(softening=0.01 * a, theta=theta).compute(s)
    Fg = g[f"F_{theta}"]
    assert np.max(np.abs(F - Fg)) <= 1e-13 * np.max(np.abs(Fg))
    assert U == pytest.approx(float(g[f"U_{theta}"]), rel=1e-13)
    F0, U0 = A.Gravity(softening=0.01 * a).compute(s)
    assert np.max(np.abs(F0 - g["F_direct"])) <= 1e-13 * np.max(np.abs(F0))


def test_argon_trajectory_matches_reference(golden):
    g = golden("argon864_vv20.npz")
    s = B.argon(6, seed=64)
    assert np.array_equal(s.pos, g["pos0"]) and np.array_equal(s.vel, g["vel0"])
    lj = A.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3)
    sim = A.Simulation(s, [lj], integrator=A.VelocityVerlet(), dt=10e-15)
    assert np.max(np.abs(sim.force - g["F0"])) <= 1e-13 * np.max(np.abs(g["F0"]))
    tr = sim.run(steps=20, positions=False)
    assert np.allclose(tr.potential, g["pe"], rtol=1e-12, atol=0)
    assert np.allclose(tr.kinetic, g["ke"], rtol=1e-12, atol=0)
    assert np.max(np.abs(s.pos - g["pos"])) <= 1e-12 * float(g["box"][0])
    assert lj.rebuilds == int(g["rebuilds"])


def test_richardson_matches_reference(golden):
    g = golden("richardson_argon108.npz")
    s = B.argon(3, seed=5)
    assert np.array_equal(s.pos, g["pos0"]) and np.array_equal(s.vel, g["vel0"])
    sim = A.Simulation(s, [A.LennardJones(B.argon_params(), cutoff=2.5, pairlist_threshold=10 ** 9)],
                       integrator=A.VelocityVerlet(), dt=10e-15)
    sim.run(steps=5, positions=False)
    status, est, p = A.timestep_convergence(sim, 50 * 10e-15)
    assert status == str(g["status"])
    assert est == pytest.approx(float(g["value"]), rel=1e-6)
    assert f"observed order {p:.1f}" in str(g["detail"])


# ---------------------------------------------------------------- cell / Verlet lists

def _random_system(n, L, periodic, seed=0, nsp=1):
    rng = np.random.default_rng(seed)
    s = A.System(3, box=L if periodic else None, periodic=periodic)
    for k in range(nsp):
        s.add_species(f"S{k}", 1.0 + k)
    pos = rng.uniform(0, L, (n, 3))
    for k in range(nsp):
        s.add(f"S{k}", pos[k::nsp])
    return s


def _brute_pairs(s, rc):
    d = s.displacements()
    r2 = np.einsum("ijk,ijk->ij", d, d)
    i, j = np.nonzero(np.triu(r2 < rc ** 2, 1))
    return set(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize("periodic,n,L,rc", [(True, 600, 10.0, 1.3), (False, 500, 8.0, 1.1),
                                             (True, 200, 3.0, 1.2), (True, 300, 5.0, 2.4),
                                             (False, 50, 1.0, 3.0)])
def test_cell_list_equals_brute_force(periodic, n, L, rc):
    s = _random_system(n, L, periodic)
    i, j = A.cell_list_pairs(s, rc)
    assert np.all(i < j)
    assert set(zip(i.tolist(), j.tolist())) == _brute_pairs(s, rc)
    assert len(set(zip(i.tolist(), j.tolist()))) == len(i)


def test_verlet_list_reuse_rule():
    s = _random_system(500, 8.0, True, seed=2)
    vl = A.VerletList()
    vl.pairs(s, 1.0, 0.4)
    assert vl.rebuilds == 1
    s.pos[3, 0] += 0.199
    vl.pairs(s, 1.0, 0.4)
    assert vl.rebuilds == 1
    s.pos[3, 0] += 0.002
    vl.pairs(s, 1.0, 0.4)
    assert vl.rebuilds == 2
    s.ids = s.ids + 1
    vl.pairs(s, 1.0, 0.4)
    assert vl.rebuilds == 3


# ---------------------------------------------------------------- Lennard-Jones

def test_lj_two_body_closed_form():
    eps, sig = 1.7e-21, 3.3e-10
    for r in (0.95 * sig, 1.3 * sig, 2.4 * sig, 2.6 * sig):
        s = A.System(3)
        s.add_species("A", 1.0)
        s.add("A", [[0, 0, 0], [r, 0, 0]])
        F, U = A.LennardJones({"A": (eps, sig)}, cutoff=2.5).compute(s)
        if r < 2.5 * sig:
            x = (sig / r) ** 6
            xc = (1 / 2.5) ** 6
            assert U == pytest.approx(4 * eps * (x * x - x - (xc * xc - xc)), rel=1e-13)
            fr = 24 * eps * (2 * x * x - x) / r      # repulsive > 0 pushes atom 1 along +x
            assert F[1, 0] == pytest.approx(fr, rel=1e-13) and F[0, 0] == pytest.approx(-fr, rel=1e-13)
        else:
            assert U == 0.0 and np.all(F == 0)


def test_lj_minimum_at_two_to_the_sixth_sigma():
    s = A.System(3)
    s.add_species("Ar", B.AR_MASS)
    sigma = 3.4e-10
    s.add("Ar", [[0, 0, 0], [2 ** (1 / 6) * sigma, 0, 0]])
    F, _ = A.LennardJones({"Ar": (1e-21, sigma)}, cutoff=10).compute(s)
    assert np.allclose(F, 0, atol=1e-20)


def _mixed_system(periodic):
    # tests/test_forces.py:8-16
    rng = np.random.default_rng(0)
    s = A.System(3, box=2e-9 if periodic else None, periodic=periodic)
    s.add_species("A", 40 * A.AMU, A.E_CHARGE)
    s.add_species("B", 20 * A.AMU, -A.E_CHARGE)
    pts = rng.uniform(0.2e-9, 1.8e-9, size=(6, 3))
    s.add("A", pts[:3])
    s.add("B", pts[3:])
    return s


LJ_MIX = {"A": (1e3 / 6.022e23, 3.4e-10), "B": (0.5e3 / 6.022e23, 2.8e-10)}
FORCES = [lambda: A.Gravity(), lambda: A.LennardJones(LJ_MIX, cutoff=5),
          lambda: A.Ewald(), lambda: A.DampedShiftedCoulomb(rcut=0.9e-9)]


@pytest.mark.parametrize("make", FORCES)
@pytest.mark.parametrize("periodic", [False, True])
def test_force_is_negative_gradient(make, periodic):
    # tests/test_forces.py:30-45
    s = _mixed_system(periodic)
    f = make()
    if isinstance(f, A.Ewald) and not periodic:
        with pytest.raises(ValueError):
            f.compute(s)
        return
    F, _ = f.compute(s)
    h = 1e-15
    for i in range(s.n):
        for k in range(3):
            s.pos[i, k] += h
            up = f.compute(s)[1]
            s.pos[i, k] -= 2 * h
            down = f.compute(s)[1]
            s.pos[i, k] += h
            assert -(up - down) / (2 * h) == pytest.approx(F[i, k], rel=1e-4, abs=1e-6 * np.max(np.abs(F)) + 1e-300)


@pytest.mark.parametrize("periodic", [False, True])
def test_lj_cell_list_matches_all_pairs(periodic):
    # tests/test_forces.py:64-76
    rng = np.random.default_rng(5)
    s = A.System(3, box=4e-9, periodic=periodic)
    s.add_species("Ar", 39.948 * A.AMU)
    s.add_species("Kr", 83.8 * A.AMU)
    s.add("Ar", rng.uniform(0, 4e-9, (300, 3)))
    s.add("Kr", rng.uniform(0, 4e-9, (200, 3)))
    params = {"Ar": (1.65e-21, 3.4e-10), "Kr": (2.3e-21, 3.6e-10)}
    F1, U1 = A.LennardJones(params, pairlist_threshold=10 ** 9).compute(s)
    F2, U2 = A.LennardJones(params, pairlist_threshold=2).compute(s)
    assert U2 == pytest.approx(U1, rel=1e-10)
    assert np.allclose(F2, F1, rtol=1e-8, atol=1e-12 * np.max(np.abs(F1)))


def test_lj_verlet_skin_matches_fresh_lists():
    s = B.argon(5, seed=3)
    lj = A.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3)
    fresh = A.LennardJones(B.argon_params(), cutoff=2.5, skin=0.0)
    sim = A.Simulation(s, [lj], integrator=A.VelocityVerlet(), dt=10e-15)
    for _ in range(5):
        sim.run(steps=20, positions=False)
        f1, u1 = lj.compute(s)
        f0, u0 = fresh.compute(s)
        assert np.max(np.abs(f1 - f0)) <= 1e-12 * np.max(np.abs(f0))
        assert abs(u1 - u0) <= 1e-12 * abs(u0)
    assert lj.rebuilds < 0.5 * 105


# ---------------------------------------------------------------- DSF / Ewald

@pytest.mark.parametrize("n_cells", [1, 2])
def test_ewald_nacl_madelung(n_cells):
    s = B.nacl_cells(n_cells)
    F, U = A.Ewald(k=1.0).compute(s)
    assert -U / (s.n / 2) == pytest.approx(1.747564594633, rel=1e-7)
    assert np.allclose(F, 0, atol=1e-6)


def test_ewald_kcut_orthorhombic_madelung():
    """A 2 x 2 x 4 box (the 1 x 1 x 2 supercell of the conventional NaCl cell) with the
    |k| <= kcut ellipsoid: the same Madelung constant and zero lattice forces; kcut equal to
    the cubic test's 2 pi 12 / 2 takes n_z up to 24 along the long axis."""
    c = B.nacl_cells(1)
    s = A.System(3, box=np.array([2.0, 2.0, 4.0]), periodic=True)
    s.add_species("Na", 1.0, 1.0)
    s.add_species("Cl", 1.0, -1.0)
    for shift in (0.0, 2.0):
        for p, sp in zip(c.pos, c.sp):
            s.add("Na" if sp == 0 else "Cl", [[p[0], p[1], p[2] + shift]])
    assert s.n == 16 and abs(float(np.sum(s.charges))) == 0.0
    F, U = A.Ewald(k=1.0, kcut=2 * np.pi * 12 / 2.0).compute(s)
    assert -U / (s.n / 2) == pytest.approx(1.747564594633, rel=1e-7)
    assert np.allclose(F, 0, atol=1e-6)
    # the pin resolves the k set: the reference's |n| <= 12 sphere reaches only half the
    # |k| along the long axis and misses the constant by far more than 1e-7
    _, Un = A.Ewald(k=1.0, kmax=12).compute(s)
    assert abs(-Un / (s.n / 2) / 1.747564594633 - 1) > 1e-5


def test_ewald_chunked_and_kcut_agree_with_dense():
    s = B.rocksalt_unit(4)
    rng = np.random.default_rng(2)
    s.pos = (s.pos + rng.normal(0, 0.25, s.pos.shape)) % s.box
    F0, U0 = A.Ewald(k=1.0).compute(s)
    F1, U1 = A.Ewald(k=1.0, chunk=13).compute(s)
    assert U1 == pytest.approx(U0, rel=1e-12) and np.allclose(F1, F0, rtol=1e-10, atol=1e-12)
    # kcut = 2 pi kmax / L selects the same sphere in a cubic box
    F2, U2 = A.Ewald(k=1.0, kcut=2 * np.pi * 12 / 4.0).compute(s)
    assert U2 == pytest.approx(U0, rel=1e-13) and np.allclose(F2, F0, rtol=1e-12, atol=1e-13)
    # the converged sum does not depend on alpha
    F3, U3 = A.Ewald(k=1.0, alpha=1.6, kmax=14).compute(s)
    assert U3 == pytest.approx(U0, rel=1e-6)


def test_dsf_madelung_close_to_ewald():
    s = B.rocksalt_unit(8)
    _, U = A.DampedShiftedCoulomb(rcut=3.9, alpha=0.6, k=1.0).compute(s)
    assert abs(-2 * U / s.n / 1.747564594633 - 1) < 2e-3


def test_dsf_gradient_and_vanishing_at_cutoff():
    # tests/test_dsf.py:24-43
    s = B.rocksalt_unit(4)
    rng = np.random.default_rng(1)
    s.pos += rng.normal(0, 0.1, s.pos.shape)
    dsf = A.DampedShiftedCoulomb(rcut=1.9, k=1.0)
    f, _ = dsf.compute(s)
    h = 1e-6
    for i, ax in [(0, 0), (5, 1), (17, 2)]:
        s.pos[i, ax] += h
        _, up = dsf.compute(s)
        s.pos[i, ax] -= 2 * h
        _, um = dsf.compute(s)
        s.pos[i, ax] += h
        assert abs(-(up - um) / (2 * h) - f[i, ax]) < 1e-6 * max(1, abs(f[i, ax]))
    assert np.allclose(f.sum(0), 0, atol=1e-10)
    two = A.System(3)
    two.add_species("A", 1.0, 1.0)
    two.add("A", [[0, 0, 0], [1.9 - 1e-9, 0, 0]])
    f2, u2 = dsf.compute(two)
    self_term = 2 * (float(A._erfc(np.array(dsf.alpha * 1.9))) / 3.8 + dsf.alpha / np.sqrt(np.pi))
    assert abs(f2[0, 0]) < 1e-7 and abs(u2 + self_term) < 1e-7


def test_dsf_vs_ewald_on_disordered_melt():
    s = B.rocksalt_unit(4)
    rng = np.random.default_rng(2)
    s.pos = (s.pos + rng.normal(0, 0.25, s.pos.shape)) % s.box
    _, ue = A.Ewald(k=1.0).compute(s)
    _, ud = A.DampedShiftedCoulomb(rcut=1.99, alpha=0.9, k=1.0).compute(s)
    assert abs(ud / ue - 1) < 3e-2


def test_dsf_cell_matches_dense_with_exclusions():
    s = B.rocksalt_unit(8)
    rng = np.random.default_rng(5)
    s.pos = (s.pos + rng.normal(0, 0.2, s.pos.shape)) % s.box
    ex = [(int(s.ids[2 * m]), int(s.ids[2 * m + 1])) for m in range(50)]
    fd, ud = A.DampedShiftedCoulomb(rcut=2.5, k=1.0, pairlist_threshold=10 ** 9, exclude=ex).compute(s)
    fc, uc = A.DampedShiftedCoulomb(rcut=2.5, k=1.0, pairlist_threshold=1, exclude=ex).compute(s)
    _, u0 = A.DampedShiftedCoulomb(rcut=2.5, k=1.0).compute(s)
    assert abs(uc / ud - 1) < 1e-12 and np.allclose(fc, fd, rtol=1e-9, atol=1e-12) and abs(u0 - ud) > 1e-6


def test_dsf_verlet_skin_matches_fresh_lists():
    # tests/test_dsf.py:119-139
    s = B.rocksalt_unit(8)
    rng = np.random.default_rng(5)
    skinned = A.DampedShiftedCoulomb(rcut=3.9, alpha=0.6, k=1.0, skin=0.6)
    for _ in range(30):
        s.pos = (s.pos + rng.normal(0, 0.05, s.pos.shape)) % s.box
        f1, u1 = skinned.compute(s)
        f0, u0 = A.DampedShiftedCoulomb(rcut=3.9, alpha=0.6, k=1.0, skin=0.0).compute(s)
        assert np.max(np.abs(f1 - f0)) <= 1e-12 * np.max(np.abs(f0))
        assert abs(u1 - u0) <= 1e-12 * abs(u0)
    assert 1 < skinned._vl.rebuilds < 15


# ---------------------------------------------------------------- gravity / Barnes-Hut

def test_gravity_two_body_and_chunked():
    s = A.System(3)
    s.add_species("a", 2.0)
    s.add_species("b", 3.0)
    s.add("a", [[0, 0, 0]])
    s.add("b", [[0, 4.0, 0]])
    F, U = A.Gravity(G=1.0, softening=3.0).compute(s)
    assert U == pytest.approx(-6.0 / 5.0, rel=1e-15)
    assert F[0, 1] == pytest.approx(6.0 * 4.0 / 125.0, rel=1e-15) and F[1, 1] == pytest.approx(-F[0, 1])
    s, a, M = B.plummer(n=257, seed=4)
    F0, U0 = A.Gravity(softening=0.01 * a).compute(s)
    F1, U1 = A.Gravity(softening=0.01 * a, chunk=50).compute(s)
    assert U1 == pytest.approx(U0, rel=1e-13) and np.allclose(F1, F0, rtol=1e-12, atol=1e-14 * np.abs(F0).max())


def test_barnes_hut_theta_zero_is_exact():
    s, a, M = B.plummer(n=300, seed=1)
    F0, U0 = A.Gravity(softening=0.01 * a).compute(s)
    F1, U1 = A.BarnesHut(softening=0.01 * a, theta=0.0).compute(s)
    assert A.rms_rel_force_error(F1, F0) < 1e-12 and abs(U1 - U0) / abs(U0) < 1e-12


@pytest.mark.parametrize("theta,tol", [(0.3, 2e-3), (0.6, 1e-2)])
def test_barnes_hut_error_controlled_by_theta(theta, tol):
    s, a, M = B.plummer(n=1500, seed=1)
    F0, U0 = A.Gravity(softening=0.01 * a).compute(s)
    F1, U1 = A.BarnesHut(softening=0.01 * a, theta=theta).compute(s)
    assert A.rms_rel_force_error(F1, F0) < tol and abs(U1 - U0) / abs(U0) < tol


# ---------------------------------------------------------------- integrators / report

def _kepler(integ, steps_per_orbit, e=0.0):
    # tests/test_dynamics.py:7-19
    s = A.System(3)
    ms, me = A.MSUN, 5.9722e24
    s.add_species("Sun", ms)
    s.add_species("Earth", me)
    GM = A.G_SI * (ms + me)
    au = 1.495978707e11
    r_p = au * (1 - e)
    v_p = np.sqrt(GM * (1 + e) / r_p)
    mu = me / (ms + me)
    s.add("Sun", [-mu * r_p, 0, 0], [0, -mu * v_p, 0])
    s.add("Earth", [(1 - mu) * r_p, 0, 0], [0, (1 - mu) * v_p, 0])
    period = 2 * np.pi * np.sqrt(au ** 3 / GM)
    return A.Simulation(s, [A.Gravity()], integrator=integ, dt=period / steps_per_orbit), period


@pytest.mark.parametrize("integ,tol", [(A.VelocityVerlet(), 2e-3), (A.Yoshida4(), 1e-6)])
def test_circular_orbit_closes(integ, tol):
    sim, period = _kepler(integ, 400)
    start = sim.system.pos[1].copy()
    sim.run(steps=400)
    assert np.linalg.norm(sim.system.pos[1] - start) / np.linalg.norm(start) < tol


@pytest.mark.parametrize("integ", [A.VelocityVerlet(), A.Yoshida4()])
def test_observed_order(integ):
    sim, period = _kepler(integ, 100, e=0.3)
    sim.run(steps=100, positions=False)
    status, est, p = A.timestep_convergence(sim, period)
    assert p == pytest.approx(integ.order, abs=0.4)


def test_drift_checks_known_answer():
    sim, period = _kepler(A.Yoshida4(), 500, e=0.5)
    sim.run(steps=500 * 5, record_every=50, positions=False)
    out = A.drift_checks(sim)
    assert out["energy"][0] == "PASS" and out["momentum"][0] == "PASS" and out["angular momentum"][0] == "PASS"
    sim2, _ = _kepler(A.VelocityVerlet(), 8, e=0.9)
    sim2.run(steps=80, positions=False)
    assert A.drift_checks(sim2)["energy"][0] in ("WARN", "FAIL")
    # hand-made trajectory: E = [10, 10.5, 9.0] with |KE| max 12 -> drift 1/12
    tr = A.Trajectory(t=[0, 1, 2], kinetic=[12.0, 12.5, 11.0], potential=[-2.0, -2.0, -2.0],
                      momentum=[np.zeros(3)] * 3, angular_momentum=[np.zeros(3)] * 3, mass=[1.0] * 3,
                      charge=[0.0] * 3)
    sim2.trajectory = tr
    assert A.drift_checks(sim2)["energy"][1] == pytest.approx(1.0 / 12.5)
