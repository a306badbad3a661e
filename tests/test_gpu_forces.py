"""GPU parity: the HIP force path against the CPU oracle on identical seeded inputs."""
import numpy as np
import pytest

from oracle import accusim_np as A
from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")


def _random_system(n, L, periodic, seed=0, nsp=1):
    rng = np.random.default_rng(seed)
    s = A.System(3, box=L if periodic else None, periodic=periodic)
    for k in range(nsp):
        s.add_species(f"S{k}", 1.0 + k)
    pos = rng.uniform(0, L, (n, 3))
    for k in range(nsp):
        s.add(f"S{k}", pos[k::nsp])
    return s


def _pairset(i, j):
    return set(zip(np.asarray(i).tolist(), np.asarray(j).tolist()))


# ---------------------------------------------------------------- cell list / Verlet list (bit-exact)

CASES = [(True, 3000, 10.0, 1.3), (False, 2500, 8.0, 1.1), (True, 200, 3.0, 1.2), (True, 300, 5.0, 2.4),
         (False, 50, 1.0, 3.0), (True, 4000, 12.0, 0.95)]


@pytest.mark.parametrize("periodic,n,L,rc", CASES)
def test_cell_assignment_and_pairs_bit_exact(periodic, n, L, rc):
    s = _random_system(n, L, periodic, seed=n)
    lo, span, ncell, cell, flat = A.cell_grid(s, rc)
    th = dp._lib.torch()
    pos_d = th.from_numpy(s.pos.copy()).cuda()
    nl = dp.NeighbourList()
    nl.build(s, pos_d, rc)
    gcell, order, nn = nl.cells()
    assert np.array_equal(gcell, flat)
    assert np.array_equal(order, np.argsort(flat, kind="stable"))
    gi, gj = nl.half_pairs()
    oi, oj = A.cell_list_pairs(s, rc)
    assert len(gi) == len(oi)
    assert _pairset(gi, gj) == _pairset(oi, oj)
    assert np.all(gi < gj)


def test_pairs_on_jittered_fcc_argon_bit_exact():
    s = B.argon(12, seed=11, jitter=0.05)     # 6912 atoms, 5^3 cells
    rc = (2.5 + 0.3) * B.AR_SIGMA
    gi, gj = dp.cell_list_pairs(s, rc)
    oi, oj = A.cell_list_pairs(s, rc)
    assert _pairset(gi, gj) == _pairset(oi, oj)


def test_verlet_list_rebuild_decisions_match_oracle():
    s = B.argon(5, seed=3)
    rng = np.random.default_rng(9)
    rc, skin = 2.5 * B.AR_SIGMA, 0.3 * B.AR_SIGMA
    vo, vg = A.VerletList(), dp.VerletList()
    for step in range(40):
        s.pos = (s.pos + rng.normal(0, 0.02 * B.AR_SIGMA, s.pos.shape)) % s.box
        oi, oj = vo.pairs(s, rc, skin)
        gi, gj = vg.pairs(s, rc, skin)
        assert vo.rebuilds == vg.rebuilds
        assert _pairset(oi, oj) == _pairset(gi, gj)
    assert 2 < vg.rebuilds < 40


def test_verlet_list_rebuilds_when_ids_change():
    """forces.py:88-89: the list is reused only while the particle ids are unchanged. The
    atoms are re-ordered (same positions, permuted ids), then one id is relabelled, as a
    reaction would: each must rebuild, in step with the oracle, and the forces of the
    LennardJones drop-in stay equal to the oracle's (a stale list would scramble them)."""
    s = B.argon(5, seed=3, jitter=0.02)
    so = s.copy()
    f, fo = dp.LennardJones(B.argon_params()), A.LennardJones(B.argon_params())
    rng = np.random.default_rng(4)
    for step in range(4):
        if step == 2:
            perm = rng.permutation(s.n)
            for t in (s, so):
                for name in ("pos", "vel", "sp", "ids"):
                    setattr(t, name, np.ascontiguousarray(getattr(t, name)[perm]))
        if step == 3:
            s.ids[7] = so.ids[7] = 10 ** 6
        Fg, Ug = f.compute(s)
        Fo, Uo = fo.compute(so)
        assert f.rebuilds == fo.rebuilds == [1, 1, 2, 3][step], (step, f.rebuilds, fo.rebuilds)
        _close(Fg, Ug, Fo, Uo)


# ---------------------------------------------------------------- Lennard-Jones

def _close(Fg, Ug, Fo, Uo, rtol=1e-10):
    assert Ug == pytest.approx(Uo, rel=rtol, abs=1e-300)
    assert np.max(np.abs(Fg - Fo)) <= rtol * np.max(np.abs(Fo))


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("threshold", [2, 10 ** 9])
def test_lj_matches_oracle_two_species(periodic, threshold):
    rng = np.random.default_rng(5)
    s = A.System(3, box=4e-9, periodic=periodic)
    s.add_species("Ar", 39.948 * A.AMU)
    s.add_species("Kr", 83.8 * A.AMU)
    s.add("Ar", rng.uniform(0, 4e-9, (300, 3)))
    s.add("Kr", rng.uniform(0, 4e-9, (200, 3)))
    params = {"Ar": (1.65e-21, 3.4e-10), "Kr": (2.3e-21, 3.6e-10)}
    Fo, Uo = A.LennardJones(params, pairlist_threshold=threshold).compute(s)
    Fg, Ug = dp.LennardJones(params, pairlist_threshold=threshold).compute(s)
    _close(Fg, Ug, Fo, Uo)


def test_lj_argon_864_matches_reference_golden(golden):
    # at the perfect lattice the net forces are cancellation noise (1e-13 of a pair
    # force), so compare there only the energy; forces at the state after 20 steps
    g = golden("argon864_vv20.npz")
    s = B.argon(6, seed=64)
    _, U0 = dp.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3).compute(s)
    assert U0 == pytest.approx(float(g["U0"]), rel=1e-13)
    s.pos = g["pos"].copy()
    Fg, Ug = dp.LennardJones(B.argon_params(), cutoff=2.5, skin=0.3).compute(s)
    _close(Fg, Ug, g["F"], float(g["pe"][-1]), rtol=1e-11)


def test_lj_mixed_precision_error_is_small():
    s = B.argon(16, seed=2, jitter=0.01)      # 16384 atoms
    Fo, Uo = dp.LennardJones(B.argon_params()).compute(s)
    Fm, Um = dp.LennardJones(B.argon_params(), precision="mixed").compute(s)
    err = A.rms_rel_force_error(Fm, Fo)
    assert err < 1e-5, err
    assert abs(Um / Uo - 1) < 1e-6


def test_lj_minimum_and_empty_list():
    s = A.System(3)
    s.add_species("Ar", B.AR_MASS)
    sigma = 3.4e-10
    s.add("Ar", [[0, 0, 0], [2 ** (1 / 6) * sigma, 0, 0]])
    F, _ = dp.LennardJones({"Ar": (1e-21, sigma)}, cutoff=10).compute(s)
    assert np.allclose(F, 0, atol=1e-20)
    far = _random_system(500, 1e-6, True, seed=1)           # dilute: no pairs within the cutoff
    lj = dp.LennardJones({"S0": (1e-21, 3.4e-10)})
    F, U = lj.compute(far)
    assert U == 0.0 and np.all(F == 0)


# ---------------------------------------------------------------- DSF / Ewald

def test_dsf_madelung():
    s = B.rocksalt_unit(8)
    _, U = dp.DampedShiftedCoulomb(rcut=3.9, alpha=0.6, k=1.0).compute(s)
    assert abs(-2 * U / s.n / 1.747564594633 - 1) < 2e-3
    _, Uo = A.DampedShiftedCoulomb(rcut=3.9, alpha=0.6, k=1.0).compute(s)
    assert U == pytest.approx(Uo, rel=1e-12)


@pytest.mark.parametrize("threshold", [1, 10 ** 9])
def test_dsf_matches_oracle_with_exclusions(threshold):
    s = B.rocksalt_unit(8)
    rng = np.random.default_rng(5)
    s.pos = (s.pos + rng.normal(0, 0.2, s.pos.shape)) % s.box
    ex = [(int(s.ids[2 * m]), int(s.ids[2 * m + 1])) for m in range(50)]
    Fo, Uo = A.DampedShiftedCoulomb(rcut=2.5, k=1.0, pairlist_threshold=threshold, exclude=ex).compute(s)
    Fg, Ug = dp.DampedShiftedCoulomb(rcut=2.5, k=1.0, pairlist_threshold=threshold, exclude=ex).compute(s)
    _close(Fg, Ug, Fo, Uo, rtol=1e-11)


def test_dsf_verlet_skin_matches_fresh_lists():
    s = B.rocksalt_unit(8)
    rng = np.random.default_rng(5)
    skinned = dp.DampedShiftedCoulomb(rcut=3.9, alpha=0.6, k=1.0, skin=0.6)
    for _ in range(30):
        s.pos = (s.pos + rng.normal(0, 0.05, s.pos.shape)) % s.box
        f1, u1 = skinned.compute(s)
        f0, u0 = dp.DampedShiftedCoulomb(rcut=3.9, alpha=0.6, k=1.0, skin=0.0).compute(s)
        assert np.max(np.abs(f1 - f0)) <= 1e-12 * np.max(np.abs(f0))
        assert abs(u1 - u0) <= 1e-12 * abs(u0)
    assert 1 < skinned.rebuilds < 15


def test_dsf_vanishes_at_cutoff_with_self_term():
    dsf = dp.DampedShiftedCoulomb(rcut=1.9, k=1.0)
    two = A.System(3)
    two.add_species("A", 1.0, 1.0)
    two.add("A", [[0, 0, 0], [1.9 - 1e-9, 0, 0]])
    f2, u2 = dsf.compute(two)
    self_term = 2 * (float(A._erfc(np.array(dsf.alpha * 1.9))) / 3.8 + dsf.alpha / np.sqrt(np.pi))
    assert abs(f2[0, 0]) < 1e-7 and abs(u2 + self_term) < 1e-7


def test_dsf_mixed_precision_error():
    s = B.rocksalt(16, 16, 16, seed=4)
    rng = np.random.default_rng(1)
    s.pos = (s.pos + rng.normal(0, 0.3e-10, s.pos.shape)) % s.box
    Fo, Uo = dp.DampedShiftedCoulomb(rcut=12e-10, alpha=0.2e10).compute(s)
    Fm, Um = dp.DampedShiftedCoulomb(rcut=12e-10, alpha=0.2e10, precision="mixed").compute(s)
    Fa, Ua = A.DampedShiftedCoulomb(rcut=12e-10, alpha=0.2e10).compute(s)
    _close(Fo, Uo, Fa, Ua, rtol=1e-10)
    assert A.rms_rel_force_error(Fm, Fo) < 1e-5


@pytest.mark.parametrize("n_cells", [1, 2])
def test_ewald_madelung(n_cells):
    s = B.nacl_cells(n_cells)
    F, U = dp.Ewald(k=1.0).compute(s)
    assert -U / (s.n / 2) == pytest.approx(1.747564594633, rel=1e-7)
    assert np.allclose(F, 0, atol=1e-6)


@pytest.mark.parametrize("threshold", [400, 10 ** 9])
def test_ewald_matches_oracle_on_melt(threshold):
    s = B.rocksalt_unit(8)
    rng = np.random.default_rng(2)
    s.pos = (s.pos + rng.normal(0, 0.25, s.pos.shape)) % s.box
    Fo, Uo = A.Ewald(k=1.0, kmax=8).compute(s)
    Fg, Ug = dp.Ewald(k=1.0, kmax=8, pairlist_threshold=threshold).compute(s)
    _close(Fg, Ug, Fo, Uo, rtol=1e-10)


def test_ewald_kcut_orthorhombic_matches_oracle():
    s = B.rocksalt(6, 6, 12, seed=3)
    rng = np.random.default_rng(3)
    s.pos = (s.pos + rng.normal(0, 0.3e-10, s.pos.shape)) % s.box
    rc = 9e-10
    a = 3.0 / rc
    kc = 2 * a * 3.0
    Fo, Uo = A.Ewald(alpha=a, rcut=rc, kcut=kc, chunk=64).compute(s)
    Fg, Ug = dp.Ewald(alpha=a, rcut=rc, kcut=kc).compute(s)
    _close(Fg, Ug, Fo, Uo, rtol=1e-10)


def test_ewald_errors_follow_reference():
    s = B.nacl_cells(1)
    s.add("Na", [[0.5, 0.5, 0.5]])
    with pytest.raises(ValueError):
        dp.Ewald(k=1.0).compute(s)
    o = A.System(3)
    o.add_species("A", 1.0, 1.0)
    o.add("A", [[0, 0, 0]])
    with pytest.raises(ValueError):
        dp.Ewald().compute(o)


# ---------------------------------------------------------------- gravity / Barnes-Hut

def test_gravity_matches_oracle_and_subset():
    s, a, M = B.plummer(n=1000, seed=4)
    Fo, Uo = A.Gravity(softening=0.01 * a).compute(s)
    g = dp.Gravity(softening=0.01 * a)
    Fg, Ug = g.compute(s)
    _close(Fg, Ug, Fo, Uo, rtol=1e-12)
    sub = np.random.default_rng(0).choice(s.n, 200, replace=False)
    Fs = dp.direct_subset(g, s, sub)
    assert np.max(np.abs(Fs - A.direct_subset(g, s, sub))) <= 1e-12 * np.max(np.abs(Fs))
    Fm, Um = dp.Gravity(softening=0.01 * a, precision="mixed").compute(s)
    assert A.rms_rel_force_error(Fm, Fo) < 1e-5 and abs(Um / Uo - 1) < 1e-6


def test_direct_subset_and_gravity_targets_out_of_range():
    # tree.py:81-87 indexes s.pos[sub]: out-of-range indices raise IndexError, negative
    # ones count from the end, as numpy takes them
    s, a, M = B.plummer(n=700, seed=6)
    g = dp.Gravity(softening=0.01 * a)
    with pytest.raises(IndexError):
        dp.direct_subset(g, s, np.array([3, 700]))
    with pytest.raises(IndexError):
        dp.direct_subset(g, s, np.array([-701]))
    Fneg = dp.direct_subset(g, s, np.array([-1, 5]))
    assert np.array_equal(Fneg, dp.direct_subset(g, s, np.array([699, 5])))
    # through the C ABI a target outside [0, n) reads nothing and comes back as a NaN row
    th = dp._lib.torch()
    L = dp._lib
    pos = th.from_numpy(s.pos.copy()).cuda()
    m = th.from_numpy(s.masses.copy()).cuda()
    tgt = th.tensor([4, 700, -3, 9], dtype=th.int64, device="cuda")
    F = th.zeros((4, 3), dtype=th.float64, device="cuda")
    L.call("accu_gravity_direct", L.ptr(pos), L.ptr(m), s.n, L.ptr(tgt), 4, float(g.G), float(g.eps2), None, 0,
           L.ACCU_FP64, None, L.ptr(F), None, L.stream())
    Fh = F.cpu().numpy()
    assert np.isnan(Fh[1]).all() and np.isnan(Fh[2]).all()
    ref = A.direct_subset(g, s, np.array([4, 9]))
    assert np.max(np.abs(Fh[[0, 3]] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_barnes_hut_theta_zero_is_exact_and_matches_reference(golden):
    s, a, M = B.plummer(n=300, seed=1)
    F0, U0 = A.Gravity(softening=0.01 * a).compute(s)
    F1, U1 = dp.BarnesHut(softening=0.01 * a, theta=0.0).compute(s)
    assert A.rms_rel_force_error(F1, F0) < 1e-12 and abs(U1 - U0) / abs(U0) < 1e-12
    g = golden("bh_theta05_n300.npz")
    for th in (0.5, 0.3):
        F, U = dp.BarnesHut(softening=0.01 * a, theta=th).compute(s)
        assert A.rms_rel_force_error(F, g[f"F_{th}"]) < 1e-12
        assert U == pytest.approx(float(g[f"U_{th}"]), rel=1e-12)


@pytest.mark.parametrize("theta,tol", [(0.3, 2e-3), (0.6, 1e-2)])
def test_barnes_hut_error_bounds_and_quadrupole(theta, tol):
    s, a, M = B.plummer(n=1500, seed=1)
    F0, U0 = A.Gravity(softening=0.01 * a).compute(s)
    e_mono = A.rms_rel_force_error(dp.BarnesHut(softening=0.01 * a, theta=theta).compute(s)[0], F0)
    e_quad = A.rms_rel_force_error(dp.BarnesHut(softening=0.01 * a, theta=theta, quadrupole=True).compute(s)[0], F0)
    assert e_mono < tol and e_quad < 0.5 * e_mono
