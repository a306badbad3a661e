"""The device-built Barnes-Hut octree against the reference's tree (tree.py:30-53, restated
as oracle.bh_tree_preorder) node for node, and its walk against the oracle (tree.py:54-78).

Structure is integer/bit-exact: node count, pre-order, cube centers and halves (the same
float64 center chain), next pointers, leaf flags, particle counts and the particle set of
every node. Mass moments differ only in summation order (1e-14). Cases cover the
reference's special geometry: particles exactly on splitting planes (`>` is strict),
n <= leaf_size (root leaf), coincident particles (the depth-48 cap), a cluster 1e-11
wide that needs more than 21 levels (the deep pass), a flat set and two particles."""
import numpy as np
import pytest

from oracle import accusim_np as A
from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")
from devinproj_amd import tree as T  # noqa: E402


def _cases():
    rng = np.random.default_rng(3)
    out = []
    s, a, M = B.plummer(n=1500, seed=1)
    out.append(("plummer1500", s.pos, s.masses, 8))
    s3, _, _ = B.plummer(n=300, seed=1)
    out.append(("plummer300_leaf1", s3.pos, s3.masses, 1))
    out.append(("plummer300_leaf3", s3.pos, s3.masses, 3))
    g = np.stack(np.meshgrid(np.arange(5.), np.arange(5.), np.arange(5.), indexing="ij"), -1).reshape(-1, 3)
    out.append(("grid_on_planes", g, np.ones(len(g)), 8))
    out.append(("n_le_leaf", g[:6].copy(), np.ones(6), 8))
    p = rng.random((400, 3))
    p[:20] = 0.5
    p[20:40] = 0.25 + rng.random((20, 3)) * 1e-11
    mm = rng.random(400)
    mm[50:60] = 0.0
    out.append(("coincident_and_tight_cluster", p, mm, 8))
    out.append(("all_same", np.zeros((10, 3)), np.ones(10), 8))
    f = rng.random((500, 3))
    f[:, 2] = 0.0
    out.append(("flat", f, np.ones(500), 8))
    out.append(("two", rng.random((2, 3)), np.array([1.0, 2.0]), 1))
    return out


CASES = _cases()


def _device(pos, m):
    th = dp._lib.torch()
    return (th.from_numpy(np.ascontiguousarray(pos, float)).cuda(),
            th.from_numpy(np.ascontiguousarray(m, float)).cuda())


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_tree_equals_reference_tree(case):
    tag, pos, m, leaf = case
    print(f"bh tree {tag} n={len(pos)} leaf={leaf}", flush=True)
    o = A.bh_tree_preorder(pos, m, leaf)
    tr = T.Octree()
    pd, md = _device(pos, m)
    tr.build(pd, md, len(pos), leaf, quadrupole=True)
    info = tr.info()
    t = tr.export(quadrupole=True)
    nn = len(o["half"])
    assert info["nodes"] == nn and info["leaves"] == int(o["leaf"].sum())
    assert info["max_depth"] == int(o["depth"].max())
    assert np.array_equal(t["cube"][:, :3], o["center"])
    assert np.array_equal(t["cube"][:, 3], o["half"])
    assert np.array_equal(t["next"], o["next"])
    assert np.array_equal(t["leaf"].astype(bool), o["leaf"])
    assert np.array_equal(t["count"], o["count"])
    perm = t["perm"]
    assert np.array_equal(np.sort(perm), np.arange(len(pos)))
    for i in range(nn):
        a = t["start"][i]
        assert np.array_equal(np.sort(perm[a:a + t["count"][i]]), o["members"][i]), (tag, i)
    assert np.allclose(t["com"][:, 3], o["mass"], rtol=1e-14, atol=0)
    scale = np.abs(pos).max() + o["half"][0]
    assert np.max(np.abs(t["com"][:, :3] - o["com"])) <= 1e-14 * scale
    # quadrupoles (extension): direct sums over each node's particles about its com
    q = t["quad"]
    for i in range(nn):
        mem = o["members"][i]
        x = pos[mem] - o["com"][i]
        r2 = (x * x).sum(1)
        w = m[mem]
        ref = np.array([np.sum(w * (3 * x[:, 0] ** 2 - r2)), np.sum(w * (3 * x[:, 1] ** 2 - r2)),
                        np.sum(w * (3 * x[:, 2] ** 2 - r2)), np.sum(w * 3 * x[:, 0] * x[:, 1]),
                        np.sum(w * 3 * x[:, 0] * x[:, 2]), np.sum(w * 3 * x[:, 1] * x[:, 2])])
        # round-off floor: moving moments between centers of mass that are known to
        # eps * |position| costs about mass * node size * eps * |position|
        tol = 1e-12 * np.sum(w * r2) + 1e-13 * np.sum(w) * scale * 2 * o["half"][i] + 1e-300
        assert np.max(np.abs(q[i] - ref)) <= tol, (tag, i, q[i], ref)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_walk_equals_oracle(case):
    tag, pos, m, leaf = case
    print(f"bh walk {tag}", flush=True)
    soft = 0.05 * (np.ptp(pos, 0).max() + 1e-3)
    for theta in (0.0, 0.5):
        of = A.BarnesHut(softening=soft, theta=theta, leaf_size=leaf)
        Fo, Uo = _oracle_bh(of, pos, m)
        f = dp.BarnesHut(softening=soft, theta=theta, leaf_size=leaf)
        pd, md = _device(pos, m)
        th = dp._lib.torch()
        F = th.empty((len(pos), 3), dtype=th.float64, device="cuda")
        U = th.empty(1, dtype=th.float64, device="cuda")
        f.forces_device(pd, md, len(pos), F, U)
        Fg = F.cpu().numpy()
        assert np.max(np.abs(Fg - Fo)) <= 1e-12 * np.max(np.abs(Fo)) + 1e-300, (tag, theta)
        assert float(U.cpu()[0]) == pytest.approx(Uo, rel=1e-12, abs=1e-300), (tag, theta)


def _oracle_bh(f, pos, m):
    class _S:
        pass
    s = _S()
    s.pos, s.masses, s.n = pos, m, len(pos)
    return f.compute(s)


def test_one_handle_rebuilt_across_sizes_and_options():
    rng = np.random.default_rng(9)
    tr = T.Octree()
    th = dp._lib.torch()
    for n, leaf, quad in ((3000, 8, True), (200, 4, False), (5000, 16, True), (50, 8, False)):
        pos = rng.normal(size=(n, 3))
        m = rng.random(n) + 0.5
        print(f"bh rebuild n={n} leaf={leaf} quad={quad}", flush=True)
        pd, md = _device(pos, m)
        tr.build(pd, md, n, leaf, quad)
        assert tr.info()["nodes"] == len(A.bh_tree_preorder(pos, m, leaf)["half"])
        F = th.empty((n, 3), dtype=th.float64, device="cuda")
        tr.walk(1.0, 0.01, 0.0, F, None)
        Fo, _ = _oracle_bh(A.BarnesHut(G=1.0, softening=0.1, theta=0.0, leaf_size=leaf), pos, m)
        assert np.max(np.abs(F.cpu().numpy() - Fo)) <= 1e-12 * np.max(np.abs(Fo))


def test_walk_without_a_tree_is_a_value_error():
    tr = T.Octree()
    th = dp._lib.torch()
    F = th.empty((4, 3), dtype=th.float64, device="cuda")
    with pytest.raises(ValueError, match="no tree"):
        tr.walk(1.0, 0.0, 0.5, F, None)
    with pytest.raises(ValueError, match="leaf_size"):
        pd, md = _device(np.zeros((4, 3)), np.ones(4))
        tr.build(pd, md, 4, 0, False)


def test_plummer_262144_tree_and_error():
    """The 262,144-star workload's tree: node count equals the restated recursion's and the
    theta = 0.5 quadrupole force error on 200 sampled stars (accuracy.py:135-146) is small."""
    s, a, M = B.plummer(n=262144, seed=5)
    f = dp.BarnesHut(softening=0.01 * a, theta=0.5, quadrupole=True)
    F, U = f.compute(s)
    info = f.tree.info()
    print(f"plummer 262144: {info}", flush=True)
    assert info["nodes"] == len(A.bh_tree_preorder(s.pos, s.masses, 8)["half"])
    sub = np.random.default_rng(0).choice(s.n, 200, replace=False)
    Fd = dp.direct_subset(f, s, sub)
    err = A.rms_rel_force_error(F[sub], Fd)
    print(f"theta 0.5 quadrupole rms rel force error {err:.3e}", flush=True)
    assert err < 1e-3
