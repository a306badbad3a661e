"""Pins for the pair-MLP oracle: analytic gradient vs central differences,
exact antisymmetry of the correction, bf16 rounding."""
import numpy as np
import pytest

from oracle import builders as B
from oracle import mlp_np as M

RC, RON, ROFF = 5.0, 1.2, 1.8


def _snapshot(seed=0):
    s = B.rocksalt_unit(4)
    rng = np.random.default_rng(seed)
    s.pos = (s.pos + rng.normal(0, 0.15, s.pos.shape)) % s.box
    q = s.charges
    n = M.coordination(s.pos, s.box, RON, ROFF)
    i, j, d, r = M.directed_pairs(s.pos, s.box, RC / 2)
    x = M.features(i, j, r, q, n, RC / 2, 1.0, 8.0)
    return s, i, j, d, r, x


def test_bf16_rounding():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1 + 2 ** -9) == 1.0                      # ties to even
    assert M.bf16(1 + 3 * 2 ** -9) == 1 + 2 ** -7
    assert M.bf16(-3.14159) == pytest.approx(-3.140625)


def test_correction_is_antisymmetric():
    s, i, j, d, r, x = _snapshot(1)
    p = M.init_params(3)
    F, sv, _ = M.correction_forces(p, s.n, i, d, r, x, RC / 2, 1.0)
    assert np.allclose(F.sum(0), 0, atol=1e-12 * np.abs(F).max())
    # s is symmetric in (i, j): features are
    key = {(a, b): k for k, (a, b) in enumerate(zip(i, j))}
    for k in range(0, len(i), 97):
        assert sv[k] == sv[key[(j[k], i[k])]]


@pytest.mark.parametrize("name", ["W1", "b1", "W2", "b2", "w3", "b3"])
def test_gradient_matches_central_differences(name):
    s, i, j, d, r, x = _snapshot(2)
    p = M.init_params(5)
    rng = np.random.default_rng(7)
    target = rng.normal(0, 0.3, (s.n, 3))
    L, g, _ = M.loss_and_grad(p, s.n, i, d, r, x, RC / 2, 1.0, target, emulate_bf16=False)
    flat = p[name].reshape(-1)
    for idx in rng.choice(flat.size, size=min(flat.size, 6), replace=False):
        h = 1e-6
        old = flat[idx]
        flat[idx] = old + h
        Lp = M.loss_and_grad(p, s.n, i, d, r, x, RC / 2, 1.0, target, emulate_bf16=False)[0]
        flat[idx] = old - h
        Lm = M.loss_and_grad(p, s.n, i, d, r, x, RC / 2, 1.0, target, emulate_bf16=False)[0]
        flat[idx] = old
        num = (Lp - Lm) / (2 * h)
        assert g[name].reshape(-1)[idx] == pytest.approx(num, rel=1e-5, abs=1e-9)
