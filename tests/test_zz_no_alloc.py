"""The allocation-free gravity entry point against the allocating one, through the C ABI.

accu_gravity_direct_noalloc runs the same kernels as accu_gravity_direct with its
scratch (energy partials, split partials, the grid-sum counter) in a fixed device
array, so the two must agree bit for bit, call after call, also when a later call is
smaller than an earlier one (the array keeps the earlier call's contents)."""
import ctypes as C

import numpy as np
import pytest

from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")
L = dp._lib

SCRATCH_DOUBLES = 1 << 20


def _inputs(n, seed):
    th = L.torch()
    s, a, M = B.plummer(n=n, seed=seed)
    pos = th.from_numpy(np.ascontiguousarray(s.pos)).cuda()
    m = th.from_numpy(np.ascontiguousarray(s.masses)).cuda()
    return s, a, pos, m


def _call(name, pos, m, n, tgt, ntgt, G, eps2, prec, origin, with_energy):
    th = L.torch()
    rows = ntgt if tgt is not None else n
    F = th.full((rows, 3), -7.0, dtype=th.float64, device="cuda")
    U = th.full((1,), -7.0, dtype=th.float64, device="cuda") if with_energy else None
    L.call(name, L.ptr(pos), L.ptr(m), n, L.ptr(tgt), ntgt, G, eps2, None, 0, prec, L.ptr(origin), L.ptr(F),
           L.ptr(U), L.stream())
    L.sync()
    return F.cpu().numpy(), (float(U.cpu().numpy()[0]) if with_energy else None)


def _compare(n, seed, prec, n_targets, with_energy):
    th = L.torch()
    s, a, pos, m = _inputs(n, seed)
    G, eps2 = float(dp.G), float((0.01 * a) ** 2)
    origin = np.ascontiguousarray(np.mean(s.pos, axis=0)) if prec else None
    tgt, nt = None, 0
    if n_targets:
        sub = np.random.default_rng(seed).choice(n, n_targets, replace=False).astype(np.int64)
        tgt, nt = th.from_numpy(sub).cuda(), n_targets
    print(f"no-alloc case n={n} seed={seed} precision={prec} targets={n_targets} energy={with_energy}", flush=True)
    Fa, Ua = _call("accu_gravity_direct", pos, m, n, tgt, nt, G, eps2, prec, origin, with_energy)
    Fb, Ub = _call("accu_gravity_direct_noalloc", pos, m, n, tgt, nt, G, eps2, prec, origin, with_energy)
    assert np.isfinite(Fa).all()
    assert np.array_equal(Fa, Fb)
    assert Ua == Ub
    return Fa, Ua


@pytest.mark.parametrize("prec", [L.ACCU_FP64, L.ACCU_FP32_PAIR])
def test_noalloc_full_sum_with_energy_equals_allocating_entry(prec):
    # grid-wide energy sum (counter and per-block partials in the fixed array); the
    # second call is smaller, so it reads fewer partials than the first one left behind
    F1, U1 = _compare(6000, 3, prec, 0, True)
    F2, U2 = _compare(2500, 4, prec, 0, True)
    assert U1 < 0 and U2 < 0 and U1 != U2


def test_noalloc_split_targets_equal_allocating_entry():
    # few targets, no energy: the sources are split over many workgroups whose force
    # partials go to the fixed array; a second, smaller call follows
    _compare(20000, 5, L.ACCU_FP64, 300, False)
    _compare(9000, 6, L.ACCU_FP64, 120, False)


def test_noalloc_targets_with_energy_then_full_sum():
    _compare(4000, 7, L.ACCU_FP64, 500, True)
    _compare(3000, 8, L.ACCU_FP64, 0, True)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    # 2^30 targets need 2^22 energy partials, beyond the 2^20 of the array: refused on
    # the host before anything is launched or read
    th = L.torch()
    dummy = th.zeros(3, dtype=th.float64, device="cuda")
    U = th.zeros(1, dtype=th.float64, device="cuda")
    n = 1 << 30
    rc = L.lib().accu_gravity_direct_noalloc(L.ptr(dummy), L.ptr(dummy), n, None, 0, 1.0, 0.0, None, 0,
                                             L.ACCU_FP64, None, L.ptr(dummy), L.ptr(U), L.stream())
    assert rc == -1
    assert "scratch" in L.lib().accu_last_error().decode()
