"""Handle-free entry points on several streams: each (device, stream) has its own scratch
(reduction partials, grid-sum counter, Ewald buffers), so calls queued on two streams
without a synchronisation between them give the results of calls made one at a time."""
import numpy as np
import pytest

from oracle import builders as B

pytestmark = pytest.mark.gpu

dp = pytest.importorskip("devinproj_amd")
from devinproj_amd import _lib as L  # noqa: E402
from devinproj_amd import forces as F  # noqa: E402


def _gravity(th, pos_d, m_d, n, F_d, U_d, stream):
    L.call("accu_gravity_direct", L.ptr(pos_d), L.ptr(m_d), n, None, 0, 1.0, 1e-4, None, 0, 0, None, L.ptr(F_d),
           L.ptr(U_d), L.C.c_void_p(stream.cuda_stream))


def test_gravity_energy_sums_on_two_streams_do_not_share_scratch():
    th = L.torch()
    rng = np.random.default_rng(5)
    sizes = (3001, 4096)
    bufs = []
    for n in sizes:
        pos = th.from_numpy(rng.normal(size=(n, 3))).cuda()
        m = th.from_numpy(rng.uniform(0.5, 1.5, n)).cuda()
        bufs.append((pos, m, n))
    # one at a time on the default stream
    ref = []
    for pos, m, n in bufs:
        Fd = th.empty((n, 3), dtype=th.float64, device="cuda")
        Ud = th.empty(1, dtype=th.float64, device="cuda")
        _gravity(th, pos, m, n, Fd, Ud, th.cuda.current_stream())
        th.cuda.synchronize()
        ref.append((Fd.cpu().numpy(), float(Ud.item())))
    # both queued, three times each, on two streams with no synchronisation in between
    streams = [th.cuda.Stream(), th.cuda.Stream()]
    outs = []
    for rep in range(3):
        for (pos, m, n), st in zip(bufs, streams):
            Fd = th.empty((n, 3), dtype=th.float64, device="cuda")
            Ud = th.empty(1, dtype=th.float64, device="cuda")
            _gravity(th, pos, m, n, Fd, Ud, st)
            outs.append((Fd, Ud))
    th.cuda.synchronize()
    for k, (Fd, Ud) in enumerate(outs):
        Fr, Ur = ref[k % 2]
        assert np.array_equal(Fd.cpu().numpy(), Fr)
        assert float(Ud.item()) == Ur


def test_ewald_on_two_streams_equals_one_at_a_time():
    th = L.torch()
    systems = [B.rocksalt_unit(4), B.rocksalt_unit(6)]
    ew = [F.Ewald(alpha=5.0 / s.box[0], rcut=s.box[0] / 2.5, kmax=6) for s in systems]
    ref = [e.compute(s) for e, s in zip(ew, systems)]
    streams = [th.cuda.Stream(), th.cuda.Stream()]
    got = [None, None]
    for rep in range(2):
        for k, (e, s, st) in enumerate(zip(ew, systems, streams)):
            with th.cuda.stream(st):
                got[k] = e.compute(s)
    for (Fg, Ug), (Fr, Ur) in zip(got, ref):
        assert np.allclose(Fg, Fr, rtol=0, atol=1e-12 * np.max(np.abs(Fr)))
        assert Ug == pytest.approx(Ur, rel=1e-13)
