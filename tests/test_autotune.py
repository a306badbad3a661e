"""Autotuner: candidate tiles fit the cell grid and the LDS (CPU), the cache round-trips
(CPU), and a tuned tile runs the same physics as the smallest tile (GPU)."""
import json

import numpy as np
import pytest

dp = pytest.importorskip("devinproj_amd")
from devinproj_amd import autotune, batch, workloads as W  # noqa: E402


def _lj(prec):
    return lambda s: [dp.LennardJones(W.argon_params(), cutoff=2.5, skin=0.3, precision=prec)]


def test_candidates_fit_grid_and_lds():
    s = W.argon(16, seed=2, jitter=0.01)  # 16 FCC cells per axis -> 9 list cells per axis
    for prec, elem in (("mixed", 16), ("fp64", 32)):
        cands = autotune.candidates_for(s, _lj(prec)(s), prec)
        assert cands, prec
        rc = autotune._list_radius(s, _lj(prec)(s))
        nc = np.floor(s.box / rc).astype(int)
        per_cell = s.n / np.prod(nc)
        for b in cands:
            assert all(b[d] <= nc[d] for d in range(3))
            assert np.prod(np.asarray(b) + 2) * per_cell * elem < autotune.LDS_BYTES
    # the float64 region is twice as large, so it admits no larger shapes than mixed
    assert set(autotune.candidates_for(s, _lj("fp64")(s), "fp64")) <= set(autotune.candidates_for(s, _lj("mixed")(s),
                                                                                                   "mixed"))


def test_cache_round_trip(tmp_path):
    s = W.argon(8, seed=2, jitter=0.01)
    key = autotune.bucket_key(s, _lj("mixed")(s), "mixed")
    f = tmp_path / "tiles.json"
    f.write_text(json.dumps({key: {"bricks": [2, 2, 3], "threads": 512}}))
    autotune._CACHE.clear()
    try:
        assert autotune.cached_tile(s, _lj("mixed")(s), "mixed", str(f)) == {"bricks": (2, 2, 3), "threads": 512}
        assert autotune.cached_tile(s, _lj("fp64")(s), "fp64", str(f)) is None
    finally:
        autotune._CACHE.clear()


@pytest.mark.gpu
def test_tuned_tile_runs_the_same_physics(tmp_path):
    s = W.argon(10, seed=5, jitter=0.02)  # 4000 atoms, 5 list cells per axis
    autotune._CACHE.clear()
    try:
        tile = autotune.choose_tile(s, _lj("mixed"), "mixed", 10e-15, cache_file=str(tmp_path / "t.json"), steps=10)
    finally:
        autotune._CACHE.clear()
    assert tile["bricks"] in autotune.candidates_for(s, _lj("mixed")(s), "mixed")
    assert tile["threads"] in autotune.THREADS
    runs = []
    for bricks, threads in ((tile["bricks"], tile["threads"]), ((1, 1, 1), 256)):
        r = s.copy()
        eng = batch.BrickEngine([r], [_lj("mixed")(r)], dp.VelocityVerlet(), 10e-15, precision="mixed",
                                   bricks=[bricks], threads=threads)
        eng.run(20)
        eng.fetch()
        runs.append((r.pos.copy(), eng.logs[0][-1, :2].copy()))
        eng.close()
    d = runs[0][0] - runs[1][0]
    d -= s.box * np.round(d / s.box)
    assert np.max(np.abs(d)) < 1e-6 * W.AR_SIGMA
    assert np.allclose(runs[0][1], runs[1][1], rtol=1e-6)
