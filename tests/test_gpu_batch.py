"""Batches of tiles on the GPU: every tile's result in a mixed batch equals
the same tile run alone (same statuses; columns equal to the last bits up to
the 1e-12 relative noise of a row re-solved by the exact search in one run
and not the other), tiles smaller than their neighborhood, empty tiles, and
the 1e-5 truth bar per tile."""
import numpy as np
import pytest

import gpupdal_amd as g
import oracle_lib as ol
from helpers import small_cloud, truth_errors

pytestmark = pytest.mark.gpu

INFO = g.ST_TIE | g.ST_INCOMPLETE


def make_tiles(sizes, seed0=100, dup=0.0):
    tiles = []
    for i, n in enumerate(sizes):
        if n == 0:
            tiles.append((np.zeros(0), np.zeros(0), np.zeros(0)))
            continue
        x, y, z = small_cloud(n, seed=seed0 + i, dup_frac=dup)
        tiles.append((x, y, z))
    return tiles


def stages_for(knn):
    return [g.stage(g.STAGE_NORMAL, knn), g.stage(g.STAGE_COVFEAT, knn, features=g.F_ALL10)]


def test_mix_equals_alone(ctx):
    sizes = [30000, 5000, 120, 60000, 2, 17, 900, 12000]
    knns = [8, 16, 32, 16, 8, 16, 32, 8]
    tiles = make_tiles(sizes, dup=0.01)
    b = ctx.batch(tiles)
    (bn, bf), (bns, bfs) = b.run(stages_for(16), tile_knn=knns, init=np.nan)
    identical = 0
    for t, (x, y, z) in enumerate(tiles):
        (an, af), (ans, afs) = ctx.cloud(x, y, z).run(stages_for(knns[t]), init=np.nan)
        assert np.array_equal(b.tile(bns, t) & ~np.uint8(INFO), ans & ~np.uint8(INFO)), t
        assert np.array_equal(b.tile(bfs, t) & ~np.uint8(INFO), afs & ~np.uint8(INFO)), t
        for mine, alone in zip(bn + bf, an + af):
            mt = b.tile(mine, t)
            np.testing.assert_allclose(mt, alone, rtol=1e-12, atol=1e-15, equal_nan=True,
                                       err_msg=f"tile {t}")
            identical += int(np.array_equal(mt, alone, equal_nan=True))
    # Nearly every column of every tile is bit-identical.
    assert identical >= 0.9 * len(tiles) * 14


def test_batch_reruns_and_device_outputs(ctx):
    tiles = make_tiles([8000, 3000, 500])
    b = ctx.batch(tiles)
    st = stages_for(10)
    (n1, f1), (s1, fs1) = b.run(st, init=np.nan)
    b.invalidate()
    (n2, f2), (s2, fs2) = b.run(st, init=np.nan)
    for u, v in zip(n1 + f1, n2 + f2):
        assert np.array_equal(u, v, equal_nan=True)
    # Device-resident outputs give the same columns.
    N = b.n
    cols = [ctx.alloc(N * 8) for _ in range(14)]
    stat = [ctx.alloc(N) for _ in range(2)]
    b.run_device(st, cols, stat)
    for i, c in enumerate(n1 + f1):
        got = ctx.to_host(cols[i], np.zeros(N))
        ok = ~np.isnan(c)
        assert np.array_equal(got[ok], c[ok])
    for p in cols + stat:
        ctx.free(p)


def test_tiny_and_empty_tiles(ctx, oracle):
    sizes = [0, 1, 3, 5, 40, 0, 2000]
    tiles = make_tiles(sizes)
    b = ctx.batch(tiles)
    st = g.stage(g.STAGE_COVFEAT, 10, stride=2, features=g.F_DIMENSIONALITY)
    (cols,), (status,) = b.run([st], init=np.nan)
    ost = ol.stage(ol.STAGE_COVFEAT, 10, 2)
    for t, (x, y, z) in enumerate(tiles):
        if not len(x):
            continue
        ocols, ostat = oracle.stage_run(oracle.kd3(x, y, z), ost, threads=4, init=np.nan)
        assert np.array_equal(b.tile(status, t) & ~np.uint8(INFO), ostat), t
        for c, oc in zip(cols, ocols):
            np.testing.assert_allclose(b.tile(c, t), oc, rtol=1e-5, atol=1e-6, equal_nan=True)


def test_batch_tiles_vs_truth(ctx, oracle):
    sizes = [20000, 7000, 15000]
    knns = [8, 32, 16]
    tiles = make_tiles(sizes, seed0=300)
    b = ctx.batch(tiles)
    st = g.stage(g.STAGE_COVFEAT, 8, features=g.F_ALL10)
    (cols,), (status,) = b.run([st], tile_knn=knns, init=np.nan)
    for t, (x, y, z) in enumerate(tiles):
        nbr = oracle.kd3(x, y, z).neighbors_all(knns[t] + 1, threads=8)
        ost = ol.stage(ol.STAGE_COVFEAT, knns[t], features=ol.F_ALL10)
        tcols, tstat, gap, ratio = oracle.truth_from_lists(x, y, z, nbr, ost, with_ratio=True)
        stt = b.tile(status, t)
        ok = ((stt & 0x3C) == 0) & (tstat == 0) & (gap > 1e-3)
        assert np.count_nonzero(ok) > 0.9 * len(x)
        errs = truth_errors(ost, [b.tile(c, t) for c in cols], tcols, ok, ratio)
        assert max(errs.values()) <= 1.0, (t, errs)


def test_batch_argument_errors(ctx):
    b = ctx.batch(make_tiles([1000, 1000]))
    with pytest.raises(g.GfError) as e:
        b.run([g.stage(g.STAGE_NORMAL, 8), g.stage(g.STAGE_COVFEAT, 10)])
    assert e.value.code == -3  # two neighborhoods in one batch run
    with pytest.raises(g.GfError):
        b.run([g.stage(g.STAGE_COVFEAT, 8)], tile_knn=[8, 70])
