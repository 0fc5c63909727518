"""Batches as a captured, replayed graph: a large batch's step (bounds, the
tiles' grids computed on the device, keys, segmented sort, column starts,
search, unflagged rows' features) is captured the second time its shape
runs and replayed after that; the exact forest is built beside it and the
flagged rows' re-solve and features read their count on the device.  Every
replay equals the launch-by-launch path bit for bit -- with new points
loaded between replays, and with the program's own eager work (another
cloud's run, a synchronisation) between them."""
import numpy as np
import pytest

import gpupdal_amd as g
from helpers import small_cloud

pytestmark = pytest.mark.gpu

SIZES = [60000, 9000, 150, 45000, 3, 0, 30000, 70000, 1200, 22000]
KNNS = [16, 8, 32, 16, 8, 16, 24, 16, 32, 8]


def tiles_for(seed):
    out = []
    for i, n in enumerate(SIZES):
        if n == 0:
            out.append((np.zeros(0), np.zeros(0), np.zeros(0)))
        else:
            out.append(small_cloud(n, seed=seed + i, dup_frac=0.02))
    return out


def stages():
    return [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 16, features=g.F_ALL10),
            g.stage(g.STAGE_EIGENVALUES, 16)]


def eager(tiles):
    c = g.Context(0)
    c.set_tuning("batch_graph", 0)
    b = c.batch(tiles)
    out = b.run(stages(), tile_knn=KNNS, init=np.nan)
    graph = c.stats()["graph"]
    b.close()
    c.close()
    assert graph == 0
    return out


def same(a, b):
    for ca, cb in zip(a[0], b[0]):
        for x, y in zip(ca, cb):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    for sa, sb in zip(a[1], b[1]):
        assert np.array_equal(sa, sb)


def test_graph_replays_equal_eager():
    t1, t2 = tiles_for(300), tiles_for(400)
    assert sum(SIZES) >= 131072
    want1, want2 = eager(t1), eager(t2)
    c = g.Context(0)
    b = c.batch(t1)
    modes = []
    for step in range(5):
        print(f"batch graph step {step}", flush=True)
        if step == 3:
            for i, (x, y, z) in enumerate(t2):
                b.set_tile(i, x, y, z)
        got = b.run(stages(), tile_knn=KNNS, init=np.nan)
        modes.append(c.stats()["graph"])
        same(got, want2 if step >= 3 else want1)
        # eager work between replays
        other = c.cloud(*t1[0])
        other.run([g.stage(g.STAGE_EIGENVALUES, 8)])
        other.close()
        c.sync()
    b.close()
    c.close()
    assert modes == [0, 2, 1, 1, 1], modes


def test_graph_device_outputs_and_invalidate():
    t1 = tiles_for(500)
    want = eager(t1)
    c = g.Context(0)
    b = c.batch(t1)
    n = b.n
    ncols = 3 + 4 + 10
    cols = [c.alloc(n * 8) for _ in range(ncols)]
    stat = [c.alloc(n) for _ in range(3)]
    for step in range(3):
        b.invalidate()
        b.run_device(stages(), cols, stat, tile_knn=KNNS)
    assert c.stats()["graph"] == 1
    host = [np.empty(n) for _ in range(ncols)]
    hs = [np.empty(n, np.uint8) for _ in range(3)]
    for h, d in zip(host, cols):
        c.copy_to_host(h, d)
    for h, d in zip(hs, stat):
        c.copy_to_host(h, d)
    got = ([host[0:4], host[4:14], host[14:17]], hs)
    # rows every stage skipped hold what the device columns held: compare
    # the rows with status 0 only
    for (cw, sw), (cg, sg) in zip(zip(want[0], want[1]), zip(got[0], got[1])):
        assert np.array_equal(sw, sg)
        ok = sw == 0
        for x, y in zip(cw, cg):
            assert np.array_equal(x[ok].view(np.uint64), y[ok].view(np.uint64))
    for p in cols + stat:
        c.free(p)
    b.close()
    c.close()
