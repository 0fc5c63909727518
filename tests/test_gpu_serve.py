"""The serving loop on the GPU: requests submitted concurrently (continuous
batching, two workers, an HBM budget small enough to force waiting) return
the same result as each tile run alone."""
import threading

import numpy as np
import pytest

import gpupdal_amd as g
from helpers import small_cloud

pytestmark = pytest.mark.gpu

INFO = g.ST_TIE | g.ST_INCOMPLETE


def stages():
    return [g.stage(g.STAGE_NORMAL, 10), g.stage(g.STAGE_COVFEAT, 10)]


def test_served_equals_alone(ctx):
    sizes = [40000, 3000, 25000, 800, 60000, 5, 12000, 30000, 7000, 50000]
    knns = [10, 16, 10, 16, 10, 16, 10, 16, 10, 16]
    tiles = [small_cloud(n, seed=500 + i) for i, n in enumerate(sizes)]
    # Budget for roughly two of the larger requests at a time.
    srv = g.Server(workers=2, hbm_budget_bytes=40 << 20)
    tickets = [None] * len(tiles)

    def submit(i):
        tickets[i] = srv.submit(*tiles[i], stages(), knns[i], init=np.nan)
    threads = [threading.Thread(target=submit, args=(i,)) for i in range(len(tiles))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    batches = set()
    for i, (x, y, z) in enumerate(tiles):
        cols, status, times = srv.result(tickets[i])
        assert times[0] <= times[1] <= times[2]
        batches.add(times[3])
        st = [g.stage(g.STAGE_NORMAL, knns[i]), g.stage(g.STAGE_COVFEAT, knns[i])]
        acols, astat = ctx.cloud(x, y, z).run(st, init=np.nan)
        for s_srv, s_alone in zip(status, astat):
            assert np.array_equal(s_srv & ~np.uint8(INFO), s_alone & ~np.uint8(INFO)), i
        for a, b in zip(cols[0] + cols[1], acols[0] + acols[1]):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15, equal_nan=True,
                                       err_msg=f"request {i}")
    srv.close()
    assert len(batches) >= 2  # the budget split the requests over several batches


def test_request_over_budget_is_refused():
    srv = g.Server(workers=1, hbm_budget_bytes=4 << 20)
    x, y, z = small_cloud(50000, seed=600)
    t = srv.submit(x, y, z, stages(), 10)
    with pytest.raises(g.GfError) as e:
        srv.result(t)
    assert e.value.code == -4
    srv.close()
