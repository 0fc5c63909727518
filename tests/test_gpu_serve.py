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


def test_bad_request_refused_at_submit():
    """A request the batch would refuse (stride 0, an unknown mode, a
    neighborhood outside the device envelope) is refused by submit itself;
    the valid requests submitted around it still run and succeed."""
    srv = g.Server(workers=1)
    x, y, z = small_cloud(20000, seed=610)
    good = [srv.submit(x, y, z, stages(), 10)]
    bad = [g.stage(g.STAGE_COVFEAT, 10, stride=0),
           g.stage(g.STAGE_COVFEAT, 10, mode=7)]
    for st in bad:
        with pytest.raises(g.GfError) as e:
            srv.submit(x, y, z, [st], 10)
        assert e.value.code == g.GF_ERR_ARG
    good.append(srv.submit(x, y, z, stages(), 16))
    for t in good:
        cols, status, times = srv.result(t)
        assert np.all(np.isfinite(cols[1][0][(status[1] & 0x3C) == 0]))
    srv.close()


def test_wait_timeout_is_distinct():
    """A wait that times out reports GF_ERR_TIMEOUT (not a failure), and the
    request can still be waited for afterwards."""
    srv = g.Server(workers=1)
    x, y, z = small_cloud(2_000_000, seed=611)
    t = srv.submit(x, y, z, stages(), 16)
    rc = srv.L.gf_server_wait(srv.h, t, 0.0)
    assert rc in (g.GF_OK, g.GF_ERR_TIMEOUT)
    cols, status, times = srv.result(t)
    assert times[2] >= times[0]
    srv.close()
