"""gf_run_noalloc: the fused stage run with every working buffer in the
library's fixed per-GPU scratch array.  Its outputs equal gf_run's on the
same points, bit for bit, on two calls in a row (the second smaller, so the
scratch array keeps the first call's contents), and sizes beyond its limit
are refused with an error before any work."""
import numpy as np
import pytest

import gpupdal_amd as g
from helpers import small_cloud

pytestmark = pytest.mark.gpu

INFO = g.ST_TIE | g.ST_INCOMPLETE


def _device_run(ctx, x, y, z, stages):
    """gf_run_noalloc on device copies of x, y, z; returns host columns and
    statuses (columns start as NaN, so skipped rows stay NaN)."""
    n = len(x)
    ncols = sum(g.stage_columns(s) for s in stages)
    dx, dy, dz = ctx.to_device(x), ctx.to_device(y), ctx.to_device(z)
    nan = np.full(n, np.nan)
    cols = [ctx.to_device(nan) for _ in range(ncols)]
    st = [ctx.to_device(np.zeros(n, np.uint8)) for _ in stages]
    try:
        rc = ctx.run_noalloc(dx, dy, dz, n, stages, cols, st)
        assert rc == g.GF_OK, ctx.L.gf_last_error(ctx.h)
        out_c = [ctx.to_host(p, np.empty(n)) for p in cols]
        out_s = [ctx.to_host(p, np.empty(n, np.uint8)) for p in st]
    finally:
        for p in [dx, dy, dz] + cols + st:
            ctx.free(p)
    return out_c, out_s


def _resident_run(ctx, x, y, z, stages):
    cl = ctx.cloud(x, y, z)
    try:
        cols, st = cl.run(stages, init=np.nan)
    finally:
        cl.close()
    return [c for group in cols for c in group], st


@pytest.mark.parametrize("stages", [
    [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 16)],
    [g.stage(g.STAGE_EIGENVALUES, 10), g.stage(g.STAGE_COVFEAT, 10, features=g.F_ALL10)],
])
def test_noalloc_equals_gf_run_twice(stages):
    ctx = g.Context(0)
    try:
        # The first call on the largest size the scratch takes, the second
        # on a smaller cloud with duplicates (ties re-solved by the tree).
        for n, seed, dup in [(g.NOALLOC_MAX_POINTS, 41, 0.0), (20000, 42, 0.03)]:
            x, y, z = small_cloud(n, seed=seed, dup_frac=dup)
            got_c, got_s = _device_run(ctx, x, y, z, stages)
            want_c, want_s = _resident_run(ctx, x, y, z, stages)
            for a, b in zip(got_s, want_s):
                assert np.array_equal(a, b), (n, "status")
            for a, b in zip(got_c, want_c):
                assert np.array_equal(a, b, equal_nan=True), (n, "columns")
            flagged = np.count_nonzero((got_s[0] & INFO) != 0)
            if dup:
                assert flagged > 0  # the exact tree ran in the scratch array
    finally:
        ctx.close()


def test_noalloc_refuses_oversize():
    ctx = g.Context(0)
    try:
        n = g.NOALLOC_MAX_POINTS + 1
        st = [g.stage(g.STAGE_COVFEAT, 8)]
        rc = ctx.run_noalloc(1, 1, 1, n, st, [1, 1, 1, 1], [1])
        assert rc == g.GF_ERR_ARG
        assert b"at most" in ctx.L.gf_last_error(ctx.h)
    finally:
        ctx.close()
