"""The exact re-solve on a KD3 tree deeper than its 64 register frames (a
line of geometrically spaced points, each twice, more than 80 levels, and
every row's boundary tied for an odd neighbor count; the oracle's
tree for it equals the reference's nanoflann, tests/test_kd3_depth.py).  The
rows whose search outgrows the frames run again with frames in device
memory: every row's neighbor set equals the oracle's KD3 lists, statuses
equal PDAL's, and the call reports how many rows took the deep pass."""
import numpy as np
import pytest

import gpupdal_amd as g
import oracle_lib as ol
from test_kd3_depth import skewed_line

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


def test_deep_tree_lists_and_stage(ctx):
    x, y, z = skewed_line()
    orc = ol.Oracle()
    t = orc.kd3(x, y, z)
    want = t.neighbors_all(15, 1, threads=8)
    print("deep tree: knn lists", flush=True)
    cl = ctx.cloud(x, y, z)
    ids, st = cl.knn(15)
    assert np.all(st != 0)  # every row re-solved by the exact search
    assert np.array_equal(np.sort(ids, axis=1), np.sort(want.astype(np.uint32), axis=1))
    print("deep tree: eigenvalues stage", flush=True)
    cols, status = cl.run([g.stage(g.STAGE_EIGENVALUES, 14)], init=np.nan)
    deep = ctx.stats()["deep_rows"]
    cl.close()
    assert deep > 0
    pc, ps = orc.stage_run(t, ol.stage(ol.STAGE_EIGENVALUES, 14), threads=8, init=np.nan)
    skip = status[0] & ~np.uint8(g.ST_TIE | g.ST_INCOMPLETE)
    assert np.array_equal(skip, ps)
    ok = skip == 0
    # Collinear neighborhoods: one eigenvalue, the other two at rounding level.
    lam = pc[2][ok]
    assert np.all(np.abs(cols[0][2][ok] - lam) <= 1e-5 * lam)
    assert np.all(np.abs(cols[0][0][ok]) <= 1e-12 * lam)
    assert np.all(np.abs(cols[0][1][ok]) <= 1e-12 * lam)
