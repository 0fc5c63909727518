"""Neighborhoods larger than the grid search's 63 entries ((knn + 1) *
stride up to 1024): every row runs the exact KD3 search with its result set
in LDS.  Lists equal the oracle's KD3Index::neighbors exactly, in order
(every row is exact); stage statuses equal PDAL's CPU path and columns are
within the 1e-5 bar of the extended-precision truth; nndistance is bit-equal.
PDAL itself runs any knn (filters/CovarianceFeaturesFilter.cpp:226-230)."""
import numpy as np
import pytest

import gpupdal_amd as g
import oracle_lib as ol
from helpers import small_cloud, truth_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle():
    return ol.Oracle()


@pytest.mark.parametrize("n,k,stride", [(20000, 100, 1), (20000, 31, 3), (6000, 1024, 1),
                                        (500, 700, 1)])
def test_large_lists_equal_kd3(ctx, oracle, n, k, stride):
    print(f"large knn lists n={n} k={k} stride={stride}", flush=True)
    x, y, z = small_cloud(n, seed=51, dup_frac=0.03)
    cl = ctx.cloud(x, y, z)
    ids, st = cl.knn(k, stride)
    cl.close()
    want = oracle.kd3(x, y, z).neighbors_all(k, stride, threads=8)
    assert np.array_equal(ids, want.astype(np.uint32))


@pytest.mark.parametrize("knn,stride", [(80, 1), (30, 3)])
def test_large_stages_vs_truth(ctx, oracle, knn, stride):
    print(f"large knn stages knn={knn} stride={stride}", flush=True)
    x, y, z = small_cloud(20000, seed=52, dup_frac=0.02)
    stages = [g.stage(g.STAGE_COVFEAT, knn, stride, features=g.F_ALL10),
              g.stage(g.STAGE_EIGENVALUES, knn, stride)]
    cl = ctx.cloud(x, y, z)
    cols, status = cl.run(stages, init=np.nan)
    cl.close()
    t = oracle.kd3(x, y, z)
    nbr = t.neighbors_all(knn + 1, stride, threads=8)
    for st_, c_, s_ in zip(stages, cols, status):
        ost = ol.stage(st_.stage, st_.knn, st_.stride, st_.mode, st_.features, st_.always_up,
                       st_.normalize)
        _, ps = oracle.stage_run(t, ost, threads=8, init=np.nan)
        skip = s_ & ~np.uint8(g.ST_TIE | g.ST_INCOMPLETE)
        assert np.array_equal(skip, ps)
        tc, ts, gap, ratio = oracle.truth_from_lists(x, y, z, nbr, ost, with_ratio=True)
        ok = (skip == 0) & (ts == 0) & (gap > 1e-3)
        assert np.count_nonzero(ok) > 0.5 * len(x)
        errs = truth_errors(ost, c_, tc, ok, ratio)
        assert max(errs.values()) <= 1.0, errs


def test_large_nndistance_and_limits(ctx, oracle):
    x, y, z = small_cloud(10000, seed=53, dup_frac=0.02)
    t = oracle.kd3(x, y, z)
    cl = ctx.cloud(x, y, z)
    for k in (100, 500):
        assert np.array_equal(cl.nndistance(k, "avg"), oracle.nn_distance(t, k, ol.NND_AVG))
    with pytest.raises(g.GfError, match="1024"):
        cl.run([g.stage(g.STAGE_EIGENVALUES, 1024)])
    with pytest.raises(g.GfError, match="1024"):
        cl.knn(1025)
    cl.close()
