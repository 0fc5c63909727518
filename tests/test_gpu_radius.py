"""Radius neighborhoods on the GPU (the stages' `radius` / `min_k` options):
neighbor counts equal KD3Index::radius's (the oracle restatement, itself
pinned against brute force and the reference's nanoflann); stage columns
within the 1e-5 bar of the extended-precision truth over the same radius
sets; rows below min_k (covariancefeatures, eigenvalues) or three (normal)
skipped exactly where the counts say; statuses equal PDAL's CPU path."""
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


def _dense_cloud(n, seed):
    """An ALS-like cloud plus a dense flat roof (quantized, constant z) and a
    duplicate stack, so that some radius boxes hold more candidates than the
    registers and are streamed."""
    x, y, z = small_cloud(n, seed=seed, dup_frac=0.02)
    rng = np.random.default_rng(seed)
    m = n // 5
    rx = np.round((x.min() + 2.0 + rng.uniform(0, 1.5, m)) / 0.01) * 0.01
    ry = np.round((y.min() + 2.0 + rng.uniform(0, 1.5, m)) / 0.01) * 0.01
    rz = np.full(m, np.round((z.max() + 3.0) / 0.01) * 0.01)
    return (np.concatenate([x, rx]), np.concatenate([y, ry]), np.concatenate([z, rz]))


@pytest.mark.parametrize("r", [0.35, 0.8, 1.6])
def test_radius_counts_equal_kd3(ctx, oracle, r):
    x, y, z = _dense_cloud(30000, seed=21)
    cl = ctx.cloud(x, y, z)
    got = cl.radius_counts(r)
    cl.close()
    want = oracle.kd3(x, y, z).radius_counts(r)
    assert np.array_equal(got, want)
    assert want.max() > 256  # the roof's boxes are streamed


@pytest.mark.parametrize("r,min_k", [(0.5, 3), (0.9, 12)])
def test_radius_stages_vs_truth(ctx, oracle, r, min_k):
    x, y, z = _dense_cloud(25000, seed=22)
    stages = [g.stage(g.STAGE_NORMAL, 8, radius=r),
              g.stage(g.STAGE_EIGENVALUES, 8, radius=r, min_k=min_k),
              g.stage(g.STAGE_COVFEAT, 8, radius=r, min_k=min_k, features=g.F_ALL10)]
    cl = ctx.cloud(x, y, z)
    cols, status = cl.run(stages, init=np.nan)
    cl.close()
    tree = oracle.kd3(x, y, z)
    counts = tree.radius_counts(r)
    rng = np.random.default_rng(3)
    rows = np.sort(rng.choice(len(x), 3000, replace=False))
    offs, ids = tree.radius_csr(rows, r)
    for st_, c_, s_ in zip(stages, cols, status):
        ost = ol.stage(st_.stage, st_.knn, st_.stride, st_.mode, st_.features, st_.always_up,
                       st_.normalize, radius=st_.radius, min_k=st_.min_k)
        need = 3 if st_.stage == g.STAGE_NORMAL else min_k
        few = counts < need
        assert np.array_equal((s_ & g.ST_FEW_NEIGHBORS) != 0, few)
        # PDAL's CPU path: the same skip / zero-covariance statuses.
        _, pst = oracle.stage_run(tree, ost, threads=8)
        assert np.array_equal(s_, pst)
        tc, ts, gap, ratio = oracle.truth_from_csr(x, y, z, offs, ids, ost, with_ratio=True)
        ok = (s_[rows] == 0) & (ts == 0) & (gap > 1e-3)
        assert np.count_nonzero(ok) > 0.5 * len(rows)
        errs = truth_errors(ost, [cc[rows] for cc in c_], tc, ok, ratio)
        assert max(errs.values()) <= 1.0, errs


def test_radius_refused_in_batches(ctx):
    x, y, z = small_cloud(5000, seed=3)
    b = ctx.batch([(x, y, z)])
    with pytest.raises(g.GfError):
        b.run([g.stage(g.STAGE_COVFEAT, 8, radius=0.5)])
    b.close()


def test_radius_argument_errors(ctx):
    x, y, z = small_cloud(2000, seed=4)
    cl = ctx.cloud(x, y, z)
    with pytest.raises(g.GfError):
        cl.run([g.stage(g.STAGE_COVFEAT, 8, radius=-1.0)])
    with pytest.raises(g.GfError):
        cl.radius_counts(0.0)
    cl.close()
