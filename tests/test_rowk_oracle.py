"""covariancefeatures' optimized=true in the oracle: each row's neighborhood is
KD3Index::neighbors(p, OptimalKNN, 1) (filters/CovarianceFeaturesFilter.cpp
:204-210).  Pinned on the CPU: a constant count equals the plain stage with
knn = count - 1; a row's list for k is the first k entries of its list for a
larger K, in the oracle's KD3 restatement and in the reference's own
nanoflann (oracle/_ref), on clouds full of exact distance ties -- the
property the device relies on when it takes each row's prefix of one list;
and a count of 0 skips the row as a zero covariance."""
import numpy as np
import pytest

import oracle_lib as ol
from helpers import small_cloud


@pytest.fixture(scope="module")
def orc():
    return ol.Oracle()


def lattice_cloud(rng, n):
    """Points on a 0.05 m lattice (many exact distance ties) with duplicates."""
    x = np.round(rng.uniform(0, 3, n) / 0.05) * 0.05 + 500000.0
    y = np.round(rng.uniform(0, 3, n) / 0.05) * 0.05 + 4000000.0
    z = np.round(rng.normal(0, 0.05, n) / 0.05) * 0.05
    d = n // 20
    x[:d], y[:d], z[:d] = x[d:2 * d], y[d:2 * d], z[d:2 * d]
    return x, y, z


def test_constant_count_equals_plain_stage(orc):
    x, y, z = small_cloud(3000, seed=3, dup_frac=0.02)
    t = orc.kd3(x, y, z)
    st = ol.stage(ol.STAGE_COVFEAT, 9, features=ol.F_ALL10)
    want, ws = orc.stage_run(t, st, threads=8, init=np.nan)
    got, gs = orc.stage_run_rowk(t, st, np.full(len(x), 10, np.uint32), init=np.nan)
    assert np.array_equal(gs, ws)
    for a, b in zip(got, want):
        assert np.array_equal(a, b, equal_nan=True)


def test_prefix_lists_oracle(orc):
    rng = np.random.default_rng(5)
    x, y, z = lattice_cloud(rng, 4000)
    t = orc.kd3(x, y, z)
    for q in rng.choice(len(x), 300, replace=False):
        big = t.neighbors(x[q], y[q], z[q], 30)
        for k in (2, 7, 11, 14, 23):
            assert np.array_equal(t.neighbors(x[q], y[q], z[q], k), big[:k])


def test_prefix_lists_reference_nanoflann():
    try:
        ref = ol.Ref()
    except FileNotFoundError:
        pytest.skip("reference harness not built (no /root/reference at build time)")
    rng = np.random.default_rng(6)
    x, y, z = lattice_cloud(rng, 4000)
    rt = ref.kd3(x, y, z)
    for q in rng.choice(len(x), 300, replace=False):
        big = rt.neighbors(x[q], y[q], z[q], 30)
        for k in (2, 7, 11, 14, 23):
            assert np.array_equal(rt.neighbors(x[q], y[q], z[q], k), big[:k])


def test_rowk_equals_prefix_lists_and_zero_skips(orc):
    rng = np.random.default_rng(8)
    x, y, z = lattice_cloud(rng, 3000)
    t = orc.kd3(x, y, z)
    row_k = rng.integers(2, 20, len(x)).astype(np.uint32)
    row_k[::97] = 0
    st = ol.stage(ol.STAGE_COVFEAT, 8, features=ol.F_ALL10)
    got, gs = orc.stage_run_rowk(t, st, row_k, init=np.nan)
    nbr = t.neighbors_all(int(row_k.max()), 1, threads=8)
    assert np.all(gs[row_k == 0] == ol.ST_ZERO_COV)
    for k in np.unique(row_k[row_k > 0]):
        rows = np.nonzero(row_k == k)[0]
        want, ws = orc.stage_from_lists(x, y, z, nbr[rows, :k], st, init=np.nan)
        assert np.array_equal(gs[rows], ws)
        for a, b in zip(got, want):
            assert np.array_equal(a[rows], b, equal_nan=True)
