"""The oracle against the reference's own nanoflann 1.3.1 and Eigen 3.4.0
(oracle/_ref/libref.so, compiled from the reference tree by
`make -C oracle ref`) on seeded random inputs.  Skipped where the reference
tree was not available to build it (the golden fixtures cover that case)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.skipif(not ol.ref_available(),
                                reason="oracle/_ref/libref.so not built (no reference tree)")


def cloud(rng, n):
    x = np.round(rng.uniform(0, 50, n) / 0.01) * 0.01 + 500000.0
    y = np.round(rng.uniform(0, 50, n) / 0.01) * 0.01 + 4000000.0
    z = np.round((np.sin((x - 500000.0) / 7.0) * 2 + rng.normal(0, 0.05, n)) / 0.01) * 0.01
    d = n // 10
    x[:d], y[:d], z[:d] = x[d:2 * d], y[d:2 * d], z[d:2 * d]  # duplicates
    return x, y, z


@pytest.mark.parametrize("n", [5, 101, 3000, 20000])
def test_kd3_tree_and_lists(oracle, n):
    ref = ol.Ref()
    rng = np.random.default_rng(n)
    x, y, z = cloud(rng, n)
    to, tr = oracle.kd3(x, y, z), ref.kd3(x, y, z)
    eo, er = to.export(), tr.export()
    assert np.array_equal(eo["vind"], er["vind"])
    leaf = eo["child1"] < 0
    assert np.array_equal(leaf.astype(np.int32), er["leaf"])
    for key in ("divfeat", "divlow", "divhigh"):
        assert np.array_equal(eo[key][~leaf], er[key][~leaf])
    step = max(1, n // 400)
    for k in (1, 5, 11, 17, 33):
        for stride in (1, 2):
            for i in range(0, n, step):
                a = to.neighbors(x[i], y[i], z[i], k, stride)
                b = tr.neighbors(x[i], y[i], z[i], k, stride)
                assert np.array_equal(a, b), (k, stride, i)


def test_kd3_threaded_reference_build_is_the_same_tree():
    # KD3 builds with several threads (KDImpl.hpp:51-60); the tree must not
    # depend on that.
    ref = ol.Ref()
    x, y, z = cloud(np.random.default_rng(9), 30000)
    a = ref.kd3(x, y, z, threads=1).export()
    b = ref.kd3(x, y, z, threads=4).export()
    assert np.array_equal(a["vind"], b["vind"])


def test_eigen_bits(oracle):
    ref = ol.Ref()
    rng = np.random.default_rng(4)
    for t in range(5000):
        m = rng.normal(size=(3, 3))
        m = m @ m.T
        if t % 4 == 1:
            m[0, 2] = m[2, 0] = 0.0
        if t % 4 == 2:
            m = np.diag(rng.uniform(0, 1, 3))
        if t % 7 == 3:
            m *= 1e-9
        i1, e1, v1 = oracle.eigen_sym3(m)
        i2, e2, v2 = ref.eigen_sym3(m)
        assert i1 == i2 and np.array_equal(e1, e2) and np.array_equal(v1, v2), t


def test_covariance_bits(oracle):
    ref = ol.Ref()
    rng = np.random.default_rng(8)
    x, y, z = cloud(rng, 64)
    for t in range(2000):
        k = 3 + t % 60
        ids = rng.permutation(64)[:k].astype(np.uint64)
        assert np.array_equal(oracle.pdal_covariance(x, y, z, ids),
                              ref.pdal_covariance(x, y, z, ids)), t
