"""The oracle's KD3Index::radius restatement (oracle/kd3.c) and its radius
stages, pinned on the CPU: counts equal brute force (d2 < r*r with the
index's distance arithmetic); the lists equal the reference's own nanoflann
radiusSearch (oracle/_ref) as sets, in distance order, with equal distances
in either order; radius stages skip rows below min_k exactly as the counts
say; and their truth equals numpy's two-pass covariance + eigh."""
import numpy as np
import pytest

import oracle_lib as ol
from helpers import small_cloud


@pytest.fixture(scope="module")
def orc():
    return ol.Oracle()


@pytest.mark.parametrize("n,r,dup", [(3000, 0.6, 0.0), (2500, 0.35, 0.05), (2000, 1.3, 0.0)])
def test_radius_counts_equal_brute_force(orc, n, r, dup):
    x, y, z = small_cloud(n, seed=5, dup_frac=dup)
    t = orc.kd3(x, y, z)
    assert np.array_equal(t.radius_counts(r, threads=4), orc.brute_radius_counts(x, y, z, r))


def test_radius_boundary_is_strict(orc):
    # Points on a 0.5 m lattice: radius 0.5 excludes the four at exactly 0.5.
    g = np.arange(5) * 0.5
    x, y = np.meshgrid(g, g)
    x, y = x.ravel(), y.ravel()
    z = np.zeros_like(x)
    t = orc.kd3(x, y, z)
    ids, d2 = t.radius(1.0, 1.0, 0.0, 0.5)
    assert list(ids) == [12]
    ids, d2 = t.radius(1.0, 1.0, 0.0, 0.5000001)
    assert sorted(ids) == [7, 11, 12, 13, 17]


def test_radius_lists_match_reference_nanoflann(orc):
    try:
        ref = ol.Ref()
    except FileNotFoundError:
        pytest.skip("reference harness not built (no /root/reference at build time)")
    x, y, z = small_cloud(4000, seed=8, dup_frac=0.03)
    t = orc.kd3(x, y, z)
    rt = ref.kd3(x, y, z)
    rng = np.random.default_rng(1)
    for q in rng.choice(len(x), 300, replace=False):
        a, da = t.radius(x[q], y[q], z[q], 0.7)
        b, db = rt.radius(x[q], y[q], z[q], 0.7)
        assert np.array_equal(da, db)  # the same distances, in order
        assert np.array_equal(np.sort(a), np.sort(b))
        # equal distances may come in another order (std::sort is not stable)
        for d in np.unique(da):
            assert set(a[da == d]) == set(b[db == d])


def test_radius_stage_min_k_and_truth(orc):
    x, y, z = small_cloud(3000, seed=12)
    t = orc.kd3(x, y, z)
    r = 0.45
    counts = t.radius_counts(r)
    st = ol.stage(ol.STAGE_EIGENVALUES, 8, radius=r, min_k=6)
    cols, status = orc.stage_run(t, st)
    few = counts < 6
    assert np.array_equal(status == ol.ST_FEW_NEIGHBORS, few)
    assert np.all(np.isnan(cols[0][few])) or np.all(cols[0][few] == 0)
    # truth over the same lists: two-pass covariance and eigh (numpy)
    rows = np.nonzero(~few)[0][:200]
    offs, ids = t.radius_csr(rows, r)
    tc, ts, gap = orc.truth_from_csr(x, y, z, offs, ids, st)
    for i, q in enumerate(rows):
        nb = ids[offs[i]:offs[i + 1]].astype(np.int64)
        p = np.stack([x[nb], y[nb], z[nb]], 1)
        c = np.cov((p - p.mean(0)).T)
        ev = np.linalg.eigvalsh(c)
        assert np.allclose([tc[k][i] for k in range(3)], ev, rtol=1e-9, atol=1e-14)
