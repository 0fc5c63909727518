"""The oracle's other kNN consumers, pinned on the CPU.

filters.nndistance: the reference's own known answers on a 10x10x10 integer
grid (test/unit/filters/NNDistanceTest.cpp:43-125, exact equality), and the
distances of brute-force neighbor lists.  filters.outlier: the reference's
counts on autzen-point-format-3.las (test/unit/OldPCLBlockTest.cpp:43-196),
the radius method against brute-force counts, the statistical threshold
against numpy's two-pass mean and sample deviation.  approximatecoplanar and
estimaterank: planes, lines and duplicates whose answer is fixed, numpy's
covariance + eigh away from the thresholds, and the rank against the
reference's own Eigen JacobiSVD (oracle/_ref) away from the threshold."""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
from helpers import small_cloud

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import las_oracle  # noqa: E402

AUTZEN = os.path.join(HERE, "golden", "las", "autzen-point-format-3.las")


@pytest.fixture(scope="module")
def orc():
    return ol.Oracle()


def faux_grid():
    """readers.faux mode=grid, bounds [0,10]^3, count 1000: the integer
    lattice 0..9 on each axis."""
    g = np.arange(10, dtype=np.float64)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def test_nndistance_reference_grid_known_answers(orc):
    x, y, z = faux_grid()
    t = orc.kd3(x, y, z)
    assert np.all(orc.nn_distance(t, 3, ol.NND_KTH) == 1.0)
    d = orc.nn_distance(t, 4, ol.NND_KTH)
    corner = np.isin(x, (0, 9)) & np.isin(y, (0, 9)) & np.isin(z, (0, 9))
    assert corner.sum() == 8
    assert np.all(d[corner] == np.sqrt(2))
    d = orc.nn_distance(t, 6, ol.NND_AVG)
    e = np.isin(x, (0, 9)).astype(int) + np.isin(y, (0, 9)) + np.isin(z, (0, 9))
    assert np.all(d[e == 3] == (1 + np.sqrt(2)) / 2.0)
    assert np.all(d[e == 2] == (2 + np.sqrt(2)) / 3.0)
    assert np.all(d[e == 1] == (5 + np.sqrt(2)) / 6.0)
    assert np.all(d[e == 0] == 1.0)


@pytest.mark.parametrize("k", [1, 8, 20])
def test_nndistance_equals_brute_force_distances(orc, k):
    x, y, z = small_cloud(3000, seed=21, dup_frac=0.02)
    t = orc.kd3(x, y, z)
    _, d2 = orc.brute_knn(x, y, z, k + 1)
    r = np.sqrt(d2)
    assert np.array_equal(orc.nn_distance(t, k, ol.NND_KTH), r[:, k])
    avg = np.zeros(len(x))
    for j in range(1, k + 1):
        avg = avg + r[:, j]
    assert np.array_equal(orc.nn_distance(t, k, ol.NND_AVG), avg / k)
    run = np.zeros(len(x))
    for j in range(1, k + 1):
        run = run + (r[:, j] - run) / j
    assert np.array_equal(orc.nn_distance(t, k, ol.NND_RUNNING_MEAN), run)


def test_nndistance_fewer_points_than_k(orc):
    # Three points, k = 5: the missing distances are 0 (PDAL zero-fills).
    x = np.array([0.0, 3.0, 0.0])
    y = np.array([0.0, 0.0, 4.0])
    z = np.zeros(3)
    t = orc.kd3(x, y, z)
    assert np.array_equal(orc.nn_distance(t, 5, ol.NND_KTH), np.zeros(3))
    avg = orc.nn_distance(t, 5, ol.NND_AVG)
    assert avg[0] == (3.0 + 4.0) / 5.0
    assert avg[1] == (3.0 + 5.0) / 5.0
    assert avg[2] == (4.0 + 5.0) / 5.0
    assert np.isnan(orc.nn_distance(t, 0, ol.NND_AVG)).all()  # 0 / 0


def autzen_xyz():
    buf = open(AUTZEN, "rb").read()
    return las_oracle.read_xyz(buf)


@pytest.mark.parametrize("mean_k,mult,inliers", [(2, 1.5, 96), (5, 0.0, 63)])
def test_outlier_statistical_reference_counts(orc, mean_k, mult, inliers):
    x, y, z = autzen_xyz()
    assert len(x) == 106
    f, thr, d = orc.outlier_statistical(orc.kd3(x, y, z), mean_k, mult)
    assert len(x) - int(f.sum()) == inliers


@pytest.mark.parametrize("r,min_k,inliers", [(200.0, 1, 60), (100.0, 2, 3)])
def test_outlier_radius_reference_counts(orc, r, min_k, inliers):
    x, y, z = autzen_xyz()
    f = orc.outlier_radius(orc.kd3(x, y, z), r, min_k)
    assert len(x) - int(f.sum()) == inliers


def test_outlier_radius_equals_brute_counts(orc):
    x, y, z = small_cloud(3000, seed=4, dup_frac=0.02)
    t = orc.kd3(x, y, z)
    for r, mk in ((0.4, 2), (0.9, 5), (0.2, 0), (0.5, -1)):
        c = orc.brute_radius_counts(x, y, z, r).astype(np.int64)
        want = (c <= mk).astype(np.uint8) if mk >= 0 else np.ones(len(x), np.uint8)
        assert np.array_equal(orc.outlier_radius(t, r, mk), want)


def test_outlier_statistical_threshold_two_pass(orc):
    x, y, z = small_cloud(4000, seed=9)
    t = orc.kd3(x, y, z)
    f, thr, d = orc.outlier_statistical(t, 8, 1.7)
    _, d2 = orc.brute_knn(x, y, z, 9)
    run = np.zeros(len(x))
    for j in range(1, 9):
        run = run + (np.sqrt(d2[:, j]) - run) / j
    assert np.array_equal(d, run)
    want = d.mean() + 1.7 * d.std(ddof=1)
    assert abs(thr - want) <= 1e-12 * abs(want)
    far = np.abs(d - want) > 1e-9 * abs(want)
    assert np.array_equal(f[far], (d[far] >= want).astype(np.uint8))
    # one point: variance 0 / 0, threshold NaN, every point an outlier
    t1 = orc.kd3(x[:1], y[:1], z[:1])
    f1, thr1, _ = orc.outlier_statistical(t1, 8, 2.0)
    assert np.isnan(thr1) and f1.tolist() == [1]


# ----------------------------------------------- coplanar and rank ---------

def lists(orc, x, y, z, k):
    t = orc.kd3(x, y, z)
    return t, t.neighbors_all(k, 1, threads=8).astype(np.int64)


def np_cov(x, y, z, nbr):
    p = np.stack([x[nbr], y[nbr], z[nbr]], axis=-1)
    p = p - p.mean(axis=1, keepdims=True)
    return np.einsum("nki,nkj->nij", p, p) / (nbr.shape[1] - 1)


def test_coplanar_against_numpy(orc):
    x, y, z = small_cloud(3000, seed=14)
    t, nbr = lists(orc, x, y, z, 10)
    st = ol.stage(ol.STAGE_COPLANAR, 10)
    cols, status = orc.stage_run(t, st, threads=8)
    ev = np.linalg.eigvalsh(np_cov(x, y, z, nbr))
    a = ev[:, 1] - 25.0 * ev[:, 0]
    b = 6.0 * ev[:, 1] - ev[:, 2]
    scale = np.abs(ev[:, 2]) + 1e-300
    clear = (np.abs(a) > 1e-9 * scale) & (np.abs(b) > 1e-9 * scale) & (status == 0)
    want = ((a > 0) & (b > 0)).astype(np.float64)
    assert clear.sum() > 0.95 * len(x)
    assert np.array_equal(cols[0][clear], want[clear])
    assert 0 < want[clear].sum() < clear.sum()  # both answers occur


def test_coplanar_fixed_answers(orc):
    rng = np.random.default_rng(3)
    n = 400
    # a tilted plane: coplanar; a line: not; an isotropic ball: not
    u, v = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    plane = (u, v, 0.3 * u - 0.2 * v)
    s = rng.uniform(-5, 5, n)
    line = (s, 0.5 * s, -0.25 * s)
    ball = tuple(rng.normal(size=n) for _ in range(3))
    for (x, y, z), want in ((plane, 1.0), (line, 0.0), (ball, 0.0)):
        t = orc.kd3(*(np.ascontiguousarray(c) for c in (x, y, z)))
        cols, status = orc.stage_run(t, ol.stage(ol.STAGE_COPLANAR, 16), threads=4)
        ok = status == 0
        assert ok.sum() > 0.9 * n
        frac = np.mean(cols[0][ok] == want)
        assert frac > (0.98 if want else 0.9)
    # all-duplicate neighborhoods are skipped (zero covariance)
    x = np.zeros(50)
    cols, status = orc.stage_run(orc.kd3(x, x, x), ol.stage(ol.STAGE_COPLANAR, 8), init=-1.0)
    assert np.all(status == ol.ST_ZERO_COV) and np.all(cols[0] == -1.0)


def test_rank_fixed_answers(orc):
    rng = np.random.default_rng(7)
    n = 300
    s = rng.uniform(-3, 3, n)
    cases = [((s, np.zeros(n), np.zeros(n)), 1),  # exactly on the x axis
             ((rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), np.zeros(n)), 2),
             (tuple(rng.normal(size=n) for _ in range(3)), 3),
             ((np.ones(n), np.ones(n), np.ones(n)), 0)]
    for (x, y, z), want in cases:
        t = orc.kd3(*(np.ascontiguousarray(c) for c in (x, y, z)))
        cols, status = orc.stage_run(t, ol.stage(ol.STAGE_RANK, 8), threads=4)
        assert np.all(status == 0)
        assert np.all(cols[0] == want), (want, np.unique(cols[0]))


def test_rank_against_reference_jacobi_svd(orc):
    try:
        ref = ol.Ref()
    except FileNotFoundError:
        pytest.skip("reference harness not built (no /root/reference at build time)")
    rng = np.random.default_rng(11)
    checked = 0
    for i in range(3000):
        k = int(rng.integers(3, 20))
        p = rng.normal(size=(k, 3)) * rng.choice([1.0, 1e-2, 1e-4, 1e-9], 3)
        if i % 5 == 0:
            p[:, 2] = 0.0
        if i % 7 == 0:
            p[:, 1:] = 0.0
        p = p @ np.linalg.qr(rng.normal(size=(3, 3)))[0] + rng.uniform(-1e3, 1e3, 3)
        x, y, z = (np.ascontiguousarray(p[:, j]) for j in range(3))
        ids = np.arange(k, dtype=np.uint64)
        cov = orc.pdal_covariance(x, y, z, ids)
        info, ev, _ = orc.eigen_sym3(cov)
        assert info == 0
        thr = float(rng.choice([0.01, 1e-3, 1e-6, 0.1]))
        mine = orc.rank_from_eigenvalues(ev, thr)
        theirs = ref.jacobi_rank(cov, thr)
        s = np.sort(np.abs(ev))[::-1]
        cut = max(s[0] * float(np.float32(thr)), np.finfo(np.float64).tiny)
        near = np.any(np.abs(s - cut) <= 1e-9 * max(s[0], cut))
        if not near:
            assert mine == theirs, (i, ev, thr)
            checked += 1
    assert checked > 2500
