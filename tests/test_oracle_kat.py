"""The oracle against the reference's own known-answer tests and against
what the mathematics fixes (closed forms, invariants, brute force, LAPACK).

Runs on the CPU.  Sources of the known answers:
  test/unit/KDIndexTest.cpp:118-222            KD3Index order and stride
  test/unit/filters/CovarianceFeaturesTest.cpp:46-188   line / plane / scattering
  test/unit/filters/NormalFilterTest.cpp:47-253          planes and ramps
"""
import math

import numpy as np
import pytest

import oracle_lib as ol


def _run_stage(orc, x, y, z, st):
    tree = orc.kd3(x, y, z)
    return orc.stage_run(tree, st, threads=1, init=np.nan)


# ------------------------------------------------------------- KD3Index --
KD_PTS = np.array([[0, 0, 0], [1, 1, 1], [3, 3, 3], [6, 6, 6], [10, 10, 10]], float)


@pytest.fixture(scope="module")
def kd5(oracle):
    return oracle.kd3(KD_PTS[:, 0], KD_PTS[:, 1], KD_PTS[:, 2])


@pytest.mark.parametrize("q,want", [((0, 0, 0), 0), ((1.1, 1.1, 1.1), 1), ((3.3, 3.3, 3.3), 2),
                                    ((6.1, 6.1, 6.1), 3), ((15, 15, 15), 4)])
def test_kd3_nearest(kd5, q, want):
    # KDIndexTest.cpp:149-153
    assert list(kd5.neighbors(*q, 1)) == [want]


@pytest.mark.parametrize("q,k,stride,want", [
    ((0, 0, 0), 5, 1, [0, 1, 2, 3, 4]),      # KDIndexTest.cpp:156-162
    ((0, 0, 0), 25, 1, [0, 1, 2, 3, 4]),     # :164-170, k clipped to n
    ((3.1, 3.1, 3.1), 5, 1, [2, 1, 3, 0, 4]),  # :172-178
    ((0, 0, 0), 2, 1, [0, 1]),               # :200-203
    ((0, 0, 0), 2, 2, [0, 2]),               # :205-208
    ((0, 0, 0), 2, 3, [0, 3]),               # :210-213
    ((0, 0, 0), 3, 2, [0, 2, 4]),            # :215-219
])
def test_kd3_neighbors_kat(kd5, q, k, stride, want):
    assert list(kd5.neighbors(*q, k, stride)) == want


def test_kd3_stride_past_end_pads_with_zero(kd5):
    # KDImpl.hpp:351-381: capacity stride*k, output zero-initialised; with
    # k=5, stride=2 only entries 0, 2, 4 exist, 6 and 8 stay 0.
    assert list(kd5.neighbors(0, 0, 0, 5, 2)) == [0, 2, 4, 0, 0]


def test_kd3_matches_brute_force_without_ties(oracle):
    orc = oracle
    # Away from boundary ties the KD3 set is the (d2, id)-ordered brute-force
    # set (tests/unit/spatial_index_test.cpp:602-647 checks the same).
    rng = np.random.default_rng(7)
    n = 3000
    x, y, z = rng.uniform(0, 30, n), rng.uniform(0, 30, n), rng.normal(0, 1, n)
    tree = orc.kd3(x, y, z)
    for k in (1, 8, 17, 33):
        kd = tree.neighbors_all(k)
        bf, d2 = orc.brute_knn(x, y, z, k)
        assert np.array_equal(kd, bf), k
        assert np.all(np.diff(d2, axis=1) >= 0)


def test_kd3_tie_keeps_first_visited(oracle):
    orc = oracle
    # Two points at exactly the same distance from the query: the result
    # holds one of them, and the same one the tree visits first every time
    # (KNNResultSet::addPoint inserts after equal entries; the leaf test is
    # strict <).  In a single leaf the visit order is the tree order, which
    # for a leaf of the unsplit root is the input order.
    x = np.array([0.0, 1.0, -1.0, 5.0])
    y = np.zeros(4)
    z = np.zeros(4)
    tree = orc.kd3(x, y, z)
    assert list(tree.neighbors(0, 0, 0, 2)) == [0, 1]
    x2 = np.array([0.0, -1.0, 1.0, 5.0])
    tree2 = orc.kd3(x2, y, z)
    assert list(tree2.neighbors(0, 0, 0, 2)) == [0, 1]


def test_kd3_tie_winner_is_not_the_smaller_id(oracle):
    # A plausible mistake -- breaking distance ties by point id instead of
    # visit order -- fails here: the tied points at x=-1 and x=+1 are ids 1
    # and 0 in one single-leaf tree, and the leaf visits id 0 (x=+1) first.
    x = np.array([1.0, -1.0, 5.0, 0.0])
    tree = oracle.kd3(x, np.zeros(4), np.zeros(4))
    assert list(tree.neighbors(0, 0, 0, 2)) == [3, 0]
    # In a split tree the closer child is searched first: with 300 points
    # on a line the root splits, and from a query just left of the cut the
    # tied partner on the left side wins even though it has the larger id.
    n = 300
    xs = np.arange(n, dtype=float)
    xs = np.concatenate([xs[n // 2:], xs[: n // 2]])  # right half gets small ids
    tree = oracle.kd3(xs, np.zeros(n), np.zeros(n))
    # Root cut at 149.5: the query x=149 (id 299) searches the left child
    # first and keeps x=148 (id 298); x=150 (id 0) ties at distance 1 but
    # arrives later and loses on the strict <.
    assert list(tree.neighbors(149.0, 0, 0, 2)) == [299, 298]


# ------------------------------------------------------------ covariance --
def test_pdal_covariance_closed_form(oracle):
    orc = oracle
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 10, 40) + 500000.0
    y = rng.uniform(0, 10, 40) + 4000000.0
    z = rng.uniform(0, 3, 40)
    ids = np.arange(0, 40, 3)
    c = orc.pdal_covariance(x, y, z, ids)
    p = np.stack([x[ids], y[ids], z[ids]])
    mean = p.mean(axis=1, keepdims=True)
    a = (p - mean).astype(np.float32).astype(np.float64)  # MathUtils.cpp:177-189
    want = a @ a.T / (len(ids) - 1)
    np.testing.assert_allclose(c, want, rtol=1e-12, atol=1e-15)
    # Without the float narrowing the relative difference is ~1e-7, so the
    # test would catch a dropped narrowing at this scale.
    exact = np.cov(p)
    assert np.max(np.abs(c - exact)) > 0 or np.allclose(c, exact)


def test_is_zero_threshold(oracle):
    orc = oracle
    z = np.zeros((3, 3))
    assert orc.is_zero(z)
    z[1, 2] = z[2, 1] = 1e-12
    assert orc.is_zero(z)
    z[1, 2] = z[2, 1] = 1.0000001e-12
    assert not orc.is_zero(z)


# ----------------------------------------------------------------- eigen --
def _random_sym(rng, i):
    m = rng.normal(size=(3, 3))
    m = m @ m.T
    if i % 5 == 1:
        m[0, 2] = m[2, 0] = 0.0
    if i % 5 == 2:
        m = np.diag(rng.uniform(0, 1, 3))
    if i % 5 == 3:
        u = rng.normal(size=3)
        m = np.outer(u, u)  # rank 1
    if i % 7 == 4:
        m *= 1e-9
    return m


def test_eigen_matches_lapack_and_invariants(oracle):
    orc = oracle
    rng = np.random.default_rng(11)
    for i in range(2000):
        m = _random_sym(rng, i)
        info, ev, vec = orc.eigen_sym3(m)
        assert info == 0
        w = np.linalg.eigvalsh(m)
        scale = max(np.abs(m).max(), 1e-300)
        np.testing.assert_allclose(ev, w, rtol=0, atol=1e-13 * scale)
        assert np.all(np.diff(ev) >= 0)
        np.testing.assert_allclose(vec.T @ vec, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(m @ vec, vec * ev, rtol=0, atol=1e-12 * scale)
        assert abs(ev.sum() - np.trace(m)) <= 1e-12 * scale * 3


def test_eigen_analytic_cases(oracle):
    orc = oracle
    info, ev, vec = orc.eigen_sym3(np.diag([3.0, 1.0, 2.0]))
    assert info == 0 and list(ev) == [1.0, 2.0, 3.0]
    np.testing.assert_array_equal(np.abs(vec[:, 0]), [0, 1, 0])
    # A plane z=0: the smallest eigenvector is the z axis.
    info, ev, vec = orc.eigen_sym3(np.array([[2.0, 0.3, 0], [0.3, 1.0, 0], [0, 0, 0]]))
    assert ev[0] == 0.0
    np.testing.assert_allclose(np.abs(vec[:, 0]), [0, 0, 1], atol=1e-15)


def test_truth_solver_matches_lapack(oracle):
    orc = oracle
    rng = np.random.default_rng(5)
    for i in range(500):
        m = _random_sym(rng, i)
        ev, vec = orc.jacobi_sym3(m)
        w = np.linalg.eigvalsh(m)
        np.testing.assert_allclose(ev, w, rtol=0, atol=1e-14 * max(np.abs(m).max(), 1e-300))


# ---------------------------------------------- covariancefeatures KATs --
ALL10 = ol.F_ALL10


def _feat(cols, st, name):
    names = [ol.FEATURE_NAMES[b] for b in range(10) if st.features & (1 << b)]
    return cols[names.index(name)]


def test_covfeat_line():
    orc = ol.Oracle()
    x, y, z = np.zeros(3), np.zeros(3), np.array([0.0, 1.0, 2.0])
    st = ol.stage(ol.STAGE_COVFEAT, 3, features=ALL10)
    cols, status = _run_stage(orc, x, y, z, st)
    f = lambda nme: _feat(cols, st, nme).astype(np.float32)  # noqa: E731
    assert np.all(status == 0)
    np.testing.assert_allclose(f("Linearity"), 1, rtol=4e-7)
    np.testing.assert_allclose(f("Verticality"), 1, rtol=4e-7)
    np.testing.assert_allclose(f("Planarity"), 0, atol=1e-6)
    np.testing.assert_allclose(f("Scattering"), 0, atol=1e-6)
    np.testing.assert_allclose(f("Omnivariance"), 0, atol=1e-6)
    np.testing.assert_allclose(f("Anisotropy"), 1, rtol=4e-7)
    assert np.all(np.isnan(f("Eigenentropy")))
    np.testing.assert_allclose(f("EigenvalueSum"), 1, rtol=4e-7)
    np.testing.assert_allclose(f("SurfaceVariation"), 0, atol=1e-6)
    np.testing.assert_allclose(f("DemantkeVerticality"), 1, rtol=4e-7)


def test_covfeat_plane():
    orc = ol.Oracle()
    x = np.array([0.0, 1, 1, 0])
    y = np.array([0.0, 0, 1, 1])
    z = np.zeros(4)
    st = ol.stage(ol.STAGE_COVFEAT, 10, features=ALL10)
    cols, status = _run_stage(orc, x, y, z, st)
    f = lambda nme: _feat(cols, st, nme).astype(np.float32)  # noqa: E731
    assert np.all(f("Linearity") <= 0.5)
    np.testing.assert_allclose(f("Verticality"), 0, atol=1e-6)
    np.testing.assert_allclose(f("Planarity"), 1, rtol=4e-7)
    np.testing.assert_allclose(f("Scattering"), 0, atol=1e-6)
    np.testing.assert_allclose(f("Omnivariance"), 0, atol=1e-6)
    np.testing.assert_allclose(f("Anisotropy"), 1, rtol=4e-7)
    assert np.all(np.isnan(f("Eigenentropy")))
    np.testing.assert_allclose(f("EigenvalueSum"), 0.667, atol=0.001)
    np.testing.assert_allclose(f("SurfaceVariation"), 0, atol=1e-6)
    np.testing.assert_allclose(f("DemantkeVerticality"), 0, atol=1e-6)


def test_covfeat_scattering():
    orc = ol.Oracle()
    x = np.array([0.0, 1, 1, 0])
    y = np.array([0.0, 0, 1, 1])
    z = np.array([0.0, 0, 1, 2])
    st = ol.stage(ol.STAGE_COVFEAT, 10, features=ALL10)
    cols, status = _run_stage(orc, x, y, z, st)
    f = lambda nme: _feat(cols, st, nme).astype(np.float32)  # noqa: E731
    assert np.all(f("Linearity") <= 0.5)
    assert np.all(f("Verticality") >= 0.5)
    assert np.all(f("Planarity") <= 1)
    assert np.all(f("Scattering") >= 0.1)
    np.testing.assert_allclose(f("Omnivariance"), 0.458, atol=0.001)
    np.testing.assert_allclose(f("Anisotropy"), 0.864, atol=0.001)
    np.testing.assert_allclose(f("Eigenentropy"), 0.488, atol=0.001)
    np.testing.assert_allclose(f("EigenvalueSum"), 1.583, atol=0.001)
    np.testing.assert_allclose(f("SurfaceVariation"), 0.095, atol=0.001)
    np.testing.assert_allclose(f("DemantkeVerticality"), 0.492, atol=0.001)


def test_covfeat_modes_closed_form():
    # One neighborhood (all 4 points), eigenvalues known from numpy: raw,
    # sqrt and normalized modes follow CovarianceFeaturesFilter.cpp:249-365.
    orc = ol.Oracle()
    x = np.array([0.0, 2, 2, 0])
    y = np.array([0.0, 0, 1, 1])
    z = np.array([0.0, 0, 0.5, 0.25])
    p = np.stack([x, y, z])
    a = (p - p.mean(axis=1, keepdims=True)).astype(np.float32).astype(np.float64)
    lam = np.sort(np.linalg.eigvalsh(a @ a.T / 3))[::-1]
    for mode, tr in ((ol.MODE_RAW, lambda v: v), (ol.MODE_SQRT, np.sqrt),
                     (ol.MODE_NORMALIZED, lambda v: v / v.sum())):
        st = ol.stage(ol.STAGE_COVFEAT, 10, mode=mode, features=ALL10)
        cols, status = _run_stage(orc, x, y, z, st)
        # The sum is taken before the mode's transform and SurfaceVariation
        # divides the transformed smallest eigenvalue by it.
        s = np.maximum(lam, 0).sum()
        l0, l1, l2 = tr(np.maximum(lam, 0))
        want = {"Linearity": (l0 - l1) / l0, "Planarity": (l1 - l2) / l0,
                "Scattering": l2 / l0, "Anisotropy": (l0 - l2) / l0,
                "Omnivariance": np.cbrt(l0 * l1 * l2),
                "EigenvalueSum": s,
                "Eigenentropy": -(l0 * np.log(l0) + l1 * np.log(l1) + l2 * np.log(l2)),
                "SurfaceVariation": l2 / s}
        for name, v in want.items():
            np.testing.assert_allclose(_feat(cols, st, name), v, rtol=1e-6,
                                       err_msg=f"mode {mode} {name}")


# ---------------------------------------------------------- normal KATs --
def _grid(bx, by, bz, count=4):
    # FauxReader mode=grid over integer bounds: points at every integer
    # coordinate with lo <= v < hi (io/FauxReader.cpp grid mode).
    xs = np.arange(bx[0], bx[1]) if bx[1] > bx[0] else np.array([bx[0]])
    ys = np.arange(by[0], by[1]) if by[1] > by[0] else np.array([by[0]])
    zs = np.arange(bz[0], bz[1]) if bz[1] > bz[0] else np.array([bz[0]])
    pts = [(a, b, c) for c in zs for b in ys for a in xs][:count]
    p = np.array(pts, float)
    return p[:, 0], p[:, 1], p[:, 2]


@pytest.mark.parametrize("bounds,want", [
    (((0, 2), (0, 2), (0, 0)), (0, 0, 1)),   # NormalFilterTest.cpp:47-82
    (((0, 2), (0, 0), (0, 2)), (0, 1, 0)),   # :84-119
    (((0, 0), (0, 2), (0, 2)), (1, 0, 0)),   # :121-156
])
def test_normal_axis_planes(bounds, want):
    orc = ol.Oracle()
    x, y, z = _grid(*bounds)
    st = ol.stage(ol.STAGE_NORMAL, 3)
    cols, status = _run_stage(orc, x, y, z, st)
    for c, w in zip(cols[:3], want):
        np.testing.assert_allclose(c.astype(np.float32), w, atol=1e-7)
    np.testing.assert_allclose(cols[3].astype(np.float32), 0, atol=1e-7)


def test_normal_ramp_plane():
    # NormalFilterTest.cpp:158-205: points on z = x; normal (-sqrt(2)/2, 0,
    # sqrt(2)/2) with always_up.
    orc = ol.Oracle()
    xs, ys = np.meshgrid(np.arange(3.0), np.arange(3.0))
    x, y = xs.ravel(), ys.ravel()
    z = x.copy()
    st = ol.stage(ol.STAGE_NORMAL, 8)
    cols, status = _run_stage(orc, x, y, z, st)
    h = math.sqrt(2) / 2
    np.testing.assert_allclose(cols[0], -h, atol=1e-7)
    np.testing.assert_allclose(cols[1], 0, atol=1e-7)
    np.testing.assert_allclose(cols[2], h, atol=1e-7)


def test_normal_always_up_flips():
    orc = ol.Oracle()
    xs, ys = np.meshgrid(np.arange(3.0), np.arange(3.0))
    x, y = xs.ravel(), ys.ravel()
    z = 0.5 * x
    up = _run_stage(orc, x, y, z, ol.stage(ol.STAGE_NORMAL, 8, always_up=1))[0]
    assert np.all(up[2] > 0)
    free = _run_stage(orc, x, y, z, ol.stage(ol.STAGE_NORMAL, 8, always_up=0))[0]
    np.testing.assert_allclose(np.abs(free[2]), up[2], rtol=1e-15)


def test_eigenvalues_stage_closed_form():
    orc = ol.Oracle()
    rng = np.random.default_rng(2)
    x, y, z = rng.uniform(0, 1, 12), rng.uniform(0, 1, 12), rng.uniform(0, 0.1, 12)
    p = np.stack([x, y, z])
    a = (p - p.mean(axis=1, keepdims=True)).astype(np.float32).astype(np.float64)
    lam = np.linalg.eigvalsh(a @ a.T / 11)
    for norm in (0, 1):
        st = ol.stage(ol.STAGE_EIGENVALUES, 20, normalize=norm)
        cols, status = _run_stage(orc, x, y, z, st)
        want = lam / lam.sum() if norm else lam
        for i in range(3):
            np.testing.assert_allclose(cols[i], want[i], rtol=1e-9, atol=1e-15)


def test_zero_covariance_rows_are_skipped():
    # Coincident points: covariance is zero, the stage leaves the row's
    # columns untouched and reports it (NormalFilter.cpp:160-165,
    # EigenvaluesFilter.cpp:150-155).
    orc = ol.Oracle()
    x = np.ones(6)
    for kind in (ol.STAGE_NORMAL, ol.STAGE_EIGENVALUES, ol.STAGE_COVFEAT):
        st = ol.stage(kind, 3)
        cols, status = _run_stage(orc, x, x, x, st)
        assert np.all(status & ol.ST_ZERO_COV)
        for c in cols:
            assert np.all(np.isnan(c))


# ------------------------------------------------ truth and list entries --
def test_truth_features_match_numpy(oracle):
    # The high-precision truth (long double two-pass covariance + Jacobi)
    # against numpy's float64 covariance + LAPACK on well-separated
    # eigenvalues, where both are accurate to ~1e-12.
    rng = np.random.default_rng(21)
    n, k = 300, 12
    x, y, z = rng.uniform(0, 5, n), rng.uniform(0, 5, n), rng.uniform(0, 2, n)
    nbr = np.stack([rng.choice(n, k, replace=False) for _ in range(n)]).astype(np.uint32)
    st = ol.stage(ol.STAGE_COVFEAT, k - 1, mode=ol.MODE_RAW, features=ALL10)
    cols, status, gap = oracle.truth_from_lists(x, y, z, nbr, st)
    for i in range(0, n, 7):
        p = np.stack([x[nbr[i]], y[nbr[i]], z[nbr[i]]])
        w, v = np.linalg.eigh(np.cov(p))
        l0, l1, l2 = w[2], w[1], w[0]
        assert abs(_feat(cols, st, "Linearity")[i] - (l0 - l1) / l0) < 1e-10
        assert abs(_feat(cols, st, "Planarity")[i] - (l1 - l2) / l0) < 1e-10
        assert abs(_feat(cols, st, "EigenvalueSum")[i] - w.sum()) < 1e-10 * w.sum()
        assert abs(_feat(cols, st, "DemantkeVerticality")[i] - (1 - abs(v[2, 0]))) < 1e-8


def test_stage_from_lists_equals_stage_run(oracle):
    rng = np.random.default_rng(22)
    n = 2000
    x, y, z = rng.uniform(0, 20, n), rng.uniform(0, 20, n), rng.normal(0, 0.3, n)
    tree = oracle.kd3(x, y, z)
    for st in (ol.stage(ol.STAGE_NORMAL, 8), ol.stage(ol.STAGE_EIGENVALUES, 8, normalize=1),
               ol.stage(ol.STAGE_COVFEAT, 10, features=ALL10),
               ol.stage(ol.STAGE_COVFEAT, 6, stride=2)):
        nbr = tree.neighbors_all(st.knn + 1, st.stride if st.stage != ol.STAGE_NORMAL else 1)
        a, sa = oracle.stage_run(tree, st)
        b, sb = oracle.stage_from_lists(x, y, z, nbr, st)
        assert np.array_equal(sa, sb)
        for u, v in zip(a, b):
            assert np.array_equal(u, v, equal_nan=True)
