"""GPU parity on the edge cases the reference's tests cover: the upstream
known-answer clouds run through the C ABI, clouds smaller than the neighbor
count, coincident points (zero covariance), every stage and mode against
PDAL's CPU semantics, the golden fixtures written by the reference's own
nanoflann / Eigen, and argument errors."""
import os

import numpy as np
import pytest

import gpupdal_amd as g
import oracle_lib as ol
from helpers import small_cloud, truth_errors

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INFO = g.ST_TIE | g.ST_INCOMPLETE  # informational bits: the row was re-solved exactly


def gpu_stage(ctx, x, y, z, st, init=np.nan):
    (cols,), (status,) = ctx.cloud(x, y, z).run([st], init=init)
    return cols, status


def oracle_stage(oracle, x, y, z, st):
    ost = ol.stage(st.stage, st.knn, st.stride, st.mode, st.features, st.always_up,
                   st.normalize)
    return oracle.stage_run(oracle.kd3(x, y, z), ost, threads=8, init=np.nan)


def assert_stage_close(ctx, oracle, x, y, z, st, rtol=1e-5, atol=1e-6):
    """Statuses equal to PDAL's CPU stage and columns close to it.  PDAL
    narrows the demeaned coordinates to float before the covariance product
    (MathUtils.cpp:177-189), so on general inputs the two differ by ~1e-7
    relative; on integer inputs they agree to the last bits."""
    cols, status = gpu_stage(ctx, x, y, z, st)
    ocols, ostatus = oracle_stage(oracle, x, y, z, st)
    assert np.array_equal(status & ~np.uint8(INFO), ostatus)
    names = ([ol.FEATURE_NAMES[b] for b in range(10) if st.features & (1 << b)]
             if st.stage == g.STAGE_COVFEAT else [None] * len(cols))
    for name, c, oc in zip(names, cols, ocols):
        if name == "Omnivariance" and "Scattering" in names:
            # On an exactly degenerate neighborhood (three points are always
            # planar) the smallest eigenvalue is rounding noise, and
            # Omnivariance = cbrt(sqrt(l0 l1 l2)) raises it to the 1/6th
            # power: ~1e-3 in any FP64 implementation.  Compared only where
            # the smallest eigenvalue is well resolved even under PDAL's
            # float-narrowed covariance (Scattering > 0.1).
            keep = ocols[names.index("Scattering")] > 0.1
            c, oc = c[keep], oc[keep]
        np.testing.assert_allclose(c, oc, rtol=rtol, atol=atol, equal_nan=True,
                                   err_msg=str(name))


# ------------------------------------------------- upstream KAT clouds --
def test_kat_line_plane_scattering(ctx, oracle):
    # CovarianceFeaturesTest.cpp:46-188 on the device path.
    line = (np.zeros(3), np.zeros(3), np.array([0.0, 1, 2]), 3)
    plane = (np.array([0.0, 1, 1, 0]), np.array([0.0, 0, 1, 1]), np.zeros(4), 10)
    scat = (np.array([0.0, 1, 1, 0]), np.array([0.0, 0, 1, 1]), np.array([0.0, 0, 1, 2]), 10)
    for x, y, z, knn in (line, plane, scat):
        st = g.stage(g.STAGE_COVFEAT, knn, features=g.F_ALL10)
        assert_stage_close(ctx, oracle, x, y, z, st, rtol=1e-12, atol=1e-14)
    cols, _ = gpu_stage(ctx, *scat[:3], g.stage(g.STAGE_COVFEAT, 10, features=g.F_ALL10))
    want = {4: 0.458, 5: 0.864, 6: 0.488, 7: 1.583, 8: 0.095, 9: 0.492}
    for ci, v in want.items():
        np.testing.assert_allclose(cols[ci].astype(np.float32), v, atol=0.001)


@pytest.mark.parametrize("pts,want", [
    ([(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)], (0, 0, 1)),   # NormalFilterTest XYPlane
    ([(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1)], (0, 1, 0)),   # XZPlane
    ([(0, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1)], (1, 0, 0)),   # YZPlane
])
def test_kat_normal_planes(ctx, oracle, pts, want):
    p = np.array(pts, float)
    cols, status = gpu_stage(ctx, p[:, 0], p[:, 1], p[:, 2], g.stage(g.STAGE_NORMAL, 3))
    for c, w in zip(cols[:3], want):
        np.testing.assert_allclose(c.astype(np.float32), w, atol=1e-7)
    np.testing.assert_allclose(cols[3], 0, atol=1e-7)


# ------------------------------------------------------ tiny clouds --
@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 16, 17, 40])
def test_clouds_smaller_than_the_neighborhood(ctx, oracle, n):
    # KD3 clips k to n and pads a strided result with id 0 (KDImpl.hpp:351-381);
    # the stages then see min(n, knn+1) neighbors.
    rng = np.random.default_rng(n)
    x, y, z = rng.uniform(0, 3, n), rng.uniform(0, 3, n), rng.uniform(0, 1, n)
    for st in (g.stage(g.STAGE_NORMAL, 8), g.stage(g.STAGE_EIGENVALUES, 16, normalize=1),
               g.stage(g.STAGE_COVFEAT, 10, features=g.F_ALL10),
               g.stage(g.STAGE_COVFEAT, 10, stride=3)):
        assert_stage_close(ctx, oracle, x, y, z, st)
    ids, status = ctx.cloud(x, y, z).knn(5, stride=3)
    tree = oracle.kd3(x, y, z)
    assert np.array_equal(ids, tree.neighbors_all(5, stride=3))


# ---------------------------------------------------- coincident points --
def test_zero_covariance_rows_keep_prior_values(ctx, oracle):
    x, y, z = small_cloud(4000, seed=40)
    # 30 copies of one point: with knn=16 those rows see only the copies.
    x = np.concatenate([x, np.full(30, x[0] + 0.123)])
    y = np.concatenate([y, np.full(30, y[0] + 0.456)])
    z = np.concatenate([z, np.full(30, 500.0)])
    for st in (g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_EIGENVALUES, 16),
               g.stage(g.STAGE_COVFEAT, 16)):
        cols, status = gpu_stage(ctx, x, y, z, st, init=-7.0)
        zero = (status & g.ST_ZERO_COV) != 0
        assert np.count_nonzero(zero[-30:]) == 30
        for c in cols:
            assert np.all(c[zero] == -7.0)
        _, ostatus = oracle_stage(oracle, x, y, z, st)
        assert np.array_equal(status & ~np.uint8(INFO), ostatus)


# ------------------------------------------------ stages and modes ----
@pytest.mark.parametrize("st", [
    g.stage(g.STAGE_EIGENVALUES, 8),
    g.stage(g.STAGE_EIGENVALUES, 12, normalize=1),
    g.stage(g.STAGE_COVFEAT, 10, mode=g.MODE_RAW, features=g.F_ALL10),
    g.stage(g.STAGE_COVFEAT, 10, mode=g.MODE_NORMALIZED, features=g.F_ALL10),
    g.stage(g.STAGE_COVFEAT, 8, stride=2, features=g.F_ALL10),
    g.stage(g.STAGE_NORMAL, 8, always_up=0),
], ids=["eig8", "eig12norm", "raw", "normalized", "stride2", "normal_free"])
def test_stage_vs_truth(ctx, oracle, st):
    """Every column within 1e-5 relative of the extended-precision truth on
    the rows whose eigenvalues are separated (gap > 1e-3); columns that depend
    on a rounding-level smallest eigenvalue are checked absolutely
    (helpers.truth_errors)."""
    x, y, z = small_cloud(20000, seed=50)
    cols, status = gpu_stage(ctx, x, y, z, st)
    stride = 1 if st.stage == g.STAGE_NORMAL else st.stride
    nbr = oracle.kd3(x, y, z).neighbors_all(st.knn + 1, stride=stride, threads=8)
    ost = ol.stage(st.stage, st.knn, st.stride, st.mode, st.features, st.always_up,
                   st.normalize)
    tcols, tstat, gap, ratio = oracle.truth_from_lists(x, y, z, nbr, ost, with_ratio=True)
    ok = ((status & 0x3C) == 0) & (tstat == 0) & (gap > 1e-3)
    assert np.count_nonzero(ok) > 0.9 * len(x)
    errs = truth_errors(ost, cols, tcols, ok, ratio)
    assert max(errs.values()) <= 1.0, errs


def test_stage_matches_pdal_path_statuses(ctx, oracle):
    # The same rows are skipped (zero covariance, too few neighbors) as in
    # PDAL's CPU stage.
    x, y, z = small_cloud(15000, seed=51, dup_frac=0.05)
    for st in (g.stage(g.STAGE_NORMAL, 2), g.stage(g.STAGE_EIGENVALUES, 4),
               g.stage(g.STAGE_COVFEAT, 4)):
        _, status = gpu_stage(ctx, x, y, z, st)
        _, ostatus = oracle_stage(oracle, x, y, z, st)
        assert np.array_equal(status & ~np.uint8(INFO), ostatus)


# ------------------------------------------ golden (reference nanoflann) --
def test_golden_kd3_lists(ctx):
    f = np.load(os.path.join(GOLDEN, "kd3_ties.npz"), allow_pickle=False)
    cl = ctx.cloud(f["x"], f["y"], f["z"])
    for key, k, stride in (("ids_k8_s1", 8, 1), ("ids_k17_s1", 17, 1), ("ids_k9_s2", 9, 2)):
        want = f[key].astype(np.uint32)
        ids, st = cl.knn(k, stride=stride)
        flagged = (st & INFO) != 0
        # Exactly re-solved rows and every strided row follow KD3 order; the
        # other rows hold the same set (ordered by distance, then id).
        exact = flagged | (stride > 1)
        assert np.array_equal(ids[exact], want[exact]), key
        assert np.array_equal(np.sort(ids, axis=1), np.sort(want, axis=1)), key
    t = cl.kdtree()
    assert np.array_equal(t["vind"], f["vind"])


def test_golden_eigen_bits(ctx):
    f = np.load(os.path.join(GOLDEN, "eigen3.npz"), allow_pickle=False)
    info, ev, vec = ctx.debug_eigen(f["mats"])
    assert np.array_equal(info, f["info"])
    assert np.array_equal(ev, f["ev"])
    assert np.array_equal(vec, f["vec"])


# ---------------------------------------------------------------- errors --
def test_argument_errors(ctx):
    x, y, z = small_cloud(1000, seed=60)
    cl = ctx.cloud(x, y, z)
    with pytest.raises(g.GfError) as e:
        cl.run([g.stage(g.STAGE_COVFEAT, 400, stride=3)])  # 401*3 > 1024
    assert e.value.code == -3
    with pytest.raises(g.GfError) as e:
        cl.run([g.stage(7, 8)])
    assert e.value.code == -1
    with pytest.raises(g.GfError):
        cl.knn(600, stride=2)  # 1200 > 1024
    # The context stays usable after an error.
    cols, status = gpu_stage(ctx, x, y, z, g.stage(g.STAGE_NORMAL, 8))
    assert np.count_nonzero(status & 0x3C) < len(x)


# ------------------------------------------------------------- autotuner --
def test_autotune_changes_speed_not_results(ctx, tmp_path):
    x, y, z = small_cloud(60000, seed=70)
    st = [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 16)]
    cl = ctx.cloud(x, y, z)
    before = cl.run(st, init=np.nan)
    f, ms = ctx.autotune(cl, st, factors=(0.45, 0.6, 0.9))
    assert f in (0.45, 0.6, 0.9) and ms > 0
    after = cl.run(st, init=np.nan)
    for sb, sa in zip(before[1], after[1]):
        assert np.array_equal(sb & ~np.uint8(INFO), sa & ~np.uint8(INFO))
    for cb, ca in zip(before[0][0] + before[0][1], after[0][0] + after[0][1]):
        np.testing.assert_allclose(ca, cb, rtol=1e-12, atol=1e-15, equal_nan=True)
    path = tmp_path / "tune.txt"
    ctx.save_tuning(path)
    text = path.read_text()
    rows = [l.split() for l in text.splitlines() if not l.startswith("#")]
    assert len(rows) >= 1 and all(len(r) == 6 for r in rows)
    ctx.load_tuning(path)


@pytest.mark.parametrize("shape", [dict(q_per_wave=16, knn_waves=1),
                                   dict(q_per_wave=32, knn_waves=2, group_z=0),
                                   dict(q_per_wave=64, knn_waves=4, group_z=4)])
def test_search_shapes_change_no_result(ctx, shape):
    # The autotuner's search-shape knobs: every row's status and columns
    # are the default shape's.
    x, y, z = small_cloud(60000, seed=71, dup_frac=0.02)
    st = [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 24, features=g.F_ALL10)]
    ref = ctx.cloud(x, y, z).run(st, init=np.nan)
    c2 = g.Context(0)
    for k, v in shape.items():
        c2.set_tuning(k, v)
    got = c2.cloud(x, y, z).run(st, init=np.nan)
    c2.close()
    for sa, sb in zip(ref[1], got[1]):
        assert np.array_equal(sa & ~np.uint8(INFO), sb & ~np.uint8(INFO))
    for ca, cb in zip(ref[0][0] + ref[0][1], got[0][0] + got[0][1]):
        np.testing.assert_allclose(cb, ca, rtol=1e-12, atol=1e-15, equal_nan=True)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("cloud", ["dups", "lattice", "dense"])
def test_set_selection_equals_sorting(ctx, cloud):
    # Feature neighborhoods with stride 1 are kept as sets, found by the
    # exact-count threshold selection; switching it off (set_select=0) runs
    # the sorting selection that lists use.  Every status bit and every
    # column bit must agree: the same sets, summed in the same (d2, id)
    # order, flagged (ties, uncertified rows) the same way.
    if cloud == "dups":
        x, y, z = small_cloud(60000, seed=73, dup_frac=0.03)
    elif cloud == "lattice":
        g1 = np.arange(40, dtype=np.float64)
        X, Y, Z = np.meshgrid(g1, g1, np.arange(5, dtype=np.float64), indexing="ij")
        x, y, z = X.ravel() * 0.5 + 500000.0, Y.ravel() * 0.5 + 4000000.0, Z.ravel() * 0.5
    else:
        x, y, z = small_cloud(80000, seed=74, side=40.0)
    for st in ([g.stage(g.STAGE_NORMAL, 8), g.stage(g.STAGE_COVFEAT, 8, features=g.F_ALL10)],
               [g.stage(g.STAGE_EIGENVALUES, 16), g.stage(g.STAGE_COVFEAT, 16)],
               [g.stage(g.STAGE_NORMAL, 32)]):
        fast = ctx.cloud(x, y, z).run(st, init=np.nan)
        c2 = g.Context(0)
        c2.set_tuning("set_select", 0)
        slow = c2.cloud(x, y, z).run(st, init=np.nan)
        c2.close()
        for sa, sb in zip(fast[1], slow[1]):
            assert np.array_equal(sa, sb), (cloud, st[0].knn)
        for ca, cb in zip([c for cs in fast[0] for c in cs], [c for cs in slow[0] for c in cs]):
            assert _bits_equal(ca, cb), (cloud, st[0].knn)


def test_tree_beside_the_search_repeats_exactly(ctx, oracle):
    """Large clouds build the exact tree on a second stream from the moment
    their bounds are known, while the index build and the search run on the
    main stream: repeated runs give bit-identical columns and statuses, the
    same as with the tree built after the search, and the tree is
    nanoflann's."""
    n = 1_000_000
    x, y, z = g.synth_cloud(n, tile=7, side=250.0)
    cl = ctx.cloud(x, y, z)
    st = [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 16)]
    ref = None
    try:
        for overlap in (131072, 131072, 131072, 0):
            ctx.set_tuning("kd_overlap_min", overlap)
            cl.invalidate()
            (nc, fc), (ns, fs) = cl.run(st, init=np.nan)
            assert ctx.stats()["ms_tree"] > 0.0
            got = [np.asarray(c) for c in nc + fc] + [ns, fs]
            if ref is None:
                ref = got
            else:
                for u, v in zip(ref, got):
                    assert np.array_equal(u, v, equal_nan=True)
    finally:
        ctx.set_tuning("kd_overlap_min", 131072)
    sub = slice(0, 200_000)
    t = ctx.cloud(x[sub], y[sub], z[sub]).kdtree()
    e = oracle.kd3(x[sub], y[sub], z[sub]).export()
    assert np.array_equal(t["vind"].astype(np.uint64), e["vind"])
