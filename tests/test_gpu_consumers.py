"""The other kNN consumers on the GPU, through the C-ABI.

nndistance and the outlier filter depend only on each row's ascending
squared distances, which do not depend on how ties were ordered: the device
values equal the oracle's to the bit (FP64 distances in KD3Index's operation
order, the same sums), and the outlier threshold and flags equal the
oracle's exactly.  The reference's own known answers (the integer grid of
NNDistanceTest.cpp, the autzen outlier counts of OldPCLBlockTest.cpp) hold on
the device.  approximatecoplanar and estimaterank are decisions on
eigenvalues: the device's covariance is exact FP64 while PDAL's narrows the
demeaned coordinates to float, so the two agree on every row whose margin
from the thresholds is clear of that rounding (1e-5 of the largest
eigenvalue), and such rows are all but a few."""
import os
import sys

import numpy as np
import pytest

import gpupdal_amd as g
import oracle_lib as ol
from helpers import small_cloud

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import las_oracle  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle():
    return ol.Oracle()


def test_nndistance_reference_grid(ctx):
    gr = np.arange(10, dtype=np.float64)
    x, y, z = (a.ravel() for a in np.meshgrid(gr, gr, gr, indexing="ij"))
    cl = ctx.cloud(x, y, z)
    assert np.all(cl.nndistance(3, "kth") == 1.0)
    corner = np.isin(x, (0, 9)) & np.isin(y, (0, 9)) & np.isin(z, (0, 9))
    assert np.all(cl.nndistance(4, "kth")[corner] == np.sqrt(2))
    d = cl.nndistance(6, "avg")
    e = np.isin(x, (0, 9)).astype(int) + np.isin(y, (0, 9)) + np.isin(z, (0, 9))
    assert np.all(d[e == 3] == (1 + np.sqrt(2)) / 2.0)
    assert np.all(d[e == 2] == (2 + np.sqrt(2)) / 3.0)
    assert np.all(d[e == 1] == (5 + np.sqrt(2)) / 6.0)
    assert np.all(d[e == 0] == 1.0)
    cl.close()


@pytest.mark.parametrize("n,k,dup", [(20000, 10, 0.03), (20000, 1, 0.0), (5000, 40, 0.05),
                                     (6, 10, 0.0), (1, 3, 0.0)])
def test_nndistance_equals_oracle(ctx, oracle, n, k, dup):
    print(f"nndistance n={n} k={k} dup={dup}", flush=True)
    x, y, z = small_cloud(n, seed=31, dup_frac=dup)
    t = oracle.kd3(x, y, z)
    cl = ctx.cloud(x, y, z)
    for mode, om in (("kth", ol.NND_KTH), ("avg", ol.NND_AVG)):
        got = cl.nndistance(k, mode)
        want = oracle.nn_distance(t, k, om)
        assert np.array_equal(got, want, equal_nan=True), mode
    cl.close()


def autzen_xyz():
    buf = open(os.path.join(HERE, "golden", "las", "autzen-point-format-3.las"), "rb").read()
    return las_oracle.read_xyz(buf)


def test_outlier_reference_counts(ctx):
    x, y, z = autzen_xyz()
    cl = ctx.cloud(x, y, z)
    for mean_k, mult, inliers in ((2, 1.5, 96), (5, 0.0, 63)):
        f, thr = cl.outlier("statistical", mean_k=mean_k, multiplier=mult)
        assert len(x) - int(f.sum()) == inliers
    for r, mk, inliers in ((200.0, 1, 60), (100.0, 2, 3)):
        f, thr = cl.outlier("radius", radius=r, min_k=mk)
        assert thr is None and len(x) - int(f.sum()) == inliers
    cl.close()


@pytest.mark.parametrize("seed", [5, 6])
def test_outlier_equals_oracle(ctx, oracle, seed):
    x, y, z = small_cloud(30000, seed=seed, dup_frac=0.02)
    t = oracle.kd3(x, y, z)
    cl = ctx.cloud(x, y, z)
    for mean_k, mult in ((8, 2.0), (12, 1.25), (3, 0.0)):
        print(f"outlier statistical seed={seed} mean_k={mean_k}", flush=True)
        f, thr = cl.outlier("statistical", mean_k=mean_k, multiplier=mult)
        wf, wthr, _ = oracle.outlier_statistical(t, mean_k, mult)
        assert thr == wthr
        assert np.array_equal(f, wf)
    for r, mk in ((0.5, 3), (1.0, 8)):
        print(f"outlier radius seed={seed} r={r}", flush=True)
        f, _ = cl.outlier("radius", radius=r, min_k=mk)
        assert np.array_equal(f, oracle.outlier_radius(t, r, mk))
    cl.close()


def test_nndistance_full_size_sampled_rows(ctx, oracle):
    # 4M points: sampled rows against brute force over the whole cloud.
    x, y, z = g.synth_cloud(4_000_000, tile=3)
    cl = ctx.cloud(x, y, z)
    d = cl.nndistance(10, "avg")
    cl.close()
    rng = np.random.default_rng(0)
    rows = np.sort(rng.choice(len(x), 200, replace=False)).astype(np.uint64)
    _, d2 = oracle.brute_knn(x, y, z, 11, queries=rows)
    r = np.sqrt(d2)
    avg = np.zeros(len(rows))
    for j in range(1, 11):
        avg = avg + r[:, j]
    assert np.array_equal(d[rows], avg / 10)


def _margins(x, y, z, nbr):
    p = np.stack([x[nbr], y[nbr], z[nbr]], axis=-1)
    p = p - p.mean(axis=1, keepdims=True)
    return np.linalg.eigvalsh(np.einsum("nki,nkj->nij", p, p) / (nbr.shape[1] - 1))


@pytest.mark.parametrize("knn", [8, 16])
def test_coplanar_and_rank_vs_pdal(ctx, oracle, knn):
    x, y, z = small_cloud(30000, seed=17, dup_frac=0.01)
    stages = [g.stage(g.STAGE_COPLANAR, knn), g.stage(g.STAGE_RANK, knn),
              g.stage(g.STAGE_RANK, knn, p1=1e-3)]
    cl = ctx.cloud(x, y, z)
    cols, status = cl.run(stages, init=np.nan)
    cl.close()
    t = oracle.kd3(x, y, z)
    nbr = t.neighbors_all(knn, 1, threads=8).astype(np.int64)
    ev = _margins(x, y, z, nbr)
    scale = np.maximum(np.abs(ev[:, 2]), 1e-300)
    for st_, c_, s_ in zip(stages, cols, status):
        ost = ol.stage(st_.stage, st_.knn, p1=st_.p1, p2=st_.p2)
        pc, ps = oracle.stage_run(t, ost, threads=8, init=np.nan)
        skip = s_ & ~np.uint8(g.ST_TIE | g.ST_INCOMPLETE)  # search flags are reports only
        assert np.array_equal(skip, ps)
        if st_.stage == g.STAGE_COPLANAR:
            a = ev[:, 1] - st_.p1 * ev[:, 0]
            b = st_.p2 * ev[:, 1] - ev[:, 2]
            clear = (np.abs(a) > 1e-5 * scale) & (np.abs(b) > 1e-5 * scale)
        else:
            cut = np.abs(ev[:, 2]) * float(np.float32(st_.p1))
            clear = np.all(np.abs(np.abs(ev) - cut[:, None]) > 1e-5 * scale[:, None], axis=1)
            clear |= np.all(ev == 0, axis=1)
        ok = clear & (skip == 0)
        assert ok.sum() > 0.98 * len(x)
        assert np.array_equal(c_[0][ok], pc[0][ok])
        assert np.all(np.isnan(c_[0][skip != 0]))
        assert not np.any(np.isnan(c_[0][skip == 0]))
