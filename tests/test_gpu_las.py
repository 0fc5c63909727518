"""LAS boundary on the GPU: the device decoder against the oracle on every
fixture (bit for bit), on large synthetic files with odd record lengths,
features from a decoded cloud against the same points uploaded as FP64, the
extra-bytes packer against the oracle, and the decoder's refusals."""
import glob
import os
import sys

import numpy as np
import pytest

import gpupdal_amd as g
from gpupdal_amd.las import las_image

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import las_oracle as lo  # noqa: E402

pytestmark = pytest.mark.gpu

LAS = os.path.join(ROOT, "tests", "golden", "las")
FILES = sorted(glob.glob(os.path.join(LAS, "*.las")) +
               glob.glob(os.path.join(LAS, "permutations", "*.las")))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_decode_matches_oracle(ctx, path):
    with open(path, "rb") as f:
        data = f.read()
    h = lo.parse_header(data)
    if lo.validate(h, len(data)):
        with pytest.raises(g.GfError) as ei:
            ctx.cloud_from_las(data)
        assert ei.value.code == g.GF_ERR_ARG
        return
    cl = ctx.cloud_from_las(data)
    assert cl.n == h["point_count"]
    want = lo.read_xyz(data)
    got = cl.xyz()
    for a, b in zip(got, want):
        assert np.array_equal(bits(a), bits(b))
    cl.close()


def test_laz_is_refused(ctx):
    with open(os.path.join(LAS, "simple.laz"), "rb") as f:
        data = f.read()
    with pytest.raises(g.GfError) as ei:
        ctx.cloud_from_las(data)
    assert ei.value.code == g.GF_ERR_UNSUPPORTED and "LAZ" in str(ei.value)


@pytest.mark.parametrize("fmt,minor,extra,n", [(1, 2, 0, 300_000), (3, 4, 3, 250_001),
                                               (0, 2, 1, 77_777), (7, 4, 29, 100_003)])
def test_decode_synthetic_odd_records(ctx, fmt, minor, extra, n):
    """Record lengths 28, 37, 21 and 65 bytes: runs of records start at every
    byte alignment, and workgroups stage their runs through LDS."""
    x, y, z = g.synth_cloud(n, tile=7, side=300.0)
    img, _ = las_image(x, y, z, offset=(500000.0, 4000000.0, 0.0), point_format=fmt,
                       version_minor=minor, extra_bytes=extra)
    cl = ctx.cloud_from_las(img)
    want = lo.read_xyz(img)
    for a, b in zip(cl.xyz(), want):
        assert np.array_equal(bits(a), bits(b))
    assert cl.las_info["point_size"] == lo.BASE_SIZE[fmt] + extra
    cl.close()


def test_features_from_las_equal_fp64_upload(ctx):
    with open(os.path.join(LAS, "1.2-with-color.las"), "rb") as f:
        data = f.read()
    x, y, z = lo.read_xyz(data)
    stages = [g.stage(g.STAGE_NORMAL, 8), g.stage(g.STAGE_COVFEAT, 10)]
    a = ctx.cloud_from_las(data).run(stages, init=-1.0)
    b = ctx.cloud(x, y, z).run(stages, init=-1.0)
    for ca, cb in zip(a[0], b[0]):
        for u, v in zip(ca, cb):
            assert np.array_equal(bits(u), bits(v))
    for u, v in zip(a[1], b[1]):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("stride,offset,n", [(61, 34, 20_001), (48, 24, 20_000), (45, 28, 3)])
def test_pack_extra_matches_oracle(ctx, stride, offset, n):
    rng = np.random.default_rng(stride)
    ncols = (stride - offset) // 8
    cols = [rng.standard_normal(n) * 1e3 for _ in range(ncols)]
    base = rng.integers(0, 256, size=n * stride, dtype=np.uint8)
    got = ctx.las_pack_extra(cols, base.copy(), stride, offset)
    want = lo.pack_extra(base.copy(), cols, stride, offset)
    assert np.array_equal(got, want)
    # bytes outside the extra dimensions are the record's own
    keep = np.ones((n, stride), bool)
    keep[:, offset:offset + 8 * ncols] = False
    assert np.array_equal(got.reshape(n, stride)[keep], base.reshape(n, stride)[keep])


def test_pack_extra_rejects_overflowing_fields(ctx):
    with pytest.raises(g.GfError):
        ctx.las_pack_extra([np.zeros(4)], np.zeros(4 * 20, np.uint8), 20, 16)
