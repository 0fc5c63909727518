"""writers.las extra_dims on the device (gf_las_write_extra): the file
image equals the LAS oracle's writer byte for byte on the reference's LAS
fixtures (a 1.2 file, a 1.4 file with an Extra Bytes VLR of its own, a file
with no points), the device decoder reads the written file's X, Y, Z back
unchanged, and a whole pass -- LAS in, features on the device, LAS out --
carries the columns to the bit."""
import os
import sys

import numpy as np
import pytest

import gpupdal_amd as g

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import las_oracle as lo  # noqa: E402

LAS = os.path.join(HERE, "golden", "las")


@pytest.fixture(scope="module")
def ctx():
    c = g.Context(0)
    yield c
    c.close()


def load(name):
    return open(os.path.join(LAS, name), "rb").read()


@pytest.mark.parametrize("name", ["1.2-with-color.las", "extrabytes.las", "test1_4.las",
                                  "simple.las", "no-points.las", "epsg_4326.las"])
def test_image_equals_oracle(ctx, name):
    src = load(name)
    n = lo.parse_header(src)["point_count"]
    rng = np.random.default_rng(11)
    cols = [rng.normal(size=n) for _ in range(4)]
    if n > 2:
        cols[1][:3] = [np.nan, -0.0, np.inf]
    names = ["Linearity", "Planarity", "Scattering", "Verticality"]
    descs = ["linearity", "", "scattering of the neighborhood", "v"]
    got = ctx.las_write_extra(src, names, cols, descs)
    want = lo.write_extra(src, names, cols, descs)
    assert got == want
    if n:
        cl = ctx.cloud_from_las(got)
        xyz = cl.xyz()
        cl.close()
        for a, b in zip(xyz, lo.read_xyz(src)):
            assert np.array_equal(a, b)


def test_las_in_features_out(ctx):
    src = load("1.2-with-color.las")
    cl = ctx.cloud_from_las(src)
    st = [g.stage(g.STAGE_COVFEAT, 10, features=g.F_DIMENSIONALITY)]
    (cols,), (status,) = cl.run(st, init=np.nan)
    cl.close()
    names = ["Linearity", "Planarity", "Scattering", "Verticality"]
    out = ctx.las_write_extra(src, names, cols)
    ex = lo.read_extra(out)
    for nm, c in zip(names, cols):
        assert np.array_equal(ex[nm].view(np.uint64), c.view(np.uint64))


def test_write_errors(ctx):
    with pytest.raises(g.GfError):
        ctx.las_write_extra(load("simple.laz"), ["A"], [np.zeros(1065)])
    with pytest.raises(g.GfError):
        ctx.las_write_extra(b"not a las file" * 20, ["A"], [np.zeros(3)])


def test_typed_columns_equal_oracle(ctx):
    src = load("1.2-with-color.las")
    n = lo.parse_header(src)["point_count"]
    rng = np.random.default_rng(5)
    cols = [rng.integers(0, 2, n).astype(float), rng.integers(10, 15, n).astype(float),
            rng.normal(size=n) * 1e3, rng.normal(size=n) * 300, rng.normal(size=n) * 1e12,
            rng.normal(size=n) * 1e5, rng.normal(size=n) * 1e5, rng.normal(size=n) * 1e5,
            rng.normal(size=n) * 1e19, rng.normal(size=n)]
    for c in cols[2:]:
        c[:5] = [np.nan, np.inf, -np.inf, 1e300, -0.5]
    types = list(range(1, 11))
    names = ["T%d" % t for t in types]
    got = ctx.las_write_extra(src, names, cols, types=types)
    assert got == lo.write_extra(src, names, cols, types=types)
