"""LAS boundary on the CPU: the oracle's header parse and XYZ decode against
the reference's own assertions on its test files, the C ABI's header parser
against the oracle on every fixture and on damaged headers, and the host LAS
writer against the oracle."""
import glob
import os
import struct
import sys

import numpy as np
import pytest

import gpupdal_amd as g
from gpupdal_amd.las import las_image

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import las_oracle as lo  # noqa: E402

LAS = os.path.join(ROOT, "tests", "golden", "las")
FILES = sorted(glob.glob(os.path.join(LAS, "*.la[sz]")) +
               glob.glob(os.path.join(LAS, "permutations", "*.las")))


def read(name):
    with open(os.path.join(LAS, name), "rb") as f:
        return f.read()


def test_oracle_header_kats():
    """test/unit/io/LasReaderTest.cpp:104-138 (simple.las)."""
    h = lo.parse_header(read("simple.las"))
    assert h["magic"] == b"LASF" and h["file_source_id"] == 0
    assert (h["version_major"], h["version_minor"]) == (1, 2)
    assert (h["creation_doy"], h["creation_year"]) == (0, 0)
    assert h["header_size"] == 227 and h["point_format"] == 3
    assert h["point_count"] == 1065 and not h["compressed"]
    assert h["scale"] == (0.01, 0.01, 0.01) and h["offset"] == (0.0, 0.0, 0.0)
    # EXPECT_DOUBLE_EQ: within 4 ulps.
    assert h["max"] == pytest.approx((638982.55, 853535.43, 586.38), rel=4 * 2.0**-52, abs=0)
    assert h["min"] == pytest.approx((635619.85, 848899.70, 406.59), rel=4 * 2.0**-52, abs=0)
    assert h["points_by_return"] == [925, 114, 21, 5, 0]
    assert lo.parse_header(read("simple.laz"))["compressed"]


def test_oracle_point_kats():
    """test/unit/Support.cpp:373-385 through LasReaderTest.cpp:178-195
    (1.2-with-color.las points 0-2 and 100-102) and the format permutations
    of LasReaderTest.cpp:228-235 (first point of each file)."""
    x, y, z = lo.read_xyz(read("1.2-with-color.las"))
    want = {0: (637012.24, 849028.31, 431.66), 1: (636896.33, 849087.70, 446.39),
            2: (636784.74, 849106.66, 426.71), 100: (636661.06, 849854.13, 424.90),
            101: (636568.18, 850179.49, 441.80), 102: (636554.63, 850040.03, 499.11)}
    for i, (a, b, c) in want.items():
        assert (x[i], y[i], z[i]) == pytest.approx((a, b, c), rel=0, abs=1e-9)
    for f in glob.glob(os.path.join(LAS, "permutations", "1.[012]_[0-3].las")):
        with open(f, "rb") as fh:
            px, py, pz = lo.read_xyz(fh.read())
        assert (px[0], py[0], pz[0]) == pytest.approx((470692.44, 4602888.90, 16.0),
                                                      rel=0, abs=1e-9)


def test_oracle_validation_messages():
    """PDAL's header errors (io/private/las/Header.cpp:118-144)."""
    good = read("simple.las")
    h = lo.parse_header(good)
    assert lo.validate(h, len(good)) is None
    assert "Invalid point count: 1065" in lo.validate(
        lo.parse_header(read("permutations/1.2-no-points.las")), 229)


def _damaged():
    b = bytearray(read("simple.las"))
    out = {"magic": bytes(b"LASX" + b[4:])}
    fmt4 = bytearray(b)
    fmt4[104] = 4
    out["format4"] = bytes(fmt4)
    small = bytearray(b)
    struct.pack_into("<H", small, 105, 20)
    out["record_too_short"] = bytes(small)
    v15 = bytearray(b)
    v15[25] = 5
    out["version_1_5"] = bytes(v15)
    far = bytearray(b)
    struct.pack_into("<I", far, 96, 10**6)
    out["offset_past_end"] = bytes(far)
    out["truncated_points"] = bytes(b[:-100])
    out["short_header"] = bytes(b[:200])
    return out


@pytest.mark.parametrize("name", sorted(_damaged().keys()))
def test_abi_parse_rejects_like_pdal(name):
    data = _damaged()[name]
    try:
        want = lo.validate(lo.parse_header(data), len(data))
    except lo.LasError as e:
        want = str(e)
    assert want
    with pytest.raises(g.GfError) as ei:
        g.las_parse(data)
    assert want in str(ei.value)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_abi_parse_matches_oracle(path):
    with open(path, "rb") as f:
        data = f.read()
    h = lo.parse_header(data)
    err = lo.validate(h, len(data))
    if err:
        with pytest.raises(g.GfError):
            g.las_parse(data)
        return
    info = g.las_parse(data)
    for k in ("version_major", "version_minor", "point_format", "header_size", "point_size",
              "point_offset", "point_count", "scale", "offset", "max", "min"):
        assert info[k] == h[k], k
    assert bool(info["compressed"]) == h["compressed"]


def test_host_writer_roundtrip():
    rng = np.random.default_rng(5)
    x = np.round(rng.uniform(500000, 500500, 5000), 2)
    y = np.round(rng.uniform(4000000, 4000500, 5000), 2)
    z = np.round(rng.uniform(0, 30, 5000), 2)
    for fmt, minor, extra in ((1, 2, 0), (3, 4, 3), (6, 4, 0), (0, 2, 16)):
        img, _ = las_image(x, y, z, offset=(500000.0, 4000000.0, 0.0), point_format=fmt,
                           version_minor=minor, extra_bytes=extra)
        h = lo.parse_header(img)
        assert lo.validate(h, len(img)) is None
        assert h["point_size"] == lo.BASE_SIZE[fmt] + extra
        rx, ry, rz = lo.read_xyz(img)
        assert np.allclose(rx, x, rtol=0, atol=1e-6) and np.allclose(rz, z, rtol=0, atol=1e-6)
