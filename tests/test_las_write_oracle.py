"""The LAS oracle's Extra Bytes reader and writer, pinned on the CPU.

Reader: the reference's assertions on extrabytes.las
(test/unit/io/LasReaderTest.cpp:330-392): Colors0..2 (uint16) equal the
point's Red, Green, Blue; Flags0/1 (int8) its return number and number of
returns; Time (uint64) its GPS time within 1 (it was truncated).  Writer:
what it writes reads back through that reader (names, types, values to the
bit, including NaN and signed zero), the source bytes of every record and
every VLR are unchanged, the header stays valid, an existing Extra Bytes VLR
is extended in place, and undocumented source bytes are declared."""
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import las_oracle as lo  # noqa: E402

LAS = os.path.join(HERE, "golden", "las")


def load(name):
    return open(os.path.join(LAS, name), "rb").read()


def base_fields(buf):
    h = lo.parse_header(buf)
    n, S = h["point_count"], h["point_size"]
    rec = np.frombuffer(buf, np.uint8, n * S, h["point_offset"]).reshape(n, S)
    rgb = np.ascontiguousarray(rec[:, 28:34]).view("<u2").reshape(n, 3)
    ret = rec[:, 14]
    gps = np.ascontiguousarray(rec[:, 20:28]).view("<f8").reshape(n)
    return rgb, ret & 7, (ret >> 3) & 7, gps


def test_reader_on_reference_extrabytes():
    buf = load("extrabytes.las")
    ex = lo.read_extra(buf)
    rgb, rn, nr, gps = base_fields(buf)
    assert len(gps) == 1065
    for i, c in enumerate(("Colors0", "Colors1", "Colors2")):
        assert ex[c].dtype == np.uint16 and np.array_equal(ex[c], rgb[:, i])
    assert ex["Flags0"].dtype == np.int8 and np.array_equal(ex["Flags0"], rn)
    assert np.array_equal(ex["Flags1"], nr)
    assert ex["Time"].dtype == np.uint64
    assert np.all(np.abs(ex["Time"].astype(np.float64) - gps) <= 1.0)


def columns(n, ncols, seed):
    rng = np.random.default_rng(seed)
    cols = [rng.normal(size=n) * 10.0 ** rng.integers(-5, 6) for _ in range(ncols)]
    if n > 3:
        cols[0][:4] = [np.nan, -0.0, np.inf, 5e-324]
    return cols


@pytest.mark.parametrize("name", ["1.2-with-color.las", "extrabytes.las", "test1_4.las",
                                  "simple.las", "no-points.las"])
def test_writer_round_trip(name):
    src = load(name)
    h = lo.parse_header(src)
    n = h["point_count"]
    names = ["Linearity", "Planarity", "AVeryLongFeatureNameThatIsTruncatedAt32"]
    cols = columns(n, 3, 7)
    out = lo.write_extra(src, names, cols, ["desc a", "", "d"])
    h2 = lo.parse_header(out)
    assert lo.validate(h2, len(out)) is None
    assert h2["point_size"] == h["point_size"] + 24
    assert h2["point_count"] == n
    ex = lo.read_extra(out)
    want_names = ["Linearity", "Planarity", names[2][:32]]
    for nm, c in zip(want_names, cols):
        assert ex[nm].dtype == np.float64
        assert np.array_equal(ex[nm].view(np.uint64), c.view(np.uint64))
    # every earlier extra dimension still reads the same
    for nm, v in lo.read_extra(src).items():
        assert np.array_equal(ex[nm], v)
    # source records unchanged in their first point_size bytes
    S, S2 = h["point_size"], h2["point_size"]
    a = np.frombuffer(src, np.uint8, n * S, h["point_offset"]).reshape(n, S)
    b = np.frombuffer(out, np.uint8, n * S2, h2["point_offset"]).reshape(n, S2)
    assert np.array_equal(a, b[:, :S])
    # the other VLRs are unchanged; one Extra Bytes VLR
    v1 = [v for v in lo.vlrs(src) if not (v["user_id"] == "LASF_Spec" and v["record_id"] == 4)]
    v2 = [v for v in lo.vlrs(out) if not (v["user_id"] == "LASF_Spec" and v["record_id"] == 4)]
    assert [(v["user_id"], v["record_id"], v["data"]) for v in v1] == \
           [(v["user_id"], v["record_id"], v["data"]) for v in v2]
    assert sum(v["user_id"] == "LASF_Spec" and v["record_id"] == 4 for v in lo.vlrs(out)) == 1
    # the bytes after the points (EVLRs) follow the wider records
    tail = len(src) - (h["point_offset"] + n * S)
    assert out[len(out) - tail:] == src[len(src) - tail:]
    if h["version_minor"] >= 4:
        (e1,) = struct.unpack_from("<Q", src, 235)
        (e2,) = struct.unpack_from("<Q", out, 235)
        if e1:
            assert out[e2:e2 + 60] == src[e1:e1 + 60]
    # the descriptions as given
    descs = {d[0]: d[3] for d in lo.extra_dims(out)}
    assert descs["Linearity"] == "desc a" and descs["Planarity"] == ""


def test_writer_declares_undocumented_bytes():
    # 1.2-with-color.las has format 3 records of 34 bytes; widen them by 3
    # undocumented bytes first: the new column must be read at offset 37.
    src = load("1.2-with-color.las")
    h = lo.parse_header(src)
    n, S = h["point_count"], h["point_size"]
    rec = np.frombuffer(src, np.uint8, n * S, h["point_offset"]).reshape(n, S)
    wide = np.zeros((n, S + 3), np.uint8)
    wide[:, :S] = rec
    wide[:, S:] = 0xAB
    hdr = bytearray(src[:h["point_offset"]])
    struct.pack_into("<H", hdr, 105, S + 3)
    src2 = bytes(hdr) + wide.tobytes() + src[h["point_offset"] + n * S:]
    col = columns(n, 1, 3)
    out = lo.write_extra(src2, ["Omnivariance"], col)
    dims = lo.extra_dims(out)
    assert [(d[0], d[2]) for d in dims] == [("Omnivariance", S + 3)]
    assert np.array_equal(lo.read_extra(out)["Omnivariance"].view(np.uint64),
                          col[0].view(np.uint64))


def test_writer_typed_columns():
    # Coplanar (uint8), OptimalKNN (uint64), Rank (uint8) as PDAL types
    # them, a float and a signed column, with values outside the ranges.
    src = load("simple.las")
    n = lo.parse_header(src)["point_count"]
    rng = np.random.default_rng(2)
    cols = [rng.integers(0, 2, n).astype(float), rng.integers(10, 15, n).astype(float),
            rng.normal(size=n) * 1e3, rng.normal(size=n) * 300, rng.normal(size=n)]
    cols[3][:4] = [np.nan, 1e30, -1e30, -2.7]
    types = [1, 7, 9, 2, 10]
    names = ["Coplanar", "OptimalKNN", "F", "I8", "D"]
    out = lo.write_extra(src, names, cols, types=types)
    ex = lo.read_extra(out)
    assert ex["Coplanar"].dtype == np.uint8 and np.array_equal(ex["Coplanar"], cols[0])
    assert ex["OptimalKNN"].dtype == np.uint64 and np.array_equal(ex["OptimalKNN"], cols[1])
    assert np.array_equal(ex["F"], cols[2].astype(np.float32))
    assert ex["I8"].dtype == np.int8 and list(ex["I8"][:4]) == [0, 127, -128, -2]
    assert np.array_equal(ex["D"], cols[4])
    assert lo.parse_header(out)["point_size"] == lo.parse_header(src)["point_size"] + 1 + 8 + 4 + 1 + 8
