"""LAS point records in and out, restated in numpy.

Test infrastructure (the oracle): only tests/, smoke() and bench.py's CPU
leg use it, as the checker of the device decoder; nothing shipped imports it.

Follows PDAL's LAS reader and writer:
  * public header block, field order and sizes: io/private/las/Header.cpp:38-75
    (227 bytes for 1.0-1.2, 235 for 1.3, 375 for 1.4: Header.hpp:75-77);
  * point format = bits & 0xF, LAZ bit 0x80: Header.hpp:78-79, 122-128;
  * point count: the 64-bit count for 1.4, else the legacy 32-bit count:
    Header.hpp:137-138;
  * base record sizes per format: Header.cpp:15-36; supported formats
    0-3 and 6-8: Utils.hpp:93-107;
  * validation and its messages: Header.cpp:118-144;
  * X, Y, Z = int32 * scale + offset in FP64, each product rounded before
    the sum: io/LasReader.cpp:926-932 (formats 0-5) and 1008-1014 (6-10);
  * extra dimensions of type double written after the base record of each
    point, little-endian (writers.las extra_dims).
"""
import struct

import numpy as np

HEADER_12 = 227
HEADER_13 = 235
HEADER_14 = 375
FORMAT_MASK = 0x0F
COMPRESSION_BIT = 0x80
BASE_SIZE = {0: 20, 1: 28, 2: 26, 3: 34, 6: 30, 7: 36, 8: 38}


class LasError(ValueError):
    pass


def parse_header(buf):
    """The header fields of a LAS file image (bytes-like)."""
    b = memoryview(buf)
    if len(b) < HEADER_12:
        raise LasError("Invalid LAS header: file shorter than the header.")
    h = {}
    h["magic"] = bytes(b[0:4])
    h["file_source_id"], h["global_encoding"] = struct.unpack_from("<HH", b, 4)
    h["version_major"], h["version_minor"] = b[24], b[25]
    h["creation_doy"], h["creation_year"] = struct.unpack_from("<HH", b, 90)
    h["header_size"], h["point_offset"], h["vlr_count"] = struct.unpack_from("<HII", b, 94)
    bits, h["point_size"], legacy = struct.unpack_from("<BHI", b, 104)
    h["point_format"] = bits & FORMAT_MASK
    h["compressed"] = bool(bits & COMPRESSION_BIT)
    h["points_by_return"] = list(struct.unpack_from("<5I", b, 111))
    h["scale"] = struct.unpack_from("<3d", b, 131)
    h["offset"] = struct.unpack_from("<3d", b, 155)
    mxx, mnx, mxy, mny, mxz, mnz = struct.unpack_from("<6d", b, 179)
    h["max"] = (mxx, mxy, mxz)
    h["min"] = (mnx, mny, mnz)
    h["point_count"] = legacy
    minor = h["version_minor"]
    if minor >= 3:
        if len(b) < (HEADER_14 if minor >= 4 else HEADER_13):
            raise LasError("Invalid LAS header: file shorter than the header.")
        if minor >= 4:
            (h["point_count"],) = struct.unpack_from("<Q", b, 247)
            h["points_by_return"] = list(struct.unpack_from("<15Q", b, 255))
    return h


def validate(h, file_size):
    """The first of PDAL's header errors, or None (Header.cpp:118-144)."""
    if h["magic"] != b"LASF":
        return ("Invalid file signature. Was expecting 'LASF', Check the first four "
                " bytes of the file.")
    if not h["compressed"] and h["point_offset"] > file_size:
        return "Invalid point offset - exceeds file size."
    if not h["compressed"] and (h["point_offset"] + h["point_count"] * h["point_size"]
                                > file_size):
        return ("Invalid point count: %d. Number of points too large for file size."
                % h["point_count"])
    if h["header_size"] > file_size:
        return "Invalid VLR offset - exceeds file size."
    if h["point_format"] not in BASE_SIZE:
        return "Unsupported LAS input point format: %d." % h["point_format"]
    if h["point_size"] < BASE_SIZE[h["point_format"]]:
        return ("Invalid point size of %d. Less than minimum required for point format %d."
                % (h["point_size"], h["point_format"]))
    if h["version_major"] != 1 or h["version_minor"] > 4:
        return "Unsupported LAS version '%d.%d'." % (h["version_major"], h["version_minor"])
    return None


def read_xyz(buf):
    """(x, y, z) FP64 arrays of an uncompressed LAS file image."""
    h = parse_header(buf)
    err = validate(h, len(buf))
    if err:
        raise LasError(err)
    if h["compressed"]:
        raise LasError("LAZ-compressed point data")
    n, size = h["point_count"], h["point_size"]
    rec = np.frombuffer(buf, dtype=np.uint8, count=n * size,
                        offset=h["point_offset"]).reshape(n, size)
    xyz = np.ascontiguousarray(rec[:, 0:12]).view("<i4").reshape(n, 3)
    out = []
    for d in range(3):
        v = xyz[:, d].astype(np.float64) * h["scale"][d]  # rounded product
        out.append(v + h["offset"][d])
    return tuple(out)


def pack_extra(records, cols, stride, offset):
    """Writes the double columns into each record of `records` (uint8, n *
    stride bytes) at byte `offset`, one after another, little-endian."""
    n = len(cols[0]) if cols else 0
    view = records.reshape(n, stride)
    for c, col in enumerate(cols):
        raw = np.ascontiguousarray(np.asarray(col, dtype="<f8")).view(np.uint8).reshape(n, 8)
        view[:, offset + 8 * c: offset + 8 * c + 8] = raw
    return records
