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


# ------------------------------------------------------------ extra bytes --
# LAS 1.4 Extra Bytes VLR (user "LASF_Spec", record 4): 192-byte descriptors
# (io/private/las/Utils.hpp:195-208), data types 1-10 = u8, i8, u16, i16,
# u32, i32, u64, i64, f32, f64 and 11-30 two or three of them
# (io/private/las/Utils.cpp:54-58, 209-227), 0 = `options` undocumented
# bytes.  PDAL names the fields of a multi-field descriptor name0, name1, ...
# and lays the described bytes out after the base record in descriptor order.
VLR_HEADER = 54
EB_SIZE = 192
EB_TYPES = [None, "u1", "i1", "<u2", "<i2", "<u4", "<i4", "<u8", "<i8", "<f4", "<f8"]


def vlrs(buf):
    """The VLRs of a LAS file image: dicts with offset (of the VLR header),
    user_id, record_id, description and data (bytes)."""
    h = parse_header(buf)
    b = memoryview(buf)
    out, at = [], h["header_size"]
    for _ in range(h["vlr_count"]):
        user = bytes(b[at + 2:at + 18]).split(b"\0")[0].decode("ascii", "replace")
        rid, length = struct.unpack_from("<HH", b, at + 18)
        desc = bytes(b[at + 22:at + 54]).split(b"\0")[0].decode("ascii", "replace")
        out.append(dict(offset=at, user_id=user, record_id=rid, description=desc,
                        data=bytes(b[at + VLR_HEADER:at + VLR_HEADER + length])))
        at += VLR_HEADER + length
    return out


def _descriptor_size(dtype, options):
    if dtype == 0:
        return options
    fields = (dtype - 1) // 10 + 1
    return fields * np.dtype(EB_TYPES[(dtype - 1) % 10 + 1]).itemsize


def extra_dims(buf):
    """The extra dimensions an Extra Bytes VLR declares: (name, numpy dtype,
    byte offset in the record, description) per field, undocumented bytes
    skipped."""
    h = parse_header(buf)
    offset = BASE_SIZE[h["point_format"]]
    out = []
    for v in vlrs(buf):
        if v["user_id"] != "LASF_Spec" or v["record_id"] != 4:
            continue
        d = v["data"]
        for i in range(len(d) // EB_SIZE):
            e = d[i * EB_SIZE:(i + 1) * EB_SIZE]
            dtype, options = e[2], e[3]
            name = e[4:36].split(b"\0")[0].decode("ascii", "replace")
            desc = e[160:192].split(b"\0")[0].decode("ascii", "replace")
            if dtype != 0:
                fields = (dtype - 1) // 10 + 1
                t = np.dtype(EB_TYPES[(dtype - 1) % 10 + 1])
                for f in range(fields):
                    out.append((name + (str(f) if fields > 1 else ""), t,
                                offset + f * t.itemsize, desc))
            offset += _descriptor_size(dtype, options)
        break
    return out


def read_extra(buf):
    """{name: values} of the declared extra dimensions."""
    h = parse_header(buf)
    n, size = h["point_count"], h["point_size"]
    rec = np.frombuffer(buf, dtype=np.uint8, count=n * size,
                        offset=h["point_offset"]).reshape(n, size)
    return {name: np.ascontiguousarray(rec[:, off:off + t.itemsize]).view(t).reshape(n)
            for name, t, off, _ in extra_dims(buf)}


def typed(col, dtype):
    """A double column as LAS data type 1..10: integer types take the value
    rounded toward zero and held to the type's range, NaN as 0."""
    t = np.dtype(EB_TYPES[dtype])
    col = np.asarray(col, dtype=np.float64)
    if t.kind == "f":
        return col.astype(t)
    info = np.iinfo(t)
    lo, hi = float(info.min), float(info.max)
    if hi >= 2.0 ** 63:
        hi = 18446744073709549568.0 if info.min == 0 else 9223372036854774784.0
    v = np.where(np.isnan(col), 0.0, np.clip(np.trunc(col), lo, hi))
    return v.astype(t)


def write_extra(src, names, cols, descriptions=None, types=None):
    """writers.las extra_dims (io/LasWriter.cpp:583-596,
    io/private/las/Utils.cpp:244-267) as a new file image: the columns (each
    as its data type, 10 = double by default) after each source record,
    descriptors (zero-padded 32-byte name and description) appended to the
    source's Extra Bytes VLR or in a new one after the last VLR; source bytes
    no descriptor covers declared first as undocumented ones; header
    offsets, VLR count and record length updated, the 1.3 waveform and 1.4
    EVLR offsets moved."""
    h = parse_header(src)
    descriptions = descriptions or [""] * len(names)
    types = list(types) if types is not None else [10] * len(cols)
    vl = vlrs(src)
    eb = next((i for i, v in enumerate(vl)
               if v["user_id"] == "LASF_Spec" and v["record_id"] == 4), None)
    S, n = h["point_size"], h["point_count"]
    described = BASE_SIZE[h["point_format"]]
    if eb is not None:
        d = vl[eb]["data"]
        for i in range(len(d) // EB_SIZE):
            described += _descriptor_size(d[i * EB_SIZE + 2], d[i * EB_SIZE + 3])

    def descriptor(dtype, options, name, desc):
        e = bytearray(EB_SIZE)
        e[2], e[3] = dtype, options
        nb, db = name.encode("ascii")[:32], desc.encode("ascii")[:32]
        e[4:4 + len(nb)] = nb
        e[160:160 + len(db)] = db
        return bytes(e)

    add = b""
    gap = S - described
    while gap > 0:
        g = min(gap, 255)
        add += descriptor(0, g, "", "")
        gap -= g
    for name, desc, t in zip(names, descriptions, types):
        add += descriptor(t, 0, name, desc)
    grow = len(add) + (0 if eb is not None else VLR_HEADER)
    S2 = S + sum(np.dtype(EB_TYPES[t]).itemsize for t in types)
    new_offset = h["point_offset"] + grow
    src_points_end = h["point_offset"] + n * S
    after = new_offset + n * S2 - src_points_end
    hdr = bytearray(src[:h["header_size"]])
    struct.pack_into("<I", hdr, 96, new_offset)
    struct.pack_into("<I", hdr, 100, h["vlr_count"] + (0 if eb is not None else 1))
    struct.pack_into("<H", hdr, 105, S2)
    if h["version_minor"] >= 3 and h["header_size"] >= 235:
        (wave,) = struct.unpack_from("<Q", src, 227)
        if wave:
            struct.pack_into("<Q", hdr, 227, wave + after)
    if h["version_minor"] >= 4 and h["header_size"] >= 243:
        (ev,) = struct.unpack_from("<Q", src, 235)
        if ev:
            struct.pack_into("<Q", hdr, 235, ev + after)
    out = bytearray(hdr)
    for i, v in enumerate(vl):
        head = bytearray(src[v["offset"]:v["offset"] + VLR_HEADER])
        data = v["data"]
        if i == eb:
            data = data + add
            struct.pack_into("<H", head, 20, len(data))
        out += head + data
    vlr_end = vl[-1]["offset"] + VLR_HEADER + len(vl[-1]["data"]) if vl else h["header_size"]
    if eb is None:
        head = bytearray(VLR_HEADER)
        head[2:11] = b"LASF_Spec"
        struct.pack_into("<HH", head, 18, 4, len(add))
        head[22:40] = b"Extra Bytes Record"
        out += head + add
    out += src[vlr_end:h["point_offset"]]
    rec = np.frombuffer(src, dtype=np.uint8, count=n * S, offset=h["point_offset"]).reshape(n, S)
    wide = np.zeros((n, S2), np.uint8)
    wide[:, :S] = rec
    at = S
    for col, t in zip(cols, types):
        size = np.dtype(EB_TYPES[t]).itemsize
        v = np.ascontiguousarray(typed(col, t)).view(np.uint8).reshape(n, size)
        wide[:, at:at + size] = v
        at += size
    out += wide.tobytes()
    out += src[src_points_end:]
    return bytes(out)
