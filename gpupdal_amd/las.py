"""LAS file images on the host: the input of the device decoder
(`Context.cloud_from_las`) and the output side (base records plus the extra
bytes `Context.las_pack_extra` writes).

The header layout is LAS 1.2/1.4's public header block as PDAL writes it
(io/private/las/Header.cpp:77-116); X, Y, Z are stored as
round((v - offset) / scale) in int32, the writer's quantization.
"""
import struct

import numpy as np

BASE_SIZE = {0: 20, 1: 28, 2: 26, 3: 34, 6: 30, 7: 36, 8: 38}


def las_image(x, y, z, scale=(0.01, 0.01, 0.01), offset=(0.0, 0.0, 0.0), point_format=1,
              version_minor=2, extra_bytes=0):
    """Bytes of an uncompressed LAS file holding the points (other point
    fields zero); each record has `extra_bytes` zero bytes after the base
    record.  Returns (image, record_offset_of_first_point)."""
    x, y, z = (np.asarray(v, dtype=np.float64) for v in (x, y, z))
    n = len(x)
    if point_format not in BASE_SIZE:
        raise ValueError("point format must be one of 0-3, 6-8")
    size = BASE_SIZE[point_format] + extra_bytes
    header = 375 if version_minor >= 4 else 227
    h = bytearray(header)
    h[0:4] = b"LASF"
    h[24], h[25] = 1, version_minor
    h[26:26 + 15] = b"gpupdal_amd    "
    struct.pack_into("<HII", h, 94, header, header, 0)
    legacy = n if (n < 2**32 and not (version_minor >= 4 and point_format >= 6)) else 0
    struct.pack_into("<BHI", h, 104, point_format, size, legacy)
    struct.pack_into("<3d", h, 131, *scale)
    struct.pack_into("<3d", h, 155, *offset)
    lo = [float(v.min()) if n else 0.0 for v in (x, y, z)]
    hi = [float(v.max()) if n else 0.0 for v in (x, y, z)]
    struct.pack_into("<6d", h, 179, hi[0], lo[0], hi[1], lo[1], hi[2], lo[2])
    if version_minor >= 4:
        struct.pack_into("<Q", h, 247, n)
    rec = np.zeros((n, size), dtype=np.uint8)
    ints = np.empty((n, 3), dtype="<i4")
    for d, v in enumerate((x, y, z)):
        ints[:, d] = np.round((v - offset[d]) / scale[d]).astype(np.int64).astype("<i4")
    rec[:, 0:12] = ints.view(np.uint8).reshape(n, 12)
    return bytes(h) + rec.tobytes(), header
