"""gpupdal_amd -- MI355X-native kNN-neighborhood features (PDAL's
filters.covariancefeatures / filters.normal / filters.eigenvalues path).

The compute lives in libgpupdal_amd.so (HIP kernels for gfx950 behind the C
ABI in include/gpupdal_amd.h).  This module is a thin ctypes binding for
Python callers, the tests and the benchmark.  There is no CPU fallback: if the
library is missing, importing the binding raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpupdal_amd.so")

GF_OK = 0
GF_ERR_ARG, GF_ERR_HIP, GF_ERR_UNSUPPORTED, GF_ERR_OOM, GF_ERR_STATE = -1, -2, -3, -4, -5
GF_ERR_TIMEOUT = -6
WHERE_HOST, WHERE_DEVICE = 0, 1
STAGE_COVFEAT, STAGE_NORMAL, STAGE_EIGENVALUES, STAGE_COPLANAR, STAGE_RANK = 0, 1, 2, 3, 4
MODE_RAW, MODE_SQRT, MODE_NORMALIZED = 0, 1, 2
F_DIMENSIONALITY = 0xF
F_ALL10 = 0x3FF
ST_TIE, ST_INCOMPLETE, ST_ZERO_COV, ST_EIGEN_FAIL = 1, 2, 4, 8
ST_FEATURE_INVALID, ST_FEW_NEIGHBORS = 16, 32

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)


class GfStage(ctypes.Structure):
    _fields_ = [("stage", ctypes.c_int32), ("knn", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("features", ctypes.c_uint32), ("always_up", ctypes.c_int32),
                ("normalize", ctypes.c_int32), ("min_k", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("radius", ctypes.c_double),
                ("p1", ctypes.c_double), ("p2", ctypes.c_double)]


class GfLasInfo(ctypes.Structure):
    _fields_ = [("version_major", ctypes.c_uint8), ("version_minor", ctypes.c_uint8),
                ("point_format", ctypes.c_uint8), ("compressed", ctypes.c_uint8),
                ("header_size", ctypes.c_uint16), ("point_size", ctypes.c_uint16),
                ("point_offset", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("point_count", ctypes.c_uint64), ("scale", ctypes.c_double * 3),
                ("offset", ctypes.c_double * 3), ("max", ctypes.c_double * 3),
                ("min", ctypes.c_double * 3), ("ms_decode", ctypes.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}
        for k in ("scale", "offset", "max", "min"):
            d[k] = tuple(d[k])
        return d


class GfStats(ctypes.Structure):
    _fields_ = [("points", ctypes.c_uint64), ("tie_rows", ctypes.c_uint64),
                ("incomplete_rows", ctypes.c_uint64), ("ms_index", ctypes.c_double),
                ("ms_search", ctypes.c_double), ("ms_features", ctypes.c_double),
                ("ms_exact", ctypes.c_double),
                ("cell", ctypes.c_double), ("cell_z", ctypes.c_double),
                ("grid_x", ctypes.c_uint32), ("grid_y", ctypes.c_uint32),
                ("grid_z", ctypes.c_uint32), ("key_bits", ctypes.c_uint32),
                ("search_passes", ctypes.c_uint64), ("candidates", ctypes.c_uint64),
                ("merges", ctypes.c_uint64), ("ms_tree", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("deep_rows", ctypes.c_uint64),
                ("graph", ctypes.c_uint64)]


class GfSynth(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("side", ctypes.c_double), ("ground_frac", ctypes.c_double),
                ("roof_frac", ctypes.c_double), ("wire_frac", ctypes.c_double),
                ("noise", ctypes.c_double), ("quantum", ctypes.c_double),
                ("offset_x", ctypes.c_double), ("offset_y", ctypes.c_double)]


EXPORTED = [
    "gf_ctx_create", "gf_ctx_destroy", "gf_last_error", "gf_sync", "gf_set_tuning",
    "gf_cloud_create", "gf_cloud_release", "gf_cloud_size", "gf_stage_columns",
    "gf_run", "gf_last_stats", "gf_knn", "gf_debug_eigen", "gf_debug_grid",
    "gf_synth_defaults", "gf_tile_seed", "gf_synth_cloud", "gf_dev_alloc", "gf_dev_free",
    "gf_memcpy", "gf_cloud_invalidate", "gf_debug_kdtree", "gf_debug_copy",
    "gf_batch_create", "gf_batch_set_tile", "gf_batch_size", "gf_batch_tile_start",
    "gf_batch_invalidate", "gf_batch_run", "gf_batch_release",
    "gf_server_create", "gf_server_submit", "gf_server_wait", "gf_server_request_info",
    "gf_server_release", "gf_server_destroy",
    "gf_autotune", "gf_tuning_save", "gf_tuning_load",
    "gf_las_parse", "gf_cloud_create_las", "gf_las_pack_extra", "gf_las_write_extra",
    "gf_debug_cloud_xyz",
    "gf_optimal_neighborhood", "gf_debug_grid_params", "gf_run_noalloc",
    "gf_radius_counts", "gf_nndistance", "gf_outlier", "gf_run_rowk",
]
NOALLOC_MAX_POINTS = 65536
NND_KTH, NND_AVG = 0, 1
OUTLIER_STATISTICAL, OUTLIER_RADIUS = 0, 1


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP library is required; there is no CPU fallback)")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.gf_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.gf_ctx_destroy.argtypes = [vp]
    L.gf_last_error.restype = ctypes.c_char_p
    L.gf_last_error.argtypes = [vp]
    L.gf_sync.argtypes = [vp]
    L.gf_set_tuning.argtypes = [vp, ctypes.c_char_p, ctypes.c_double]
    L.gf_cloud_create.argtypes = [vp, vp, vp, vp, ctypes.c_uint64, ctypes.c_int,
                                  ctypes.POINTER(vp)]
    L.gf_cloud_release.argtypes = [vp]
    L.gf_cloud_size.restype = ctypes.c_uint64
    L.gf_cloud_size.argtypes = [vp]
    L.gf_cloud_invalidate.argtypes = [vp]
    L.gf_stage_columns.argtypes = [ctypes.POINTER(GfStage)]
    L.gf_run.argtypes = [vp, vp, ctypes.POINTER(GfStage), ctypes.c_int,
                         ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int]
    L.gf_last_stats.argtypes = [vp, ctypes.POINTER(GfStats)]
    L.gf_radius_counts.argtypes = [vp, vp, ctypes.c_double, vp, ctypes.c_int]
    L.gf_run_noalloc.argtypes = [vp, vp, vp, vp, ctypes.c_uint64, ctypes.POINTER(GfStage),
                                 ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.gf_knn.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, ctypes.c_int]
    L.gf_run_rowk.argtypes = [vp, vp, ctypes.POINTER(GfStage), ctypes.c_int, vp,
                              ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int]
    L.gf_nndistance.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_int, vp, ctypes.c_int]
    L.gf_outlier.argtypes = [vp, vp, ctypes.c_int, ctypes.c_uint32, ctypes.c_double,
                             ctypes.c_double, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_double)]
    L.gf_debug_eigen.argtypes = [vp, _dp, ctypes.c_uint64, _dp, _dp, _i32p]
    L.gf_debug_grid.argtypes = [vp, vp, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, _u32p]
    L.gf_debug_grid_params.argtypes = [vp, vp, ctypes.c_uint32, _dp, _u32p]
    L.gf_debug_kdtree.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.gf_debug_copy.argtypes = [vp, vp, vp, ctypes.c_uint64]
    L.gf_las_parse.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(GfLasInfo), ctypes.c_char_p,
                               ctypes.c_uint64]
    L.gf_cloud_create_las.argtypes = [vp, vp, ctypes.c_uint64, ctypes.POINTER(GfLasInfo),
                                      ctypes.POINTER(vp)]
    L.gf_las_write_extra.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_char_p),
                                     ctypes.POINTER(ctypes.c_char_p), vp, ctypes.POINTER(vp),
                                     ctypes.c_int, vp, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint64)]
    L.gf_las_pack_extra.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_int, ctypes.c_uint64, vp,
                                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    L.gf_debug_cloud_xyz.argtypes = [vp, vp, vp, vp, vp]
    L.gf_optimal_neighborhood.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp,
                                          ctypes.c_int]
    L.gf_dev_alloc.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.gf_dev_free.argtypes = [vp, vp]
    L.gf_memcpy.argtypes = [vp, vp, vp, ctypes.c_uint64, ctypes.c_int]
    L.gf_synth_defaults.argtypes = [ctypes.POINTER(GfSynth)]
    L.gf_tile_seed.restype = ctypes.c_uint64
    L.gf_tile_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.gf_synth_cloud.argtypes = [ctypes.POINTER(GfSynth), _dp, _dp, _dp]
    L.gf_batch_create.argtypes = [vp, ctypes.c_uint32, vp, ctypes.POINTER(vp)]
    L.gf_batch_set_tile.argtypes = [vp, ctypes.c_uint32, vp, vp, vp, ctypes.c_int]
    L.gf_batch_size.restype = ctypes.c_uint64
    L.gf_batch_size.argtypes = [vp]
    L.gf_batch_tile_start.restype = ctypes.c_uint64
    L.gf_batch_tile_start.argtypes = [vp, ctypes.c_uint32]
    L.gf_batch_invalidate.argtypes = [vp]
    L.gf_batch_run.argtypes = [vp, vp, ctypes.POINTER(GfStage), ctypes.c_int, vp,
                               ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int]
    L.gf_batch_release.argtypes = [vp]
    L.gf_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.POINTER(vp)]
    L.gf_server_submit.argtypes = [vp, vp, vp, vp, ctypes.c_uint64, ctypes.POINTER(GfStage),
                                   ctypes.c_int, ctypes.c_int32, ctypes.POINTER(vp),
                                   ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint64)]
    L.gf_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_double]
    L.gf_server_request_info.argtypes = [vp, ctypes.c_uint64, _dp, _u32p, _u32p]
    L.gf_server_release.argtypes = [vp, ctypes.c_uint64]
    L.gf_server_destroy.argtypes = [vp]
    L.gf_autotune.argtypes = [vp, vp, ctypes.POINTER(GfStage), ctypes.c_int, _dp, ctypes.c_int,
                              _dp, _dp]
    L.gf_tuning_save.argtypes = [vp, ctypes.c_char_p]
    L.gf_tuning_load.argtypes = [vp, ctypes.c_char_p]
    return L


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load_library()
    return _LIB


class GfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gpupdal_amd error {code}: {msg}")
        self.code = code


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------ generator --
def synth_params(n, seed=None, tile=0, side=500.0, ground_frac=0.7, roof_frac=0.0,
                 wire_frac=0.0, noise=0.05, quantum=0.01):
    L = lib()
    p = GfSynth()
    L.gf_synth_defaults(ctypes.byref(p))
    p.seed = L.gf_tile_seed(0x5EED, tile) if seed is None else seed
    p.n = n
    p.side = side
    p.ground_frac = ground_frac
    p.roof_frac = roof_frac
    p.wire_frac = wire_frac
    p.noise = noise
    p.quantum = quantum
    return p


def las_parse(data):
    """The header of a LAS file image (bytes-like) as a dict; GfError with
    PDAL's validation message for a file PDAL's reader rejects.  Host only."""
    L = lib()
    buf = np.frombuffer(data, dtype=np.uint8)
    info = GfLasInfo()
    msg = ctypes.create_string_buffer(256)
    rc = L.gf_las_parse(buf.ctypes.data if len(buf) else None, len(buf), ctypes.byref(info),
                        msg, 256)
    if rc != GF_OK:
        raise GfError(rc, "gf_las_parse: " + msg.value.decode())
    return info.as_dict()


def synth_cloud(n, **kw):
    """Seeded synthetic ALS tile (x, y, z float64 arrays)."""
    L = lib()
    p = synth_params(n, **kw)
    x = np.empty(n, np.float64)
    y = np.empty(n, np.float64)
    z = np.empty(n, np.float64)
    rc = L.gf_synth_cloud(ctypes.byref(p), x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                          z.ctypes.data_as(_dp))
    if rc != GF_OK:
        raise GfError(rc, "gf_synth_cloud")
    return x, y, z


def stage(kind, knn, stride=1, mode=MODE_SQRT, features=F_DIMENSIONALITY, always_up=1,
          normalize=0, radius=0.0, min_k=3, p1=None, p2=None):
    """A stage of a chain.  p1/p2 default to PDAL's: approximatecoplanar
    thresh1 25, thresh2 6; estimaterank thresh 0.01."""
    if p1 is None:
        p1 = {STAGE_COPLANAR: 25.0, STAGE_RANK: 0.01}.get(kind, 0.0)
    if p2 is None:
        p2 = 6.0 if kind == STAGE_COPLANAR else 0.0
    return GfStage(kind, knn, stride, mode, features, always_up, normalize, min_k, 0, radius,
                   p1, p2)


def stage_columns(st):
    return lib().gf_stage_columns(ctypes.byref(st))


# -------------------------------------------------------------- context --
class Context:
    def __init__(self, device=0):
        self.L = lib()
        h = ctypes.c_void_p()
        rc = self.L.gf_ctx_create(device, ctypes.byref(h))
        if rc != GF_OK:
            raise GfError(rc, "gf_ctx_create failed (is a GPU visible?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.gf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != GF_OK:
            raise GfError(rc, f"{what}: {self.L.gf_last_error(self.h).decode()}")

    def set_tuning(self, key, value):
        self._check(self.L.gf_set_tuning(self.h, key.encode(), float(value)), "gf_set_tuning")

    def sync(self):
        self._check(self.L.gf_sync(self.h), "gf_sync")

    def stats(self):
        s = GfStats()
        self.L.gf_last_stats(self.h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in GfStats._fields_}

    def cloud(self, x, y, z):
        return Cloud(self, x, y, z)

    def cloud_from_las(self, data):
        """A resident cloud decoded on the device from the point records of an
        uncompressed LAS file image (bytes-like); its header is `.las_info`."""
        buf = np.frombuffer(data, dtype=np.uint8)
        info = GfLasInfo()
        h = ctypes.c_void_p()
        self._check(self.L.gf_cloud_create_las(self.h, buf.ctypes.data, len(buf),
                                               ctypes.byref(info), ctypes.byref(h)),
                    "gf_cloud_create_las")
        cl = Cloud.__new__(Cloud)
        cl.ctx, cl.h, cl.n = self, h, int(info.point_count)
        cl.las_info = info.as_dict()
        return cl

    def las_pack_extra(self, cols, records, stride, offset):
        """Writes the double columns into each stride-byte record of the
        uint8 array `records` at byte `offset` (host arrays)."""
        cols = [_c64(c) for c in cols]
        n = len(cols[0]) if cols else 0
        cp = (ctypes.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
        self._check(self.L.gf_las_pack_extra(self.h, cp, len(cols), n, records.ctypes.data,
                                             stride, offset, WHERE_HOST), "gf_las_pack_extra")
        return records

    def las_write_extra(self, data, names, cols, descriptions=None, types=None):
        """writers.las extra_dims: a new LAS file image (bytes) from the
        uncompressed image `data` with the columns (host arrays of doubles,
        one value per point) appended to every record, each as its LAS data
        type (types: 1..10 = u8 .. f64; None: all f64), and described in the
        Extra Bytes VLR."""
        src = np.frombuffer(data, dtype=np.uint8)
        cols = [_c64(c) for c in cols]
        k = len(cols)
        if k != len(names):
            raise ValueError("one name per column")
        descriptions = list(descriptions) if descriptions is not None else [""] * k
        nm = (ctypes.c_char_p * k)(*[str(v).encode("ascii") for v in names])
        ds = (ctypes.c_char_p * k)(*[str(v).encode("ascii") for v in descriptions])
        cp = (ctypes.c_void_p * max(1, k))(*[c.ctypes.data for c in cols])
        tp = None if types is None else np.ascontiguousarray(types, dtype=np.uint8)
        tptr = None if tp is None else tp.ctypes.data
        size = ctypes.c_uint64()
        self._check(self.L.gf_las_write_extra(self.h, src.ctypes.data, len(src), k, nm, ds, tptr,
                                              cp, WHERE_HOST, None, 0, ctypes.byref(size)),
                    "gf_las_write_extra")
        out = np.empty(size.value, np.uint8)
        self._check(self.L.gf_las_write_extra(self.h, src.ctypes.data, len(src), k, nm, ds, tptr,
                                              cp, WHERE_HOST, out.ctypes.data, len(out),
                                              ctypes.byref(size)), "gf_las_write_extra")
        return out.tobytes()

    def autotune(self, cloud, stages, factors=(0.4, 0.5, 0.6, 0.75, 0.9)):
        """Times the run for each column-size factor, then each search-shape
        knob (z grouping, queries per wave, waves per workgroup) one at a
        time; the fastest settings are kept for the cloud's (density, k)
        bucket.  Returns (factor, ms of the fastest run)."""
        stages = list(stages)
        arr = (GfStage * len(stages))(*stages)
        f = np.ascontiguousarray(factors, dtype=np.float64)
        bf = ctypes.c_double()
        bm = ctypes.c_double()
        self._check(self.L.gf_autotune(self.h, cloud.h, arr, len(stages), f.ctypes.data_as(_dp),
                                       len(f), ctypes.byref(bf), ctypes.byref(bm)), "gf_autotune")
        return bf.value, bm.value

    def save_tuning(self, path):
        self._check(self.L.gf_tuning_save(self.h, str(path).encode()), "gf_tuning_save")

    def load_tuning(self, path):
        self._check(self.L.gf_tuning_load(self.h, str(path).encode()), "gf_tuning_load")

    def batch(self, tiles):
        """A batch of tiles [(x, y, z), ...] resident on the device."""
        return Batch(self, tiles)

    def run_noalloc(self, x, y, z, n, stages, cols, status):
        """gf_run_noalloc: the fused stage run on device arrays (x, y, z:
        n doubles each; cols, status as Cloud.run_device) with no device
        allocation; n <= NOALLOC_MAX_POINTS."""
        stages = list(stages)
        arr = (GfStage * len(stages))(*stages)
        cp = (ctypes.c_void_p * max(1, len(cols)))(*cols)
        sp = (ctypes.c_void_p * len(status))(*status)
        return self.L.gf_run_noalloc(self.h, x, y, z, n, arr, len(stages), cp, sp)

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._check(self.L.gf_dev_alloc(self.h, nbytes, ctypes.byref(p)), "gf_dev_alloc")
        return p.value

    def free(self, ptr):
        self._check(self.L.gf_dev_free(self.h, ptr), "gf_dev_free")

    def copy(self, dst, src, count):
        """Plain device copy of count doubles (HBM counter calibration)."""
        self._check(self.L.gf_debug_copy(self.h, dst, src, count), "gf_debug_copy")

    def copy_to_host(self, arr, ptr):
        """Fills the host array from device memory at ptr (arr.nbytes bytes)."""
        self._check(self.L.gf_memcpy(self.h, arr.ctypes.data, ptr, arr.nbytes, 2), "gf_memcpy")
        return arr

    def to_device(self, arr):
        """A new device copy of a host array (free it with free())."""
        arr = np.ascontiguousarray(arr)
        p = self.alloc(max(arr.nbytes, 8))
        self._check(self.L.gf_memcpy(self.h, p, arr.ctypes.data, arr.nbytes, 1), "gf_memcpy")
        return p

    def to_host(self, ptr, arr):
        self._check(self.L.gf_memcpy(self.h, arr.ctypes.data, ptr, arr.nbytes, 2), "gf_memcpy")
        return arr

    def debug_eigen(self, covs):
        covs = _c64(np.asarray(covs).reshape(-1, 9))
        m = len(covs)
        ev = np.zeros((m, 3))
        vec = np.zeros((m, 9))
        info = np.zeros(m, np.int32)
        self._check(self.L.gf_debug_eigen(self.h, covs.ctypes.data_as(_dp), m,
                                          ev.ctypes.data_as(_dp), vec.ctypes.data_as(_dp),
                                          info.ctypes.data_as(_i32p)), "gf_debug_eigen")
        return info, ev, vec.reshape(m, 3, 3).transpose(0, 2, 1)  # [:, :, j] = vector j


class Cloud:
    def __init__(self, ctx, x, y, z):
        self.ctx = ctx
        x, y, z = _c64(x), _c64(y), _c64(z)
        self.n = len(x)
        h = ctypes.c_void_p()
        ctx._check(ctx.L.gf_cloud_create(ctx.h, x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                         self.n, WHERE_HOST, ctypes.byref(h)),
                   "gf_cloud_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.gf_cloud_release(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def optimal_neighborhood(self, min_k=10, max_k=14):
        """filters.optimalneighborhood: (OptimalKNN uint64, OptimalRadius)."""
        kopt = np.zeros(self.n, np.uint64)
        ropt = np.zeros(self.n, np.float64)
        self.ctx._check(self.ctx.L.gf_optimal_neighborhood(self.ctx.h, self.h, min_k, max_k,
                                                           kopt.ctypes.data, ropt.ctypes.data,
                                                           WHERE_HOST),
                        "gf_optimal_neighborhood")
        return kopt, ropt

    def xyz(self):
        """The cloud's resident coordinates, copied to the host."""
        out = [np.empty(self.n, np.float64) for _ in range(3)]
        self.ctx._check(self.ctx.L.gf_debug_cloud_xyz(self.ctx.h, self.h, out[0].ctypes.data,
                                                      out[1].ctypes.data, out[2].ctypes.data),
                        "gf_debug_cloud_xyz")
        return tuple(out)

    def run(self, stages, init=0.0):
        """Runs a chain of stages; returns (list of per-stage column lists,
        list of per-stage status arrays)."""
        stages = list(stages)
        arr = (GfStage * len(stages))(*stages)
        cols, per_stage = [], []
        for st in stages:
            c = [np.full(self.n, init, np.float64) for _ in range(stage_columns(st))]
            per_stage.append(c)
            cols.extend(c)
        status = [np.zeros(self.n, np.uint8) for _ in stages]
        cp = (ctypes.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
        sp = (ctypes.c_void_p * len(status))(*[s.ctypes.data for s in status])
        self.ctx._check(self.ctx.L.gf_run(self.ctx.h, self.h, arr, len(stages), cp, sp,
                                          WHERE_HOST), "gf_run")
        return per_stage, status

    def run_rowk(self, stages, row_k, init=0.0):
        """Runs stages over each row's own neighbor count (row_k[i] nearest
        points in KD3Index order; covariancefeatures' optimized=true)."""
        stages = list(stages)
        row_k = np.ascontiguousarray(row_k, dtype=np.uint32)
        if row_k.shape != (self.n,):
            raise ValueError("row_k needs one count per point")
        arr = (GfStage * len(stages))(*stages)
        cols, per_stage = [], []
        for st in stages:
            c = [np.full(self.n, init, np.float64) for _ in range(stage_columns(st))]
            per_stage.append(c)
            cols.extend(c)
        status = [np.zeros(self.n, np.uint8) for _ in stages]
        cp = (ctypes.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
        sp = (ctypes.c_void_p * len(status))(*[s.ctypes.data for s in status])
        self.ctx._check(self.ctx.L.gf_run_rowk(self.ctx.h, self.h, arr, len(stages),
                                               row_k.ctypes.data, cp, sp, WHERE_HOST),
                        "gf_run_rowk")
        return per_stage, status

    def invalidate(self):
        """Drops the cached index; the next run rebuilds it."""
        self.ctx.L.gf_cloud_invalidate(self.h)

    def run_device(self, stages, cols, status):
        """Runs a chain with outputs in device memory (cols: device column
        pointers in stage order; status: one device pointer per stage)."""
        stages = list(stages)
        arr = (GfStage * len(stages))(*stages)
        cp = (ctypes.c_void_p * max(1, len(cols)))(*cols)
        sp = (ctypes.c_void_p * len(status))(*status)
        self.ctx._check(self.ctx.L.gf_run(self.ctx.h, self.h, arr, len(stages), cp, sp,
                                          WHERE_DEVICE), "gf_run")

    def radius_counts(self, radius):
        """Points within `radius` of each point (itself included), strictly."""
        out = np.zeros(self.n, np.uint32)
        self.ctx._check(self.ctx.L.gf_radius_counts(self.ctx.h, self.h, radius,
                                                    out.ctypes.data, WHERE_HOST),
                        "gf_radius_counts")
        return out

    def nndistance(self, k=10, mode="kth"):
        """filters.nndistance (NNDistance): the distance to the k-th nearest
        neighbor ("kth") or the mean distance to the k nearest ("avg")."""
        m = {"kth": NND_KTH, "avg": NND_AVG}.get(str(mode).lower())
        if m is None:
            raise ValueError("nndistance mode must be 'kth' or 'avg'")
        out = np.zeros(self.n, np.float64)
        self.ctx._check(self.ctx.L.gf_nndistance(self.ctx.h, self.h, k, m, out.ctypes.data,
                                                 WHERE_HOST), "gf_nndistance")
        return out

    def outlier(self, method="statistical", mean_k=8, multiplier=2.0, radius=1.0, min_k=2):
        """filters.outlier: (uint8 flags, 1 = outlier; the statistical
        threshold, or None for the radius method)."""
        m = {"statistical": OUTLIER_STATISTICAL, "radius": OUTLIER_RADIUS}.get(
            str(method).lower())
        if m is None:
            raise ValueError("outlier method must be 'statistical' or 'radius'")
        if mean_k < 0:
            raise ValueError("mean_k must be >= 0")
        out = np.zeros(max(self.n, 1), np.uint8)
        thr = ctypes.c_double(float("nan"))
        self.ctx._check(self.ctx.L.gf_outlier(self.ctx.h, self.h, m, mean_k, multiplier, radius,
                                              min_k, out.ctypes.data, ctypes.byref(thr)),
                        "gf_outlier")
        return out[:self.n], (thr.value if m == OUTLIER_STATISTICAL else None)

    def knn(self, k, stride=1):
        kk = min(k, self.n)
        ids = np.zeros((self.n, kk), np.uint32)
        st = np.zeros(self.n, np.uint8)
        self.ctx._check(self.ctx.L.gf_knn(self.ctx.h, self.h, k, stride, ids.ctypes.data,
                                          st.ctypes.data, WHERE_HOST), "gf_knn")
        return ids, st

    def kdtree(self):
        """The exact KD3Index tree: vind and level-order nodes."""
        m = ctypes.c_uint32()
        L = self.ctx.L
        self.ctx._check(L.gf_debug_kdtree(self.ctx.h, self.h, None, ctypes.byref(m), None, None,
                                          None, None, None), "gf_debug_kdtree")
        k = m.value
        vind = np.zeros(self.n, np.uint32)
        lo = np.zeros(k, np.uint32)
        hi = np.zeros(k, np.uint32)
        feat = np.zeros(k, np.int32)
        dl = np.zeros(k)
        dh = np.zeros(k)
        self.ctx._check(L.gf_debug_kdtree(self.ctx.h, self.h, vind.ctypes.data, ctypes.byref(m),
                                          lo.ctypes.data, hi.ctypes.data, feat.ctypes.data,
                                          dl.ctypes.data, dh.ctypes.data), "gf_debug_kdtree")
        return dict(vind=vind, left=lo, right=hi, divfeat=feat, divlow=dl, divhigh=dh)

    def grid(self, k):
        ncols = ctypes.c_uint32()
        bits = ctypes.c_uint32()
        self.ctx._check(self.ctx.L.gf_debug_grid(self.ctx.h, self.h, k, None, None, None,
                                                 ctypes.byref(ncols), ctypes.byref(bits)),
                        "gf_debug_grid")
        perm = np.zeros(self.n, np.uint32)
        keys = np.zeros(self.n, np.uint32)
        cs = np.zeros(ncols.value + 1, np.uint32)
        self.ctx._check(self.ctx.L.gf_debug_grid(self.ctx.h, self.h, k,
                                                 perm.ctypes.data_as(_u32p),
                                                 keys.ctypes.data_as(_u32p),
                                                 cs.ctypes.data_as(_u32p),
                                                 ctypes.byref(ncols), ctypes.byref(bits)),
                        "gf_debug_grid")
        return perm, keys, cs, bits.value

    def grid_params(self, k):
        """The grid of the index built for k: a dict with origin (ox, oy,
        oz), edges (c, cz), inverse edges (inv_c, inv_cz) and integer fields
        gx, gy, gz, tiles_x, bz, key_bits, ncols."""
        f = np.zeros(7)
        u = np.zeros(8, np.uint32)
        self.ctx._check(self.ctx.L.gf_debug_grid_params(self.ctx.h, self.h, k,
                                                        f.ctypes.data_as(_dp),
                                                        u.ctypes.data_as(_u32p)),
                        "gf_debug_grid_params")
        d = dict(zip(["ox", "oy", "oz", "c", "cz", "inv_c", "inv_cz"], f.tolist()))
        d.update(zip(["gx", "gy", "gz", "tiles_x", "bz", "key_bits", "ncols"],
                     [int(v) for v in u[:7]]))
        return d


class Batch:
    """Independent tiles run together (one index build, one search, one
    exact re-solve and one feature pass); each tile's rows equal the tile run
    alone.  Points of tile t are rows starts[t]:starts[t+1] of every output."""

    def __init__(self, ctx, tiles):
        self.ctx = ctx
        L = ctx.L
        counts = np.array([len(t[0]) for t in tiles], dtype=np.uint64)
        h = ctypes.c_void_p()
        ctx._check(L.gf_batch_create(ctx.h, len(tiles), counts.ctypes.data, ctypes.byref(h)),
                   "gf_batch_create")
        self.h = h
        self.ntiles = len(tiles)
        for i, (x, y, z) in enumerate(tiles):
            x, y, z = _c64(x), _c64(y), _c64(z)
            ctx._check(L.gf_batch_set_tile(h, i, x.ctypes.data, y.ctypes.data, z.ctypes.data,
                                           WHERE_HOST), "gf_batch_set_tile")
        self.starts = [int(L.gf_batch_tile_start(h, i)) for i in range(self.ntiles + 1)]
        self.n = int(L.gf_batch_size(h))

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.gf_batch_release(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def invalidate(self):
        self.ctx.L.gf_batch_invalidate(self.h)

    def set_tile(self, t, x, y, z):
        """New coordinates for tile t (its point count is fixed)."""
        x, y, z = _c64(x), _c64(y), _c64(z)
        self.ctx._check(self.ctx.L.gf_batch_set_tile(self.h, t, x.ctypes.data, y.ctypes.data,
                                                     z.ctypes.data, WHERE_HOST),
                        "gf_batch_set_tile")

    def _knn_arg(self, tile_knn):
        if tile_knn is None:
            return None, None
        a = np.ascontiguousarray(tile_knn, dtype=np.int32)
        assert len(a) == self.ntiles
        return a, a.ctypes.data

    def run(self, stages, tile_knn=None, init=0.0):
        """Returns (per-stage column lists over the whole batch, per-stage
        status arrays)."""
        stages = list(stages)
        arr = (GfStage * len(stages))(*stages)
        cols, per_stage = [], []
        for st in stages:
            c = [np.full(self.n, init, np.float64) for _ in range(stage_columns(st))]
            per_stage.append(c)
            cols.extend(c)
        status = [np.zeros(self.n, np.uint8) for _ in stages]
        cp = (ctypes.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
        sp = (ctypes.c_void_p * len(status))(*[s.ctypes.data for s in status])
        keep, kp = self._knn_arg(tile_knn)
        self.ctx._check(self.ctx.L.gf_batch_run(self.ctx.h, self.h, arr, len(stages), kp, cp, sp,
                                                WHERE_HOST), "gf_batch_run")
        return per_stage, status

    def run_device(self, stages, cols, status, tile_knn=None):
        stages = list(stages)
        arr = (GfStage * len(stages))(*stages)
        cp = (ctypes.c_void_p * max(1, len(cols)))(*cols)
        sp = (ctypes.c_void_p * len(status))(*status)
        keep, kp = self._knn_arg(tile_knn)
        self.ctx._check(self.ctx.L.gf_batch_run(self.ctx.h, self.h, arr, len(stages), kp, cp, sp,
                                                WHERE_DEVICE), "gf_batch_run")

    def tile(self, arr, t):
        """Rows of tile t in an output array."""
        return arr[self.starts[t]:self.starts[t + 1]]


class Server:
    """The serving loop: submit tiles at any time, wait for their results.
    Host arrays passed to submit() are kept alive here until the request is
    released."""

    def __init__(self, device=0, workers=2, hbm_budget_bytes=0, max_batch_points=0):
        self.L = lib()
        h = ctypes.c_void_p()
        rc = self.L.gf_server_create(device, workers, hbm_budget_bytes, max_batch_points,
                                     ctypes.byref(h))
        if rc != GF_OK:
            raise GfError(rc, "gf_server_create failed (is a GPU visible?)")
        self.h = h
        self._held = {}

    def submit(self, x, y, z, stages, knn, init=0.0):
        """Queues one tile; returns a ticket.  Outputs come from result()."""
        x, y, z = _c64(x), _c64(y), _c64(z)
        n = len(x)
        stages = list(stages)
        arr = (GfStage * len(stages))(*stages)
        cols, per_stage = [], []
        for st in stages:
            c = [np.full(n, init, np.float64) for _ in range(stage_columns(st))]
            per_stage.append(c)
            cols.extend(c)
        status = [np.zeros(n, np.uint8) for _ in stages]
        cp = (ctypes.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
        sp = (ctypes.c_void_p * len(status))(*[s.ctypes.data for s in status])
        t = ctypes.c_uint64()
        rc = self.L.gf_server_submit(self.h, x.ctypes.data, y.ctypes.data, z.ctypes.data, n, arr,
                                     len(stages), int(knn), cp, sp, ctypes.byref(t))
        if rc != GF_OK:
            raise GfError(rc, "gf_server_submit")
        self._held[t.value] = (x, y, z, arr, cp, sp, per_stage, status)
        return t.value

    def result(self, ticket, timeout=-1.0):
        """Waits; returns (per-stage columns, statuses, times) with times =
        (submit, batch start, done, batch id, tiles in the batch)."""
        rc = self.L.gf_server_wait(self.h, ticket, float(timeout))
        if rc == GF_ERR_TIMEOUT:
            raise GfError(rc, f"request {ticket}: still running after {timeout} s")
        if rc != GF_OK:
            # Finished with an error: the request and its arrays are let go.
            self.L.gf_server_release(self.h, ticket)
            self._held.pop(ticket, None)
            raise GfError(rc, f"request {ticket}")
        tm = np.zeros(3)
        bid = ctypes.c_uint32()
        bt = ctypes.c_uint32()
        self.L.gf_server_request_info(self.h, ticket, tm.ctypes.data_as(_dp), ctypes.byref(bid),
                                      ctypes.byref(bt))
        held = self._held.pop(ticket)
        self.L.gf_server_release(self.h, ticket)
        return held[6], held[7], (tm[0], tm[1], tm[2], bid.value, bt.value)

    def close(self):
        if getattr(self, "h", None):
            self.L.gf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
