"""ctypes bindings for the oracle (oracle/liboracle.so) and, where it was
built, oracle/_ref/libref.so (the reference's own nanoflann and Eigen).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

STAGE_COVFEAT, STAGE_NORMAL, STAGE_EIGENVALUES, STAGE_COPLANAR, STAGE_RANK = 0, 1, 2, 3, 4
MODE_RAW, MODE_SQRT, MODE_NORMALIZED = 0, 1, 2
F_LINEARITY = 1 << 0
F_PLANARITY = 1 << 1
F_SCATTERING = 1 << 2
F_VERTICALITY = 1 << 3
F_OMNIVARIANCE = 1 << 4
F_ANISOTROPY = 1 << 5
F_EIGENENTROPY = 1 << 6
F_EIGENVALUESUM = 1 << 7
F_SURFACEVARIATION = 1 << 8
F_DEMANTKEVERTICALITY = 1 << 9
F_DIMENSIONALITY = F_LINEARITY | F_PLANARITY | F_SCATTERING | F_VERTICALITY
F_ALL10 = (1 << 10) - 1
FEATURE_NAMES = ["Linearity", "Planarity", "Scattering", "Verticality",
                 "Omnivariance", "Anisotropy", "Eigenentropy", "EigenvalueSum",
                 "SurfaceVariation", "DemantkeVerticality"]

ST_TIE, ST_INCOMPLETE, ST_ZERO_COV, ST_EIGEN_FAIL = 1, 2, 4, 8
ST_FEATURE_INVALID, ST_FEW_NEIGHBORS = 16, 32


class Stage(ctypes.Structure):
    _fields_ = [("stage", ctypes.c_int32), ("knn", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("features", ctypes.c_uint32), ("always_up", ctypes.c_int32),
                ("normalize", ctypes.c_int32), ("min_k", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("radius", ctypes.c_double),
                ("p1", ctypes.c_double), ("p2", ctypes.c_double)]


def stage(kind, knn, stride=1, mode=MODE_SQRT, features=F_DIMENSIONALITY,
          always_up=1, normalize=0, radius=0.0, min_k=3, p1=None, p2=None):
    if p1 is None:
        p1 = {STAGE_COPLANAR: 25.0, STAGE_RANK: 0.01}.get(kind, 0.0)
    if p2 is None:
        p2 = 6.0 if kind == STAGE_COPLANAR else 0.0
    return Stage(kind, knn, stride, mode, features, always_up, normalize, min_k, 0, radius,
                 p1, p2)


def stage_columns(st):
    if st.stage in (STAGE_COPLANAR, STAGE_RANK):
        return 1
    if st.stage == STAGE_NORMAL:
        return 4
    if st.stage == STAGE_EIGENVALUES:
        return 3
    return bin(st.features & F_ALL10).count("1")


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.orc_kd3_build.restype = ctypes.c_void_p
        L.orc_kd3_build.argtypes = [_dp, _dp, _dp, ctypes.c_uint64]
        L.orc_kd3_free.argtypes = [ctypes.c_void_p]
        L.orc_optimal_neighborhood.restype = ctypes.c_uint64
        L.orc_optimal_neighborhood.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_uint64,
                                               ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                               _u64p, _dp, _dp]
        L.orc_kd3_neighbors.restype = ctypes.c_uint64
        L.orc_kd3_neighbors.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_uint64,
                                        _u64p, _dp]
        L.orc_kd3_neighbors_all.argtypes = [ctypes.c_void_p, ctypes.c_uint64,
                                            ctypes.c_uint64, ctypes.c_int, _u32p]
        L.orc_kd3_neighbors_rows.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64,
                                             ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                             _u32p]
        L.orc_kd3_node_count.restype = ctypes.c_uint64
        L.orc_kd3_node_count.argtypes = [ctypes.c_void_p]
        L.orc_kd3_export.argtypes = [ctypes.c_void_p, _u64p, _i32p, _i32p,
                                     _u64p, _u64p, _i32p, _dp, _dp]
        L.orc_brute_knn.argtypes = [_dp, _dp, _dp, ctypes.c_uint64, _u64p,
                                    ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_int, _u32p, _dp]
        L.orc_pdal_covariance.argtypes = [_dp, _dp, _dp, _u64p,
                                          ctypes.c_uint64, _dp]
        L.orc_is_zero.restype = ctypes.c_int
        L.orc_is_zero.argtypes = [_dp]
        L.orc_eigen_sym3.restype = ctypes.c_int
        L.orc_eigen_sym3.argtypes = [_dp, _dp, _dp]
        L.orc_jacobi_sym3.restype = ctypes.c_int
        L.orc_jacobi_sym3.argtypes = [_dp, _dp, _dp]
        sp = ctypes.POINTER(Stage)
        L.orc_stage_run.argtypes = [ctypes.c_void_p, _dp, _dp, _dp,
                                    ctypes.c_uint64, sp, ctypes.c_int,
                                    ctypes.POINTER(_dp), _u8p]
        L.orc_stage_run_range.argtypes = [ctypes.c_void_p, _dp, _dp, _dp,
                                          ctypes.c_uint64, sp, ctypes.c_int,
                                          ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.POINTER(_dp), _u8p]
        L.orc_stage_from_lists.argtypes = [_dp, _dp, _dp, ctypes.c_uint64,
                                           _u32p, ctypes.c_uint64, sp,
                                           ctypes.c_int, ctypes.POINTER(_dp),
                                           _u8p]
        L.orc_truth_from_lists.argtypes = [_dp, _dp, _dp, ctypes.c_uint64,
                                           _u32p, ctypes.c_uint64, sp,
                                           ctypes.c_int, ctypes.POINTER(_dp),
                                           _u8p, _dp, _dp]
        L.orc_rank_from_eigenvalues.restype = ctypes.c_int
        L.orc_rank_from_eigenvalues.argtypes = [_dp, ctypes.c_double]
        L.orc_kd3_radius.restype = ctypes.c_uint64
        L.orc_kd3_radius.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                     _u64p, _dp]
        L.orc_kd3_radius_counts.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                            _u32p]
        L.orc_brute_radius_counts.argtypes = [_dp, _dp, _dp, ctypes.c_uint64, ctypes.c_double,
                                              _u64p, ctypes.c_uint64, _u32p]
        L.orc_truth_from_csr.argtypes = [_dp, _dp, _dp, ctypes.c_uint64, _u64p, _u32p, sp,
                                         ctypes.c_int, ctypes.POINTER(_dp), _u8p, _dp, _dp]
        L.orc_stage_run_rowk.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_uint64, sp,
                                         _u32p, ctypes.c_int, ctypes.POINTER(_dp), _u8p]
        L.orc_nn_distance.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_uint64,
                                      ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_outlier_statistical.restype = ctypes.c_uint64
        L.orc_outlier_statistical.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_uint64,
                                              ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                              _u8p, _dp, _dp]
        L.orc_outlier_radius.restype = ctypes.c_uint64
        L.orc_outlier_radius.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double,
                                         ctypes.c_int64, ctypes.c_int, _u8p]
        self.L = L

    # --------------------------------------------------------------- KD3
    class Tree:
        def __init__(self, orc, x, y, z):
            self.orc = orc
            self.x, self.y, self.z = (_c(v, np.float64) for v in (x, y, z))
            self.n = len(self.x)
            self.h = orc.L.orc_kd3_build(_ptr(self.x, _dp), _ptr(self.y, _dp),
                                         _ptr(self.z, _dp), self.n)

        def __del__(self):
            if getattr(self, "h", None):
                self.orc.L.orc_kd3_free(self.h)
                self.h = None

        def neighbors(self, qx, qy, qz, k, stride=1):
            cap = max(1, stride * min(k, self.n))
            ids = np.zeros(cap, np.uint64)
            d2 = np.zeros(cap, np.float64)
            c = self.orc.L.orc_kd3_neighbors(self.h, qx, qy, qz, k, stride,
                                             _ptr(ids, _u64p), _ptr(d2, _dp))
            return ids[:c].copy()

        def neighbors_all(self, k, stride=1, threads=1):
            kk = min(k, self.n)
            out = np.zeros((self.n, kk), np.uint32)
            self.orc.L.orc_kd3_neighbors_all(self.h, k, stride, threads,
                                             _ptr(out, _u32p))
            return out

        def neighbors_rows(self, rows, k, stride=1, threads=1):
            """neighbors(point r, k, stride) for each r in rows."""
            rows = _c(rows, np.uint64)
            kk = min(k, self.n)
            out = np.zeros((len(rows), kk), np.uint32)
            self.orc.L.orc_kd3_neighbors_rows(self.h, _ptr(rows, _u64p), len(rows), k, stride,
                                              threads, _ptr(out, _u32p))
            return out

        def radius(self, qx, qy, qz, r):
            """KD3Index::radius: (ids, d2) sorted by distance."""
            cnt = self.orc.L.orc_kd3_radius(self.h, qx, qy, qz, r, 0, None, None)
            ids = np.zeros(max(cnt, 1), np.uint64)
            d2 = np.zeros(max(cnt, 1), np.float64)
            self.orc.L.orc_kd3_radius(self.h, qx, qy, qz, r, cnt, _ptr(ids, _u64p), _ptr(d2, _dp))
            return ids[:cnt], d2[:cnt]

        def radius_counts(self, r, threads=8):
            out = np.zeros(self.n, np.uint32)
            self.orc.L.orc_kd3_radius_counts(self.h, r, threads, _ptr(out, _u32p))
            return out

        def radius_csr(self, rows, r):
            """CSR (offsets, ids) of KD3Index::radius for each row."""
            offs = [0]
            parts = []
            for q in rows:
                ids, _ = self.radius(self.x[q], self.y[q], self.z[q], r)
                parts.append(ids.astype(np.uint32))
                offs.append(offs[-1] + len(ids))
            ids = np.concatenate(parts) if parts else np.zeros(0, np.uint32)
            return np.array(offs, np.uint64), ids

        def export(self):
            m = self.orc.L.orc_kd3_node_count(self.h)
            vind = np.zeros(self.n, np.uint64)
            c1 = np.zeros(m, np.int32)
            c2 = np.zeros(m, np.int32)
            lo = np.zeros(m, np.uint64)
            hi = np.zeros(m, np.uint64)
            feat = np.zeros(m, np.int32)
            dl = np.zeros(m, np.float64)
            dh = np.zeros(m, np.float64)
            self.orc.L.orc_kd3_export(self.h, _ptr(vind, _u64p), _ptr(c1, _i32p),
                                      _ptr(c2, _i32p), _ptr(lo, _u64p),
                                      _ptr(hi, _u64p), _ptr(feat, _i32p),
                                      _ptr(dl, _dp), _ptr(dh, _dp))
            return dict(vind=vind, child1=c1, child2=c2, left=lo, right=hi,
                        divfeat=feat, divlow=dl, divhigh=dh)

    def kd3(self, x, y, z):
        return Oracle.Tree(self, x, y, z)

    def brute_knn(self, x, y, z, kk, queries=None, threads=8):
        x, y, z = (_c(v, np.float64) for v in (x, y, z))
        n = len(x)
        q = np.arange(n, dtype=np.uint64) if queries is None else _c(queries, np.uint64)
        ids = np.zeros((len(q), kk), np.uint32)
        d2 = np.zeros((len(q), kk), np.float64)
        self.L.orc_brute_knn(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp), n,
                             _ptr(q, _u64p), len(q), kk, threads,
                             _ptr(ids, _u32p), _ptr(d2, _dp))
        return ids, d2

    # ------------------------------------------------------ arithmetic
    def pdal_covariance(self, x, y, z, ids):
        x, y, z = (_c(v, np.float64) for v in (x, y, z))
        ids = _c(ids, np.uint64)
        cov = np.zeros(9)
        self.L.orc_pdal_covariance(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                                   _ptr(ids, _u64p), len(ids), _ptr(cov, _dp))
        return cov.reshape(3, 3)

    def rank_from_eigenvalues(self, ev, thresh):
        e = _c(np.asarray(ev, np.float64).reshape(3), np.float64)
        return self.L.orc_rank_from_eigenvalues(_ptr(e, _dp), thresh)

    def is_zero(self, cov):
        c = _c(np.asarray(cov).reshape(9), np.float64)
        return bool(self.L.orc_is_zero(_ptr(c, _dp)))

    def eigen_sym3(self, cov):
        c = _c(np.asarray(cov).reshape(9), np.float64)
        ev = np.zeros(3)
        vec = np.zeros(9)
        info = self.L.orc_eigen_sym3(_ptr(c, _dp), _ptr(ev, _dp), _ptr(vec, _dp))
        return info, ev, vec.reshape(3, 3).T  # columns = eigenvectors

    def jacobi_sym3(self, cov):
        c = _c(np.asarray(cov).reshape(9), np.float64)
        ev = np.zeros(3)
        vec = np.zeros(9)
        self.L.orc_jacobi_sym3(_ptr(c, _dp), _ptr(ev, _dp), _ptr(vec, _dp))
        return ev, vec.reshape(3, 3).T

    # ---------------------------------------------------------- stages
    def _cols(self, st, n, init=0.0):
        cols = [np.full(n, init, np.float64) for _ in range(stage_columns(st))]
        arr = (_dp * max(1, len(cols)))(*[_ptr(c, _dp) for c in cols])
        return cols, arr

    def stage_run(self, tree, st, threads=1, init=0.0):
        n = tree.n
        cols, arr = self._cols(st, n, init)
        status = np.zeros(n, np.uint8)
        self.L.orc_stage_run(tree.h, _ptr(tree.x, _dp), _ptr(tree.y, _dp),
                             _ptr(tree.z, _dp), n, ctypes.byref(st), threads,
                             arr, _ptr(status, _u8p))
        return cols, status

    def stage_run_rowk(self, tree, st, row_k, threads=8, init=0.0):
        """The stage over each row's own neighbor count (optimized=true)."""
        n = tree.n
        row_k = _c(row_k, np.uint32)
        cols, arr = self._cols(st, n, init)
        status = np.zeros(n, np.uint8)
        self.L.orc_stage_run_rowk(tree.h, _ptr(tree.x, _dp), _ptr(tree.y, _dp),
                                  _ptr(tree.z, _dp), n, ctypes.byref(st), _ptr(row_k, _u32p),
                                  threads, arr, _ptr(status, _u8p))
        return cols, status

    def optimal_neighborhood(self, tree, kmin=10, kmax=14, threads=8):
        """filters.optimalneighborhood: (OptimalKNN, OptimalRadius, entropy
        gap of the choice, failed rows)."""
        n = tree.n
        kopt = np.zeros(n, np.uint64)
        ropt = np.zeros(n, np.float64)
        gap = np.zeros(n, np.float64)
        failed = self.L.orc_optimal_neighborhood(tree.h, _ptr(tree.x, _dp), _ptr(tree.y, _dp),
                                                 _ptr(tree.z, _dp), n, kmin, kmax, threads,
                                                 _ptr(kopt, _u64p), _ptr(ropt, _dp),
                                                 _ptr(gap, _dp))
        return kopt, ropt, gap, int(failed)

    def stage_run_range(self, tree, st, begin, end, threads=1, init=0.0):
        """PDAL's CPU stage over rows [begin, end) of the tree's cloud."""
        n = tree.n
        cols, arr = self._cols(st, n, init)
        status = np.zeros(n, np.uint8)
        self.L.orc_stage_run_range(tree.h, _ptr(tree.x, _dp), _ptr(tree.y, _dp),
                                   _ptr(tree.z, _dp), n, ctypes.byref(st), threads,
                                   begin, end, arr, _ptr(status, _u8p))
        return cols, status

    def stage_from_lists(self, x, y, z, nbr, st, threads=8, init=0.0):
        """Rows are nbr's rows (ids index x, y, z)."""
        x, y, z = (_c(v, np.float64) for v in (x, y, z))
        nbr = _c(nbr, np.uint32)
        n = nbr.shape[0]
        cols, arr = self._cols(st, n, init)
        status = np.zeros(n, np.uint8)
        self.L.orc_stage_from_lists(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp), n,
                                    _ptr(nbr, _u32p), nbr.shape[1],
                                    ctypes.byref(st), threads, arr,
                                    _ptr(status, _u8p))
        return cols, status

    def truth_from_lists(self, x, y, z, nbr, st, threads=8, init=np.nan,
                         with_ratio=False):
        """Extended-precision truth; returns (cols, status, gap) or, with
        with_ratio, (cols, status, gap, ratio) where ratio = lambda_min /
        lambda_max of the row's covariance.  Rows are nbr's rows (ids index
        x, y, z)."""
        x, y, z = (_c(v, np.float64) for v in (x, y, z))
        nbr = _c(nbr, np.uint32)
        n = nbr.shape[0]
        cols, arr = self._cols(st, n, init)
        status = np.zeros(n, np.uint8)
        gap = np.zeros(n, np.float64)
        ratio = np.zeros(n, np.float64)
        self.L.orc_truth_from_lists(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp), n,
                                    _ptr(nbr, _u32p), nbr.shape[1],
                                    ctypes.byref(st), threads, arr,
                                    _ptr(status, _u8p), _ptr(gap, _dp),
                                    _ptr(ratio, _dp))
        if with_ratio:
            return cols, status, gap, ratio
        return cols, status, gap


def _brute_radius_counts(orc, x, y, z, r, rows=None):
    x, y, z = (_c(v, np.float64) for v in (x, y, z))
    n = len(x)
    q = None if rows is None else _c(rows, np.uint64)
    nq = n if rows is None else len(q)
    out = np.zeros(nq, np.uint32)
    orc.L.orc_brute_radius_counts(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp), n, r,
                                  None if q is None else _ptr(q, _u64p), nq, _ptr(out, _u32p))
    return out


Oracle.brute_radius_counts = _brute_radius_counts


def _truth_from_csr(orc, x, y, z, offsets, ids, st, threads=8, init=np.nan, with_ratio=False):
    """Extended-precision truth over CSR neighbor lists (rows in order)."""
    x, y, z = (_c(v, np.float64) for v in (x, y, z))
    offsets = _c(offsets, np.uint64)
    ids = _c(ids, np.uint32)
    n = len(offsets) - 1
    cols, arr = orc._cols(st, n, init)
    status = np.zeros(n, np.uint8)
    gap = np.zeros(n, np.float64)
    ratio = np.zeros(n, np.float64)
    orc.L.orc_truth_from_csr(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp), n,
                             _ptr(offsets, _u64p), _ptr(ids if len(ids) else np.zeros(1, np.uint32), _u32p),
                             ctypes.byref(st), threads, arr, _ptr(status, _u8p), _ptr(gap, _dp),
                             _ptr(ratio, _dp))
    if with_ratio:
        return cols, status, gap, ratio
    return cols, status, gap


Oracle.truth_from_csr = _truth_from_csr

NND_KTH, NND_AVG, NND_RUNNING_MEAN = 0, 1, 2


def _nn_distance(orc, tree, k, mode=NND_KTH, threads=8):
    """filters.nndistance per point (mode 2: the outlier running mean)."""
    n = len(tree.x)
    out = np.zeros(n)
    orc.L.orc_nn_distance(tree.h, _ptr(tree.x, _dp), _ptr(tree.y, _dp), _ptr(tree.z, _dp), n,
                          k, mode, threads, _ptr(out, _dp))
    return out


def _outlier_statistical(orc, tree, mean_k=8, multiplier=2.0, threads=8):
    """(outlier flags, threshold, per-point mean distances)."""
    n = len(tree.x)
    out = np.zeros(max(n, 1), np.uint8)
    dist = np.zeros(max(n, 1))
    thr = ctypes.c_double(0.0)
    orc.L.orc_outlier_statistical(tree.h, _ptr(tree.x, _dp), _ptr(tree.y, _dp),
                                  _ptr(tree.z, _dp), n, mean_k, multiplier, threads,
                                  _ptr(out, _u8p), ctypes.byref(thr), _ptr(dist, _dp))
    return out[:n], thr.value, dist[:n]


def _outlier_radius(orc, tree, radius=1.0, min_k=2, threads=8):
    n = len(tree.x)
    out = np.zeros(max(n, 1), np.uint8)
    orc.L.orc_outlier_radius(tree.h, n, radius, min_k, threads, _ptr(out, _u8p))
    return out[:n]


Oracle.nn_distance = _nn_distance
Oracle.outlier_statistical = _outlier_statistical
Oracle.outlier_radius = _outlier_radius


class Ref:
    """The reference's own nanoflann + Eigen (oracle/_ref).  Present only
    where /root/reference was available when build() ran."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.ref_kd3_build.restype = ctypes.c_void_p
        L.ref_kd3_build.argtypes = [_dp, _dp, _dp, ctypes.c_uint64, ctypes.c_int]
        L.ref_kd3_free.argtypes = [ctypes.c_void_p]
        L.ref_kd3_neighbors.restype = ctypes.c_uint64
        L.ref_kd3_neighbors.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_uint64, _u64p]
        L.ref_kd3_node_count.restype = ctypes.c_uint64
        L.ref_kd3_node_count.argtypes = [ctypes.c_void_p]
        L.ref_kd3_export.argtypes = [ctypes.c_void_p, _u64p, _u64p, _u64p, _i32p,
                                     _dp, _dp, _i32p]
        L.ref_eigen_sym3.restype = ctypes.c_int
        L.ref_eigen_sym3.argtypes = [_dp, _dp, _dp]
        L.ref_pdal_covariance.argtypes = [_dp, _dp, _dp, _u64p, ctypes.c_uint64, _dp]
        L.ref_jacobi_rank.restype = ctypes.c_int
        L.ref_jacobi_rank.argtypes = [_dp, ctypes.c_double]
        L.ref_kd3_radius.restype = ctypes.c_uint64
        L.ref_kd3_radius.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                     _u64p, _dp]
        self.L = L

    class Tree:
        def __init__(self, ref, x, y, z, threads=1):
            self.ref = ref
            self.x, self.y, self.z = (_c(v, np.float64) for v in (x, y, z))
            self.n = len(self.x)
            self.h = ref.L.ref_kd3_build(_ptr(self.x, _dp), _ptr(self.y, _dp),
                                         _ptr(self.z, _dp), self.n, threads)

        def __del__(self):
            if getattr(self, "h", None):
                self.ref.L.ref_kd3_free(self.h)
                self.h = None

        def neighbors(self, qx, qy, qz, k, stride=1):
            ids = np.zeros(max(1, min(k, self.n)), np.uint64)
            c = self.ref.L.ref_kd3_neighbors(self.h, qx, qy, qz, k, stride,
                                             _ptr(ids, _u64p))
            return ids[:c].copy()

        def radius(self, qx, qy, qz, r):
            cnt = self.ref.L.ref_kd3_radius(self.h, qx, qy, qz, r, 0, None, None)
            ids = np.zeros(max(cnt, 1), np.uint64)
            d2 = np.zeros(max(cnt, 1), np.float64)
            self.ref.L.ref_kd3_radius(self.h, qx, qy, qz, r, cnt, _ptr(ids, _u64p),
                                      _ptr(d2, _dp))
            return ids[:cnt], d2[:cnt]

        def export(self):
            m = self.ref.L.ref_kd3_node_count(self.h)
            vind = np.zeros(self.n, np.uint64)
            lo = np.zeros(m, np.uint64)
            hi = np.zeros(m, np.uint64)
            feat = np.zeros(m, np.int32)
            dl = np.zeros(m)
            dh = np.zeros(m)
            leaf = np.zeros(m, np.int32)
            self.ref.L.ref_kd3_export(self.h, _ptr(vind, _u64p), _ptr(lo, _u64p),
                                      _ptr(hi, _u64p), _ptr(feat, _i32p),
                                      _ptr(dl, _dp), _ptr(dh, _dp),
                                      _ptr(leaf, _i32p))
            return dict(vind=vind, left=lo, right=hi, divfeat=feat, divlow=dl,
                        divhigh=dh, leaf=leaf)

    def kd3(self, x, y, z, threads=1):
        return Ref.Tree(self, x, y, z, threads)

    def jacobi_rank(self, cov, threshold):
        c = _c(np.asarray(cov).reshape(9), np.float64)
        return self.L.ref_jacobi_rank(_ptr(c, _dp), threshold)

    def eigen_sym3(self, cov):
        c = _c(np.asarray(cov).reshape(9), np.float64)
        ev = np.zeros(3)
        vec = np.zeros(9)
        info = self.L.ref_eigen_sym3(_ptr(c, _dp), _ptr(ev, _dp), _ptr(vec, _dp))
        return info, ev, vec.reshape(3, 3).T

    def pdal_covariance(self, x, y, z, ids):
        x, y, z = (_c(v, np.float64) for v in (x, y, z))
        ids = _c(ids, np.uint64)
        cov = np.zeros(9)
        self.L.ref_pdal_covariance(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                                   _ptr(ids, _u64p), len(ids), _ptr(cov, _dp))
        return cov.reshape(3, 3)


def ref_available():
    return os.path.exists(REF_SO)
