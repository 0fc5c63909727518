/*
 * gpupdal_amd.h -- C ABI of the MI355X kNN-neighborhood feature library.
 *
 * The library computes what PDAL's filters.covariancefeatures,
 * filters.normal and filters.eigenvalues compute (k nearest neighbors per
 * point, the neighborhood covariance, its eigensystem and the stage
 * columns) on one AMD GPU (gfx950).  Plain C types only; no function throws
 * across this boundary.  Every entry returns GF_OK or a negative GF_ERR_*
 * code; gf_last_error() gives the message of the context's last failure.
 *
 * Ownership: the library owns device memory (clouds, indices, scratch).  The
 * caller owns host arrays; they are read or written only during the call.
 * Device pointers passed with GF_WHERE_DEVICE must stay valid until the call
 * returns.  A context is used by one thread at a time; the serving queue
 * (gf_serve_*) is safe to submit to from several threads.
 *
 * Which reference interface each entry replaces is stated beside it
 * (paths relative to the GPUPDAL tree).
 */
#ifndef GPUPDAL_AMD_H
#define GPUPDAL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GF_OK 0
#define GF_ERR_ARG (-1)
#define GF_ERR_HIP (-2)
#define GF_ERR_UNSUPPORTED (-3)
#define GF_ERR_OOM (-4)
#define GF_ERR_STATE (-5)
#define GF_ERR_TIMEOUT (-6)

#define GF_WHERE_HOST 0
#define GF_WHERE_DEVICE 1

/* Stage kinds, modes and feature bits.  The feature bit order is the
 * reference's CovarianceFeature enum (include/pdg/stages/Neighborhood.hpp:153-165). */
#define GF_STAGE_COVFEAT 0
#define GF_STAGE_NORMAL 1
#define GF_STAGE_EIGENVALUES 2
/* filters.approximatecoplanar (filters/ApproximateCoplanarFilter.cpp:67-124):
 * one column, Coplanar 1/0, from ev[1] > thresh1 * ev[0] && thresh2 * ev[1] >
 * ev[2] with thresh1 = p1 (25), thresh2 = p2 (6); the search asks knn (the
 * point itself included), not knn + 1. */
#define GF_STAGE_COPLANAR 3
/* filters.estimaterank (filters/EstimateRankFilter.cpp:62-84,
 * pdal/private/MathUtils.cpp:197-208): one column, Rank, the number of the
 * covariance's singular values at or above (float)thresh times the largest,
 * thresh = p1 (0.01); the search asks knn. */
#define GF_STAGE_RANK 4

#define GF_MODE_RAW 0
#define GF_MODE_SQRT 1
#define GF_MODE_NORMALIZED 2

#define GF_F_LINEARITY (1u << 0)
#define GF_F_PLANARITY (1u << 1)
#define GF_F_SCATTERING (1u << 2)
#define GF_F_VERTICALITY (1u << 3)
#define GF_F_OMNIVARIANCE (1u << 4)
#define GF_F_ANISOTROPY (1u << 5)
#define GF_F_EIGENENTROPY (1u << 6)
#define GF_F_EIGENVALUESUM (1u << 7)
#define GF_F_SURFACEVARIATION (1u << 8)
#define GF_F_DEMANTKEVERTICALITY (1u << 9)
#define GF_F_DIMENSIONALITY 0xFu

/* Row status bits; 1..16 keep the meanings of the reference's per-row status
 * (include/pdg/index/SpatialIndex.hpp:121-131).  GF_ST_TIE and
 * GF_ST_INCOMPLETE are informational on return: such rows were re-solved by
 * the exact KD3-order search and are exact. */
#define GF_ST_TIE 1u
#define GF_ST_INCOMPLETE 2u
#define GF_ST_ZERO_COV 4u
#define GF_ST_EIGEN_FAIL 8u
#define GF_ST_FEATURE_INVALID 16u
#define GF_ST_FEW_NEIGHBORS 32u

typedef struct gf_ctx gf_ctx;
typedef struct gf_cloud gf_cloud;

/* One stage of a chain.  Fields mirror the stages' options
 * (filters/CovarianceFeaturesFilter.cpp:102-113, filters/NormalFilter.cpp:85-97,
 * filters/EigenvaluesFilter.cpp:77-87): knn is the user's option (the search
 * asks knn+1 to include the point itself). */
typedef struct gf_stage {
    int32_t stage;     /* GF_STAGE_* */
    int32_t knn;
    int32_t stride;    /* covfeat / eigenvalues; 1 for normal */
    int32_t mode;      /* covfeat: GF_MODE_* */
    uint32_t features; /* covfeat: GF_F_* mask */
    int32_t always_up; /* normal */
    int32_t normalize; /* eigenvalues */
    int32_t min_k;     /* radius mode, covfeat / eigenvalues: rows with fewer
                          neighbors are skipped (GF_ST_FEW_NEIGHBORS); PDAL's
                          default is 3 */
    int32_t reserved;  /* 0 */
    double radius;     /* > 0: the neighborhood is every point within this
                          distance (KD3Index::radius, strict), knn and stride
                          are ignored -- the stages' `radius` option */
    double p1, p2;     /* stage parameters: approximatecoplanar thresh1 and
                          thresh2, estimaterank thresh (p1) */
} gf_stage;

typedef struct gf_stats {
    uint64_t points;
    uint64_t tie_rows;        /* rows re-solved for a boundary tie */
    uint64_t incomplete_rows; /* rows re-solved after the grid budget */
    double ms_index;          /* grid build, device time */
    double ms_search;         /* kNN search + covariance kernel, device time */
    double ms_features;       /* eigensolve + stage columns of the rows the search
                                 did not flag (run right after the search) */
    double ms_exact;          /* from the end of those features to the end of the
                                 flagged rows' features: the part of the tree build
                                 not hidden behind search and features, the exact
                                 KD3 re-solve and the flagged rows' features */
    double cell;              /* column edge used (m) */
    double cell_z;            /* z cell height used (m) */
    uint32_t grid_x, grid_y, grid_z;
    uint32_t key_bits;
    /* search counters, filled when tuning "collect_stats" is 1 */
    uint64_t search_passes;   /* box searches (1 per row when the first box certifies) */
    uint64_t candidates;      /* candidate points evaluated */
    uint64_t merges;          /* wave-wide sorted merges */
    double ms_tree;           /* exact KD3 tree build on its own stream (it overlaps
                                 the search; 0 when the tree was cached) */
    double ms_total;          /* the call's device time, index build to features */
    uint64_t deep_rows;       /* flagged rows whose exact search needed more than 64
                                 recursion frames (a tree deeper than 64 levels) and
                                 ran again with frames in device memory */
    uint64_t graph;           /* batches: 0 launched kernel by kernel, 1 the step was a
                                 replay of a captured graph, 2 captured in this call
                                 and replayed */
} gf_stats;

/* ---- context: replaces pdg::CudaMemoryResource + stream ownership
 *      (src/core/allocator.cu:135-170; include/pdg/Memory.hpp:12-81). */
int gf_ctx_create(int device, gf_ctx **out);
int gf_ctx_destroy(gf_ctx *ctx);
const char *gf_last_error(const gf_ctx *ctx);
int gf_sync(gf_ctx *ctx);
/* Tuning knobs: "cell_factor", "cell_z_factor", "ring0", "ring_max",
 * "small_col", "group_z", "q_per_wave", "knn_waves", "collect_stats" (and
 * the exact tree's).  None of them changes any row's result. */
int gf_set_tuning(gf_ctx *ctx, const char *key, double value);

/* ---- autotuning: replaces the reference's grid-configuration heuristics
 *      (makeKnnGridConfig / makeAdaptiveKnnConfig, src/index/SpatialIndex.cpp
 *      :434-631).  gf_autotune times the whole run over a cloud for each
 *      candidate column-size factor (column edge = factor * sqrt(k / (pi *
 *      2-D density))), keeps the fastest, then tries the search's launch
 *      shape one knob at a time -- z cells a query group spans (0..4),
 *      queries per wave (16, 32, 64), waves per workgroup (1, 2, 4) --
 *      keeping each change that is faster, and stores the result for the
 *      cloud's (density, k) bucket.  Later grids in that bucket (clouds and
 *      batch tiles) use its factor; clouds use its shape, and a batch's one
 *      search launch the shape of its largest tile's bucket.  No row's result
 *      depends on any of them.  The table is saved and loaded as text (one
 *      line per bucket: rho_bucket k_bucket factor group_z+1 q_per_wave
 *      knn_waves, 0 = the context's setting; three-field lines are read as
 *      factor only).  The same knobs can be set for a context with
 *      gf_set_tuning ("group_z", "q_per_wave", "knn_waves"). */
int gf_autotune(gf_ctx *ctx, gf_cloud *cloud, const gf_stage *stages, int nstages,
                const double *factors, int nfactors, double *best_factor, double *best_ms);
int gf_tuning_save(gf_ctx *ctx, const char *path);
int gf_tuning_load(gf_ctx *ctx, const char *path);

/* ---- resident clouds: replaces the PointBatch XYZ upload of
 *      makeResidentNeighborhood (src/pdal/PdgNeighborhood.cpp:1383-1550). */
int gf_cloud_create(gf_ctx *ctx, const double *x, const double *y,
                    const double *z, uint64_t n, int where, gf_cloud **out);
int gf_cloud_release(gf_cloud *cloud);
uint64_t gf_cloud_size(const gf_cloud *cloud);
/* Drops the cached index (the next run rebuilds it), as the reference's
 * SpatialIndex::invalidate (include/pdg/index/SpatialIndex.hpp:241). */
int gf_cloud_invalidate(gf_cloud *cloud);

/* ---- one chain of stages on a resident cloud: replaces
 *      pdg::knnEigenSystems + repairAmbiguousEigenSystems +
 *      project{Normal,Eigenvalue,CovarianceFeature}Columns
 *      (include/pdg/index/SpatialIndex.hpp:499-500;
 *      src/pdal/PdgNeighborhood.cpp:1233-1370;
 *      include/pdg/stages/Neighborhood.hpp:187-199).
 * cols: the stages' output columns in order (gf_stage_columns each), n FP64
 * values per column, original point order.  Rows a stage skips keep their
 * prior values.  status: one n-byte array per stage.  where says whether
 * cols/status are host or device pointers.  A neighborhood of (knn + 1) *
 * stride up to 63 entries runs the grid search (rows it cannot certify go to
 * the exact KD3 search); larger ones, up to 1024, run every row through the
 * exact search; beyond that GF_ERR_UNSUPPORTED. */
int gf_stage_columns(const gf_stage *st);
int gf_run(gf_ctx *ctx, gf_cloud *cloud, const gf_stage *stages, int nstages,
           double *const *cols, uint8_t *const *status, int where);
int gf_last_stats(const gf_ctx *ctx, gf_stats *out);

/* ---- the same fused run with no device allocation.  No counterpart in the
 *      reference: for callers that cannot allocate device memory (code under
 *      HIP graph capture, tight serving loops, harnesses that forbid
 *      hipMalloc).  x, y, z: n FP64 device arrays (original order); cols and
 *      status: device arrays as gf_run's with where = GF_WHERE_DEVICE.  Every
 *      working buffer comes from one fixed array per GPU (24 MiB, held until
 *      the process ends), which limits n to GF_NOALLOC_MAX_POINTS; larger n
 *      returns GF_ERR_ARG.  The call synchronises the context's stream, as
 *      gf_run does.  Calls that overlap on one GPU (from any context) share
 *      the array and race: callers serialise them. */
#define GF_NOALLOC_MAX_POINTS 65536u
int gf_run_noalloc(gf_ctx *ctx, const double *x, const double *y, const double *z,
                   uint64_t n, const gf_stage *stages, int nstages, double *const *cols,
                   uint8_t *const *status);

/* ---- neighbor lists: replaces pdg::knnGather
 *      (include/pdg/index/SpatialIndex.hpp:480-490) and KD3Index::neighbors(id,
 *      k, stride) (pdal/KDIndex.hpp:103-106).
 * ids: n x min(n,k) uint32 (row i = neighbors of point i).  Every row holds
 * exactly KD3Index::neighbors' set.  The order: rows the exact search
 * re-solved are in KD3Index's order -- those with status != 0 (a tie at the
 * boundary or within a strided list, or an uncertified grid search), and
 * every row when k * stride is above 63 or above the cloud's size; the other
 * rows are ascending by (squared distance, point id), which is KD3Index's
 * order wherever no two kept distances are equal.  k * stride up to 1024. */
int gf_knn(gf_ctx *ctx, gf_cloud *cloud, uint32_t k, uint32_t stride,
           uint32_t *ids, uint8_t *status, int where);

/* ---- radius neighbor counts: replaces pdg::detail::radiusCountsDevice
 *      (include/pdg/index/SpatialIndex.hpp:53; src/index/SpatialIndexKernels.cu:
 *      127-228) and the size of KD3Index::radius(id, r)
 *      (pdal/private/KDImpl.hpp:497-516).  counts: n uint32, row i = the number
 *      of points (point i included) within r of point i, strictly. */
int gf_radius_counts(gf_ctx *ctx, gf_cloud *cloud, double radius, uint32_t *counts,
                     int where);

/* ---- batches of independent tiles: replaces running the stages tile by
 *      tile (the reference processes one PointView / one SpatialTile at a time,
 *      include/pdg/index/SpatialTile.hpp:29-40; src/pdal/PdgNeighborhood.cpp
 *      :1383-1550).  The tiles of a batch share one index build, one search
 *      launch, one exact re-solve and one feature launch; every tile's result
 *      equals the tile run alone.  Points of tile t sit at
 *      [gf_batch_tile_start(t), gf_batch_tile_start(t+1)) of the batch, and
 *      gf_batch_run's cols/status use that order. */
typedef struct gf_batch gf_batch;
int gf_batch_create(gf_ctx *ctx, uint32_t ntiles, const uint64_t *counts, gf_batch **out);
int gf_batch_set_tile(gf_batch *batch, uint32_t tile, const double *x, const double *y,
                      const double *z, int where);
uint64_t gf_batch_size(const gf_batch *batch);
uint64_t gf_batch_tile_start(const gf_batch *batch, uint32_t tile);
int gf_batch_invalidate(gf_batch *batch);
/* All stages share one neighborhood; tile_knn (NULL: the stages' knn) gives
 * each tile its own knn, overriding every stage's. */
int gf_batch_run(gf_ctx *ctx, gf_batch *batch, const gf_stage *stages, int nstages,
                 const int32_t *tile_knn, double *const *cols, uint8_t *const *status,
                 int where);
int gf_batch_release(gf_batch *batch);

/* ---- serving: no counterpart in the reference, which runs one pipeline
 *      per process (src/cli/main.cpp:1981-2207).  Requests (one tile each,
 *      host arrays that stay valid until the request is done) are admitted
 *      under an HBM budget by worker threads, each with its own context and
 *      stream, and run as batches (continuous batching: a request that arrives
 *      while batches are on the device joins the next batch a free worker
 *      forms).  A request's result equals the tile run alone.  Requests of a
 *      batch share the stage shape (kinds, features, modes, stride); knn may
 *      differ.  hbm_budget_bytes 0 = 200 GB; max_batch_points 0 = 64 Mi. */
typedef struct gf_server gf_server;
int gf_server_create(int device, int workers, uint64_t hbm_budget_bytes,
                     uint64_t max_batch_points, gf_server **out);
int gf_server_submit(gf_server *server, const double *x, const double *y, const double *z,
                     uint64_t n, const gf_stage *stages, int nstages, int32_t knn,
                     double *const *cols, uint8_t *const *status, uint64_t *ticket);
/* Blocks until the request is done (timeout_s < 0: no limit); returns its
 * result code, or GF_ERR_TIMEOUT if the timeout expired first (the request
 * is still queued or running).  Requests are checked when submitted: a stage
 * gf_run would refuse (bad kind, mode, knn or stride; knn and stride outside
 * the device envelope; more than 32 columns) is refused by gf_server_submit
 * and never joins a batch. */
int gf_server_wait(gf_server *server, uint64_t ticket, double timeout_s);
/* times: submit, batch start, done (seconds, monotonic clock). */
int gf_server_request_info(gf_server *server, uint64_t ticket, double *times,
                           uint32_t *batch_id, uint32_t *batch_tiles);
int gf_server_release(gf_server *server, uint64_t ticket);
int gf_server_destroy(gf_server *server);

/* ---- device buffers for callers without their own HIP allocator (the
 *      reference's MemoryResource::allocate/deallocate,
 *      include/pdg/Memory.hpp:12-81).  Copies are synchronous. */
int gf_dev_alloc(gf_ctx *ctx, uint64_t bytes, void **out);
int gf_dev_free(gf_ctx *ctx, void *ptr);
int gf_memcpy(gf_ctx *ctx, void *dst, const void *src, uint64_t bytes, int kind);
#define GF_COPY_H2D 1
#define GF_COPY_D2H 2
#define GF_COPY_D2D 3

/* filters.optimalneighborhood (filters/OptimalNeighborhoodFilter.cpp:60-163):
 * for every point the k in [min_k, max_k] (defaults 10, 14; neighbors in
 * KD3Index order, the point itself first) whose covariance has the smallest
 * eigenentropy -> optimal_knn (OptimalKNN), and the distance to that k-th
 * neighbor -> optimal_radius (OptimalRadius); n values each.  Needs
 * 2 <= min_k <= max_k <= 63 and at least max_k points.  An eigen solve that
 * does not converge: GF_ERR_UNSUPPORTED with PDAL's message. */
int gf_optimal_neighborhood(gf_ctx *ctx, gf_cloud *cloud, uint32_t min_k, uint32_t max_k,
                            uint64_t *optimal_knn, double *optimal_radius, int where);

/* covariancefeatures' optimized=true (filters/CovarianceFeaturesFilter.cpp:
 * 204-210: KD3Index::neighbors(p, OptimalKNN, 1)), for any stage kinds:
 * row i's neighborhood is its row_k[i] nearest points in KD3Index order
 * (clamped to the cloud's size).  stages' knn and stride are not used
 * (radius must be 0); row_k holds n counts (host or device, as `where`), at
 * most 1024.  cols/status as gf_run's; no search flags are reported (the lists
 * are exact).  A row with k = 0 has a zero covariance (skipped); k = 1 gives
 * 0/0 covariances, as PDAL's division by k - 1 does. */
int gf_run_rowk(gf_ctx *ctx, gf_cloud *cloud, const gf_stage *stages, int nstages,
                const uint32_t *row_k, double *const *cols, uint8_t *const *status, int where);

/* filters.nndistance (filters/NNDistanceFilter.cpp:97-133): per point, from
 * its k + 1 nearest points (itself included, KD3Index::knnSearch) in
 * ascending distance: GF_NND_KTH the distance to the last of them, GF_NND_AVG
 * the mean distance to all but the first.  out: n FP64 (NNDistance).  With
 * fewer than k + 1 points the missing distances count as 0, as in PDAL. */
#define GF_NND_KTH 0
#define GF_NND_AVG 1
int gf_nndistance(gf_ctx *ctx, gf_cloud *cloud, uint32_t k, int mode, double *out, int where);

/* filters.outlier (filters/OutlierFilter.cpp:78-172).  GF_OUTLIER_STATISTICAL
 * (mean_k, multiplier): each point's running mean distance to its mean_k
 * nearest neighbors; the mean and sample standard deviation of those over the
 * cloud (online moments in point order) give threshold = mean +
 * multiplier * stdev, and a point at or above it is an outlier.
 * GF_OUTLIER_RADIUS (radius, min_k): a point with at most min_k points
 * within radius (itself included) is an outlier.  outlier: n host bytes,
 * 1 = outlier (PDAL gives these the noise class); threshold (host, may be
 * NULL) receives the statistical threshold. */
#define GF_OUTLIER_STATISTICAL 0
#define GF_OUTLIER_RADIUS 1
int gf_outlier(gf_ctx *ctx, gf_cloud *cloud, int method, uint32_t mean_k, double multiplier,
               double radius, int32_t min_k, uint8_t *outlier, double *threshold);

/* ---- LAS point records in and out (the on-disk format either side of the
 * path).  Replaces GPUPDAL's device LAS hydration, src/pdal/PdgNeighborhood.cpp:
 * 1067-1193 with src/io/las/LasKernels.cu:306-319, for uncompressed LAS. */

/* A LAS file's public header as PDAL reads it (io/private/las/Header.cpp:
 * 38-75; point count and format bits, Header.hpp:122-138). */
typedef struct gf_las_info {
    uint8_t version_major, version_minor;
    uint8_t point_format;      /* format bits & 0xF */
    uint8_t compressed;        /* LAZ bit 0x80 */
    uint16_t header_size;
    uint16_t point_size;       /* point record length in bytes */
    uint32_t point_offset;     /* offset of the first point record */
    uint32_t reserved;
    uint64_t point_count;      /* LAS 1.4: the 64-bit count; else the 32-bit one */
    double scale[3], offset[3];
    double max[3], min[3];     /* bounds as stored in the header */
    double ms_decode;          /* gf_cloud_create_las: device time of the decode
                                  (tuning "timing" 1), else 0 */
} gf_las_info;

/* Parses and validates the header of a LAS file image (host memory; no GPU
 * needed).  GF_ERR_ARG with PDAL's validation message in msg (if msg_size >
 * 0) for a file PDAL's reader rejects (io/private/las/Header.cpp:118-144:
 * signature, point offset and count against the file size, point formats
 * 0-3 and 6-8, record length at least the format's, version 1.0-1.4). */
int gf_las_parse(const void *bytes, uint64_t nbytes, gf_las_info *info, char *msg,
                 uint64_t msg_size);
/* A resident cloud from the point records of an uncompressed LAS file image
 * in host memory: the records are copied to HBM once and X, Y, Z decoded
 * there as int32 * scale + offset (io/LasReader.cpp:926-932, 1008-1014), in
 * file order.  LAZ-compressed data: GF_ERR_UNSUPPORTED.  info may be NULL. */
int gf_cloud_create_las(gf_ctx *ctx, const void *bytes, uint64_t nbytes, gf_las_info *info,
                        gf_cloud **out);
/* Writes ncols double columns (n rows each) into n point records of stride
 * bytes at byte `offset` of each record, one after another, little-endian:
 * the extra bytes that writers.las extra_dims appends after the base record
 * (io/LasWriter.cpp).  where: GF_WHERE_HOST (host columns and records) or
 * GF_WHERE_DEVICE.  offset + 8 * ncols must not exceed stride; ncols <= 16. */
int gf_las_pack_extra(gf_ctx *ctx, const double *const *cols, int ncols, uint64_t n,
                      uint8_t *records, uint32_t stride, uint32_t offset, int where);

/* writers.las extra_dims (io/LasWriter.cpp:583-596 with
 * io/private/las/Utils.cpp:244-267): a new LAS file image from an
 * uncompressed source image (host memory), every point record followed by
 * the ncols (1..16) columns, little-endian, each as its LAS data type
 * (types[c] in 1..10 = u8, i8, u16, i16, u32, i32, u64, i64, f32, f64;
 * types NULL: all f64; integer types take the value rounded toward zero,
 * held to the type's range, NaN as 0), described by Extra Bytes descriptors
 * (names and descriptions zero-padded to 32 bytes, longer ones truncated)
 * appended to the source's Extra Bytes VLR, or in a new one ("LASF_Spec",
 * record 4) after the last VLR.  Source record
 * bytes that no descriptor covers are declared as undocumented extra bytes
 * first, so each column is described at its offset.  The header's point
 * offset, VLR count and record length are updated, and the waveform and
 * EVLR offsets (LAS 1.3/1.4) moved past the wider records; everything else
 * is copied.  The records are widened and packed on the device.  cols: n
 * (the source's point count) values each, host or device as `where`.
 * *dst_nbytes receives the image's size; dst NULL asks for the size only,
 * and a dst_capacity below it is GF_ERR_ARG. */
int gf_las_write_extra(gf_ctx *ctx, const void *src, uint64_t src_nbytes, int ncols,
                       const char *const *names, const char *const *descriptions,
                       const uint8_t *types, const double *const *cols, int where, void *dst,
                       uint64_t dst_capacity, uint64_t *dst_nbytes);

/* ---- parity hooks (no reference counterpart; used by the test suite). */
/* The device eigensolver on m row-major 3x3 matrices. */
int gf_debug_eigen(gf_ctx *ctx, const double *cov, uint64_t m, double *evals,
                   double *evecs, int32_t *info);
/* The grid a cloud gets for a given k: sorted permutation (n), sorted keys
 * (n) and column starts (ncols+1; pass NULL to query sizes via *ncols). */
int gf_debug_grid(gf_ctx *ctx, gf_cloud *cloud, uint32_t k, uint32_t *perm,
                  uint32_t *keys, uint32_t *colstart, uint32_t *ncols,
                  uint32_t *key_bits);
/* The parameters of that grid: f = {origin x, y, z, column edge, z cell
 * height, 1/column edge, 1/z cell height}, u = {columns along x, along y,
 * z cells, 8x8 column tiles along x, z bits of the key, key bits, padded
 * column count, 0}.  A point's key is (column << zbits) | zcell with
 * column = tile * 64 + the Morton interleave of the low 3 bits of its x and
 * y columns (x in the even bits), tile = (ycol >> 3) * tiles_x + (xcol >> 3),
 * and each cell coordinate floor((v - origin) * inverse edge) clamped to
 * [0, cells - 1] (NaN: 0). */
int gf_debug_grid_params(gf_ctx *ctx, gf_cloud *cloud, uint32_t k, double *f, uint32_t *u);

/* The cloud's resident coordinates (n each, original order) copied to host
 * arrays. */
int gf_debug_cloud_xyz(gf_ctx *ctx, gf_cloud *cloud, double *x, double *y, double *z);
/* A plain device-to-device copy of count doubles (the reference kernel that
 * calibrates HBM byte counters). */
int gf_debug_copy(gf_ctx *ctx, double *dst, const double *src, uint64_t count);
/* The cloud's exact KD3Index tree (built on demand): vind (n, tree order ->
 * point id) and the node records; pass NULL arrays to query the record
 * count.  Leaves have feat = -1; records with an empty range [left, right)
 * are ids reserved during the build and not used by any node. */
int gf_debug_kdtree(gf_ctx *ctx, gf_cloud *cloud, uint32_t *vind, uint32_t *nodes,
                    uint32_t *left, uint32_t *right, int32_t *feat, double *divlow,
                    double *divhigh);

/* ---- seeded synthetic clouds (the benchmark generator). */
typedef struct gf_synth {
    uint64_t seed;          /* tile seed (splitmix64(0x5EED ^ tile) by convention) */
    uint64_t n;             /* points */
    double side;            /* tile edge in metres (XY uniform over side^2) */
    double ground_frac;     /* fraction on the ground surface (0.7) */
    double roof_frac;       /* fraction on exactly planar roof patches */
    double wire_frac;       /* fraction on exactly collinear wires */
    double noise;           /* Gaussian sigma on xyz (0.05) */
    double quantum;         /* LAS scale (0.01) */
    double offset_x, offset_y; /* (500000, 4000000) */
} gf_synth;
void gf_synth_defaults(gf_synth *p);
uint64_t gf_tile_seed(uint64_t base, uint64_t tile);
int gf_synth_cloud(const gf_synth *p, double *x, double *y, double *z);

#ifdef __cplusplus
}
#endif
#endif
