/*
 * oracle.h -- CPU restatement of PDAL's kNN-neighborhood feature path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU path is
 * compared against, and the scalar CPU baseline that bench.py times beside the
 * GPU.  It is loaded only by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg.  The product (libgpupdal_amd.so and the C++
 * stage layer) never links, loads or calls it.
 *
 * What it restates (reference paths are relative to the GPUPDAL tree):
 *   - KD3Index = nanoflann 1.3.1 KDTreeSingleIndexAdaptor over PDAL's KD3
 *     adaptor: leaf size 100, PointView bounds as the root box, middleSplit_,
 *     planeSplit, findNeighbors/searchLevel and KNNResultSet
 *     (pdal/private/KDImpl.hpp:209-381; vendor/nanoflann/nanoflann.hpp:145-208,
 *     619-875, 1012-1080, 1154-1248).
 *   - math::computeCentroid / computeCovariance (pdal/private/MathUtils.cpp:139-192).
 *   - Eigen 3.4 SelfAdjointEigenSolver<Matrix3d>::compute for a 3x3 real
 *     matrix: scaling, the 3x3 tridiagonalization, implicit symmetric QR with
 *     Wilkinson shift, ascending selection sort
 *     (vendor/eigen/Eigen/src/Eigenvalues/SelfAdjointEigenSolver.h:420-575,
 *     839-890; Tridiagonalization.h:464-504; Jacobi/Jacobi.h:231-266).
 *   - The stage formulas of filters.covariancefeatures, filters.normal and
 *     filters.eigenvalues (filters/CovarianceFeaturesFilter.cpp:204-367,
 *     filters/NormalFilter.cpp:140-182, pdal/private/MathUtils.cpp:572-611,
 *     filters/EigenvaluesFilter.cpp:115-172).
 *   - An extended-precision "truth" path (long double two-pass covariance
 *     without float narrowing, cyclic Jacobi to convergence) that the GPU's
 *     features are held to within 1e-5 relative.
 *   - Brute-force kNN ordered by (squared distance, point index).
 *
 * Parity pins (each checked in tests/):
 *   kd3_*       : KATs of test/unit/KDIndexTest.cpp:118-222; brute force on
 *                 tie-free clouds; oracle/_ref (the reference's own nanoflann
 *                 compiled from vendor/) on tie-heavy clouds, node by node.
 *   pdal_cov    : closed forms; oracle/_ref (Eigen product) bit for bit.
 *   eigen_sym3  : oracle/_ref (the reference's Eigen) bit for bit; numpy eigh.
 *   stage_*     : KATs of test/unit/filters/CovarianceFeaturesTest.cpp:46-188
 *                 and test/unit/filters/NormalFilterTest.cpp:47-253.
 *   truth_*     : numpy/LAPACK eigh in float64, analytic planes and lines.
 */
#ifndef GPUPDAL_ORACLE_H
#define GPUPDAL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- KD3 --- */
typedef struct orc_kd3 orc_kd3;

/* Builds the tree over n points (SoA, FP64).  Arrays must outlive the tree. */
orc_kd3 *orc_kd3_build(const double *x, const double *y, const double *z,
                       uint64_t n);
void orc_kd3_free(orc_kd3 *t);

/* KD3ImplT::neighbors(x, y, z, k, stride): returns the count written (the
 * clamped k).  ids/d2 need room for stride*min(n,k) entries; ids receives the
 * strided selection, d2 the un-strided squared distances of the result set. */
uint64_t orc_kd3_neighbors(const orc_kd3 *t, double qx, double qy, double qz,
                           uint64_t k, uint64_t stride, uint64_t *ids,
                           double *d2);

/* neighbors(point i, k, stride) for every point; out is n x min(n,k),
 * row-major.  threads > 1 splits the rows into contiguous chunks. */
void orc_kd3_neighbors_all(const orc_kd3 *t, uint64_t k, uint64_t stride,
                           int threads, uint32_t *out);
/* The same for the points rows[0..nrows): out is nrows x min(n,k). */
void orc_kd3_neighbors_rows(const orc_kd3 *t, const uint64_t *rows, uint64_t nrows,
                            uint64_t k, uint64_t stride, int threads, uint32_t *out);

/* Tree introspection (nodes in allocation order = DFS preorder). */
uint64_t orc_kd3_node_count(const orc_kd3 *t);
void orc_kd3_export(const orc_kd3 *t, uint64_t *vind, int32_t *child1,
                    int32_t *child2, uint64_t *left, uint64_t *right,
                    int32_t *divfeat, double *divlow, double *divhigh);

/* ------------------------------------------------------- brute force --- */
/* For each query index q in [0,nq): the kk nearest points of the cloud to
 * point q (query = the point's own coordinates), ordered by (d2, index).
 * ids/d2 are nq x kk. */
void orc_brute_knn(const double *x, const double *y, const double *z,
                   uint64_t n, const uint64_t *queries, uint64_t nq,
                   uint64_t kk, int threads, uint32_t *ids, double *d2);

/* ------------------------------------------------------ PDAL arithmetic --- */
/* math::computeCovariance over ids[0..n): full 3x3, row-major. */
void orc_pdal_covariance(const double *x, const double *y, const double *z,
                         const uint64_t *ids, uint64_t n, double cov[9]);
/* math::computeCovariance for 32-bit id lists (same arithmetic). */
void orc_pdal_covariance_u32(const double *x, const double *y, const double *z,
                             const uint32_t *ids, uint64_t n, double cov[9]);
/* Eigen::MatrixBase::isZero() with the double dummy precision 1e-12. */
int orc_is_zero(const double cov[9]);
/* SelfAdjointEigenSolver<Matrix3d>(cov): returns 0 (Success) or 1
 * (NoConvergence).  evals ascending; evecs column-major (column j is the
 * eigenvector of evals[j], evecs[3*j + axis]). */
int orc_eigen_sym3(const double cov[9], double evals[3], double evecs[9]);

/* ----------------------------------------------------------- stages --- */
enum {
    ORC_STAGE_COVFEAT = 0,
    ORC_STAGE_NORMAL = 1,
    ORC_STAGE_EIGENVALUES = 2,
    ORC_STAGE_COPLANAR = 3, /* filters.approximatecoplanar (neighbors(p, knn)) */
    ORC_STAGE_RANK = 4      /* filters.estimaterank (neighbors(p, knn))       */
};
enum { ORC_MODE_RAW = 0, ORC_MODE_SQRT = 1, ORC_MODE_NORMALIZED = 2 };

/* Feature bit order of include/pdg/stages/Neighborhood.hpp:153-165. */
enum {
    ORC_F_LINEARITY = 1u << 0,
    ORC_F_PLANARITY = 1u << 1,
    ORC_F_SCATTERING = 1u << 2,
    ORC_F_VERTICALITY = 1u << 3,
    ORC_F_OMNIVARIANCE = 1u << 4,
    ORC_F_ANISOTROPY = 1u << 5,
    ORC_F_EIGENENTROPY = 1u << 6,
    ORC_F_EIGENVALUESUM = 1u << 7,
    ORC_F_SURFACEVARIATION = 1u << 8,
    ORC_F_DEMANTKEVERTICALITY = 1u << 9
};

/* Row status bits (shared with the product's C ABI). */
enum {
    ORC_ST_TIE = 1u,           /* boundary distance tie (kNN only)       */
    ORC_ST_INCOMPLETE = 2u,    /* search not certified (GPU only)        */
    ORC_ST_ZERO_COV = 4u,      /* isZero covariance: row skipped         */
    ORC_ST_EIGEN_FAIL = 8u,    /* solver NoConvergence                   */
    ORC_ST_FEATURE_INVALID = 16u, /* covfeat: largest clamped lambda == 0 */
    ORC_ST_FEW_NEIGHBORS = 32u /* normal: fewer than 3 neighbors        */
};

typedef struct {
    int32_t stage;     /* ORC_STAGE_* */
    int32_t knn;       /* the user's knn option (the search asks knn+1) */
    int32_t stride;    /* covfeat/eigenvalues stride option (normal: 1) */
    int32_t mode;      /* covfeat: ORC_MODE_* */
    uint32_t features; /* covfeat: ORC_F_* mask */
    int32_t always_up; /* normal */
    int32_t normalize; /* eigenvalues */
    int32_t min_k;     /* radius mode, covfeat/eigenvalues: fewer neighbors -> skip */
    int32_t pad;
    double radius;     /* > 0: neighbors within this radius (knn, stride ignored) */
    double p1, p2;     /* approximatecoplanar thresh1/thresh2; estimaterank thresh */
} orc_stage;

/* Eigen's JacobiSVD::rank() of a symmetric PSD 3x3 from its eigenvalues (the
 * singular values are their magnitudes), threshold narrowed to float as
 * math::computeRank does (pdal/private/MathUtils.cpp:197-208). */
int orc_rank_from_eigenvalues(const double ev[3], double thresh);

/* Number of output columns a stage writes, in PDAL's registration order. */
int orc_stage_columns(const orc_stage *st);

/* PDAL's CPU path for one stage over every point of the cloud, using the KD3
 * tree (built once per view and shared by chained stages).  cols[c] is an
 * n-long FP64 column; rows PDAL skips keep their prior values.  status gets
 * one byte per row.  threads > 1 uses contiguous chunks of rows. */
void orc_stage_run(const orc_kd3 *t, const double *x, const double *y,
                   const double *z, uint64_t n, const orc_stage *st,
                   int threads, double *const *cols, uint8_t *status);

/* orc_stage_run restricted to rows [row_begin, row_end) (the CPU baseline's
 * bounded sample); the tree still covers the whole cloud. */
void orc_stage_run_range(const orc_kd3 *t, const double *x, const double *y,
                         const double *z, uint64_t n, const orc_stage *st,
                         int threads, uint64_t row_begin, uint64_t row_end,
                         double *const *cols, uint8_t *status);

/* The same stage arithmetic applied to caller-given neighbor lists
 * (nbr: n x kk, uint32).  Used to replay GPU neighbor sets on the CPU. */
void orc_stage_from_lists(const double *x, const double *y, const double *z,
                          uint64_t n, const uint32_t *nbr, uint64_t kk,
                          const orc_stage *st, int threads,
                          double *const *cols, uint8_t *status);

/* KD3Index::radius (pdal/private/KDImpl.hpp:398-415): the points with
 * squared distance strictly below r*r, sorted by distance (equal distances in
 * walk order); fills up to cap ids/d2, returns the full count. */
uint64_t orc_kd3_radius(const orc_kd3 *t, double qx, double qy, double qz, double r,
                        uint64_t cap, uint64_t *ids, double *d2);
/* Neighbor counts within r of every point (the point itself included). */
void orc_kd3_radius_counts(const orc_kd3 *t, double r, int threads, uint32_t *counts);
/* The same counts by brute force over all points, for nq query rows
 * (queries NULL: rows 0..nq-1). */
void orc_brute_radius_counts(const double *x, const double *y, const double *z, uint64_t n,
                             double r, const uint64_t *queries, uint64_t nq, uint32_t *counts);

/* The truth (below) from variable-length neighbor lists: row i's neighbors
 * are ids[offsets[i] .. offsets[i+1]).  Radius stages apply min_k. */
void orc_truth_from_csr(const double *x, const double *y, const double *z, uint64_t n,
                        const uint64_t *offsets, const uint32_t *ids, const orc_stage *st,
                        int threads, double *const *cols, uint8_t *status, double *gap,
                        double *ratio);

/* Extended-precision truth for the same stage from neighbor lists: long
 * double two-pass covariance without float narrowing, cyclic Jacobi to
 * convergence.  gap (optional, n entries) receives a conditioning figure per
 * row: min over adjacent raw eigenvalue gaps divided by the largest; ratio
 * (optional) the smallest eigenvalue (clamped at 0) over the largest. */
void orc_truth_from_lists(const double *x, const double *y, const double *z,
                          uint64_t n, const uint32_t *nbr, uint64_t kk,
                          const orc_stage *st, int threads,
                          double *const *cols, uint8_t *status, double *gap,
                          double *ratio);

/* filters.optimalneighborhood (filters/OptimalNeighborhoodFilter.cpp:72-163)
 * for every point: KD3Index::knnSearch with kmax neighbors, the running
 * mean and scatter matrix in PDAL's order, the Eigen-semantics solve of
 * B / (n - 1) for k in [kmin, kmax], eigenentropy with the C library's log;
 * kopt[i] = OptimalKNN, ropt[i] = OptimalRadius; gap (optional) the
 * smallest |entropy - best| over the other k with a finite entropy,
 * relative to |best| (how close the choice was).  Returns the number of rows
 * whose solve failed (PDAL throws). */
uint64_t orc_optimal_neighborhood(const orc_kd3 *t, const double *x, const double *y,
                                  const double *z, uint64_t n, uint64_t kmin, uint64_t kmax,
                                  int threads, uint64_t *kopt, double *ropt, double *gap);

/* The stage with each row's own neighbor count: KD3Index::neighbors(p,
 * row_k[i], 1) (covariancefeatures' optimized=true); the stage's knn and
 * stride are not used. */
void orc_stage_run_rowk(const orc_kd3 *t, const double *x, const double *y, const double *z,
                        uint64_t n, const orc_stage *st, const uint32_t *row_k, int threads,
                        double *const *cols, uint8_t *status);

/* filters.nndistance per point from its k + 1 nearest (mode 0 kth, 1 avg),
 * or mode 2: filters.outlier's statistical running mean distance. */
void orc_nn_distance(const orc_kd3 *t, const double *x, const double *y, const double *z,
                     uint64_t n, uint64_t k, int mode, int threads, double *out);
/* filters.outlier, statistical and radius: outlier[i] = 1 for an outlier;
 * returns the number of outliers.  threshold and dist (n) may be NULL. */
uint64_t orc_outlier_statistical(const orc_kd3 *t, const double *x, const double *y,
                                 const double *z, uint64_t n, uint64_t mean_k,
                                 double multiplier, int threads, uint8_t *outlier,
                                 double *threshold, double *dist);
uint64_t orc_outlier_radius(const orc_kd3 *t, uint64_t n, double r, int64_t min_k,
                            int threads, uint8_t *outlier);

/* Extended-precision symmetric 3x3 eigensolve (cyclic Jacobi): ascending
 * values, column-major vectors.  Returns the sweep count. */
int orc_jacobi_sym3(const double cov[9], double evals[3], double evecs[9]);

#ifdef __cplusplus
}
#endif
#endif
