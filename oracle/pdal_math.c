/*
 * pdal_math.c -- restatement of PDAL's covariance, Eigen's 3x3 self-adjoint
 * eigensolver and the three stages' formulas, plus the extended-precision
 * truth path.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------- covariance --- */
/* math::computeCentroid (MathUtils.cpp:139-166): ordered online mean,
 * avg += (v - avg) / n with n counted from 1.
 * math::computeCovariance (MathUtils.cpp:168-192): demeaned coordinates
 * narrowed to float, A * A^T accumulated in double from the first term in id
 * order, divided by (n - 1). */
#define PDAL_COV_BODY(IDT)                                                    \
    double mx = 0.0, my = 0.0, mz = 0.0;                                      \
    for (uint64_t j = 0; j < n; ++j) {                                        \
        double cnt = (double)(j + 1);                                         \
        IDT id = ids[j];                                                      \
        mx = mx + (x[id] - mx) / cnt;                                         \
        my = my + (y[id] - my) / cnt;                                         \
        mz = mz + (z[id] - mz) / cnt;                                         \
    }                                                                         \
    double s[6] = {0, 0, 0, 0, 0, 0};                                         \
    for (uint64_t j = 0; j < n; ++j) {                                        \
        IDT id = ids[j];                                                      \
        double a0 = (double)(float)(x[id] - mx);                              \
        double a1 = (double)(float)(y[id] - my);                              \
        double a2 = (double)(float)(z[id] - mz);                              \
        double p[6] = {a0 * a0, a0 * a1, a0 * a2, a1 * a1, a1 * a2, a2 * a2}; \
        for (int c = 0; c < 6; ++c) s[c] = (j == 0) ? p[c] : s[c] + p[c];     \
    }                                                                         \
    double den = (double)(n - 1);                                             \
    cov[0] = s[0] / den;                                                      \
    cov[1] = s[1] / den;                                                      \
    cov[2] = s[2] / den;                                                      \
    cov[3] = cov[1];                                                          \
    cov[4] = s[3] / den;                                                      \
    cov[5] = s[4] / den;                                                      \
    cov[6] = cov[2];                                                          \
    cov[7] = cov[5];                                                          \
    cov[8] = s[5] / den;

void orc_pdal_covariance(const double *x, const double *y, const double *z,
                         const uint64_t *ids, uint64_t n, double cov[9])
{
    PDAL_COV_BODY(uint64_t)
}

void orc_pdal_covariance_u32(const double *x, const double *y, const double *z,
                             const uint32_t *ids, uint64_t n, double cov[9])
{
    PDAL_COV_BODY(uint32_t)
}

int orc_is_zero(const double cov[9])
{
    for (int i = 0; i < 9; ++i)
        if (!(fabs(cov[i]) <= 1e-12)) return 0;
    return 1;
}

/* --------------------------------------------------- Eigen 3x3 solver --- */
/* JacobiRotation::makeGivens, real case (Jacobi.h:231-266). */
static void givens(double p, double q, double *c, double *s)
{
    if (q == 0.0) {
        *c = p < 0.0 ? -1.0 : 1.0;
        *s = 0.0;
    } else if (p == 0.0) {
        *c = 0.0;
        *s = q < 0.0 ? 1.0 : -1.0;
    } else if (fabs(p) > fabs(q)) {
        double t = q / p;
        double u = sqrt(1.0 + t * t);
        if (p < 0.0) u = -u;
        *c = 1.0 / u;
        *s = -t * *c;
    } else {
        double t = p / q;
        double u = sqrt(1.0 + t * t);
        if (q < 0.0) u = -u;
        *s = -1.0 / u;
        *c = -t * *s;
    }
}

/* numext::hypot -> positive_real_hypot(|x|, |y|) (MathFunctionsImpl.h:80-95). */
static double eig_hypot(double x, double y)
{
    x = fabs(x);
    y = fabs(y);
    if (isinf(x) || isinf(y)) return INFINITY;
    if (isnan(x) || isnan(y)) return NAN;
    double p = x < y ? y : x; /* maxi */
    double m = y < x ? y : x; /* mini */
    if (p == 0.0) return 0.0;
    double qp = m / p;
    return p * sqrt(1.0 + qp * qp);
}

/* tridiagonal_qr_step (SelfAdjointEigenSolver.h:839-890); q column-major. */
static void qr_step(double *diag, double *sub, int start, int end, double *q)
{
    double td = (diag[end - 1] - diag[end]) * 0.5;
    double e = sub[end - 1];
    double mu = diag[end];
    if (td == 0.0) {
        mu -= fabs(e);
    } else if (e != 0.0) {
        double e2 = e * e;
        double h = eig_hypot(td, e);
        if (e2 == 0.0)
            mu -= e / ((td + (td > 0.0 ? h : -h)) / e);
        else
            mu -= e2 / (td + (td > 0.0 ? h : -h));
    }
    double xx = diag[start] - mu;
    double zz = sub[start];
    for (int k = start; k < end && zz != 0.0; ++k) {
        double c, s;
        givens(xx, zz, &c, &s);
        double sdk = s * diag[k] + c * sub[k];
        double dkp1 = s * sub[k] + c * diag[k + 1];
        diag[k] = c * (c * diag[k] - s * sub[k]) - s * (c * sub[k] - s * diag[k + 1]);
        diag[k + 1] = s * sdk + c * dkp1;
        sub[k] = c * sdk - s * dkp1;
        if (k > start) sub[k - 1] = c * sub[k - 1] - s * zz;
        xx = sub[k];
        if (k < end - 1) {
            zz = -s * sub[k + 1];
            sub[k + 1] = c * sub[k + 1];
        }
        /* Q = Q * G: columns k and k+1 with the transposed rotation. */
        for (int i = 0; i < 3; ++i) {
            double xi = q[3 * k + i], yi = q[3 * (k + 1) + i];
            q[3 * k + i] = c * xi + (-s) * yi;
            q[3 * (k + 1) + i] = s * xi + c * yi;
        }
    }
}

int orc_eigen_sym3(const double cov[9], double evals[3], double evecs[9])
{
    /* Lower triangle, row-major m[r][c] = cov[3r+c] for c <= r. */
    double m00 = cov[0], m10 = cov[3], m11 = cov[4], m20 = cov[6],
           m21 = cov[7], m22 = cov[8];
    for (int i = 0; i < 9; ++i)
        if (isnan(cov[i])) {
            for (int j = 0; j < 3; ++j) evals[j] = NAN;
            for (int j = 0; j < 9; ++j) evecs[j] = NAN;
            return 1;
        }
    /* scale = max |coeff| of the lower-triangular copy (upper part zero). */
    double scale = 0.0;
    double lo[6] = {m00, m10, m11, m20, m21, m22};
    for (int i = 0; i < 6; ++i)
        if (fabs(lo[i]) > scale) scale = fabs(lo[i]);
    if (scale == 0.0) scale = 1.0;
    m00 /= scale;
    m10 /= scale;
    m11 /= scale;
    m20 /= scale;
    m21 /= scale;
    m22 /= scale;

    /* tridiagonalization_inplace_selector<Matrix3d,3,false>. */
    double diag[3], sub[2], q[9];
    diag[0] = m00;
    double v1norm2 = m20 * m20;
    if (v1norm2 <= DBL_MIN) {
        diag[1] = m11;
        diag[2] = m22;
        sub[0] = m10;
        sub[1] = m21;
        double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        memcpy(q, id, sizeof(q));
    } else {
        double beta = sqrt(m10 * m10 + v1norm2);
        double inv = 1.0 / beta;
        double m01 = m10 * inv;
        double m02 = m20 * inv;
        double qq = 2.0 * m01 * m21 + m02 * (m22 - m11);
        diag[1] = m11 + m02 * qq;
        diag[2] = m22 - m02 * qq;
        sub[0] = beta;
        sub[1] = m21 - m01 * qq;
        /* rows (1,0,0), (0,m01,m02), (0,m02,-m01); stored column-major. */
        q[0] = 1; q[3] = 0;   q[6] = 0;
        q[1] = 0; q[4] = m01; q[7] = m02;
        q[2] = 0; q[5] = m02; q[8] = -m01;
    }

    /* computeFromTridiagonal_impl (SelfAdjointEigenSolver.h:496-575). */
    const double tiny = DBL_MIN;
    const double prec_inv = 1.0 / DBL_EPSILON;
    int end = 2, start = 0, iter = 0;
    const int max_iter = 30;
    while (end > 0) {
        for (int i = start; i < end; ++i) {
            if (fabs(sub[i]) < tiny) {
                sub[i] = 0.0;
            } else {
                double sc = prec_inv * sub[i];
                if (sc * sc <= fabs(diag[i]) + fabs(diag[i + 1])) sub[i] = 0.0;
            }
        }
        while (end > 0 && sub[end - 1] == 0.0) end--;
        if (end <= 0) break;
        iter++;
        if (iter > max_iter * 3) break;
        start = end - 1;
        while (start > 0 && sub[start - 1] != 0.0) start--;
        qr_step(diag, sub, start, end, q);
    }
    int info = (iter <= max_iter * 3) ? 0 : 1;
    if (info == 0) {
        for (int i = 0; i < 2; ++i) {
            int k = 0;
            double best = diag[i];
            for (int j = 1; j < 3 - i; ++j)
                if (diag[i + j] < best) {
                    best = diag[i + j];
                    k = j;
                }
            if (k > 0) {
                double t = diag[i];
                diag[i] = diag[i + k];
                diag[i + k] = t;
                for (int r = 0; r < 3; ++r) {
                    double u = q[3 * i + r];
                    q[3 * i + r] = q[3 * (i + k) + r];
                    q[3 * (i + k) + r] = u;
                }
            }
        }
    }
    for (int i = 0; i < 3; ++i) evals[i] = diag[i] * scale;
    memcpy(evecs, q, sizeof(q));
    return info;
}

/* ------------------------------------------------------------ stages --- */
int orc_stage_columns(const orc_stage *st)
{
    switch (st->stage) {
    case ORC_STAGE_NORMAL: return 4;
    case ORC_STAGE_EIGENVALUES: return 3;
    case ORC_STAGE_COPLANAR: return 1;
    case ORC_STAGE_RANK: return 1;
    default: {
        int c = 0;
        for (int b = 0; b < 10; ++b)
            if (st->features & (1u << b)) ++c;
        return c;
    }
    }
}

/* Feature column values for one row from an ascending eigensystem; the
 * arithmetic follows CovarianceFeaturesFilter::setDimensionality
 * (CovarianceFeaturesFilter.cpp:249-365) operation by operation. */
static uint8_t covfeat_row(const orc_stage *st, const double ev[3],
                           const double vec[9], double *vals)
{
    double lam[3];
    /* std::max(ev, 0.0) keeps ev when !(ev < 0): NaN stays NaN. */
    for (int i = 0; i < 3; ++i) {
        double e = ev[2 - i];
        lam[i] = (e < 0.0) ? 0.0 : e;
    }
    double sum = 0.0;
    sum = sum + lam[0];
    sum = sum + lam[1];
    sum = sum + lam[2];
    if (lam[0] == 0.0) return ORC_ST_FEATURE_INVALID;
    if (st->mode == ORC_MODE_SQRT)
        for (int i = 0; i < 3; ++i) lam[i] = sqrt(lam[i]);
    else if (st->mode == ORC_MODE_NORMALIZED)
        for (int i = 0; i < 3; ++i) lam[i] = lam[i] / sum;
    double v1[3], v2[3], v3[3];
    for (int i = 0; i < 3; ++i) {
        v1[i] = vec[3 * 2 + i];
        v2[i] = vec[3 * 1 + i];
        v3[i] = vec[3 * 0 + i];
    }
    int c = 0;
    for (int b = 0; b < 10; ++b) {
        if (!(st->features & (1u << b))) continue;
        double v = 0.0;
        switch (1u << b) {
        case ORC_F_LINEARITY: v = (lam[0] - lam[1]) / lam[0]; break;
        case ORC_F_PLANARITY: v = (lam[1] - lam[2]) / lam[0]; break;
        case ORC_F_SCATTERING: v = lam[2] / lam[0]; break;
        case ORC_F_VERTICALITY: {
            double u[3], norm = 0.0;
            for (int i = 0; i < 3; ++i) {
                u[i] = lam[0] * fabs(v1[i]) + lam[1] * fabs(v2[i]) +
                       lam[2] * fabs(v3[i]);
                norm += u[i] * u[i];
            }
            norm = sqrt(norm);
            v = u[2] / norm;
            break;
        }
        case ORC_F_OMNIVARIANCE: v = cbrt(lam[2] * lam[1] * lam[0]); break;
        case ORC_F_ANISOTROPY: v = (lam[0] - lam[2]) / lam[0]; break;
        case ORC_F_EIGENENTROPY:
            v = -(lam[2] * log(lam[2]) + lam[1] * log(lam[1]) +
                  lam[0] * log(lam[0]));
            break;
        case ORC_F_EIGENVALUESUM: v = sum; break;
        case ORC_F_SURFACEVARIATION: v = lam[2] / sum; break;
        case ORC_F_DEMANTKEVERTICALITY: v = 1 - fabs(v3[2]); break;
        }
        vals[c++] = v;
    }
    return 0;
}

/* One row of a stage given its neighbor list (already in KD3 order). */
typedef void (*cov_fn)(const double *, const double *, const double *,
                       const void *, uint64_t, double *);

static void cov_u64(const double *x, const double *y, const double *z,
                    const void *ids, uint64_t n, double *cov)
{
    orc_pdal_covariance(x, y, z, (const uint64_t *)ids, n, cov);
}
static void cov_u32(const double *x, const double *y, const double *z,
                    const void *ids, uint64_t n, double *cov)
{
    orc_pdal_covariance_u32(x, y, z, (const uint32_t *)ids, n, cov);
}

int orc_rank_from_eigenvalues(const double ev[3], double thresh)
{
    /* SVDBase::rank (Eigen/src/SVD/SVDBase.h:148-157): singular values in
     * decreasing order; the leading ones at or above max(thresh * largest,
     * DBL_MIN), among the nonzero ones. */
    double s[3] = {fabs(ev[0]), fabs(ev[1]), fabs(ev[2])};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (s[j] > s[i]) {
                double t = s[i];
                s[i] = s[j];
                s[j] = t;
            }
    const double thr = (double)(float)thresh;
    const double pre = fmax(s[0] * thr, DBL_MIN);
    int r = 0;
    while (r < 3 && s[r] != 0.0 && s[r] >= pre) ++r;
    return r;
}

static uint8_t stage_row(const orc_stage *st, const double *x, const double *y,
                         const double *z, const void *ids, uint64_t cnt,
                         cov_fn covf, double *vals)
{
    double cov[9], ev[3], vec[9];
    if (st->stage == ORC_STAGE_RANK) {
        /* EstimateRankFilter.cpp:77-83: computeRank, no isZero skip. */
        covf(x, y, z, ids, cnt, cov);
        if (orc_eigen_sym3(cov, ev, vec) != 0) return ORC_ST_EIGEN_FAIL;
        vals[0] = (double)orc_rank_from_eigenvalues(ev, st->p1);
        return 0;
    }
    if (st->stage == ORC_STAGE_COPLANAR) {
        /* ApproximateCoplanarFilter.cpp:90-121. */
        covf(x, y, z, ids, cnt, cov);
        if (orc_is_zero(cov)) return ORC_ST_ZERO_COV;
        if (orc_eigen_sym3(cov, ev, vec) != 0) return ORC_ST_EIGEN_FAIL;
        vals[0] = (ev[1] > st->p1 * ev[0] && st->p2 * ev[1] > ev[2]) ? 1.0 : 0.0;
        return 0;
    }
    if (st->stage == ORC_STAGE_NORMAL) {
        /* math::findNormal (MathUtils.cpp:572-611). */
        if (cnt < 3) return ORC_ST_FEW_NEIGHBORS;
        covf(x, y, z, ids, cnt, cov);
        if (orc_is_zero(cov)) return ORC_ST_ZERO_COV;
        if (orc_eigen_sym3(cov, ev, vec) != 0) return ORC_ST_EIGEN_FAIL;
        double sum = ev[0] + ev[1] + ev[2];
        double curv = sum != 0.0 ? fabs(ev[0] / sum) : 0.0;
        double nx = vec[0], ny = vec[1], nz = vec[2];
        if (st->always_up && nz < 0.0) {
            nx = -nx;
            ny = -ny;
            nz = -nz;
        }
        vals[0] = nx;
        vals[1] = ny;
        vals[2] = nz;
        vals[3] = curv;
        return 0;
    }
    covf(x, y, z, ids, cnt, cov);
    if (orc_is_zero(cov)) return ORC_ST_ZERO_COV;
    if (orc_eigen_sym3(cov, ev, vec) != 0) return ORC_ST_EIGEN_FAIL;
    if (st->stage == ORC_STAGE_EIGENVALUES) {
        if (st->normalize) {
            double sum = ev[0] + ev[1] + ev[2];
            for (int i = 0; i < 3; ++i) ev[i] = ev[i] / sum;
        }
        vals[0] = ev[0];
        vals[1] = ev[1];
        vals[2] = ev[2];
        return 0;
    }
    return covfeat_row(st, ev, vec, vals);
}

typedef struct {
    const orc_kd3 *t;
    const double *x, *y, *z;
    uint64_t n;
    const orc_stage *st;
    const uint32_t *lists;
    const uint64_t *offsets; /* with lists: CSR rows of any length */
    uint64_t kk;
    int truth;
    uint64_t begin, end;
    double *const *cols;
    uint8_t *status;
    double *gap;
    double *ratio;
    const uint32_t *row_k; /* per-row neighbor counts (optimized=true) */
} stage_job;

static uint8_t truth_row(const orc_stage *st, const double *x, const double *y,
                         const double *z, const uint32_t *ids, uint64_t cnt,
                         double *vals, double *gap, double *ratio);

static void *stage_worker(void *arg)
{
    stage_job *j = arg;
    const orc_stage *st = j->st;
    /* approximatecoplanar and estimaterank ask neighbors(p, knn); the other
     * stages knn + 1. */
    uint64_t K = (uint64_t)st->knn +
                 ((st->stage == ORC_STAGE_COPLANAR || st->stage == ORC_STAGE_RANK) ? 0 : 1);
    uint64_t stride = st->stage == ORC_STAGE_NORMAL ? 1 : (uint64_t)st->stride;
    if (stride < 1) stride = 1;
    uint64_t kc = K < j->n ? K : j->n;
    uint64_t cap = stride * kc;
    uint64_t *ids = malloc((cap ? cap : 1) * sizeof(uint64_t));
    double *d2 = malloc((cap ? cap : 1) * sizeof(double));
    int ncol = orc_stage_columns(st);
    double vals[16];
    /* Radius stages: growable id/distance buffers. */
    uint64_t rcap = 0;
    uint64_t *rids = NULL;
    double *rd2 = NULL;
    for (uint64_t i = j->begin; i < j->end; ++i) {
        uint8_t s;
        if (j->lists && j->offsets) {
            const uint32_t *row = j->lists + j->offsets[i];
            uint64_t cnt = j->offsets[i + 1] - j->offsets[i];
            if (st->radius > 0 && st->stage != ORC_STAGE_NORMAL && cnt < (uint64_t)st->min_k)
                s = ORC_ST_FEW_NEIGHBORS;
            else if (j->truth)
                s = truth_row(st, j->x, j->y, j->z, row, cnt, vals,
                              j->gap ? j->gap + i : NULL, j->ratio ? j->ratio + i : NULL);
            else
                s = stage_row(st, j->x, j->y, j->z, row, cnt, cov_u32, vals);
        } else if (!j->lists && st->radius > 0) {
            /* KD3Index::radius, then the stage (CovarianceFeaturesFilter.cpp:
             * 216-223, EigenvaluesFilter.cpp:123-136: fewer than min_k
             * neighbors skips the row; NormalFilter.cpp:149-150: findNormal
             * skips fewer than three). */
            uint64_t cnt = orc_kd3_radius(j->t, j->x[i], j->y[i], j->z[i], st->radius, rcap,
                                          rids, rd2);
            if (cnt > rcap) {
                rcap = cnt + cnt / 2 + 16;
                free(rids);
                free(rd2);
                rids = malloc(rcap * sizeof(uint64_t));
                rd2 = malloc(rcap * sizeof(double));
                cnt = orc_kd3_radius(j->t, j->x[i], j->y[i], j->z[i], st->radius, rcap, rids,
                                     rd2);
            }
            if (st->stage != ORC_STAGE_NORMAL && cnt < (uint64_t)st->min_k)
                s = ORC_ST_FEW_NEIGHBORS;
            else
                s = stage_row(st, j->x, j->y, j->z, rids, cnt, cov_u64, vals);
        } else if (j->row_k) {
            /* CovarianceFeaturesFilter.cpp:204-210 (optimized=true):
             * KD3Index::neighbors(p, OptimalKNN, 1). */
            uint64_t kr = j->row_k[i] < j->n ? j->row_k[i] : j->n;
            if (kr > rcap) {
                rcap = kr;
                free(rids);
                free(rd2);
                rids = malloc(rcap * sizeof(uint64_t));
                rd2 = malloc(rcap * sizeof(double));
            }
            uint64_t cnt = kr ? orc_kd3_neighbors(j->t, j->x[i], j->y[i], j->z[i], kr, 1, rids,
                                                  rd2)
                              : 0;
            s = stage_row(st, j->x, j->y, j->z, rids, cnt, cov_u64, vals);
        } else if (j->lists) {
            const uint32_t *row = j->lists + i * j->kk;
            if (j->truth)
                s = truth_row(st, j->x, j->y, j->z, row, j->kk, vals,
                              j->gap ? j->gap + i : NULL, j->ratio ? j->ratio + i : NULL);
            else
                s = stage_row(st, j->x, j->y, j->z, row, j->kk, cov_u32, vals);
        } else {
            uint64_t cnt = orc_kd3_neighbors(j->t, j->x[i], j->y[i], j->z[i],
                                             K, stride, ids, d2);
            s = stage_row(st, j->x, j->y, j->z, ids, cnt, cov_u64, vals);
        }
        j->status[i] = s;
        if (s == 0)
            for (int c = 0; c < ncol; ++c) j->cols[c][i] = vals[c];
    }
    free(ids);
    free(d2);
    free(rids);
    free(rd2);
    return NULL;
}

static void run_jobs_range(stage_job proto, uint64_t b0, uint64_t e0, int threads)
{
    if (threads < 1) threads = 1;
    uint64_t n = e0 - b0;
    uint64_t chunk = n / (uint64_t)threads + ((n % (uint64_t)threads) ? 1 : 0);
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    stage_job *jobs = malloc(sizeof(stage_job) * (size_t)threads);
    for (int i = 0; i < threads; ++i) {
        jobs[i] = proto;
        uint64_t b = (uint64_t)i * chunk;
        uint64_t e = (i + 1 == threads) ? n : (uint64_t)(i + 1) * chunk;
        jobs[i].begin = b0 + (b > n ? n : b);
        jobs[i].end = b0 + (e > n ? n : e);
        if (threads == 1)
            stage_worker(&jobs[i]);
        else
            pthread_create(&th[i], NULL, stage_worker, &jobs[i]);
    }
    if (threads > 1)
        for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}

static void run_jobs(stage_job proto, uint64_t n, int threads)
{
    if (threads < 1) threads = 1;
    uint64_t chunk = n / (uint64_t)threads + ((n % (uint64_t)threads) ? 1 : 0);
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    stage_job *jobs = malloc(sizeof(stage_job) * (size_t)threads);
    for (int i = 0; i < threads; ++i) {
        jobs[i] = proto;
        uint64_t b = (uint64_t)i * chunk;
        uint64_t e = (i + 1 == threads) ? n : (uint64_t)(i + 1) * chunk;
        jobs[i].begin = b > n ? n : b;
        jobs[i].end = e > n ? n : e;
        if (threads == 1)
            stage_worker(&jobs[i]);
        else
            pthread_create(&th[i], NULL, stage_worker, &jobs[i]);
    }
    if (threads > 1)
        for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}

void orc_stage_run_rowk(const orc_kd3 *t, const double *x, const double *y, const double *z,
                        uint64_t n, const orc_stage *st, const uint32_t *row_k, int threads,
                        double *const *cols, uint8_t *status)
{
    stage_job p;
    memset(&p, 0, sizeof(p));
    p.t = t;
    p.x = x;
    p.y = y;
    p.z = z;
    p.n = n;
    p.st = st;
    p.cols = cols;
    p.status = status;
    p.row_k = row_k;
    run_jobs(p, n, threads);
}

void orc_stage_run(const orc_kd3 *t, const double *x, const double *y,
                   const double *z, uint64_t n, const orc_stage *st,
                   int threads, double *const *cols, uint8_t *status)
{
    stage_job p;
    memset(&p, 0, sizeof(p));
    p.t = t;
    p.x = x;
    p.y = y;
    p.z = z;
    p.n = n;
    p.st = st;
    p.cols = cols;
    p.status = status;
    run_jobs(p, n, threads);
}

void orc_stage_run_range(const orc_kd3 *t, const double *x, const double *y,
                         const double *z, uint64_t n, const orc_stage *st,
                         int threads, uint64_t row_begin, uint64_t row_end,
                         double *const *cols, uint8_t *status)
{
    stage_job p;
    memset(&p, 0, sizeof(p));
    p.t = t;
    p.x = x;
    p.y = y;
    p.z = z;
    p.n = n;
    p.st = st;
    p.cols = cols;
    p.status = status;
    if (row_end > n) row_end = n;
    if (row_begin > row_end) row_begin = row_end;
    run_jobs_range(p, row_begin, row_end, threads);
}

void orc_stage_from_lists(const double *x, const double *y, const double *z,
                          uint64_t n, const uint32_t *nbr, uint64_t kk,
                          const orc_stage *st, int threads,
                          double *const *cols, uint8_t *status)
{
    stage_job p;
    memset(&p, 0, sizeof(p));
    p.x = x;
    p.y = y;
    p.z = z;
    p.n = n;
    p.st = st;
    p.lists = nbr;
    p.kk = kk;
    p.cols = cols;
    p.status = status;
    run_jobs(p, n, threads);
}

/* ------------------------------------------------------------- truth --- */
typedef long double ld;

/* Cyclic Jacobi on a symmetric 3x3 in long double, to convergence. */
static int jacobi3(ld a[3][3], ld v[3][3])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v[i][j] = (i == j) ? 1.0L : 0.0L;
    int sweep;
    for (sweep = 0; sweep < 64; ++sweep) {
        int rotated = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                ld apq = a[p][q];
                if (apq == 0.0L) continue;
                /* Negligible against the geometric mean of the pivots: the
                 * relative-accuracy stopping test for definite matrices. */
                if (fabsl(apq) <= 1e-22L * sqrtl(fabsl(a[p][p] * a[q][q])) ||
                    fabsl(apq) < 1e-300L) {
                    a[p][q] = a[q][p] = 0.0L;
                    continue;
                }
                rotated = 1;
                ld theta = (a[q][q] - a[p][p]) / (2.0L * apq);
                ld t = (theta >= 0 ? 1.0L : -1.0L) /
                       (fabsl(theta) + sqrtl(theta * theta + 1.0L));
                ld c = 1.0L / sqrtl(t * t + 1.0L), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    ld akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    ld apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                a[p][q] = a[q][p] = 0.0L;
                for (int k = 0; k < 3; ++k) {
                    ld vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
        if (!rotated) break;
    }
    return sweep;
}

static int jacobi_sorted(const ld m[3][3], ld ev[3], ld vec[3][3])
{
    ld a[3][3], v[3][3];
    memcpy(a, m, sizeof(a));
    int sweeps = jacobi3(a, v);
    int ord[3] = {0, 1, 2};
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (a[ord[j]][ord[j]] < a[ord[i]][ord[i]]) {
                int t = ord[i];
                ord[i] = ord[j];
                ord[j] = t;
            }
    for (int i = 0; i < 3; ++i) {
        ev[i] = a[ord[i]][ord[i]];
        for (int r = 0; r < 3; ++r) vec[i][r] = v[r][ord[i]];
    }
    return sweeps;
}

int orc_jacobi_sym3(const double cov[9], double evals[3], double evecs[9])
{
    ld m[3][3], ev[3], vec[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = cov[3 * i + j];
    int s = jacobi_sorted(m, ev, vec);
    for (int i = 0; i < 3; ++i) {
        evals[i] = (double)ev[i];
        for (int r = 0; r < 3; ++r) evecs[3 * i + r] = (double)vec[i][r];
    }
    return s;
}

static uint8_t truth_row(const orc_stage *st, const double *x, const double *y,
                         const double *z, const uint32_t *ids, uint64_t cnt,
                         double *vals, double *gap, double *ratio)
{
    if (st->stage == ORC_STAGE_NORMAL && cnt < 3) return ORC_ST_FEW_NEIGHBORS;
    ld m[3] = {0, 0, 0};
    for (uint64_t j = 0; j < cnt; ++j) {
        m[0] += (ld)x[ids[j]];
        m[1] += (ld)y[ids[j]];
        m[2] += (ld)z[ids[j]];
    }
    for (int i = 0; i < 3; ++i) m[i] /= (ld)cnt;
    ld c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (uint64_t j = 0; j < cnt; ++j) {
        ld d[3] = {(ld)x[ids[j]] - m[0], (ld)y[ids[j]] - m[1],
                   (ld)z[ids[j]] - m[2]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) c[a][b] += d[a] * d[b];
    }
    int allzero = 1;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            c[a][b] /= (ld)(cnt - 1);
            if (fabsl(c[a][b]) > 1e-12L) allzero = 0;
        }
    if (allzero && st->stage != ORC_STAGE_RANK) return ORC_ST_ZERO_COV;
    ld ev[3], vec[3][3];
    jacobi_sorted(c, ev, vec);
    if (st->stage == ORC_STAGE_RANK) {
        double evd[3] = {(double)ev[0], (double)ev[1], (double)ev[2]};
        vals[0] = (double)orc_rank_from_eigenvalues(evd, st->p1);
        return 0;
    }
    if (st->stage == ORC_STAGE_COPLANAR) {
        vals[0] = (ev[1] > (ld)st->p1 * ev[0] && (ld)st->p2 * ev[1] > ev[2]) ? 1.0 : 0.0;
        return 0;
    }
    if (gap) {
        ld big = fabsl(ev[2]) > 0 ? fabsl(ev[2]) : 1.0L;
        ld g1 = ev[1] - ev[0], g2 = ev[2] - ev[1];
        *gap = (double)((g1 < g2 ? g1 : g2) / big);
    }
    if (ratio) {
        ld big = fabsl(ev[2]) > 0 ? fabsl(ev[2]) : 1.0L;
        *ratio = (double)((ev[0] > 0 ? ev[0] : 0) / big);
    }
    if (st->stage == ORC_STAGE_NORMAL) {
        ld sum = ev[0] + ev[1] + ev[2];
        ld curv = sum != 0 ? fabsl(ev[0] / sum) : 0;
        ld n[3] = {vec[0][0], vec[0][1], vec[0][2]};
        if (st->always_up && n[2] < 0)
            for (int i = 0; i < 3; ++i) n[i] = -n[i];
        for (int i = 0; i < 3; ++i) vals[i] = (double)n[i];
        vals[3] = (double)curv;
        return 0;
    }
    if (st->stage == ORC_STAGE_EIGENVALUES) {
        ld sum = ev[0] + ev[1] + ev[2];
        for (int i = 0; i < 3; ++i)
            vals[i] = (double)(st->normalize ? ev[i] / sum : ev[i]);
        return 0;
    }
    ld lam[3];
    for (int i = 0; i < 3; ++i) lam[i] = ev[2 - i] < 0 ? 0 : ev[2 - i];
    ld sum = lam[0] + lam[1] + lam[2];
    if (lam[0] == 0) return ORC_ST_FEATURE_INVALID;
    if (st->mode == ORC_MODE_SQRT)
        for (int i = 0; i < 3; ++i) lam[i] = sqrtl(lam[i]);
    else if (st->mode == ORC_MODE_NORMALIZED)
        for (int i = 0; i < 3; ++i) lam[i] /= sum;
    int k = 0;
    for (int b = 0; b < 10; ++b) {
        if (!(st->features & (1u << b))) continue;
        ld v = 0;
        switch (1u << b) {
        case ORC_F_LINEARITY: v = (lam[0] - lam[1]) / lam[0]; break;
        case ORC_F_PLANARITY: v = (lam[1] - lam[2]) / lam[0]; break;
        case ORC_F_SCATTERING: v = lam[2] / lam[0]; break;
        case ORC_F_VERTICALITY: {
            ld u[3], nn = 0;
            for (int i = 0; i < 3; ++i) {
                u[i] = lam[0] * fabsl(vec[2][i]) + lam[1] * fabsl(vec[1][i]) +
                       lam[2] * fabsl(vec[0][i]);
                nn += u[i] * u[i];
            }
            v = u[2] / sqrtl(nn);
            break;
        }
        case ORC_F_OMNIVARIANCE: v = cbrtl(lam[2] * lam[1] * lam[0]); break;
        case ORC_F_ANISOTROPY: v = (lam[0] - lam[2]) / lam[0]; break;
        case ORC_F_EIGENENTROPY:
            v = -(lam[2] * logl(lam[2]) + lam[1] * logl(lam[1]) +
                  lam[0] * logl(lam[0]));
            break;
        case ORC_F_EIGENVALUESUM: v = sum; break;
        case ORC_F_SURFACEVARIATION: v = lam[2] / sum; break;
        case ORC_F_DEMANTKEVERTICALITY: v = 1 - fabsl(vec[0][2]); break;
        }
        vals[k++] = (double)v;
    }
    return 0;
}

void orc_truth_from_csr(const double *x, const double *y, const double *z, uint64_t n,
                        const uint64_t *offsets, const uint32_t *ids, const orc_stage *st,
                        int threads, double *const *cols, uint8_t *status, double *gap,
                        double *ratio)
{
    stage_job p;
    memset(&p, 0, sizeof(p));
    p.x = x;
    p.y = y;
    p.z = z;
    p.n = n;
    p.st = st;
    p.lists = ids;
    p.offsets = offsets;
    p.truth = 1;
    p.cols = cols;
    p.status = status;
    p.gap = gap;
    p.ratio = ratio;
    run_jobs(p, n, threads);
}

void orc_truth_from_lists(const double *x, const double *y, const double *z,
                          uint64_t n, const uint32_t *nbr, uint64_t kk,
                          const orc_stage *st, int threads,
                          double *const *cols, uint8_t *status, double *gap,
                          double *ratio)
{
    stage_job p;
    memset(&p, 0, sizeof(p));
    p.x = x;
    p.y = y;
    p.z = z;
    p.n = n;
    p.st = st;
    p.lists = nbr;
    p.kk = kk;
    p.truth = 1;
    p.cols = cols;
    p.status = status;
    p.gap = gap;
    p.ratio = ratio;
    run_jobs(p, n, threads);
}

/* ------------------------------------------ optimal neighborhood ------ */
/* filters/OptimalNeighborhoodFilter.cpp:72-163, one point at a time. */
typedef struct {
    const orc_kd3 *t;
    const double *x, *y, *z;
    uint64_t kmin, kmax, begin, end;
    uint64_t *kopt;
    double *ropt, *gap;
    uint64_t failed;
} optn_job;

static void *optn_worker(void *arg)
{
    optn_job *j = (optn_job *)arg;
    uint64_t ids[64];
    double d2[64];
    for (uint64_t i = j->begin; i < j->end; ++i) {
        orc_kd3_neighbors(j->t, j->x[i], j->y[i], j->z[i], j->kmax, 1, ids, d2);
        double minentropy = DBL_MAX;
        uint64_t kopt = 0;
        double ropt = 0.0;
        double mx = 0.0, my = 0.0, mz = 0.0;
        double B[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        double ent[64];
        int nent = 0;
        for (uint64_t k = 0; k < j->kmax; ++k) {
            const uint64_t q = ids[k];
            double dx = j->x[q] - mx;
            double dy = j->y[q] - my;
            double dz = j->z[q] - mz;
            double n = (double)(k + 1);
            mx += dx / n;
            my += dy / n;
            mz += dz / n;
            double s = (n - 1) / n;
            B[0][0] = B[0][0] + s * dx * dx;
            B[1][1] = B[1][1] + s * dy * dy;
            B[2][2] = B[2][2] + s * dz * dz;
            B[1][0] = B[0][1] = B[0][1] + s * dx * dy;
            B[2][0] = B[0][2] = B[0][2] + s * dx * dz;
            B[1][2] = B[2][1] = B[2][1] + s * dy * dz;
            if (k + 1 < j->kmin) continue;
            double cov[9], ev[3], vec[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) cov[3 * r + c] = B[r][c] / (n - 1);
            if (orc_eigen_sym3(cov, ev, vec) != 0) {
                j->failed++;
                break;
            }
            double lambda[3] = {ev[2] > 0.0 ? ev[2] : 0.0, ev[1] > 0.0 ? ev[1] : 0.0,
                                ev[0] > 0.0 ? ev[0] : 0.0};
            double sum = 0.0;
            for (int m = 0; m < 3; ++m) sum += lambda[m];
            for (int m = 0; m < 3; ++m) lambda[m] = lambda[m] / sum;
            double entropy = -(lambda[2] * log(lambda[2]) + lambda[1] * log(lambda[1]) +
                               lambda[0] * log(lambda[0]));
            if (entropy == entropy) ent[nent++] = entropy;
            if (entropy < minentropy) {
                minentropy = entropy;
                kopt = k + 1;
                ropt = d2[k];
            }
        }
        j->kopt[i] = kopt;
        j->ropt[i] = sqrt(ropt);
        if (j->gap) {
            double g = INFINITY;
            int seen_best = 0;
            for (int m = 0; m < nent; ++m) {
                if (ent[m] == minentropy && !seen_best) {
                    seen_best = 1;
                    continue;
                }
                double dd = fabs(ent[m] - minentropy) / (fabs(minentropy) > 0 ? fabs(minentropy) : 1.0);
                if (dd < g) g = dd;
            }
            j->gap[i] = g;
        }
    }
    return NULL;
}

uint64_t orc_optimal_neighborhood(const orc_kd3 *t, const double *x, const double *y,
                                  const double *z, uint64_t n, uint64_t kmin, uint64_t kmax,
                                  int threads, uint64_t *kopt, double *ropt, double *gap)
{
    if (threads < 1) threads = 1;
    uint64_t chunk = n / (uint64_t)threads + ((n % (uint64_t)threads) ? 1 : 0);
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    optn_job *jobs = malloc(sizeof(optn_job) * (size_t)threads);
    for (int i = 0; i < threads; ++i) {
        uint64_t b = (uint64_t)i * chunk, e = (i + 1 == threads) ? n : (uint64_t)(i + 1) * chunk;
        if (b > n) b = n;
        if (e > n) e = n;
        jobs[i] = (optn_job){t, x, y, z, kmin, kmax, b, e, kopt, ropt, gap, 0};
        if (threads == 1)
            optn_worker(&jobs[i]);
        else
            pthread_create(&th[i], NULL, optn_worker, &jobs[i]);
    }
    uint64_t failed = 0;
    for (int i = 0; i < threads; ++i) {
        if (threads > 1) pthread_join(th[i], NULL);
        failed += jobs[i].failed;
    }
    free(th);
    free(jobs);
    return failed;
}

/* ------------------------------------------- kNN distance consumers --- */
/* filters.nndistance (NNDistanceFilter.cpp:113-132): K = k + 1 squared
 * distances from KD3Index::knnSearch, zero-initialised so that entries past
 * the cloud's size stay 0 (KDImpl.hpp:383-396 clamps k to the point count);
 * kth = sqrt(d2[K-1]), avg = sum_{i=1}^{K-1} sqrt(d2[i]) / (K - 1).
 * mode 2 is OutlierFilter::processStatistical's per-point running mean
 * (OutlierFilter.cpp:122-136): dist += (sqrt(d2[j]) - dist) / j. */
typedef struct {
    const orc_kd3 *t;
    const double *x, *y, *z;
    uint64_t k;
    int mode;
    uint64_t b, e;
    double *out;
} nnd_job;

static void *nnd_worker(void *arg)
{
    nnd_job *jb = (nnd_job *)arg;
    const uint64_t K = jb->k + 1;
    uint64_t *ids = malloc(sizeof(uint64_t) * (size_t)K);
    double *d2 = malloc(sizeof(double) * (size_t)K);
    for (uint64_t i = jb->b; i < jb->e; ++i) {
        memset(d2, 0, sizeof(double) * (size_t)K);
        orc_kd3_neighbors(jb->t, jb->x[i], jb->y[i], jb->z[i], K, 1, ids, d2);
        double val = 0.0;
        if (jb->mode == 0) {
            val = sqrt(d2[K - 1]);
        } else if (jb->mode == 1) {
            for (uint64_t q = 1; q < K; ++q) val += sqrt(d2[q]);
            val /= (double)(K - 1);
        } else {
            for (uint64_t q = 1; q < K; ++q) {
                double delta = sqrt(d2[q]) - val;
                val += delta / (double)q;
            }
        }
        jb->out[i] = val;
    }
    free(ids);
    free(d2);
    return NULL;
}

void orc_nn_distance(const orc_kd3 *t, const double *x, const double *y, const double *z,
                     uint64_t n, uint64_t k, int mode, int threads, double *out)
{
    if (threads < 1) threads = 1;
    uint64_t chunk = n / (uint64_t)threads + ((n % (uint64_t)threads) ? 1 : 0);
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    nnd_job *jobs = malloc(sizeof(nnd_job) * (size_t)threads);
    for (int i = 0; i < threads; ++i) {
        uint64_t b = (uint64_t)i * chunk, e = (i + 1 == threads) ? n : (uint64_t)(i + 1) * chunk;
        if (b > n) b = n;
        if (e > n) e = n;
        jobs[i] = (nnd_job){t, x, y, z, k, mode, b, e, out};
        if (threads == 1)
            nnd_worker(&jobs[i]);
        else
            pthread_create(&th[i], NULL, nnd_worker, &jobs[i]);
    }
    if (threads > 1)
        for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}

/* OutlierFilter::processStatistical (OutlierFilter.cpp:98-170): the online
 * moments over the per-point distances in point order, variance M2 / (n - 1),
 * threshold = mean + multiplier * stdev; outlier iff distance >= threshold
 * (the inlier test is distance < threshold). */
uint64_t orc_outlier_statistical(const orc_kd3 *t, const double *x, const double *y,
                                 const double *z, uint64_t n, uint64_t mean_k,
                                 double multiplier, int threads, uint8_t *outlier,
                                 double *threshold, double *dist)
{
    double *d = dist ? dist : malloc(sizeof(double) * (size_t)(n ? n : 1));
    orc_nn_distance(t, x, y, z, n, mean_k, 2, threads, d);
    uint64_t cnt = 0;
    double M1 = 0.0, M2 = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t n1 = cnt;
        cnt++;
        double delta = d[i] - M1;
        double delta_n = delta / (double)cnt;
        M1 += delta_n;
        M2 += delta * delta_n * (double)n1;
    }
    double variance = M2 / ((double)cnt - 1.0);
    double thr = M1 + multiplier * sqrt(variance);
    uint64_t nout = 0;
    for (uint64_t i = 0; i < n; ++i) {
        outlier[i] = d[i] < thr ? 0 : 1;
        nout += outlier[i];
    }
    if (threshold) *threshold = thr;
    if (!dist) free(d);
    return nout;
}

/* OutlierFilter::processRadius (OutlierFilter.cpp:78-96): inlier iff
 * |KD3Index::radius(i, r)| > size_t(min_k). */
uint64_t orc_outlier_radius(const orc_kd3 *t, uint64_t n, double r, int64_t min_k,
                            int threads, uint8_t *outlier)
{
    uint32_t *counts = malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    orc_kd3_radius_counts(t, r, threads, counts);
    uint64_t nout = 0;
    for (uint64_t i = 0; i < n; ++i) {
        outlier[i] = (uint64_t)counts[i] > (uint64_t)min_k ? 0 : 1;
        nout += outlier[i];
    }
    free(counts);
    return nout;
}
