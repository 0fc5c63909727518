/*
 * kd3.c -- restatement of PDAL's KD3Index (nanoflann 1.3.1 as vendored by
 * the reference) for the oracle.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Semantics followed, with reference citations:
 *   - leaf size 100 (pdal/private/KDImpl.hpp:216); the root box is the
 *     PointView bounds (KDImpl.hpp:263-276), i.e. the data min/max;
 *   - divideTree: leaves compute their box from their points, internal nodes
 *     store divlow = max of child1's points and divhigh = min of child2's
 *     points along the cut axis (nanoflann.hpp:640-690);
 *   - middleSplit_: cut axis = widest actual spread among the axes whose box
 *     span exceeds (1-1e-5)*max span; cut value = box middle clamped to the
 *     actual [min,max]; split index from lim1/lim2/count/2
 *     (nanoflann.hpp:779-827);
 *   - planeSplit: two Hoare passes (< cutval, then <= cutval) with unsigned
 *     indices (nanoflann.hpp:837-875);
 *   - findNeighbors/searchLevel: incremental box distance, closer child first,
 *     other child visited when mindist <= worst; leaves compare against the
 *     worst distance read once per leaf, strict < (nanoflann.hpp:1059-1080,
 *     1184-1248);
 *   - KNNResultSet::addPoint without NANOFLANN_FIRST_MATCH: equal distances
 *     are inserted after existing ones (nanoflann.hpp:145-208);
 *   - KD3ImplT::neighbors: k = min(n, k), capacity stride*k, then every
 *     stride-th id (KDImpl.hpp:351-381);
 *   - distance d0*d0 + d1*d1 + d2*d2 accumulated left to right from 0
 *     (nanoflann.hpp:303-311 L2_Simple_Adaptor::evalMetric), no FMA.
 */
#include "oracle.h"

#include <float.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define KD3_LEAF 100u

struct orc_kd3 {
    const double *pt[3];
    uint64_t n;
    uint64_t *vind;
    /* node arrays, DFS preorder (allocation order of divideTree) */
    int32_t *c1, *c2, *feat;
    uint64_t *lo, *hi;
    double *divlow, *divhigh;
    uint64_t nodes, cap;
    double root_box[3][2];
};

static int32_t kd_alloc(struct orc_kd3 *t)
{
    if (t->nodes == t->cap) {
        uint64_t nc = t->cap ? 2 * t->cap : 1024;
        t->c1 = realloc(t->c1, nc * sizeof(int32_t));
        t->c2 = realloc(t->c2, nc * sizeof(int32_t));
        t->feat = realloc(t->feat, nc * sizeof(int32_t));
        t->lo = realloc(t->lo, nc * sizeof(uint64_t));
        t->hi = realloc(t->hi, nc * sizeof(uint64_t));
        t->divlow = realloc(t->divlow, nc * sizeof(double));
        t->divhigh = realloc(t->divhigh, nc * sizeof(double));
        t->cap = nc;
    }
    int32_t id = (int32_t)t->nodes++;
    t->c1[id] = t->c2[id] = -1;
    t->feat[id] = 0;
    t->lo[id] = t->hi[id] = 0;
    t->divlow[id] = t->divhigh[id] = 0.0;
    return id;
}

static inline double kd_get(const struct orc_kd3 *t, uint64_t idx, int dim)
{
    return t->pt[dim][idx];
}

/* computeMinMax over ind[0..count) along one axis. */
static void kd_minmax(const struct orc_kd3 *t, const uint64_t *ind,
                      uint64_t count, int dim, double *mn, double *mx)
{
    double a = kd_get(t, ind[0], dim), b = a;
    for (uint64_t i = 1; i < count; ++i) {
        double v = kd_get(t, ind[i], dim);
        if (v < a) a = v;
        if (v > b) b = v;
    }
    *mn = a;
    *mx = b;
}

/* The two unsigned Hoare passes of planeSplit. */
static void kd_plane_split(const struct orc_kd3 *t, uint64_t *ind,
                           uint64_t count, int dim, double cutval,
                           uint64_t *lim1, uint64_t *lim2)
{
    uint64_t l = 0, r = count - 1;
    for (;;) {
        while (l <= r && kd_get(t, ind[l], dim) < cutval) ++l;
        while (r && l <= r && kd_get(t, ind[r], dim) >= cutval) --r;
        if (l > r || !r) break;
        uint64_t tmp = ind[l];
        ind[l] = ind[r];
        ind[r] = tmp;
        ++l;
        --r;
    }
    *lim1 = l;
    r = count - 1;
    for (;;) {
        while (l <= r && kd_get(t, ind[l], dim) <= cutval) ++l;
        while (r && l <= r && kd_get(t, ind[r], dim) > cutval) --r;
        if (l > r || !r) break;
        uint64_t tmp = ind[l];
        ind[l] = ind[r];
        ind[r] = tmp;
        ++l;
        --r;
    }
    *lim2 = l;
}

static void kd_middle_split(const struct orc_kd3 *t, uint64_t *ind,
                            uint64_t count, double box[3][2], uint64_t *index,
                            int *cutfeat, double *cutval)
{
    const double eps = 0.00001;
    double max_span = box[0][1] - box[0][0];
    for (int i = 1; i < 3; ++i) {
        double span = box[i][1] - box[i][0];
        if (span > max_span) max_span = span;
    }
    double max_spread = -1;
    int feat = 0;
    for (int i = 0; i < 3; ++i) {
        double span = box[i][1] - box[i][0];
        if (span > (1 - eps) * max_span) {
            double mn, mx;
            kd_minmax(t, ind, count, i, &mn, &mx);
            double spread = mx - mn;
            if (spread > max_spread) {
                feat = i;
                max_spread = spread;
            }
        }
    }
    double split = (box[feat][0] + box[feat][1]) / 2;
    double mn, mx;
    kd_minmax(t, ind, count, feat, &mn, &mx);
    double cv = split;
    if (split < mn)
        cv = mn;
    else if (split > mx)
        cv = mx;
    uint64_t lim1, lim2;
    kd_plane_split(t, ind, count, feat, cv, &lim1, &lim2);
    if (lim1 > count / 2)
        *index = lim1;
    else if (lim2 < count / 2)
        *index = lim2;
    else
        *index = count / 2;
    *cutfeat = feat;
    *cutval = cv;
}

static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }

static int32_t kd_divide(struct orc_kd3 *t, uint64_t left, uint64_t right,
                         double box[3][2])
{
    int32_t node = kd_alloc(t);
    if (right - left <= KD3_LEAF) {
        t->lo[node] = left;
        t->hi[node] = right;
        for (int i = 0; i < 3; ++i)
            box[i][0] = box[i][1] = kd_get(t, t->vind[left], i);
        for (uint64_t k = left + 1; k < right; ++k)
            for (int i = 0; i < 3; ++i) {
                double v = kd_get(t, t->vind[k], i);
                if (box[i][0] > v) box[i][0] = v;
                if (box[i][1] < v) box[i][1] = v;
            }
        return node;
    }
    uint64_t idx;
    int feat;
    double cv;
    kd_middle_split(t, t->vind + left, right - left, box, &idx, &feat, &cv);
    t->feat[node] = feat;
    t->lo[node] = left;
    t->hi[node] = right;

    double lbox[3][2], rbox[3][2];
    memcpy(lbox, box, sizeof(lbox));
    lbox[feat][1] = cv;
    int32_t a = kd_divide(t, left, left + idx, lbox);
    memcpy(rbox, box, sizeof(rbox));
    rbox[feat][0] = cv;
    int32_t b = kd_divide(t, left + idx, right, rbox);
    t->c1[node] = a;
    t->c2[node] = b;
    t->divlow[node] = lbox[feat][1];
    t->divhigh[node] = rbox[feat][0];
    for (int i = 0; i < 3; ++i) {
        box[i][0] = dmin(lbox[i][0], rbox[i][0]);
        box[i][1] = dmax(lbox[i][1], rbox[i][1]);
    }
    return node;
}

orc_kd3 *orc_kd3_build(const double *x, const double *y, const double *z,
                       uint64_t n)
{
    struct orc_kd3 *t = calloc(1, sizeof(*t));
    t->pt[0] = x;
    t->pt[1] = y;
    t->pt[2] = z;
    t->n = n;
    t->vind = malloc((n ? n : 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) t->vind[i] = i;
    if (n == 0) return t;
    /* PointView::calculateBounds: plain min/max over the points. */
    for (int d = 0; d < 3; ++d) {
        double a = t->pt[d][0], b = a;
        for (uint64_t i = 1; i < n; ++i) {
            double v = t->pt[d][i];
            a = dmin(a, v);
            b = dmax(b, v);
        }
        t->root_box[d][0] = a;
        t->root_box[d][1] = b;
    }
    kd_divide(t, 0, n, t->root_box);
    return t;
}

void orc_kd3_free(orc_kd3 *t)
{
    if (!t) return;
    free(t->vind);
    free(t->c1);
    free(t->c2);
    free(t->feat);
    free(t->lo);
    free(t->hi);
    free(t->divlow);
    free(t->divhigh);
    free(t);
}

uint64_t orc_kd3_node_count(const orc_kd3 *t) { return t->nodes; }

void orc_kd3_export(const orc_kd3 *t, uint64_t *vind, int32_t *child1,
                    int32_t *child2, uint64_t *left, uint64_t *right,
                    int32_t *divfeat, double *divlow, double *divhigh)
{
    if (vind) memcpy(vind, t->vind, t->n * sizeof(uint64_t));
    for (uint64_t i = 0; i < t->nodes; ++i) {
        if (child1) child1[i] = t->c1[i];
        if (child2) child2[i] = t->c2[i];
        if (left) left[i] = t->lo[i];
        if (right) right[i] = t->hi[i];
        if (divfeat) divfeat[i] = t->feat[i];
        if (divlow) divlow[i] = t->divlow[i];
        if (divhigh) divhigh[i] = t->divhigh[i];
    }
}

/* ------------------------------------------------------------ search --- */
typedef struct {
    uint64_t cap, count;
    uint64_t *ids;
    double *d2;
} knnset;

static void knn_add(knnset *s, double d, uint64_t id)
{
    uint64_t i;
    for (i = s->count; i > 0; --i) {
        if (s->d2[i - 1] > d) {
            if (i < s->cap) {
                s->d2[i] = s->d2[i - 1];
                s->ids[i] = s->ids[i - 1];
            }
        } else
            break;
    }
    if (i < s->cap) {
        s->d2[i] = d;
        s->ids[i] = id;
    }
    if (s->count < s->cap) s->count++;
}

static inline double kd_dist(const struct orc_kd3 *t, const double q[3],
                             uint64_t idx)
{
    double r = 0.0;
    for (int i = 0; i < 3; ++i) {
        double diff = q[i] - t->pt[i][idx];
        r += diff * diff;
    }
    return r;
}

static void kd_search(const struct orc_kd3 *t, knnset *s, const double q[3],
                      int32_t node, double mindistsq, double dists[3])
{
    if (t->c1[node] < 0 && t->c2[node] < 0) {
        double worst = s->d2[s->cap - 1];
        for (uint64_t i = t->lo[node]; i < t->hi[node]; ++i) {
            uint64_t idx = t->vind[i];
            double d = kd_dist(t, q, idx);
            if (d < worst) knn_add(s, d, idx);
        }
        return;
    }
    int f = t->feat[node];
    double val = q[f];
    double diff1 = val - t->divlow[node];
    double diff2 = val - t->divhigh[node];
    int32_t best, other;
    double cut;
    if ((diff1 + diff2) < 0) {
        best = t->c1[node];
        other = t->c2[node];
        cut = (val - t->divhigh[node]) * (val - t->divhigh[node]);
    } else {
        best = t->c2[node];
        other = t->c1[node];
        cut = (val - t->divlow[node]) * (val - t->divlow[node]);
    }
    kd_search(t, s, q, best, mindistsq, dists);
    double dst = dists[f];
    mindistsq = mindistsq + cut - dst;
    dists[f] = cut;
    /* epsError = 1 + eps with eps = 0 (SearchParams default). */
    if (mindistsq * (double)1.0f <= s->d2[s->cap - 1])
        kd_search(t, s, q, other, mindistsq, dists);
    dists[f] = dst;
}

/* KNNResultSet capacity cap, query q: fills ids/d2, returns the count. */
static uint64_t kd_find(const struct orc_kd3 *t, const double q[3],
                        uint64_t cap, uint64_t *ids, double *d2)
{
    knnset s;
    s.cap = cap;
    s.count = 0;
    s.ids = ids;
    s.d2 = d2;
    if (cap) d2[cap - 1] = DBL_MAX;
    if (t->n == 0) return 0;
    double dists[3] = {0.0, 0.0, 0.0};
    double distsq = 0.0;
    for (int i = 0; i < 3; ++i) {
        if (q[i] < t->root_box[i][0]) {
            dists[i] = (q[i] - t->root_box[i][0]) * (q[i] - t->root_box[i][0]);
            distsq += dists[i];
        }
        if (q[i] > t->root_box[i][1]) {
            dists[i] = (q[i] - t->root_box[i][1]) * (q[i] - t->root_box[i][1]);
            distsq += dists[i];
        }
    }
    kd_search(t, &s, q, 0, distsq, dists);
    return s.count;
}

uint64_t orc_kd3_neighbors(const orc_kd3 *t, double qx, double qy, double qz,
                           uint64_t k, uint64_t stride, uint64_t *ids,
                           double *d2)
{
    if (k > t->n) k = t->n;
    uint64_t k2 = stride * k;
    /* PointIdList output(k2) and the distance vector start zeroed. */
    memset(ids, 0, k2 * sizeof(uint64_t));
    memset(d2, 0, k2 * sizeof(double));
    double q[3] = {qx, qy, qz};
    kd_find(t, q, k2, ids, d2);
    if (stride > 1)
        for (uint64_t i = 1; i < k; ++i) ids[i] = ids[i * stride];
    return k;
}

/* KD3Index::radius (pdal/private/KDImpl.hpp:398-415): nanoflann's
 * radiusSearch (nanoflann.hpp:1116-1138) with RadiusResultSet (:229-258):
 * the same searchLevel walk with the worst distance fixed at r*r, a point
 * kept when its squared distance is strictly below it, then sorted by
 * distance (params.sorted).  The reference sorts with std::sort, whose order
 * among exactly equal distances is the library's; here equal distances keep
 * the walk's order (a stable sort). */
typedef struct {
    double r2;
    uint64_t count, cap;
    uint64_t *ids;
    double *d2;
} radset;

static void kd_radius_walk(const struct orc_kd3 *t, radset *s, const double q[3], int32_t node,
                           double mindistsq, double dists[3])
{
    if (t->c1[node] < 0 && t->c2[node] < 0) {
        for (uint64_t i = t->lo[node]; i < t->hi[node]; ++i) {
            uint64_t idx = t->vind[i];
            double d = kd_dist(t, q, idx);
            if (d < s->r2) {
                if (s->count < s->cap) {
                    s->ids[s->count] = idx;
                    s->d2[s->count] = d;
                }
                s->count++;
            }
        }
        return;
    }
    int f = t->feat[node];
    double val = q[f];
    double diff1 = val - t->divlow[node];
    double diff2 = val - t->divhigh[node];
    int32_t best, other;
    double cut;
    if ((diff1 + diff2) < 0) {
        best = t->c1[node];
        other = t->c2[node];
        cut = (val - t->divhigh[node]) * (val - t->divhigh[node]);
    } else {
        best = t->c2[node];
        other = t->c1[node];
        cut = (val - t->divlow[node]) * (val - t->divlow[node]);
    }
    kd_radius_walk(t, s, q, best, mindistsq, dists);
    double dst = dists[f];
    mindistsq = mindistsq + cut - dst;
    dists[f] = cut;
    if (mindistsq * (double)1.0f <= s->r2) kd_radius_walk(t, s, q, other, mindistsq, dists);
    dists[f] = dst;
}

uint64_t orc_kd3_radius(const orc_kd3 *t, double qx, double qy, double qz, double r,
                        uint64_t cap, uint64_t *ids, double *d2)
{
    radset s;
    s.r2 = r * r;
    s.count = 0;
    s.cap = cap;
    s.ids = ids;
    s.d2 = d2;
    if (t->n == 0) return 0;
    double q[3] = {qx, qy, qz};
    double dists[3] = {0.0, 0.0, 0.0};
    double distsq = 0.0;
    for (int i = 0; i < 3; ++i) {
        if (q[i] < t->root_box[i][0]) {
            dists[i] = (q[i] - t->root_box[i][0]) * (q[i] - t->root_box[i][0]);
            distsq += dists[i];
        }
        if (q[i] > t->root_box[i][1]) {
            dists[i] = (q[i] - t->root_box[i][1]) * (q[i] - t->root_box[i][1]);
            distsq += dists[i];
        }
    }
    kd_radius_walk(t, &s, q, 0, distsq, dists);
    /* Stable insertion sort by distance of the entries held. */
    uint64_t m = s.count < cap ? s.count : cap;
    for (uint64_t i = 1; i < m; ++i) {
        double d = d2[i];
        uint64_t id = ids[i];
        uint64_t j = i;
        while (j > 0 && d2[j - 1] > d) {
            d2[j] = d2[j - 1];
            ids[j] = ids[j - 1];
            --j;
        }
        d2[j] = d;
        ids[j] = id;
    }
    return s.count;
}

typedef struct {
    const orc_kd3 *t;
    double r;
    uint64_t begin, end;
    uint32_t *counts;
} rad_job;

static void *rad_worker(void *arg)
{
    rad_job *j = arg;
    for (uint64_t i = j->begin; i < j->end; ++i)
        j->counts[i] = (uint32_t)orc_kd3_radius(j->t, j->t->pt[0][i], j->t->pt[1][i],
                                                j->t->pt[2][i], j->r, 0, NULL, NULL);
    return NULL;
}

void orc_kd3_radius_counts(const orc_kd3 *t, double r, int threads, uint32_t *counts)
{
    if (threads < 1) threads = 1;
    uint64_t n = t->n, chunk = n / (uint64_t)threads + 1;
    pthread_t th[64];
    rad_job jobs[64];
    if (threads > 64) threads = 64;
    for (int i = 0; i < threads; ++i) {
        jobs[i].t = t;
        jobs[i].r = r;
        jobs[i].begin = (uint64_t)i * chunk < n ? (uint64_t)i * chunk : n;
        jobs[i].end = (uint64_t)(i + 1) * chunk < n ? (uint64_t)(i + 1) * chunk : n;
        jobs[i].counts = counts;
        pthread_create(&th[i], NULL, rad_worker, &jobs[i]);
    }
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
}

/* Brute-force radius counts (d2 < r*r, KD3Index's distance arithmetic). */
void orc_brute_radius_counts(const double *x, const double *y, const double *z, uint64_t n,
                             double r, const uint64_t *queries, uint64_t nq, uint32_t *counts)
{
    const double r2 = r * r;
    for (uint64_t a = 0; a < nq; ++a) {
        uint64_t q = queries ? queries[a] : a;
        uint32_t c = 0;
        for (uint64_t i = 0; i < n; ++i) {
            double d = 0.0, diff;
            diff = x[q] - x[i];
            d += diff * diff;
            diff = y[q] - y[i];
            d += diff * diff;
            diff = z[q] - z[i];
            d += diff * diff;
            if (d < r2) ++c;
        }
        counts[a] = c;
    }
}

typedef struct {
    const orc_kd3 *t;
    uint64_t k, stride, begin, end;
    uint32_t *out;
    const uint64_t *rows; /* NULL: row i is point i */
} kd_job;

static void *kd_worker(void *arg)
{
    kd_job *j = arg;
    uint64_t kk = j->k < j->t->n ? j->k : j->t->n;
    uint64_t k2 = j->stride * kk;
    uint64_t *ids = malloc((k2 ? k2 : 1) * sizeof(uint64_t));
    double *d2 = malloc((k2 ? k2 : 1) * sizeof(double));
    for (uint64_t i = j->begin; i < j->end; ++i) {
        uint64_t p = j->rows ? j->rows[i] : i;
        orc_kd3_neighbors(j->t, j->t->pt[0][p], j->t->pt[1][p], j->t->pt[2][p],
                          j->k, j->stride, ids, d2);
        for (uint64_t c = 0; c < kk; ++c)
            j->out[i * kk + c] = (uint32_t)ids[c];
    }
    free(ids);
    free(d2);
    return NULL;
}

static void kd_run_rows(const orc_kd3 *t, const uint64_t *rows, uint64_t n, uint64_t k,
                        uint64_t stride, int threads, uint32_t *out);

void orc_kd3_neighbors_all(const orc_kd3 *t, uint64_t k, uint64_t stride,
                           int threads, uint32_t *out)
{
    kd_run_rows(t, NULL, t->n, k, stride, threads, out);
}

void orc_kd3_neighbors_rows(const orc_kd3 *t, const uint64_t *rows, uint64_t nrows,
                            uint64_t k, uint64_t stride, int threads, uint32_t *out)
{
    kd_run_rows(t, rows, nrows, k, stride, threads, out);
}

static void kd_run_rows(const orc_kd3 *t, const uint64_t *rows, uint64_t n, uint64_t k,
                        uint64_t stride, int threads, uint32_t *out)
{
    if (threads < 1) threads = 1;
    uint64_t chunk = n / (uint64_t)threads + ((n % (uint64_t)threads) ? 1 : 0);
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    kd_job *jobs = malloc(sizeof(kd_job) * (size_t)threads);
    for (int i = 0; i < threads; ++i) {
        uint64_t b = (uint64_t)i * chunk;
        uint64_t e = (i + 1 == threads) ? n : (uint64_t)(i + 1) * chunk;
        if (b > n) b = n;
        if (e > n) e = n;
        jobs[i] = (kd_job){t, k, stride, b, e, out, rows};
        if (threads == 1)
            kd_worker(&jobs[i]);
        else
            pthread_create(&th[i], NULL, kd_worker, &jobs[i]);
    }
    if (threads > 1)
        for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}

/* ------------------------------------------------------- brute force --- */
typedef struct {
    const double *x, *y, *z;
    uint64_t n;
    const uint64_t *q;
    uint64_t begin, end, kk;
    uint32_t *ids;
    double *d2;
} bf_job;

static void *bf_worker(void *arg)
{
    bf_job *j = arg;
    for (uint64_t qi = j->begin; qi < j->end; ++qi) {
        uint64_t p = j->q[qi];
        double qx = j->x[p], qy = j->y[p], qz = j->z[p];
        uint32_t *oi = j->ids + qi * j->kk;
        double *od = j->d2 + qi * j->kk;
        uint64_t cnt = 0;
        for (uint64_t c = 0; c < j->n; ++c) {
            double dx = qx - j->x[c], dy = qy - j->y[c], dz = qz - j->z[c];
            double d = 0.0;
            d += dx * dx;
            d += dy * dy;
            d += dz * dz;
            if (cnt == j->kk && !(d < od[j->kk - 1])) continue;
            /* insert (d, c) ordered by (d, index); c increases so equal
             * distances go after existing entries. */
            uint64_t pos = cnt < j->kk ? cnt : j->kk - 1;
            while (pos > 0 && od[pos - 1] > d) {
                if (pos < j->kk) {
                    od[pos] = od[pos - 1];
                    oi[pos] = oi[pos - 1];
                }
                --pos;
            }
            od[pos] = d;
            oi[pos] = (uint32_t)c;
            if (cnt < j->kk) ++cnt;
        }
        for (; cnt < j->kk; ++cnt) {
            od[cnt] = DBL_MAX;
            oi[cnt] = 0xffffffffu;
        }
    }
    return NULL;
}

void orc_brute_knn(const double *x, const double *y, const double *z,
                   uint64_t n, const uint64_t *queries, uint64_t nq,
                   uint64_t kk, int threads, uint32_t *ids, double *d2)
{
    if (threads < 1) threads = 1;
    uint64_t chunk = nq / (uint64_t)threads + ((nq % (uint64_t)threads) ? 1 : 0);
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
    bf_job *jobs = malloc(sizeof(bf_job) * (size_t)threads);
    for (int i = 0; i < threads; ++i) {
        uint64_t b = (uint64_t)i * chunk, e = b + chunk;
        if (b > nq) b = nq;
        if (e > nq) e = nq;
        jobs[i] = (bf_job){x, y, z, n, queries, b, e, kk, ids, d2};
        pthread_create(&th[i], NULL, bf_worker, &jobs[i]);
    }
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(jobs);
}
