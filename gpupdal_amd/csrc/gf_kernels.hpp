// gf_kernels.hpp -- host-side launchers of the HIP kernels.
#pragma once

#include "gf_internal.hpp"

namespace gf {

constexpr uint32_t kBoundsMaxParts = 1024;

// grid.hip
void launch_bounds(const double* x, const double* y, const double* z,
                   uint64_t n, double* part, double* out, hipStream_t s);
void launch_keys(const double* x, const double* y, const double* z, uint64_t n,
                 const GridParams& g, uint32_t* keys, uint32_t* ids, hipStream_t s);
void launch_colstart(const uint32_t* keys, uint64_t n, const GridParams& g,
                     uint32_t* colstart, hipStream_t s);
uint64_t scan_scratch_words(uint64_t n);
void launch_excl_scan(const uint32_t* in, uint32_t* out, uint64_t n,
                      uint32_t* scratch, hipStream_t s);
uint64_t radix_hist_words(uint64_t n);
void launch_radix_sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                       uint64_t n, uint32_t bits, uint32_t* hist,
                       uint32_t* scan_scratch, uint32_t** kres, uint32_t** vres,
                       hipStream_t s);
void launch_gather_sorted(const uint32_t* perm, const double* x, const double* y,
                          const double* z, uint64_t n, double* sx, double* sy,
                          double* sz, hipStream_t s);
void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
void launch_copy_f64(const double* src, double* dst, uint64_t n, hipStream_t s);

// knn.hip
struct KnnOutputs {
    double* col[16];         // device output columns (original point order)
    uint8_t* status[kMaxStages];
    uint32_t* ids;           // neighbor-list mode: n x k (original order)
    uint32_t* flagged;       // sorted indices of tie/incomplete rows
    uint32_t* nflagged;      // [0] flagged rows, [1] tie rows, [2] incomplete rows
    double* cov;             // 6 x n FP64 covariance, sorted order (SoA)
    uint8_t* flags;          // n tie/incomplete flags, sorted order
    unsigned long long* counters; // optional: passes, candidates, merges, queries
};

struct KnnTuning {
    int32_t ring0;       // first ring (box half-width in columns)
    int32_t ring_max;    // budget before a row is marked incomplete
    int32_t small_col;   // columns up to this size are taken whole
    int32_t group_z;     // z cells beyond the first a query group may span
};

struct CloudView {
    const double* sx;
    const double* sy;
    const double* sz;
    const uint32_t* perm;     // sorted position -> original id
    const uint32_t* keys;     // sorted keys
    const uint32_t* colstart; // ncols + 1
    uint64_t n;
    GridParams g;
};

// mode 0: neighborhood covariance into out.cov / out.flags (sorted order);
// mode 1: neighbor lists into out.ids.
void launch_knn(const CloudView& cv, const FusedProgram& prog, int mode,
                const KnnOutputs& out, const KnnTuning& tune, hipStream_t s);
// Eigensystem + stage columns from out.cov for every row.
void launch_features(const CloudView& cv, const FusedProgram& prog,
                     const KnnOutputs& out, hipStream_t s);

// kdtree.hip -- the exact KD3Index tree (nanoflann 1.3.1 semantics).
struct KdView {
    uint32_t* vind;          // n: tree order -> original id
    double *px, *py, *pz;    // n: coordinates in tree order
    uint32_t* segid;         // n: deepest node holding the position (build)
    uint32_t *nl, *nr;       // node segment [nl, nr)
    uint32_t* split;         // split position of internal nodes
    int32_t *c1, *c2;        // children (-1: leaf)
    int32_t* feat;           // cut axis (-1: leaf)
    int32_t* parent;
    uint8_t* which;          // 0: child1 of its parent, 1: child2
    double *divlow, *divhigh, *cutval;
    double* box;             // 6 per node: the box passed down (lo xyz, hi xyz)
    unsigned long long *mn, *mx;  // 3 per node: order-encoded extents (build)
    uint32_t cap;            // node capacity
    uint32_t pad;
    double root[6];          // root box (data bounds)
};

struct KdScratch {
    uint32_t *flag, *s1, *s2, *tmp;  // n + 1 each
    uint32_t *is_split, *rank;       // capacity + 1 each
    uint32_t* scan_tmp;
    uint32_t* overflow;
};

uint32_t kd_node_capacity(uint64_t n);
// Sends every row to the exact search: used when the cloud holds fewer
// points than the result capacity stride*min(n, k).
void launch_flag_all(uint64_t n, const uint32_t* perm, int mode, const KnnOutputs& out,
                     hipStream_t s);
bool launch_kd_build(const double* x, const double* y, const double* z, uint64_t n,
                     const double* bounds, KdView t, KdScratch w, uint32_t* nodes,
                     hipStream_t s);
void launch_kd_resolve(const KdView& t, const CloudView& cv, const FusedProgram& prog, int mode,
                       const KnnOutputs& out, const double* x, const double* y, const double* z,
                       uint32_t nflagged, uint32_t* failed, hipStream_t s);

// Debug/parity hook: device eigen solver over m matrices (row-major 3x3).
void launch_eigen_batch(const double* cov, uint64_t m, double* evals,
                        double* evecs, int32_t* info, hipStream_t s);

} // namespace gf
