// gf_kernels.hpp -- host-side launchers of the HIP kernels.
#pragma once

#include "gf_internal.hpp"

#include <vector>
#include "gpupdal_amd.h"

namespace gf {

constexpr uint32_t kBoundsMaxParts = 1024;

// grid.hip
void launch_bounds(const double* x, const double* y, const double* z,
                   uint64_t n, double* part, double* out, hipStream_t s);
// Keys and ids, plus an interleaved copy of the points (xyz: 3 doubles per
// point) for the gather after the sort.
void launch_keys(const double* x, const double* y, const double* z, uint64_t n,
                 const GridParams& g, uint32_t* keys, uint32_t* ids, double* xyz,
                 hipStream_t s);
void launch_colstart(const uint32_t* keys, uint64_t n, const GridParams& g,
                     uint32_t* colstart, hipStream_t s);
uint64_t scan_scratch_words(uint64_t n);
// Exclusive scan of nb per-tile sums in place (one block).
void launch_scan_top(uint32_t* sums, uint32_t nb, hipStream_t s);
void launch_excl_scan(const uint32_t* in, uint32_t* out, uint64_t n,
                      uint32_t* scratch, hipStream_t s);
uint64_t radix_hist_words(uint64_t n);
void launch_radix_sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                       uint64_t n, uint32_t bits, uint32_t* hist,
                       uint32_t* scan_scratch, uint32_t** kres, uint32_t** vres,
                       hipStream_t s);
// Sorted coordinates from the interleaved copy (one cache line per point
// read at random instead of three), and the inverse permutation.
void launch_gather_sorted(const uint32_t* perm, const double* xyz, uint64_t n, double* sx,
                          double* sy, double* sz, uint32_t* inv, hipStream_t s);
void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);

// Tiled batches: segment blocks of <= 4096 points that never cross a tile,
// with each block's slots in the tile-major radix histogram.
constexpr uint32_t kSegBlockPoints = 4096;
struct SegBlock {
    uint32_t begin;  // first point
    uint32_t count;  // <= kSegBlockPoints
    uint32_t tile;
    uint32_t hbase;  // 256 * (first block of the tile)
    uint32_t nbt;    // blocks of the tile
    uint32_t lb;     // index of this block within the tile
};
// A tile's grid and its place in the batch (device copy for the index kernels).
struct TileGrid {
    GridParams g;
    uint64_t start, n;
    uint64_t colbase;
};
struct TileInfo;
// Autotuned column factor of one (density, capacity) bucket.
struct TuneFactor {
    int32_t rb, kb;
    double factor;
};
// Column starts of the tiles' real columns when the tiles' column tables
// are laid out by capacity (grids[t].colbase): pre (T + 1 entries) receives
// the prefix of the real entries; the launch shape depends only on
// max_entries.
void launch_tile_colstart_real(const uint32_t* keys, const TileGrid* grids, uint32_t T,
                               uint64_t* pre, uint32_t* colstart, uint64_t max_entries,
                               hipStream_t s);
// Every tile's bounds encoding reset (min slots all ones, max slots zero).
void launch_bounds_init(unsigned long long* benc, uint32_t T, hipStream_t s);
// Every tile's grid from its encoded bounds on the device (grid_params with
// the factor looked up in `table`, else default_factor): grids[t] (colbase
// from `colbase`), bounds (6 per tile, decoded) and, if given, the grid,
// column base and certification margin of infos[t].
void launch_tile_grids(const unsigned long long* benc, const uint64_t* starts,
                       const uint32_t* fetch, const uint64_t* colbase, uint32_t T,
                       const TuneFactor* table, int ntable, double default_factor,
                       double cell_z_factor, TileGrid* grids, double* bounds, TileInfo* infos,
                       hipStream_t s);
void launch_tile_bounds(const double* x, const double* y, const double* z,
                        const SegBlock* blocks, uint32_t nblocks, unsigned long long* benc,
                        hipStream_t s);
void launch_tile_keys(const double* x, const double* y, const double* z, const SegBlock* blocks,
                      uint32_t nblocks, const TileGrid* grids, uint32_t* keys, uint32_t* ids,
                      double* xyz, hipStream_t s);
void launch_seg_radix_sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                           const SegBlock* blocks, uint32_t nblocks, uint32_t bits,
                           uint32_t* hist, uint32_t* scan_scratch, uint32_t** kres,
                           uint32_t** vres, hipStream_t s);
void launch_tile_colstart(const uint32_t* keys, const TileGrid* grids, uint32_t T,
                          uint64_t entries, uint32_t* colstart, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
void launch_copy_f64(const double* src, double* dst, uint64_t n, hipStream_t s);

// knn.hip
struct KnnOutputs {
    double* col[kMaxOutColumns];  // device output columns (original point order)
    uint8_t* status[kMaxStages];
    uint32_t* ids;           // neighbor-list mode: n x k (original order)
    uint32_t* flagged;       // sorted indices of tie/incomplete rows
    uint32_t* nflagged;      // [0] flagged rows, [1] tie rows, [2] incomplete rows
    double* cov;             // per sorted row a kCovRecord-double record: the 6 unique
                             // covariance entries, then the row flags (as a uint64)
    unsigned long long* counters; // optional: passes, candidates, merges, queries
};

struct KnnTuning {
    int32_t ring0;       // first ring (box half-width in columns)
    int32_t ring_max;    // budget before a row is marked incomplete
    int32_t small_col;   // columns up to this size are taken whole
    int32_t group_z;     // z cells beyond the first a query group may span
    int32_t q_per_wave;  // queries per wave (0: by k)
    int32_t waves;       // waves per workgroup of the single-cloud search (0: 4)
    int32_t set_select;  // exact-count selection for set neighborhoods (1; 0: sorting)
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

// Doubles per covariance record (64 bytes: one cache line per row).
constexpr uint32_t kCovRecord = 8;

// One tile (an independent cloud) of a batch.  Tiles are contiguous in the
// batch's point order, and each tile's sorted rows stay in its own range
// [start, start + n): sorting never crosses a tile boundary.
struct TileInfo {
    GridParams g;
    FusedProgram prog;     // k, stride, fetch of this tile (stages are shared)
    uint64_t start;        // first point of the tile
    uint64_t n;            // points
    uint64_t colbase;      // first entry of the tile's column starts
    double margin;         // certification margin of its grid
    uint32_t root;         // root node of its exact tree
    uint32_t all_exact;    // fewer points than the result capacity: every row exact
};

// The queries of one wavefront in the tiled search: [q0, q0 + nq) of tile.
// Batch tables (device memory); tiles == nullptr means a single cloud whose
// grid and program come from the CloudView / FusedProgram arguments.
// las.hip -- LAS point records.
struct LasXform {
    double scale[3], offset[3];
};
struct LasColumns {
    const double* col[16];
};
// LAS Extra Bytes data types (1..10) of the columns.
struct LasTypes {
    int32_t type[16];
};
// Bytes of LAS Extra Bytes data type t (1..10), else 0.
uint32_t las_type_size(int t);
// Row i's columns, each converted to its type, one after another from
// out + i * stride + offset, little-endian.
void launch_las_pack_typed(const LasColumns& cols, const LasTypes& types, int ncols, uint64_t n,
                           uint8_t* out, uint32_t stride, uint32_t offset, hipStream_t s);
bool las_parse(const uint8_t* b, uint64_t nbytes, gf_las_info& h, std::string& err);
// Base record size of a supported point format (0, 1, 2, 3, 6, 7, 8), else 0.
uint32_t las_base_size(int format);
// rec: n records of L bytes in device memory, padded by at least 4 bytes.
void launch_las_xyz(const uint8_t* rec, uint64_t n, uint32_t L, const LasXform& xf, double* x,
                    double* y, double* z, hipStream_t s);
void launch_las_pack(const LasColumns& cols, int ncols, uint64_t n, uint8_t* out,
                     uint32_t stride, uint32_t offset, hipStream_t s);

struct TileTables {
    const TileInfo* tiles;
    const uint64_t* starts;  // T + 1 tile starts (for row -> tile lookups)
    uint32_t ntiles;
    // Wave tasks: tile t's rows are cut into runs of q_per_wave queries,
    // tasks task0[t] .. task0[t+1]-1 (T + 1 entries; 0 tasks for tiles
    // searched only exactly).
    const uint32_t* task0;
    uint32_t ntasks;
    int32_t max_k;           // largest k over the tiles (LDS sizing)
    int32_t q_per_wave;      // queries per task
};

// Queries per wavefront task for a neighborhood of k (LDS lists scale with k).
inline int knn_queries_per_wave(int max_k) { return max_k <= 32 ? 64 : 32; }

// mode 0: neighborhood covariance into out.cov / out.flags (sorted order);
// mode 1: neighbor lists into out.ids; mode 2: the same, every row with two
// equal distances in its list flagged, so that the lists also follow
// KD3Index's order among equal distances after the exact re-solve.
void launch_knn(const CloudView& cv, const FusedProgram& prog, int mode,
                const KnnOutputs& out, const KnnTuning& tune, hipStream_t s,
                const TileTables* tt = nullptr);
// Radius neighborhoods: mode 0 writes each sorted row's covariance record
// (FP64 two-pass over the points with d2 < radius^2, neighbor count in word
// 7); mode 1 writes the counts (original order) to `counts`.
void launch_radius(const CloudView& cv, double radius, int mode, double* cov, uint32_t* counts,
                   int group_z, hipStream_t s);
// Eigensystem + stage columns from out.cov, in original point order (inv:
// original id -> sorted position), so outputs are written contiguously.
// rows: kFeaturesAll every row; kFeaturesUnflagged every row the search did
// not flag (launched right after the search, beside the exact tree);
// kFeaturesFlagged the nflagged rows of out.flagged, after their re-solve.
constexpr int kFeaturesAll = 0, kFeaturesUnflagged = 1, kFeaturesFlagged = 2;
// The flagged rows' features with their count read on the device (at most
// max_rows; the launch shape depends only on max_rows).
void launch_features_flagged(const CloudView& cv, const FusedProgram& prog,
                             const KnnOutputs& out, const uint32_t* inv, uint64_t max_rows,
                             hipStream_t s, const TileTables* tt = nullptr);
void launch_features(const CloudView& cv, const FusedProgram& prog,
                     const KnnOutputs& out, const uint32_t* inv, hipStream_t s,
                     const TileTables* tt = nullptr, int rows = kFeaturesAll,
                     uint64_t nflagged = 0);

// kdtree.hip -- the exact KD3Index tree (nanoflann 1.3.1 semantics).
struct KdView {
    uint32_t* vind;          // n: tree order -> original id
    double *px, *py, *pz;    // n: coordinates in tree order
    uint32_t *nl, *nr;       // node segment [nl, nr)
    uint32_t* split;         // split position of internal nodes
    int32_t *c1, *c2;        // children (-1: leaf)
    int32_t* feat;           // cut axis (-1: leaf)
    int32_t* parent;
    uint8_t* which;          // 0: child1 of its parent, 1: child2
    double *divlow, *divhigh, *cutval;
    double* box;             // 6 per node: the box passed down (lo xyz, hi xyz)
    unsigned long long *mn, *mx;  // 3 per node: order-encoded extents (build)
    uint32_t* aux;           // per node: index in the next big-node list (build)
    uint32_t cap;            // node capacity
    uint32_t pad;
};

struct KdScratch {
    uint32_t *s1, *s2, *tmp;   // n + 1 each: pass prefixes (active index), pairing
    uint32_t* sums;            // per-tile sums of the prefix scans
    uint32_t *act[2], *off[2]; // big nodes of a level and their point-count prefix
    uint32_t* small;           // nodes handed to the per-workgroup subtree builder
    uint32_t list_cap;         // entries of act/small (off: list_cap + 1)
    uint32_t* hdr;             // 8 words: level header (counts, overflow, nodes)
};

// Build strategy knobs (the tree does not depend on them).
struct KdTuning {
    uint32_t subtree_points;    // nodes up to this size go to one workgroup (<= 4096)
    uint32_t minmax_lds_nodes;  // levels up to this many big nodes stage extents in LDS
    uint32_t wave_node;         // subtree nodes up to this size are built by one wave
    uint32_t small_blocks;      // one-wave subtree workgroups of <= 512 points
};

// Words of the exact-tree build header at the front of the node-list buffer.
constexpr uint32_t kKdHeaderWords = 16;
uint32_t kd_node_capacity(uint64_t n);
uint64_t kd_sums_words(uint64_t n);  // per-block sums of the level prefix scans
uint32_t kd_list_capacity(uint64_t n);
// Sends every row to the exact search: used when the cloud holds fewer
// points than the result capacity stride*min(n, k).
void launch_flag_all(uint64_t n, const uint32_t* perm, int mode, const KnnOutputs& out,
                     hipStream_t s);
// The same for the tiles of a batch marked all_exact.
void launch_flag_tiles(uint64_t n, const uint32_t* perm, int mode, const KnnOutputs& out,
                       const TileTables& tt, hipStream_t s);
// 0 on success, else 1 (node capacity) or 2 (node lists) to grow and retry,
// 3 if a subtree level overflowed its table (cannot happen for leaf size 100).
// A forest: tile r ([starts[r], starts[r+1]) of the points, bounds at
// 6 r) is rooted at node r.
uint32_t launch_kd_build(const double* x, const double* y, const double* z, uint64_t n,
                         const uint64_t* starts, uint32_t ntiles, const double* bounds,
                         KdView t, KdScratch w, const KdTuning& kt, uint32_t* nodes,
                         hipStream_t s);
// failed[0] counts rows whose search needs more than 64 recursion frames (a
// tree deeper than 64 levels); deep_rows (nflagged entries, may be null)
// receives their flagged indices for launch_kd_resolve_deep.
void launch_kd_resolve(const KdView& t, const CloudView& cv, const FusedProgram& prog, int mode,
                       const KnnOutputs& out, const double* x, const double* y, const double* z,
                       uint32_t nflagged, uint32_t* failed, uint32_t* deep_rows, hipStream_t s,
                       const TileTables* tt = nullptr);
// The listed rows (ndeep: device count) with `depth` frames per wave in
// global memory (frames: waves * depth * kd_deep_frame_bytes() bytes);
// failed2 counts rows that still overflow.
size_t kd_deep_frame_bytes();
// Largest result capacity ((knn + 1) * stride) of the exact search; above 64
// its result set lives in LDS.
uint32_t kd_max_capacity();
void launch_kd_resolve_deep(const KdView& t, const CloudView& cv, const FusedProgram& prog,
                            int mode, const KnnOutputs& out, const double* x, const double* y,
                            const double* z, const uint32_t* deep_rows, const uint32_t* ndeep,
                            void* frames, int depth, uint32_t waves, uint32_t* failed2,
                            hipStream_t s, const TileTables* tt = nullptr);
// Levels of the deepest of the trees rooted at nodes [0, nroots).
int kd_tree_depth(const std::vector<int32_t>& c1, const std::vector<int32_t>& c2,
                  uint32_t nroots);

// optnbhd.hip -- filters.optimalneighborhood from n x kmax KD3-ordered lists
// of sorted positions (rows in original order; x, y, z original order, sx,
// sy, sz sorted); failed counts rows whose eigen solve did not converge.
void launch_optimal_nbhd(const uint32_t* ids, uint64_t n, int kmin, int kmax, const double* x,
                         const double* y, const double* z, const double* sx, const double* sy,
                         const double* sz, uint64_t* kopt, double* ropt, uint32_t* failed,
                         hipStream_t s);
// nndistance (kNndKth, kNndAvg) and the outlier filter's running mean
// distance (kNndRunningMean) per row from n x kk lists of sorted positions of
// the K = k + 1 nearest points (kk = min(n, K)); rows in original order.
void launch_nn_distance(const uint32_t* ids, int kk, int K, uint64_t n, int mode,
                        const double* x, const double* y, const double* z, const double* sx,
                        const double* sy, const double* sz, double* out, hipStream_t s);
// Stages over row i's first min(row_k[i], kk) neighbors of an n x kk list of
// sorted positions (KD3Index order); rows in original order, no search flags.
void launch_prefix_stages(const uint32_t* ids, int kk, const uint32_t* row_k, uint64_t n,
                          const double* x, const double* y, const double* z, const double* sx,
                          const double* sy, const double* sz, const FusedProgram& prog,
                          const KnnOutputs& out, hipStream_t s);
// out[0] = max of v[0..n) (out cleared first).
void launch_max_u32(const uint32_t* v, uint64_t n, uint32_t* out, hipStream_t s);
// The flagged rows' lists (point ids from the exact re-solve) to sorted
// positions.
void launch_lists_to_sorted(uint32_t* ids, int kmax, const uint32_t* flagged, uint32_t nflagged,
                            const uint32_t* perm, const uint32_t* inv, hipStream_t s);

// Debug/parity hook: device eigen solver over m matrices (row-major 3x3).
void launch_eigen_batch(const double* cov, uint64_t m, double* evals,
                        double* evecs, int32_t* info, hipStream_t s);

} // namespace gf
