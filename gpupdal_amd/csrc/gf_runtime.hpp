// gf_runtime.hpp -- host runtime objects behind the C ABI: grow-only device
// buffers, the context (device, one default-priority stream, timing events,
// tuning, scratch) and resident clouds with their cached index.
#pragma once

#include "gpupdal_amd.h"

#include "gf_internal.hpp"
#include "gf_kernels.hpp"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace gf {

// Grow-only device allocation.  Hot paths call ensure() with the size they
// need; memory is allocated only when a larger size is first requested.
struct DevBuf {
    void* ptr = nullptr;
    uint64_t size = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        size = 0;
    }
    void ensure(uint64_t bytes)
    {
        if (bytes <= size && ptr) return;
        release();
        uint64_t b = bytes ? bytes : 8;
        hipError_t e = hipMalloc(&ptr, b);
        if (e != hipSuccess) {
            ptr = nullptr;
            throw Error(GF_ERR_OOM, std::string("hipMalloc(") + std::to_string(b) +
                                        "): " + hipGetErrorString(e));
        }
        size = b;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Autotuned column-size factors, per (2-D density, result capacity) bucket:
// rho_bucket = floor(log2(points per m^2)), k_bucket = ceil(log2(fetch)).
// One autotuned (density, capacity) bucket: the column-size factor and the
// search's launch shape (0: the context's setting).
struct TuneEntry {
    int rho_bucket;
    int k_bucket;
    double cell_factor;
    int group_z = 0;      // 0: default; else z cells + 1
    int q_per_wave = 0;
    int knn_waves = 0;
};


struct Tuning {
    double cell_factor = 0.6;    // column edge = factor * sqrt(k / (pi * rho2d))
    std::vector<TuneEntry> table;  // autotuned factors (override cell_factor)
    double forced_factor = 0.0;    // > 0: used for every grid (autotuning runs)
    double cell_z_factor = 1.0;  // z cell height = factor * column edge
    int ring0 = 1;
    int ring_max = 8;
    int small_col = 0;
    int group_z = 3;             // measured: flat for knn 16, 4% faster search for knn 32
    int q_per_wave = 0;          // queries per search wave (0: 64 up to k 32, else 32)
    int knn_waves = 0;           // waves per search workgroup (0: 4)
    int set_select = 1;          // exact-count selection for set neighborhoods (0: sorting)
    double radius_cell = 0.5;    // radius grids: column edge = radius_cell * radius
    // Exact tree: subtree-per-workgroup threshold (points; 2048-point
    // workgroups, two per CU, measured fastest alone and beside the search).
    int kd_subtree = 2048;
    int kd_wave_node = 256;      // exact tree: one-wave subtree threshold (points)
    // Exact tree: one-wave subtree workgroups of <= 512 points (-1: when built
    // beside the search, 0: never, 1: always).
    int kd_small_blocks = 0;
    int kd_minmax_lds = 128;     // exact tree: LDS-staged extents up to this many nodes (<= 128)
    // Clouds (batches) of at least this many points build the exact tree on
    // a side stream during the search, before knowing whether a row needs
    // it; smaller ones build it only when a row is flagged.  0: never early.
    uint64_t kd_overlap_min = 131072;
    // Batches of at least kd_overlap_min points run their step as a
    // captured, replayed graph (0: launch by launch).
    int batch_graph = 1;
    bool collect_stats = false;

    void set(const char* key, double v)
    {
        if (!std::strcmp(key, "cell_factor") && v > 0) cell_factor = v;
        else if (!std::strcmp(key, "cell_z_factor") && v > 0) cell_z_factor = v;
        else if (!std::strcmp(key, "ring0") && v >= 1) ring0 = static_cast<int>(v);
        else if (!std::strcmp(key, "ring_max") && v >= 1) ring_max = static_cast<int>(v);
        else if (!std::strcmp(key, "small_col") && v >= 0) small_col = static_cast<int>(v);
        else if (!std::strcmp(key, "group_z") && v >= 0) group_z = static_cast<int>(v);
        else if (!std::strcmp(key, "q_per_wave") && (v == 0 || v == 16 || v == 32 || v == 64))
            q_per_wave = static_cast<int>(v);
        else if (!std::strcmp(key, "knn_waves") && (v == 0 || v == 1 || v == 2 || v == 4))
            knn_waves = static_cast<int>(v);
        else if (!std::strcmp(key, "set_select") && (v == 0 || v == 1))
            set_select = static_cast<int>(v);
        else if (!std::strcmp(key, "radius_cell") && v >= 0.2 && v <= 4.0) radius_cell = v;
        else if (!std::strcmp(key, "kd_subtree") && v >= 101 && v <= 4096)
            kd_subtree = static_cast<int>(v);
        else if (!std::strcmp(key, "kd_wave_node") && v >= 101 && v <= 1024)
            kd_wave_node = static_cast<int>(v);
        else if (!std::strcmp(key, "kd_small_blocks") && v >= -1 && v <= 1)
            kd_small_blocks = static_cast<int>(v);
        else if (!std::strcmp(key, "kd_minmax_lds") && v >= 0) kd_minmax_lds = static_cast<int>(v);
        else if (!std::strcmp(key, "kd_overlap_min") && v >= 0)
            kd_overlap_min = static_cast<uint64_t>(v);
        else if (!std::strcmp(key, "collect_stats")) collect_stats = v != 0.0;
        else if (!std::strcmp(key, "batch_graph")) batch_graph = v != 0.0 ? 1 : 0;
        else throw Error(GF_ERR_ARG, std::string("unknown or invalid tuning key: ") + key);
        if (ring_max < ring0) ring_max = ring0;
    }
    // The factor for a grid of this density and result capacity.
    double factor_for(double rho, uint32_t fetch) const
    {
        if (forced_factor > 0) return forced_factor;
        const int rb = tune_rho_bucket(rho), kb = tune_k_bucket(fetch);
        for (const TuneEntry& e : table)
            if (e.rho_bucket == rb && e.k_bucket == kb) return e.cell_factor;
        return cell_factor;
    }
    // The autotuned entry of a bucket, or null.
    const TuneEntry* entry_for(double rho, uint32_t fetch) const
    {
        const int rb = tune_rho_bucket(rho), kb = tune_k_bucket(fetch);
        for (const TuneEntry& e : table)
            if (e.rho_bucket == rb && e.k_bucket == kb) return &e;
        return nullptr;
    }
    void remember(double rho, uint32_t fetch, const TuneEntry& v)
    {
        const int rb = tune_rho_bucket(rho), kb = tune_k_bucket(fetch);
        TuneEntry e = v;
        e.rho_bucket = rb;
        e.k_bucket = kb;
        for (TuneEntry& x : table)
            if (x.rho_bucket == rb && x.k_bucket == kb) {
                x = e;
                return;
            }
        table.push_back(e);
    }
    // Grid-shaping knobs only (the ring knobs do not change the index).
    uint64_t signature() const
    {
        uint64_t a, b, c;
        std::memcpy(&a, &cell_factor, 8);
        std::memcpy(&b, &cell_z_factor, 8);
        std::memcpy(&c, &forced_factor, 8);
        uint64_t h = a * 1000003ull ^ b ^ (c * 7919ull);
        for (const TuneEntry& e : table) {
            uint64_t f;
            std::memcpy(&f, &e.cell_factor, 8);
            h = h * 1099511628211ull ^ (static_cast<uint64_t>(e.rho_bucket + 4096) << 20) ^
                static_cast<uint64_t>(e.k_bucket) ^ f;
        }
        return h;
    }
};

struct Context;

// Device storage of a cloud's exact KD3Index tree.
struct KdBuffers {
    DevBuf vind, px, py, pz, nodes_u32, nodes_i32, which, nodes_f64, ext;
    uint32_t cap = 0;
    uint32_t list_cap = 0;
    uint32_t nodes = 0;
    KdView view{};
};

struct Cloud {
    Context* ctx;
    uint64_t n = 0;
    DevBuf x, y, z;              // original order, FP64
    // index (valid for idx_fetch / idx_tune)
    bool idx_valid = false;
    uint32_t idx_fetch = 0;
    double idx_cell = 0.0;       // forced column edge of a radius grid (0: density rule)
    uint64_t idx_tune = 0;
    GridParams g{};
    DevBuf keys_a, vals_a, keys_b, vals_b, colstart, sx, sy, sz;
    uint32_t* keys = nullptr;    // sorted keys (into keys_a/keys_b)
    uint32_t* perm = nullptr;    // sorted position -> original id
    DevBuf inv;                  // original id -> sorted position
    double bounds[6] = {0, 0, 0, 0, 0, 0};  // min xyz, max xyz
    DevBuf dbounds;
    DevBuf dstarts;              // {0, n}: the cloud as a one-tile forest
    bool kd_valid = false;       // exact tree (built on demand)
    KdBuffers kd;
    explicit Cloud(Context* c) : ctx(c) {}
};

// A batch of independent tiles held back to back on the device: one index
// build, one search, one exact re-solve and one feature pass serve them all.
struct Batch {
    Context* ctx;
    uint32_t T = 0;
    uint64_t n = 0;
    std::vector<uint64_t> starts;        // T + 1
    std::vector<SegBlock> blocks;        // segment blocks (never cross a tile)
    DevBuf x, y, z;
    DevBuf dstarts, dblocks;
    // index, valid for idx_fetch (per tile) and idx_tune
    bool idx_valid = false;
    std::vector<uint32_t> idx_fetch;
    uint64_t idx_tune = 0;
    std::vector<TileGrid> grids;
    std::vector<double> bounds;          // 6 per tile
    DevBuf dgrids, dbounds, benc, hist;
    DevBuf keys_a, vals_a, keys_b, vals_b, colstart, sx, sy, sz, inv;
    uint32_t* keys = nullptr;
    uint32_t* perm = nullptr;
    uint64_t col_entries = 0;
    // run tables
    DevBuf dtiles, dtasks;
    bool kd_valid = false;
    KdBuffers kd;
    // Graph mode: the step from the bounds through the unflagged rows'
    // features captured once per shape (two parts: bounds and grids, then
    // the rest, so that the exact forest can start beside the second), and
    // replayed; gkey holds what the capture depends on.
    hipGraphExec_t gexec_a = nullptr, gexec_b = nullptr;
    std::vector<uint64_t> gkey, seen_key;
    DevBuf dfetch, dcolbase, dcolpre, dtune;
    explicit Batch(Context* c) : ctx(c) {}
    ~Batch();
    void drop_graph();
    void reset(uint32_t ntiles, const uint64_t* counts, hipStream_t s);
};

// Points per unit of xy extent of bounds b (min xyz, max xyz) holding n
// points: the density the grid rule and the autotune buckets use.
double cloud_rho(const double b[6], uint64_t n);
// cell > 0: that column edge instead of the density rule (radius grids).
GridParams make_grid(const double b[6], uint64_t n, uint32_t fetch, const Tuning& t,
                     double cell = 0.0);
int stage_ncols(const gf_stage& s);
// The stages' device programs (columns laid out in stage order); returns the
// column count, GF_ERR_UNSUPPORTED above kMaxOutColumns.
int fill_stage_programs(FusedProgram& prog, const gf_stage* st, int ns);
void validate_stage(const gf_stage& s);
// Throws GF_ERR_UNSUPPORTED when a cloud (or tile) of n points whose search
// asks K neighbors with this stride holds more than `limit` entries: 63 for
// the grid search (batches, serving, the no-allocation path), the exact
// search's capacity for gf_run.
void check_neighborhood(uint64_t n, int64_t K, int stride, uint32_t limit);
// Neighbors a stage's search asks (knn + 1, or knn for approximatecoplanar
// and estimaterank).
int64_t stage_neighbors(const gf_stage& s);

// A device-time interval between two events, read after the call's final
// synchronisation (timing adds no host wait between kernels).
struct Span {
    hipEvent_t a, b;
    double* dst;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    // Second default-priority stream: the exact tree, which depends only on
    // the points, is built on it while the search runs on `stream`.
    hipStream_t side = nullptr;
    hipEvent_t join = nullptr;
    // join was recorded on `stream` right after the current points' bounds
    // reached the device (the tree needs nothing else), so the side stream
    // may start during the rest of the index build.
    bool points_mark = false;
    std::vector<hipEvent_t> evpool;
    size_t evnext = 0;
    std::vector<Span> spans;
    bool timing = true;
    Tuning tune;
    gf_stats stats{};
    std::string last_error;
    DevBuf part, bounds, hist, scan_tmp, outbuf, statbuf, flagged, counters, covbuf;
    DevBuf xyz_stage;  // interleaved copy of the points between the keys and the gather
    DevBuf lasbuf;     // LAS point records (decode input, extra-bytes packing)
    DevBuf listbuf, liststat;  // kNN lists (gf_knn host copies, nndistance, outlier)
    DevBuf deep_rows, deep_frames;  // exact re-solve of trees deeper than 64 levels
    // Exact-tree scratch, its own: the tree may be built on the side stream
    // while the index build uses scan_tmp on the main stream.
    DevBuf kd_s1, kd_s2, kd_tmp, kd_lists, kd_sums;
    double* host_bounds = nullptr;

    explicit Context(int dev) : device(dev)
    {
        int count = 0;
        GF_HIP(hipGetDeviceCount(&count));
        if (dev < 0 || dev >= count)
            throw Error(GF_ERR_ARG, "device index out of range");
        GF_HIP(hipSetDevice(dev));
        GF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        GF_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        GF_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        GF_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_bounds), 64, hipHostMallocDefault));
    }
    ~Context()
    {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        if (host_bounds) (void)hipHostFree(host_bounds);
        for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    KnnTuning knn_tuning() const
    {
        KnnTuning k{};
        k.ring0 = tune.ring0;
        k.ring_max = tune.ring_max;
        k.small_col = tune.small_col;
        k.group_z = tune.group_z;
        k.q_per_wave = tune.q_per_wave;
        k.waves = tune.knn_waves;
        k.set_select = tune.set_select;
        return k;
    }
    // The search shape for a cloud and result capacity: the context's
    // settings, with an autotuned bucket's overriding them.
    KnnTuning knn_tuning(const Cloud& cl, uint32_t fetch) const;

    void build_index(Cloud& cl, uint32_t fetch, double cell = 0.0);
    // The grid of the cloud's current index (host copy).
    GridParams grid_of(const Cloud& cl) const { return cl.g; }
    CloudView view(const Cloud& cl) const;
    void run_group(Cloud& cl, const gf_stage* st, int ns, double* const* cols,
                   uint8_t* const* status, int where);
    void run_group_radius(Cloud& cl, const gf_stage* st, int ns, double* const* cols,
                          uint8_t* const* status, int where);
    // beside: built on the side stream while the search runs.
    void build_kdtree(Cloud& cl, hipStream_t s, bool beside);
    void build_kdtree(Cloud& cl) { build_kdtree(cl, stream, false); }
    void build_forest(KdBuffers& b, const double* x, const double* y, const double* z,
                      uint64_t n, const uint64_t* dstarts, uint32_t T, const double* dbounds,
                      hipStream_t s, bool beside);
    // Makes the side stream wait for the points and bounds (points_mark), or
    // for all work queued on `stream` so far.
    void fork_side();
    // KD3Index::neighbors(id, k, stride) lists of every point (n x min(n, k),
    // rows in original order) and each row's status; mode 1 writes point
    // ids, mode 2 sorted positions (stride 1 only).  Queued on the stream.
    void knn_lists(Cloud& cl, uint32_t k, uint32_t stride, int mode, uint32_t* ids,
                   uint8_t* status);
    // gf_run_rowk: the stages over each row's first row_k[i] KD3Index
    // neighbors (row_k: device array).
    void run_rowk(Cloud& cl, const gf_stage* st, int ns, const uint32_t* row_k,
                  double* const* cols, uint8_t* const* status, int where);
    // Per-row distance statistic (kNndKth, kNndAvg, kNndRunningMean) of the
    // k + 1 nearest points into out (device, original order).
    void nn_distance(Cloud& cl, uint32_t k, int mode, double* out);
    // Re-solves the rows the first exact pass listed in deep_rows (count
    // nfail on the host, dcount on the device) with frames for the whole
    // tree depth; trees rooted at nodes [0, nroots).
    void resolve_deep(KdBuffers& kd, uint32_t nroots, const CloudView& cv,
                      const FusedProgram& prog, int mode, const KnnOutputs& out,
                      const double* x, const double* y, const double* z, uint32_t nfail,
                      const uint32_t* dcount, const TileTables* tt);
    // gf_las_write_extra: the widened file image into dst (or its size only,
    // dst null); cols host or device as `where`.
    void las_write_extra(const uint8_t* src, uint64_t src_nbytes, int ncols,
                         const char* const* names, const char* const* descriptions,
                         const uint8_t* types, const double* const* cols, int where,
                         uint8_t* dst, uint64_t dst_capacity, uint64_t* dst_nbytes);
    // Returns the number of flagged rows.
    uint32_t resolve_flagged(Cloud& cl, const CloudView& cv, const FusedProgram& prog,
                             int mode, const KnnOutputs& out, hipEvent_t from);
    // The device address of this GPU's fixed scratch array (the no-allocation
    // entry point's only working memory), bound when the context is created.
    void* static_scratch = nullptr;
    // gf_run_noalloc: one group of stages sharing a neighborhood, every
    // working buffer carved from static_scratch.
    void run_group_static(const double* x, const double* y, const double* z, uint64_t n,
                          const gf_stage* st, int ns, double* const* cols,
                          uint8_t* const* status);
    void build_batch_index(Batch& b, const std::vector<uint32_t>& fetch);
    void run_batch(Batch& b, const gf_stage* st, int ns, const int32_t* tile_knn,
                   double* const* cols, uint8_t* const* status, int where);
    // The batch step as a replayed graph (large batches; tuning
    // "batch_graph").  tiles/fetch: the host tables run_batch prepared.
    // Returns false, having run nothing, the first time a shape is seen.
    bool run_batch_graph(Batch& b, int ns, std::vector<TileInfo>& tiles,
                         const std::vector<uint32_t>& fetch, const FusedProgram& base,
                         int ncol_total, int max_k, bool any_exact_tile, double* const* cols,
                         uint8_t* const* status, int where);
    // Timing: mark() records a pooled event on a stream (nullptr when timing
    // is off); span() registers an interval; collect_spans(), after the
    // call's final synchronisation, fills every registered interval.
    hipEvent_t mark(hipStream_t s);
    hipEvent_t mark() { return mark(stream); }
    void span(hipEvent_t a, hipEvent_t b, double* dst);
    void collect_spans();
    void reset_spans();
};

} // namespace gf
