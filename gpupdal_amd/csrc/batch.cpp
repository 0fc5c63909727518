// batch.cpp -- batches of independent tiles behind the C ABI: tiles of
// different sizes and neighborhood sizes share one index build (per-tile
// grids, a tile-segmented radix sort), one search launch, one exact-tree
// forest and one feature launch.  A tile's rows depend only on that tile, so
// every result equals the tile run alone.
#include "gpupdal_amd.h"

#include "gf_internal.hpp"
#include "gf_kernels.hpp"
#include "gf_runtime.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace gf {

namespace {

double dec_order(unsigned long long u)
{
    const unsigned long long b = (u >> 63) ? (u ^ 0x8000000000000000ull) : ~u;
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

} // namespace

Batch::~Batch() { drop_graph(); }

void Batch::drop_graph()
{
    if (gexec_a) (void)hipGraphExecDestroy(gexec_a);
    if (gexec_b) (void)hipGraphExecDestroy(gexec_b);
    gexec_a = gexec_b = nullptr;
    gkey.clear();
    seen_key.clear();
}

// New tile sizes: tile starts, segment blocks (each tile cut into blocks of
// at most 4096 points) and grow-only point buffers.  The index is dropped.
void Batch::reset(uint32_t ntiles, const uint64_t* counts, hipStream_t s)
{
    drop_graph();
    T = ntiles;
    starts.assign(static_cast<size_t>(ntiles) + 1, 0);
    for (uint32_t t = 0; t < ntiles; ++t) starts[t + 1] = starts[t] + counts[t];
    n = starts[ntiles];
    if (n >= (1ull << 32) - 1)
        throw Error(GF_ERR_UNSUPPORTED, "batches are limited to 2^32-2 points");
    blocks.clear();
    for (uint32_t t = 0; t < ntiles; ++t) {
        const uint64_t nt = counts[t];
        const uint32_t nbt = static_cast<uint32_t>(div_up(nt, kSegBlockPoints));
        const uint32_t first = static_cast<uint32_t>(blocks.size());
        for (uint32_t i = 0; i < nbt; ++i) {
            SegBlock sb{};
            sb.begin = static_cast<uint32_t>(starts[t] + static_cast<uint64_t>(i) * kSegBlockPoints);
            sb.count = static_cast<uint32_t>(
                std::min<uint64_t>(kSegBlockPoints, nt - static_cast<uint64_t>(i) * kSegBlockPoints));
            sb.tile = t;
            sb.hbase = 256u * first;
            sb.nbt = nbt;
            sb.lb = i;
            blocks.push_back(sb);
        }
    }
    x.ensure(n * 8 + 8);
    y.ensure(n * 8 + 8);
    z.ensure(n * 8 + 8);
    dstarts.ensure(starts.size() * 8 + 8);
    GF_HIP(hipMemcpyAsync(dstarts.ptr, starts.data(), starts.size() * 8, hipMemcpyHostToDevice, s));
    dblocks.ensure(blocks.size() * sizeof(SegBlock) + 8);
    if (!blocks.empty())
        GF_HIP(hipMemcpyAsync(dblocks.ptr, blocks.data(), blocks.size() * sizeof(SegBlock),
                              hipMemcpyHostToDevice, s));
    GF_HIP(hipStreamSynchronize(s));
    idx_valid = false;
    kd_valid = false;
}

void Context::build_batch_index(Batch& b, const std::vector<uint32_t>& fetch)
{
    if (b.idx_valid && b.idx_fetch == fetch && b.idx_tune == tune.signature()) return;
    b.idx_valid = false;
    b.kd_valid = false;
    const uint32_t T = b.T;
    const uint64_t N = b.n;
    const uint32_t nb = static_cast<uint32_t>(b.blocks.size());
    hipEvent_t e0 = mark();

    // Per-tile bounds (order-encoded min/max), read back for the grids.
    std::vector<unsigned long long> init(static_cast<size_t>(T) * 6);
    for (uint32_t t = 0; t < T; ++t)
        for (int d = 0; d < 3; ++d) {
            init[6 * t + d] = ~0ull;
            init[6 * t + 3 + d] = 0ull;
        }
    b.benc.ensure(init.size() * 8 + 8);
    GF_HIP(hipMemcpyAsync(b.benc.ptr, init.data(), init.size() * 8, hipMemcpyHostToDevice,
                          stream));
    launch_tile_bounds(b.x.as<double>(), b.y.as<double>(), b.z.as<double>(),
                       b.dblocks.as<SegBlock>(), nb, b.benc.as<unsigned long long>(), stream);
    GF_HIP(hipMemcpyAsync(init.data(), b.benc.ptr, init.size() * 8, hipMemcpyDeviceToHost,
                          stream));
    GF_HIP(hipStreamSynchronize(stream));

    // Grids on the host (the same rule as a single cloud), column bases.
    b.grids.assign(T, TileGrid{});
    b.bounds.assign(static_cast<size_t>(T) * 6, 0.0);
    uint64_t colbase = 0;
    uint32_t key_bits = 0;
    for (uint32_t t = 0; t < T; ++t) {
        double* bd = &b.bounds[6 * t];
        const uint64_t nt = b.starts[t + 1] - b.starts[t];
        for (int i = 0; i < 6; ++i) bd[i] = nt ? dec_order(init[6 * t + i]) : 0.0;
        TileGrid& tg = b.grids[t];
        tg.g = make_grid(bd, nt, fetch[t], tune);
        tg.start = b.starts[t];
        tg.n = nt;
        tg.colbase = colbase;
        colbase += static_cast<uint64_t>(tg.g.ncols) + 1;
        key_bits = std::max(key_bits, tg.g.key_bits);
    }
    b.col_entries = colbase;
    if (colbase >= (1ull << 32))
        throw Error(GF_ERR_UNSUPPORTED, "batch column table above 2^32 entries");
    b.dgrids.ensure(static_cast<uint64_t>(T) * sizeof(TileGrid) + 8);
    GF_HIP(hipMemcpyAsync(b.dgrids.ptr, b.grids.data(), static_cast<uint64_t>(T) * sizeof(TileGrid),
                          hipMemcpyHostToDevice, stream));
    b.dbounds.ensure(b.bounds.size() * 8 + 8);
    GF_HIP(hipMemcpyAsync(b.dbounds.ptr, b.bounds.data(), b.bounds.size() * 8,
                          hipMemcpyHostToDevice, stream));
    GF_HIP(hipEventRecord(join, stream));
    points_mark = true;

    // Keys, tile-segmented sort, column starts, sorted coordinates.
    b.keys_a.ensure(N * 4 + 8);
    b.vals_a.ensure(N * 4 + 8);
    b.keys_b.ensure(N * 4 + 8);
    b.vals_b.ensure(N * 4 + 8);
    b.colstart.ensure(colbase * 4 + 8);
    b.sx.ensure(N * 8 + 8);
    b.sy.ensure(N * 8 + 8);
    b.sz.ensure(N * 8 + 8);
    b.hist.ensure(static_cast<uint64_t>(nb) * 256 * 4 + 8);
    scan_tmp.ensure(scan_scratch_words(static_cast<uint64_t>(nb) * 256 + 1) * 4);
    xyz_stage.ensure(N * 24 + 8);
    launch_tile_keys(b.x.as<double>(), b.y.as<double>(), b.z.as<double>(),
                     b.dblocks.as<SegBlock>(), nb, b.dgrids.as<TileGrid>(),
                     b.keys_a.as<uint32_t>(), b.vals_a.as<uint32_t>(), xyz_stage.as<double>(),
                     stream);
    uint32_t *kres, *vres;
    launch_seg_radix_sort(b.keys_a.as<uint32_t>(), b.vals_a.as<uint32_t>(),
                          b.keys_b.as<uint32_t>(), b.vals_b.as<uint32_t>(),
                          b.dblocks.as<SegBlock>(), nb, key_bits, b.hist.as<uint32_t>(),
                          scan_tmp.as<uint32_t>(), &kres, &vres, stream);
    launch_tile_colstart(kres, b.dgrids.as<TileGrid>(), T, colbase, b.colstart.as<uint32_t>(),
                         stream);
    b.inv.ensure(N * 4 + 8);
    launch_gather_sorted(vres, xyz_stage.as<double>(), N, b.sx.as<double>(), b.sy.as<double>(),
                         b.sz.as<double>(), b.inv.as<uint32_t>(), stream);
    b.keys = kres;
    b.perm = vres;
    b.idx_valid = true;
    b.idx_fetch = fetch;
    b.idx_tune = tune.signature();
    span(e0, mark(), &stats.ms_index);
}

void Context::run_batch(Batch& b, const gf_stage* st, int ns, const int32_t* tile_knn,
                        double* const* cols, uint8_t* const* status, int where)
{
    const uint32_t T = b.T;
    const uint64_t N = b.n;
    // One neighborhood group: every stage shares the stride and, per tile,
    // the knn (tile_knn overrides the stages' knn).
    const int stride = st[0].stage == GF_STAGE_NORMAL ? 1 : st[0].stride;
    for (int i = 1; i < ns; ++i) {
        const int si = st[i].stage == GF_STAGE_NORMAL ? 1 : st[i].stride;
        if (si != stride || (!tile_knn && st[i].knn != st[0].knn))
            throw Error(GF_ERR_UNSUPPORTED,
                        "a batch runs stages that share one neighborhood (knn and stride)");
    }
    FusedProgram base{};
    base.nstages = ns;
    base.stride = stride;
    const int ncol_total = fill_stage_programs(base, st, ns);
    reset_spans();
    hipEvent_t t_all = mark();

    // Per-tile programs.
    std::vector<TileInfo> tiles(T);
    std::vector<uint32_t> fetch(T);
    int max_k = 1;
    bool any_exact_tile = false;
    for (uint32_t t = 0; t < T; ++t) {
        const uint64_t nt = b.starts[t + 1] - b.starts[t];
        const int64_t knn = tile_knn ? tile_knn[t] : st[0].knn;
        if (knn < 1) throw Error(GF_ERR_ARG, "knn must be at least 1");
        const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(nt, static_cast<uint64_t>(knn) + 1));
        const uint32_t f = static_cast<uint32_t>(stride) * k;
        check_neighborhood(nt, static_cast<int64_t>(knn) + 1, stride, kMaxFetch - 1);
        fetch[t] = std::max<uint32_t>(f, 1);
        TileInfo& ti = tiles[t];
        ti.prog = base;
        ti.prog.k = static_cast<int32_t>(k);
        ti.prog.fetch = static_cast<int32_t>(f);
        ti.start = b.starts[t];
        ti.n = nt;
        ti.root = t;
        ti.all_exact = (nt > 0 && f > nt) ? 1u : 0u;
        any_exact_tile = any_exact_tile || ti.all_exact;
        max_k = std::max<int>(max_k, static_cast<int>(k));
    }
    if (tune.batch_graph && tune.kd_overlap_min && N >= tune.kd_overlap_min && T > 0 &&
        run_batch_graph(b, ns, tiles, fetch, base, ncol_total, max_k, any_exact_tile, cols,
                        status, where))
        return;
    points_mark = false;
    build_batch_index(b, fetch);
    // One search launch serves every tile: it takes the shape autotuned for
    // the bucket of the batch's largest tile (the context's otherwise).
    KnnTuning kt = knn_tuning();
    uint32_t big = 0;
    for (uint32_t t = 1; t < T; ++t)
        if (b.starts[t + 1] - b.starts[t] > b.starts[big + 1] - b.starts[big]) big = t;
    if (T) {
        const uint64_t nb = b.starts[big + 1] - b.starts[big];
        if (const TuneEntry* e = tune.entry_for(cloud_rho(&b.bounds[6 * big], nb), fetch[big])) {
            if (e->group_z > 0) kt.group_z = e->group_z - 1;
            if (e->q_per_wave > 0) kt.q_per_wave = e->q_per_wave;
        }
    }
    const int Q = kt.q_per_wave > 0 ? kt.q_per_wave : knn_queries_per_wave(max_k);
    std::vector<uint32_t> task0(static_cast<size_t>(T) + 1, 0u);
    uint64_t ntasks = 0;
    for (uint32_t t = 0; t < T; ++t) {
        TileInfo& ti = tiles[t];
        ti.g = b.grids[t].g;
        ti.colbase = b.grids[t].colbase;
        const GridParams& g = ti.g;
        ti.margin = 1e-6 * std::fmin(g.c, g.cz) +
                    1e-15 * (std::fabs(g.ox) + std::fabs(g.oy) + std::fabs(g.oz)) + 1e-300;
        task0[t] = static_cast<uint32_t>(ntasks);
        if (!ti.all_exact) ntasks += (ti.n + Q - 1) / Q;
    }
    task0[T] = static_cast<uint32_t>(ntasks);
    if (ntasks > 0xFFFFFF00ull)
        throw Error(GF_ERR_UNSUPPORTED, "batch above the search's task range");
    b.dtiles.ensure(static_cast<uint64_t>(T) * sizeof(TileInfo) + 8);
    GF_HIP(hipMemcpyAsync(b.dtiles.ptr, tiles.data(), static_cast<uint64_t>(T) * sizeof(TileInfo),
                          hipMemcpyHostToDevice, stream));
    b.dtasks.ensure(task0.size() * sizeof(uint32_t) + 16);
    GF_HIP(hipMemcpyAsync(b.dtasks.ptr, task0.data(), task0.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice, stream));
    TileTables tt{};
    tt.tiles = b.dtiles.as<TileInfo>();
    tt.starts = b.dstarts.as<uint64_t>();
    tt.ntiles = T;
    tt.task0 = b.dtasks.as<uint32_t>();
    tt.ntasks = static_cast<uint32_t>(ntasks);
    tt.max_k = max_k;
    tt.q_per_wave = Q;

    KnnOutputs out{};
    if (where == GF_WHERE_DEVICE) {
        for (int c = 0; c < ncol_total; ++c) out.col[c] = cols[c];
        for (int i = 0; i < ns; ++i) out.status[i] = status[i];
    } else {
        outbuf.ensure(static_cast<uint64_t>(ncol_total) * N * 8 + 8);
        statbuf.ensure(static_cast<uint64_t>(ns) * N + 8);
        for (int c = 0; c < ncol_total; ++c) {
            out.col[c] = outbuf.as<double>() + static_cast<uint64_t>(c) * N;
            GF_HIP(hipMemcpyAsync(out.col[c], cols[c], N * 8, hipMemcpyHostToDevice, stream));
        }
        for (int i = 0; i < ns; ++i) out.status[i] = statbuf.as<uint8_t>() + i * N;
    }
    flagged.ensure(N * 4 + 4);
    counters.ensure(64);
    covbuf.ensure(N * 8 * kCovRecord + 8);
    out.flagged = flagged.as<uint32_t>();
    out.nflagged = counters.as<uint32_t>();
    out.cov = covbuf.as<double>();
    launch_fill_u32(out.nflagged, 4, 0u, stream);
    if (tune.collect_stats) {
        out.counters = reinterpret_cast<unsigned long long*>(counters.as<uint32_t>() + 8);
        launch_fill_u32(counters.as<uint32_t>() + 8, 8, 0u, stream);
    }

    CloudView cv{};
    cv.sx = b.sx.as<double>();
    cv.sy = b.sy.as<double>();
    cv.sz = b.sz.as<double>();
    cv.perm = b.perm;
    cv.keys = b.keys;
    cv.colstart = b.colstart.as<uint32_t>();
    cv.n = N;
    cv.g = b.grids.empty() ? GridParams{} : b.grids[0].g;
    FusedProgram prog0 = tiles.empty() ? base : tiles[0].prog;

    // The tiles' exact trees depend only on the points: on large batches
    // they are built on the side stream while the search runs.
    const bool early = !b.kd_valid && tune.kd_overlap_min && N >= tune.kd_overlap_min;
    if (early) fork_side();
    hipEvent_t e0 = mark();
    launch_knn(cv, prog0, 0, out, kt, stream, &tt);
    if (any_exact_tile) launch_flag_tiles(N, cv.perm, 0, out, tt, stream);
    hipEvent_t e1 = mark();
    span(e0, e1, &stats.ms_search);
    // Rows the search did not flag are final: their features run now, beside
    // the rest of the tiles' trees; the flagged rows' after their re-solve.
    launch_features(cv, prog0, out, b.inv.as<uint32_t>(), stream, &tt, kFeaturesUnflagged);
    hipEvent_t e2 = mark();
    span(e1, e2, &stats.ms_features);
    auto forest = [&](hipStream_t s, bool beside) {
        if (b.kd_valid) return;
        build_forest(b.kd, b.x.as<double>(), b.y.as<double>(), b.z.as<double>(), N,
                     b.dstarts.as<uint64_t>(), T, b.dbounds.as<double>(), s, beside);
        b.kd_valid = true;
    };
    if (early) forest(side, true);

    // Exact re-solve of flagged rows on the tiles' KD3 trees.
    uint32_t h[4] = {0, 0, 0, 0};
    GF_HIP(hipMemcpyAsync(h, out.nflagged, sizeof(h), hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    stats.tie_rows = h[1];
    stats.incomplete_rows = h[2];
    if (h[0]) {
        forest(stream, false);
        uint32_t* failed = counters.as<uint32_t>() + 5;
        launch_fill_u32(failed, 1, 0u, stream);
        deep_rows.ensure(static_cast<uint64_t>(h[0]) * 4 + 4);
        launch_kd_resolve(b.kd.view, cv, prog0, 0, out, b.x.as<double>(), b.y.as<double>(),
                          b.z.as<double>(), h[0], failed, deep_rows.as<uint32_t>(), stream, &tt);
        uint32_t nfail = 0;
        GF_HIP(hipMemcpyAsync(&nfail, failed, 4, hipMemcpyDeviceToHost, stream));
        GF_HIP(hipStreamSynchronize(stream));
        if (nfail)
            resolve_deep(b.kd, T, cv, prog0, 0, out, b.x.as<double>(), b.y.as<double>(),
                         b.z.as<double>(), nfail, failed, &tt);
        launch_features(cv, prog0, out, b.inv.as<uint32_t>(), stream, &tt, kFeaturesFlagged,
                        h[0]);
    }
    e1 = mark();
    if (h[0]) span(e2, e1, &stats.ms_exact);
    span(t_all, e1, &stats.ms_total);

    unsigned long long ctr[4] = {0, 0, 0, 0};
    if (out.counters)
        GF_HIP(hipMemcpyAsync(ctr, out.counters, sizeof(ctr), hipMemcpyDeviceToHost, stream));
    if (where == GF_WHERE_HOST) {
        for (int c = 0; c < ncol_total; ++c)
            GF_HIP(hipMemcpyAsync(cols[c], out.col[c], N * 8, hipMemcpyDeviceToHost, stream));
        for (int i = 0; i < ns; ++i)
            GF_HIP(hipMemcpyAsync(status[i], out.status[i], N, hipMemcpyDeviceToHost, stream));
    }
    GF_HIP(hipStreamSynchronize(stream));
    collect_spans();
    stats.points = N;
    stats.search_passes = ctr[0];
    stats.candidates = ctr[1];
    stats.merges = ctr[2];
}

// The batch step with no host synchronisation before the exact tail: the
// bounds, every tile's grid (grid_params on the device, the same rule and
// factor table as the host's), keys, segmented sort, column starts, search
// and the unflagged rows' features are captured once per shape into two
// graphs and replayed.  Column tables are laid out by each tile's column
// capacity instead of its column count, the sort always takes four 8-bit
// passes, and the flagged rows' re-solve and features read their count on
// the device, so no launch depends on a value the step computes.  The exact
// forest is built on the side stream (its build reads back its own
// counters) while the second graph runs.  Every row's result equals the
// launch-by-launch path's.
bool Context::run_batch_graph(Batch& b, int ns, std::vector<TileInfo>& tiles,
                              const std::vector<uint32_t>& fetch, const FusedProgram& base,
                              int ncol_total, int max_k, bool any_exact_tile,
                              double* const* cols, uint8_t* const* status, int where)
{
    const uint32_t T = b.T;
    const uint64_t N = b.n;
    const uint32_t nb = static_cast<uint32_t>(b.blocks.size());
    // Search shape: the context's, or the autotuned one of the bucket of the
    // largest tile as its last bounds put it.
    KnnTuning kt = knn_tuning();
    uint32_t big = 0;
    for (uint32_t t = 1; t < T; ++t)
        if (b.starts[t + 1] - b.starts[t] > b.starts[big + 1] - b.starts[big]) big = t;
    if (b.bounds.size() == static_cast<size_t>(T) * 6) {
        const uint64_t nbig = b.starts[big + 1] - b.starts[big];
        if (const TuneEntry* e = tune.entry_for(cloud_rho(&b.bounds[6 * big], nbig), fetch[big])) {
            if (e->group_z > 0) kt.group_z = e->group_z - 1;
            if (e->q_per_wave > 0) kt.q_per_wave = e->q_per_wave;
        }
    }
    const int Q = kt.q_per_wave > 0 ? kt.q_per_wave : knn_queries_per_wave(max_k);
    // Host-known tables: column bases by capacity, wave tasks.
    std::vector<uint64_t> colbase(T);
    uint64_t centries = 0;
    for (uint32_t t = 0; t < T; ++t) {
        colbase[t] = centries;
        centries += grid_col_cap(b.starts[t + 1] - b.starts[t]);
    }
    if (centries >= (1ull << 32))
        throw Error(GF_ERR_UNSUPPORTED, "batch column table above 2^32 entries");
    std::vector<uint32_t> task0(static_cast<size_t>(T) + 1, 0u);
    uint64_t ntasks = 0;
    for (uint32_t t = 0; t < T; ++t) {
        tiles[t].colbase = colbase[t];
        task0[t] = static_cast<uint32_t>(ntasks);
        if (!tiles[t].all_exact) ntasks += (tiles[t].n + Q - 1) / Q;
    }
    task0[T] = static_cast<uint32_t>(ntasks);
    if (ntasks > 0xFFFFFF00ull)
        throw Error(GF_ERR_UNSUPPORTED, "batch above the search's task range");
    std::vector<TuneFactor> tf;
    if (!(tune.forced_factor > 0))
        for (const TuneEntry& e : tune.table) tf.push_back(TuneFactor{e.rho_bucket, e.k_bucket, e.cell_factor});
    const double deff = tune.forced_factor > 0 ? tune.forced_factor : tune.cell_factor;

    // Every buffer the graphs touch, sized before capture.
    b.keys_a.ensure(N * 4 + 8);
    b.vals_a.ensure(N * 4 + 8);
    b.keys_b.ensure(N * 4 + 8);
    b.vals_b.ensure(N * 4 + 8);
    b.colstart.ensure(centries * 4 + 8);
    b.sx.ensure(N * 8 + 8);
    b.sy.ensure(N * 8 + 8);
    b.sz.ensure(N * 8 + 8);
    b.inv.ensure(N * 4 + 8);
    b.hist.ensure(static_cast<uint64_t>(nb) * 256 * 4 + 8);
    scan_tmp.ensure(scan_scratch_words(static_cast<uint64_t>(nb) * 256 + 1) * 4);
    xyz_stage.ensure(N * 24 + 8);
    b.benc.ensure(static_cast<uint64_t>(T) * 6 * 8 + 8);
    b.dgrids.ensure(static_cast<uint64_t>(T) * sizeof(TileGrid) + 8);
    b.dbounds.ensure(static_cast<uint64_t>(T) * 6 * 8 + 8);
    b.dfetch.ensure(static_cast<uint64_t>(T) * 4 + 8);
    b.dcolbase.ensure(static_cast<uint64_t>(T) * 8 + 8);
    b.dcolpre.ensure(static_cast<uint64_t>(T) * 8 + 16);
    b.dtune.ensure(std::max<size_t>(tf.size(), 1) * sizeof(TuneFactor) + 8);
    b.dtiles.ensure(static_cast<uint64_t>(T) * sizeof(TileInfo) + 8);
    b.dtasks.ensure(task0.size() * sizeof(uint32_t) + 16);
    KnnOutputs out{};
    if (where == GF_WHERE_DEVICE) {
        for (int c = 0; c < ncol_total; ++c) out.col[c] = cols[c];
        for (int i = 0; i < ns; ++i) out.status[i] = status[i];
    } else {
        outbuf.ensure(static_cast<uint64_t>(ncol_total) * N * 8 + 8);
        statbuf.ensure(static_cast<uint64_t>(ns) * N + 8);
        for (int c = 0; c < ncol_total; ++c)
            out.col[c] = outbuf.as<double>() + static_cast<uint64_t>(c) * N;
        for (int i = 0; i < ns; ++i) out.status[i] = statbuf.as<uint8_t>() + i * N;
    }
    flagged.ensure(N * 4 + 4);
    counters.ensure(128);
    covbuf.ensure(N * 8 * kCovRecord + 8);
    deep_rows.ensure(N * 4 + 4);
    out.flagged = flagged.as<uint32_t>();
    out.nflagged = counters.as<uint32_t>();
    out.cov = covbuf.as<double>();
    if (tune.collect_stats)
        out.counters = reinterpret_cast<unsigned long long*>(counters.as<uint32_t>() + 8);

    TileTables tt{};
    tt.tiles = b.dtiles.as<TileInfo>();
    tt.starts = b.dstarts.as<uint64_t>();
    tt.ntiles = T;
    tt.task0 = b.dtasks.as<uint32_t>();
    tt.ntasks = static_cast<uint32_t>(ntasks);
    tt.max_k = max_k;
    tt.q_per_wave = Q;
    FusedProgram prog0 = tiles[0].prog;

    // What the capture depends on.
    std::vector<uint64_t> key = {T, N, static_cast<uint64_t>(ns),
                                 static_cast<uint64_t>(ncol_total),
                                 static_cast<uint64_t>(where), static_cast<uint64_t>(Q),
                                 static_cast<uint64_t>(kt.group_z), static_cast<uint64_t>(kt.ring0),
                                 static_cast<uint64_t>(kt.ring_max),
                                 static_cast<uint64_t>(kt.small_col), any_exact_tile ? 1u : 0u,
                                 tune.collect_stats ? 1u : 0u, tf.size()};
    auto put = [&](const void* p, size_t bytes) {
        for (size_t o = 0; o < bytes; o += 8) {
            uint64_t v = 0;
            std::memcpy(&v, static_cast<const char*>(p) + o, std::min<size_t>(8, bytes - o));
            key.push_back(v);
        }
    };
    put(&deff, 8);
    put(&tune.cell_z_factor, 8);
    put(&base, sizeof(base));
    put(fetch.data(), fetch.size() * 4);
    put(tf.data(), tf.size() * sizeof(TuneFactor));
    const std::vector<uint64_t> shape = key;  // without the buffers' addresses
    const void* ptrs[] = {b.x.ptr, b.y.ptr, b.z.ptr, b.dblocks.ptr, b.dstarts.ptr, b.benc.ptr,
                          b.dgrids.ptr, b.dbounds.ptr, b.dfetch.ptr, b.dcolbase.ptr, b.dcolpre.ptr,
                          b.dtune.ptr, b.dtiles.ptr, b.dtasks.ptr, b.keys_a.ptr, b.vals_a.ptr,
                          b.keys_b.ptr, b.vals_b.ptr, b.colstart.ptr, b.sx.ptr, b.sy.ptr,
                          b.sz.ptr, b.inv.ptr, b.hist.ptr, scan_tmp.ptr, xyz_stage.ptr,
                          flagged.ptr, counters.ptr, covbuf.ptr};
    put(ptrs, sizeof(ptrs));
    put(out.col, sizeof(out.col));
    put(out.status, sizeof(out.status));

    stats.graph = 1;
    if (key != b.gkey || !b.gexec_a || !b.gexec_b) {
        // A shape is captured the second time it runs (and again when a
        // buffer it uses has moved): one-off shapes (a serving loop's
        // batches) run launch by launch, capture-free.
        if (shape != b.seen_key) {
            b.seen_key = shape;
            stats.graph = 0;
            return false;
        }
        b.drop_graph();
        b.seen_key = shape;
        GF_HIP(hipMemcpyAsync(b.dfetch.ptr, fetch.data(), fetch.size() * 4,
                              hipMemcpyHostToDevice, stream));
        GF_HIP(hipMemcpyAsync(b.dcolbase.ptr, colbase.data(), colbase.size() * 8,
                              hipMemcpyHostToDevice, stream));
        if (!tf.empty())
            GF_HIP(hipMemcpyAsync(b.dtune.ptr, tf.data(), tf.size() * sizeof(TuneFactor),
                                  hipMemcpyHostToDevice, stream));
        GF_HIP(hipMemcpyAsync(b.dtiles.ptr, tiles.data(),
                              static_cast<uint64_t>(T) * sizeof(TileInfo), hipMemcpyHostToDevice,
                              stream));
        GF_HIP(hipMemcpyAsync(b.dtasks.ptr, task0.data(), task0.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, stream));
        GF_HIP(hipStreamSynchronize(stream));
        auto capture = [&](hipGraphExec_t* exec, auto&& body) {
            hipGraph_t g = nullptr;
            GF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            GF_HIP(hipStreamEndCapture(stream, &g));
            const hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            GF_HIP(e);
        };
        capture(&b.gexec_a, [&] {
            launch_bounds_init(b.benc.as<unsigned long long>(), T, stream);
            launch_tile_bounds(b.x.as<double>(), b.y.as<double>(), b.z.as<double>(),
                               b.dblocks.as<SegBlock>(), nb, b.benc.as<unsigned long long>(),
                               stream);
            launch_tile_grids(b.benc.as<unsigned long long>(), b.dstarts.as<uint64_t>(),
                              b.dfetch.as<uint32_t>(), b.dcolbase.as<uint64_t>(), T,
                              b.dtune.as<TuneFactor>(), static_cast<int>(tf.size()), deff,
                              tune.cell_z_factor, b.dgrids.as<TileGrid>(), b.dbounds.as<double>(),
                              b.dtiles.as<TileInfo>(), stream);
        });
        capture(&b.gexec_b, [&] {
            launch_fill_u32(out.nflagged, 4, 0u, stream);
            if (out.counters) launch_fill_u32(counters.as<uint32_t>() + 8, 8, 0u, stream);
            launch_tile_keys(b.x.as<double>(), b.y.as<double>(), b.z.as<double>(),
                             b.dblocks.as<SegBlock>(), nb, b.dgrids.as<TileGrid>(),
                             b.keys_a.as<uint32_t>(), b.vals_a.as<uint32_t>(),
                             xyz_stage.as<double>(), stream);
            uint32_t *kres, *vres;
            launch_seg_radix_sort(b.keys_a.as<uint32_t>(), b.vals_a.as<uint32_t>(),
                                  b.keys_b.as<uint32_t>(), b.vals_b.as<uint32_t>(),
                                  b.dblocks.as<SegBlock>(), nb, 32, b.hist.as<uint32_t>(),
                                  scan_tmp.as<uint32_t>(), &kres, &vres, stream);
            launch_tile_colstart_real(kres, b.dgrids.as<TileGrid>(), T,
                                      b.dcolpre.as<uint64_t>(), b.colstart.as<uint32_t>(),
                                      centries, stream);
            launch_gather_sorted(vres, xyz_stage.as<double>(), N, b.sx.as<double>(),
                                 b.sy.as<double>(), b.sz.as<double>(), b.inv.as<uint32_t>(),
                                 stream);
            b.keys = kres;
            b.perm = vres;
            CloudView cv{};
            cv.sx = b.sx.as<double>();
            cv.sy = b.sy.as<double>();
            cv.sz = b.sz.as<double>();
            cv.perm = b.perm;
            cv.keys = b.keys;
            cv.colstart = b.colstart.as<uint32_t>();
            cv.n = N;
            launch_knn(cv, prog0, 0, out, kt, stream, &tt);
            if (any_exact_tile) launch_flag_tiles(N, cv.perm, 0, out, tt, stream);
            launch_features(cv, prog0, out, b.inv.as<uint32_t>(), stream, &tt,
                            kFeaturesUnflagged);
        });
        b.gkey = key;
        stats.graph = 2;
    }
    b.idx_valid = false;  // the eager path's host grids do not describe this index

    CloudView cv{};
    cv.sx = b.sx.as<double>();
    cv.sy = b.sy.as<double>();
    cv.sz = b.sz.as<double>();
    cv.perm = b.perm;
    cv.keys = b.keys;
    cv.colstart = b.colstart.as<uint32_t>();
    cv.n = N;
    reset_spans();
    hipEvent_t t_all = mark();
    if (where != GF_WHERE_DEVICE)  // rows a stage skips keep their prior values
        for (int c = 0; c < ncol_total; ++c)
            GF_HIP(hipMemcpyAsync(out.col[c], cols[c], N * 8, hipMemcpyHostToDevice, stream));
    GF_HIP(hipGraphLaunch(b.gexec_a, stream));
    const bool need_forest = !b.kd_valid;
    if (need_forest) GF_HIP(hipEventRecord(join, stream));
    GF_HIP(hipGraphLaunch(b.gexec_b, stream));
    if (need_forest) {
        // The forest reads the points and the tiles' bounds the first graph
        // wrote; its build returns when it is done.
        GF_HIP(hipStreamWaitEvent(side, join, 0));
        build_forest(b.kd, b.x.as<double>(), b.y.as<double>(), b.z.as<double>(), N,
                     b.dstarts.as<uint64_t>(), T, b.dbounds.as<double>(), side, true);
        b.kd_valid = true;
    }
    uint32_t* failed = counters.as<uint32_t>() + 5;
    launch_fill_u32(failed, 1, 0u, stream);
    launch_kd_resolve(b.kd.view, cv, prog0, 0, out, b.x.as<double>(), b.y.as<double>(),
                      b.z.as<double>(), static_cast<uint32_t>(std::min<uint64_t>(N, 0xFFFFFFFFu)),
                      failed, deep_rows.as<uint32_t>(), stream, &tt);
    launch_features_flagged(cv, prog0, out, b.inv.as<uint32_t>(), N, stream, &tt);
    span(t_all, mark(), &stats.ms_total);
    uint32_t h[16];
    GF_HIP(hipMemcpyAsync(h, counters.ptr, sizeof(h), hipMemcpyDeviceToHost, stream));
    unsigned long long ctr[4] = {0, 0, 0, 0};
    if (out.counters)
        GF_HIP(hipMemcpyAsync(ctr, out.counters, sizeof(ctr), hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    if (h[5]) {
        // Trees deeper than the register frames: those rows again with deeper
        // frames, then every flagged row's features again.
        resolve_deep(b.kd, T, cv, prog0, 0, out, b.x.as<double>(), b.y.as<double>(),
                     b.z.as<double>(), h[5], failed, &tt);
        launch_features_flagged(cv, prog0, out, b.inv.as<uint32_t>(), N, stream, &tt);
    }
    if (where == GF_WHERE_HOST) {
        for (int c = 0; c < ncol_total; ++c)
            GF_HIP(hipMemcpyAsync(cols[c], out.col[c], N * 8, hipMemcpyDeviceToHost, stream));
        for (int i = 0; i < ns; ++i)
            GF_HIP(hipMemcpyAsync(status[i], out.status[i], N, hipMemcpyDeviceToHost, stream));
    }
    GF_HIP(hipStreamSynchronize(stream));
    collect_spans();
    stats.tie_rows = h[1];
    stats.incomplete_rows = h[2];
    stats.points = N;
    stats.search_passes = ctr[0];
    stats.candidates = ctr[1];
    stats.merges = ctr[2];
    return true;
}

} // namespace gf

// ------------------------------------------------------------------ ABI ----
using namespace gf;

namespace {

template <class F>
int bguarded(gf_ctx* ctx, F&& f)
{
    try {
        if (ctx) reinterpret_cast<Context*>(ctx)->reset_spans();
        f();
        return GF_OK;
    } catch (const gf::Error& e) {
        if (ctx) reinterpret_cast<Context*>(ctx)->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) reinterpret_cast<Context*>(ctx)->last_error = "out of host memory";
        return GF_ERR_OOM;
    } catch (const std::exception& e) {
        if (ctx) reinterpret_cast<Context*>(ctx)->last_error = e.what();
        return GF_ERR_STATE;
    }
}

Batch* B(gf_batch* b) { return reinterpret_cast<Batch*>(b); }

} // namespace

extern "C" {

int gf_batch_create(gf_ctx* ctx, uint32_t ntiles, const uint64_t* counts, gf_batch** out)
{
    if (!ctx || !out || (ntiles && !counts)) return GF_ERR_ARG;
    *out = nullptr;
    return bguarded(ctx, [&] {
        Context* c = reinterpret_cast<Context*>(ctx);
        GF_HIP(hipSetDevice(c->device));
        Batch* b = new Batch(c);
        try {
            b->reset(ntiles, counts, c->stream);
        } catch (...) {
            delete b;
            throw;
        }
        *out = reinterpret_cast<gf_batch*>(b);
    });
}

int gf_batch_set_tile(gf_batch* batch, uint32_t tile, const double* x, const double* y,
                      const double* z, int where)
{
    if (!batch) return GF_ERR_ARG;
    Batch* b = B(batch);
    gf_ctx* ctx = reinterpret_cast<gf_ctx*>(b->ctx);
    if (tile >= b->T) return GF_ERR_ARG;
    const uint64_t nt = b->starts[tile + 1] - b->starts[tile];
    if (nt && (!x || !y || !z)) return GF_ERR_ARG;
    return bguarded(ctx, [&] {
        Context* c = b->ctx;
        const hipMemcpyKind kind =
            where == GF_WHERE_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const uint64_t off = b->starts[tile] * 8;
        if (nt) {
            GF_HIP(hipMemcpyAsync(static_cast<char*>(b->x.ptr) + off, x, nt * 8, kind, c->stream));
            GF_HIP(hipMemcpyAsync(static_cast<char*>(b->y.ptr) + off, y, nt * 8, kind, c->stream));
            GF_HIP(hipMemcpyAsync(static_cast<char*>(b->z.ptr) + off, z, nt * 8, kind, c->stream));
        }
        GF_HIP(hipStreamSynchronize(c->stream));
        b->idx_valid = false;
        b->kd_valid = false;
    });
}

uint64_t gf_batch_size(const gf_batch* batch)
{
    return batch ? reinterpret_cast<const Batch*>(batch)->n : 0;
}

uint64_t gf_batch_tile_start(const gf_batch* batch, uint32_t tile)
{
    if (!batch) return 0;
    const Batch* b = reinterpret_cast<const Batch*>(batch);
    return tile <= b->T ? b->starts[tile] : b->n;
}

int gf_batch_invalidate(gf_batch* batch)
{
    if (!batch) return GF_ERR_ARG;
    B(batch)->idx_valid = false;
    B(batch)->kd_valid = false;
    return GF_OK;
}

int gf_batch_run(gf_ctx* ctx, gf_batch* batch, const gf_stage* stages, int nstages,
                 const int32_t* tile_knn, double* const* cols, uint8_t* const* status,
                 int where)
{
    if (!ctx || !batch || !stages || nstages < 1 || nstages > kMaxStages || !status)
        return GF_ERR_ARG;
    return bguarded(ctx, [&] {
        Context* c = reinterpret_cast<Context*>(ctx);
        Batch* b = B(batch);
        if (b->ctx != c) throw Error(GF_ERR_ARG, "batch belongs to another context");
        c->stats = gf_stats{};
        int total_cols = 0;
        for (int i = 0; i < nstages; ++i) {
            validate_stage(stages[i]);
            if (stages[i].radius > 0.0)
                throw Error(GF_ERR_UNSUPPORTED,
                            "radius neighborhoods run through gf_run, not batches");
            if (stages[i].stage == GF_STAGE_COPLANAR || stages[i].stage == GF_STAGE_RANK)
                throw Error(GF_ERR_UNSUPPORTED,
                            "approximatecoplanar and estimaterank run through gf_run");
            total_cols += stage_ncols(stages[i]);
        }
        if (total_cols && !cols) throw Error(GF_ERR_ARG, "null column array");
        if (b->n == 0) return;
        GF_HIP(hipSetDevice(c->device));
        c->run_batch(*b, stages, nstages, tile_knn, cols, status, where);
    });
}

int gf_batch_release(gf_batch* batch)
{
    if (!batch) return GF_ERR_ARG;
    delete B(batch);
    return GF_OK;
}

} // extern "C"
