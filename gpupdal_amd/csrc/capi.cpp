// capi.cpp -- the C ABI (include/gpupdal_amd.h): contexts, resident clouds,
// index builds and fused stage runs on one HIP stream.
#include "gpupdal_amd.h"

#include "gf_internal.hpp"
#include "gf_kernels.hpp"
#include "gf_runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

using namespace gf;

namespace gf {

// Points per unit of the cloud's xy extent (the density rule's rho).
double cloud_rho(const double b[6], uint64_t n) { return grid_rho(b, n); }

KnnTuning Context::knn_tuning(const Cloud& cl, uint32_t fetch) const
{
    KnnTuning k = knn_tuning();
    if (tune.forced_factor > 0) return k;  // an autotuning run: the settings under test
    if (const TuneEntry* e = tune.entry_for(cloud_rho(cl.bounds, cl.n), fetch)) {
        if (e->group_z > 0) k.group_z = e->group_z - 1;
        if (e->q_per_wave > 0) k.q_per_wave = e->q_per_wave;
        if (e->knn_waves > 0) k.waves = e->knn_waves;
    }
    return k;
}

GridParams make_grid(const double b[6], uint64_t n, uint32_t fetch,
                     const Tuning& t, double cell)
{
    return grid_params(b, n, fetch, t.factor_for(grid_rho(b, n), fetch), t.cell_z_factor, cell);
}

void Context::build_index(Cloud& cl, uint32_t fetch, double cell)
{
    const uint64_t n = cl.n;
    if (cl.idx_valid && cl.idx_fetch == fetch && cl.idx_cell == cell &&
        cl.idx_tune == tune.signature())
        return;
    cl.idx_valid = false;
    hipEvent_t e0 = mark();
    part.ensure(kBoundsMaxParts * 6 * sizeof(double));
    bounds.ensure(6 * sizeof(double));
    launch_bounds(cl.x.as<double>(), cl.y.as<double>(), cl.z.as<double>(), n,
                  part.as<double>(), bounds.as<double>(), stream);
    GF_HIP(hipMemcpyAsync(host_bounds, bounds.ptr, 6 * sizeof(double),
                          hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    GridParams g = make_grid(host_bounds, n, fetch, tune, cell);
    for (int i = 0; i < 6; ++i) cl.bounds[i] = host_bounds[i];
    cl.dbounds.ensure(64);
    GF_HIP(hipMemcpyAsync(cl.dbounds.ptr, bounds.ptr, 6 * sizeof(double),
                          hipMemcpyDeviceToDevice, stream));
    GF_HIP(hipEventRecord(join, stream));
    points_mark = true;

    cl.keys_a.ensure(n * 4);
    cl.vals_a.ensure(n * 4);
    cl.keys_b.ensure(n * 4);
    cl.vals_b.ensure(n * 4);
    cl.colstart.ensure((static_cast<uint64_t>(g.ncols) + 1) * 4);
    cl.sx.ensure(n * 8);
    cl.sy.ensure(n * 8);
    cl.sz.ensure(n * 8);
    hist.ensure(radix_hist_words(n) * 4);
    scan_tmp.ensure(scan_scratch_words(std::max<uint64_t>(radix_hist_words(n), g.ncols)) * 4);

    xyz_stage.ensure(n * 24 + 8);
    launch_keys(cl.x.as<double>(), cl.y.as<double>(), cl.z.as<double>(), n, g,
                cl.keys_a.as<uint32_t>(), cl.vals_a.as<uint32_t>(), xyz_stage.as<double>(),
                stream);
    uint32_t *kres, *vres;
    launch_radix_sort(cl.keys_a.as<uint32_t>(), cl.vals_a.as<uint32_t>(),
                      cl.keys_b.as<uint32_t>(), cl.vals_b.as<uint32_t>(), n,
                      g.key_bits, hist.as<uint32_t>(), scan_tmp.as<uint32_t>(), &kres,
                      &vres, stream);
    launch_colstart(kres, n, g, cl.colstart.as<uint32_t>(), stream);
    cl.inv.ensure(n * 4 + 8);
    launch_gather_sorted(vres, xyz_stage.as<double>(), n, cl.sx.as<double>(),
                         cl.sy.as<double>(), cl.sz.as<double>(), cl.inv.as<uint32_t>(), stream);
    cl.keys = kres;
    cl.perm = vres;
    cl.g = g;
    cl.idx_valid = true;
    cl.idx_fetch = fetch;
    cl.idx_cell = cell;
    cl.idx_tune = tune.signature();
    span(e0, mark(), &stats.ms_index);
    stats.cell = g.c;
    stats.cell_z = g.cz;
    stats.grid_x = g.gx;
    stats.grid_y = g.gy;
    stats.grid_z = g.gz;
    stats.key_bits = g.key_bits;
}

CloudView Context::view(const Cloud& cl) const
{
    CloudView v;
    v.sx = cl.sx.as<double>();
    v.sy = cl.sy.as<double>();
    v.sz = cl.sz.as<double>();
    v.perm = cl.perm;
    v.keys = cl.keys;
    v.colstart = cl.colstart.as<uint32_t>();
    v.n = cl.n;
    v.g = cl.g;
    return v;
}

int stage_ncols(const gf_stage& s)
{
    switch (s.stage) {
    case GF_STAGE_NORMAL: return 4;
    case GF_STAGE_EIGENVALUES: return 3;
    case GF_STAGE_COPLANAR: return 1;
    case GF_STAGE_RANK: return 1;
    case GF_STAGE_COVFEAT: return __builtin_popcount(s.features & 0x3FFu);
    default: return -1;
    }
}

int fill_stage_programs(FusedProgram& prog, const gf_stage* st, int ns)
{
    int ncol_total = 0;
    prog.nstages = ns;
    for (int i = 0; i < ns; ++i) {
        StageProgram& sp = prog.st[i];
        sp.stage = st[i].stage;
        sp.mode = st[i].mode;
        sp.features = st[i].features;
        sp.always_up = st[i].always_up;
        sp.normalize = st[i].normalize;
        sp.p1 = st[i].p1;
        sp.p2 = st[i].p2;
        sp.ncols = stage_ncols(st[i]);
        sp.col_offset = ncol_total;
        ncol_total += sp.ncols;
    }
    if (ncol_total > kMaxOutColumns)
        throw Error(GF_ERR_UNSUPPORTED,
                    "more than " + std::to_string(kMaxOutColumns) + " output columns");
    return ncol_total;
}

void validate_stage(const gf_stage& s)
{
    if (stage_ncols(s) < 0) throw Error(GF_ERR_ARG, "unknown stage kind");
    if (!(s.radius >= 0.0) || s.radius == INFINITY)
        throw Error(GF_ERR_ARG, "radius must be a finite number, 0 for knn neighborhoods");
    if (s.radius > 0.0) return;  // knn and stride are not used
    if (s.knn < 1) throw Error(GF_ERR_ARG, "knn must be at least 1");
    if (s.stage != GF_STAGE_NORMAL && s.stride < 1)
        throw Error(GF_ERR_ARG, "stride must be at least 1");
    if (s.stage == GF_STAGE_COVFEAT) {
        if (s.mode < GF_MODE_RAW || s.mode > GF_MODE_NORMALIZED)
            throw Error(GF_ERR_ARG, "unknown eigenvalue mode");
        if ((s.features & ~0x3FFu) != 0)
            throw Error(GF_ERR_ARG, "unknown feature bits");
    }
}

int64_t stage_neighbors(const gf_stage& s)
{
    // approximatecoplanar and estimaterank call KD3Index::neighbors(p, knn)
    // (ApproximateCoplanarFilter.cpp:96, EstimateRankFilter.cpp:80); the other
    // stages ask knn + 1, the point itself included.
    return s.stage == GF_STAGE_COPLANAR || s.stage == GF_STAGE_RANK
               ? static_cast<int64_t>(s.knn)
               : static_cast<int64_t>(s.knn) + 1;
}

void check_neighborhood(uint64_t n, int64_t K, int stride, uint32_t limit)
{
    const uint64_t k = std::min<uint64_t>(n, static_cast<uint64_t>(K));
    if (static_cast<uint64_t>(stride) * k > limit)
        throw Error(GF_ERR_UNSUPPORTED, "(knn+1)*stride above " + std::to_string(limit) +
                                            " is outside the device kNN envelope");
}

// Runs one group of stages that share the neighborhood (knn, stride).
void Context::run_group(Cloud& cl, const gf_stage* st, int ns, double* const* cols,
                        uint8_t* const* status, int where)
{
    if (st[0].radius > 0.0) {
        run_group_radius(cl, st, ns, cols, status, where);
        return;
    }
    const uint64_t n = cl.n;
    const int stride = st[0].stage == GF_STAGE_NORMAL ? 1 : st[0].stride;
    const uint64_t K = static_cast<uint64_t>(stage_neighbors(st[0]));
    const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(n, K));
    const uint32_t fetch = static_cast<uint32_t>(stride) * k;
    // Neighborhoods of up to 63 entries (knn + 1) * stride run the grid
    // search; larger ones (up to the exact search's capacity) run every row
    // through the exact search.
    check_neighborhood(n, stage_neighbors(st[0]), stride, kd_max_capacity());
    FusedProgram prog{};
    prog.nstages = ns;
    prog.k = static_cast<int32_t>(k);
    prog.stride = stride;
    prog.fetch = static_cast<int32_t>(fetch);
    const int ncol_total = fill_stage_programs(prog, st, ns);
    const bool all_exact = fetch > n || fetch > static_cast<uint32_t>(kMaxFetch - 1);

    reset_spans();
    hipEvent_t t_all = mark();
    points_mark = false;
    build_index(cl, fetch);

    KnnOutputs out{};
    if (where == GF_WHERE_DEVICE) {
        for (int c = 0; c < ncol_total; ++c) out.col[c] = cols[c];
        for (int i = 0; i < ns; ++i) out.status[i] = status[i];
    } else {
        outbuf.ensure(static_cast<uint64_t>(ncol_total) * n * 8 + 8);
        statbuf.ensure(static_cast<uint64_t>(ns) * n + 8);
        for (int c = 0; c < ncol_total; ++c) {
            out.col[c] = outbuf.as<double>() + static_cast<uint64_t>(c) * n;
            // Rows a stage skips keep their prior values: start from them.
            GF_HIP(hipMemcpyAsync(out.col[c], cols[c], n * 8, hipMemcpyHostToDevice,
                                  stream));
        }
        for (int i = 0; i < ns; ++i) out.status[i] = statbuf.as<uint8_t>() + i * n;
    }
    flagged.ensure(n * 4 + 4);
    counters.ensure(64);
    covbuf.ensure(n * 8 * kCovRecord + 8);
    out.flagged = flagged.as<uint32_t>();
    out.nflagged = counters.as<uint32_t>();
    out.cov = covbuf.as<double>();
    launch_fill_u32(out.nflagged, 4, 0u, stream);
    if (tune.collect_stats) {
        out.counters = reinterpret_cast<unsigned long long*>(counters.as<uint32_t>() + 8);
        launch_fill_u32(counters.as<uint32_t>() + 8, 8, 0u, stream);
    }

    const CloudView cv = view(cl);
    // The exact tree depends only on the points: on large clouds, which
    // nearly always have a flagged row, it is built on the side stream while
    // the search runs.
    const bool early = !all_exact && !cl.kd_valid && tune.kd_overlap_min &&
                       n >= tune.kd_overlap_min;
    if (early) fork_side();
    hipEvent_t e0 = mark();
    if (all_exact)  // fewer points than the result capacity, or a large one
        launch_flag_all(n, cv.perm, 0, out, stream);
    else
        launch_knn(cv, prog, 0, out, knn_tuning(cl, fetch), stream);
    hipEvent_t e1 = mark();
    span(e0, e1, &stats.ms_search);
    // Rows the search did not flag are final: their features run now, beside
    // the rest of the exact tree; the flagged rows' after their re-solve.
    launch_features(cv, prog, out, cl.inv.as<uint32_t>(), stream, nullptr,
                    kFeaturesUnflagged);
    hipEvent_t e2 = mark();
    span(e1, e2, &stats.ms_features);
    if (early) build_kdtree(cl, side, true);
    const uint32_t nflagged = resolve_flagged(cl, cv, prog, 0, out, e2);
    launch_features(cv, prog, out, cl.inv.as<uint32_t>(), stream, nullptr, kFeaturesFlagged,
                    nflagged);
    e1 = mark();
    if (nflagged) span(e2, e1, &stats.ms_exact);
    span(t_all, e1, &stats.ms_total);

    unsigned long long ctr[4] = {0, 0, 0, 0};
    if (out.counters)
        GF_HIP(hipMemcpyAsync(ctr, out.counters, sizeof(ctr), hipMemcpyDeviceToHost, stream));
    if (where == GF_WHERE_HOST) {
        for (int c = 0; c < ncol_total; ++c)
            GF_HIP(hipMemcpyAsync(cols[c], out.col[c], n * 8, hipMemcpyDeviceToHost,
                                  stream));
        for (int i = 0; i < ns; ++i)
            GF_HIP(hipMemcpyAsync(status[i], out.status[i], n, hipMemcpyDeviceToHost,
                                  stream));
    }
    GF_HIP(hipStreamSynchronize(stream));
    collect_spans();
    stats.points = n;
    stats.search_passes = ctr[0];
    stats.candidates = ctr[1];
    stats.merges = ctr[2];
}

// The exact KD3Index trees of T tiles (points [starts[r], starts[r+1]) of
// x, y, z; bounds 6 per tile), tile r rooted at node r.
void Context::build_forest(KdBuffers& b, const double* x, const double* y, const double* z,
                           uint64_t n, const uint64_t* dstarts, uint32_t T,
                           const double* dbounds, hipStream_t s, bool beside)
{
    hipEvent_t t0 = mark(s);
    uint32_t cap = std::max(kd_node_capacity(n) + T, b.cap);
    uint32_t lcap = std::max(kd_list_capacity(n) + T, b.list_cap);
    for (int attempt = 0;; ++attempt) {
        b.vind.ensure(n * 4 + 8);
        b.px.ensure(n * 8 + 8);
        b.py.ensure(n * 8 + 8);
        b.pz.ensure(n * 8 + 8);
        b.nodes_u32.ensure(static_cast<uint64_t>(cap) * 4 * 4);
        b.nodes_i32.ensure(static_cast<uint64_t>(cap) * 4 * 4);
        b.which.ensure(cap + 8);
        b.nodes_f64.ensure(static_cast<uint64_t>(cap) * 8 * 9);
        b.ext.ensure(static_cast<uint64_t>(cap) * 8 * 6);
        kd_s1.ensure((n + 1) * 4);
        kd_s2.ensure((n + 1) * 4);
        kd_tmp.ensure((n + 1) * 4);
        kd_lists.ensure((static_cast<uint64_t>(lcap) * 5 + 2 + kKdHeaderWords) * 4);
        kd_sums.ensure(kd_sums_words(n) * 4);
        KdView t{};
        t.vind = b.vind.as<uint32_t>();
        t.px = b.px.as<double>();
        t.py = b.py.as<double>();
        t.pz = b.pz.as<double>();
        t.nl = b.nodes_u32.as<uint32_t>();
        t.nr = t.nl + cap;
        t.split = t.nr + cap;
        t.aux = t.split + cap;
        t.c1 = b.nodes_i32.as<int32_t>();
        t.c2 = t.c1 + cap;
        t.feat = t.c2 + cap;
        t.parent = t.feat + cap;
        t.which = b.which.as<uint8_t>();
        t.divlow = b.nodes_f64.as<double>();
        t.divhigh = t.divlow + cap;
        t.cutval = t.divhigh + cap;
        t.box = t.cutval + cap;
        t.mn = b.ext.as<unsigned long long>();
        t.mx = t.mn + 3 * static_cast<uint64_t>(cap);
        t.cap = cap;
        KdScratch w{};
        w.s1 = kd_s1.as<uint32_t>();
        w.s2 = kd_s2.as<uint32_t>();
        w.tmp = kd_tmp.as<uint32_t>();
        w.sums = kd_sums.as<uint32_t>();
        uint32_t* L = kd_lists.as<uint32_t>();
        w.hdr = L;
        w.act[0] = L + kKdHeaderWords;
        w.act[1] = w.act[0] + lcap;
        w.off[0] = w.act[1] + lcap;
        w.off[1] = w.off[0] + lcap + 1;
        w.small = w.off[1] + lcap + 1;
        w.list_cap = lcap;
        uint32_t nodes = 0;
        KdTuning kt{};
        kt.subtree_points = static_cast<uint32_t>(tune.kd_subtree);
        kt.wave_node = static_cast<uint32_t>(tune.kd_wave_node);
        kt.small_blocks = tune.kd_small_blocks < 0 ? (beside ? 1u : 0u)
                                                   : static_cast<uint32_t>(tune.kd_small_blocks);
        kt.minmax_lds_nodes = static_cast<uint32_t>(tune.kd_minmax_lds);
        const uint32_t rc =
            launch_kd_build(x, y, z, n, dstarts, T, dbounds, t, w, kt, &nodes, s);
        if (rc == 0) {
            b.view = t;
            b.cap = cap;
            b.list_cap = lcap;
            b.nodes = nodes;
            span(t0, mark(s), &stats.ms_tree);
            return;
        }
        if (rc == 3 || attempt >= 8)
            throw Error(GF_ERR_UNSUPPORTED, "exact KD3 build: table overflow (code " +
                                                std::to_string(rc) + ")");
        if (rc == 1) cap *= 2;  // node capacity: grow and build again
        if (rc == 2) lcap *= 2; // node lists
    }
}

// The exact KD3Index tree of a cloud (built once, reused until
// invalidated), complete on return.
void Context::build_kdtree(Cloud& cl, hipStream_t s, bool beside)
{
    if (cl.kd_valid) return;
    build_forest(cl.kd, cl.x.as<double>(), cl.y.as<double>(), cl.z.as<double>(), cl.n,
                 cl.dstarts.as<uint64_t>(), 1, cl.dbounds.as<double>(), s, beside);
    cl.kd_valid = true;
}

void Context::fork_side()
{
    if (!points_mark) GF_HIP(hipEventRecord(join, stream));
    GF_HIP(hipStreamWaitEvent(side, join, 0));
    points_mark = false;
}

// Re-solves the rows the grid search flagged (boundary ties, uncertified
// searches) with the exact KD3Index search.  ms_exact runs from `from` to the
// end of the re-solve.
uint32_t Context::resolve_flagged(Cloud& cl, const CloudView& cv, const FusedProgram& prog,
                                  int mode, const KnnOutputs& out, hipEvent_t from)
{
    uint32_t h[4] = {0, 0, 0, 0};
    GF_HIP(hipMemcpyAsync(h, out.nflagged, sizeof(h), hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    stats.tie_rows = h[1];
    stats.incomplete_rows = h[2];
    if (!h[0]) return 0;
    hipEvent_t e0 = from ? from : mark();
    build_kdtree(cl);
    uint32_t* failed = counters.as<uint32_t>() + 5;
    launch_fill_u32(failed, 1, 0u, stream);
    deep_rows.ensure(static_cast<uint64_t>(h[0]) * 4 + 4);
    launch_kd_resolve(cl.kd.view, cv, prog, mode, out, cl.x.as<double>(), cl.y.as<double>(),
                      cl.z.as<double>(), h[0], failed, deep_rows.as<uint32_t>(), stream);
    uint32_t nfail = 0;
    GF_HIP(hipMemcpyAsync(&nfail, failed, 4, hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    if (nfail)
        resolve_deep(cl.kd, 1, cv, prog, mode, out, cl.x.as<double>(), cl.y.as<double>(),
                     cl.z.as<double>(), nfail, failed, nullptr);
    span(e0, mark(), &stats.ms_exact);
    return h[0];
}

// Rows whose exact search outgrew the 64 register frames (a tree deeper
// than 64 levels, as very unevenly spread points build): again, with frames
// for the tree's whole depth in device memory.  The depth comes from the
// tree's child arrays; at most kDeepFrameBudget bytes of frames are held.
void Context::resolve_deep(KdBuffers& kd, uint32_t nroots, const CloudView& cv,
                           const FusedProgram& prog, int mode, const KnnOutputs& out,
                           const double* x, const double* y, const double* z, uint32_t nfail,
                           const uint32_t* dcount, const TileTables* tt)
{
    constexpr uint64_t kDeepFrameBudget = 1ull << 30;
    const uint32_t nodes = kd.nodes;
    std::vector<int32_t> c1(nodes), c2(nodes);
    GF_HIP(hipMemcpyAsync(c1.data(), kd.view.c1, nodes * 4ull, hipMemcpyDeviceToHost, stream));
    GF_HIP(hipMemcpyAsync(c2.data(), kd.view.c2, nodes * 4ull, hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    const int depth = kd_tree_depth(c1, c2, nroots) + 1;
    const uint64_t per_wave = static_cast<uint64_t>(depth) * kd_deep_frame_bytes();
    if (per_wave > kDeepFrameBudget)
        throw Error(GF_ERR_UNSUPPORTED, "exact KD3 tree of " + std::to_string(depth) +
                                            " levels is deeper than the re-solve supports");
    const uint32_t waves = static_cast<uint32_t>(
        std::min<uint64_t>({static_cast<uint64_t>(nfail), kDeepFrameBudget / per_wave, 1024}));
    deep_frames.ensure(waves * per_wave + 8);
    counters.ensure(128);
    uint32_t* failed2 = counters.as<uint32_t>() + 20;
    launch_fill_u32(failed2, 1, 0u, stream);
    launch_kd_resolve_deep(kd.view, cv, prog, mode, out, x, y, z, deep_rows.as<uint32_t>(), dcount,
                           deep_frames.ptr, depth, waves, failed2, stream, tt);
    uint32_t nfail2 = 0;
    GF_HIP(hipMemcpyAsync(&nfail2, failed2, 4, hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    stats.deep_rows = nfail;
    if (nfail2)
        throw Error(GF_ERR_UNSUPPORTED, "exact KD3 search exceeded the tree's depth on " +
                                            std::to_string(nfail2) + " rows");
}

hipEvent_t Context::mark(hipStream_t s)
{
    if (!timing) return nullptr;
    if (evnext == evpool.size()) {
        hipEvent_t e;
        GF_HIP(hipEventCreate(&e));
        evpool.push_back(e);
    }
    hipEvent_t e = evpool[evnext++];
    GF_HIP(hipEventRecord(e, s));
    return e;
}

void Context::span(hipEvent_t a, hipEvent_t b, double* dst)
{
    if (a && b) spans.push_back(Span{a, b, dst});
}

void Context::collect_spans()
{
    for (const Span& sp : spans) {
        float ms = 0.f;
        GF_HIP(hipEventSynchronize(sp.b));
        GF_HIP(hipEventElapsedTime(&ms, sp.a, sp.b));
        *sp.dst = ms;
    }
    reset_spans();
}

void Context::reset_spans()
{
    spans.clear();
    evnext = 0;
}

// Stages with a radius neighborhood (their `radius` option): one radius
// search writes every row's covariance record and neighbor count, then the
// stage columns.  Every row is exact; there is nothing to re-solve.
void Context::run_group_radius(Cloud& cl, const gf_stage* st, int ns, double* const* cols,
                               uint8_t* const* status, int where)
{
    const uint64_t n = cl.n;
    const double r = st[0].radius;
    FusedProgram prog{};
    prog.nstages = ns;
    prog.k = 0;
    prog.stride = 1;
    prog.fetch = 0;
    prog.radius = r;
    prog.min_k = st[0].min_k;
    const int ncol_total = fill_stage_programs(prog, st, ns);
    reset_spans();
    hipEvent_t t_all = mark();
    points_mark = false;
    build_index(cl, 0, r * tune.radius_cell);
    KnnOutputs out{};
    if (where == GF_WHERE_DEVICE) {
        for (int c = 0; c < ncol_total; ++c) out.col[c] = cols[c];
        for (int i = 0; i < ns; ++i) out.status[i] = status[i];
    } else {
        outbuf.ensure(static_cast<uint64_t>(ncol_total) * n * 8 + 8);
        statbuf.ensure(static_cast<uint64_t>(ns) * n + 8);
        for (int c = 0; c < ncol_total; ++c) {
            out.col[c] = outbuf.as<double>() + static_cast<uint64_t>(c) * n;
            GF_HIP(hipMemcpyAsync(out.col[c], cols[c], n * 8, hipMemcpyHostToDevice, stream));
        }
        for (int i = 0; i < ns; ++i) out.status[i] = statbuf.as<uint8_t>() + i * n;
    }
    covbuf.ensure(n * 8 * kCovRecord + 8);
    out.cov = covbuf.as<double>();
    const CloudView cv = view(cl);
    hipEvent_t e0 = mark();
    launch_radius(cv, r, 0, out.cov, nullptr, tune.group_z, stream);
    hipEvent_t e1 = mark();
    span(e0, e1, &stats.ms_search);
    launch_features(cv, prog, out, cl.inv.as<uint32_t>(), stream, nullptr, kFeaturesAll);
    hipEvent_t e2 = mark();
    span(e1, e2, &stats.ms_features);
    span(t_all, e2, &stats.ms_total);
    if (where == GF_WHERE_HOST) {
        for (int c = 0; c < ncol_total; ++c)
            GF_HIP(hipMemcpyAsync(cols[c], out.col[c], n * 8, hipMemcpyDeviceToHost, stream));
        for (int i = 0; i < ns; ++i)
            GF_HIP(hipMemcpyAsync(status[i], out.status[i], n, hipMemcpyDeviceToHost, stream));
    }
    GF_HIP(hipStreamSynchronize(stream));
    collect_spans();
    stats.points = n;
}

void Context::knn_lists(Cloud& cl, uint32_t k, uint32_t stride, int mode, uint32_t* ids,
                        uint8_t* status)
{
    const uint64_t n = cl.n;
    const uint32_t kk = static_cast<uint32_t>(std::min<uint64_t>(n, k));
    const uint64_t fetch64 = static_cast<uint64_t>(kk) * stride;
    if (fetch64 > kd_max_capacity())
        throw Error(GF_ERR_UNSUPPORTED, "k*stride above " + std::to_string(kd_max_capacity()) +
                                            " is outside the device kNN envelope");
    const uint32_t fetch = static_cast<uint32_t>(fetch64);
    if (mode == 2 && stride != 1) throw Error(GF_ERR_ARG, "sorted-position lists need stride 1");
    const bool all_exact = fetch > n || fetch > static_cast<uint32_t>(kMaxFetch - 1);
    points_mark = false;
    build_index(cl, fetch);
    FusedProgram prog{};
    prog.k = static_cast<int32_t>(kk);
    prog.stride = static_cast<int32_t>(stride);
    prog.fetch = static_cast<int32_t>(fetch);
    KnnOutputs out{};
    out.ids = ids;
    out.status[0] = status;
    flagged.ensure(n * 4 + 4);
    counters.ensure(128);
    out.flagged = flagged.as<uint32_t>();
    out.nflagged = counters.as<uint32_t>();
    launch_fill_u32(out.nflagged, 4, 0u, stream);
    const CloudView cv = view(cl);
    const bool early = !all_exact && !cl.kd_valid && tune.kd_overlap_min &&
                       n >= tune.kd_overlap_min;
    if (early) fork_side();
    if (all_exact)
        launch_flag_all(n, cv.perm, mode, out, stream);
    else
        launch_knn(cv, prog, mode, out, knn_tuning(cl, fetch), stream);
    if (early) build_kdtree(cl, side, true);
    const uint32_t nflagged = resolve_flagged(cl, cv, prog, mode, out, nullptr);
    if (mode == 2)
        launch_lists_to_sorted(ids, static_cast<int>(kk), out.flagged, nflagged, cv.perm,
                               cl.inv.as<uint32_t>(), stream);
}

// filters.nndistance (filters/NNDistanceFilter.cpp:103-133): the distance to
// the k-th neighbor, or the mean distance to the k neighbors, from each
// point's k + 1 nearest (itself first) in ascending distance order.
void Context::nn_distance(Cloud& cl, uint32_t k, int mode, double* out)
{
    const uint64_t n = cl.n;
    const uint32_t K = k + 1;
    const uint32_t kk = static_cast<uint32_t>(std::min<uint64_t>(n, K));
    listbuf.ensure(n * kk * 4 + 8);
    liststat.ensure(n + 8);
    knn_lists(cl, K, 1, 2, listbuf.as<uint32_t>(), liststat.as<uint8_t>());
    const CloudView cv = view(cl);
    launch_nn_distance(listbuf.as<uint32_t>(), static_cast<int>(kk), static_cast<int>(K), n,
                       mode, cl.x.as<double>(), cl.y.as<double>(), cl.z.as<double>(), cv.sx,
                       cv.sy, cv.sz, out, stream);
}

// Stages whose neighborhood size is each row's own (row_k, device array):
// one KD3Index-ordered list of the largest size, each row's prefix.
void Context::run_rowk(Cloud& cl, const gf_stage* st, int ns, const uint32_t* row_k,
                       double* const* cols, uint8_t* const* status, int where)
{
    const uint64_t n = cl.n;
    FusedProgram prog{};
    const int ncol_total = fill_stage_programs(prog, st, ns);
    counters.ensure(128);
    uint32_t* dmax = counters.as<uint32_t>() + 16;
    launch_max_u32(row_k, n, dmax, stream);
    uint32_t K = 0;
    GF_HIP(hipMemcpyAsync(&K, dmax, 4, hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    const uint32_t kk = static_cast<uint32_t>(std::min<uint64_t>(n, K));
    if (kk > kd_max_capacity())
        throw Error(GF_ERR_UNSUPPORTED, "a per-row neighbor count above " +
                                            std::to_string(kd_max_capacity()) +
                                            " is outside the device kNN envelope");
    prog.k = static_cast<int32_t>(kk);
    prog.stride = 1;
    prog.fetch = static_cast<int32_t>(kk);
    listbuf.ensure(n * std::max<uint32_t>(kk, 1) * 4 + 8);
    liststat.ensure(n + 8);
    if (kk) knn_lists(cl, kk, 1, 2, listbuf.as<uint32_t>(), liststat.as<uint8_t>());
    else {
        points_mark = false;
        build_index(cl, 1);  // sorted coordinates only: no row reads a neighbor
    }
    KnnOutputs out{};
    if (where == GF_WHERE_DEVICE) {
        for (int c = 0; c < ncol_total; ++c) out.col[c] = cols[c];
        for (int i = 0; i < ns; ++i) out.status[i] = status[i];
    } else {
        outbuf.ensure(static_cast<uint64_t>(ncol_total) * n * 8 + 8);
        statbuf.ensure(static_cast<uint64_t>(ns) * n + 8);
        for (int c = 0; c < ncol_total; ++c) {
            out.col[c] = outbuf.as<double>() + static_cast<uint64_t>(c) * n;
            GF_HIP(hipMemcpyAsync(out.col[c], cols[c], n * 8, hipMemcpyHostToDevice, stream));
        }
        for (int i = 0; i < ns; ++i) out.status[i] = statbuf.as<uint8_t>() + i * n;
    }
    const CloudView cv = view(cl);
    launch_prefix_stages(listbuf.as<uint32_t>(), static_cast<int>(kk), row_k, n,
                         cl.x.as<double>(), cl.y.as<double>(), cl.z.as<double>(), cv.sx, cv.sy,
                         cv.sz, prog, out, stream);
    if (where != GF_WHERE_DEVICE) {
        for (int c = 0; c < ncol_total; ++c)
            GF_HIP(hipMemcpyAsync(cols[c], out.col[c], n * 8, hipMemcpyDeviceToHost, stream));
        for (int i = 0; i < ns; ++i)
            GF_HIP(hipMemcpyAsync(status[i], out.status[i], n, hipMemcpyDeviceToHost, stream));
    }
    GF_HIP(hipStreamSynchronize(stream));
}

// ------------------------------------------------- no-allocation path ----
//
// gf_run_noalloc runs the same fused pass as gf_run on caller-owned device
// arrays, with every working buffer (grid index, covariance records, flagged
// rows, exact tree) carved from one fixed array per GPU, `scratch`, so that it
// calls no hipMalloc, hipFree or other allocating routine.  Each GPU holds its
// own copy of the array until the process ends; its address on a device is
// taken once, with that device current, when the first context on it is
// created.  Calls that overlap on one GPU share the array and race.

constexpr uint64_t kStaticScratchBytes = 24ull << 20;
constexpr uint64_t kStaticMaxPoints = GF_NOALLOC_MAX_POINTS;
constexpr int kStaticMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kStaticMaxDevices] = {};
}  // namespace

// The device address of `scratch` on `dev` (the current device), looked up
// once per device.
void* bind_static_scratch(int dev)
{
    if (dev < 0 || dev >= kStaticMaxDevices)
        throw Error(GF_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(-2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

namespace {

// Consecutive 256-byte-aligned pieces of the scratch array.
struct Arena {
    unsigned char* base;
    uint64_t off = 0;
    uint64_t cap;
    template <class T> T* take(uint64_t bytes)
    {
        off = (off + 255) & ~uint64_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += bytes;
        if (off > cap)
            throw Error(GF_ERR_ARG, "gf_run_noalloc: the working set of " + std::to_string(off) +
                                        " bytes exceeds the fixed scratch of " +
                                        std::to_string(cap) + " bytes");
        return p;
    }
};

}  // namespace

void Context::run_group_static(const double* x, const double* y, const double* z, uint64_t n,
                               const gf_stage* st, int ns, double* const* cols,
                               uint8_t* const* status)
{
    if (n > kStaticMaxPoints)
        throw Error(GF_ERR_ARG, "gf_run_noalloc takes at most " + std::to_string(kStaticMaxPoints) +
                                    " points (the fixed scratch size)");
    if (st[0].radius > 0.0)
        throw Error(GF_ERR_UNSUPPORTED, "gf_run_noalloc takes knn neighborhoods; radius "
                                        "neighborhoods run through gf_run");
    if (!static_scratch) throw Error(GF_ERR_STATE, "no scratch array bound on this device");
    const int stride = st[0].stage == GF_STAGE_NORMAL ? 1 : st[0].stride;
    const uint64_t K = static_cast<uint64_t>(stage_neighbors(st[0]));
    const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(n, K));
    const uint32_t fetch = static_cast<uint32_t>(stride) * k;
    check_neighborhood(n, stage_neighbors(st[0]), stride, kMaxFetch - 1);
    FusedProgram prog{};
    prog.nstages = ns;
    prog.k = static_cast<int32_t>(k);
    prog.stride = stride;
    prog.fetch = static_cast<int32_t>(fetch);
    const int ncol_total = fill_stage_programs(prog, st, ns);

    // Every piece is laid out before the first launch, so a call that does
    // not fit fails before touching the device.
    Arena ar{static_cast<unsigned char*>(static_scratch), 0, kStaticScratchBytes};
    double* part = ar.take<double>(kBoundsMaxParts * 6 * sizeof(double));
    double* dbounds = ar.take<double>(6 * sizeof(double));
    uint64_t* dstarts = ar.take<uint64_t>(2 * sizeof(uint64_t));
    uint32_t* keys_a = ar.take<uint32_t>(n * 4 + 4);
    uint32_t* vals_a = ar.take<uint32_t>(n * 4 + 4);
    uint32_t* keys_b = ar.take<uint32_t>(n * 4 + 4);
    uint32_t* vals_b = ar.take<uint32_t>(n * 4 + 4);
    const uint64_t col_cap = std::max<uint64_t>(4 * n, 4096);  // make_grid's bound on columns
    uint32_t* colstart = ar.take<uint32_t>((col_cap + 1) * 4);
    double* sx = ar.take<double>(n * 8 + 8);
    double* sy = ar.take<double>(n * 8 + 8);
    double* sz = ar.take<double>(n * 8 + 8);
    uint32_t* inv = ar.take<uint32_t>(n * 4 + 8);
    double* xyz = ar.take<double>(n * 24 + 8);
    uint32_t* hist = ar.take<uint32_t>(radix_hist_words(n) * 4);
    uint32_t* scan = ar.take<uint32_t>(
        scan_scratch_words(std::max<uint64_t>(radix_hist_words(n), col_cap)) * 4);
    uint32_t* flg = ar.take<uint32_t>(n * 4 + 4);
    uint32_t* cnt = ar.take<uint32_t>(64);
    double* cov = ar.take<double>(n * 8 * kCovRecord + 8);
    const uint32_t cap = kd_node_capacity(n) + 1;
    const uint32_t lcap = kd_list_capacity(n) + 1;
    KdView t{};
    t.vind = ar.take<uint32_t>(n * 4 + 8);
    t.px = ar.take<double>(n * 8 + 8);
    t.py = ar.take<double>(n * 8 + 8);
    t.pz = ar.take<double>(n * 8 + 8);
    t.nl = ar.take<uint32_t>(static_cast<uint64_t>(cap) * 16);
    t.nr = t.nl + cap;
    t.split = t.nr + cap;
    t.aux = t.split + cap;
    t.c1 = ar.take<int32_t>(static_cast<uint64_t>(cap) * 16);
    t.c2 = t.c1 + cap;
    t.feat = t.c2 + cap;
    t.parent = t.feat + cap;
    t.which = ar.take<uint8_t>(cap + 8);
    t.divlow = ar.take<double>(static_cast<uint64_t>(cap) * 8 * 9);
    t.divhigh = t.divlow + cap;
    t.cutval = t.divhigh + cap;
    t.box = t.cutval + cap;
    t.mn = ar.take<unsigned long long>(static_cast<uint64_t>(cap) * 8 * 6);
    t.mx = t.mn + 3 * static_cast<uint64_t>(cap);
    t.cap = cap;
    KdScratch w{};
    w.s1 = ar.take<uint32_t>((n + 1) * 4);
    w.s2 = ar.take<uint32_t>((n + 1) * 4);
    w.tmp = ar.take<uint32_t>((n + 1) * 4);
    w.sums = ar.take<uint32_t>(kd_sums_words(n) * 4);
    uint32_t* L = ar.take<uint32_t>((static_cast<uint64_t>(lcap) * 5 + 2 + kKdHeaderWords) * 4);
    w.hdr = L;
    w.act[0] = L + kKdHeaderWords;
    w.act[1] = w.act[0] + lcap;
    w.off[0] = w.act[1] + lcap;
    w.off[1] = w.off[0] + lcap + 1;
    w.small = w.off[1] + lcap + 1;
    w.list_cap = lcap;

    // Grid index (the same kernels as the resident-cloud path).
    launch_bounds(x, y, z, n, part, dbounds, stream);
    GF_HIP(hipMemcpyAsync(host_bounds, dbounds, 6 * sizeof(double), hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    const GridParams g = make_grid(host_bounds, n, fetch, tune);
    launch_keys(x, y, z, n, g, keys_a, vals_a, xyz, stream);
    uint32_t *kres, *vres;
    launch_radix_sort(keys_a, vals_a, keys_b, vals_b, n, g.key_bits, hist, scan, &kres, &vres,
                      stream);
    launch_colstart(kres, n, g, colstart, stream);
    launch_gather_sorted(vres, xyz, n, sx, sy, sz, inv, stream);
    CloudView cv{};
    cv.sx = sx;
    cv.sy = sy;
    cv.sz = sz;
    cv.perm = vres;
    cv.keys = kres;
    cv.colstart = colstart;
    cv.n = n;
    cv.g = g;

    KnnOutputs out{};
    for (int c = 0; c < ncol_total; ++c) out.col[c] = cols[c];
    for (int i = 0; i < ns; ++i) out.status[i] = status[i];
    out.flagged = flg;
    out.nflagged = cnt;
    out.cov = cov;
    launch_fill_u32(cnt, 16, 0u, stream);
    if (fetch > n)
        launch_flag_all(n, cv.perm, 0, out, stream);
    else
        launch_knn(cv, prog, 0, out, knn_tuning(), stream);
    launch_features(cv, prog, out, inv, stream, nullptr, kFeaturesUnflagged);
    uint32_t h[4] = {0, 0, 0, 0};
    GF_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
    stats.tie_rows = h[1];
    stats.incomplete_rows = h[2];
    if (h[0]) {
        // The exact tree for the flagged rows, in the scratch array: its
        // tables are sized for the worst case and never grown.
        launch_fill_u32(reinterpret_cast<uint32_t*>(dstarts), 4, 0u, stream);
        launch_fill_u32(reinterpret_cast<uint32_t*>(dstarts) + 2, 1, static_cast<uint32_t>(n),
                        stream);
        KdTuning kt{};
        kt.subtree_points = static_cast<uint32_t>(tune.kd_subtree);
        kt.wave_node = static_cast<uint32_t>(tune.kd_wave_node);
        kt.small_blocks = tune.kd_small_blocks < 0 ? 0u : static_cast<uint32_t>(tune.kd_small_blocks);
        kt.minmax_lds_nodes = static_cast<uint32_t>(tune.kd_minmax_lds);
        uint32_t nodes = 0;
        const uint32_t rc = launch_kd_build(x, y, z, n, dstarts, 1, dbounds, t, w, kt, &nodes, stream);
        if (rc != 0)
            throw Error(GF_ERR_UNSUPPORTED, "gf_run_noalloc: exact KD3 tree tables overflowed (code " +
                                                std::to_string(rc) + ")");
        uint32_t* failed = cnt + 5;
        launch_fill_u32(failed, 1, 0u, stream);
        launch_kd_resolve(t, cv, prog, 0, out, x, y, z, h[0], failed, nullptr, stream);
        launch_features(cv, prog, out, inv, stream, nullptr, kFeaturesFlagged, h[0]);
        uint32_t nfail = 0;
        GF_HIP(hipMemcpyAsync(&nfail, failed, 4, hipMemcpyDeviceToHost, stream));
        GF_HIP(hipStreamSynchronize(stream));
        if (nfail)
            throw Error(GF_ERR_UNSUPPORTED, "exact KD3 search exceeded its depth bound on " +
                                                std::to_string(nfail) + " rows");
    }
    GF_HIP(hipStreamSynchronize(stream));
    stats.points = n;
    stats.cell = g.c;
    stats.cell_z = g.cz;
    stats.grid_x = g.gx;
    stats.grid_y = g.gy;
    stats.grid_z = g.gz;
    stats.key_bits = g.key_bits;
}

} // namespace gf

// ------------------------------------------------------------------ ABI ----
namespace {

template <class F>
int guarded(gf_ctx* ctx, F&& f)
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

Context* C(gf_ctx* c) { return reinterpret_cast<Context*>(c); }
Cloud* CL(gf_cloud* c) { return reinterpret_cast<Cloud*>(c); }

} // namespace

extern "C" {

int gf_ctx_create(int device, gf_ctx** out)
{
    if (!out) return GF_ERR_ARG;
    *out = nullptr;
    try {
        Context* c = new Context(device);
        try {
            c->static_scratch = bind_static_scratch(device);
        } catch (...) {
            delete c;
            throw;
        }
        *out = reinterpret_cast<gf_ctx*>(c);
        return GF_OK;
    } catch (const gf::Error& e) {
        std::fprintf(stderr, "gf_ctx_create: %s\n", e.what());
        return e.code;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gf_ctx_create: %s\n", e.what());
        return GF_ERR_STATE;
    }
}

int gf_ctx_destroy(gf_ctx* ctx)
{
    if (!ctx) return GF_ERR_ARG;
    delete C(ctx);
    return GF_OK;
}

const char* gf_last_error(const gf_ctx* ctx)
{
    return ctx ? reinterpret_cast<const Context*>(ctx)->last_error.c_str() : "null context";
}

int gf_sync(gf_ctx* ctx)
{
    if (!ctx) return GF_ERR_ARG;
    return guarded(ctx, [&] { GF_HIP(hipStreamSynchronize(C(ctx)->stream)); });
}

int gf_set_tuning(gf_ctx* ctx, const char* key, double value)
{
    if (!ctx || !key) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        if (!std::strcmp(key, "timing"))
            C(ctx)->timing = value != 0.0;  // per-phase device timers (they synchronize)
        else
            C(ctx)->tune.set(key, value);
    });
}

int gf_cloud_create(gf_ctx* ctx, const double* x, const double* y, const double* z,
                    uint64_t n, int where, gf_cloud** out)
{
    if (!ctx || !out || (n && (!x || !y || !z))) return GF_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        if (n >= (1ull << 32) - 1)
            throw Error(GF_ERR_UNSUPPORTED, "clouds are limited to 2^32-2 points");
        Cloud* cl = new Cloud(C(ctx));
        cl->n = n;
        cl->x.ensure(n * 8 + 8);
        cl->y.ensure(n * 8 + 8);
        cl->z.ensure(n * 8 + 8);
        cl->dstarts.ensure(16);
        const uint64_t se[2] = {0, n};
        GF_HIP(hipMemcpy(cl->dstarts.ptr, se, sizeof(se), hipMemcpyHostToDevice));
        hipMemcpyKind kind = where == GF_WHERE_DEVICE ? hipMemcpyDeviceToDevice
                                                      : hipMemcpyHostToDevice;
        if (n) {
            GF_HIP(hipMemcpyAsync(cl->x.ptr, x, n * 8, kind, C(ctx)->stream));
            GF_HIP(hipMemcpyAsync(cl->y.ptr, y, n * 8, kind, C(ctx)->stream));
            GF_HIP(hipMemcpyAsync(cl->z.ptr, z, n * 8, kind, C(ctx)->stream));
        }
        GF_HIP(hipStreamSynchronize(C(ctx)->stream));
        *out = reinterpret_cast<gf_cloud*>(cl);
    });
}

int gf_cloud_release(gf_cloud* cloud)
{
    if (!cloud) return GF_ERR_ARG;
    delete CL(cloud);
    return GF_OK;
}

uint64_t gf_cloud_size(const gf_cloud* cloud)
{
    return cloud ? reinterpret_cast<const Cloud*>(cloud)->n : 0;
}

int gf_cloud_invalidate(gf_cloud* cloud)
{
    if (!cloud) return GF_ERR_ARG;
    CL(cloud)->idx_valid = false;
    CL(cloud)->kd_valid = false;
    return GF_OK;
}

int gf_stage_columns(const gf_stage* st)
{
    if (!st) return GF_ERR_ARG;
    if (st->stage == GF_STAGE_COVFEAT && (st->features & ~0x3FFu)) return GF_ERR_ARG;
    const int n = stage_ncols(*st);
    return n < 0 ? GF_ERR_ARG : n;
}

int gf_run(gf_ctx* ctx, gf_cloud* cloud, const gf_stage* stages, int nstages,
           double* const* cols, uint8_t* const* status, int where)
{
    if (!ctx || !cloud || !stages || nstages < 1 || nstages > kMaxStages || !status)
        return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        c->stats = gf_stats{};
        int total_cols = 0;
        for (int i = 0; i < nstages; ++i) {
            validate_stage(stages[i]);
            total_cols += stage_ncols(stages[i]);
        }
        if (total_cols && !cols) throw Error(GF_ERR_ARG, "null column array");
        if (cl->n == 0) return;
        // Stages with the same neighborhood run in one fused pass.
        int col0 = 0;
        int i = 0;
        while (i < nstages) {
            int j = i + 1, ncol = stage_ncols(stages[i]);
            auto same = [&](const gf_stage& a, const gf_stage& b) {
                if (a.radius > 0.0 || b.radius > 0.0)
                    return a.radius == b.radius && a.min_k == b.min_k;
                int sa = a.stage == GF_STAGE_NORMAL ? 1 : a.stride;
                int sb = b.stage == GF_STAGE_NORMAL ? 1 : b.stride;
                return stage_neighbors(a) == stage_neighbors(b) && sa == sb;
            };
            while (j < nstages && same(stages[i], stages[j])) {
                ncol += stage_ncols(stages[j]);
                ++j;
            }
            c->run_group(*cl, stages + i, j - i, cols + col0, status + i, where);
            col0 += ncol;
            i = j;
        }
    });
}

int gf_run_noalloc(gf_ctx* ctx, const double* x, const double* y, const double* z, uint64_t n,
                   const gf_stage* stages, int nstages, double* const* cols,
                   uint8_t* const* status)
{
    if (!ctx || !stages || nstages < 1 || nstages > kMaxStages || !status) return GF_ERR_ARG;
    if (n && (!x || !y || !z)) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        c->stats = gf_stats{};
        int total_cols = 0;
        for (int i = 0; i < nstages; ++i) {
            validate_stage(stages[i]);
            total_cols += stage_ncols(stages[i]);
        }
        if (total_cols && !cols) throw Error(GF_ERR_ARG, "null column array");
        if (n == 0) return;
        int col0 = 0;
        int i = 0;
        while (i < nstages) {
            int j = i + 1, ncol = stage_ncols(stages[i]);
            auto same = [&](const gf_stage& a, const gf_stage& b) {
                if (a.radius > 0.0 || b.radius > 0.0)
                    return a.radius == b.radius && a.min_k == b.min_k;
                int sa = a.stage == GF_STAGE_NORMAL ? 1 : a.stride;
                int sb = b.stage == GF_STAGE_NORMAL ? 1 : b.stride;
                return stage_neighbors(a) == stage_neighbors(b) && sa == sb;
            };
            while (j < nstages && same(stages[i], stages[j])) {
                ncol += stage_ncols(stages[j]);
                ++j;
            }
            c->run_group_static(x, y, z, n, stages + i, j - i, cols + col0, status + i);
            col0 += ncol;
            i = j;
        }
    });
}

int gf_radius_counts(gf_ctx* ctx, gf_cloud* cloud, double radius, uint32_t* counts, int where)
{
    if (!ctx || !cloud || !counts) return GF_ERR_ARG;
    if (!(radius > 0.0) || radius == INFINITY) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        const uint64_t n = cl->n;
        if (!n) return;
        c->points_mark = false;
        c->build_index(*cl, 0, radius * c->tune.radius_cell);
        uint32_t* dst = counts;
        if (where == GF_WHERE_HOST) {
            c->outbuf.ensure(n * 4 + 8);
            dst = c->outbuf.as<uint32_t>();
        }
        launch_radius(c->view(*cl), radius, 1, nullptr, dst, c->tune.group_z, c->stream);
        if (where == GF_WHERE_HOST)
            GF_HIP(hipMemcpyAsync(counts, dst, n * 4, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_run_rowk(gf_ctx* ctx, gf_cloud* cloud, const gf_stage* stages, int nstages,
                const uint32_t* row_k, double* const* cols, uint8_t* const* status, int where)
{
    if (!ctx || !cloud || !stages || !row_k || !status || nstages < 1 ||
        nstages > kMaxStages)
        return GF_ERR_ARG;
    return guarded(ctx, [&] {
        for (int i = 0; i < nstages; ++i) {
            gf_stage s = stages[i];
            if (s.radius != 0.0)
                throw Error(GF_ERR_ARG, "per-row neighbor counts take knn neighborhoods only");
            s.knn = 1;  // knn and stride are not used: each row has its own count
            s.stride = 1;
            validate_stage(s);
            if (!status[i]) throw Error(GF_ERR_ARG, "null status array");
        }
        int ncols = 0;
        for (int i = 0; i < nstages; ++i) ncols += stage_ncols(stages[i]);
        if (ncols && !cols) throw Error(GF_ERR_ARG, "null column array");
        for (int c = 0; c < ncols; ++c)
            if (!cols[c]) throw Error(GF_ERR_ARG, "null column array");
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        const uint64_t n = cl->n;
        if (!n) return;
        const uint32_t* drow = row_k;
        DevBuf tmp;
        if (where != GF_WHERE_DEVICE) {
            tmp.ensure(n * 4 + 8);
            GF_HIP(hipMemcpyAsync(tmp.ptr, row_k, n * 4, hipMemcpyHostToDevice, c->stream));
            drow = tmp.as<uint32_t>();
        }
        c->run_rowk(*cl, stages, nstages, drow, cols, status, where);
    });
}

int gf_nndistance(gf_ctx* ctx, gf_cloud* cloud, uint32_t k, int mode, double* out, int where)
{
    if (!ctx || !cloud || !out || (mode != GF_NND_KTH && mode != GF_NND_AVG)) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        const uint64_t n = cl->n;
        if (!n) return;
        double* dst = out;
        if (where != GF_WHERE_DEVICE) {
            c->outbuf.ensure(n * 8 + 8);
            dst = c->outbuf.as<double>();
        }
        c->nn_distance(*cl, k, mode, dst);
        if (where != GF_WHERE_DEVICE)
            GF_HIP(hipMemcpyAsync(out, dst, n * 8, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_outlier(gf_ctx* ctx, gf_cloud* cloud, int method, uint32_t mean_k, double multiplier,
               double radius, int32_t min_k, uint8_t* outlier, double* threshold)
{
    if (!ctx || !cloud || !outlier) return GF_ERR_ARG;
    if (method != GF_OUTLIER_STATISTICAL && method != GF_OUTLIER_RADIUS) return GF_ERR_ARG;
    if (method == GF_OUTLIER_RADIUS && (!(radius > 0.0) || radius == INFINITY)) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        const uint64_t n = cl->n;
        if (!n) return;
        if (method == GF_OUTLIER_RADIUS) {
            // processRadius (filters/OutlierFilter.cpp:80-96): an inlier has
            // more than min_k points within radius (itself included).
            c->points_mark = false;
            c->build_index(*cl, 0, radius * c->tune.radius_cell);
            c->outbuf.ensure(n * 4 + 8);
            launch_radius(c->view(*cl), radius, 1, nullptr, c->outbuf.as<uint32_t>(),
                          c->tune.group_z, c->stream);
            std::vector<uint32_t> counts(n);
            GF_HIP(hipMemcpyAsync(counts.data(), c->outbuf.ptr, n * 4, hipMemcpyDeviceToHost,
                                  c->stream));
            GF_HIP(hipStreamSynchronize(c->stream));
            // ids.size() > size_t(min_k): a negative min_k converts to a
            // count no row reaches.
            const uint64_t need = static_cast<uint64_t>(static_cast<int64_t>(min_k));
            for (uint64_t i = 0; i < n; ++i) outlier[i] = counts[i] > need ? 0 : 1;
            return;
        }
        // processStatistical (filters/OutlierFilter.cpp:98-170): each
        // point's mean distance to its mean_k neighbors (a running mean over
        // them in ascending order, on the device), then the running moments
        // over all points in point order (on the host, in that order, so the
        // threshold is the reference's to the bit).
        c->outbuf.ensure(n * 8 + 8);
        c->nn_distance(*cl, mean_k, kNndRunningMean, c->outbuf.as<double>());
        std::vector<double> d(n);
        GF_HIP(hipMemcpyAsync(d.data(), c->outbuf.ptr, n * 8, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipStreamSynchronize(c->stream));
        uint64_t cnt = 0;
        double M1 = 0.0, M2 = 0.0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t n1 = cnt;
            cnt++;
            const double delta = d[i] - M1;
            const double delta_n = delta / static_cast<double>(cnt);
            M1 += delta_n;
            M2 += delta * delta_n * static_cast<double>(n1);
        }
        const double variance = M2 / (static_cast<double>(cnt) - 1.0);
        const double thr = M1 + multiplier * std::sqrt(variance);
        if (threshold) *threshold = thr;
        for (uint64_t i = 0; i < n; ++i) outlier[i] = d[i] < thr ? 0 : 1;
    });
}

int gf_last_stats(const gf_ctx* ctx, gf_stats* out)
{
    if (!ctx || !out) return GF_ERR_ARG;
    *out = reinterpret_cast<const Context*>(ctx)->stats;
    return GF_OK;
}

int gf_knn(gf_ctx* ctx, gf_cloud* cloud, uint32_t k, uint32_t stride, uint32_t* ids,
           uint8_t* status, int where)
{
    if (!ctx || !cloud || !ids || !status || k < 1 || stride < 1) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        const uint64_t n = cl->n;
        if (!n) return;
        const uint32_t kk = static_cast<uint32_t>(std::min<uint64_t>(n, k));
        uint32_t* dids = ids;
        uint8_t* dst = status;
        if (where != GF_WHERE_DEVICE) {
            c->listbuf.ensure(n * kk * 4 + 8);
            c->liststat.ensure(n + 8);
            dids = c->listbuf.as<uint32_t>();
            dst = c->liststat.as<uint8_t>();
        }
        c->knn_lists(*cl, k, stride, 1, dids, dst);
        if (where != GF_WHERE_DEVICE) {
            GF_HIP(hipMemcpyAsync(ids, dids, n * kk * 4, hipMemcpyDeviceToHost, c->stream));
            GF_HIP(hipMemcpyAsync(status, dst, n, hipMemcpyDeviceToHost, c->stream));
        }
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_optimal_neighborhood(gf_ctx* ctx, gf_cloud* cloud, uint32_t min_k, uint32_t max_k,
                            uint64_t* optimal_knn, double* optimal_radius, int where)
{
    if (!ctx || !cloud || !optimal_knn || !optimal_radius) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        if (min_k < 2 || min_k > max_k || max_k > static_cast<uint32_t>(kMaxFetch - 1))
            throw Error(GF_ERR_ARG, "optimalneighborhood needs 2 <= min_k <= max_k <= 63");
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        const uint64_t n = cl->n;
        if (!n) return;
        if (n < max_k) throw Error(GF_ERR_ARG, "fewer points than max_k");
        c->points_mark = false;
        c->build_index(*cl, max_k);
        FusedProgram prog{};
        prog.k = static_cast<int32_t>(max_k);
        prog.stride = 1;
        prog.fetch = static_cast<int32_t>(max_k);
        KnnOutputs out{};
        c->outbuf.ensure(n * max_k * 4 + n + 8);
        out.ids = c->outbuf.as<uint32_t>();
        out.status[0] = reinterpret_cast<uint8_t*>(out.ids + n * max_k);
        c->flagged.ensure(n * 4 + 4);
        c->counters.ensure(64);
        out.flagged = c->flagged.as<uint32_t>();
        out.nflagged = c->counters.as<uint32_t>();
        launch_fill_u32(out.nflagged, 8, 0u, c->stream);
        const CloudView cv = c->view(*cl);
        const bool early = !cl->kd_valid && c->tune.kd_overlap_min && n >= c->tune.kd_overlap_min;
        if (early) c->fork_side();
        launch_knn(cv, prog, 2, out, c->knn_tuning(*cl, max_k), c->stream);
        if (early) c->build_kdtree(*cl, c->side, true);
        const uint32_t nflagged = c->resolve_flagged(*cl, cv, prog, 2, out, nullptr);
        launch_lists_to_sorted(out.ids, static_cast<int>(max_k), out.flagged, nflagged, cv.perm,
                               cl->inv.as<uint32_t>(), c->stream);
        DevBuf res;
        uint64_t* kopt = optimal_knn;
        double* ropt = optimal_radius;
        if (where != GF_WHERE_DEVICE) {
            res.ensure(n * 16 + 16);
            kopt = res.as<uint64_t>();
            ropt = reinterpret_cast<double*>(kopt + n);
        }
        uint32_t* failed = c->counters.as<uint32_t>() + 6;
        launch_optimal_nbhd(out.ids, n, static_cast<int>(min_k), static_cast<int>(max_k),
                            cl->x.as<double>(), cl->y.as<double>(), cl->z.as<double>(),
                            cv.sx, cv.sy, cv.sz, kopt, ropt, failed, c->stream);
        uint32_t nfail = 0;
        GF_HIP(hipMemcpyAsync(&nfail, failed, 4, hipMemcpyDeviceToHost, c->stream));
        if (where != GF_WHERE_DEVICE) {
            GF_HIP(hipMemcpyAsync(optimal_knn, kopt, n * 8, hipMemcpyDeviceToHost, c->stream));
            GF_HIP(hipMemcpyAsync(optimal_radius, ropt, n * 8, hipMemcpyDeviceToHost,
                                  c->stream));
        }
        GF_HIP(hipStreamSynchronize(c->stream));
        if (nfail) throw Error(GF_ERR_UNSUPPORTED, "Cannot perform eigen decomposition.");
    });
}

int gf_dev_alloc(gf_ctx* ctx, uint64_t bytes, void** out)
{
    if (!ctx || !out) return GF_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        GF_HIP(hipSetDevice(C(ctx)->device));
        hipError_t e = hipMalloc(out, bytes ? bytes : 8);
        if (e != hipSuccess) {
            *out = nullptr;
            throw Error(GF_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
    });
}

int gf_dev_free(gf_ctx* ctx, void* ptr)
{
    if (!ctx) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        if (ptr) GF_HIP(hipFree(ptr));
    });
}

int gf_memcpy(gf_ctx* ctx, void* dst, const void* src, uint64_t bytes, int kind)
{
    if (!ctx || (bytes && (!dst || !src))) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        hipMemcpyKind k = kind == GF_COPY_H2D   ? hipMemcpyHostToDevice
                          : kind == GF_COPY_D2H ? hipMemcpyDeviceToHost
                          : kind == GF_COPY_D2D ? hipMemcpyDeviceToDevice
                                                : throw Error(GF_ERR_ARG, "copy kind");
        if (bytes) GF_HIP(hipMemcpyAsync(dst, src, bytes, k, C(ctx)->stream));
        GF_HIP(hipStreamSynchronize(C(ctx)->stream));
    });
}

int gf_debug_eigen(gf_ctx* ctx, const double* cov, uint64_t m, double* evals,
                   double* evecs, int32_t* info)
{
    if (!ctx || (m && (!cov || !evals || !evecs || !info))) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        DevBuf dc, de, dv, di;
        dc.ensure(m * 72 + 8);
        de.ensure(m * 24 + 8);
        dv.ensure(m * 72 + 8);
        di.ensure(m * 4 + 8);
        GF_HIP(hipMemcpyAsync(dc.ptr, cov, m * 72, hipMemcpyHostToDevice, c->stream));
        launch_eigen_batch(dc.as<double>(), m, de.as<double>(), dv.as<double>(),
                           di.as<int32_t>(), c->stream);
        GF_HIP(hipMemcpyAsync(evals, de.ptr, m * 24, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipMemcpyAsync(evecs, dv.ptr, m * 72, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipMemcpyAsync(info, di.ptr, m * 4, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_debug_grid(gf_ctx* ctx, gf_cloud* cloud, uint32_t k, uint32_t* perm,
                  uint32_t* keys, uint32_t* colstart, uint32_t* ncols,
                  uint32_t* key_bits)
{
    if (!ctx || !cloud || !ncols) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        if (!cl->n) throw Error(GF_ERR_ARG, "empty cloud");
        c->build_index(*cl, k);
        *ncols = cl->g.ncols;
        if (key_bits) *key_bits = cl->g.key_bits;
        if (perm)
            GF_HIP(hipMemcpyAsync(perm, cl->perm, cl->n * 4, hipMemcpyDeviceToHost, c->stream));
        if (keys)
            GF_HIP(hipMemcpyAsync(keys, cl->keys, cl->n * 4, hipMemcpyDeviceToHost, c->stream));
        if (colstart)
            GF_HIP(hipMemcpyAsync(colstart, cl->colstart.ptr,
                                  (static_cast<uint64_t>(cl->g.ncols) + 1) * 4,
                                  hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_debug_grid_params(gf_ctx* ctx, gf_cloud* cloud, uint32_t k, double* f, uint32_t* u)
{
    if (!ctx || !cloud || !f || !u) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        if (!cl->n) throw Error(GF_ERR_ARG, "empty cloud");
        c->build_index(*cl, k);
        const GridParams g = c->grid_of(*cl);
        const double fv[7] = {g.ox, g.oy, g.oz, g.c, g.cz, g.inv_c, g.inv_cz};
        const uint32_t uv[8] = {g.gx, g.gy, g.gz, g.tiles_x, g.bz, g.key_bits, g.ncols, 0};
        std::memcpy(f, fv, sizeof(fv));
        std::memcpy(u, uv, sizeof(uv));
    });
}

int gf_las_parse(const void* bytes, uint64_t nbytes, gf_las_info* info, char* msg,
                 uint64_t msg_size)
{
    if (!bytes || !info) return GF_ERR_ARG;
    std::string err;
    const bool ok = gf::las_parse(static_cast<const uint8_t*>(bytes), nbytes, *info, err);
    if (msg && msg_size) std::snprintf(msg, msg_size, "%s", err.c_str());
    return ok ? GF_OK : GF_ERR_ARG;
}

int gf_cloud_create_las(gf_ctx* ctx, const void* bytes, uint64_t nbytes, gf_las_info* info,
                        gf_cloud** out)
{
    if (!ctx || !bytes || !out) return GF_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        gf_las_info h;
        std::string err;
        if (!las_parse(static_cast<const uint8_t*>(bytes), nbytes, h, err))
            throw Error(GF_ERR_ARG, err);
        if (h.compressed)
            throw Error(GF_ERR_UNSUPPORTED,
                        "LAZ-compressed point data: the device decoder reads uncompressed LAS");
        const uint64_t n = h.point_count;
        if (n >= (1ull << 32) - 1)
            throw Error(GF_ERR_UNSUPPORTED, "clouds are limited to 2^32-2 points");
        Context* c = C(ctx);
        std::unique_ptr<Cloud> cl(new Cloud(c));
        cl->n = n;
        cl->x.ensure(n * 8 + 8);
        cl->y.ensure(n * 8 + 8);
        cl->z.ensure(n * 8 + 8);
        cl->dstarts.ensure(16);
        const uint64_t se[2] = {0, n};
        GF_HIP(hipMemcpy(cl->dstarts.ptr, se, sizeof(se), hipMemcpyHostToDevice));
        const uint64_t L = h.point_size;
        c->lasbuf.ensure(n * L + 16);  // the decoder reads up to 4 bytes past the last record
        LasXform xf{};
        for (int d = 0; d < 3; ++d) {
            xf.scale[d] = h.scale[d];
            xf.offset[d] = h.offset[d];
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (n) {
            GF_HIP(hipMemcpyAsync(c->lasbuf.ptr,
                                  static_cast<const uint8_t*>(bytes) + h.point_offset, n * L,
                                  hipMemcpyHostToDevice, c->stream));
            c->reset_spans();
            e0 = c->mark();
            launch_las_xyz(c->lasbuf.as<uint8_t>(), n, static_cast<uint32_t>(L), xf,
                           cl->x.as<double>(), cl->y.as<double>(), cl->z.as<double>(),
                           c->stream);
            e1 = c->mark();
        }
        GF_HIP(hipStreamSynchronize(c->stream));
        h.ms_decode = 0.0;
        if (e0 && e1) {
            float ms = 0.f;
            GF_HIP(hipEventElapsedTime(&ms, e0, e1));
            h.ms_decode = ms;
        }
        c->reset_spans();
        if (info) *info = h;
        *out = reinterpret_cast<gf_cloud*>(cl.release());
    });
}

int gf_las_pack_extra(gf_ctx* ctx, const double* const* cols, int ncols, uint64_t n,
                      uint8_t* records, uint32_t stride, uint32_t offset, int where)
{
    if (!ctx || ncols < 0 || ncols > 16 || (n && ncols && (!cols || !records)))
        return GF_ERR_ARG;
    return guarded(ctx, [&] {
        if (where != GF_WHERE_HOST && where != GF_WHERE_DEVICE)
            throw Error(GF_ERR_ARG, "where must be GF_WHERE_HOST or GF_WHERE_DEVICE");
        if (static_cast<uint64_t>(offset) + 8ull * static_cast<uint64_t>(ncols) > stride)
            throw Error(GF_ERR_ARG, "the extra dimensions do not fit in the record");
        if (!n || !ncols) return;
        for (int i = 0; i < ncols; ++i)
            if (!cols[i]) throw Error(GF_ERR_ARG, "null column");
        Context* c = C(ctx);
        LasColumns lc{};
        uint8_t* dst = records;
        const uint64_t bytes = n * stride;
        if (where == GF_WHERE_HOST) {
            c->outbuf.ensure(n * 8 * static_cast<uint64_t>(ncols) + 8);
            for (int i = 0; i < ncols; ++i) {
                double* d = c->outbuf.as<double>() + static_cast<uint64_t>(i) * n;
                GF_HIP(hipMemcpyAsync(d, cols[i], n * 8, hipMemcpyHostToDevice, c->stream));
                lc.col[i] = d;
            }
            c->lasbuf.ensure(bytes + 16);
            GF_HIP(hipMemcpyAsync(c->lasbuf.ptr, records, bytes, hipMemcpyHostToDevice,
                                  c->stream));
            dst = c->lasbuf.as<uint8_t>();
        } else {
            for (int i = 0; i < ncols; ++i) lc.col[i] = cols[i];
        }
        launch_las_pack(lc, ncols, n, dst, stride, offset, c->stream);
        if (where == GF_WHERE_HOST)
            GF_HIP(hipMemcpyAsync(records, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_las_write_extra(gf_ctx* ctx, const void* src, uint64_t src_nbytes, int ncols,
                       const char* const* names, const char* const* descriptions,
                       const uint8_t* types, const double* const* cols, int where, void* dst,
                       uint64_t dst_capacity, uint64_t* dst_nbytes)
{
    if (!ctx || !src || !dst_nbytes || ncols < 1 || ncols > 16 || !cols) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        if (where != GF_WHERE_HOST && where != GF_WHERE_DEVICE)
            throw Error(GF_ERR_ARG, "where must be GF_WHERE_HOST or GF_WHERE_DEVICE");
        for (int i = 0; i < ncols; ++i)
            if (!cols[i]) throw Error(GF_ERR_ARG, "null column");
        C(ctx)->las_write_extra(static_cast<const uint8_t*>(src), src_nbytes, ncols, names,
                                descriptions, types, cols, where, static_cast<uint8_t*>(dst),
                                dst_capacity, dst_nbytes);
    });
}

int gf_debug_cloud_xyz(gf_ctx* ctx, gf_cloud* cloud, double* x, double* y, double* z)
{
    if (!ctx || !cloud || !x || !y || !z) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        if (!cl->n) return;
        GF_HIP(hipMemcpyAsync(x, cl->x.ptr, cl->n * 8, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipMemcpyAsync(y, cl->y.ptr, cl->n * 8, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipMemcpyAsync(z, cl->z.ptr, cl->n * 8, hipMemcpyDeviceToHost, c->stream));
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_debug_copy(gf_ctx* ctx, double* dst, const double* src, uint64_t count)
{
    if (!ctx || (count && (!dst || !src))) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        launch_copy_f64(src, dst, count, C(ctx)->stream);
        GF_HIP(hipStreamSynchronize(C(ctx)->stream));
    });
}

int gf_debug_kdtree(gf_ctx* ctx, gf_cloud* cloud, uint32_t* vind, uint32_t* nodes,
                    uint32_t* left, uint32_t* right, int32_t* feat, double* divlow,
                    double* divhigh)
{
    if (!ctx || !cloud || !nodes) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        if (!cl->n) throw Error(GF_ERR_ARG, "empty cloud");
        if (!cl->idx_valid) c->build_index(*cl, 1);
        c->build_kdtree(*cl);
        const KdView& t = cl->kd.view;
        const uint32_t m = cl->kd.nodes;
        *nodes = m;
        auto get = [&](void* dst, const void* src, uint64_t bytes) {
            if (dst) GF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        };
        get(vind, t.vind, cl->n * 4);
        get(left, t.nl, static_cast<uint64_t>(m) * 4);
        get(right, t.nr, static_cast<uint64_t>(m) * 4);
        get(feat, t.feat, static_cast<uint64_t>(m) * 4);
        get(divlow, t.divlow, static_cast<uint64_t>(m) * 8);
        get(divhigh, t.divhigh, static_cast<uint64_t>(m) * 8);
        GF_HIP(hipStreamSynchronize(c->stream));
    });
}

int gf_autotune(gf_ctx* ctx, gf_cloud* cloud, const gf_stage* stages, int nstages,
                const double* factors, int nfactors, double* best_factor, double* best_ms)
{
    if (!ctx || !cloud || !stages || nstages < 1 || nstages > kMaxStages || !factors ||
        nfactors < 1)
        return GF_ERR_ARG;
    return guarded(ctx, [&] {
        Context* c = C(ctx);
        Cloud* cl = CL(cloud);
        for (int i = 0; i < nstages; ++i) validate_stage(stages[i]);
        if (!cl->n) throw Error(GF_ERR_ARG, "empty cloud");
        const uint64_t n = cl->n;
        int ncols = 0;
        for (int i = 0; i < nstages; ++i) ncols += stage_ncols(stages[i]);
        DevBuf dcols, dstat;
        dcols.ensure(static_cast<uint64_t>(std::max(ncols, 1)) * n * 8);
        dstat.ensure(static_cast<uint64_t>(nstages) * n);
        std::vector<double*> cp(std::max(ncols, 1));
        std::vector<uint8_t*> sp(nstages);
        for (int k = 0; k < ncols; ++k) cp[k] = dcols.as<double>() + static_cast<uint64_t>(k) * n;
        for (int k = 0; k < nstages; ++k) sp[k] = dstat.as<uint8_t>() + static_cast<uint64_t>(k) * n;
        const bool timing = c->timing;
        c->timing = false;
        const Tuning saved = c->tune;
        // One measurement: the whole run, index rebuilt, best of two.
        auto measure = [&]() {
            double t_min = 1e300;
            for (int rep = 0; rep < 2; ++rep) {
                cl->idx_valid = false;
                cl->kd_valid = false;
                GF_HIP(hipStreamSynchronize(c->stream));
                const auto t0 = std::chrono::steady_clock::now();
                if (gf_run(ctx, cloud, stages, nstages, cp.data(), sp.data(), GF_WHERE_DEVICE) !=
                    GF_OK)
                    throw Error(GF_ERR_STATE, c->last_error);
                const double ms = std::chrono::duration<double, std::milli>(
                                      std::chrono::steady_clock::now() - t0)
                                      .count();
                t_min = std::min(t_min, ms);
            }
            return t_min;
        };
        double best = 1e300, bestf = factors[0];
        TuneEntry win{};
        try {
            // Coordinate descent: the column-size factor, then the search's
            // z grouping, queries per wave and waves per workgroup.
            for (int fi = 0; fi < nfactors; ++fi) {
                if (!(factors[fi] > 0)) throw Error(GF_ERR_ARG, "factors must be positive");
                c->tune.forced_factor = factors[fi];
                const double ms = measure();
                if (ms < best) {
                    best = ms;
                    bestf = factors[fi];
                }
            }
            c->tune.forced_factor = bestf;
            win.cell_factor = bestf;
            int bestgz = saved.group_z;
            for (int gz : {0, 1, 2, 3, 4}) {
                if (gz == saved.group_z) continue;  // measured above
                c->tune.group_z = gz;
                const double ms = measure();
                if (ms < best) {
                    best = ms;
                    bestgz = gz;
                }
            }
            c->tune.group_z = bestgz;
            win.group_z = bestgz + 1;
            const int q0 = saved.q_per_wave;
            int bestq = q0;
            for (int q : {16, 32, 64}) {
                if (q == q0) continue;
                c->tune.q_per_wave = q;
                const double ms = measure();
                if (ms < best) {
                    best = ms;
                    bestq = q;
                }
            }
            c->tune.q_per_wave = bestq;
            win.q_per_wave = bestq;
            int bestw = saved.knn_waves;
            for (int w : {1, 2, 4}) {
                if (w == saved.knn_waves || (saved.knn_waves == 0 && w == 4)) continue;
                c->tune.knn_waves = w;
                const double ms = measure();
                if (ms < best) {
                    best = ms;
                    bestw = w;
                }
            }
            win.knn_waves = bestw;
        } catch (...) {
            c->tune = saved;
            c->timing = timing;
            throw;
        }
        c->tune = saved;
        c->timing = timing;
        // Remember the winner for this cloud's bucket.
        const int stride = stages[0].stage == GF_STAGE_NORMAL ? 1 : stages[0].stride;
        const uint32_t fetch = static_cast<uint32_t>(stride) *
                               static_cast<uint32_t>(std::min<uint64_t>(n, stage_neighbors(stages[0])));
        c->tune.remember(cloud_rho(cl->bounds, n), fetch, win);
        cl->idx_valid = false;
        cl->kd_valid = false;
        if (best_factor) *best_factor = bestf;
        if (best_ms) *best_ms = best;
    });
}

int gf_tuning_save(gf_ctx* ctx, const char* path)
{
    if (!ctx || !path) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        FILE* f = std::fopen(path, "w");
        if (!f) throw Error(GF_ERR_ARG, std::string("cannot write ") + path);
        std::fprintf(f, "# gpupdal_amd autotune: rho_bucket k_bucket cell_factor group_z+1 "
                        "q_per_wave knn_waves (0: the context's setting)\n");
        for (const TuneEntry& e : C(ctx)->tune.table)
            std::fprintf(f, "%d %d %.17g %d %d %d\n", e.rho_bucket, e.k_bucket, e.cell_factor,
                         e.group_z, e.q_per_wave, e.knn_waves);
        std::fclose(f);
    });
}

int gf_tuning_load(gf_ctx* ctx, const char* path)
{
    if (!ctx || !path) return GF_ERR_ARG;
    return guarded(ctx, [&] {
        FILE* f = std::fopen(path, "r");
        if (!f) throw Error(GF_ERR_ARG, std::string("cannot read ") + path);
        char line[256];
        std::vector<TuneEntry> t;
        while (std::fgets(line, sizeof(line), f)) {
            if (line[0] == '#') continue;
            TuneEntry e{};
            const int got = std::sscanf(line, "%d %d %lf %d %d %d", &e.rho_bucket, &e.k_bucket,
                                        &e.cell_factor, &e.group_z, &e.q_per_wave,
                                        &e.knn_waves);
            const bool shape_ok = e.group_z >= 0 && e.group_z <= 65 &&
                                  (e.q_per_wave == 0 || e.q_per_wave == 16 ||
                                   e.q_per_wave == 32 || e.q_per_wave == 64) &&
                                  (e.knn_waves == 0 || e.knn_waves == 1 || e.knn_waves == 2 ||
                                   e.knn_waves == 4);
            if ((got == 3 || got == 6) && e.cell_factor > 0 && shape_ok) t.push_back(e);
        }
        std::fclose(f);
        for (const TuneEntry& e : t) {
            bool found = false;
            for (TuneEntry& x : C(ctx)->tune.table)
                if (x.rho_bucket == e.rho_bucket && x.k_bucket == e.k_bucket) {
                    x = e;
                    found = true;
                }
            if (!found) C(ctx)->tune.table.push_back(e);
        }
    });
}

} // extern "C"
