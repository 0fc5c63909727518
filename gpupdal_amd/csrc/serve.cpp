// serve.cpp -- a serving loop over batches: requests (one tile each) arrive
// at any time; worker threads, each with its own context and stream, admit
// pending requests under an HBM budget, run them as one batch, and hand the
// results back.  Requests that arrive while batches are on the device wait
// only until a worker is free and the budget allows; two workers keep two
// batches in flight, so uploads, searches and copies of different batches
// overlap on the device.  A request's result equals the tile run alone.
#include "gpupdal_amd.h"

#include "gf_internal.hpp"
#include "gf_kernels.hpp"
#include "gf_runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gf {

namespace {

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Device bytes a request of n points holds while in flight: the point
// buffers, the index, the exact tree, the covariance and flags, and the
// stage outputs.
uint64_t request_bytes(uint64_t n, int ncols)
{
    return n * (256ull + 8ull * static_cast<uint64_t>(ncols)) + (1ull << 20);
}

// Page-locked host memory (allocated as such, grow-only), the staging
// area of a worker's result downloads.
struct PinnedBuf {
    void* ptr = nullptr;
    uint64_t size = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf()
    {
        if (ptr) (void)hipHostFree(ptr);
    }
    // Page-locking is slow (hundreds of ms for a few GB), so the buffer grows
    // at least twofold.
    void ensure(uint64_t bytes)
    {
        if (bytes <= size && ptr) return;
        const uint64_t want = std::max<uint64_t>(bytes, 2 * size);
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        size = 0;
        GF_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
        size = want;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Bytes of one column-piece staging slot (two per worker).
constexpr uint64_t kStageSlotBytes = 32ull << 20;

constexpr uint8_t kSkipped =
    GF_ST_ZERO_COV | GF_ST_EIGEN_FAIL | GF_ST_FEATURE_INVALID | GF_ST_FEW_NEIGHBORS;

struct Request {
    uint64_t id = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;
    uint64_t n = 0;
    gf_stage st[kMaxStages];
    int ns = 0;
    int32_t knn = 0;
    std::vector<double*> cols;
    std::vector<uint8_t*> status;
    int ncols = 0;
    uint64_t bytes = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    uint32_t batch_id = 0, batch_tiles = 0;
    bool done = false;
    int rc = GF_OK;
    std::string error;
    int waiters = 0;        // threads inside gf_server_wait on this request
    bool released = false;  // released while waited on: the last waiter frees it
};

bool same_shape(const Request& a, const Request& b)
{
    if (a.ns != b.ns) return false;
    for (int i = 0; i < a.ns; ++i) {
        const gf_stage &p = a.st[i], &q = b.st[i];
        if (p.stage != q.stage || p.mode != q.mode || p.features != q.features ||
            p.always_up != q.always_up || p.normalize != q.normalize ||
            (p.stage != GF_STAGE_NORMAL && p.stride != q.stride))
            return false;
    }
    return true;
}

} // namespace

struct Server {
    int device = 0;
    uint64_t budget = 0;
    uint64_t max_batch_points = 0;
    uint32_t max_batch_tiles = 256;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_id = 1;
    uint64_t in_use = 0;
    uint32_t next_batch = 1;
    bool stopping = false;
    std::vector<std::thread> workers;
    int ready = 0;  // workers that created their context (or failed to)
    std::string create_error;

    void worker_main(int wi);
};

void Server::worker_main(int)
{
    std::unique_ptr<Context> ctx;
    // Page-locked staging: two fixed slots for the column pieces, and the
    // statuses (3 bytes a point for three stages: grown twofold when needed).
    PinnedBuf slots, hstat_buf;
    hipEvent_t slot_ev[2] = {nullptr, nullptr};
    struct EventGuard {
        hipEvent_t* ev;
        ~EventGuard()
        {
            for (int i = 0; i < 2; ++i)
                if (ev[i]) (void)hipEventDestroy(ev[i]);
        }
    } event_guard{slot_ev};
    try {
        ctx.reset(new Context(device));
        ctx->timing = false;
        slots.ensure(2 * kStageSlotBytes);
        hstat_buf.ensure(16ull << 20);
        GF_HIP(hipEventCreateWithFlags(&slot_ev[0], hipEventDisableTiming));
        GF_HIP(hipEventCreateWithFlags(&slot_ev[1], hipEventDisableTiming));
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu);
        create_error = e.what();
        ++ready;
        cv_done.notify_all();
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        ++ready;
    }
    cv_done.notify_all();
    Batch batch(ctx.get());
    DevBuf dcols, dstat;
    for (;;) {
        std::vector<Request*> take;
        uint64_t bytes = 0;
        uint32_t bid = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (stopping && pending.empty()) return;
                if (!pending.empty()) {
                    // Admission: requests in arrival order while they fit the
                    // budget (a request alone larger than the whole budget is
                    // refused), share the first one's stage shape, and the
                    // batch stays within its point and tile limits.
                    Request* first = pending.front();
                    if (first->bytes > budget) {
                        pending.pop_front();
                        first->rc = GF_ERR_OOM;
                        first->error = "request exceeds the device memory budget";
                        first->t_done = now_s();
                        first->done = true;
                        cv_done.notify_all();
                        continue;
                    }
                    uint64_t pts = 0;
                    for (auto it = pending.begin(); it != pending.end();) {
                        Request* r = *it;
                        if (!same_shape(*r, *first)) {
                            ++it;
                            continue;
                        }
                        if (in_use + bytes + r->bytes > budget) break;
                        if (!take.empty() && (pts + r->n > max_batch_points ||
                                              take.size() >= max_batch_tiles))
                            break;
                        take.push_back(r);
                        bytes += r->bytes;
                        pts += r->n;
                        it = pending.erase(it);
                    }
                    if (!take.empty()) {
                        in_use += bytes;
                        bid = next_batch++;
                        break;
                    }
                }
                cv_work.wait(lk);
            }
        }
        // Run the batch on this worker's stream.
        const double t0 = now_s();
        int rc = GF_OK;
        std::string err;
        try {
            GF_HIP(hipSetDevice(device));
            const uint32_t T = static_cast<uint32_t>(take.size());
            std::vector<uint64_t> counts(T);
            std::vector<int32_t> knn(T);
            for (uint32_t t = 0; t < T; ++t) {
                counts[t] = take[t]->n;
                knn[t] = take[t]->knn;
            }
            batch.reset(T, counts.data(), ctx->stream);
            const uint64_t N = batch.n;
            const int ncols = take[0]->ncols;
            const int ns = take[0]->ns;
            dcols.ensure(static_cast<uint64_t>(ncols) * N * 8 + 8);
            dstat.ensure(static_cast<uint64_t>(ns) * N + 8);
            std::vector<double*> cp(ncols);
            std::vector<uint8_t*> sp(ns);
            for (int c = 0; c < ncols; ++c) cp[c] = dcols.as<double>() + static_cast<uint64_t>(c) * N;
            for (int i = 0; i < ns; ++i) sp[i] = dstat.as<uint8_t>() + static_cast<uint64_t>(i) * N;
            for (uint32_t t = 0; t < T; ++t) {
                Request* r = take[t];
                const uint64_t o = batch.starts[t];
                if (!r->n) continue;
                GF_HIP(hipMemcpyAsync(batch.x.as<double>() + o, r->x, r->n * 8,
                                      hipMemcpyHostToDevice, ctx->stream));
                GF_HIP(hipMemcpyAsync(batch.y.as<double>() + o, r->y, r->n * 8,
                                      hipMemcpyHostToDevice, ctx->stream));
                GF_HIP(hipMemcpyAsync(batch.z.as<double>() + o, r->z, r->n * 8,
                                      hipMemcpyHostToDevice, ctx->stream));
            }
            if (N) {
                ctx->stats = gf_stats{};
                ctx->run_batch(batch, take[0]->st, ns, knn.data(), cp.data(), sp.data(),
                               GF_WHERE_DEVICE);
            }
            // Results come back through page-locked staging: the statuses
            // first, then the columns in pieces through two fixed slots, each
            // piece's DMA running while the host copies the previous piece
            // out.  Rows a stage skipped keep the caller's prior values (put
            // back from a copy taken just before the piece lands), so those
            // never travel to the device.
            if (N) {
                hstat_buf.ensure(static_cast<uint64_t>(ns) * N);
                GF_HIP(hipMemcpyAsync(hstat_buf.ptr, dstat.ptr, static_cast<uint64_t>(ns) * N,
                                      hipMemcpyDeviceToHost, ctx->stream));
                GF_HIP(hipStreamSynchronize(ctx->stream));
            }
            const uint8_t* hstat = hstat_buf.as<uint8_t>();
            struct Piece {
                double* dst;
                const double* src;
                uint64_t count;
                const uint8_t* status;
            };
            std::vector<Piece> pieces;
            const uint64_t slot_elems = kStageSlotBytes / 8;
            for (uint32_t t = 0; t < T; ++t) {
                Request* r = take[t];
                const uint64_t o = batch.starts[t];
                int c = 0;
                for (int i = 0; i < ns; ++i) {
                    const uint8_t* srow = hstat + static_cast<uint64_t>(i) * N + o;
                    const int nc = stage_ncols(r->st[i]);
                    for (int cc = 0; cc < nc; ++cc, ++c)
                        for (uint64_t j = 0; j < r->n; j += slot_elems)
                            pieces.push_back({r->cols[c] + j, cp[c] + o + j,
                                              std::min<uint64_t>(slot_elems, r->n - j), srow + j});
                    if (r->n) std::memcpy(r->status[i], srow, r->n);
                }
            }
            std::vector<std::pair<uint64_t, double>> keep;
            auto copy_out = [&](const Piece& pc, const double* staged) {
                keep.clear();
                for (uint64_t j = 0; j < pc.count; ++j)
                    if (pc.status[j] & kSkipped) keep.emplace_back(j, pc.dst[j]);
                std::memcpy(pc.dst, staged, pc.count * 8);
                for (const auto& kv : keep) pc.dst[kv.first] = kv.second;
            };
            for (size_t q = 0; q <= pieces.size(); ++q) {
                if (q < pieces.size()) {
                    GF_HIP(hipMemcpyAsync(slots.as<double>() + (q & 1) * slot_elems, pieces[q].src,
                                          pieces[q].count * 8, hipMemcpyDeviceToHost,
                                          ctx->stream));
                    GF_HIP(hipEventRecord(slot_ev[q & 1], ctx->stream));
                }
                if (q > 0) {
                    GF_HIP(hipEventSynchronize(slot_ev[(q - 1) & 1]));
                    copy_out(pieces[q - 1], slots.as<double>() + ((q - 1) & 1) * slot_elems);
                }
            }
        } catch (const Error& e) {
            rc = e.code;
            err = e.what();
        } catch (const std::exception& e) {
            rc = GF_ERR_STATE;
            err = e.what();
        }
        const double t1 = now_s();
        {
            std::lock_guard<std::mutex> lk(mu);
            in_use -= bytes;
            for (Request* r : take) {
                r->t_start = t0;
                r->t_done = t1;
                r->batch_id = bid;
                r->batch_tiles = static_cast<uint32_t>(take.size());
                r->rc = rc;
                r->error = err;
                r->done = true;
            }
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

} // namespace gf

using namespace gf;

extern "C" {

int gf_server_create(int device, int workers, uint64_t hbm_budget_bytes,
                     uint64_t max_batch_points, gf_server** out)
{
    if (!out || workers < 1 || workers > 8) return GF_ERR_ARG;
    *out = nullptr;
    Server* s = new Server();
    s->device = device;
    s->budget = hbm_budget_bytes ? hbm_budget_bytes : (200ull << 30);
    s->max_batch_points = max_batch_points ? max_batch_points : (64ull << 20);
    for (int i = 0; i < workers; ++i) s->workers.emplace_back(&Server::worker_main, s, i);
    bool failed;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->ready == workers; });
        failed = !s->create_error.empty();
        if (failed) {
            std::fprintf(stderr, "gf_server_create: %s\n", s->create_error.c_str());
            s->stopping = true;
        }
    }
    if (failed) {
        s->cv_work.notify_all();
        for (auto& t : s->workers) t.join();
        delete s;
        return GF_ERR_HIP;
    }
    *out = reinterpret_cast<gf_server*>(s);
    return GF_OK;
}

int gf_server_submit(gf_server* server, const double* x, const double* y, const double* z,
                     uint64_t n, const gf_stage* stages, int nstages, int32_t knn,
                     double* const* cols, uint8_t* const* status, uint64_t* ticket)
{
    if (!server || !stages || nstages < 1 || nstages > kMaxStages || !status || !ticket ||
        (n && (!x || !y || !z)) || knn < 1)
        return GF_ERR_ARG;
    Server* s = reinterpret_cast<Server*>(server);
    std::unique_ptr<Request> r(new Request());
    int ncols = 0;
    // A request is checked here, before it is queued: one the batch would
    // refuse must not fail the requests batched with it.
    for (int i = 0; i < nstages; ++i) {
        r->st[i] = stages[i];
        r->st[i].knn = knn;
        try {
            validate_stage(r->st[i]);
        } catch (const Error& e) {
            return e.code;
        }
        if (r->st[i].radius > 0.0) return GF_ERR_UNSUPPORTED;  // served stages use knn
        if (r->st[i].stage == GF_STAGE_COPLANAR || r->st[i].stage == GF_STAGE_RANK)
            return GF_ERR_UNSUPPORTED;
        if (i && (r->st[i].stage == GF_STAGE_NORMAL ? 1 : r->st[i].stride) !=
                     (r->st[0].stage == GF_STAGE_NORMAL ? 1 : r->st[0].stride))
            return GF_ERR_ARG;  // the stages of one request share one neighborhood
        ncols += gf_stage_columns(&r->st[i]);
    }
    if (ncols > kMaxOutColumns) return GF_ERR_UNSUPPORTED;
    try {
        check_neighborhood(n, static_cast<int64_t>(knn) + 1,
                           r->st[0].stage == GF_STAGE_NORMAL ? 1 : r->st[0].stride,
                           kMaxFetch - 1);
    } catch (const Error& e) {
        return e.code;
    }
    if (ncols && !cols) return GF_ERR_ARG;
    r->x = x;
    r->y = y;
    r->z = z;
    r->n = n;
    r->ns = nstages;
    r->knn = knn;
    r->ncols = ncols;
    r->cols.assign(cols, cols + ncols);
    r->status.assign(status, status + nstages);
    r->bytes = request_bytes(n, ncols);
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return GF_ERR_STATE;
        r->id = s->next_id++;
        *ticket = r->id;
        s->pending.push_back(r.get());
        s->requests[r->id] = std::move(r);
    }
    s->cv_work.notify_one();
    return GF_OK;
}

int gf_server_wait(gf_server* server, uint64_t ticket, double timeout_s)
{
    if (!server) return GF_ERR_ARG;
    Server* s = reinterpret_cast<Server*>(server);
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return GF_ERR_ARG;
    Request* r = it->second.get();
    if (r->released) return GF_ERR_ARG;
    // The request stays alive while this thread waits on it: a release in
    // the meantime only marks it, and the last waiter frees it.
    ++r->waiters;
    const auto pred = [&] { return r->done; };
    bool finished = true;
    if (timeout_s < 0)
        s->cv_done.wait(lk, pred);
    else
        finished = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    const int rc = finished ? r->rc : GF_ERR_TIMEOUT;
    if (--r->waiters == 0 && r->released) s->requests.erase(r->id);
    return rc;
}

int gf_server_request_info(gf_server* server, uint64_t ticket, double* times,
                           uint32_t* batch_id, uint32_t* batch_tiles)
{
    if (!server) return GF_ERR_ARG;
    Server* s = reinterpret_cast<Server*>(server);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return GF_ERR_ARG;
    const Request* r = it->second.get();
    if (r->released) return GF_ERR_ARG;
    if (times) {
        times[0] = r->t_submit;
        times[1] = r->t_start;
        times[2] = r->t_done;
    }
    if (batch_id) *batch_id = r->batch_id;
    if (batch_tiles) *batch_tiles = r->batch_tiles;
    return r->done ? r->rc : GF_ERR_STATE;
}

int gf_server_release(gf_server* server, uint64_t ticket)
{
    if (!server) return GF_ERR_ARG;
    Server* s = reinterpret_cast<Server*>(server);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end() || !it->second->done || it->second->released) return GF_ERR_ARG;
    if (it->second->waiters > 0)
        it->second->released = true;  // freed by its last waiter
    else
        s->requests.erase(it);
    return GF_OK;
}

int gf_server_destroy(gf_server* server)
{
    if (!server) return GF_ERR_ARG;
    Server* s = reinterpret_cast<Server*>(server);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    delete s;
    return GF_OK;
}

} // extern "C"
