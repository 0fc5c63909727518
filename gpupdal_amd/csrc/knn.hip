// knn.hip -- wavefront-per-query k-nearest-neighbor search over the column
// grid, fused with the neighborhood covariance, the eigensolve and the stage
// outputs.
//
// Replaces the reference's thread-per-query knnGatherKernel with its fused
// covariance, the device eigen kernel and the projection kernels
// (src/index/SpatialIndexKernels.cu:716-1086, 1716-1769;
// src/stages/NeighborhoodKernels.cu:67-119, 422-485).
//
// Per query, the 64 lanes of one wave evaluate the candidates of a box of
// (2r+1)^2 columns x (2r+1) z cells around the query's cell in parallel,
// with FP64 squared distances accumulated exactly as KD3Index does
// (d0*d0 + d1*d1 + d2*d2, left to right, no FMA).  Survivors are merged into
// a wave-wide list of the 64 smallest (d2, point id) pairs by a bitonic
// network over lanes.  The search is certified when the fetch-th distance is
// below the distance from the query to the nearest face of the box that
// borders unexamined space (minus a rounding margin); otherwise the box grows
// to the radius the current fetch-th distance demands, up to a budget.
// Rows whose fetch-th and (fetch+1)-th distances are equal have a boundary
// tie, which KD3Index breaks by tree visit order rather than by index; they
// and uncertified rows are flagged for the exact KD3 search.
//
// Each wave runs Q queries back to back and parks each neighbor list in LDS;
// then every lane takes one of those queries and computes its FP64
// covariance (two-pass mean), the eigensystem and all requested stage
// columns, so the per-row solve runs on all lanes.
#include "gf_eigen.hpp"
#include "gf_internal.hpp"
#include "gf_kernels.hpp"

namespace gf {

namespace {

constexpr uint64_t kInfD = ~0ull;
constexpr uint32_t kInfU = ~0u;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

struct Ent {
    uint64_t d;   // bit pattern of the non-negative FP64 squared distance
    uint32_t id;  // original point index (tie order between equal distances)
    uint32_t j;   // sorted position (where the coordinates live)
};

__device__ __forceinline__ Ent ent_inf() { return Ent{kInfD, kInfU, kInfU}; }

__device__ __forceinline__ bool ent_less(const Ent& a, const Ent& b)
{
    return a.d < b.d || (a.d == b.d && a.id < b.id);
}

__device__ __forceinline__ Ent ent_shfl_xor(const Ent& e, int m)
{
    Ent r;
    r.d = __shfl_xor(static_cast<unsigned long long>(e.d), m);
    r.id = __shfl_xor(e.id, m);
    r.j = __shfl_xor(e.j, m);
    return r;
}

__device__ __forceinline__ Ent ent_shfl(const Ent& e, int src)
{
    Ent r;
    r.d = __shfl(static_cast<unsigned long long>(e.d), src);
    r.id = __shfl(e.id, src);
    r.j = __shfl(e.j, src);
    return r;
}

// Ascending bitonic sort of one entry per lane across the wave.
__device__ __forceinline__ void bitonic_sort64(Ent& e, int lane)
{
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            Ent o = ent_shfl_xor(e, j);
            bool up = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            bool take = (lower == up) ? ent_less(o, e) : ent_less(e, o);
            if (take) e = o;
        }
    }
}

// best (ascending) <- the 64 smallest of best U cand (cand ascending).
__device__ __forceinline__ void merge_sorted(Ent& best, const Ent& cand, int lane)
{
    Ent rev = ent_shfl(cand, 63 - lane);
    if (ent_less(rev, best)) best = rev;
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) {
        Ent o = ent_shfl_xor(best, j);
        bool lower = (lane & j) == 0;
        bool take = lower ? ent_less(o, best) : ent_less(best, o);
        if (take) best = o;
    }
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* __restrict__ a,
                                                    uint32_t lo, uint32_t hi,
                                                    uint32_t key)
{
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Block index -> logical block, so that each XCD (blocks are dealt to the 8
// XCDs round-robin) works on one contiguous run of sorted queries and its L2
// sees neighboring columns.
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t nb)
{
    const uint32_t x = 8;
    uint32_t xcd = b % x, idx = b / x, per = nb / x, rem = nb % x;
    return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

struct KnnArgs {
    CloudView cv;
    FusedProgram prog;
    KnnOutputs out;
    KnnTuning tune;
    int32_t mode;     // 0 features, 1 neighbor lists
    int32_t q_per_wave;
    int32_t lds_stride;
    double margin;
};

// The k-NN search of one query by the whole wave.  On return lane l holds the
// l-th smallest (d2, id) pair (64 kept); *incomplete tells whether the
// budget ran out before certification.
__device__ void wave_knn(const KnnArgs& a, double qx, double qy, double qz,
                         uint32_t qcx, uint32_t qcy, uint32_t qcz, int lane,
                         Ent& best, bool& incomplete, uint64_t* cnt)
{
    const CloudView& cv = a.cv;
    const GridParams& g = cv.g;
    const int fetch = a.prog.fetch;
    int r = a.tune.ring0;
    incomplete = false;
    for (;;) {
        best = ent_inf();
        cnt[0] += 1;
        const int span = 2 * r + 1;
        const int ncb = span * span;
        const int za = max(static_cast<int>(qcz) - r, 0);
        const int zb = min(static_cast<int>(qcz) + r, static_cast<int>(g.gz) - 1);
        const bool zfull = (za == 0) && (zb == static_cast<int>(g.gz) - 1);
        for (int cb = 0; cb < ncb; cb += 64) {
            const int ci = cb + lane;
            uint32_t lo = 0, len = 0;
            if (ci < ncb) {
                int cx = static_cast<int>(qcx) + ci % span - r;
                int cy = static_cast<int>(qcy) + ci / span - r;
                if (cx >= 0 && cy >= 0 && cx < static_cast<int>(g.gx) &&
                    cy < static_cast<int>(g.gy)) {
                    uint32_t col = column_index(g, cx, cy);
                    uint32_t b0 = cv.colstart[col], b1 = cv.colstart[col + 1];
                    if (!zfull && b1 - b0 > static_cast<uint32_t>(a.tune.small_col)) {
                        uint32_t klo = (col << g.bz) + static_cast<uint32_t>(za);
                        uint32_t khi = (col << g.bz) + static_cast<uint32_t>(zb + 1);
                        b0 = lower_bound_u32(cv.keys, b0, b1, klo);
                        b1 = lower_bound_u32(cv.keys, b0, b1, khi);
                    }
                    lo = b0;
                    len = b1 - b0;
                }
            }
            const uint32_t incl = wave_incl_scan_u32(len, lane);
            const uint32_t tot = __shfl(incl, 63);
            const uint32_t excl = incl - len;
            cnt[1] += tot;
            for (uint32_t t0 = 0; t0 < tot; t0 += 64) {
                const uint32_t t = t0 + lane;
                const bool valid = t < tot;
                int src = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    int p = src + step;
                    uint32_t e = __shfl(excl, p & 63);
                    if (p < 64 && e <= t) src = p;
                }
                const uint32_t slo = __shfl(lo, src);
                const uint32_t sex = __shfl(excl, src);
                Ent c = ent_inf();
                if (valid) {
                    const uint32_t j = slo + (t - sex);
                    const double dx = qx - cv.sx[j];
                    const double dy = qy - cv.sy[j];
                    const double dz = qz - cv.sz[j];
                    double d2 = dx * dx;
                    d2 = d2 + dy * dy;
                    d2 = d2 + dz * dz;
                    c.d = static_cast<uint64_t>(__double_as_longlong(d2));
                    c.id = cv.perm[j];
                    c.j = j;
                }
                const Ent worst = ent_shfl(best, fetch);
                const bool surv = valid && ent_less(c, worst);
                if (__ballot(surv)) {
                    cnt[2] += 1;
                    if (!surv) c = ent_inf();
                    bitonic_sort64(c, lane);
                    merge_sorted(best, c, lane);
                }
            }
        }
        // Certification against the faces of the box that border cells of
        // the grid (faces on or beyond the grid's edge bound nothing).
        const Ent kth = ent_shfl(best, fetch - 1);
        double R = INFINITY;
        if (static_cast<int>(qcx) - r > 0)
            R = fmin(R, qx - (g.ox + static_cast<double>(static_cast<int>(qcx) - r) * g.c));
        if (static_cast<int>(qcx) + r < static_cast<int>(g.gx) - 1)
            R = fmin(R, (g.ox + static_cast<double>(static_cast<int>(qcx) + r + 1) * g.c) - qx);
        if (static_cast<int>(qcy) - r > 0)
            R = fmin(R, qy - (g.oy + static_cast<double>(static_cast<int>(qcy) - r) * g.c));
        if (static_cast<int>(qcy) + r < static_cast<int>(g.gy) - 1)
            R = fmin(R, (g.oy + static_cast<double>(static_cast<int>(qcy) + r + 1) * g.c) - qy);
        if (za > 0)
            R = fmin(R, qz - (g.oz + static_cast<double>(za) * g.cz));
        if (zb < static_cast<int>(g.gz) - 1)
            R = fmin(R, (g.oz + static_cast<double>(zb + 1) * g.cz) - qz);
        bool certified;
        double kd2 = __longlong_as_double(static_cast<long long>(kth.d));
        if (R == INFINITY) {
            certified = true;
        } else if (kth.d != kInfD) {
            double Rm = R - a.margin;
            certified = Rm > 0.0 && kd2 < Rm * Rm;
        } else {
            certified = false;
        }
        if (certified) return;
        if (r >= a.tune.ring_max) {
            incomplete = true;
            return;
        }
        int rn = r + 1;
        if (kth.d != kInfD) {
            double need = sqrt(kd2) + a.margin;
            double rr = ceil(need / g.c);
            if (rr > rn) rn = rr > 1e6 ? 1000000 : static_cast<int>(rr);
        } else {
            rn = max(rn, 2 * r);
        }
        r = min(rn, a.tune.ring_max);
    }
}

__global__ __launch_bounds__(kBlock) void k_knn(KnnArgs a)
{
    extern __shared__ uint32_t s_nbr[];
    const int lane = __lane_id();
    const int w = threadIdx.x / kWave;
    const CloudView& cv = a.cv;
    const uint64_t n = cv.n;
    const int Q = a.q_per_wave;
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint64_t q0 = (static_cast<uint64_t>(lb) * kWaves + w) * Q;
    if (q0 >= n) return;
    const int nq = static_cast<int>(n - q0 < static_cast<uint64_t>(Q) ? n - q0 : Q);
    const int fetch = a.prog.fetch;
    const int k = a.prog.k;
    const int stride = a.prog.stride;
    uint32_t* nbr = s_nbr + static_cast<size_t>(w) * Q * a.lds_stride;

    uint8_t my_flags = 0;
    uint64_t cnt[3] = {0, 0, 0};  // search passes, candidates, merges
    for (int qi = 0; qi < nq; ++qi) {
        const uint64_t s = q0 + qi;
        const double qx = cv.sx[s], qy = cv.sy[s], qz = cv.sz[s];
        const uint32_t qcx = cell_coord(qx, cv.g.ox, cv.g.inv_c, cv.g.gx);
        const uint32_t qcy = cell_coord(qy, cv.g.oy, cv.g.inv_c, cv.g.gy);
        const uint32_t qcz = cell_coord(qz, cv.g.oz, cv.g.inv_cz, cv.g.gz);
        Ent best;
        bool incomplete;
        wave_knn(a, qx, qy, qz, qcx, qcy, qcz, lane, best, incomplete, cnt);

        // Boundary tie: the fetch-th and (fetch+1)-th distances are equal.
        // With stride > 1 every adjacent equal pair inside the fetched list
        // can change which ids are kept, so any such pair flags the row.
        const Ent nxt = ent_shfl(best, lane + 1 < 64 ? lane + 1 : 63);
        bool eq = (lane < fetch) && (lane + 1 < 64) && best.d == nxt.d && best.d != kInfD;
        bool tie = stride > 1 ? (__ballot(eq) != 0) : (__shfl(static_cast<int>(eq), fetch - 1) != 0);
        uint8_t flags = (tie ? ST_TIE : 0) | (incomplete ? ST_INCOMPLETE : 0);

        // Kept list in canonical (d2, id) order: entries 0, stride, 2*stride...
        const Ent kept = ent_shfl(best, (lane * stride) & 63);
        if (a.mode == 1) {
            const uint32_t qid = cv.perm[s];
            if (lane < k) a.out.ids[static_cast<uint64_t>(qid) * k + lane] = kept.id;
            if (lane == 0) a.out.status[0][qid] = flags;
            if (lane == 0 && flags) {
                uint32_t slot = atomicAdd(a.out.nflagged, 1u);
                a.out.flagged[slot] = static_cast<uint32_t>(s);
            }
        } else {
            // An uncertified row can hold fewer than fetch entries; keep its
            // list in bounds (the exact search replaces the row anyway).
            if (lane < k) nbr[qi * a.lds_stride + lane] = kept.j == kInfU ? 0u : kept.j;
            if (lane == qi) my_flags = flags;
        }
    }
    if (a.out.counters && lane == 0) {
        atomicAdd(a.out.counters + 0, static_cast<unsigned long long>(cnt[0]));
        atomicAdd(a.out.counters + 1, static_cast<unsigned long long>(cnt[1]));
        atomicAdd(a.out.counters + 2, static_cast<unsigned long long>(cnt[2]));
        atomicAdd(a.out.counters + 3, static_cast<unsigned long long>(nq));
    }
    if (a.mode == 1) return;
    // The lists were written by other lanes of this same wave: order those
    // LDS stores before the loads below (no block barrier: waves of this
    // block may already have exited).
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();

    // One query per lane: FP64 covariance of its neighbor list (two-pass
    // mean), stored in sorted order for the eigen/feature kernel.
    if (lane >= nq) return;
    const uint64_t s = q0 + lane;
    const uint32_t* list = nbr + lane * a.lds_stride;
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (int l = 0; l < k; ++l) {
        const uint32_t j = list[l];
        mx += cv.sx[j];
        my += cv.sy[j];
        mz += cv.sz[j];
    }
    const double inv_k = 1.0 / static_cast<double>(k);
    mx *= inv_k;
    my *= inv_k;
    mz *= inv_k;
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (int l = 0; l < k; ++l) {
        const uint32_t j = list[l];
        const double dx = cv.sx[j] - mx, dy = cv.sy[j] - my, dz = cv.sz[j] - mz;
        c[0] += dx * dx;
        c[1] += dx * dy;
        c[2] += dx * dz;
        c[3] += dy * dy;
        c[4] += dy * dz;
        c[5] += dz * dz;
    }
    const double den = static_cast<double>(k - 1);
    for (int i = 0; i < 6; ++i) a.out.cov[static_cast<uint64_t>(i) * n + s] = c[i] / den;
    a.out.flags[s] = my_flags;
    if (my_flags) {
        uint32_t slot = atomicAdd(a.out.nflagged, 1u);
        a.out.flagged[slot] = static_cast<uint32_t>(s);
    }
}

// Per sorted row: eigensystem of the stored covariance and every requested
// stage column, written at the row's original index.
__global__ __launch_bounds__(256) void k_features(const double* __restrict__ cov,
                                                  const uint8_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ perm,
                                                  uint64_t n, FusedProgram prog,
                                                  KnnOutputs out)
{
    const uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    double c[6];
    for (int i = 0; i < 6; ++i) c[i] = cov[static_cast<uint64_t>(i) * n + s];
    const uint8_t f = flags[s];
    const uint32_t qid = perm[s];
    const bool zero = cov_is_zero(c);
    double ev[3] = {0, 0, 0}, vec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int info = 0;
    if (!zero) info = eigen_sym3(c, ev, vec);
    for (int si = 0; si < prog.nstages; ++si) {
        const StageProgram& sp = prog.st[si];
        double vals[10];
        uint8_t st = stage_outputs(sp, prog.k, c, zero, info, ev, vec, vals);
        if (st == 0)
            for (int cc = 0; cc < sp.ncols; ++cc) out.col[sp.col_offset + cc][qid] = vals[cc];
        out.status[si][qid] = st | f;
    }
}

__global__ __launch_bounds__(256) void k_eigen_batch(const double* __restrict__ cov,
                                                     uint64_t m, double* __restrict__ ev,
                                                     double* __restrict__ vec,
                                                     int32_t* __restrict__ info)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double* a = cov + 9 * i;
    double c[6] = {a[0], a[3], a[6], a[4], a[7], a[8]};
    double e[3], v[9];
    info[i] = eigen_sym3(c, e, v);
    for (int t = 0; t < 3; ++t) ev[3 * i + t] = e[t];
    for (int t = 0; t < 9; ++t) vec[9 * i + t] = v[t];
}

} // namespace

void launch_knn(const CloudView& cv, const FusedProgram& prog, int mode,
                const KnnOutputs& out, const KnnTuning& tune, hipStream_t s)
{
    if (cv.n == 0) return;
    KnnArgs a;
    a.cv = cv;
    a.prog = prog;
    a.out = out;
    a.tune = tune;
    a.mode = mode;
    a.q_per_wave = prog.k <= 32 ? 64 : 32;
    a.lds_stride = prog.k | 1;  // odd stride: lanes hit distinct LDS banks
    const GridParams& g = cv.g;
    a.margin = 1e-6 * fmin(g.c, g.cz) +
               1e-15 * (fabs(g.ox) + fabs(g.oy) + fabs(g.oz)) + 1e-300;
    const uint64_t per_block = static_cast<uint64_t>(kWaves) * a.q_per_wave;
    const uint32_t blocks = div_up(cv.n, per_block);
    const size_t lds = mode == 1 ? 0
                                 : static_cast<size_t>(kWaves) * a.q_per_wave *
                                       a.lds_stride * sizeof(uint32_t);
    hipLaunchKernelGGL(k_knn, dim3(blocks), dim3(kBlock), lds, s, a);
}

void launch_features(const CloudView& cv, const FusedProgram& prog,
                     const KnnOutputs& out, hipStream_t s)
{
    if (cv.n == 0) return;
    hipLaunchKernelGGL(k_features, dim3(div_up(cv.n, 256)), dim3(256), 0, s, out.cov,
                       out.flags, cv.perm, cv.n, prog, out);
}

void launch_eigen_batch(const double* cov, uint64_t m, double* evals,
                        double* evecs, int32_t* info, hipStream_t s)
{
    if (!m) return;
    hipLaunchKernelGGL(k_eigen_batch, dim3(div_up(m, 256)), dim3(256), 0, s, cov, m,
                       evals, evecs, info);
}

} // namespace gf
