// knn.hip -- wavefront-per-query k-nearest-neighbor search over the column
// grid, fused with the neighborhood covariance; plus the per-row eigensolve
// and stage-output kernel.
//
// Replaces the reference's thread-per-query knnGatherKernel with its fused
// covariance, the device eigen kernel and the projection kernels
// (src/index/SpatialIndexKernels.cu:716-1086, 1716-1769;
// src/stages/NeighborhoodKernels.cu:67-119, 422-485).
//
// Work split.  A wave owns Q consecutive sorted points (queries).  Sorted
// order is (column, z cell), so consecutive queries form groups that share a
// column and a short z range; every query of a group has the same candidate
// box (the 3x3 columns around it, z cells [first-1, last+1]).  The wave loads
// the box once -- column starts, z-window bounds, the candidates' FP64
// coordinates and ids -- into registers (up to 256 candidates, 4 per lane),
// then for each query of the group computes the FP64 squared distances
// exactly as KD3Index does (d0*d0 + d1*d1 + d2*d2, left to right, no FMA)
// and selects the fetch+1 smallest (d2, point id) pairs:  a bisection on d2
// counted with ballots finds a threshold that keeps between fetch+1 and 64
// candidates, which are compacted through LDS and sorted once by a bitonic
// network across the lanes.
//
// Certification.  A query's list is exact when its fetch-th distance is below
// the distance from the query to the nearest face of the examined box that
// borders unexamined grid cells (minus a rounding margin).  Otherwise the
// query alone is searched again with a growing box (streaming candidates
// through sort-and-merge) up to a budget; a query still uncertified is
// flagged incomplete.  Rows whose fetch-th and (fetch+1)-th distances are
// equal have a boundary tie, which KD3Index breaks by tree visit order rather
// than by index; they are flagged for the exact KD3 search.
//
// Covariance.  Each query's kept list is parked in LDS; after the wave's
// queries every lane takes one of them and forms its FP64 covariance
// (two-pass mean), written in sorted order for the eigen/feature kernel.
#include "gf_eigen.hpp"
#include "gf_internal.hpp"
#include "gf_kernels.hpp"
#include "gf_wave.hpp"

namespace gf {

namespace {

// Waves per workgroup.  One cloud: four, so the tree built beside the search
// finds the slots a workgroup's early waves leave idle until its last wave
// ends.  Tiles of a batch: one; a wave's slot and LDS free as soon as it ends
// (measured: the batch's search 6% faster, the step 7% faster).
// Waves per workgroup: one for the tiled (batch) search, 1, 2 or 4 (tuning
// "knn_waves", default 4) for a single cloud.
constexpr int kKnnWavesDefault = 4;
constexpr int kRegBatches = 4;
constexpr uint32_t kRegCand = kRegBatches * 64;
// Per-wave LDS scratch: 64 compacted entries (16 B), 64 ranked entries, and a
// 64-entry list of queries left for the growing search.
constexpr int kScratchWords = 256 + 256 + 64;

// First positions in a[lo, hi) (sorted ascending) of keys >= k0 and of keys
// >= k1 (k0 <= k1): two binary searches advanced together, so that their
// loads are issued side by side (half the dependent load chain of running
// them one after the other).
__device__ __forceinline__ void lower_bound2_u32(const uint32_t* __restrict__ a, uint32_t lo,
                                                 uint32_t hi, uint32_t k0, uint32_t k1,
                                                 uint32_t& r0, uint32_t& r1)
{
    uint32_t lo0 = lo, hi0 = hi, lo1 = lo, hi1 = hi;
    while (lo0 < hi0 || lo1 < hi1) {
        const uint32_t m0 = lo0 + ((hi0 - lo0) >> 1), m1 = lo1 + ((hi1 - lo1) >> 1);
        const uint32_t v0 = lo0 < hi0 ? a[m0] : 0u;
        const uint32_t v1 = lo1 < hi1 ? a[m1] : 0u;
        if (lo0 < hi0) {
            if (v0 < k0)
                lo0 = m0 + 1;
            else
                hi0 = m0;
        }
        if (lo1 < hi1) {
            if (v1 < k1)
                lo1 = m1 + 1;
            else
                hi1 = m1;
        }
    }
    r0 = lo0;
    r1 = lo1;
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
    CloudView cv;       // tiled: global arrays; colstart of all tiles
    FusedProgram prog;  // tiled: replaced per wave by the tile's
    KnnOutputs out;
    KnnTuning tune;
    int32_t mode;     // 0 features, 1 neighbor lists, 2 lists in KD3Index order
    int32_t q_per_wave;
    int32_t lds_stride;
    double margin;
    const TileInfo* tiles;   // tiled search only
    const uint32_t* task0;   // tiled: per-tile first task (ntiles + 1)
    uint32_t ntasks;
    uint32_t ntiles;
};

// What a wavefront needs of its cloud or tile: the grid, the column starts,
// the certification margin and the neighborhood sizes.
struct WaveGrid {
    GridParams g;
    const uint32_t* colstart;
    double margin;
    int fetch, k, stride;
};

// The examined region: columns [x0, x1] x [y0, y1], z cells [z0, z1].
struct Box {
    int x0, x1, y0, y1, z0, z1;
};

// Distance from the query to the nearest face of the box that borders grid
// cells outside it (faces on the grid's edge bound nothing: +inf).
__device__ __forceinline__ double box_clearance(const GridParams& g, const Box& b,
                                                double qx, double qy, double qz)
{
    double R = INFINITY;
    if (b.x0 > 0) R = fmin(R, qx - (g.ox + static_cast<double>(b.x0) * g.c));
    if (b.x1 < static_cast<int>(g.gx) - 1)
        R = fmin(R, (g.ox + static_cast<double>(b.x1 + 1) * g.c) - qx);
    if (b.y0 > 0) R = fmin(R, qy - (g.oy + static_cast<double>(b.y0) * g.c));
    if (b.y1 < static_cast<int>(g.gy) - 1)
        R = fmin(R, (g.oy + static_cast<double>(b.y1 + 1) * g.c) - qy);
    if (b.z0 > 0) R = fmin(R, qz - (g.oz + static_cast<double>(b.z0) * g.cz));
    if (b.z1 < static_cast<int>(g.gz) - 1)
        R = fmin(R, (g.oz + static_cast<double>(b.z1 + 1) * g.cz) - qz);
    return R;
}

__device__ __forceinline__ bool certified(const WaveGrid& wg, const Box& b, const Ent& kth,
                                          double qx, double qy, double qz)
{
    const double R = box_clearance(wg.g, b, qx, qy, qz);
    if (R == INFINITY) return true;
    if (kth.d == kInfD) return false;
    const double Rm = R - wg.margin;
    return Rm > 0.0 && ent_d2(kth) < Rm * Rm;
}

// Column ranges of a box, one column per lane (box columns must be <= 64):
// lo / len of the column's sorted points restricted to z cells [z0, z1]
// (whole columns when the box spans every z cell or the column is small).
__device__ __forceinline__ void box_ranges(const KnnArgs& a, const WaveGrid& wg, const Box& b,
                                           int cb, int lane, uint32_t& lo, uint32_t& len)
{
    const CloudView& cv = a.cv;
    const GridParams& g = wg.g;
    const int w = b.x1 - b.x0 + 1;
    const int ncb = w * (b.y1 - b.y0 + 1);
    const int ci = cb + lane;
    lo = 0;
    len = 0;
    if (ci >= ncb) return;
    const int cx = b.x0 + ci % w;
    const int cy = b.y0 + ci / w;
    if (cx < 0 || cy < 0 || cx >= static_cast<int>(g.gx) || cy >= static_cast<int>(g.gy))
        return;
    const bool zfull = b.z0 <= 0 && b.z1 >= static_cast<int>(g.gz) - 1;
    const uint32_t col = column_index(g, cx, cy);
    uint32_t b0 = wg.colstart[col], b1 = wg.colstart[col + 1];
    if (!zfull && b1 - b0 > static_cast<uint32_t>(a.tune.small_col)) {
        const uint32_t klo = (col << g.bz) + static_cast<uint32_t>(max(b.z0, 0));
        const uint32_t khi = (col << g.bz) + static_cast<uint32_t>(min(b.z1, static_cast<int>(g.gz) - 1) + 1);
        lower_bound2_u32(cv.keys, b0, b1, klo, khi, b0, b1);
    }
    lo = b0;
    len = b1 - b0;
}

// Sorted position of candidate t of the packed ranges: the lane whose range
// holds t is the last lane with excl <= t.
__device__ __forceinline__ uint32_t candidate_pos(uint32_t lo, uint32_t excl, uint32_t t)
{
    int src = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        int p = src + step;
        uint32_t e = __shfl(excl, p & 63);
        if (p < 64 && e <= t) src = p;
    }
    const uint32_t slo = __shfl(lo, src);
    const uint32_t sex = __shfl(excl, src);
    return slo + (t - sex);
}

// Bit pattern of the FP64 squared distance, in KD3Index's operation order.
__device__ __forceinline__ uint64_t dist_bits(double qx, double qy, double qz, double px,
                                              double py, double pz)
{
    const double dx = qx - px;
    const double dy = qy - py;
    const double dz = qz - pz;
    double d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    return static_cast<uint64_t>(__double_as_longlong(d2));
}

__device__ __forceinline__ double bits_d2(uint64_t d)
{
    return __longlong_as_double(static_cast<long long>(d));
}

__device__ __forceinline__ Ent make_ent(double qx, double qy, double qz, double px,
                                        double py, double pz, uint32_t id, uint32_t j)
{
    const double dx = qx - px;
    const double dy = qy - py;
    const double dz = qz - pz;
    double d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    return Ent{static_cast<uint64_t>(__double_as_longlong(d2)), id, j};
}

// 1 when the compacted entry e = (d2 bits low, high, id, payload) precedes
// (d, id) in (d2, id) order.  Evaluated without branches, so that unrolled
// iterations issue their LDS loads together (short-circuit forms compile to
// one conditional load, and one wait, per entry).  Non-negative doubles order
// as their bit patterns.
__device__ __forceinline__ uint32_t rank_step(const uint4 e, uint64_t d, uint32_t id)
{
    const uint64_t od = static_cast<uint64_t>(e.x) | (static_cast<uint64_t>(e.y) << 32);
    return static_cast<uint32_t>(od < d) |
           (static_cast<uint32_t>(od == d) & static_cast<uint32_t>(e.z < id));
}

// Selection over at most 256 candidate entries held in registers (4 per
// lane).  First a threshold E with need <= #{d2 <= E} <= cap (need = fetch +
// 1, so the (fetch+1)-th entry that decides ties is kept), each trial counted
// exactly by ballots: the caller's estimate (the previous query's need-th d2)
// first, then two steps that interpolate the count linearly in d2 (points on
// a surface: count grows with the area, i.e. with d2) inside the bracket, then
// bisection.  The survivors are compacted into LDS and ranked: each
// lane counts the survivors that precede its own entry in (d2, id) order,
// reading them as LDS broadcasts, and stores itself at that rank.  Every
// candidate with d2 equal to the boundary distance is below E, so the first
// fetch+1 ranked survivors are exactly the fetch+1 smallest (d2, id) pairs.
// When no such E exists (more than 64 points at one distance) the register
// batches are merged one by one with bitonic networks instead.
//
// The lane's candidate of register batch b is (cd[b], cid[b], cj[b]): squared
// distance bits (kInfD: none), point id and sorted position.
__device__ __forceinline__ void select_entries(const uint64_t (&cd)[kRegBatches],
                                               const uint32_t (&cid)[kRegBatches],
                                               const uint32_t (&cj)[kRegBatches], uint32_t tot,
                                               int fetch, int lane, uint32_t* scratch,
                                               double hint, Ent& best, uint64_t* cnt)
{
    const uint32_t need = min(static_cast<uint32_t>(fetch + 1), tot);
    // Survivor window: few enough to rank cheaply, at most one per lane.
    const uint32_t cap = min(need + 15u, 64u);
    bool ok = true;
    double E = INFINITY;
    cnt[2] += 1;
    if (tot > cap) {
        const uint32_t nbat = (tot + 63u) >> 6;  // register batches holding candidates
        auto count_le = [&](double t) {
            uint32_t k = 0;
#pragma unroll
            for (int b = 0; b < kRegBatches; ++b)
                if (static_cast<uint32_t>(b) < nbat)
                    k += __popcll(__ballot(cd[b] != kInfD && bits_d2(cd[b]) <= t));
            return k;
        };
        double lo_e = 0.0, hi_e = INFINITY;
        ok = false;
        if (hint < INFINITY) {
            const uint32_t k = count_le(hint);
            if (k >= need && k <= cap) {
                E = hint;
                ok = true;
            } else if (k > cap) {
                hi_e = hint;
            } else {
                lo_e = hint;
            }
        }
        if (!ok && hi_e == INFINITY) {
            double mx = 0.0;
#pragma unroll
            for (int b = 0; b < kRegBatches; ++b)
                if (static_cast<uint32_t>(b) < nbat && cd[b] != kInfD) mx = fmax(mx, bits_d2(cd[b]));
#pragma unroll
            for (int m = 32; m > 0; m >>= 1) mx = fmax(mx, __shfl_xor(mx, m));
            hi_e = mx;  // counts every candidate (> cap)
        }
        for (int it = 0; it < 80 && !ok; ++it) {
            const double mid = 0.5 * (lo_e + hi_e);
            const uint32_t k = count_le(mid);
            if (k < need) {
                if (!(mid > lo_e)) break;
                lo_e = mid;
            } else if (k > cap) {
                if (!(mid < hi_e)) break;
                hi_e = mid;
            } else {
                E = mid;
                ok = true;
            }
        }
    }
    if (ok) {
        uint32_t S = 0;
        const uint32_t nbat_s = (tot + 63u) >> 6;
#pragma unroll
        for (int b = 0; b < kRegBatches; ++b) {
            if (static_cast<uint32_t>(b) >= nbat_s) break;
            const bool s = cd[b] != kInfD && bits_d2(cd[b]) <= E;
            const uint64_t m = __ballot(s);
            if (s) {
                // Survivors of lower lanes in this batch (mbcnt: popcount of
                // the ballot below this lane).
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                    static_cast<uint32_t>(m >> 32),
                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
                uint32_t* e = scratch + 4 * (S + below);
                e[0] = static_cast<uint32_t>(cd[b]);
                e[1] = static_cast<uint32_t>(cd[b] >> 32);
                e[2] = cid[b];
                e[3] = cj[b];
            }
            S += __popcll(m);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        Ent v = ent_inf();
        uint32_t rank = 0;
        if (static_cast<uint32_t>(lane) < S) {
            const uint32_t* e = scratch + 4 * lane;
            v.d = static_cast<uint64_t>(e[0]) | (static_cast<uint64_t>(e[1]) << 32);
            v.id = e[2];
            v.j = e[3];
        }
#pragma unroll 4
        for (uint32_t i = 0; i < S; ++i)
            rank += rank_step(reinterpret_cast<const uint4*>(scratch)[i], v.d, v.id);
        uint32_t* sorted = scratch + 256;
        if (static_cast<uint32_t>(lane) < S) {
            uint32_t* e = sorted + 4 * rank;
            e[0] = static_cast<uint32_t>(v.d);
            e[1] = static_cast<uint32_t>(v.d >> 32);
            e[2] = v.id;
            e[3] = v.j;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        best = ent_inf();
        if (static_cast<uint32_t>(lane) < S) {
            const uint32_t* e = sorted + 4 * lane;
            best.d = static_cast<uint64_t>(e[0]) | (static_cast<uint64_t>(e[1]) << 32);
            best.id = e[2];
            best.j = e[3];
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        return;
    }
    best = ent_inf();
#pragma unroll
    for (int b = 0; b < kRegBatches; ++b) {
        if (static_cast<uint32_t>(b * 64) >= tot) break;
        Ent c = cd[b] != kInfD ? Ent{cd[b], cid[b], cj[b]} : ent_inf();
        bitonic_sort64(c, lane);
        merge_sorted(best, c, lane);
        cnt[2] += 1;
    }
}

// Candidates whose d2 bits are <= t (below t when kStrict); kInfD entries
// never count.  Non-negative doubles order as their bit patterns.
template <bool kStrict>
__device__ __forceinline__ uint32_t count_bits(const uint64_t (&cd)[kRegBatches], uint32_t nbat,
                                               uint64_t t)
{
    uint32_t c = 0;
#pragma unroll
    for (int b = 0; b < kRegBatches; ++b)
        if (static_cast<uint32_t>(b) < nbat)
            c += __popcll(__ballot(kStrict ? cd[b] < t : cd[b] <= t));
    return c;
}

__device__ __forceinline__ uint64_t d2_bits(double d)
{
    return static_cast<uint64_t>(__double_as_longlong(d));
}

// Exact-count selection, for neighborhoods kept as a set (features, stride
// 1): a threshold E with exactly `fetch` candidates at d2 <= E.  Such an E
// exists exactly when the fetch-th and (fetch+1)-th smallest distances
// differ (no boundary tie), and the kept set is then {d2 <= E} whatever the
// order among equal distances inside it.  Counts are exact (ballots).  First
// round: kRound thresholds around the hint (the previous query's E) counted
// together; one with count fetch is E, else the two around fetch bracket it,
// and when few candidates lie inside the bracket each ranks itself among
// them and the one at the wanted rank, checked by its count, is E.  Failing
// that: linear interpolation of the count in d2 inside the bracket, then
// candidates strictly inside it as pivots.  A bracket with no candidate
// strictly inside means the fetch-th and (fetch+1)-th distances are equal:
// false, and the caller takes the sorting selection (as when the trials run
// out).
//
// Thresholds of the first round, as factors of the hint (for points on a
// surface the count grows about linearly with d2).
constexpr int kRound = 7;

__device__ __forceinline__ float round_factor(int i)
{
    constexpr float f[kRound] = {0.6f, 0.75f, 0.87f, 1.0f, 1.15f, 1.33f, 1.6f};
    return f[i];
}

__device__ __forceinline__ bool select_count(const uint64_t (&cd)[kRegBatches], uint32_t nbat,
                                             uint32_t fetch, double hint, uint64_t& E)
{
    const double h = uniform_f64(hint);
    // First round: counts at kRound thresholds around the hint, independent
    // of each other (one latency for all of them).
    uint64_t tb[kRound];
    uint32_t c[kRound];
#pragma unroll
    for (int i = 0; i < kRound; ++i) {
        tb[i] = d2_bits(h * static_cast<double>(round_factor(i)));
        c[i] = 0;
    }
#pragma unroll
    for (int b = 0; b < kRegBatches; ++b) {
        if (static_cast<uint32_t>(b) >= nbat) break;
#pragma unroll
        for (int i = 0; i < kRound; ++i) c[i] += __popcll(__ballot(cd[b] <= tb[i]));
    }
    // The bracket: the last threshold below fetch, the first above.
    int ilo = -1, ihi = kRound;
#pragma unroll
    for (int i = 0; i < kRound; ++i) {
        if (c[i] == fetch) {
            E = tb[i];
            return true;
        }
        if (c[i] < fetch) ilo = i;
        if (c[i] > fetch && ihi == kRound) ihi = i;
    }
    double lo = 0.0, hi = INFINITY;
    uint32_t clo = 0, chi = 0;
    uint64_t lob = 0, hib = kInfD;
#pragma unroll
    for (int i = 0; i < kRound; ++i) {
        if (i == ilo) {
            lob = tb[i];
            clo = c[i];
        }
        if (i == ihi) {
            hib = tb[i];
            chi = c[i];
        }
    }
    if (ilo >= 0) lo = bits_d2(lob);
    if (ihi < kRound) hi = bits_d2(hib);
    if (ihi < kRound && chi - clo <= 16u) {
        // Few candidates inside the bracket: each ranks itself among them
        // (strictly smaller ones), and the one at rank fetch - clo - 1 is the
        // fetch-th distance when it is alone there; checked by its count.
        uint64_t mb[kRegBatches];
        uint32_t rk[kRegBatches];
#pragma unroll
        for (int b = 0; b < kRegBatches; ++b) {
            rk[b] = 0;
            mb[b] = static_cast<uint32_t>(b) < nbat
                        ? __ballot((ilo < 0 || cd[b] > lob) && cd[b] <= hib)
                        : 0ull;
        }
#pragma unroll
        for (int b = 0; b < kRegBatches; ++b) {
            uint64_t m = mb[b];
            while (m) {
                const int l = __builtin_ctzll(m);
                m &= m - 1;
                const uint64_t v = lane_u64(cd[b], l);
#pragma unroll
                for (int b2 = 0; b2 < kRegBatches; ++b2)
                    if (static_cast<uint32_t>(b2) < nbat) rk[b2] += v < cd[b2] ? 1u : 0u;
            }
        }
        const uint32_t want = fetch - clo - 1;
        uint64_t pick = kInfD;
        uint32_t npick = 0;
#pragma unroll
        for (int b = 0; b < kRegBatches; ++b) {
            if (static_cast<uint32_t>(b) >= nbat) break;
            const uint64_t m = mb[b] & __ballot(rk[b] == want);
            npick += static_cast<uint32_t>(__popcll(m));
            if (m && pick == kInfD) pick = lane_u64(cd[b], __builtin_ctzll(m));
        }
        if (npick == 1) {
                    if (count_bits<false>(cd, nbat, pick) == fetch) {
                E = pick;
                return true;
            }
        }
    }
    // Otherwise: linear interpolation of the count in d2 inside the bracket,
    // then candidates strictly inside the bracket as pivots.  When the
    // bracket holds no candidate strictly inside, the fetch-th and
    // (fetch+1)-th distances are equal: false.
    double t;
    if (hi == INFINITY) {
        t = (ilo >= 0 ? lo : h) * static_cast<double>((static_cast<float>(fetch) + 0.5f) /
                                                    static_cast<float>(clo > 0 ? clo : 1));
    } else {
        const float f = (static_cast<float>(fetch - clo) + 0.5f) / static_cast<float>(chi - clo);
        t = lo + (hi - lo) * static_cast<double>(f);
    }
#pragma unroll 1
    for (int trial = 0; trial < 16; ++trial) {
        if (trial >= 3 || !(t > lo && t < hi)) {
            // Pivot: the first candidate strictly inside (lo, hi).
            const uint64_t lb = d2_bits(lo), hb = d2_bits(hi);
            uint64_t piv = kInfD;
#pragma unroll
            for (int b = 0; b < kRegBatches; ++b) {
                if (static_cast<uint32_t>(b) >= nbat || piv != kInfD) continue;
                const uint64_t m = __ballot(cd[b] > lb && cd[b] < hb);
                if (m) piv = lane_u64(cd[b], __builtin_ctzll(m));
            }
            if (piv == kInfD) return false;
            t = bits_d2(piv);
        }
        const uint64_t tbt = d2_bits(t);
            const uint32_t ct = count_bits<false>(cd, nbat, tbt);
        if (ct == fetch) {
            E = tbt;
            return true;
        }
        if (ct < fetch) {
            lo = t;
            clo = ct;
        } else {
            hi = t;
            chi = ct;
        }
        if (hi == INFINITY) {
            // Count ~ d2 through the origin.
            t = lo * static_cast<double>((static_cast<float>(fetch) + 0.5f) /
                                         static_cast<float>(clo > 0 ? clo : 1));
        } else {
            const float f = (static_cast<float>(fetch - clo) + 0.5f) /
                            static_cast<float>(chi - clo);
            t = lo + (hi - lo) * static_cast<double>(f);
        }
    }
    return false;
}

// Lane l's value of v, for every lane (l wave-uniform).
__device__ __forceinline__ double lane_value(double v, int l)
{
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(b), l);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(b >> 32), l);
    return __longlong_as_double(
        static_cast<long long>(static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32)));
}

__device__ __forceinline__ uint32_t lanes_below64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// The n (<= 64) entries waiting in scratch, sorted and merged into best.
__device__ __forceinline__ void stream_flush(const uint32_t* scratch, uint32_t n, int lane,
                                             Ent& best)
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    Ent c = ent_inf();
    if (static_cast<uint32_t>(lane) < n) {
        const uint32_t* e = scratch + 4 * lane;
        c.d = static_cast<uint64_t>(e[0]) | (static_cast<uint64_t>(e[1]) << 32);
        c.id = e[2];
        c.j = e[3];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    bitonic_sort64(c, lane);
    merge_sorted(best, c, lane);
}

// One query alone, with a growing box starting at half-width r (columns and
// z cells around the query's cell): candidates stream through the registers
// (<= 256, one box of <= 64 columns) or through sort-and-merge otherwise.
__device__ __forceinline__ void search_single(const KnnArgs& a, const WaveGrid& wg, double qx,
                              double qy, double qz, int qcx, int qcy, int qcz, int r, int lane,
                              uint32_t* scratch, Ent& best, bool& incomplete,
                              uint64_t* cnt)
{
    const CloudView& cv = a.cv;
    const int fetch = wg.fetch;
    incomplete = false;
    uint32_t nbuf = 0;  // streamed candidates waiting in LDS
    for (;;) {
        cnt[0] += 1;
        const Box b{qcx - r, qcx + r, qcy - r, qcy + r, qcz - r, qcz + r};
        const int ncb = (2 * r + 1) * (2 * r + 1);
        best = ent_inf();
        for (int cb = 0; cb < ncb; cb += 64) {
            uint32_t lo, len;
            box_ranges(a, wg, b, cb, lane, lo, len);
            const uint32_t incl = wave_incl_scan_u32(len, lane);
            const uint32_t tot = lane_u32(incl, 63);
            const uint32_t excl = incl - len;
            cnt[1] += tot;
            if (ncb <= 64 && tot <= kRegCand) {
                uint64_t cd[kRegBatches];
                uint32_t cid[kRegBatches], cj[kRegBatches];
#pragma unroll
                for (int bb = 0; bb < kRegBatches; ++bb) {
                    cd[bb] = kInfD;
                    cid[bb] = cj[bb] = kInfU;
                    if (static_cast<uint32_t>(bb * 64) < tot) {
                        const uint32_t t = bb * 64 + lane;
                        const uint32_t j = candidate_pos(lo, excl, t);
                        if (t < tot) {
                            cd[bb] = dist_bits(qx, qy, qz, cv.sx[j], cv.sy[j], cv.sz[j]);
                            cid[bb] = cv.perm[j];
                            cj[bb] = j;
                        }
                    }
                }
                select_entries(cd, cid, cj, tot, fetch, lane, scratch, INFINITY, best, cnt);
                break;
            }
            // Streamed: candidates below the current (fetch+1)-th entry are
            // collected in LDS and sorted into the list 64 at a time.  That
            // entry only decreases, so a candidate not below it when it
            // is seen cannot be among the final fetch+1.
            for (uint32_t t0 = 0; t0 < tot; t0 += 64) {
                const uint32_t t = t0 + lane;
                const uint32_t j = candidate_pos(lo, excl, t);
                Ent c = ent_inf();
                if (t < tot)
                    c = make_ent(qx, qy, qz, cv.sx[j], cv.sy[j], cv.sz[j], cv.perm[j], j);
                const Ent worst = ent_lane(best, fetch);
                const bool surv = t < tot && ent_less(c, worst);
                const uint64_t m = __ballot(surv);
                if (!m) continue;
                const uint32_t ns = static_cast<uint32_t>(__popcll(m));
                if (nbuf + ns > 64u) {
                    stream_flush(scratch, nbuf, lane, best);
                    cnt[2] += 1;
                    nbuf = 0;
                }
                if (surv) {
                    uint32_t* e = scratch + 4 * (nbuf + lanes_below64(m));
                    e[0] = static_cast<uint32_t>(c.d);
                    e[1] = static_cast<uint32_t>(c.d >> 32);
                    e[2] = c.id;
                    e[3] = c.j;
                }
                nbuf += ns;
            }
        }
        if (nbuf) {
            stream_flush(scratch, nbuf, lane, best);
            cnt[2] += 1;
            nbuf = 0;
        }
        const Ent kth = ent_lane(best, fetch - 1);
        if (certified(wg, b, kth, qx, qy, qz)) return;
        if (r >= a.tune.ring_max) {
            incomplete = true;
            return;
        }
        int rn = r + 1;
        if (kth.d != kInfD) {
            const double rr = ceil((sqrt(ent_d2(kth)) + wg.margin) / wg.g.c);
            if (rr > rn) rn = rr > 1e6 ? 1000000 : static_cast<int>(rr);
        } else {
            rn = max(rn, 2 * r);
        }
        r = min(rn, a.tune.ring_max);
    }
}

// Tie test and hand-off of one query's result: in neighbor-list mode the kept
// ids go straight to global memory; in feature mode the kept sorted positions
// are parked in LDS for the covariance pass.
__device__ __forceinline__ void finish_query(const KnnArgs& a, const WaveGrid& wg, const Ent& best,
                                             bool incomplete, uint64_t s, int qslot,
                                             int lane, uint32_t* nbr, uint8_t& my_flags)
{
    const int fetch = wg.fetch;
    const int k = wg.k;
    const int stride = wg.stride;
    // Boundary tie: the fetch-th and (fetch+1)-th distances are equal.  With
    // stride > 1 every adjacent equal pair inside the fetched list can change
    // which ids are kept, and ordered lists (mode 2) need KD3Index's order
    // among equal distances, so there any such pair flags the row.
    const Ent nxt = ent_shfl(best, lane + 1 < 64 ? lane + 1 : 63);
    const bool eq = (lane < fetch) && (lane + 1 < 64) && best.d == nxt.d && best.d != kInfD;
    const bool tie = (stride > 1 || a.mode == 2)
                         ? (__ballot(eq) != 0)
                         : (lane_u32(eq ? 1u : 0u, fetch - 1) != 0);
    const uint8_t flags = (tie ? ST_TIE : 0) | (incomplete ? ST_INCOMPLETE : 0);
    // Kept list in canonical (d2, id) order: entries 0, stride, 2*stride...
    const Ent kept = ent_shfl(best, (lane * stride) & 63);
    if (a.mode != 0) {
        const uint32_t qid = a.cv.perm[s];
        // Mode 2 keeps sorted positions (the consumer reads the sorted
        // coordinates, which neighbors share cache lines in).
        if (lane < k)
            a.out.ids[static_cast<uint64_t>(qid) * k + lane] = a.mode == 2 ? kept.j : kept.id;
        if (lane == 0) a.out.status[0][qid] = flags;
        if (lane == 0 && flags) {
            uint32_t slot = atomicAdd(a.out.nflagged, 1u);
            a.out.flagged[slot] = static_cast<uint32_t>(s);
            atomicAdd(a.out.nflagged + ((flags & ST_INCOMPLETE) ? 2 : 1), 1u);
        }
    } else {
        // An uncertified row can hold fewer than fetch entries; keep its list
        // in bounds (the exact search replaces the row).
        if (lane < k) nbr[qslot * a.lds_stride + lane] = kept.j == kInfU ? 0u : kept.j;
        if (lane == qslot) my_flags = flags;
    }
}

template <bool kTiled, int kWaves>
// Four waves per SIMD: the selection is a chain of ballots and LDS
// round trips, so more resident waves hide more of its latency (measured
// faster than three waves even with the register spills four waves cost).
__global__ __launch_bounds__(kWaves * kWave) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_knn(KnnArgs a_in)
{
    extern __shared__ uint32_t s_lds[];
    const int lane = __lane_id();
    // Wave-uniform by construction; said so, so that per-wave tables load
    // into scalar registers.
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
    const int Q = a_in.q_per_wave;
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint64_t task = static_cast<uint64_t>(lb) * kWaves + w;
    const KnnArgs& a = a_in;
    WaveGrid wg;
    uint64_t q0;
    int nq;
    if (kTiled) {
        // The wave's queries and its tile's grid and program.
        if (task >= a_in.ntasks) return;
        // The task's tile: the last t with task0[t] <= task (tiles with no
        // task share the next tile's prefix).
        uint32_t tlo = 0, thi = a_in.ntiles;
        while (thi - tlo > 1) {
            const uint32_t mid = (tlo + thi) >> 1;
            if (a_in.task0[mid] <= task)
                tlo = mid;
            else
                thi = mid;
        }
        const TileInfo& ti = a_in.tiles[tlo];
        const uint64_t first = static_cast<uint64_t>(task - a_in.task0[tlo]) * Q;
        wg.g = ti.g;
        wg.colstart = a_in.cv.colstart + ti.colbase;
        wg.margin = ti.margin;
        wg.fetch = ti.prog.fetch;
        wg.k = ti.prog.k;
        wg.stride = ti.prog.stride;
        q0 = ti.start + first;
        nq = static_cast<int>(ti.n - first < static_cast<uint64_t>(Q) ? ti.n - first : Q);
    } else {
        wg.g = a_in.cv.g;
        wg.colstart = a_in.cv.colstart;
        wg.margin = a_in.margin;
        wg.fetch = a_in.prog.fetch;
        wg.k = a_in.prog.k;
        wg.stride = a_in.prog.stride;
        const uint64_t n = a_in.cv.n;
        q0 = task * Q;
        if (q0 >= n) return;
        nq = static_cast<int>(n - q0 < static_cast<uint64_t>(Q) ? n - q0 : Q);
    }
    const CloudView& cv = a.cv;
    const GridParams& g = wg.g;
    const int fetch = wg.fetch;
    const int k = wg.k;
    // LDS: per-wave neighbor lists (features mode), then per wave a 64-entry
    // compaction scratch of 16 bytes per entry and a 64-entry pending list.
    const size_t list_words = a.mode != 0 ? 0 : static_cast<size_t>(kWaves) * Q * a.lds_stride;
    uint32_t* nbr = s_lds + static_cast<size_t>(w) * Q * a.lds_stride;  // lds_stride >= k
    uint32_t* scratch = s_lds + list_words + static_cast<size_t>(w) * kScratchWords;
    uint32_t* pending = scratch + 512;  // queries left for the growing search
    int npend = 0;
    const uint32_t zmask = (1u << g.bz) - 1u;

    uint8_t my_flags = 0;
    uint64_t cnt[3] = {0, 0, 0};  // search passes, candidates, selections
    int qi = 0;
    int rg_prev = 1;  // box half-width of the previous group (density guess)
    // Features with stride 1 keep each neighborhood as a set: the exact-count
    // selection serves them; lists, and strided neighborhoods, need the
    // (d2, id) order.
    const bool as_set = a.mode == 0 && wg.stride == 1 && a.tune.set_select != 0;
    // The previous query's threshold, the first one tried (queries in sorted
    // order are close, with similar neighborhood radii).
    double hint = INFINITY;
    // The wave's queries, one per lane, loaded once: coordinates and keys.
    const bool lane_q = lane < nq;
    const double wx = lane_q ? cv.sx[q0 + lane] : 0.0;
    const double wy = lane_q ? cv.sy[q0 + lane] : 0.0;
    const double wz = lane_q ? cv.sz[q0 + lane] : 0.0;
    const uint32_t wkey = lane_q ? cv.keys[q0 + lane] : 0u;
    while (qi < nq) {
        // The group: the run of queries from qi that lie in one 2x2 block of
        // columns (one column if the block is too dense) with z cells within
        // group_z of the first query's.  All of them share one candidate box:
        // the block grown by rg columns on each side, z cells of the members
        // grown by rg.
        const uint64_t s0 = q0 + qi;
        const bool in_range = qi + lane < nq;
        // (shuffled by every lane: a lane outside the wave's queries is still
        // the source of another's key)
        const uint32_t key_sh =
            static_cast<uint32_t>(__shfl(static_cast<int>(wkey), (qi + lane) & 63));
        const uint32_t key_l = in_range ? key_sh : 0u;
        const uint32_t key0 = lane_u32(key_l, 0);
        const int cz0 = static_cast<int>(key0 & zmask);
        const int cz_l = static_cast<int>(key_l & zmask);
        const bool zok = in_range && abs(cz_l - cz0) <= a.tune.group_z;
        const double fx = lane_value(wx, qi), fy = lane_value(wy, qi);
        const int qcx = static_cast<int>(cell_coord(fx, g.ox, g.inv_c, g.gx));
        const int qcy = static_cast<int>(cell_coord(fy, g.oy, g.inv_c, g.gy));
        const uint32_t need = static_cast<uint32_t>(fetch + 1);
        int level = 1;
        int gm = 0, zmin = cz0, zmax = cz0, rg = rg_prev;
        Box box{};
        uint32_t lo = 0, len = 0, tot = 0, excl = 0;
        for (int attempt = 0; attempt < 6; ++attempt) {
            const uint32_t shift = g.bz + 2u * static_cast<uint32_t>(level);
            const bool member = zok && (key_l >> shift) == (key0 >> shift);
            const uint64_t mm = __ballot(member);
            gm = (~mm == 0ull) ? 64 : __builtin_ctzll(~mm);
            int zl = lane < gm ? cz_l : cz0, zh = zl;
#pragma unroll
            for (int m = 32; m > 0; m >>= 1) {
                zl = min(zl, __shfl_xor(zl, m));
                zh = max(zh, __shfl_xor(zh, m));
            }
            zmin = uniform_i32(zl);
            zmax = uniform_i32(zh);
            const int bx0 = level ? (qcx & ~1) : qcx, by0 = level ? (qcy & ~1) : qcy;
            const int bw = level ? 2 : 1;
            box = Box{bx0 - rg, bx0 + bw - 1 + rg, by0 - rg, by0 + bw - 1 + rg, zmin - rg,
                      zmax + rg};
            box_ranges(a, wg, box, 0, lane, lo, len);
            const uint32_t incl = wave_incl_scan_u32(len, lane);
            tot = lane_u32(incl, 63);
            excl = incl - len;
            // Half-width the local density asks for: the need-th neighbor
            // radius of a surface of this 2-D density, with 10% slack.
            const int width = bw + 2 * rg;
            const double area = static_cast<double>(width * width) * g.c * g.c;
            int est = 3;
            if (tot > 0) {
                const double rk = sqrt(static_cast<double>(need) * area /
                                       (3.141592653589793 * static_cast<double>(tot)));
                est = static_cast<int>(ceil(1.1 * rk / g.c));
                // rg <= 3 keeps the box within 64 columns ((2 + 2 rg)^2),
                // the one box_ranges chunk this path reads.
                est = est < 1 ? 1 : (est > 3 ? 3 : est);
            }
            if (tot > kRegCand) {
                if (rg > 1) {
                    rg = min(rg - 1, est);
                    continue;
                }
                if (level == 1) {
                    level = 0;
                    continue;
                }
                break;  // one dense column: searched query by query
            }
            if (est > rg) {
                rg = est;
                continue;
            }
            break;
        }
        rg_prev = rg;
        const bool cached = tot <= kRegCand;

        // Shared candidates of the group's box, cached in registers.
        double px[kRegBatches], py[kRegBatches], pz[kRegBatches];
        uint32_t pid[kRegBatches], pj[kRegBatches];
#pragma unroll
        for (int b = 0; b < kRegBatches; ++b) {
            px[b] = py[b] = pz[b] = 0.0;
            pid[b] = pj[b] = kInfU;
            if (cached && static_cast<uint32_t>(b * 64) < tot) {
                const uint32_t t = b * 64 + lane;
                const uint32_t j = candidate_pos(lo, excl, t);
                if (t < tot) {
                    px[b] = cv.sx[j];
                    py[b] = cv.sy[j];
                    pz[b] = cv.sz[j];
                    pid[b] = cv.perm[j];
                    pj[b] = j;
                }
            }
        }

        for (int gi = 0; gi < gm; ++gi) {
            const uint64_t s = s0 + gi;
            const int qslot = qi + gi;
            bool done = false;
            if (cached) {
                const double qx = lane_value(wx, qslot), qy = lane_value(wy, qslot),
                             qz = lane_value(wz, qslot);
                cnt[0] += 1;
                cnt[1] += tot;
                // The candidates' ids and positions are the cache's own.
                uint64_t cd[kRegBatches];
#pragma unroll
                for (int b = 0; b < kRegBatches; ++b)
                    cd[b] = (static_cast<uint32_t>(b) * 64u < tot && pj[b] != kInfU)
                                ? dist_bits(qx, qy, qz, px[b], py[b], pz[b])
                                : kInfD;
                const uint32_t nbat = (tot + 63u) >> 6;
                uint64_t E = 0;
                if (as_set && hint < INFINITY && tot >= static_cast<uint32_t>(fetch) &&
                    select_count(cd, nbat, static_cast<uint32_t>(fetch), hint, E)) {
                    // The kept set is {d2 <= E}, with no boundary tie.  It is
                    // certified when its fetch-th distance (<= E) is below
                    // the box clearance less the margin, i.e. when E is, or
                    // when fetch candidates are.
                    cnt[2] += 1;
                    hint = bits_d2(E);
                    const double R = box_clearance(g, box, qx, qy, qz);
                    done = R == INFINITY;
                    if (!done) {
                        const double Rm = R - wg.margin;
                        if (Rm > 0.0) {
                            const double R2 = Rm * Rm;
                            done = bits_d2(E) < R2 ||
                                   count_bits<true>(cd, nbat, d2_bits(R2)) >=
                                       static_cast<uint32_t>(fetch);
                        }
                    }
                    if (done) {
                        // The list in (d2, id) order, as the sorting path
                        // keeps it (so the covariance sums in one order
                        // whatever the grid and launch shape): the kept
                        // entries are compacted into the scratch, and each
                        // of the first fetch lanes counts the entries before
                        // its own and stores its sorted position there.
                        uint32_t base = 0;
#pragma unroll
                        for (int b = 0; b < kRegBatches; ++b) {
                            if (static_cast<uint32_t>(b) >= nbat) break;
                            const bool kept = cd[b] <= E;
                            const uint64_t m = __ballot(kept);
                            if (kept)
                                reinterpret_cast<uint4*>(scratch)[base + lanes_below64(m)] =
                                    make_uint4(static_cast<uint32_t>(cd[b]),
                                               static_cast<uint32_t>(cd[b] >> 32), pid[b], pj[b]);
                            base += static_cast<uint32_t>(__popcll(m));
                        }
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        if (lane < fetch) {
                            const uint4 me = reinterpret_cast<const uint4*>(scratch)[lane];
                            const uint64_t vd = static_cast<uint64_t>(me.x) |
                                                (static_cast<uint64_t>(me.y) << 32);
                            uint32_t rank = 0;
#pragma unroll 4
                            for (int i = 0; i < fetch; ++i)
                                rank += rank_step(reinterpret_cast<const uint4*>(scratch)[i],
                                                  vd, me.z);
                            nbr[qslot * a.lds_stride + rank] = me.w;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                    }
                } else {
                    // Neighboring queries of a group have similar need-th
                    // distances: the previous one is the first threshold tried.
                    Ent best;
                    select_entries(cd, pid, pj, tot, fetch, lane, scratch, hint, best, cnt);
                    const Ent nth = ent_lane(best, static_cast<int>(min(need, tot)) - 1);
                    hint = nth.d == kInfD ? INFINITY : ent_d2(nth);
                    done = certified(wg, box, ent_lane(best, fetch - 1), qx, qy, qz);
                    if (done) finish_query(a, wg, best, false, s, qslot, lane, nbr, my_flags);
                }
            }
            if (!done) {
                // Searched alone after the groups, so the group's register
                // cache is not live at the same time as the growing search.
                if (lane == 0)
                    pending[npend] = static_cast<uint32_t>(qslot) |
                                     (static_cast<uint32_t>(cached ? rg + 1 : 1) << 8);
                ++npend;
            }
        }
        qi += gm;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int p = 0; p < npend; ++p) {
        const uint32_t word = pending[p];
        const int qslot = static_cast<int>(word & 0xFFu);
        const uint64_t s = q0 + qslot;
        const double qx = lane_value(wx, qslot), qy = lane_value(wy, qslot),
                     qz = lane_value(wz, qslot);
        const int qcx = static_cast<int>(cell_coord(qx, g.ox, g.inv_c, g.gx));
        const int qcy = static_cast<int>(cell_coord(qy, g.oy, g.inv_c, g.gy));
        const int qcz = static_cast<int>(cell_coord(qz, g.oz, g.inv_cz, g.gz));
        Ent best;
        bool incomplete = false;
        search_single(a, wg, qx, qy, qz, qcx, qcy, qcz, static_cast<int>(word >> 8), lane,
                      scratch, best, incomplete, cnt);
        finish_query(a, wg, best, incomplete, s, qslot, lane, nbr, my_flags);
    }
    if (a.out.counters && lane == 0) {
        atomicAdd(a.out.counters + 0, static_cast<unsigned long long>(cnt[0]));
        atomicAdd(a.out.counters + 1, static_cast<unsigned long long>(cnt[1]));
        atomicAdd(a.out.counters + 2, static_cast<unsigned long long>(cnt[2]));
        atomicAdd(a.out.counters + 3, static_cast<unsigned long long>(nq));
    }
    if (a.mode != 0) return;
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
    // Coordinates relative to the query point (exact differences at any
    // offset), so the mean carries no rounding of the cloud's absolute
    // position: at UTM scale that rounding (~1e-9 m) would add a rank-one
    // term to the covariance of flat and collinear neighborhoods.
    const double ox = cv.sx[s], oy = cv.sy[s], oz = cv.sz[s];
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (int l = 0; l < k; ++l) {
        const uint32_t j = list[l];
        mx += cv.sx[j] - ox;
        my += cv.sy[j] - oy;
        mz += cv.sz[j] - oz;
    }
    mx /= static_cast<double>(k);
    my /= static_cast<double>(k);
    mz /= static_cast<double>(k);
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (int l = 0; l < k; ++l) {
        const uint32_t j = list[l];
        const double dx = (cv.sx[j] - ox) - mx, dy = (cv.sy[j] - oy) - my,
                     dz = (cv.sz[j] - oz) - mz;
        c[0] += dx * dx;
        c[1] += dx * dy;
        c[2] += dx * dz;
        c[3] += dy * dy;
        c[4] += dy * dz;
        c[5] += dz * dz;
    }
    const double den = static_cast<double>(k - 1);
    double* rec = a.out.cov + kCovRecord * s;
    for (int i = 0; i < 6; ++i) rec[i] = c[i] / den;
    reinterpret_cast<uint64_t*>(rec)[6] = my_flags;
    if (my_flags) {
        uint32_t slot = atomicAdd(a.out.nflagged, 1u);
        a.out.flagged[slot] = static_cast<uint32_t>(s);
        atomicAdd(a.out.nflagged + ((my_flags & ST_INCOMPLETE) ? 2 : 1), 1u);
    }
}

// ---------------------------------------------------------------------------
// Radius neighborhoods (the stages' `radius` option; KD3Index::radius,
// pdal/private/KDImpl.hpp:398-415, nanoflann radiusSearch): every point whose
// FP64 squared distance (KD3Index's operation order) is strictly below r*r.
// The set is decided by that comparison alone -- no boundary tie, no exact
// re-solve.  A wavefront takes a group of consecutive sorted queries (one
// 2x2 block of columns with nearby z cells, as the kNN search groups them)
// and the box of every cell that can hold a point within r of any of them:
// the block grown by R = ceil(r / c) columns, the group's z cells by
// ceil(r / cz) (cell edges chosen so that R is 2).  Per query the
// wavefront masks the box's candidates by d2 < r*r and forms the count and
// the FP64 two-pass covariance (coordinates relative to the query point)
// from masked wavefront sums.  Boxes of up to 256 candidates and 64 columns
// are held in registers for the whole group; larger boxes are streamed from
// memory per query.
struct RadiusArgs {
    CloudView cv;
    double r2;          // r * r, the strict bound on d2
    int32_t rx, rz;     // box growth in columns and z cells
    int32_t group_z;
    int32_t mode;       // 0: covariance records; 1: counts only
    double* cov;        // mode 0: records in sorted order (count in word 7)
    uint32_t* counts;   // mode 1: counts in original order
};

// Candidates of one chunk: d2 bits per register batch (kInfD outside r).
__device__ __forceinline__ void radius_pass1(const double (&px)[kRegBatches],
                                             const double (&py)[kRegBatches],
                                             const double (&pz)[kRegBatches],
                                             const uint32_t (&pj)[kRegBatches], uint32_t n1,
                                             double qx, double qy, double qz, double r2,
                                             uint32_t& cnt, double& sx, double& sy, double& sz)
{
#pragma unroll
    for (int b = 0; b < kRegBatches; ++b) {
        if (static_cast<uint32_t>(b) * 64u >= n1) break;
        const bool valid = pj[b] != kInfU;
        const double d2 = bits_d2(dist_bits(qx, qy, qz, px[b], py[b], pz[b]));
        const bool in = valid && d2 < r2;
        cnt += static_cast<uint32_t>(__popcll(__ballot(in)));
        if (in) {
            sx += px[b] - qx;
            sy += py[b] - qy;
            sz += pz[b] - qz;
        }
    }
}

__device__ __forceinline__ void radius_pass2(const double (&px)[kRegBatches],
                                             const double (&py)[kRegBatches],
                                             const double (&pz)[kRegBatches],
                                             const uint32_t (&pj)[kRegBatches], uint32_t n1,
                                             double qx, double qy, double qz, double r2,
                                             double mx, double my, double mz, double (&c)[6])
{
#pragma unroll
    for (int b = 0; b < kRegBatches; ++b) {
        if (static_cast<uint32_t>(b) * 64u >= n1) break;
        const bool valid = pj[b] != kInfU;
        const double d2 = bits_d2(dist_bits(qx, qy, qz, px[b], py[b], pz[b]));
        if (valid && d2 < r2) {
            const double dx = (px[b] - qx) - mx, dy = (py[b] - qy) - my, dz = (pz[b] - qz) - mz;
            c[0] += dx * dx;
            c[1] += dx * dy;
            c[2] += dx * dz;
            c[3] += dy * dy;
            c[4] += dy * dz;
            c[5] += dz * dz;
        }
    }
}

__global__ __launch_bounds__(256) void k_radius(RadiusArgs a)
{
    const int lane = __lane_id();
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const CloudView& cv = a.cv;
    const GridParams& g = cv.g;
    const uint64_t q0 = (static_cast<uint64_t>(lb) * 4 + w) * kWave;
    if (q0 >= cv.n) return;
    const int nq = static_cast<int>(cv.n - q0 < 64 ? cv.n - q0 : 64);
    const uint32_t zmask = (1u << g.bz) - 1u;
    // box_ranges reads only these fields of the kNN arguments.
    KnnArgs ka{};
    ka.cv = cv;
    ka.tune.small_col = 0;
    WaveGrid wg{};
    wg.g = g;
    wg.colstart = cv.colstart;
    int qi = 0;
    while (qi < nq) {
        const uint64_t s0 = q0 + qi;
        const bool in_range = qi + lane < nq;
        const uint32_t key_l = in_range ? cv.keys[s0 + lane] : 0u;
        const uint32_t key0 = lane_u32(key_l, 0);
        const int cz0 = static_cast<int>(key0 & zmask);
        const int cz_l = static_cast<int>(key_l & zmask);
        const bool zok = in_range && abs(cz_l - cz0) <= a.group_z;
        const uint32_t shift = g.bz + 2u;
        const uint64_t mm = __ballot(zok && (key_l >> shift) == (key0 >> shift));
        const int gm = (~mm == 0ull) ? 64 : __builtin_ctzll(~mm);
        int zl = lane < gm ? cz_l : cz0, zh = zl;
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) {
            zl = min(zl, __shfl_xor(zl, m));
            zh = max(zh, __shfl_xor(zh, m));
        }
        const double fx = cv.sx[s0], fy = cv.sy[s0];
        const int qcx = static_cast<int>(cell_coord(fx, g.ox, g.inv_c, g.gx));
        const int qcy = static_cast<int>(cell_coord(fy, g.oy, g.inv_c, g.gy));
        const int bx0 = qcx & ~1, by0 = qcy & ~1;
        const Box box{bx0 - a.rx, bx0 + 1 + a.rx, by0 - a.rx, by0 + 1 + a.rx, zl - a.rz, zh + a.rz};
        const int ncb = (2 + 2 * a.rx) * (2 + 2 * a.rx);
        uint32_t tot = 0;
        for (int cb = 0; cb < ncb; cb += kWave) {
            uint32_t lo, len;
            box_ranges(ka, wg, box, cb, lane, lo, len);
            tot += lane_u32(wave_incl_scan_u32(len, lane), 63);
        }
        const bool cached = ncb <= kWave && tot <= kRegCand;
        double px[kRegBatches], py[kRegBatches], pz[kRegBatches];
        uint32_t pj[kRegBatches];
        uint32_t lo0 = 0, excl0 = 0;
        if (cached) {
            uint32_t len;
            box_ranges(ka, wg, box, 0, lane, lo0, len);
            excl0 = wave_incl_scan_u32(len, lane) - len;
#pragma unroll
            for (int b = 0; b < kRegBatches; ++b) {
                px[b] = py[b] = pz[b] = 0.0;
                pj[b] = kInfU;
                if (static_cast<uint32_t>(b * 64) < tot) {
                    const uint32_t t = b * 64 + lane;
                    const uint32_t j = candidate_pos(lo0, excl0, t);
                    if (t < tot) {
                        px[b] = cv.sx[j];
                        py[b] = cv.sy[j];
                        pz[b] = cv.sz[j];
                        pj[b] = j;
                    }
                }
            }
        }
        for (int gi = 0; gi < gm; ++gi) {
            const uint64_t s = s0 + gi;
            const double qx = cv.sx[s], qy = cv.sy[s], qz = cv.sz[s];
            uint32_t cnt = 0;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            double c[6] = {0, 0, 0, 0, 0, 0};
            for (int pass = 0; pass < (a.mode == 0 ? 2 : 1); ++pass) {
                double mx = 0.0, my = 0.0, mz = 0.0;
                if (pass == 1) {
                    const double nn = static_cast<double>(cnt);
                    mx = wave_sum_f64(sx) / nn;
                    my = wave_sum_f64(sy) / nn;
                    mz = wave_sum_f64(sz) / nn;
                }
                if (cached) {
                    if (pass == 0)
                        radius_pass1(px, py, pz, pj, tot, qx, qy, qz, a.r2, cnt, sx, sy, sz);
                    else
                        radius_pass2(px, py, pz, pj, tot, qx, qy, qz, a.r2, mx, my, mz, c);
                    continue;
                }
                // Streamed: chunks of 64 columns and 256 candidates.
                for (int cb = 0; cb < ncb; cb += kWave) {
                    uint32_t lo, len;
                    box_ranges(ka, wg, box, cb, lane, lo, len);
                    const uint32_t incl = wave_incl_scan_u32(len, lane);
                    const uint32_t ctot = lane_u32(incl, 63);
                    const uint32_t excl = incl - len;
                    for (uint32_t t0 = 0; t0 < ctot; t0 += kRegCand) {
                        double cx[kRegBatches], cy[kRegBatches], czv[kRegBatches];
                        uint32_t cj[kRegBatches];
                        const uint32_t n1 = min(ctot - t0, kRegCand);
#pragma unroll
                        for (int b = 0; b < kRegBatches; ++b) {
                            cx[b] = cy[b] = czv[b] = 0.0;
                            cj[b] = kInfU;
                            if (static_cast<uint32_t>(b * 64) < n1) {
                                const uint32_t t = t0 + b * 64 + lane;
                                const uint32_t j = candidate_pos(lo, excl, t);
                                if (t < ctot) {
                                    cx[b] = cv.sx[j];
                                    cy[b] = cv.sy[j];
                                    czv[b] = cv.sz[j];
                                    cj[b] = j;
                                }
                            }
                        }
                        if (pass == 0)
                            radius_pass1(cx, cy, czv, cj, n1, qx, qy, qz, a.r2, cnt, sx, sy, sz);
                        else
                            radius_pass2(cx, cy, czv, cj, n1, qx, qy, qz, a.r2, mx, my, mz, c);
                    }
                }
            }
            if (a.mode == 1) {
                if (lane == 0) a.counts[cv.perm[s]] = cnt;
                continue;
            }
            const double den = static_cast<double>(cnt) - 1.0;
            double v[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) v[i] = wave_sum_f64(c[i]) / den;
            if (lane == 0) {
                double* rec = a.cov + kCovRecord * s;
#pragma unroll
                for (int i = 0; i < 6; ++i) rec[i] = v[i];
                reinterpret_cast<uint64_t*>(rec)[6] = 0;
                reinterpret_cast<uint64_t*>(rec)[7] = cnt;
            }
        }
        qi += gm;
    }
}

// Per sorted row: eigensystem of the stored covariance and every requested
// stage column, written at the row's original index.
__device__ __forceinline__ uint32_t tile_of(const uint64_t* __restrict__ starts, uint32_t T,
                                            uint64_t s)
{
    uint32_t lo = 0, hi = T;  // largest t with starts[t] <= s
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (starts[mid] <= s)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// One thread per row: by original id (kFeaturesAll, kFeaturesUnflagged: a
// row whose record carries a search flag is skipped), or the i-th row of
// out.flagged by sorted position (kFeaturesFlagged).

__device__ __forceinline__ void features_row(uint64_t i, const double* __restrict__ cov,
                                             const uint32_t* __restrict__ inv,
                                             const uint32_t* __restrict__ perm, int rows,
                                             const FusedProgram& prog, const KnnOutputs& out,
                                             const TileInfo* __restrict__ tiles,
                                             const uint64_t* __restrict__ starts,
                                             uint32_t ntiles)
{
    uint64_t qid = i, s;
    if (rows == kFeaturesFlagged) {
        s = out.flagged[i];
        qid = perm[s];
    } else {
        s = inv[qid];  // the row's sorted position: one record read
    }
    // The neighborhood size: the row's own count in radius mode, else its
    // tile's k (stages are the same in every tile).
    const double* rec = cov + kCovRecord * s;
    const int k = prog.radius > 0.0
                      ? static_cast<int>(reinterpret_cast<const uint64_t*>(rec)[7])
                      : (tiles ? tiles[tile_of(starts, ntiles, s)].prog.k : prog.k);
    double c[6];
    for (int i = 0; i < 6; ++i) c[i] = rec[i];
    const uint8_t f = static_cast<uint8_t>(reinterpret_cast<const uint64_t*>(rec)[6]);
    if (rows == kFeaturesUnflagged && f) return;  // written after its exact re-solve
    const bool zero = cov_is_zero(c);
    double ev[3] = {0, 0, 0}, vec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int info = 0;
    // estimaterank has no isZero skip: it takes the singular values of any
    // covariance.
    bool rank_stage = false;
    for (int si = 0; si < prog.nstages; ++si) rank_stage |= prog.st[si].stage == STAGE_RANK;
    if (!zero || rank_stage) info = eigen_sym3(c, ev, vec);
    for (int si = 0; si < prog.nstages; ++si) {
        const StageProgram& sp = prog.st[si];
        double vals[10];
        // Radius mode: covariancefeatures and eigenvalues skip rows with
        // fewer than min_k neighbors (CovarianceFeaturesFilter.cpp:216-223,
        // EigenvaluesFilter.cpp:125-135); normal skips fewer than three.
        uint8_t st = (prog.radius > 0.0 && sp.stage != STAGE_NORMAL && k < prog.min_k)
                         ? static_cast<uint8_t>(ST_FEW_NEIGHBORS)
                         : stage_outputs(sp, k, c, zero, info, ev, vec, vals);
        if (st == 0)
            for (int cc = 0; cc < sp.ncols; ++cc) out.col[sp.col_offset + cc][qid] = vals[cc];
        out.status[si][qid] = st | (f & kReportedSearchFlags);
    }
}

__global__ __launch_bounds__(256) void k_features(const double* __restrict__ cov,
                                                  const uint32_t* __restrict__ inv,
                                                  const uint32_t* __restrict__ perm,
                                                  uint64_t n, int rows, FusedProgram prog,
                                                  KnnOutputs out, const TileInfo* __restrict__ tiles,
                                                  const uint64_t* __restrict__ starts,
                                                  uint32_t ntiles)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    features_row(i, cov, inv, perm, rows, prog, out, tiles, starts, ntiles);
}

// The flagged rows, their count read on the device (a launch whose shape
// does not depend on it, as a captured graph needs).
__global__ __launch_bounds__(256) void k_features_flagged(const double* __restrict__ cov,
                                                          const uint32_t* __restrict__ inv,
                                                          const uint32_t* __restrict__ perm,
                                                          FusedProgram prog, KnnOutputs out,
                                                          const TileInfo* __restrict__ tiles,
                                                          const uint64_t* __restrict__ starts,
                                                          uint32_t ntiles)
{
    const uint64_t n = *out.nflagged;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        features_row(i, cov, inv, perm, kFeaturesFlagged, prog, out, tiles, starts, ntiles);
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
                const KnnOutputs& out, const KnnTuning& tune, hipStream_t s,
                const TileTables* tt)
{
    if (cv.n == 0) return;
    KnnArgs a;
    a.cv = cv;
    a.prog = prog;
    a.out = out;
    a.tune = tune;
    a.mode = mode;
    a.tiles = nullptr;
    a.task0 = nullptr;
    a.ntasks = 0;
    a.ntiles = 0;
    const int max_k = tt ? tt->max_k : prog.k;
    a.q_per_wave = tt ? tt->q_per_wave
                      : (tune.q_per_wave > 0 ? tune.q_per_wave : knn_queries_per_wave(prog.k));
    a.lds_stride = max_k | 1;  // odd stride: lanes hit distinct LDS banks
    const GridParams& g = cv.g;
    a.margin = 1e-6 * fmin(g.c, g.cz) +
               1e-15 * (fabs(g.ox) + fabs(g.oy) + fabs(g.oz)) + 1e-300;
    const int wv = tune.waves == 1 || tune.waves == 2 ? tune.waves : kKnnWavesDefault;
    const size_t waves = static_cast<size_t>(tt ? 1 : wv);
    const size_t list_words = mode != 0 ? 0 : waves * a.q_per_wave * a.lds_stride;
    const size_t lds = (list_words + waves * kScratchWords) * sizeof(uint32_t);
    if (tt) {
        if (!tt->ntasks) return;
        a.tiles = tt->tiles;
        a.task0 = tt->task0;
        a.ntasks = tt->ntasks;
        a.ntiles = tt->ntiles;
        const uint32_t blocks = div_up(tt->ntasks, waves);
        hipLaunchKernelGGL((k_knn<true, 1>), dim3(blocks), dim3(waves * kWave), lds, s, a);
        return;
    }
    const uint64_t per_block = static_cast<uint64_t>(waves) * a.q_per_wave;
    const uint32_t blocks = div_up(cv.n, per_block);
    if (wv == 1)
        hipLaunchKernelGGL((k_knn<false, 1>), dim3(blocks), dim3(waves * kWave), lds, s, a);
    else if (wv == 2)
        hipLaunchKernelGGL((k_knn<false, 2>), dim3(blocks), dim3(waves * kWave), lds, s, a);
    else
        hipLaunchKernelGGL((k_knn<false, 4>), dim3(blocks), dim3(waves * kWave), lds, s, a);
}

void launch_radius(const CloudView& cv, double radius, int mode, double* cov, uint32_t* counts,
                   int group_z, hipStream_t s)
{
    if (cv.n == 0) return;
    RadiusArgs a{};
    a.cv = cv;
    a.r2 = radius * radius;
    // Every point within r of a query lies within these cells: the cell
    // coordinates' own rounding (relative 4 eps of the grid's span) is
    // covered.
    const GridParams& g = cv.g;
    const double span = fmax(fmax(static_cast<double>(g.gx) * g.c, static_cast<double>(g.gy) * g.c),
                             static_cast<double>(g.gz) * g.cz);
    const double slack = 1e-15 * (span + fabs(g.ox) + fabs(g.oy) + fabs(g.oz));
    a.rx = static_cast<int32_t>(ceil((radius + slack) / g.c));
    a.rz = static_cast<int32_t>(ceil((radius + slack) / g.cz));
    a.group_z = group_z;
    a.mode = mode;
    a.cov = cov;
    a.counts = counts;
    const uint32_t blocks = div_up(cv.n, 256);
    hipLaunchKernelGGL(k_radius, dim3(blocks), dim3(256), 0, s, a);
}

void launch_features(const CloudView& cv, const FusedProgram& prog,
                     const KnnOutputs& out, const uint32_t* inv, hipStream_t s,
                     const TileTables* tt, int rows, uint64_t nflagged)
{
    const uint64_t n = rows == kFeaturesFlagged ? nflagged : cv.n;
    if (n == 0) return;
    hipLaunchKernelGGL(k_features, dim3(div_up(n, 256)), dim3(256), 0, s, out.cov, inv, cv.perm,
                       n, rows, prog, out, tt ? tt->tiles : nullptr, tt ? tt->starts : nullptr,
                       tt ? tt->ntiles : 0u);
}

void launch_features_flagged(const CloudView& cv, const FusedProgram& prog,
                             const KnnOutputs& out, const uint32_t* inv, uint64_t max_rows,
                             hipStream_t s, const TileTables* tt)
{
    if (!max_rows) return;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(div_up(max_rows, 256), 2048));
    hipLaunchKernelGGL(k_features_flagged, dim3(blocks), dim3(256), 0, s, out.cov, inv, cv.perm,
                       prog, out, tt ? tt->tiles : nullptr, tt ? tt->starts : nullptr,
                       tt ? tt->ntiles : 0u);
}

void launch_eigen_batch(const double* cov, uint64_t m, double* evals,
                        double* evecs, int32_t* info, hipStream_t s)
{
    if (!m) return;
    hipLaunchKernelGGL(k_eigen_batch, dim3(div_up(m, 256)), dim3(256), 0, s, cov, m,
                       evals, evecs, info);
}

} // namespace gf
