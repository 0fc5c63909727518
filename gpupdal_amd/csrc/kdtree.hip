// kdtree.hip -- the exact KD3Index tree on the GPU and its search, used to
// re-solve the rows the grid search cannot decide alone: boundary distance
// ties (KD3Index keeps the tied point it visits first) and searches the grid
// budget did not certify.
//
// The tree is the one nanoflann 1.3.1 builds for PDAL's KD3 adaptor (leaf
// size 100, the cloud's bounds as the root box, middleSplit_, planeSplit;
// pdal/private/KDImpl.hpp:209-309, vendor/nanoflann/nanoflann.hpp:619-875,
// 1012-1032), built level by level instead of recursively:
//   - per node, the min/max of its points on each axis (a segmented
//     reduction) gives the split axis and the clamped middle cut value, and
//     the parent's divlow/divhigh (max of child1 / min of child2 on the cut
//     axis);
//   - planeSplit's two Hoare passes (< cut, then <= cut on the rest) are
//     applied in closed form: with L the count of "left" elements, the j-th
//     misplaced element of [0, L) counted from the left is swapped with the
//     j-th misplaced element of [L, n) counted from the right, found through
//     prefix counts; this yields the same permutation as the sequential
//     two-pointer loop, so leaves hold their points in the same order;
//   - the split index is lim1 / lim2 / count/2 by the same rule.
// The search restates findNeighbors / searchLevel / KNNResultSet (strict <
// against the worst distance read at leaf entry, closer child first, the
// other child when the incremental box distance is <= the worst,
// nanoflann.hpp:145-208, 1059-1080, 1184-1248) one row per wavefront.
#include "gf_internal.hpp"
#include "gf_kernels.hpp"
#include "gf_wave.hpp"

#include <algorithm>
#include <utility>
#include <vector>

namespace gf {

namespace {

constexpr uint32_t kLeafMax = 100;
constexpr int kKdThreads = 256;

__device__ __forceinline__ unsigned long long enc_order(double d)
{
    unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(d));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double dec_order(unsigned long long u)
{
    unsigned long long b = (u >> 63) ? (u ^ 0x8000000000000000ull) : ~u;
    return __longlong_as_double(static_cast<long long>(b));
}

__device__ __forceinline__ const double* axis(const KdView& t, int f)
{
    return f == 0 ? t.px : (f == 1 ? t.py : t.pz);
}

// ------------------------------------------------------------- building --
// Two regimes.  Nodes of more than kSub points ("big") are split level by
// level over the whole device: every kernel of a level walks only the points
// of the level's big nodes (an "active index" t maps to a node by binary
// search over the nodes' point-count prefix, then to its position).  A node
// of at most kSub points ("small") is handed to one workgroup, which loads
// its points into LDS and builds its whole subtree there, level by level
// with block barriers.
constexpr uint32_t kSub = 4096;           // small-node capacity (points)
constexpr int kSubThreads = 1024;
// Smaller subtree sizes use workgroups sized to them (four points per
// thread): 2048 points in about 80 KB of LDS and 1024 in about 41 KB, two
// and four workgroups per CU.
// While the search runs beside the build, its blocks hold every CU; a
// one-wave workgroup on at most kSubSmall points (about 22 KB of LDS) fits in
// the slot a finished search block frees, a kSub one (155 KB) does not.
constexpr uint32_t kSubSmall = 512;
constexpr int kSubSmallThreads = 64;
// Subtree nodes of at most wave_node (<= kWaveNodeMax, >= kLeafMax + 1)
// points are built by one wave each; bigger ones level by level by the
// whole workgroup.  Block-level nodes then hold more than kLeafMax points,
// so a level has fewer than kSub / (kLeafMax + 1) of them.
constexpr uint32_t kWaveNodeMax = 1024;
constexpr int kSubNodes = 48;             // block-level nodes per subtree level (> 4096 / 101)
constexpr int kWaveStack = 12;            // pending nodes of one wave (see wave_build)
constexpr int kWaveQ = 96;                // nodes handed to the waves per block level
// The level kernels run in one-wave workgroups: beside the search (whose
// workgroups are single waves too) a level kernel's workgroup then needs one
// free wave slot, not four on one CU at once, and is not held back until the
// search drains.
constexpr int kLvlThreads = 64;
constexpr int kPredThreads = kLvlThreads;
constexpr int kPredItems = 16;
constexpr uint32_t kPredTile = kPredThreads * kPredItems;
// Levels of up to this many prediction blocks (512K active points) compute
// each block's prefix in k_kd_pred_down from the raw block totals (at most
// eight loads per thread); larger levels pay one scan launch instead, which
// costs less than thousands of blocks each adding up thousands of totals.
constexpr uint32_t kRawSumBlocks = 512;
constexpr int kChildThreads = 256;       // one block; small enough to run beside the search
constexpr uint32_t kMinMaxLds = 128;      // big children staged per block in LDS (6 KB)

// Build header (device): the last level's big-node count and point count,
// subtree list length, overflow code, node count, and per level parity
// (A, P) at kHdrLvl + 2 * parity, so levels run back to back without the
// host reading their sizes.
enum : int {
    kHdrA = 0, kHdrP = 1, kHdrSmall = 2, kHdrOverflow = 3, kHdrNodes = 4, kHdrLvl = 8,
    kHdrWords = static_cast<int>(kKdHeaderWords)
};
enum : uint32_t { kOvfNodes = 1, kOvfLists = 2, kOvfSubtree = 3 };

// Exclusive scan of one value per thread across a block of NT threads.
template <int NT> __device__ __forceinline__ uint32_t kd_block_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t s_w[NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        const uint32_t x = s_w[i];
        if (i < w) off += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

__device__ __forceinline__ double pick3(const double v[3], int f)
{
    return f == 0 ? v[0] : (f == 1 ? v[1] : v[2]);  // no dynamic register indexing
}

// middleSplit_: among the axes whose box span is within 1e-5 of the widest,
// the one with the largest data spread; cut at the box middle clamped to the
// data range (nanoflann.hpp:779-827).
__device__ __forceinline__ void kd_choose_split(const double* b, const double mn[3],
                                                const double mx[3], int& f, double& cv)
{
    const double eps = 0.00001;
    double max_span = b[3] - b[0];
    for (int i = 1; i < 3; ++i) {
        const double span = b[3 + i] - b[i];
        if (span > max_span) max_span = span;
    }
    double max_spread = -1;
    f = 0;
    for (int i = 0; i < 3; ++i) {
        const double span = b[3 + i] - b[i];
        if (span > (1 - eps) * max_span) {
            const double spread = mx[i] - mn[i];
            if (spread > max_spread) {
                f = i;
                max_spread = spread;
            }
        }
    }
    const double split = (b[f] + b[3 + f]) / 2;
    const double lo = pick3(mn, f), hi = pick3(mx, f);
    cv = split;
    if (split < lo)
        cv = lo;
    else if (split > hi)
        cv = hi;
}

// The parent's divlow (max of child1) or divhigh (min of child2) on its cut
// axis, from this node's extent (nanoflann.hpp:672-683).
__device__ __forceinline__ void kd_parent_div(const KdView& t, uint32_t nd, const double mn[3],
                                              const double mx[3])
{
    const int par = t.parent[nd];
    if (par < 0) return;
    const int f = t.feat[par];
    if (t.which[nd] == 0)
        t.divlow[par] = pick3(mx, f);
    else
        t.divhigh[par] = pick3(mn, f);
}

// planeSplit's index rule: lim1 if more than half are strictly left, lim2 if
// fewer than half are left-or-equal, else the middle (nanoflann.hpp:663-670).
__device__ __forceinline__ uint32_t kd_split_index(uint32_t cnt, uint32_t lim1, uint32_t lim2)
{
    if (lim1 > cnt / 2) return lim1;
    if (lim2 < cnt / 2) return lim2;
    return cnt / 2;
}

// Records of the two children of nd: segments [lo, lo+idx) and [lo+idx, hi),
// boxes cut at the node's cut value, no children yet.
__device__ __forceinline__ void kd_make_children(const KdView& t, uint32_t nd, uint32_t lo,
                                                 uint32_t idx, uint32_t hi, uint32_t c1, int f,
                                                 double cv, const double* b)
{
    const uint32_t c2 = c1 + 1;
    t.c1[nd] = static_cast<int32_t>(c1);
    t.c2[nd] = static_cast<int32_t>(c2);
    t.split[nd] = lo + idx;
    double* b1 = t.box + 6 * static_cast<size_t>(c1);
    double* b2 = t.box + 6 * static_cast<size_t>(c2);
    for (int i = 0; i < 6; ++i) {
        b1[i] = b[i];
        b2[i] = b[i];
    }
    b1[3 + f] = cv;
    b2[f] = cv;
    t.nl[c1] = lo;
    t.nr[c1] = lo + idx;
    t.nl[c2] = lo + idx;
    t.nr[c2] = hi;
    t.parent[c1] = static_cast<int32_t>(nd);
    t.parent[c2] = static_cast<int32_t>(nd);
    t.which[c1] = 0;
    t.which[c2] = 1;
    t.c1[c1] = t.c2[c1] = t.c1[c2] = t.c2[c2] = -1;
    t.feat[c1] = t.feat[c2] = -1;
}

// The same with nd's split and box read from its record.
__device__ __forceinline__ void kd_make_children(const KdView& t, uint32_t nd, uint32_t lo,
                                                 uint32_t idx, uint32_t hi, uint32_t c1)
{
    kd_make_children(t, nd, lo, idx, hi, c1, t.feat[nd], t.cutval[nd],
                     t.box + 6 * static_cast<size_t>(nd));
}

// The parent's div as in kd_parent_div, the parent (par, its cut axis pf)
// and the side (which) known to the caller.
__device__ __forceinline__ void kd_set_div(const KdView& t, int par, int which, int pf,
                                           const double mn[3], const double mx[3])
{
    if (which == 0)
        t.divlow[par] = pick3(mx, pf);
    else
        t.divhigh[par] = pick3(mn, pf);
}

__global__ __launch_bounds__(kKdThreads) void k_kd_init(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, uint32_t n, KdView t)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    t.vind[p] = p;
    t.px[p] = x[p];
    t.py[p] = y[p];
    t.pz[p] = z[p];
}

// Node t = the root of tile t (the whole tile; its extent is the tile's
// bounds).  Roots of more than sub points start the big-node levels, in tile
// order; the others go straight to the subtree builder.
__global__ __launch_bounds__(kChildThreads) void k_kd_roots(KdView t, const uint64_t* __restrict__ starts,
                                                            uint32_t T,
                                                            const double* __restrict__ bounds,
                                                            KdScratch w, uint32_t sub)
{
    uint32_t nbig = 0, nsmall = 0, psum = 0;
    for (uint32_t r0 = 0; r0 < T; r0 += kChildThreads) {
        const uint32_t r = r0 + threadIdx.x;
        const bool ok = r < T;
        uint32_t cnt = 0;
        if (ok) {
            const uint32_t lo = static_cast<uint32_t>(starts[r]);
            const uint32_t hi = static_cast<uint32_t>(starts[r + 1]);
            cnt = hi - lo;
            t.nl[r] = lo;
            t.nr[r] = hi;
            t.parent[r] = -1;
            t.which[r] = 0;
            t.c1[r] = t.c2[r] = -1;
            t.feat[r] = -1;
            const double* b = bounds + 6 * static_cast<size_t>(r);
            for (int i = 0; i < 6; ++i) t.box[6 * static_cast<size_t>(r) + i] = b[i];
            for (int d = 0; d < 3; ++d) {
                t.mn[3 * r + d] = enc_order(b[d]);
                t.mx[3 * r + d] = enc_order(b[3 + d]);
            }
        }
        const bool big = ok && cnt > sub;
        uint32_t tb, ts, tz;
        const uint32_t eb = kd_block_scan<kChildThreads>(big ? 1u : 0u, &tb);
        const uint32_t es = kd_block_scan<kChildThreads>((ok && !big) ? 1u : 0u, &ts);
        const uint32_t ez = kd_block_scan<kChildThreads>(big ? cnt : 0u, &tz);
        if (big) {
            if (nbig + eb < w.list_cap) {
                w.act[0][nbig + eb] = r;
                w.off[0][nbig + eb] = psum + ez;
            } else {
                w.hdr[kHdrOverflow] = kOvfLists;
            }
        } else if (ok) {
            if (nsmall + es < w.list_cap)
                w.small[nsmall + es] = r;
            else
                w.hdr[kHdrOverflow] = kOvfLists;
        }
        nbig += tb;
        nsmall += ts;
        psum += tz;
    }
    if (threadIdx.x == 0) {
        if (nbig <= w.list_cap) w.off[0][nbig] = psum;
        w.hdr[kHdrA] = nbig;
        w.hdr[kHdrP] = psum;
        w.hdr[kHdrSmall] = nsmall;
        w.hdr[kHdrNodes] = T;
        const bool ovf = w.hdr[kHdrOverflow] != 0;
        w.hdr[kHdrLvl] = ovf ? 0u : nbig;
        w.hdr[kHdrLvl + 1] = ovf ? 0u : psum;
    }
}

// Largest k in [0, A) with off[k] <= t (big nodes are never empty).
__device__ __forceinline__ uint32_t act_find(const uint32_t* __restrict__ off, uint32_t A,
                                             uint32_t t)
{
    uint32_t lo = 0, hi = A;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// planeSplit's two Hoare passes as predicates -- pass 1 "coord < cut" over
// the node, pass 2 "coord <= cut" over [L, right) where L = lo + #pass 1 --
// over increasing active indices of one thread: the node's data (range, axis,
// cut, pass-2 start) stays in registers until the index crosses into the next
// node, so each item costs one coordinate load.
struct PredCursor {
    uint32_t k, t0, t1, lo, L;
    const double* ax;
    double cv;

    uint32_t nd;
    int f;

    // Pass 1 of the split-fused reduction (kSplit) chooses the node's split
    // from its extents and box here; the thread at the node's first active
    // index stores it (see pred).  Later kernels of the level read it.
    template <bool kSplit = false>
    __device__ __forceinline__ void load(const KdView& t, const uint32_t* act,
                                         const uint32_t* off, const uint32_t* s1, int pass,
                                         uint32_t kk)
    {
        k = kk;
        nd = act[k];
        lo = t.nl[nd];
        t0 = off[k];
        t1 = off[k + 1];
        if (kSplit) {
            double mn[3], mx[3];
            for (int d = 0; d < 3; ++d) {
                mn[d] = dec_order(t.mn[3 * nd + d]);
                mx[d] = dec_order(t.mx[3 * nd + d]);
            }
            kd_choose_split(t.box + 6 * static_cast<size_t>(nd), mn, mx, f, cv);
        } else {
            f = t.feat[nd];
            cv = t.cutval[nd];
        }
        ax = axis(t, f);
        L = pass == 1 ? 0u : lo + (s1[t1] - s1[t0]);
    }

    template <bool kSplit = false>
    __device__ __forceinline__ bool pred(const KdView& t, const uint32_t* act,
                                         const uint32_t* off, const uint32_t* s1, int pass,
                                         uint32_t A, uint32_t ti)
    {
        if (ti >= t1) {
            uint32_t kk = k;
            while (kk + 1 < A && off[kk + 1] <= ti) ++kk;
            load<kSplit>(t, act, off, s1, pass, kk);
        }
        if (kSplit && ti == t0) {
            // The node's one writer: its split, and the parent's div from
            // the node's extents.
            double mn[3], mx[3];
            for (int d = 0; d < 3; ++d) {
                mn[d] = dec_order(t.mn[3 * nd + d]);
                mx[d] = dec_order(t.mx[3 * nd + d]);
            }
            kd_parent_div(t, nd, mn, mx);
            t.feat[nd] = f;
            t.cutval[nd] = cv;
        }
        const uint32_t p = lo + (ti - t0);
        const double v = ax[p];
        return p >= L && (pass == 1 ? v < cv : v <= cv);
    }
};

__global__ __launch_bounds__(kPredThreads) void k_kd_pred_reduce(
    KdView t, int pass, const uint32_t* __restrict__ act, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ lvl, const uint32_t* __restrict__ s1,
    uint32_t* __restrict__ sums)
{
    const uint32_t A = lvl[0], P = lvl[1];
    const uint32_t base = blockIdx.x * kPredTile;
    if (base >= P) {  // past this level's points (grids are sized for a bound)
        if (threadIdx.x == 0) sums[blockIdx.x] = 0;
        return;
    }
    PredCursor c;
    uint32_t acc = 0;
    if (pass == 1) {
        // The level's splits are chosen here, by the blocks that cover each
        // node's points.
        c.load<true>(t, act, off, s1, pass, act_find(off, A, base));  // one search per block
        for (int i = 0; i < kPredItems; ++i) {
            const uint32_t ti = base + i * kPredThreads + threadIdx.x;
            if (ti < P && c.pred<true>(t, act, off, s1, pass, A, ti)) ++acc;
        }
    } else {
        c.load(t, act, off, s1, pass, act_find(off, A, base));
        for (int i = 0; i < kPredItems; ++i) {
            const uint32_t ti = base + i * kPredThreads + threadIdx.x;
            if (ti < P && c.pred(t, act, off, s1, pass, A, ti)) ++acc;
        }
    }
    uint32_t tot;
    kd_block_scan<kPredThreads>(acc, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// In-place exclusive scan of the nb per-block sums, one workgroup, NT * ITEMS
// sums per step: one wave (one free slot beside the search) while the sums
// take at most a few steps, four waves for larger levels.
template <int NT, int ITEMS>
__global__ __launch_bounds__(NT) void k_kd_scan_top(uint32_t* __restrict__ sums, uint32_t nb)
{
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += NT * ITEMS) {
        const uint32_t i0 = base + threadIdx.x * ITEMS;
        uint32_t v[ITEMS];
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            v[i] = i0 + i < nb ? sums[i0 + i] : 0u;
            acc += v[i];
        }
        uint32_t tot;
        uint32_t ex = kd_block_scan<NT>(acc, &tot) + carry;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (i0 + i < nb) sums[i0 + i] = ex;
            ex += v[i];
        }
        carry += tot;
    }
}

void launch_kd_scan_top(uint32_t* sums, uint32_t nb, hipStream_t s)
{
    if (nb <= 4u * 64u * 64u)
        hipLaunchKernelGGL((k_kd_scan_top<64, 64>), dim3(1), dim3(64), 0, s, sums, nb);
    else
        hipLaunchKernelGGL((k_kd_scan_top<256, 16>), dim3(1), dim3(256), 0, s, sums, nb);
}

// Exclusive prefix of the predicate over active indices [0, P], written to
// out (P + 1 entries; the predicate is 0 at P).
// With raw_sums the block adds up the block totals before it itself (levels
// of few blocks: no separate scan launch); otherwise sums holds their
// exclusive scan.
__global__ __launch_bounds__(kPredThreads) void k_kd_pred_down(
    KdView t, int pass, const uint32_t* __restrict__ act, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ lvl, const uint32_t* __restrict__ s1,
    const uint32_t* __restrict__ sums, uint32_t* __restrict__ out, int raw_sums)
{
    const uint32_t A = lvl[0], P = lvl[1];
    __shared__ uint32_t s[kPredTile];
    const uint32_t base = blockIdx.x * kPredTile;
    if (base > P) return;  // the block holding index P writes out[P]
    if (P == 0) {  // a queued level with no big node
        if (threadIdx.x == 0) out[0] = 0;
        return;
    }
    PredCursor c;
    c.load(t, act, off, s1, pass, act_find(off, A, base < P ? base : (P ? P - 1 : 0)));
    for (int i = 0; i < kPredItems; ++i) {
        const uint32_t l = i * kPredThreads + threadIdx.x;
        const uint32_t ti = base + l;
        s[l] = (ti < P && c.pred(t, act, off, s1, pass, A, ti)) ? 1u : 0u;
    }
    __syncthreads();
    uint32_t v[kPredItems];
    uint32_t acc = 0;
    for (int i = 0; i < kPredItems; ++i) {
        v[i] = s[threadIdx.x * kPredItems + i];
        acc += v[i];
    }
    uint32_t before = 0;
    if (raw_sums) {
        uint32_t part = 0;
        for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kPredThreads) part += sums[b];
        kd_block_scan<kPredThreads>(part, &before);
    } else {
        before = sums[blockIdx.x];
    }
    uint32_t tot;
    uint32_t ex = kd_block_scan<kPredThreads>(acc, &tot) + before;
    for (int i = 0; i < kPredItems; ++i) {
        s[threadIdx.x * kPredItems + i] = ex;
        ex += v[i];
    }
    __syncthreads();
    for (int i = 0; i < kPredItems; ++i) {
        const uint32_t l = i * kPredThreads + threadIdx.x;
        const uint32_t ti = base + l;
        if (ti <= P) out[ti] = s[l];
    }
}

// A node's data for one pass, kept in registers while a thread's active
// indices stay in the node: range start lo and active indices [t0, t1), the
// pass's range [a, e) and split m = a + #predicate, the cut.
struct LevelCursor {
    uint32_t k, t0, t1, lo, a, m, sct0, sct1;
    const double* ax;
    double cv;

    __device__ __forceinline__ void load(const KdView& t, const uint32_t* act,
                                         const uint32_t* off, const uint32_t* s1,
                                         const uint32_t* sc, int pass, uint32_t kk)
    {
        k = kk;
        const uint32_t nd = act[k];
        lo = t.nl[nd];
        t0 = off[k];
        t1 = off[k + 1];
        ax = axis(t, t.feat[nd]);
        cv = t.cutval[nd];
        a = pass == 1 ? lo : lo + (s1[t1] - s1[t0]);
        sct0 = sc[t0];
        sct1 = sc[t1];
        m = a + (sct1 - sct0);
    }

    __device__ __forceinline__ void seek(const KdView& t, const uint32_t* act,
                                         const uint32_t* off, const uint32_t* s1,
                                         const uint32_t* sc, int pass, uint32_t A, uint32_t ti)
    {
        if (ti < t1) return;
        uint32_t kk = k;
        while (kk + 1 < A && off[kk + 1] <= ti) ++kk;
        load(t, act, off, s1, sc, pass, kk);
    }

    // planeSplit's predicate at position p: pass 1 "coord < cut", pass 2
    // "coord <= cut" over [a, e).
    __device__ __forceinline__ bool pred(int pass, uint32_t p, double v) const
    {
        return pass == 1 ? v < cv : (p >= a && v <= cv);
    }
};

// Publish and swap take kLvlItems active indices per thread, kLvlThreads
// apart: a block covers kLvlTile consecutive indices, and a thread reloads
// node data only where its indices cross into the next node.
constexpr int kLvlItems = 8;
constexpr uint32_t kLvlTile = kLvlThreads * kLvlItems;

// Closed-form Hoare pairing: the r-th misplaced element of [a, m) counted
// from the left swaps with the r-th misplaced element of [m, e) counted from
// the right.  Misplaced right elements publish their position under their
// rank; misplaced left elements then swap with their partner.
__global__ __launch_bounds__(kLvlThreads) void k_kd_publish(
    KdView t, int pass, const uint32_t* __restrict__ act, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ lvl, const uint32_t* __restrict__ s1,
    const uint32_t* __restrict__ sc, uint32_t* __restrict__ tmp)
{
    const uint32_t A = lvl[0], P = lvl[1];
    const uint32_t base = blockIdx.x * kLvlTile;
    if (base >= P) return;
    LevelCursor c;
    c.load(t, act, off, s1, sc, pass, act_find(off, A, base));  // one search per block
#pragma unroll
    for (int i = 0; i < kLvlItems; ++i) {
        const uint32_t ti = base + i * kLvlThreads + threadIdx.x;
        if (ti >= P) break;
        c.seek(t, act, off, s1, sc, pass, A, ti);
        const uint32_t p = c.lo + (ti - c.t0);
        if (p < c.m) continue;
        if (c.pred(pass, p, c.ax[p])) tmp[c.a + (c.sct1 - sc[ti + 1])] = p;
    }
}

__global__ __launch_bounds__(kLvlThreads) void k_kd_swap(
    KdView t, int pass, const uint32_t* __restrict__ act, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ lvl, const uint32_t* __restrict__ s1,
    const uint32_t* __restrict__ sc, const uint32_t* __restrict__ tmp)
{
    const uint32_t A = lvl[0], P = lvl[1];
    const uint32_t base = blockIdx.x * kLvlTile;
    if (base >= P) return;
    LevelCursor c;
    c.load(t, act, off, s1, sc, pass, act_find(off, A, base));
#pragma unroll
    for (int i = 0; i < kLvlItems; ++i) {
        const uint32_t ti = base + i * kLvlThreads + threadIdx.x;
        if (ti >= P) break;
        c.seek(t, act, off, s1, sc, pass, A, ti);
        const uint32_t p = c.lo + (ti - c.t0);
        // Only positions in [a, m) are read here, and each only by its own
        // thread, before that thread writes it: swaps write [a, m) at the
        // writer's own position and [m, e) at the partner's.
        if (p < c.a || p >= c.m) continue;
        if (c.pred(pass, p, c.ax[p])) continue;
        // The predicate is 0 on [lo, a) in pass 2, so sc at a equals sc at t0.
        const uint32_t rank = (p - c.a) - (sc[ti] - c.sct0);
        const uint32_t q = tmp[c.a + rank];
        const uint32_t vi = t.vind[p];
        const double x = t.px[p], y = t.py[p], z = t.pz[p];
        t.vind[p] = t.vind[q];
        t.px[p] = t.px[q];
        t.py[p] = t.py[q];
        t.pz[p] = t.pz[q];
        t.vind[q] = vi;
        t.px[q] = x;
        t.py[q] = y;
        t.pz[q] = z;
    }
}

// One block: children of every big node of the level (ids base + 2k,
// base + 2k + 1), then the next level's big list with its point-count
// prefix, and the small children appended to the subtree list -- both in
// child order, so the tree's numbering is deterministic.
__global__ __launch_bounds__(kChildThreads) void k_kd_big_children(
    KdView t, const uint32_t* __restrict__ act, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ s1, const uint32_t* __restrict__ s2, KdScratch w, int nxt,
    uint32_t sub)
{
    const uint32_t A = w.hdr[kHdrLvl + 2 * (nxt ^ 1)];
    const uint32_t base = w.hdr[kHdrNodes];
    const uint32_t small0 = w.hdr[kHdrSmall];
    uint32_t nbig = 0, nsmall = 0, pnext = 0;
    for (uint32_t k0 = 0; k0 < A; k0 += kChildThreads) {
        const uint32_t k = k0 + threadIdx.x;
        uint32_t sa = 0, sb = 0, ca = 0;
        bool ok = k < A;
        if (ok) {
            const uint32_t nd = act[k];
            const uint32_t lo = t.nl[nd], hi = t.nr[nd], cnt = hi - lo;
            const uint32_t t0 = off[k], t1 = off[k + 1];
            const uint32_t lim1 = s1[t1] - s1[t0];
            const uint32_t lim2 = lim1 + (s2[t1] - s2[t0]);
            const uint32_t idx = kd_split_index(cnt, lim1, lim2);
            ca = base + 2 * k;
            if (ca + 1 >= t.cap) {
                w.hdr[kHdrOverflow] = kOvfNodes;
                ok = false;
            } else {
                kd_make_children(t, nd, lo, idx, hi, ca);
                sa = idx;
                sb = cnt - idx;
            }
        }
        const bool ba = ok && sa > sub, bb = ok && sb > sub;
        const uint32_t cb = (ba ? 1u : 0u) + (bb ? 1u : 0u);
        const uint32_t cs = ok ? 2u - cb : 0u;
        const uint32_t sz = (ba ? sa : 0u) + (bb ? sb : 0u);
        uint32_t tb, ts, tz;
        const uint32_t eb = kd_block_scan<kChildThreads>(cb, &tb);
        const uint32_t es = kd_block_scan<kChildThreads>(cs, &ts);
        const uint32_t ez = kd_block_scan<kChildThreads>(sz, &tz);
        if (ok) {
            uint32_t ib = nbig + eb, is = small0 + nsmall + es, pz = pnext + ez;
            const uint32_t ch[2] = {ca, ca + 1};
            const uint32_t chs[2] = {sa, sb};
            const bool chb[2] = {ba, bb};
            for (int c = 0; c < 2; ++c) {
                if (chb[c]) {
                    if (ib >= w.list_cap) {
                        w.hdr[kHdrOverflow] = kOvfLists;
                    } else {
                        w.act[nxt][ib] = ch[c];
                        w.off[nxt][ib] = pz;
                        t.aux[ch[c]] = ib;
                        for (int d = 0; d < 3; ++d) {
                            t.mn[3 * ch[c] + d] = ~0ull;
                            t.mx[3 * ch[c] + d] = 0ull;
                        }
                    }
                    ++ib;
                    pz += chs[c];
                } else {
                    if (is >= w.list_cap)
                        w.hdr[kHdrOverflow] = kOvfLists;
                    else
                        w.small[is] = ch[c];
                    t.aux[ch[c]] = ~0u;
                    ++is;
                }
            }
        }
        nbig += tb;
        nsmall += ts;
        pnext += tz;
    }
    __syncthreads();  // every thread's overflow store is visible below
    if (threadIdx.x == 0) {
        if (nbig <= w.list_cap) w.off[nxt][nbig] = pnext;
        w.hdr[kHdrA] = nbig;
        w.hdr[kHdrP] = pnext;
        w.hdr[kHdrSmall] = small0 + nsmall;
        w.hdr[kHdrNodes] = base + 2 * A;
        // After an overflow the next levels see no work (the build is redone).
        const bool ovf = w.hdr[kHdrOverflow] != 0;
        w.hdr[kHdrLvl + 2 * nxt] = ovf ? 0u : nbig;
        w.hdr[kHdrLvl + 2 * nxt + 1] = ovf ? 0u : pnext;
    }
}

// Extents of the next level's big nodes over this level's active points.
// Each wave walks a contiguous chunk of kMinMaxChunk active indices (children
// are contiguous runs of them), accumulating each lane's points while the
// lane's child stays the same; the lanes' accumulators are combined by a
// segmented reduction only at the chunk's end and where a lane crosses into
// another child, then one atomic per run goes to a per-block LDS table (few
// big children) or to global memory.
constexpr uint32_t kMinMaxIter = 16;
constexpr uint32_t kMinMaxChunk = 64 * kMinMaxIter;

__device__ __forceinline__ void minmax_flush(uint32_t j, unsigned long long (&lo)[3],
                                             unsigned long long (&hi)[3], int lane,
                                             bool use_lds, unsigned long long* s_lo,
                                             unsigned long long* s_hi, const KdView& t,
                                             const uint32_t* act_next)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t oj = __shfl_down(j, o);
        const bool same = lane + o < 64 && oj == j;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const unsigned long long ol = __shfl_down(lo[d], o);
            const unsigned long long oh = __shfl_down(hi[d], o);
            if (same) {
                lo[d] = ol < lo[d] ? ol : lo[d];
                hi[d] = oh > hi[d] ? oh : hi[d];
            }
        }
    }
    const uint32_t pj = __shfl_up(j, 1);
    if (j != ~0u && (lane == 0 || pj != j)) {
        if (use_lds) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                atomicMin(&s_lo[j * 3 + d], lo[d]);
                atomicMax(&s_hi[j * 3 + d], hi[d]);
            }
        } else {
            const uint32_t nd = act_next[j];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                atomicMin(&t.mn[3 * nd + d], lo[d]);
                atomicMax(&t.mx[3 * nd + d], hi[d]);
            }
        }
    }
}

// A lane's big node of this level, kept while its active indices stay in
// it: the index end t1, position = index + delta, the split position and the
// next-level list slots of the two children (~0u: a small child).
struct MinMaxCursor {
    uint32_t k, t1, delta, split, j1, j2;

    __device__ __forceinline__ void load(const KdView& t, const uint32_t* act,
                                         const uint32_t* off, uint32_t kk)
    {
        k = kk;
        const uint32_t nd = act[k];
        t1 = off[k + 1];
        delta = t.nl[nd] - off[k];
        split = t.split[nd];
        j1 = t.aux[t.c1[nd]];
        j2 = t.aux[t.c2[nd]];
    }
};

__global__ __launch_bounds__(kLvlThreads) void k_kd_big_minmax(
    KdView t, const uint32_t* __restrict__ act, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ lvl, const uint32_t* __restrict__ hdr,
    const uint32_t* __restrict__ act_next, uint32_t lds_max)
{
    __shared__ unsigned long long s_lo[kMinMaxLds * 3], s_hi[kMinMaxLds * 3];
    const uint32_t A = lvl[0], P = lvl[1];
    const uint32_t An = hdr[kHdrA];
    // After an overflow the children are incomplete; the build is redone.
    if (An == 0 || hdr[kHdrOverflow]) return;
    const bool use_lds = An <= min(lds_max, kMinMaxLds);
    if (use_lds) {
        for (uint32_t i = threadIdx.x; i < An * 3; i += blockDim.x) {
            s_lo[i] = ~0ull;
            s_hi[i] = 0ull;
        }
        __syncthreads();
    }
    const int lane = __lane_id();
    const uint32_t wpb = blockDim.x / 64;
    const uint32_t wave = blockIdx.x * wpb + threadIdx.x / 64;
    const uint32_t nwaves = gridDim.x * wpb;
    const uint32_t nchunks = (P + kMinMaxChunk - 1) / kMinMaxChunk;
    for (uint32_t c = wave; c < nchunks; c += nwaves) {
        const uint32_t c0 = c * kMinMaxChunk;
        MinMaxCursor cur;
        cur.load(t, act, off, act_find(off, A, c0));  // walked forward per lane
        uint32_t jc = ~0u;
        unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
        for (uint32_t r = 0; r < kMinMaxIter; ++r) {
            const uint32_t ti = c0 + r * 64 + lane;
            uint32_t j = ~0u;
            uint32_t p = 0;
            if (ti < P) {
                if (ti >= cur.t1) {
                    uint32_t kk = cur.k;
                    while (kk + 1 < A && off[kk + 1] <= ti) ++kk;
                    cur.load(t, act, off, kk);
                }
                p = ti + cur.delta;
                j = p < cur.split ? cur.j1 : cur.j2;
            }
            // A lane entering another child first hands its run to the wave.
            if (__ballot(jc != ~0u && j != jc)) {
                minmax_flush(jc, lo, hi, lane, use_lds, s_lo, s_hi, t, act_next);
                jc = ~0u;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    lo[d] = ~0ull;
                    hi[d] = 0ull;
                }
            }
            if (j != ~0u) {
                jc = j;
                const unsigned long long v[3] = {enc_order(t.px[p]), enc_order(t.py[p]),
                                                 enc_order(t.pz[p])};
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    lo[d] = v[d] < lo[d] ? v[d] : lo[d];
                    hi[d] = v[d] > hi[d] ? v[d] : hi[d];
                }
            }
        }
        if (__ballot(jc != ~0u)) minmax_flush(jc, lo, hi, lane, use_lds, s_lo, s_hi, t, act_next);
    }
    if (use_lds) {
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < An; j += blockDim.x) {
            if (s_hi[j * 3] == 0ull) continue;  // no point of this node in the block
            const uint32_t nd = act_next[j];
            for (int d = 0; d < 3; ++d) {
                atomicMin(&t.mn[3 * nd + d], s_lo[j * 3 + d]);
                atomicMax(&t.mx[3 * nd + d], s_hi[j * 3 + d]);
            }
        }
    }
}

// ----------------------------------------------------- small subtrees --
// One workgroup per small node: its points in LDS (coordinates, ids), the
// node's subtree built level by level -- extents by segmented reduction and
// LDS atomics, the two Hoare passes by block prefix sums and the same
// closed-form pairing, children records written to global memory -- and the
// permuted points written back.
// A node waiting on a wave's stack: its box travels with it, so a wave never
// reads back node records that another lane stored.
struct WaveFrame {
    double box[6];
    uint32_t nd;
    uint16_t lo, hi;
    int32_t par, pf;  // parent and its cut axis (the node is its child2)
};

// Parent argument of wave_build for a node whose parent is to be read from
// its record.
constexpr int kParentLookup = -2;

// LDS of one subtree workgroup: SUB points, NT threads.
template <uint32_t SUB, int NT> struct SubShared {
    static constexpr uint32_t kCap = SUB;
    static constexpr int kThreads = NT;
    double x[SUB], y[SUB], z[SUB];
    uint16_t v[SUB];                   // the point's slot at load (its id is read back at the end)
    uint16_t pre[SUB + 1];
    uint16_t tmp[SUB];
    uint8_t node[SUB];                 // level-local node index, 0xFF: not at block level
    unsigned long long mn[kSubNodes * 3], mx[kSubNodes * 3];
    double cut[kSubNodes];
    uint32_t nid[kSubNodes];
    uint16_t lo[kSubNodes], hi[kSubNodes], L[kSubNodes], L2[kSubNodes], sp[kSubNodes];
    uint8_t child[kSubNodes], child2[kSubNodes];
    int8_t feat[kSubNodes];
    uint32_t nnid[kSubNodes];
    uint16_t nlo[kSubNodes], nhi[kSubNodes];
    uint32_t wq_nd[kWaveQ];
    uint16_t wq_lo[kWaveQ], wq_hi[kWaveQ];
    int32_t wq_par[kWaveQ];                 // parent (kParentLookup: read it)
    int8_t wq_which[kWaveQ], wq_pf[kWaveQ];  // side, parent's cut axis
    WaveFrame frames[NT / 64][kWaveStack];
    uint32_t nc, nnext, nwq, wq_next;
    uint32_t id_base, id_next, id_cap;  // node ids reserved for this subtree
};

// Node ids.  One device-wide counter handing out two ids per split would be
// hit by every split of every workgroup; on a chip of eight XCDs each with
// its own L2, atomics on one address are performed one after another past the
// L2s, so each workgroup instead reserves about what its subtree needs
// (KD3 trees of leaf size <= 100 have ~0.035 nodes per point; the reservation
// is 0.042 per point + 2) and hands ids out from LDS.  A subtree that needs
// more takes the rest from the device counter, one pair at a time.
__device__ __forceinline__ uint32_t sub_reservation(uint32_t m)
{
    return m > kLeafMax ? 2u * ((m + 47u) / 48u) + 2u : 0u;
}

template <class SS> __device__ __forceinline__ uint32_t sub_alloc2(SS& S, uint32_t* hdr)
{
    const uint32_t k = atomicAdd(&S.id_next, 2u);
    return k + 2u <= S.id_cap ? S.id_base + k : atomicAdd(&hdr[kHdrNodes], 2u);
}

// Reserved ids the subtree did not use: records that no node links to, with
// an empty point range.
template <class SS>
__device__ __forceinline__ void sub_blank_unused(const SS& S, const KdView& t, uint32_t tid)
{
    for (uint32_t i = min(S.id_next, S.id_cap) + tid; i < S.id_cap; i += SS::kThreads) {
        const uint32_t nd = S.id_base + i;
        if (nd >= t.cap) break;
        t.nl[nd] = t.nr[nd] = 0;
        t.parent[nd] = -1;
        t.c1[nd] = t.c2[nd] = -1;
        t.feat[nd] = -1;
    }
}

template <class SS> __device__ __forceinline__ double sub_coord(const SS& S, int f, uint32_t i)
{
    return f == 0 ? S.x[i] : (f == 1 ? S.y[i] : S.z[i]);
}

// Exclusive block prefix of pred(i) over i in [0, SUB) into S.pre (pre[SUB]
// = total).  Thread tid owns items [16 tid, 16 tid + 16).
template <class SS, class Pred> __device__ __forceinline__ void sub_prefix(SS& S, Pred pred)
{
    constexpr int kItems = SS::kCap / SS::kThreads;
    const uint32_t i0 = threadIdx.x * kItems;
    uint32_t bits = 0, cnt = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i)
        if (pred(i0 + i)) {
            bits |= 1u << i;
            ++cnt;
        }
    uint32_t tot;
    uint32_t ex = kd_block_scan<SS::kThreads>(cnt, &tot);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        S.pre[i0 + i] = static_cast<uint16_t>(ex);
        ex += (bits >> i) & 1u;
    }
    if (threadIdx.x == 0) S.pre[SS::kCap] = static_cast<uint16_t>(tot);
    __syncthreads();
}

template <class SS> __device__ __forceinline__ void sub_swap(SS& S, uint32_t i, uint32_t q)
{
    const double x = S.x[i], y = S.y[i], z = S.z[i];
    const uint16_t v = S.v[i];
    S.x[i] = S.x[q];
    S.y[i] = S.y[q];
    S.z[i] = S.z[q];
    S.v[i] = S.v[q];
    S.x[q] = x;
    S.y[q] = y;
    S.z[q] = z;
    S.v[q] = v;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Min/max of points [lo, hi) of S on each axis, in every lane.
template <class SS>
__device__ __forceinline__ void wave_extents(const SS& S, uint32_t lo, uint32_t hi,
                                             int lane, double mn[3], double mx[3])
{
    unsigned long long l[3] = {~0ull, ~0ull, ~0ull}, h[3] = {0ull, 0ull, 0ull};
    for (uint32_t i = lo + lane; i < hi; i += 64) {
        const unsigned long long v[3] = {enc_order(S.x[i]), enc_order(S.y[i]), enc_order(S.z[i])};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            l[d] = v[d] < l[d] ? v[d] : l[d];
            h[d] = v[d] > h[d] ? v[d] : h[d];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const unsigned long long ol = __shfl_xor(l[d], o);
            const unsigned long long oh = __shfl_xor(h[d], o);
            l[d] = ol < l[d] ? ol : l[d];
            h[d] = oh > h[d] ? oh : h[d];
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        mn[d] = dec_order(l[d]);
        mx[d] = dec_order(h[d]);
    }
}

template <class SS>
__device__ __forceinline__ bool wave_pred(const SS& S, int f, double cv, int pass,
                                          uint32_t i)
{
    const double v = sub_coord(S, f, i);
    return pass == 1 ? v < cv : v <= cv;
}

// One planeSplit pass over [a, e) by one wave: returns c = #pred and leaves
// the predicate-true elements in [a, a + c), with the sequential two-pointer
// loop's swaps: the r-th misplaced element of [a, a + c) from the left
// trades places with the r-th misplaced element of [a + c, e) from the right.
template <class SS>
__device__ __forceinline__ uint32_t wave_pass(SS& S, int f, double cv, int pass,
                                              uint32_t a, uint32_t e, int lane)
{
    uint32_t c = 0;
    for (uint32_t i0 = a; i0 < e; i0 += 64) {
        const uint32_t i = i0 + lane;
        c += static_cast<uint32_t>(__popcll(__ballot(i < e && wave_pred(S, f, cv, pass, i))));
    }
    const uint32_t m = a + c;
    // Misplaced right elements, numbered from the left.
    uint32_t R = 0;
    for (uint32_t i0 = m; i0 < e; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool p = i < e && wave_pred(S, f, cv, pass, i);
        const uint64_t b = __ballot(p);
        if (p) S.tmp[a + R + lanes_below(b)] = static_cast<uint16_t>(i);
        R += static_cast<uint32_t>(__popcll(b));
    }
    if (R == 0) return c;
    wave_sync();
    uint32_t q = 0;
    for (uint32_t i0 = a; i0 < m; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool mis = i < m && !wave_pred(S, f, cv, pass, i);
        const uint64_t b = __ballot(mis);
        if (mis) sub_swap(S, i, S.tmp[a + (R - 1 - (q + lanes_below(b)))]);
        q += static_cast<uint32_t>(__popcll(b));
    }
    wave_sync();
    return c;
}

// The subtree of node nd (points [lo, hi) of S, which start at a0 in the
// tree arrays; box bx) built by one wave, depth first.  The wave continues
// into the first child and keeps the second on its stack unless it is a
// leaf (whose extents, for the parent's divhigh, are taken at once), so the
// stack holds only disjoint nodes of more than kLeafMax points: at most
// kWaveNodeMax / (kLeafMax + 1) < kWaveStack of them.  The node's parent
// par (kParentLookup: read from the record), its side and the parent's cut
// axis pf travel with it, as do box, cut and axis of the node being split,
// so the records written here are never read back.
template <class SS>
__device__ void wave_build(SS& S, const KdView& t, uint32_t* hdr, uint32_t a0,
                           uint32_t nd, uint32_t lo, uint32_t hi, const double* bx, int par,
                           int which, int pf, int lane, WaveFrame* stk)
{
    double box[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) box[i] = bx[i];
    int sp = 0;
    for (;;) {
        double mn[3], mx[3];
        wave_extents(S, lo, hi, lane, mn, mx);
        if (lane == 0) {
            if (par == kParentLookup)
                kd_parent_div(t, nd, mn, mx);
            else if (par >= 0)
                kd_set_div(t, par, which, pf, mn, mx);
        }
        const uint32_t cnt = hi - lo;
        bool down = false;
        if (cnt > kLeafMax) {
            // The children's ids: the atomic's round trip overlaps the passes.
            uint32_t c1 = 0;
            if (lane == 0) c1 = sub_alloc2(S, hdr);
            int f;
            double cv;
            kd_choose_split(box, mn, mx, f, cv);
            if (lane == 0) {
                t.feat[nd] = f;
                t.cutval[nd] = cv;
            }
            const uint32_t lim1 = wave_pass(S, f, cv, 1, lo, hi, lane);
            const uint32_t lim2 = lim1 + wave_pass(S, f, cv, 2, lo + lim1, hi, lane);
            const uint32_t idx = kd_split_index(cnt, lim1, lim2);
            c1 = __shfl(c1, 0);
            if (c1 + 1 >= t.cap) {
                if (lane == 0) hdr[kHdrOverflow] = kOvfNodes;  // the build is redone
            } else {
                if (lane == 0) kd_make_children(t, nd, a0 + lo, idx, a0 + hi, c1, f, cv, box);
                const uint32_t m2 = lo + idx;
                if (hi - m2 > kLeafMax) {
                    if (sp < kWaveStack) {
                        if (lane == 0) {
                            WaveFrame& fr = stk[sp];
                            for (int i = 0; i < 6; ++i) fr.box[i] = box[i];
                            fr.box[f] = cv;
                            fr.nd = c1 + 1;
                            fr.lo = static_cast<uint16_t>(m2);
                            fr.hi = static_cast<uint16_t>(hi);
                            fr.par = static_cast<int32_t>(nd);
                            fr.pf = f;
                        }
                        ++sp;
                    } else if (lane == 0) {
                        hdr[kHdrOverflow] = kOvfSubtree;
                    }
                } else {
                    double lmn[3], lmx[3];
                    wave_extents(S, m2, hi, lane, lmn, lmx);
                    if (lane == 0) kd_set_div(t, static_cast<int>(nd), 1, f, lmn, lmx);
                }
                box[3 + f] = cv;  // first child: the upper bound on the cut axis
                par = static_cast<int>(nd);
                which = 0;
                pf = f;
                nd = c1;
                hi = m2;
                down = true;
            }
        }
        if (down) continue;
        if (sp == 0) break;
        --sp;
        wave_sync();
        const WaveFrame& fr = stk[sp];
#pragma unroll
        for (int i = 0; i < 6; ++i) box[i] = fr.box[i];
        nd = fr.nd;
        lo = fr.lo;
        hi = fr.hi;
        par = fr.par;
        which = 1;
        pf = fr.pf;
        wave_sync();
    }
}

// The nodes queued for the waves, one wave per node until the queue is empty.
template <class SS>
__device__ __forceinline__ void wave_drain(SS& S, const KdView& t, uint32_t* hdr,
                                           uint32_t a0, int lane, int wv)
{
    for (;;) {
        uint32_t q = 0;
        if (lane == 0) q = atomicAdd(&S.wq_next, 1u);
        q = __shfl(q, 0);
        if (q >= S.nwq) break;
        const uint32_t nd = S.wq_nd[q];
        double bx[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) bx[i] = t.box[6 * static_cast<size_t>(nd) + i];
        wave_build(S, t, hdr, a0, nd, S.wq_lo[q], S.wq_hi[q], bx, S.wq_par[q], S.wq_which[q],
                   S.wq_pf[q], lane, S.frames[wv]);
    }
}

// At most 128 VGPRs per lane: a one-wave workgroup then fits in the registers
// one finished search wave frees.
template <uint32_t SUB, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_num_vgpr(128))) void k_kd_subtree(
    KdView t,
                                                           const uint32_t* __restrict__ small,
                                                           uint32_t* __restrict__ hdr,
                                                           uint32_t wave_node)
{
    __shared__ SubShared<SUB, NT> S;
    const uint32_t tid = threadIdx.x;
    const int lane = __lane_id(), wv = tid >> 6;
    const uint32_t nd0 = small[blockIdx.x];
    const uint32_t a0 = t.nl[nd0];
    const uint32_t m = t.nr[nd0] - a0;  // <= SUB
    if (tid == 0) {
        S.id_cap = sub_reservation(m);
        S.id_base = S.id_cap ? atomicAdd(&hdr[kHdrNodes], S.id_cap) : 0u;
        S.id_next = 0;
    }
    for (uint32_t i = tid; i < SUB; i += NT) {
        if (i < m) {
            S.x[i] = t.px[a0 + i];
            S.y[i] = t.py[a0 + i];
            S.z[i] = t.pz[a0 + i];
            S.v[i] = static_cast<uint16_t>(i);
            S.node[i] = 0;
        } else {
            S.node[i] = 0xFF;
        }
    }
    if constexpr (NT == 64) {
        // One wave: the node's whole subtree, depth first.
        __syncthreads();
        double bx[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) bx[i] = t.box[6 * static_cast<size_t>(nd0) + i];
        wave_build(S, t, hdr, a0, nd0, 0u, m, bx, kParentLookup, 0, 0, lane, S.frames[0]);
        __syncthreads();
    } else {
    if (tid == 0) {
        S.wq_next = 0;
        if (m <= wave_node) {  // small enough for one wave from the start
            S.nc = 0;
            S.nwq = 1;
            S.wq_nd[0] = nd0;
            S.wq_par[0] = kParentLookup;
            S.wq_which[0] = 0;
            S.wq_pf[0] = 0;
            S.wq_lo[0] = 0;
            S.wq_hi[0] = static_cast<uint16_t>(m);
        } else {
            S.nc = 1;
            S.nwq = 0;
            S.nid[0] = nd0;
            S.lo[0] = 0;
            S.hi[0] = static_cast<uint16_t>(m);  // m <= 4096 < 65536
        }
    }
    __syncthreads();
    for (;;) {
        // Nodes of at most wave_node points: one wave each, no block barrier.
        wave_drain(S, t, hdr, a0, lane, wv);
        __syncthreads();
        const uint32_t nc = S.nc;
        if (nc == 0) break;
        // Block level: nodes of more than wave_node points, all split.
        // (a) extents
        for (uint32_t i = tid; i < nc * 3; i += NT) {
            S.mn[i] = ~0ull;
            S.mx[i] = 0ull;
        }
        if (tid == 0) {
            S.nnext = 0;
            S.nwq = 0;
            S.wq_next = 0;
        }
        __syncthreads();
        for (uint32_t c = wv; c * 64 < m; c += NT / 64) {
            const uint32_t i = c * 64 + lane;
            const uint32_t j = i < m ? S.node[i] : 0xFFu;
            unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
            if (j != 0xFFu) {
                lo[0] = hi[0] = enc_order(S.x[i]);
                lo[1] = hi[1] = enc_order(S.y[i]);
                lo[2] = hi[2] = enc_order(S.z[i]);
            }
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t oj = __shfl_down(j, o);
                const bool same = lane + o < 64 && oj == j;
                for (int d = 0; d < 3; ++d) {
                    const unsigned long long ol = __shfl_down(lo[d], o);
                    const unsigned long long oh = __shfl_down(hi[d], o);
                    if (same) {
                        lo[d] = ol < lo[d] ? ol : lo[d];
                        hi[d] = oh > hi[d] ? oh : hi[d];
                    }
                }
            }
            const uint32_t pj = __shfl_up(j, 1);
            if (j != 0xFFu && (lane == 0 || pj != j))
                for (int d = 0; d < 3; ++d) {
                    atomicMin(&S.mn[j * 3 + d], lo[d]);
                    atomicMax(&S.mx[j * 3 + d], hi[d]);
                }
        }
        __syncthreads();
        // (b) one thread per node: parent's div and the split
        if (tid < nc) {
            const uint32_t nd = S.nid[tid];
            double mn[3], mx[3];
            for (int d = 0; d < 3; ++d) {
                mn[d] = dec_order(S.mn[tid * 3 + d]);
                mx[d] = dec_order(S.mx[tid * 3 + d]);
            }
            kd_parent_div(t, nd, mn, mx);
            int f;
            double cv;
            kd_choose_split(t.box + 6 * static_cast<size_t>(nd), mn, mx, f, cv);
            t.feat[nd] = f;
            t.cutval[nd] = cv;
            S.feat[tid] = static_cast<int8_t>(f);
            S.cut[tid] = cv;
        }
        __syncthreads();
        // (c) pass 1: coord < cut
        auto p1 = [&](uint32_t i) {
            const uint32_t j = S.node[i];
            if (j == 0xFFu) return false;
            return sub_coord(S, S.feat[j], i) < S.cut[j];
        };
        sub_prefix(S, p1);
        if (tid < nc)
            S.L[tid] = static_cast<uint16_t>(S.lo[tid] + (S.pre[S.hi[tid]] - S.pre[S.lo[tid]]));
        __syncthreads();
        for (uint32_t i = tid; i < m; i += NT) {
            const uint32_t j = S.node[i];
            if (j == 0xFFu) continue;
            if (i >= S.L[j] && p1(i)) S.tmp[S.lo[j] + (S.pre[S.hi[j]] - S.pre[i + 1])] = i;
        }
        __syncthreads();
        for (uint32_t i = tid; i < m; i += NT) {
            const uint32_t j = S.node[i];
            if (j == 0xFFu) continue;
            if (i < S.L[j] && !p1(i)) {
                const uint32_t r = (i - S.lo[j]) - (S.pre[i] - S.pre[S.lo[j]]);
                sub_swap(S, i, S.tmp[S.lo[j] + r]);
            }
        }
        __syncthreads();
        // (d) pass 2: coord <= cut over [L, hi)
        auto p2 = [&](uint32_t i) {
            const uint32_t j = S.node[i];
            if (j == 0xFFu || i < S.L[j]) return false;
            return sub_coord(S, S.feat[j], i) <= S.cut[j];
        };
        sub_prefix(S, p2);
        if (tid < nc)
            S.L2[tid] = static_cast<uint16_t>(S.L[tid] + (S.pre[S.hi[tid]] - S.pre[S.lo[tid]]));
        __syncthreads();
        for (uint32_t i = tid; i < m; i += NT) {
            const uint32_t j = S.node[i];
            if (j == 0xFFu) continue;
            if (i >= S.L2[j] && p2(i)) S.tmp[S.L[j] + (S.pre[S.hi[j]] - S.pre[i + 1])] = i;
        }
        __syncthreads();
        for (uint32_t i = tid; i < m; i += NT) {
            const uint32_t j = S.node[i];
            if (j == 0xFFu) continue;
            if (i >= S.L[j] && i < S.L2[j] && !p2(i)) {
                const uint32_t r = (i - S.L[j]) - (S.pre[i] - S.pre[S.L[j]]);
                sub_swap(S, i, S.tmp[S.L[j] + r]);
            }
        }
        __syncthreads();
        // (e) children: small ones to the waves' queue, big ones to the next
        // block level
        if (tid < nc) {
            const uint32_t nd = S.nid[tid];
            const uint32_t lo = S.lo[tid], hi = S.hi[tid];
            const uint32_t idx = kd_split_index(hi - lo, S.L[tid] - lo, S.L2[tid] - lo);
            const uint32_t c1 = sub_alloc2(S, hdr);
            S.sp[tid] = static_cast<uint16_t>(lo + idx);
            S.child[tid] = S.child2[tid] = 0xFF;
            if (c1 + 1 >= t.cap) {
                hdr[kHdrOverflow] = kOvfNodes;  // the build is redone
            } else {
                kd_make_children(t, nd, a0 + lo, idx, a0 + hi, c1, S.feat[tid], S.cut[tid],
                                 t.box + 6 * static_cast<size_t>(nd));
                const uint32_t clo[2] = {lo, lo + idx}, chi[2] = {lo + idx, hi};
                for (int c = 0; c < 2; ++c) {
                    if (chi[c] - clo[c] <= wave_node) {
                        const uint32_t q = atomicAdd(&S.nwq, 1u);
                        if (q < static_cast<uint32_t>(kWaveQ)) {
                            S.wq_nd[q] = c1 + c;
                            S.wq_par[q] = static_cast<int32_t>(nd);
                            S.wq_which[q] = static_cast<int8_t>(c);
                            S.wq_pf[q] = S.feat[tid];
                            S.wq_lo[q] = static_cast<uint16_t>(clo[c]);
                            S.wq_hi[q] = static_cast<uint16_t>(chi[c]);
                        } else {
                            hdr[kHdrOverflow] = kOvfSubtree;
                        }
                    } else {
                        const uint32_t slot = atomicAdd(&S.nnext, 1u);
                        if (slot < static_cast<uint32_t>(kSubNodes)) {
                            (c == 0 ? S.child : S.child2)[tid] = static_cast<uint8_t>(slot);
                            S.nnid[slot] = c1 + c;
                            S.nlo[slot] = static_cast<uint16_t>(clo[c]);
                            S.nhi[slot] = static_cast<uint16_t>(chi[c]);
                        } else {
                            hdr[kHdrOverflow] = kOvfSubtree;
                        }
                    }
                }
            }
        }
        __syncthreads();
        // (f) points move to their child's block-level slot (0xFF: a wave's)
        for (uint32_t i = tid; i < m; i += NT) {
            const uint32_t j = S.node[i];
            if (j == 0xFFu) continue;
            S.node[i] = i < S.sp[j] ? S.child[j] : S.child2[j];
        }
        __syncthreads();
        const uint32_t nn = min(S.nnext, static_cast<uint32_t>(kSubNodes));
        if (tid < nn) {
            S.nid[tid] = S.nnid[tid];
            S.lo[tid] = S.nlo[tid];
            S.hi[tid] = S.nhi[tid];
        }
        if (tid == 0) {
            S.nc = nn;
            S.nwq = min(S.nwq, static_cast<uint32_t>(kWaveQ));
        }
        __syncthreads();
    }
    }  // one wave or the whole workgroup
    sub_blank_unused(S, t, tid);
    // The workgroup owns [a0, a0 + m) of the point arrays: every old id is
    // read (through the slot each position now holds) before any is written.
    constexpr int kPer = static_cast<int>(SUB) / NT;
    uint32_t ids[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t i = tid + static_cast<uint32_t>(r * NT);
        ids[r] = i < m ? t.vind[a0 + S.v[i]] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t i = tid + static_cast<uint32_t>(r * NT);
        if (i >= m) continue;
        t.px[a0 + i] = S.x[i];
        t.py[a0 + i] = S.y[i];
        t.pz[a0 + i] = S.z[i];
        t.vind[a0 + i] = ids[r];
    }
}

// -------------------------------------------------------------- search --
// KD3Index::neighbors for one flagged row per wave, the row's own
// coordinates as the query.  The traversal (closer child first, the other
// child when the incremental box distance is <= the worst kept distance) is
// uniform across the wave; the recursion stack lives in lane-indexed
// registers (frame i in lane i).  A leaf's points are evaluated in parallel
// and those below the worst distance read at leaf entry are merged into the
// lane-held sorted list keyed by (d2, visit order): the capacity smallest by
// (d2, visit order) is exactly what KNNResultSet::addPoint keeps when the
// points arrive one by one with strict < (earlier points win ties).
__device__ __forceinline__ double lane_get(double v, int i)
{
    return __shfl(v, i);
}

__device__ __forceinline__ int lane_get(int v, int i)
{
    return __builtin_amdgcn_readlane(v, i);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

__device__ __forceinline__ uint32_t kd_tile_of(const uint64_t* __restrict__ starts, uint32_t T,
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

// Recursion frames of the exact search: the node, the child not taken, the
// cut axis and its two incremental distances, the node's box distance and
// the frame's stage (0 enter, 1 other child pending, 2 restore).
// RegFrames: frame i in lane i's registers, 64 frames (the tree's depth).
struct RegFrames {
    int f_node, f_other, f_idx, f_stage;
    double f_mind, f_cut, f_dst;
    int lane;
    __device__ int cap() const { return 64; }
    __device__ void init(int root, double mind)
    {
        f_node = root;
        f_other = -1;
        f_idx = 0;
        f_stage = 0;
        f_mind = mind;
        f_cut = 0.0;
        f_dst = 0.0;
    }
    __device__ int node(int i) const { return lane_get(f_node, i); }
    __device__ int stage(int i) const { return lane_get(f_stage, i); }
    __device__ int other(int i) const { return lane_get(f_other, i); }
    __device__ int idx(int i) const { return lane_get(f_idx, i); }
    __device__ double mind(int i) const { return lane_get(f_mind, i); }
    __device__ double cut(int i) const { return lane_get(f_cut, i); }
    __device__ double dst(int i) const { return lane_get(f_dst, i); }
    __device__ void push(int i, int nd, double m)
    {
        if (lane == i) {
            f_node = nd;
            f_stage = 0;
            f_mind = m;
        }
    }
    __device__ void set_split(int i, int oth, double c, int ix)
    {
        if (lane == i) {
            f_other = oth;
            f_cut = c;
            f_idx = ix;
            f_stage = 1;
        }
    }
    __device__ void set_dst(int i, double d)
    {
        if (lane == i) {
            f_dst = d;
            f_stage = 2;
        }
    }
};

// MemFrames: the frames in global memory, each lane keeping its own copy of
// every frame (the values are uniform across the wave, so every lane reads
// back what it wrote itself), for trees deeper than 64 levels.
struct MemFrame {
    int node, other, idx, stage;
    double mind, cut, dst;
};

struct MemFrames {
    MemFrame* base;  // depth x 64 frames of this wave
    int depth;
    int lane;
    __device__ int cap() const { return depth; }
    __device__ MemFrame& at(int i) const { return base[static_cast<size_t>(i) * 64 + lane]; }
    __device__ void init(int root, double mind)
    {
        MemFrame& f = at(0);
        f.node = root;
        f.other = -1;
        f.idx = 0;
        f.stage = 0;
        f.mind = mind;
        f.cut = 0.0;
        f.dst = 0.0;
    }
    __device__ int node(int i) const { return at(i).node; }
    __device__ int stage(int i) const { return at(i).stage; }
    __device__ int other(int i) const { return at(i).other; }
    __device__ int idx(int i) const { return at(i).idx; }
    __device__ double mind(int i) const { return at(i).mind; }
    __device__ double cut(int i) const { return at(i).cut; }
    __device__ double dst(int i) const { return at(i).dst; }
    __device__ void push(int i, int nd, double m)
    {
        MemFrame& f = at(i);
        f.node = nd;
        f.stage = 0;
        f.mind = m;
    }
    __device__ void set_split(int i, int oth, double c, int ix)
    {
        MemFrame& f = at(i);
        f.other = oth;
        f.cut = c;
        f.idx = ix;
        f.stage = 1;
    }
    __device__ void set_dst(int i, double d)
    {
        MemFrame& f = at(i);
        f.dst = d;
        f.stage = 2;
    }
};

// Result sets of the exact search, both keyed by (d2, visit order): the
// capacity smallest by that key is exactly what KNNResultSet::addPoint keeps
// when the points arrive one by one with strict < (earlier points win ties).
// RegResults: one entry per lane, capacity <= 64.
struct RegResults {
    Ent best;
    uint32_t cap;
    int lane;
    __device__ void init(uint32_t c, int l)
    {
        best = ent_inf();
        cap = c;
        lane = l;
    }
    __device__ double worst() const
    {
        const Ent w = ent_shfl(best, static_cast<int>(cap) - 1);
        return w.d == kInfD ? INFINITY : ent_d2(w);
    }
    // c: at most one candidate per lane (ent_inf() for none).
    __device__ void add(Ent c)
    {
        bitonic_sort64(c, lane);
        merge_sorted(best, c, lane);
    }
};

// LdsResults: a sorted array of up to cap (<= kBigCap) entries in LDS, merged
// out of place (buf -> other) with every entry's final rank: an existing
// entry moves up by the candidates below it, a candidate lands after the
// entries below it.  The workgroup is one wave.
constexpr uint32_t kBigCap = 1024;

struct LdsResults {
    Ent* buf;
    Ent* other;
    uint32_t cap;
    uint32_t count;
    int lane;
    __device__ void init(Ent* lds, uint32_t c, int l)
    {
        buf = lds;
        other = lds + c;
        cap = c;
        count = 0;
        lane = l;
    }
    __device__ double worst() const
    {
        return count < cap ? INFINITY : ent_d2(buf[cap - 1]);
    }
    __device__ void add(Ent c)
    {
        bitonic_sort64(c, lane);
        const uint32_t nc = static_cast<uint32_t>(__popcll(__ballot(c.d != kInfD)));
        // Existing entries: rank among the sorted candidates by binary search
        // over the lanes.
        for (uint32_t j0 = 0; j0 < count; j0 += 64) {
            const uint32_t j = j0 + static_cast<uint32_t>(lane);
            const Ent e = j < count ? buf[j] : ent_inf();
            uint32_t lo = 0, hi = nc;  // candidates [0, lo) are below e
            for (int step = 0; step < 7; ++step) {  // 2^7 > 64: converged
                const uint32_t mid = (lo + hi) >> 1;
                const Ent cm = ent_shfl(c, static_cast<int>(mid < 64 ? mid : 63));
                const bool below = ent_less(cm, e);
                if (lo < hi) {
                    if (below)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
            }
            const uint32_t pos = j + lo;
            if (j < count && pos < cap) other[pos] = e;
        }
        // Candidates: rank among the existing entries.
        if (static_cast<uint32_t>(lane) < nc) {
            uint32_t lo = 0, hi = count;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ent_less(buf[mid], c))
                    lo = mid + 1;
                else
                    hi = mid;
            }
            const uint32_t pos = static_cast<uint32_t>(lane) + lo;
            if (pos < cap) other[pos] = c;
        }
        __syncthreads();
        Ent* t = buf;
        buf = other;
        other = t;
        count = min(cap, count + nc);
    }
};

// The exact KD3Index search of one query (a uniform walk of the tree rooted
// at `root`, closer child first, the other child when the incremental box
// distance is <= the worst kept distance); returns true when the tree is
// deeper than the frames hold.
template <class Frames, class Results>
__device__ bool kd_traverse(const KdView& t, uint32_t root, const double q[3], Frames& fr,
                            Results& res)
{
    const int lane = __lane_id();
    const double* rb = t.box + 6 * static_cast<size_t>(root);  // the tile's bounds
    uint32_t seq = 0;
    double dists[3] = {0.0, 0.0, 0.0};
    double mind = 0.0;
    for (int i = 0; i < 3; ++i) {
        if (q[i] < rb[i]) {
            dists[i] = (q[i] - rb[i]) * (q[i] - rb[i]);
            mind += dists[i];
        }
        if (q[i] > rb[3 + i]) {
            dists[i] = (q[i] - rb[3 + i]) * (q[i] - rb[3 + i]);
            mind += dists[i];
        }
    }
    fr.init(static_cast<int>(root), mind);
    const int depth_cap = fr.cap();
    int sp = 1;
    while (sp > 0) {
        const int top = sp - 1;
        const int nd = fr.node(top);
        const int stage = fr.stage(top);
        if (stage == 0) {
            if (t.c1[nd] < 0) {
                const double worst = res.worst();
                const uint32_t l0 = t.nl[nd], l1 = t.nr[nd];
                for (uint32_t i0 = l0; i0 < l1; i0 += 64) {
                    const uint32_t i = i0 + lane;
                    Ent c = ent_inf();
                    if (i < l1) {
                        double d = 0.0;
                        const double d0 = q[0] - t.px[i];
                        d += d0 * d0;
                        const double d1 = q[1] - t.py[i];
                        d += d1 * d1;
                        const double d2 = q[2] - t.pz[i];
                        d += d2 * d2;
                        if (d < worst)
                            c = Ent{static_cast<uint64_t>(__double_as_longlong(d)),
                                    seq + (i - l0), t.vind[i]};
                    }
                    if (__ballot(c.d != kInfD)) res.add(c);
                }
                seq += l1 - l0;
                --sp;
                continue;
            }
            const int idx = t.feat[nd];
            const double val = q[idx];
            const double dl = t.divlow[nd], dh = t.divhigh[nd];
            const double diff1 = val - dl;
            const double diff2 = val - dh;
            int bestc, other;
            double cut;
            if ((diff1 + diff2) < 0) {
                bestc = t.c1[nd];
                other = t.c2[nd];
                cut = (val - dh) * (val - dh);
            } else {
                bestc = t.c2[nd];
                other = t.c1[nd];
                cut = (val - dl) * (val - dl);
            }
            const double fm = fr.mind(top);
            fr.set_split(top, other, cut, idx);
            if (sp >= depth_cap) return true;
            fr.push(sp, bestc, fm);
            ++sp;
        } else if (stage == 1) {
            const int idx = fr.idx(top);
            const double cut = fr.cut(top);
            const double dst = dists[idx];
            const double m2 = fr.mind(top) + cut - dst;
            dists[idx] = cut;
            fr.set_dst(top, dst);
            if (m2 * 1.0 <= res.worst()) {
                if (sp >= depth_cap) return true;
                fr.push(sp, fr.other(top), m2);
                ++sp;
            }
        } else {
            dists[fr.idx(top)] = fr.dst(top);
            --sp;
        }
    }
    return false;
}

// A flagged row's tile: its root node (tile r's tree is rooted at node r),
// the id short strided lists are padded with, and its neighborhood size.
struct RowShape {
    uint32_t root, pad_id, cap, k, stride;
};

__device__ __forceinline__ RowShape row_shape(const FusedProgram& prog0,
                                              const TileInfo* __restrict__ tiles,
                                              const uint64_t* __restrict__ starts,
                                              uint32_t ntiles, uint32_t s)
{
    RowShape r{0, 0, static_cast<uint32_t>(prog0.fetch), static_cast<uint32_t>(prog0.k),
               static_cast<uint32_t>(prog0.stride)};
    if (tiles) {
        r.root = kd_tile_of(starts, ntiles, s);
        r.pad_id = static_cast<uint32_t>(starts[r.root]);  // KD3Index pads with the view's first point
        const FusedProgram& pr = tiles[r.root].prog;
        r.cap = static_cast<uint32_t>(pr.fetch);
        r.k = static_cast<uint32_t>(pr.k);
        r.stride = static_cast<uint32_t>(pr.stride);
    }
    return r;
}

// The covariance record of a row from its neighbors' point ids (one per
// lane and chunk of 64), relative to the query point (exact differences),
// so the mean carries no rounding of the cloud's absolute offset.
template <class IdOf>
__device__ void write_cov(const KnnOutputs& out, uint32_t s, uint32_t k,
                          const double* __restrict__ x, const double* __restrict__ y,
                          const double* __restrict__ z, const double q[3], IdOf id_of)
{
    const int lane = __lane_id();
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (uint32_t o0 = 0; o0 < k; o0 += 64) {
        const uint32_t o = o0 + static_cast<uint32_t>(lane);
        if (o < k) {
            const uint32_t id = id_of(o);
            sx += x[id] - q[0];
            sy += y[id] - q[1];
            sz += z[id] - q[2];
        }
    }
    const double mx = wave_sum(sx) / static_cast<double>(k),
                 my = wave_sum(sy) / static_cast<double>(k),
                 mz = wave_sum(sz) / static_cast<double>(k);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0;
    for (uint32_t o0 = 0; o0 < k; o0 += 64) {
        const uint32_t o = o0 + static_cast<uint32_t>(lane);
        if (o < k) {
            const uint32_t id = id_of(o);
            const double dx = (x[id] - q[0]) - mx, dy = (y[id] - q[1]) - my,
                         dz = (z[id] - q[2]) - mz;
            a0 += dx * dx;
            a1 += dx * dy;
            a2 += dx * dz;
            a3 += dy * dy;
            a4 += dy * dz;
            a5 += dz * dz;
        }
    }
    const double den = static_cast<double>(k - 1);
    const double c0 = wave_sum(a0) / den, c1 = wave_sum(a1) / den, c2 = wave_sum(a2) / den,
                 c3 = wave_sum(a3) / den, c4 = wave_sum(a4) / den, c5 = wave_sum(a5) / den;
    if (lane == 0) {
        double* rec = out.cov + kCovRecord * static_cast<uint64_t>(s);
        rec[0] = c0;
        rec[1] = c1;
        rec[2] = c2;
        rec[3] = c3;
        rec[4] = c4;
        rec[5] = c5;
    }
}

// One flagged row (flagged[f]) re-solved by one wave with a register result
// set (capacity <= 64); returns true when the tree is deeper than the frames
// hold (nothing written for the row).
template <class Frames>
__device__ bool kd_resolve_row(const KdView& t, const CloudView& cv, const FusedProgram& prog0,
                               int mode, const KnnOutputs& out, const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               const TileInfo* __restrict__ tiles,
                               const uint64_t* __restrict__ starts, uint32_t ntiles,
                               uint32_t f, Frames& fr, Ent*)
{
    const int lane = __lane_id();
    const uint32_t s = out.flagged[f];
    const RowShape sh = row_shape(prog0, tiles, starts, ntiles, s);
    const uint32_t qid = cv.perm[s];
    const double q[3] = {x[qid], y[qid], z[qid]};
    RegResults res;
    res.init(sh.cap, lane);
    if (kd_traverse(t, sh.root, q, fr, res)) return true;
    // KD3ImplT::neighbors: every stride-th entry of the result set; slots
    // past the points found keep the output's zero initialisation.
    const Ent kept = ent_shfl(res.best, static_cast<int>((static_cast<uint32_t>(lane) * sh.stride) & 63u));
    const uint32_t kept_id = kept.d == kInfD ? sh.pad_id : kept.j;
    if (mode != 0) {  // neighbor lists
        if (static_cast<uint32_t>(lane) < sh.k)
            out.ids[static_cast<uint64_t>(qid) * sh.k + lane] = kept_id;
        return false;
    }
    // k <= 64: neighbor o is lane o's.
    write_cov(out, s, sh.k, x, y, z, q, [&](uint32_t) { return kept_id; });
    return false;
}

// The same with the result set in LDS (capacity up to kBigCap): the
// neighborhoods of more than 63 neighbors (knn + 1) * stride.
template <class Frames>
__device__ bool kd_resolve_row_big(const KdView& t, const CloudView& cv,
                                   const FusedProgram& prog0, int mode, const KnnOutputs& out,
                                   const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ z,
                                   const TileInfo* __restrict__ tiles,
                                   const uint64_t* __restrict__ starts, uint32_t ntiles,
                                   uint32_t f, Frames& fr, Ent* lds)
{
    const int lane = __lane_id();
    const uint32_t s = out.flagged[f];
    const RowShape sh = row_shape(prog0, tiles, starts, ntiles, s);
    const uint32_t qid = cv.perm[s];
    const double q[3] = {x[qid], y[qid], z[qid]};
    LdsResults res;
    res.init(lds, sh.cap, lane);
    if (kd_traverse(t, sh.root, q, fr, res)) {
        __syncthreads();
        return true;
    }
    const Ent* kept = res.buf;
    const uint32_t found = res.count;
    auto id_of = [&](uint32_t o) {
        const uint32_t e = o * sh.stride;
        return e < found ? kept[e].j : sh.pad_id;
    };
    if (mode != 0) {
        for (uint32_t o = static_cast<uint32_t>(lane); o < sh.k; o += 64)
            out.ids[static_cast<uint64_t>(qid) * sh.k + o] = id_of(o);
    } else {
        write_cov(out, s, sh.k, x, y, z, q, id_of);
    }
    __syncthreads();  // the next row reuses the LDS
    return false;
}

// Rows whose search needs more than 64 frames are counted in failed[0] and,
// with deep_rows, listed there for k_kd_resolve_deep.  BIG: result sets in
// LDS (one-wave workgroups, 2 * cap entries of dynamic LDS).
template <bool BIG>
__global__ __launch_bounds__(256) void k_kd_resolve(KdView t, CloudView cv, FusedProgram prog0,
                                                    int mode, KnnOutputs out,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z,
                                                    uint32_t* __restrict__ failed,
                                                    uint32_t* __restrict__ deep_rows,
                                                    const TileInfo* __restrict__ tiles,
                                                    const uint64_t* __restrict__ starts,
                                                    uint32_t ntiles)
{
    extern __shared__ Ent lds_ents[];
    const int lane = __lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    const uint32_t nflag = *out.nflagged;
    for (uint32_t f = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; f < nflag; f += waves) {
        RegFrames fr;
        fr.lane = lane;
        const bool over =
            BIG ? kd_resolve_row_big(t, cv, prog0, mode, out, x, y, z, tiles, starts, ntiles, f,
                                     fr, lds_ents)
                : kd_resolve_row(t, cv, prog0, mode, out, x, y, z, tiles, starts, ntiles, f, fr,
                                 lds_ents);
        if (over && lane == 0) {
            const uint32_t slot = atomicAdd(failed, 1u);
            if (deep_rows) deep_rows[slot] = f;
        }
    }
}

// The rows k_kd_resolve listed, with frames for the tree's whole depth in
// global memory (frames: waves x depth x 64); rows that still overflow are
// counted in failed2.
template <bool BIG>
__global__ __launch_bounds__(64) void k_kd_resolve_deep(KdView t, CloudView cv,
                                                        FusedProgram prog0, int mode,
                                                        KnnOutputs out,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const uint32_t* __restrict__ deep_rows,
                                                        const uint32_t* __restrict__ ndeep,
                                                        MemFrame* __restrict__ frames, int depth,
                                                        uint32_t* __restrict__ failed2,
                                                        const TileInfo* __restrict__ tiles,
                                                        const uint64_t* __restrict__ starts,
                                                        uint32_t ntiles)
{
    extern __shared__ Ent lds_ents[];
    const int lane = __lane_id();
    const uint32_t nd = *ndeep;
    MemFrames fr;
    fr.base = frames + static_cast<size_t>(blockIdx.x) * static_cast<size_t>(depth) * 64;
    fr.depth = depth;
    fr.lane = lane;
    for (uint32_t i = blockIdx.x; i < nd; i += gridDim.x) {
        const uint32_t f = deep_rows[i];
        const bool over =
            BIG ? kd_resolve_row_big(t, cv, prog0, mode, out, x, y, z, tiles, starts, ntiles, f,
                                     fr, lds_ents)
                : kd_resolve_row(t, cv, prog0, mode, out, x, y, z, tiles, starts, ntiles, f, fr,
                                 lds_ents);
        if (over && lane == 0) atomicAdd(failed2, 1u);
    }
}

// Every row to the exact search (clouds smaller than the result capacity).
__global__ __launch_bounds__(256) void k_flag_all(uint64_t n, const uint32_t* __restrict__ perm,
                                                  int mode, KnnOutputs out)
{
    const uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (s == 0) out.nflagged[0] = static_cast<uint32_t>(n);
    if (s >= n) return;
    out.flagged[s] = static_cast<uint32_t>(s);
    if (mode != 0)
        out.status[0][perm[s]] = 0;
    else
        reinterpret_cast<uint64_t*>(out.cov)[kCovRecord * s + 6] = ST_EXACT_ONLY;
}

// Rows of the tiles with fewer points than their result capacity go to the
// exact search.
__global__ __launch_bounds__(256) void k_flag_tiles(uint64_t n, const uint32_t* __restrict__ perm,
                                                    int mode, KnnOutputs out,
                                                    const TileInfo* __restrict__ tiles,
                                                    const uint64_t* __restrict__ starts,
                                                    uint32_t ntiles)
{
    const uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (s >= n) return;
    const uint32_t tl = kd_tile_of(starts, ntiles, s);
    if (!tiles[tl].all_exact) return;
    const uint32_t slot = atomicAdd(out.nflagged, 1u);
    out.flagged[slot] = static_cast<uint32_t>(s);
    if (mode != 0)
        out.status[0][perm[s]] = 0;
    else
        reinterpret_cast<uint64_t*>(out.cov)[kCovRecord * s + 6] = ST_EXACT_ONLY;
}

} // namespace

void launch_flag_tiles(uint64_t n, const uint32_t* perm, int mode, const KnnOutputs& out,
                       const TileTables& tt, hipStream_t s)
{
    hipLaunchKernelGGL(k_flag_tiles, dim3(div_up(std::max<uint64_t>(n, 1), 256)), dim3(256), 0,
                       s, n, perm, mode, out, tt.tiles, tt.starts, tt.ntiles);
}

void launch_flag_all(uint64_t n, const uint32_t* perm, int mode, const KnnOutputs& out,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_flag_all, dim3(div_up(std::max<uint64_t>(n, 1), 256)), dim3(256), 0, s,
                       n, perm, mode, out);
}

uint64_t kd_sums_words(uint64_t n) { return div_up(n + 1, kPredTile) + 2; }

uint32_t kd_node_capacity(uint64_t n) { return static_cast<uint32_t>(n / 12 + 1024); }

// Builds the tree; returns 0, or an overflow code (kOvfNodes: node
// capacity, kOvfLists: big/small node lists) after which the caller grows
// that capacity and builds again.
uint32_t launch_kd_build(const double* x, const double* y, const double* z, uint64_t n,
                         const uint64_t* starts, uint32_t ntiles, const double* bounds,
                         KdView t, KdScratch w, const KdTuning& kt, uint32_t* nodes,
                         hipStream_t s)
{
    const uint32_t sub = std::min<uint32_t>(std::max<uint32_t>(kt.subtree_points, kLeafMax + 1),
                                            kt.small_blocks ? kSubSmall : kSub);
    hipLaunchKernelGGL(k_kd_init, dim3(div_up(n, kKdThreads)), dim3(kKdThreads), 0, s, x, y, z,
                       static_cast<uint32_t>(n), t);
    launch_fill_u32(w.hdr, kHdrWords, 0u, s);
    hipLaunchKernelGGL(k_kd_roots, dim3(1), dim3(kChildThreads), 0, s, t, starts, ntiles,
                       bounds, w, sub);
    uint32_t h[kHdrWords];
    GF_HIP(hipMemcpyAsync(h, w.hdr, sizeof(h), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (h[kHdrOverflow]) return h[kHdrOverflow];
    uint32_t A = h[kHdrA], P = h[kHdrP], nsmall = h[kHdrSmall];
    int cur = 0;
    // Levels are queued kLevelsPerRead at a time; each kernel reads its
    // level's (A, P) from the header.  Grids are sized from bounds known on
    // the host: P never grows from one level to the next, A at most doubles.
    constexpr int kLevelsPerRead = 4;
    while (A) {
        uint64_t Ab = A;
        const uint32_t Pb = P;
        for (int l = 0; l < kLevelsPerRead; ++l) {
            const uint32_t* lvl = w.hdr + kHdrLvl + 2 * cur;
            // Pass 1's reduction chooses the level's splits (no split
            // launch); levels of at most kRawSumBlocks blocks skip the scan
            // launch (each block adds up the block totals before it).
            for (int pass = 1; pass <= 2; ++pass) {
                uint32_t* sc = pass == 1 ? w.s1 : w.s2;
                const uint32_t nb = div_up(static_cast<uint64_t>(Pb) + 1, kPredTile);
                hipLaunchKernelGGL(k_kd_pred_reduce, dim3(nb), dim3(kPredThreads), 0, s, t,
                                   pass, w.act[cur], w.off[cur], lvl, w.s1, w.sums);
                const int raw = nb <= kRawSumBlocks ? 1 : 0;
                if (!raw) launch_kd_scan_top(w.sums, nb, s);
                hipLaunchKernelGGL(k_kd_pred_down, dim3(nb), dim3(kPredThreads), 0, s, t,
                                   pass, w.act[cur], w.off[cur], lvl, w.s1, w.sums, sc, raw);
                hipLaunchKernelGGL(k_kd_publish, dim3(div_up(Pb, kLvlTile)), dim3(kLvlThreads),
                                   0, s, t, pass, w.act[cur], w.off[cur], lvl, w.s1, sc, w.tmp);
                hipLaunchKernelGGL(k_kd_swap, dim3(div_up(Pb, kLvlTile)), dim3(kLvlThreads), 0,
                                   s, t, pass, w.act[cur], w.off[cur], lvl, w.s1, sc, w.tmp);
            }
            hipLaunchKernelGGL(k_kd_big_children, dim3(1), dim3(kChildThreads), 0, s, t,
                               w.act[cur], w.off[cur], w.s1, w.s2, w, cur ^ 1, sub);
            hipLaunchKernelGGL(
                k_kd_big_minmax,
                dim3(std::min<uint32_t>(div_up(Pb, kMinMaxChunk), 4096)),
                dim3(kLvlThreads), 0, s, t, w.act[cur], w.off[cur], lvl, w.hdr, w.act[cur ^ 1],
                kt.minmax_lds_nodes);
            cur ^= 1;
            Ab *= 2;
        }
        GF_HIP(hipMemcpyAsync(h, w.hdr, sizeof(h), hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        if (h[kHdrOverflow]) return h[kHdrOverflow];
        A = h[kHdrLvl + 2 * cur];
        P = h[kHdrLvl + 2 * cur + 1];
        nsmall = h[kHdrSmall];
    }
    if (nsmall) {
        const uint32_t wave_node = std::min<uint32_t>(
            std::max<uint32_t>(kt.wave_node, kLeafMax + 1), kWaveNodeMax);
        if (kt.small_blocks)
            hipLaunchKernelGGL((k_kd_subtree<kSubSmall, kSubSmallThreads>), dim3(nsmall),
                               dim3(kSubSmallThreads), 0, s, t, w.small, w.hdr, wave_node);
        else if (sub <= kSub / 4)
            hipLaunchKernelGGL((k_kd_subtree<kSub / 4, kSubThreads / 4>), dim3(nsmall),
                               dim3(kSubThreads / 4), 0, s, t, w.small, w.hdr, wave_node);
        else if (sub <= kSub / 2)
            hipLaunchKernelGGL((k_kd_subtree<kSub / 2, kSubThreads / 2>), dim3(nsmall),
                               dim3(kSubThreads / 2), 0, s, t, w.small, w.hdr, wave_node);
        else
            hipLaunchKernelGGL((k_kd_subtree<kSub, kSubThreads>), dim3(nsmall),
                               dim3(kSubThreads), 0, s, t, w.small, w.hdr, wave_node);
        GF_HIP(hipGetLastError());
    }
    GF_HIP(hipMemcpyAsync(h, w.hdr, sizeof(h), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (h[kHdrOverflow]) return h[kHdrOverflow];
    // Subtree reservations may run past the capacity without using those ids.
    *nodes = std::min<uint32_t>(h[kHdrNodes], t.cap);
    return 0;
}

uint32_t kd_list_capacity(uint64_t n) { return static_cast<uint32_t>(n / kSubSmall * 4 + 4096); }

// The rows' result capacity decides between register and LDS result sets.
static uint32_t max_capacity(const FusedProgram& prog, const TileTables* tt)
{
    return tt ? 64u : static_cast<uint32_t>(prog.fetch);  // batch tiles: fetch <= 63
}

void launch_kd_resolve(const KdView& t, const CloudView& cv, const FusedProgram& prog, int mode,
                       const KnnOutputs& out, const double* x, const double* y, const double* z,
                       uint32_t nflagged, uint32_t* failed, uint32_t* deep_rows, hipStream_t s,
                       const TileTables* tt)
{
    if (!nflagged) return;
    const uint32_t cap = max_capacity(prog, tt);
    const TileInfo* tiles = tt ? tt->tiles : nullptr;
    const uint64_t* starts = tt ? tt->starts : nullptr;
    const uint32_t ntiles = tt ? tt->ntiles : 0u;
    if (cap <= 64) {
        const uint32_t blocks = std::min<uint32_t>(div_up(nflagged, 4), 4096);
        hipLaunchKernelGGL(k_kd_resolve<false>, dim3(blocks), dim3(256), 0, s, t, cv, prog, mode,
                           out, x, y, z, failed, deep_rows, tiles, starts, ntiles);
    } else {
        const uint32_t blocks = std::min<uint32_t>(nflagged, 16384);
        const size_t lds = sizeof(Ent) * 2 * cap;
        hipLaunchKernelGGL(k_kd_resolve<true>, dim3(blocks), dim3(64), lds, s, t, cv, prog, mode,
                           out, x, y, z, failed, deep_rows, tiles, starts, ntiles);
    }
}

uint32_t kd_max_capacity() { return kBigCap; }

size_t kd_deep_frame_bytes() { return sizeof(MemFrame) * 64; }

void launch_kd_resolve_deep(const KdView& t, const CloudView& cv, const FusedProgram& prog,
                            int mode, const KnnOutputs& out, const double* x, const double* y,
                            const double* z, const uint32_t* deep_rows, const uint32_t* ndeep,
                            void* frames, int depth, uint32_t waves, uint32_t* failed2,
                            hipStream_t s, const TileTables* tt)
{
    if (!waves) return;
    const uint32_t cap = max_capacity(prog, tt);
    const TileInfo* tiles = tt ? tt->tiles : nullptr;
    const uint64_t* starts = tt ? tt->starts : nullptr;
    const uint32_t ntiles = tt ? tt->ntiles : 0u;
    MemFrame* fp = static_cast<MemFrame*>(frames);
    if (cap <= 64)
        hipLaunchKernelGGL(k_kd_resolve_deep<false>, dim3(waves), dim3(64), 0, s, t, cv, prog,
                           mode, out, x, y, z, deep_rows, ndeep, fp, depth, failed2, tiles,
                           starts, ntiles);
    else
        hipLaunchKernelGGL(k_kd_resolve_deep<true>, dim3(waves), dim3(64), sizeof(Ent) * 2 * cap,
                           s, t, cv, prog, mode, out, x, y, z, deep_rows, ndeep, fp, depth,
                           failed2, tiles, starts, ntiles);
}

// Depth of the deepest node (root 0) of a tree given its child arrays, or of
// the deepest of several roots [0, nroots).
int kd_tree_depth(const std::vector<int32_t>& c1, const std::vector<int32_t>& c2,
                  uint32_t nroots)
{
    int best = 0;
    std::vector<std::pair<int32_t, int>> st;
    for (uint32_t r = 0; r < nroots; ++r) {
        st.push_back({static_cast<int32_t>(r), 1});
        while (!st.empty()) {
            const auto [nd, d] = st.back();
            st.pop_back();
            best = std::max(best, d);
            if (nd < 0 || static_cast<size_t>(nd) >= c1.size()) continue;
            if (c1[nd] >= 0) st.push_back({c1[nd], d + 1});
            if (c2[nd] >= 0) st.push_back({c2[nd], d + 1});
        }
    }
    return best;
}

} // namespace gf
