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
__global__ void k_kd_root(KdView t, uint32_t n, const double* __restrict__ bounds)
{
    t.nl[0] = 0;
    t.nr[0] = n;
    t.parent[0] = -1;
    t.which[0] = 0;
    for (int i = 0; i < 6; ++i) t.box[i] = bounds[i];
}

__global__ __launch_bounds__(kKdThreads) void k_kd_init(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, uint32_t n, KdView t)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    t.vind[p] = p;
    t.px[p] = x[p];
    t.py[p] = y[p];
    t.pz[p] = z[p];
    t.segid[p] = 0;
}

__global__ __launch_bounds__(kKdThreads) void k_kd_minmax_reset(KdView t, uint32_t lv0,
                                                               uint32_t lv1)
{
    uint32_t nd = lv0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= lv1) return;
    for (int d = 0; d < 3; ++d) {
        t.mn[3 * nd + d] = ~0ull;
        t.mx[3 * nd + d] = 0ull;
    }
}

// Min/max of every level node's points: a segmented reduction inside each
// wave (runs of equal node ids are contiguous), then per run head an atomic:
// into a per-block LDS table when the level has few nodes (the top levels,
// where every wave would otherwise hit the same global address), else
// straight to global memory.
constexpr uint32_t kLdsNodes = 256;

__global__ __launch_bounds__(kKdThreads) void k_kd_minmax(KdView t, uint32_t n, uint32_t lv0,
                                                          uint32_t lv1)
{
    __shared__ unsigned long long s_lo[kLdsNodes * 3], s_hi[kLdsNodes * 3];
    const uint32_t nl = lv1 - lv0;
    const bool use_lds = nl <= kLdsNodes;
    if (use_lds) {
        for (uint32_t i = threadIdx.x; i < nl * 3; i += blockDim.x) {
            s_lo[i] = ~0ull;
            s_hi[i] = 0ull;
        }
        __syncthreads();
    }
    const int lane = __lane_id();
    const double* a[3] = {t.px, t.py, t.pz};
    for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * blockDim.x; base < n;
         base += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint64_t p = base + threadIdx.x;
        const bool act = p < n && t.segid[p] >= lv0;
        const uint32_t nd = act ? t.segid[p] : 0xFFFFFFFFu;
        unsigned long long lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            unsigned long long e = act ? enc_order(a[d][p]) : 0ull;
            lo[d] = act ? e : ~0ull;
            hi[d] = act ? e : 0ull;
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t ond = __shfl_down(nd, off);
            const bool same = lane + off < 64 && ond == nd;
            for (int d = 0; d < 3; ++d) {
                unsigned long long ol = __shfl_down(lo[d], off);
                unsigned long long oh = __shfl_down(hi[d], off);
                if (same) {
                    lo[d] = ol < lo[d] ? ol : lo[d];
                    hi[d] = oh > hi[d] ? oh : hi[d];
                }
            }
        }
        const uint32_t prev = __shfl_up(nd, 1);
        const bool head = act && (lane == 0 || prev != nd);
        if (head) {
            if (use_lds) {
                for (int d = 0; d < 3; ++d) {
                    atomicMin(&s_lo[(nd - lv0) * 3 + d], lo[d]);
                    atomicMax(&s_hi[(nd - lv0) * 3 + d], hi[d]);
                }
            } else {
                for (int d = 0; d < 3; ++d) {
                    atomicMin(&t.mn[3 * nd + d], lo[d]);
                    atomicMax(&t.mx[3 * nd + d], hi[d]);
                }
            }
        }
    }
    if (use_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x) {
            if (s_hi[i * 3] == 0ull) continue;  // no point of node i in this block
            for (int d = 0; d < 3; ++d) {
                atomicMin(&t.mn[3 * (lv0 + i) + d], s_lo[i * 3 + d]);
                atomicMax(&t.mx[3 * (lv0 + i) + d], s_hi[i * 3 + d]);
            }
        }
    }
}

// One thread per level node: the parent's divlow/divhigh from this node's
// extent, then leaf or split parameters (middleSplit_).
__global__ __launch_bounds__(kKdThreads) void k_kd_split(KdView t, uint32_t lv0, uint32_t lv1,
                                                         uint32_t* __restrict__ is_split)
{
    const uint32_t nd = lv0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= lv1) return;
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        mn[d] = dec_order(t.mn[3 * nd + d]);
        mx[d] = dec_order(t.mx[3 * nd + d]);
    }
    const int par = t.parent[nd];
    if (par >= 0) {
        const int f = t.feat[par];
        if (t.which[nd] == 0)
            t.divlow[par] = mx[f];
        else
            t.divhigh[par] = mn[f];
    }
    const uint32_t cnt = t.nr[nd] - t.nl[nd];
    t.c1[nd] = -1;
    t.c2[nd] = -1;
    if (cnt <= kLeafMax) {
        t.feat[nd] = -1;
        is_split[nd - lv0] = 0;
        return;
    }
    const double* b = t.box + 6 * static_cast<size_t>(nd);
    const double eps = 0.00001;
    double max_span = b[3] - b[0];
    for (int i = 1; i < 3; ++i) {
        double span = b[3 + i] - b[i];
        if (span > max_span) max_span = span;
    }
    double max_spread = -1;
    int f = 0;
    for (int i = 0; i < 3; ++i) {
        double span = b[3 + i] - b[i];
        if (span > (1 - eps) * max_span) {
            double spread = mx[i] - mn[i];
            if (spread > max_spread) {
                f = i;
                max_spread = spread;
            }
        }
    }
    const double split = (b[f] + b[3 + f]) / 2;
    double cv = split;
    if (split < mn[f])
        cv = mn[f];
    else if (split > mx[f])
        cv = mx[f];
    t.feat[nd] = f;
    t.cutval[nd] = cv;
    is_split[nd - lv0] = 1;
}

// flag[p] = 1 when p belongs to a splitting node and sits on the "left" side
// of the pass: pass 1 tests coord < cut over the node; pass 2 tests
// coord <= cut over [L, right) of the node (L from pass 1's counts).
__global__ __launch_bounds__(kKdThreads) void k_kd_flags(KdView t, uint32_t n, uint32_t lv0,
                                                         int pass,
                                                         const uint32_t* __restrict__ s1,
                                                         uint32_t* __restrict__ flag)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n) return;
    uint32_t f = 0;
    if (p < n) {
        const uint32_t nd = t.segid[p];
        if (nd >= lv0 && t.feat[nd] >= 0) {
            const double v = axis(t, t.feat[nd])[p];
            const double cv = t.cutval[nd];
            if (pass == 1) {
                f = v < cv ? 1u : 0u;
            } else {
                const uint32_t s0 = t.nl[nd], e = t.nr[nd];
                const uint32_t L = s0 + (s1[e] - s1[s0]);
                f = (p >= L && v <= cv) ? 1u : 0u;
            }
        }
    }
    flag[p] = f;
}

// The pairing of the closed-form Hoare pass: every misplaced right-side
// element publishes its position under its rank counted from the node's
// end; every misplaced left-side element then swaps with the partner of the
// same rank.  [a, e) is the pass's range, split at m = a + #flag in [a, e).
__device__ __forceinline__ void kd_pass_range(const KdView& t, uint32_t nd, int pass,
                                              const uint32_t* s1, const uint32_t* sc,
                                              uint32_t& a, uint32_t& e, uint32_t& m)
{
    const uint32_t s0 = t.nl[nd];
    e = t.nr[nd];
    a = pass == 1 ? s0 : s0 + (s1[e] - s1[s0]);
    m = a + (sc[e] - sc[a]);
}

__global__ __launch_bounds__(kKdThreads) void k_kd_publish(KdView t, uint32_t n, uint32_t lv0,
                                                           int pass,
                                                           const uint32_t* __restrict__ s1,
                                                           const uint32_t* __restrict__ sc,
                                                           const uint32_t* __restrict__ flag,
                                                           uint32_t* __restrict__ tmp)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t nd = t.segid[p];
    if (nd < lv0 || t.feat[nd] < 0 || !flag[p]) return;
    uint32_t a, e, m;
    kd_pass_range(t, nd, pass, s1, sc, a, e, m);
    if (p >= m) tmp[a + (sc[e] - sc[p + 1])] = p;
}

__global__ __launch_bounds__(kKdThreads) void k_kd_swap(KdView t, uint32_t n, uint32_t lv0,
                                                        int pass,
                                                        const uint32_t* __restrict__ s1,
                                                        const uint32_t* __restrict__ sc,
                                                        const uint32_t* __restrict__ flag,
                                                        const uint32_t* __restrict__ tmp)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t nd = t.segid[p];
    if (nd < lv0 || t.feat[nd] < 0 || flag[p]) return;
    uint32_t a, e, m;
    kd_pass_range(t, nd, pass, s1, sc, a, e, m);
    if (p < a || p >= m) return;
    const uint32_t rank = (p - a) - (sc[p] - sc[a]);
    const uint32_t q = tmp[a + rank];
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

// Children of every splitting node: ids from the exclusive scan of the
// level's split flags, segments from the lim1/lim2/count/2 rule, boxes from
// the cut value.
__global__ __launch_bounds__(kKdThreads) void k_kd_children(KdView t, uint32_t lv0, uint32_t lv1,
                                                            const uint32_t* __restrict__ is_split,
                                                            const uint32_t* __restrict__ rank,
                                                            const uint32_t* __restrict__ s1,
                                                            const uint32_t* __restrict__ s2,
                                                            uint32_t* __restrict__ overflow)
{
    const uint32_t nd = lv0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= lv1 || !is_split[nd - lv0]) return;
    const uint32_t s0 = t.nl[nd], e = t.nr[nd], cnt = e - s0;
    const uint32_t L = s0 + (s1[e] - s1[s0]);
    const uint32_t L2 = L + (s2[e] - s2[L]);
    const uint32_t lim1 = L - s0, lim2 = L2 - s0;
    uint32_t idx;
    if (lim1 > cnt / 2)
        idx = lim1;
    else if (lim2 < cnt / 2)
        idx = lim2;
    else
        idx = cnt / 2;
    const uint32_t c1 = lv1 + 2 * rank[nd - lv0];
    const uint32_t c2 = c1 + 1;
    if (c2 >= t.cap) {
        *overflow = 1;
        return;
    }
    t.c1[nd] = static_cast<int32_t>(c1);
    t.c2[nd] = static_cast<int32_t>(c2);
    t.split[nd] = s0 + idx;
    const int f = t.feat[nd];
    const double cv = t.cutval[nd];
    const double* b = t.box + 6 * static_cast<size_t>(nd);
    double* b1 = t.box + 6 * static_cast<size_t>(c1);
    double* b2 = t.box + 6 * static_cast<size_t>(c2);
    for (int i = 0; i < 6; ++i) {
        b1[i] = b[i];
        b2[i] = b[i];
    }
    b1[3 + f] = cv;
    b2[f] = cv;
    t.nl[c1] = s0;
    t.nr[c1] = s0 + idx;
    t.nl[c2] = s0 + idx;
    t.nr[c2] = e;
    t.parent[c1] = static_cast<int32_t>(nd);
    t.parent[c2] = static_cast<int32_t>(nd);
    t.which[c1] = 0;
    t.which[c2] = 1;
}

__global__ __launch_bounds__(kKdThreads) void k_kd_segid(KdView t, uint32_t n, uint32_t lv0)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t nd = t.segid[p];
    if (nd < lv0 || t.c1[nd] < 0) return;
    t.segid[p] = p < t.split[nd] ? static_cast<uint32_t>(t.c1[nd])
                                 : static_cast<uint32_t>(t.c2[nd]);
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

__global__ __launch_bounds__(256) void k_kd_resolve(KdView t, CloudView cv, FusedProgram prog,
                                                    int mode, KnnOutputs out,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z,
                                                    uint32_t* __restrict__ failed)
{
    const int lane = __lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    const uint32_t nflag = *out.nflagged;
    const uint32_t cap = static_cast<uint32_t>(prog.fetch);
    const uint32_t k = static_cast<uint32_t>(prog.k);
    const uint32_t stride = static_cast<uint32_t>(prog.stride);
    for (uint32_t f = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; f < nflag; f += waves) {
        const uint32_t s = out.flagged[f];
        const uint32_t qid = cv.perm[s];
        const double q0 = x[qid], q1 = y[qid], q2 = z[qid];
        const double q[3] = {q0, q1, q2};
        Ent best = ent_inf();
        uint32_t seq = 0;
        double dists[3] = {0.0, 0.0, 0.0};
        double mind = 0.0;
        for (int i = 0; i < 3; ++i) {
            if (q[i] < t.root[i]) {
                dists[i] = (q[i] - t.root[i]) * (q[i] - t.root[i]);
                mind += dists[i];
            }
            if (q[i] > t.root[3 + i]) {
                dists[i] = (q[i] - t.root[3 + i]) * (q[i] - t.root[3 + i]);
                mind += dists[i];
            }
        }
        // Frame registers: lane i holds stack frame i.
        int f_node = 0, f_other = -1, f_idx = 0, f_stage = 0;
        double f_mind = mind, f_cut = 0.0, f_dst = 0.0;
        int sp = 1;
        bool overflow = false;
        while (sp > 0) {
            const int top = sp - 1;
            const int nd = lane_get(f_node, top);
            const int stage = lane_get(f_stage, top);
            if (stage == 0) {
                if (t.c1[nd] < 0) {
                    const Ent wst = ent_shfl(best, static_cast<int>(cap) - 1);
                    const double worst = wst.d == kInfD ? INFINITY : ent_d2(wst);
                    const uint32_t l0 = t.nl[nd], l1 = t.nr[nd];
                    for (uint32_t i0 = l0; i0 < l1; i0 += 64) {
                        const uint32_t i = i0 + lane;
                        Ent c = ent_inf();
                        if (i < l1) {
                            double d = 0.0;
                            const double d0 = q0 - t.px[i];
                            d += d0 * d0;
                            const double d1 = q1 - t.py[i];
                            d += d1 * d1;
                            const double d2 = q2 - t.pz[i];
                            d += d2 * d2;
                            if (d < worst)
                                c = Ent{static_cast<uint64_t>(__double_as_longlong(d)),
                                        seq + (i - l0), t.vind[i]};
                        }
                        if (__ballot(c.d != kInfD)) {
                            bitonic_sort64(c, lane);
                            merge_sorted(best, c, lane);
                        }
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
                const double fm = lane_get(f_mind, top);
                if (lane == top) {
                    f_other = other;
                    f_cut = cut;
                    f_idx = idx;
                    f_stage = 1;
                }
                if (sp >= 64) {
                    overflow = true;
                    break;
                }
                if (lane == sp) {
                    f_node = bestc;
                    f_stage = 0;
                    f_mind = fm;
                }
                ++sp;
            } else if (stage == 1) {
                const int idx = lane_get(f_idx, top);
                const double cut = lane_get(f_cut, top);
                const double dst = dists[idx];
                const double m2 = lane_get(f_mind, top) + cut - dst;
                dists[idx] = cut;
                if (lane == top) {
                    f_dst = dst;
                    f_stage = 2;
                }
                const Ent wst = ent_shfl(best, static_cast<int>(cap) - 1);
                const double worst = wst.d == kInfD ? INFINITY : ent_d2(wst);
                if (m2 * 1.0 <= worst) {
                    if (sp >= 64) {
                        overflow = true;
                        break;
                    }
                    const int other = lane_get(f_other, top);
                    if (lane == sp) {
                        f_node = other;
                        f_stage = 0;
                        f_mind = m2;
                    }
                    ++sp;
                }
            } else {
                dists[lane_get(f_idx, top)] = lane_get(f_dst, top);
                --sp;
            }
        }
        if (overflow) {
            if (lane == 0) atomicAdd(failed, 1u);
            continue;
        }
        // KD3ImplT::neighbors: every stride-th entry of the result set; slots
        // past the points found keep the output's zero initialisation.
        const Ent kept = ent_shfl(best, static_cast<int>((static_cast<uint32_t>(lane) * stride) & 63u));
        const bool mine = static_cast<uint32_t>(lane) < k;
        const uint32_t kept_id = kept.d == kInfD ? 0u : kept.j;
        if (mode == 1) {
            if (mine) out.ids[static_cast<uint64_t>(qid) * k + lane] = kept_id;
            continue;
        }
        const uint32_t id = mine ? kept_id : 0u;
        const double px = mine ? x[id] : 0.0, py = mine ? y[id] : 0.0, pz = mine ? z[id] : 0.0;
        const double inv_k = 1.0 / static_cast<double>(k);
        const double mx = wave_sum(px) * inv_k, my = wave_sum(py) * inv_k,
                     mz = wave_sum(pz) * inv_k;
        const double dx = mine ? px - mx : 0.0, dy = mine ? py - my : 0.0,
                     dz = mine ? pz - mz : 0.0;
        const double den = static_cast<double>(k - 1);
        const double c0 = wave_sum(dx * dx) / den, c1 = wave_sum(dx * dy) / den,
                     c2 = wave_sum(dx * dz) / den, c3 = wave_sum(dy * dy) / den,
                     c4 = wave_sum(dy * dz) / den, c5 = wave_sum(dz * dz) / den;
        if (lane == 0) {
            const uint64_t n = cv.n;
            out.cov[0 * n + s] = c0;
            out.cov[1 * n + s] = c1;
            out.cov[2 * n + s] = c2;
            out.cov[3 * n + s] = c3;
            out.cov[4 * n + s] = c4;
            out.cov[5 * n + s] = c5;
        }
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
    if (mode == 1)
        out.status[0][perm[s]] = 0;
    else
        out.flags[s] = 0;
}

} // namespace

void launch_flag_all(uint64_t n, const uint32_t* perm, int mode, const KnnOutputs& out,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_flag_all, dim3(div_up(std::max<uint64_t>(n, 1), 256)), dim3(256), 0, s,
                       n, perm, mode, out);
}

uint32_t kd_node_capacity(uint64_t n) { return static_cast<uint32_t>(n / 16 + 1024); }

// Builds the tree level by level; returns false if the node capacity was
// too small (the caller grows it and builds again).
bool launch_kd_build(const double* x, const double* y, const double* z, uint64_t n,
                     const double* bounds, KdView t, KdScratch w, uint32_t* nodes,
                     hipStream_t s)
{
    const uint32_t nn = static_cast<uint32_t>(n);
    const uint32_t nb = div_up(n + 1, kKdThreads);
    hipLaunchKernelGGL(k_kd_init, dim3(div_up(n, kKdThreads)), dim3(kKdThreads), 0, s, x, y, z,
                       nn, t);
    hipLaunchKernelGGL(k_kd_root, dim3(1), dim3(1), 0, s, t, nn, bounds);
    launch_fill_u32(w.overflow, 1, 0u, s);
    uint32_t lv0 = 0, lv1 = 1;
    while (lv1 > lv0) {
        const uint32_t nl = lv1 - lv0;
        const uint32_t gnodes = div_up(nl, kKdThreads);
        hipLaunchKernelGGL(k_kd_minmax_reset, dim3(gnodes), dim3(kKdThreads), 0, s, t, lv0, lv1);
        hipLaunchKernelGGL(k_kd_minmax, dim3(nl <= kLdsNodes ? std::min<uint32_t>(nb, 2048) : nb),
                           dim3(kKdThreads), 0, s, t, nn, lv0, lv1);
        launch_fill_u32(w.is_split, nl + 1, 0u, s);
        hipLaunchKernelGGL(k_kd_split, dim3(gnodes), dim3(kKdThreads), 0, s, t, lv0, lv1,
                           w.is_split);
        // pass 1: coord < cut over each splitting node
        hipLaunchKernelGGL(k_kd_flags, dim3(nb), dim3(kKdThreads), 0, s, t, nn, lv0, 1, w.s1,
                           w.flag);
        launch_excl_scan(w.flag, w.s1, n + 1, w.scan_tmp, s);
        hipLaunchKernelGGL(k_kd_publish, dim3(nb), dim3(kKdThreads), 0, s, t, nn, lv0, 1, w.s1,
                           w.s1, w.flag, w.tmp);
        hipLaunchKernelGGL(k_kd_swap, dim3(nb), dim3(kKdThreads), 0, s, t, nn, lv0, 1, w.s1,
                           w.s1, w.flag, w.tmp);
        // pass 2: coord <= cut over [L, right)
        hipLaunchKernelGGL(k_kd_flags, dim3(nb), dim3(kKdThreads), 0, s, t, nn, lv0, 2, w.s1,
                           w.flag);
        launch_excl_scan(w.flag, w.s2, n + 1, w.scan_tmp, s);
        hipLaunchKernelGGL(k_kd_publish, dim3(nb), dim3(kKdThreads), 0, s, t, nn, lv0, 2, w.s1,
                           w.s2, w.flag, w.tmp);
        hipLaunchKernelGGL(k_kd_swap, dim3(nb), dim3(kKdThreads), 0, s, t, nn, lv0, 2, w.s1,
                           w.s2, w.flag, w.tmp);
        // children
        launch_excl_scan(w.is_split, w.rank, nl + 1, w.scan_tmp, s);
        hipLaunchKernelGGL(k_kd_children, dim3(gnodes), dim3(kKdThreads), 0, s, t, lv0, lv1,
                           w.is_split, w.rank, w.s1, w.s2, w.overflow);
        hipLaunchKernelGGL(k_kd_segid, dim3(div_up(n, kKdThreads)), dim3(kKdThreads), 0, s, t,
                           nn, lv0);
        uint32_t h[2];
        GF_HIP(hipMemcpyAsync(&h[0], w.rank + nl, 4, hipMemcpyDeviceToHost, s));
        GF_HIP(hipMemcpyAsync(&h[1], w.overflow, 4, hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        if (h[1]) return false;
        lv0 = lv1;
        lv1 = lv1 + 2 * h[0];
    }
    *nodes = lv1;
    return true;
}

void launch_kd_resolve(const KdView& t, const CloudView& cv, const FusedProgram& prog, int mode,
                       const KnnOutputs& out, const double* x, const double* y, const double* z,
                       uint32_t nflagged, uint32_t* failed, hipStream_t s)
{
    if (!nflagged) return;
    const uint32_t blocks = std::min<uint32_t>(div_up(nflagged, 4), 4096);
    hipLaunchKernelGGL(k_kd_resolve, dim3(blocks), dim3(256), 0, s, t, cv, prog, mode, out, x, y,
                       z, failed);
}

} // namespace gf
