// grid.hip -- uniform column grid over one cloud: bounds, cell keys, an
// LDS-staged LSD radix sort of (key, point id), column start offsets and the
// sorted FP64 coordinate copy.
//
// Replaces the reference's buildUniformGridDevice (generateCellsKernel, CUB
// SortPairs / RunLengthEncode / ExclusiveSum, gatherSortedCoordinatesKernel;
// src/index/SpatialIndexKernels.cu:89-108, 301-325, 1772-1906).  The layout
// differs: a dense 2-D column table (8x8-tiled, Morton inside a tile) with the
// z cell in the low key bits, so a query finds a column with one load and a
// z window with a short binary search, and tall z extents cost key bits only.
#include "gf_internal.hpp"
#include "gf_kernels.hpp"

#include <algorithm>

namespace gf {

// ------------------------------------------------------------- bounds ----
constexpr int kBoundsThreads = 256;

__global__ __launch_bounds__(kBoundsThreads) void k_bounds_partial(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, uint64_t n, double* __restrict__ part)
{
    __shared__ double s[6][kBoundsThreads];
    double mn[3] = {INFINITY, INFINITY, INFINITY};
    double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    const double* a[3] = {x, y, z};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double v = a[d][i];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    }
    for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = mn[d];
        s[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int w = kBoundsThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] =
                    fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(kBoundsThreads) void k_bounds_final(
    const double* __restrict__ part, uint32_t nparts, double* __restrict__ out)
{
    __shared__ double s[6][kBoundsThreads];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < nparts; b += kBoundsThreads)
        for (int d = 0; d < 6; ++d)
            v[d] = d < 3 ? fmin(v[d], part[b * 6 + d]) : fmax(v[d], part[b * 6 + d]);
    for (int d = 0; d < 6; ++d) s[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int w = kBoundsThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int d = 0; d < 6; ++d)
                s[d][threadIdx.x] = d < 3 ? fmin(s[d][threadIdx.x], s[d][threadIdx.x + w])
                                          : fmax(s[d][threadIdx.x], s[d][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

void launch_bounds(const double* x, const double* y, const double* z,
                   uint64_t n, double* part, double* out, hipStream_t s)
{
    uint32_t blocks = n ? div_up(n, kBoundsThreads * 8) : 1;
    if (blocks > kBoundsMaxParts) blocks = kBoundsMaxParts;
    hipLaunchKernelGGL(k_bounds_partial, dim3(blocks), dim3(kBoundsThreads), 0, s,
                       x, y, z, n, part);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBoundsThreads), 0, s, part, blocks,
                       out);
}

// ---------------------------------------------------------------- keys ----
__global__ __launch_bounds__(256) void k_keys(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, uint64_t n, GridParams g,
    uint32_t* __restrict__ keys, uint32_t* __restrict__ ids, double* __restrict__ xyz)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    uint32_t cx = cell_coord(px, g.ox, g.inv_c, g.gx);
    uint32_t cy = cell_coord(py, g.oy, g.inv_c, g.gy);
    uint32_t cz = cell_coord(pz, g.oz, g.inv_cz, g.gz);
    uint32_t col = column_index(g, cx, cy);
    keys[i] = (col << g.bz) | cz;
    ids[i] = static_cast<uint32_t>(i);
    xyz[3 * i] = px;
    xyz[3 * i + 1] = py;
    xyz[3 * i + 2] = pz;
}

void launch_keys(const double* x, const double* y, const double* z, uint64_t n,
                 const GridParams& g, uint32_t* keys, uint32_t* ids, double* xyz,
                 hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_keys, dim3(div_up(n, 256)), dim3(256), 0, s, x, y, z, n,
                       g, keys, ids, xyz);
}

// colstart[c] = first sorted position whose column is >= c (c in [0, ncols]),
// by binary search over the sorted keys: no atomics, every column written once.
__global__ __launch_bounds__(256) void k_colstart(const uint32_t* __restrict__ keys,
                                                  uint64_t n, uint32_t ncols, uint32_t bz,
                                                  uint32_t* __restrict__ colstart)
{
    uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (c > ncols) return;
    if (c == ncols) {
        colstart[c] = static_cast<uint32_t>(n);
        return;
    }
    const uint64_t key = static_cast<uint64_t>(c) << bz;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (static_cast<uint64_t>(keys[mid]) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    colstart[c] = static_cast<uint32_t>(lo);
}

void launch_colstart(const uint32_t* keys, uint64_t n, const GridParams& g,
                     uint32_t* colstart, hipStream_t s)
{
    hipLaunchKernelGGL(k_colstart, dim3(div_up(static_cast<uint64_t>(g.ncols) + 1, 256)),
                       dim3(256), 0, s, keys, n, g.ncols, g.bz, colstart);
}

// --------------------------------------------------------------- scan -----
// Exclusive scan of uint32 in three launches: per-tile sums, a single-block
// scan of the tile sums, and a per-tile scan seeded with its offset.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread over a 256-thread block; returns
// the exclusive prefix and writes the block total to *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t s_w[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan(v);
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        uint32_t t = s_w[i];
        if (i < w) off += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(
    const uint32_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ sums)
{
    uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t acc = 0;
    for (int i = 0; i < kScanItems; ++i) {
        uint64_t p = base + (uint64_t)i * kScanThreads + threadIdx.x;
        if (p < n) acc += in[p];
    }
    uint32_t tot;
    block_excl_scan(acc, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_top(uint32_t* __restrict__ sums,
                                                           uint32_t nb)
{
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kScanThreads) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(
    const uint32_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ sums,
    uint32_t* __restrict__ out)
{
    __shared__ uint32_t s[kScanTile];
    uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    for (int i = 0; i < kScanItems; ++i) {
        uint32_t l = i * kScanThreads + threadIdx.x;
        uint64_t p = base + l;
        s[l] = p < n ? in[p] : 0u;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t acc = 0;
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = s[threadIdx.x * kScanItems + i];
        acc += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(acc, &tot) + sums[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        s[threadIdx.x * kScanItems + i] = ex;
        ex += v[i];
    }
    __syncthreads();
    for (int i = 0; i < kScanItems; ++i) {
        uint32_t l = i * kScanThreads + threadIdx.x;
        uint64_t p = base + l;
        if (p < n) out[p] = s[l];
    }
}

uint64_t scan_scratch_words(uint64_t n) { return div_up(n ? n : 1, kScanTile) + 1; }

void launch_scan_top(uint32_t* sums, uint32_t nb, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, sums, nb);
}

void launch_excl_scan(const uint32_t* in, uint32_t* out, uint64_t n,
                      uint32_t* scratch, hipStream_t s)
{
    if (!n) return;
    uint32_t nb = div_up(n, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, s, in, n, scratch);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, scratch, nb);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanThreads), 0, s, in, n, scratch,
                       out);
}

// ---------------------------------------------------------- radix sort -----
// One 8-bit LSD pass = histogram per 4096-key tile, an exclusive scan of the
// digit-major (digit, tile) histogram, and a stable scatter.  The scatter
// ranks keys inside a tile with wave ballots (lanes sharing a digit), stages
// the tile in LDS in digit order and writes runs of equal digits coalesced.
constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;
constexpr int kRsBins = 256;
constexpr int kRsWaves = kRsThreads / 64;

__global__ __launch_bounds__(kRsThreads) void k_radix_up(
    const uint32_t* __restrict__ kin, uint64_t n, uint32_t shift, uint32_t nb,
    uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[kRsBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * kRsTile;
    for (int i = 0; i < kRsItems; ++i) {
        uint64_t p = base + (uint64_t)i * kRsThreads + threadIdx.x;
        if (p < n) atomicAdd(&h[(kin[p] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kRsThreads) void k_radix_down(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
    const uint32_t* __restrict__ goff, uint64_t n, uint32_t shift, uint32_t nb)
{
    __shared__ uint32_t s_cnt[kRsWaves][kRsBins];
    __shared__ uint32_t s_start[kRsBins];
    __shared__ uint32_t s_goff[kRsBins];
    __shared__ uint32_t s_key[kRsTile];
    __shared__ uint32_t s_val[kRsTile];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
    for (int i = threadIdx.x; i < kRsWaves * kRsBins; i += kRsThreads)
        (&s_cnt[0][0])[i] = 0;
    s_goff[threadIdx.x] = goff[(uint64_t)threadIdx.x * nb + blockIdx.x];
    __syncthreads();

    uint32_t key[kRsItems], val[kRsItems], rank[kRsItems];
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int it = 0; it < kRsItems; ++it) {
        uint64_t pos = base + (uint64_t)(w * kRsItems + it) * 64 + lane;
        bool valid = pos < n;
        key[it] = valid ? kin[pos] : 0u;
        val[it] = valid ? vin[pos] : 0u;
        uint32_t d = (key[it] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t r = __popcll(m & lt);
        uint32_t cnt = __popcll(m);
        uint32_t prior = valid ? s_cnt[w][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && r == cnt - 1) s_cnt[w][d] = prior + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[it] = prior + r;
    }
    __syncthreads();
    {
        const int t = threadIdx.x;  // one digit per thread
        uint32_t c[kRsWaves], tot = 0;
        for (int i = 0; i < kRsWaves; ++i) {
            c[i] = s_cnt[i][t];
            s_cnt[i][t] = tot;
            tot += c[i];
        }
        uint32_t all;
        s_start[t] = block_excl_scan(tot, &all);
    }
    __syncthreads();
    for (int it = 0; it < kRsItems; ++it) {
        uint64_t pos = base + (uint64_t)(w * kRsItems + it) * 64 + lane;
        if (pos < n) {
            uint32_t d = (key[it] >> shift) & 255u;
            uint32_t p = s_start[d] + s_cnt[w][d] + rank[it];
            s_key[p] = key[it];
            s_val[p] = val[it];
        }
    }
    __syncthreads();
    uint64_t cnt = n - base < (uint64_t)kRsTile ? n - base : (uint64_t)kRsTile;
    for (uint32_t p = threadIdx.x; p < cnt; p += kRsThreads) {
        uint32_t k = s_key[p];
        uint32_t d = (k >> shift) & 255u;
        uint32_t g = s_goff[d] + (p - s_start[d]);
        kout[g] = k;
        vout[g] = s_val[p];
    }
}

uint64_t radix_hist_words(uint64_t n) { return (uint64_t)kRsBins * div_up(n ? n : 1, kRsTile); }

// Sorts (keys, vals) by the low `bits` key bits.  Buffers ping-pong; the
// result pointers are returned through *kres / *vres.
void launch_radix_sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                       uint64_t n, uint32_t bits, uint32_t* hist,
                       uint32_t* scan_scratch, uint32_t** kres, uint32_t** vres,
                       hipStream_t s)
{
    uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
    if (n > 1) {
        uint32_t nb = div_up(n, kRsTile);
        uint64_t hw = (uint64_t)kRsBins * nb;
        for (uint32_t shift = 0; shift < bits; shift += 8) {
            hipLaunchKernelGGL(k_radix_up, dim3(nb), dim3(kRsThreads), 0, s, ka, n,
                               shift, nb, hist);
            launch_excl_scan(hist, hist, hw, scan_scratch, s);
            hipLaunchKernelGGL(k_radix_down, dim3(nb), dim3(kRsThreads), 0, s, ka, va,
                               kb, vb, hist, n, shift, nb);
            uint32_t* t = ka; ka = kb; kb = t;
            t = va; va = vb; vb = t;
        }
    }
    *kres = ka;
    *vres = va;
}

// ------------------------------------------------------ tiled batches -----
// A batch holds independent tiles back to back.  Work is cut into segment
// blocks of at most 4096 points that never cross a tile, so every per-block
// histogram belongs to one tile; laid out tile-major (then digit, then block
// within the tile), one exclusive scan of all histograms gives each tile's
// scatter offsets inside its own range.

__device__ __forceinline__ unsigned long long ord_enc(double d)
{
    unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(d));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Per segment block: min/max of x, y, z, then one atomic per value into the
// tile's order-encoded bounds (6 per tile: min xyz, max xyz).
__global__ __launch_bounds__(kRsThreads) void k_tile_bounds(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const SegBlock* __restrict__ blocks,
    unsigned long long* __restrict__ benc)
{
    const SegBlock b = blocks[blockIdx.x];
    unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
    for (uint32_t i = threadIdx.x; i < b.count; i += kRsThreads) {
        const uint64_t p = static_cast<uint64_t>(b.begin) + i;
        const unsigned long long v[3] = {ord_enc(x[p]), ord_enc(y[p]), ord_enc(z[p])};
        for (int d = 0; d < 3; ++d) {
            lo[d] = v[d] < lo[d] ? v[d] : lo[d];
            hi[d] = v[d] > hi[d] ? v[d] : hi[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) {
            const unsigned long long ol = __shfl_xor(lo[d], m);
            const unsigned long long oh = __shfl_xor(hi[d], m);
            lo[d] = ol < lo[d] ? ol : lo[d];
            hi[d] = oh > hi[d] ? oh : hi[d];
        }
    }
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 3; ++d) {
            atomicMin(&benc[6 * static_cast<uint64_t>(b.tile) + d], lo[d]);
            atomicMax(&benc[6 * static_cast<uint64_t>(b.tile) + 3 + d], hi[d]);
        }
}

__global__ __launch_bounds__(kRsThreads) void k_tile_keys(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const SegBlock* __restrict__ blocks,
    const TileGrid* __restrict__ grids, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ ids, double* __restrict__ xyz)
{
    const SegBlock b = blocks[blockIdx.x];
    const GridParams g = grids[b.tile].g;
    for (uint32_t i = threadIdx.x; i < b.count; i += kRsThreads) {
        const uint64_t p = static_cast<uint64_t>(b.begin) + i;
        const double px = x[p], py = y[p], pz = z[p];
        const uint32_t cx = cell_coord(px, g.ox, g.inv_c, g.gx);
        const uint32_t cy = cell_coord(py, g.oy, g.inv_c, g.gy);
        const uint32_t cz = cell_coord(pz, g.oz, g.inv_cz, g.gz);
        keys[p] = (column_index(g, cx, cy) << g.bz) | cz;
        ids[p] = static_cast<uint32_t>(p);
        xyz[3 * p] = px;
        xyz[3 * p + 1] = py;
        xyz[3 * p + 2] = pz;
    }
}

__global__ __launch_bounds__(kRsThreads) void k_seg_radix_up(
    const uint32_t* __restrict__ kin, const SegBlock* __restrict__ blocks, uint32_t shift,
    uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[kRsBins];
    const SegBlock b = blocks[blockIdx.x];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < b.count; i += kRsThreads)
        atomicAdd(&h[(kin[static_cast<uint64_t>(b.begin) + i] >> shift) & 255u], 1u);
    __syncthreads();
    hist[static_cast<uint64_t>(b.hbase) + static_cast<uint64_t>(threadIdx.x) * b.nbt + b.lb] =
        h[threadIdx.x];
}

// The stable scatter of one segment block (the scheme of k_radix_down with
// the block's own range and histogram slots).
__global__ __launch_bounds__(kRsThreads) void k_seg_radix_down(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
    const uint32_t* __restrict__ goff, const SegBlock* __restrict__ blocks, uint32_t shift)
{
    __shared__ uint32_t s_cnt[kRsWaves][kRsBins];
    __shared__ uint32_t s_start[kRsBins];
    __shared__ uint32_t s_goff[kRsBins];
    __shared__ uint32_t s_key[kRsTile];
    __shared__ uint32_t s_val[kRsTile];
    const SegBlock b = blocks[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t base = b.begin;
    const uint64_t n = base + b.count;
    for (int i = threadIdx.x; i < kRsWaves * kRsBins; i += kRsThreads)
        (&s_cnt[0][0])[i] = 0;
    s_goff[threadIdx.x] =
        goff[static_cast<uint64_t>(b.hbase) + static_cast<uint64_t>(threadIdx.x) * b.nbt + b.lb];
    __syncthreads();

    uint32_t key[kRsItems], val[kRsItems], rank[kRsItems];
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int it = 0; it < kRsItems; ++it) {
        uint64_t pos = base + (uint64_t)(w * kRsItems + it) * 64 + lane;
        bool valid = pos < n;
        key[it] = valid ? kin[pos] : 0u;
        val[it] = valid ? vin[pos] : 0u;
        uint32_t d = (key[it] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bb = 0; bb < 8; ++bb) {
            uint64_t bits = __ballot((d >> bb) & 1u);
            m &= ((d >> bb) & 1u) ? bits : ~bits;
        }
        uint32_t r = __popcll(m & lt);
        uint32_t cnt = __popcll(m);
        uint32_t prior = valid ? s_cnt[w][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && r == cnt - 1) s_cnt[w][d] = prior + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[it] = prior + r;
    }
    __syncthreads();
    {
        const int t = threadIdx.x;
        uint32_t c[kRsWaves], tot = 0;
        for (int i = 0; i < kRsWaves; ++i) {
            c[i] = s_cnt[i][t];
            s_cnt[i][t] = tot;
            tot += c[i];
        }
        uint32_t all;
        s_start[t] = block_excl_scan(tot, &all);
    }
    __syncthreads();
    for (int it = 0; it < kRsItems; ++it) {
        uint64_t pos = base + (uint64_t)(w * kRsItems + it) * 64 + lane;
        if (pos < n) {
            uint32_t d = (key[it] >> shift) & 255u;
            uint32_t p = s_start[d] + s_cnt[w][d] + rank[it];
            s_key[p] = key[it];
            s_val[p] = val[it];
        }
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < b.count; p += kRsThreads) {
        uint32_t k = s_key[p];
        uint32_t d = (k >> shift) & 255u;
        uint32_t g = s_goff[d] + (p - s_start[d]);
        kout[g] = k;
        vout[g] = s_val[p];
    }
}

// Column starts of every tile: entry e of tile t (colbase <= e <= colbase +
// ncols) is the first sorted position of the tile whose column is >= e -
// colbase (the tile's end for the last entry).
__global__ __launch_bounds__(256) void k_tile_colstart(
    const uint32_t* __restrict__ keys, const TileGrid* __restrict__ grids, uint32_t T,
    uint64_t entries, uint32_t* __restrict__ colstart)
{
    const uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (e >= entries) return;
    uint32_t lo_t = 0, hi_t = T;  // largest t with colbase <= e
    while (hi_t - lo_t > 1) {
        const uint32_t mid = (lo_t + hi_t) >> 1;
        if (grids[mid].colbase <= e)
            lo_t = mid;
        else
            hi_t = mid;
    }
    const TileGrid& tg = grids[lo_t];
    const uint64_t c = e - tg.colbase;
    if (c >= tg.g.ncols) {
        colstart[e] = static_cast<uint32_t>(tg.start + tg.n);
        return;
    }
    const uint64_t key = c << tg.g.bz;
    uint64_t lo = tg.start, hi = tg.start + tg.n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (static_cast<uint64_t>(keys[mid]) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    colstart[e] = static_cast<uint32_t>(lo);
}

// -------------------------------------------------------------- gather -----
__global__ __launch_bounds__(256) void k_gather_sorted(
    const uint32_t* __restrict__ perm, const double* __restrict__ xyz, uint64_t n,
    double* __restrict__ sx, double* __restrict__ sy, double* __restrict__ sz,
    uint32_t* __restrict__ inv)
{
    uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = perm[j];
    const double* p = xyz + 3 * static_cast<uint64_t>(i);
    sx[j] = p[0];
    sy[j] = p[1];
    sz[j] = p[2];
    inv[i] = static_cast<uint32_t>(j);
}

void launch_gather_sorted(const uint32_t* perm, const double* xyz, uint64_t n, double* sx,
                          double* sy, double* sz, uint32_t* inv, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_gather_sorted, dim3(div_up(n, 256)), dim3(256), 0, s, perm, xyz, n,
                       sx, sy, sz, inv);
}

// Writes colstart[ncols] = n after the exclusive scan of the column counts.
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, p, v);
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// A plain streaming copy: the reference point for HBM byte counters.
__global__ __launch_bounds__(256) void k_copy_f64(const double* __restrict__ src,
                                                  double* __restrict__ dst, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy_f64(const double* src, double* dst, uint64_t n, hipStream_t s)
{
    if (!n) return;
    uint32_t b = div_up(n, 256);
    if (b > 65536) b = 65536;
    hipLaunchKernelGGL(k_copy_f64, dim3(b), dim3(256), 0, s, src, dst, n);
}

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s)
{
    if (!n) return;
    uint32_t b = div_up(n, 256);
    if (b > 65536) b = 65536;
    hipLaunchKernelGGL(k_fill_u32, dim3(b), dim3(256), 0, s, p, n, v);
}


namespace {

__global__ __launch_bounds__(256) void k_bounds_init(unsigned long long* __restrict__ benc,
                                                     uint32_t T)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * T) return;
    benc[i] = (i % 6) < 3 ? ~0ull : 0ull;
}

// The inverse of ord_enc: order-encoded bits back to the double.
__device__ __forceinline__ double ord_dec(unsigned long long u)
{
    const unsigned long long b = (u >> 63) ? (u ^ 0x8000000000000000ull) : ~u;
    return __longlong_as_double(static_cast<long long>(b));
}

__global__ __launch_bounds__(64) void k_tile_grids(const unsigned long long* __restrict__ benc,
                                                   const uint64_t* __restrict__ starts,
                                                   const uint32_t* __restrict__ fetch,
                                                   const uint64_t* __restrict__ colbase,
                                                   uint32_t T, const TuneFactor* __restrict__ table,
                                                   int ntable, double default_factor,
                                                   double cell_z_factor,
                                                   TileGrid* __restrict__ grids,
                                                   double* __restrict__ bounds,
                                                   TileInfo* __restrict__ infos)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const uint64_t nt = starts[t + 1] - starts[t];
    double bd[6];
    for (int i = 0; i < 6; ++i) bd[i] = nt ? ord_dec(benc[6ull * t + i]) : 0.0;
    const int rb = tune_rho_bucket(grid_rho(bd, nt)), kb = tune_k_bucket(fetch[t]);
    double f = default_factor;
    for (int e = 0; e < ntable; ++e)
        if (table[e].rb == rb && table[e].kb == kb) {
            f = table[e].factor;
            break;
        }
    const GridParams g = grid_params(bd, nt, fetch[t], f, cell_z_factor, 0.0);
    TileGrid& tg = grids[t];
    tg.g = g;
    tg.start = starts[t];
    tg.n = nt;
    tg.colbase = colbase[t];
    for (int i = 0; i < 6; ++i) bounds[6ull * t + i] = bd[i];
    if (infos) {
        TileInfo& ti = infos[t];
        ti.g = g;
        ti.colbase = colbase[t];
        ti.margin = 1e-6 * fmin(g.c, g.cz) + 1e-15 * (fabs(g.ox) + fabs(g.oy) + fabs(g.oz)) +
                    1e-300;
    }
}

// Exclusive prefix of the tiles' real column entries (ncols + 1 each) from
// their device grids: one workgroup, T small.
__global__ __launch_bounds__(256) void k_col_prefix(const TileGrid* __restrict__ grids,
                                                    uint32_t T, uint64_t* __restrict__ pre)
{
    __shared__ uint64_t part[256];
    const uint32_t per = (T + 255) / 256;
    const uint32_t lo = threadIdx.x * per, hi = min(T, lo + per);
    uint64_t sum = 0;
    for (uint32_t t = lo; t < hi; ++t) sum += static_cast<uint64_t>(grids[t].g.ncols) + 1;
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
        for (int i = 0; i < 256; ++i) {
            const uint64_t v = part[i];
            part[i] = run;
            run += v;
        }
        pre[T] = run;
    }
    __syncthreads();
    uint64_t run = part[threadIdx.x];
    for (uint32_t t = lo; t < hi; ++t) {
        pre[t] = run;
        run += static_cast<uint64_t>(grids[t].g.ncols) + 1;
    }
}

// Column starts of every tile's real columns (tiles laid out by capacity
// at grids[t].colbase), a fixed grid striding over the real entries:
// entry e of the real-entry order is column e - pre[t] of tile t.
__global__ __launch_bounds__(256) void k_tile_colstart_real(const uint32_t* __restrict__ keys,
                                                            const TileGrid* __restrict__ grids,
                                                            const uint64_t* __restrict__ pre,
                                                            uint32_t T,
                                                            uint32_t* __restrict__ colstart)
{
    const uint64_t total = pre[T];
    for (uint64_t e = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        uint32_t lo_t = 0, hi_t = T;  // largest t with pre[t] <= e
        while (hi_t - lo_t > 1) {
            const uint32_t mid = (lo_t + hi_t) >> 1;
            if (pre[mid] <= e)
                lo_t = mid;
            else
                hi_t = mid;
        }
        const TileGrid& tg = grids[lo_t];
        const uint64_t c = e - pre[lo_t];
        uint32_t v;
        if (c >= tg.g.ncols) {
            v = static_cast<uint32_t>(tg.start + tg.n);
        } else {
            const uint64_t key = c << tg.g.bz;
            uint64_t lo = tg.start, hi = tg.start + tg.n;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (static_cast<uint64_t>(keys[mid]) < key)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            v = static_cast<uint32_t>(lo);
        }
        colstart[tg.colbase + c] = v;
    }
}

} // namespace

void launch_tile_colstart_real(const uint32_t* keys, const TileGrid* grids, uint32_t T,
                               uint64_t* pre, uint32_t* colstart, uint64_t max_entries,
                               hipStream_t s)
{
    if (!T) return;
    hipLaunchKernelGGL(k_col_prefix, dim3(1), dim3(256), 0, s, grids, T, pre);
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(div_up(max_entries, 256), 8192));
    hipLaunchKernelGGL(k_tile_colstart_real, dim3(blocks), dim3(256), 0, s, keys, grids, pre, T,
                       colstart);
}

void launch_bounds_init(unsigned long long* benc, uint32_t T, hipStream_t s)
{
    if (!T) return;
    hipLaunchKernelGGL(k_bounds_init, dim3(div_up(6ull * T, 256)), dim3(256), 0, s, benc, T);
}

void launch_tile_grids(const unsigned long long* benc, const uint64_t* starts,
                       const uint32_t* fetch, const uint64_t* colbase, uint32_t T,
                       const TuneFactor* table, int ntable, double default_factor,
                       double cell_z_factor, TileGrid* grids, double* bounds, TileInfo* infos,
                       hipStream_t s)
{
    if (!T) return;
    hipLaunchKernelGGL(k_tile_grids, dim3(div_up(T, 64)), dim3(64), 0, s, benc, starts, fetch,
                       colbase, T, table, ntable, default_factor, cell_z_factor, grids, bounds,
                       infos);
}

void launch_tile_bounds(const double* x, const double* y, const double* z,
                        const SegBlock* blocks, uint32_t nblocks, unsigned long long* benc,
                        hipStream_t s)
{
    if (!nblocks) return;
    hipLaunchKernelGGL(k_tile_bounds, dim3(nblocks), dim3(kRsThreads), 0, s, x, y, z, blocks,
                       benc);
}

void launch_tile_keys(const double* x, const double* y, const double* z, const SegBlock* blocks,
                      uint32_t nblocks, const TileGrid* grids, uint32_t* keys, uint32_t* ids,
                      double* xyz, hipStream_t s)
{
    if (!nblocks) return;
    hipLaunchKernelGGL(k_tile_keys, dim3(nblocks), dim3(kRsThreads), 0, s, x, y, z, blocks,
                       grids, keys, ids, xyz);
}

void launch_seg_radix_sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                           const SegBlock* blocks, uint32_t nblocks, uint32_t bits,
                           uint32_t* hist, uint32_t* scan_scratch, uint32_t** kres,
                           uint32_t** vres, hipStream_t s)
{
    uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
    if (nblocks) {
        const uint64_t hw = static_cast<uint64_t>(kRsBins) * nblocks;
        for (uint32_t shift = 0; shift < bits; shift += 8) {
            hipLaunchKernelGGL(k_seg_radix_up, dim3(nblocks), dim3(kRsThreads), 0, s, ka, blocks,
                               shift, hist);
            launch_excl_scan(hist, hist, hw, scan_scratch, s);
            hipLaunchKernelGGL(k_seg_radix_down, dim3(nblocks), dim3(kRsThreads), 0, s, ka, va,
                               kb, vb, hist, blocks, shift);
            uint32_t* t = ka; ka = kb; kb = t;
            t = va; va = vb; vb = t;
        }
    }
    *kres = ka;
    *vres = va;
}

void launch_tile_colstart(const uint32_t* keys, const TileGrid* grids, uint32_t T,
                          uint64_t entries, uint32_t* colstart, hipStream_t s)
{
    if (!entries) return;
    hipLaunchKernelGGL(k_tile_colstart, dim3(div_up(entries, 256)), dim3(256), 0, s, keys, grids,
                       T, entries, colstart);
}

} // namespace gf
