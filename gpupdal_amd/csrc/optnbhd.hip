// Consumers of whole kNN lists: optimalneighborhood, nndistance and the
// outlier filter's mean distances.
// filters.optimalneighborhood: for each point, the k in [min_k, max_k] whose
// neighborhood has the smallest eigenentropy (filters/OptimalNeighborhoodFilter.cpp:
// 72-163).  Input: every row's max_k neighbors in KD3Index order (the
// point itself first) as sorted positions, from the ordered-list search and
// its exact re-solve; a row's neighbors are near each other in the sorted
// coordinates, so their reads share cache lines.
// One thread per row walks its list once: the running mean and scatter
// matrix are updated neighbor by neighbor in PDAL's order and rounding
// (:94-129), and from min_k on each prefix's covariance B / (n - 1) goes
// through the Eigen-semantics solver; the eigenvalues, clamped at zero and
// normalized by their sum, give -sum(l ln l) (:131-152).  The first k with
// the strictly smallest entropy wins; OptimalRadius is the square root of
// that neighbor's squared distance (:154-161).
#include "gf_eigen.hpp"
#include "gf_internal.hpp"
#include "gf_kernels.hpp"

namespace gf {

namespace {

// The exact re-solve writes point ids: its rows' lists become sorted
// positions through the inverse permutation.
__global__ __launch_bounds__(256) void k_lists_to_sorted(uint32_t* __restrict__ ids, int kmax,
                                                         const uint32_t* __restrict__ flagged,
                                                         uint32_t nflagged,
                                                         const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ inv)
{
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= static_cast<uint64_t>(nflagged) * kmax) return;
    const uint32_t f = static_cast<uint32_t>(t / kmax);
    const uint32_t j = static_cast<uint32_t>(t % kmax);
    const uint64_t row = perm[flagged[f]];
    uint32_t& e = ids[row * kmax + j];
    e = inv[e];
}

__global__ __launch_bounds__(256) void k_optimal_nbhd(const uint32_t* __restrict__ ids,
                                                      uint64_t n, int kmin, int kmax,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      const double* __restrict__ sx,
                                                      const double* __restrict__ sy,
                                                      const double* __restrict__ sz,
                                                      uint64_t* __restrict__ kopt_out,
                                                      double* __restrict__ ropt_out,
                                                      uint32_t* __restrict__ failed)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const uint32_t* nb = ids + i * static_cast<uint64_t>(kmax);
    const double qx = x[i], qy = y[i], qz = z[i];
    double mx = 0.0, my = 0.0, mz = 0.0;
    double b00 = 0.0, b01 = 0.0, b02 = 0.0, b11 = 0.0, b12 = 0.0, b22 = 0.0;
    double minentropy = 1.7976931348623157e308;  // DBL_MAX
    uint64_t kopt = 0;
    double ropt = 0.0;
    for (int k = 0; k < kmax; ++k) {
        const uint32_t j = nb[k];
        const double px = sx[j], py = sy[j], pz = sz[j];
        const double dx = px - mx, dy = py - my, dz = pz - mz;
        const double cnt = static_cast<double>(k + 1);
        mx += dx / cnt;
        my += dy / cnt;
        mz += dz / cnt;
        const double s = (cnt - 1.0) / cnt;
        b00 = b00 + s * dx * dx;
        b11 = b11 + s * dy * dy;
        b22 = b22 + s * dz * dz;
        b01 = b01 + s * dx * dy;
        b02 = b02 + s * dx * dz;
        b12 = b12 + s * dy * dz;
        if (k < kmin - 1) continue;
        const double den = cnt - 1.0;
        const double c[6] = {b00 / den, b01 / den, b02 / den, b11 / den, b12 / den, b22 / den};
        double ev[3], vec[9];
        if (eigen_sym3(c, ev, vec) != 0) {
            atomicAdd(failed, 1u);
            return;
        }
        double l0 = fmax(ev[2], 0.0), l1 = fmax(ev[1], 0.0), l2 = fmax(ev[0], 0.0);
        const double sum = ((0.0 + l0) + l1) + l2;
        l0 /= sum;
        l1 /= sum;
        l2 /= sum;
        const double entropy = -((l2 * log(l2) + l1 * log(l1)) + l0 * log(l0));
        if (entropy < minentropy) {
            minentropy = entropy;
            kopt = static_cast<uint64_t>(k + 1);
            // nanoflann's squared distance: query minus point, per axis.
            const double ex = qx - px, ey = qy - py, ez = qz - pz;
            ropt = ex * ex + ey * ey + ez * ez;
        }
    }
    kopt_out[i] = kopt;
    ropt_out[i] = sqrt(ropt);
}

// filters.nndistance (filters/NNDistanceFilter.cpp:113-132) and the
// per-point mean distance of filters.outlier's statistical method
// (filters/OutlierFilter.cpp:122-136), from each row's list of the K = k + 1
// nearest points as sorted positions.  KD3Index::knnSearch returns the
// squared distances in ascending order, which do not depend on how ties
// between equal distances were ordered; with fewer than K points in the cloud
// its tail entries stay 0 (:119-120, KDImpl.hpp:386).
__global__ __launch_bounds__(256) void k_nn_distance(const uint32_t* __restrict__ ids, int kk,
                                                     int K, uint64_t n, int mode,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ z,
                                                     const double* __restrict__ sx,
                                                     const double* __restrict__ sy,
                                                     const double* __restrict__ sz,
                                                     double* __restrict__ out)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const uint32_t* nb = ids + i * static_cast<uint64_t>(kk);
    const double qx = x[i], qy = y[i], qz = z[i];
    // nanoflann's squared distance: query minus point, per axis, summed in
    // axis order.
    auto dist = [&](int j) -> double {
        if (j >= kk) return 0.0;
        const uint32_t p = nb[j];
        const double ex = qx - sx[p], ey = qy - sy[p], ez = qz - sz[p];
        return std::sqrt(ex * ex + ey * ey + ez * ez);
    };
    double val = 0.0;
    if (mode == kNndKth) {
        val = dist(K - 1);
    } else if (mode == kNndAvg) {
        for (int j = 1; j < K; ++j) val += dist(j);
        val /= static_cast<double>(K - 1);
    } else {
        // Running mean over the neighbors in ascending order.
        for (int j = 1; j < K; ++j) {
            const double delta = dist(j) - val;
            val += delta / static_cast<double>(j);
        }
    }
    out[i] = val;
}

// Stages over a per-row neighbor count (covariancefeatures' optimized=true:
// KD3Index::neighbors(p, OptimalKNN, 1), filters/CovarianceFeaturesFilter.cpp
// :204-210): row i's neighborhood is the first row_k[i] entries (at most the
// cloud's size) of its KD3Index-ordered list, whose prefixes are exactly the
// smaller queries' lists.  The covariance is formed as the fused search's
// (relative to the query point, mean first, divided by k - 1), then the same
// eigen solve and stage formulas as the feature kernel.
__global__ __launch_bounds__(256) void k_prefix_stages(const uint32_t* __restrict__ ids, int kk,
                                                       const uint32_t* __restrict__ row_k,
                                                       uint64_t n, const double* __restrict__ x,
                                                       const double* __restrict__ y,
                                                       const double* __restrict__ z,
                                                       const double* __restrict__ sx,
                                                       const double* __restrict__ sy,
                                                       const double* __restrict__ sz,
                                                       FusedProgram prog, KnnOutputs out)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int k = static_cast<int>(min(row_k[i], static_cast<uint32_t>(kk)));
    const uint32_t* nb = ids + i * static_cast<uint64_t>(kk);
    const double ox = x[i], oy = y[i], oz = z[i];
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (int l = 0; l < k; ++l) {
        const uint32_t j = nb[l];
        mx += sx[j] - ox;
        my += sy[j] - oy;
        mz += sz[j] - oz;
    }
    mx /= static_cast<double>(k);
    my /= static_cast<double>(k);
    mz /= static_cast<double>(k);
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (int l = 0; l < k; ++l) {
        const uint32_t j = nb[l];
        const double dx = (sx[j] - ox) - mx, dy = (sy[j] - oy) - my, dz = (sz[j] - oz) - mz;
        c[0] += dx * dx;
        c[1] += dx * dy;
        c[2] += dx * dz;
        c[3] += dy * dy;
        c[4] += dy * dz;
        c[5] += dz * dz;
    }
    const double den = static_cast<double>(k - 1);
    for (int t = 0; t < 6; ++t) c[t] /= den;
    const bool zero = cov_is_zero(c);
    double ev[3] = {0, 0, 0}, vec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int info = 0;
    bool rank_stage = false;
    for (int si = 0; si < prog.nstages; ++si) rank_stage |= prog.st[si].stage == STAGE_RANK;
    if (!zero || rank_stage) info = eigen_sym3(c, ev, vec);
    for (int si = 0; si < prog.nstages; ++si) {
        const StageProgram& sp = prog.st[si];
        double vals[10];
        const uint8_t st = stage_outputs(sp, k, c, zero, info, ev, vec, vals);
        if (st == 0)
            for (int cc = 0; cc < sp.ncols; ++cc) out.col[sp.col_offset + cc][i] = vals[cc];
        out.status[si][i] = st;
    }
}

__global__ __launch_bounds__(256) void k_max_u32(const uint32_t* __restrict__ v, uint64_t n,
                                                 uint32_t* __restrict__ out)
{
    uint32_t m = 0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        m = max(m, v[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(m, o)));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

} // namespace

void launch_prefix_stages(const uint32_t* ids, int kk, const uint32_t* row_k, uint64_t n,
                          const double* x, const double* y, const double* z, const double* sx,
                          const double* sy, const double* sz, const FusedProgram& prog,
                          const KnnOutputs& out, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_prefix_stages, dim3(static_cast<uint32_t>(div_up(n, 256))), dim3(256),
                       0, s, ids, kk, row_k, n, x, y, z, sx, sy, sz, prog, out);
}

void launch_max_u32(const uint32_t* v, uint64_t n, uint32_t* out, hipStream_t s)
{
    launch_fill_u32(out, 1, 0u, s);
    if (!n) return;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(div_up(n, 256), 1024));
    hipLaunchKernelGGL(k_max_u32, dim3(blocks), dim3(256), 0, s, v, n, out);
}

void launch_nn_distance(const uint32_t* ids, int kk, int K, uint64_t n, int mode,
                        const double* x, const double* y, const double* z, const double* sx,
                        const double* sy, const double* sz, double* out, hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_nn_distance, dim3(static_cast<uint32_t>(div_up(n, 256))), dim3(256), 0,
                       s, ids, kk, K, n, mode, x, y, z, sx, sy, sz, out);
}

void launch_lists_to_sorted(uint32_t* ids, int kmax, const uint32_t* flagged, uint32_t nflagged,
                            const uint32_t* perm, const uint32_t* inv, hipStream_t s)
{
    if (!nflagged) return;
    const uint64_t total = static_cast<uint64_t>(nflagged) * static_cast<uint64_t>(kmax);
    hipLaunchKernelGGL(k_lists_to_sorted, dim3(static_cast<uint32_t>(div_up(total, 256))),
                       dim3(256), 0, s, ids, kmax, flagged, nflagged, perm, inv);
}

void launch_optimal_nbhd(const uint32_t* ids, uint64_t n, int kmin, int kmax, const double* x,
                         const double* y, const double* z, const double* sx, const double* sy,
                         const double* sz, uint64_t* kopt, double* ropt, uint32_t* failed,
                         hipStream_t s)
{
    if (!n) return;
    hipLaunchKernelGGL(k_optimal_nbhd, dim3(static_cast<uint32_t>(div_up(n, 256))), dim3(256),
                       0, s, ids, n, kmin, kmax, x, y, z, sx, sy, sz, kopt, ropt, failed);
}

} // namespace gf
