// gf_internal.hpp -- shared definitions of the HIP neighborhood-feature
// library (device structs, launch helpers, error handling).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

namespace gf {

// Row status bits.  The meanings of 1/2/4/8/16 follow the reference's
// per-row contract (include/pdg/index/SpatialIndex.hpp:121-131); 32 is
// filters.normal's "fewer than three neighbors" skip (MathUtils.cpp:576-580).
enum : uint8_t {
    ST_TIE = 1,           // boundary distance tie: resolved by the KD3 search
    ST_INCOMPLETE = 2,    // grid search not certified: resolved by KD3 search
    ST_ZERO_COV = 4,      // Eigen isZero covariance: PDAL skips the row
    ST_EIGEN_FAIL = 8,    // solver NoConvergence
    ST_FEATURE_INVALID = 16, // covfeat: largest clamped eigenvalue is 0
    ST_FEW_NEIGHBORS = 32, // normal: fewer than 3 neighbors
    // Internal, in a covariance record's flag word only: the row goes to the
    // exact search without a grid search (a tile or cloud with fewer points
    // than the result capacity).  Never reported.
    ST_EXACT_ONLY = 128
};
constexpr uint8_t kReportedSearchFlags = ST_TIE | ST_INCOMPLETE;

enum : int32_t {
    STAGE_COVFEAT = 0, STAGE_NORMAL = 1, STAGE_EIGENVALUES = 2, STAGE_COPLANAR = 3, STAGE_RANK = 4
};
enum : int32_t { MODE_RAW = 0, MODE_SQRT = 1, MODE_NORMALIZED = 2 };

constexpr int kWave = 64;
constexpr int kMaxStages = 4;
constexpr int kMaxFetch = 64;     // K*stride held by the wave-wide top list
// Output columns of one fused run: four stages write at most 10 + 10 + 4 + 3.
constexpr int kMaxOutColumns = 32;
// Per-row distance statistics of a kNN list: the public GF_NND_KTH and
// GF_NND_AVG, and the outlier filter's running mean.
constexpr int kNndKth = 0;
constexpr int kNndAvg = 1;
constexpr int kNndRunningMean = 2;

// Uniform column grid over one cloud (or one tile of a batch).  Columns are
// (cx, cy) cells of edge `c`, stored in 8x8 tiles (row-major tiles, Morton
// order inside a tile); points are sorted by key = (column << bz) | cz.
struct GridParams {
    double ox, oy, oz;  // origin: the cloud's minimum corner
    double c, cz;       // column edge (x, y) and z cell height
    double inv_c, inv_cz;
    uint32_t gx, gy, gz;
    uint32_t tiles_x;
    uint32_t bz;        // bits of cz in the key
    uint32_t key_bits;  // total significant key bits
    uint32_t ncols;     // padded column count (tiles * 64)
    uint32_t pad;
};

// Per-stage output program (one fused pass can serve several stages that
// share the same neighborhood size).
struct StageProgram {
    int32_t stage;
    int32_t mode;
    uint32_t features;
    int32_t always_up;
    int32_t normalize;
    int32_t ncols;       // output columns of this stage
    int32_t col_offset;  // first output column
    int32_t pad;
    double p1, p2;       // approximatecoplanar thresh1/thresh2; estimaterank thresh
};

struct FusedProgram {
    int32_t nstages;
    int32_t k;          // neighbors used (min(n, knn+1))
    int32_t stride;     // fetch stride*k, keep every stride-th
    int32_t fetch;      // stride*k
    StageProgram st[kMaxStages];
    // Radius neighborhoods (radius > 0): every point within `radius`; each
    // row's neighbor count is in its covariance record, and covfeat /
    // eigenvalues skip rows with fewer than min_k.
    double radius;
    int32_t min_k;
    int32_t pad;
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define GF_HIP(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess)                                                 \
            throw ::gf::Error(-2, std::string(#expr) + ": " +                 \
                                      hipGetErrorString(e_));                 \
    } while (0)

inline uint32_t div_up(uint64_t a, uint64_t b) {
    return static_cast<uint32_t>((a + b - 1) / b);
}

// Count of bits needed to represent values in [0, v].
__host__ __device__ inline uint32_t bits_for(uint64_t v) {
    uint32_t b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

// Points per unit of xy extent of bounds b (min xyz, max xyz) holding n
// points, each extent held at least at 1e-9 of the largest (and 1e-9).
__host__ __device__ inline double grid_rho(const double b[6], uint64_t n) {
    const double sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
    const double big = fmax(fmax(sx, sy), fmax(sz, 0.0));
    const double tiny = fmax(big * 1e-9, 1e-9);
    return static_cast<double>(n > 1 ? n : 1) / (fmax(sx, tiny) * fmax(sy, tiny));
}

// Autotune buckets: floor(log2 rho) and ceil(log2 capacity), exactly (frexp
// and integer bits, the same on host and device).
__host__ __device__ inline int tune_rho_bucket(double rho) {
    int e = 0;
    frexp(rho > 1e-300 ? rho : 1e-300, &e);
    return e - 1;
}
__host__ __device__ inline int tune_k_bucket(uint32_t fetch) {
    return fetch <= 1 ? 0 : static_cast<int>(bits_for(static_cast<uint64_t>(fetch) - 1));
}

// The grid of a cloud (or tile) with bounds b and n points for result
// capacity `fetch`: column edge factor * sqrt(fetch / (pi * rho)) (or
// `cell` when > 0), grown until the padded column count fits 4n (and 2^30),
// z cells of cell_z_factor * edge, the z cells' bits cut so that a key fits
// 32 bits.  The same function on host and device, so both build the same
// grid.
__host__ __device__ inline GridParams grid_params(const double b[6], uint64_t n, uint32_t fetch,
                                                  double factor, double cell_z_factor,
                                                  double cell) {
    GridParams g{};
    g.ox = b[0];
    g.oy = b[1];
    g.oz = b[2];
    const double sx = b[3] - b[0], sy = b[4] - b[1], sz = b[5] - b[2];
    const double big = fmax(fmax(sx, sy), fmax(sz, 0.0));
    const double tiny = fmax(big * 1e-9, 1e-9);
    const double rho = grid_rho(b, n);
    double c = cell > 0.0 ? cell
                          : factor * sqrt(static_cast<double>(fetch > 1 ? fetch : 1) /
                                          (3.141592653589793 * rho));
    if (!(c > tiny)) c = tiny;
    const uint64_t col_cap = 4 * n > 4096 ? 4 * n : 4096;
    for (;;) {
        g.c = c;
        g.inv_c = 1.0 / c;
        // Column counts beyond 2^30 per axis are refused before the cast (the
        // column cap below refuses them anyway).
        const double fx = floor(sx * g.inv_c) + 1.0, fy = floor(sy * g.inv_c) + 1.0;
        if (!(fx <= 1073741824.0 && fy <= 1073741824.0)) {
            c *= 2.0;
            continue;
        }
        g.gx = static_cast<uint32_t>(fx);
        g.gy = static_cast<uint32_t>(fy);
        g.tiles_x = (g.gx + 7) / 8;
        const uint64_t tiles_y = (g.gy + 7) / 8;
        const uint64_t ncols = static_cast<uint64_t>(g.tiles_x) * tiles_y * 64;
        if (ncols <= col_cap && ncols < (1ull << 30)) {
            g.ncols = static_cast<uint32_t>(ncols);
            break;
        }
        c *= 1.25;
    }
    const uint32_t bcol = bits_for(g.ncols - 1);
    double cz = c * cell_z_factor;
    g.cz = cz;
    g.inv_cz = 1.0 / cz;
    uint64_t gz = static_cast<uint64_t>(floor(sz * g.inv_cz)) + 1;
    uint32_t bz = bits_for(gz - 1);
    if (bcol + bz > 32) {
        bz = 32 - bcol;
        const uint64_t gzmax = 1ull << bz;
        cz = sz / static_cast<double>(gzmax - 1) * (1.0 + 1e-9);
        for (;;) {
            g.cz = cz;
            g.inv_cz = 1.0 / cz;
            gz = static_cast<uint64_t>(floor(sz * g.inv_cz)) + 1;
            if (gz <= gzmax) break;
            cz *= 1.0 + 1e-6;
        }
    }
    g.gz = static_cast<uint32_t>(gz);
    g.bz = bz;
    g.key_bits = bcol + bz;
    return g;
}

// Column entries a grid of a cloud (or tile) of n points may hold: the
// padded columns are at most max(4n, 4096); plus the end entry.
__host__ __device__ inline uint64_t grid_col_cap(uint64_t n) {
    return (4 * n > 4096 ? 4 * n : 4096) + 1;
}

// ---------------------------------------------------------------- device --
__device__ __forceinline__ uint32_t morton_tile6(uint32_t x, uint32_t y) {
    // interleave 3 bits of x and y: x0 y0 x1 y1 x2 y2 (x in even bits)
    uint32_t r = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        r |= ((x >> b) & 1u) << (2 * b);
        r |= ((y >> b) & 1u) << (2 * b + 1);
    }
    return r;
}

__host__ __device__ __forceinline__ uint32_t column_index(const GridParams& g,
                                                          uint32_t cx,
                                                          uint32_t cy) {
    uint32_t tile = (cy >> 3) * g.tiles_x + (cx >> 3);
    uint32_t in = 0;
    for (int b = 0; b < 3; ++b) {
        in |= ((cx >> b) & 1u) << (2 * b);
        in |= ((cy >> b) & 1u) << (2 * b + 1);
    }
    return tile * 64u + in;
}

// Cell coordinate along one axis.  Points and queries use this same
// function, so a query always finds itself in its own cell.
__host__ __device__ __forceinline__ uint32_t cell_coord(double v, double o,
                                                        double inv,
                                                        uint32_t dim) {
    double f = floor((v - o) * inv);
    if (!(f >= 0.0)) return 0u;  // also catches NaN
    if (f >= static_cast<double>(dim - 1)) return dim - 1;
    return static_cast<uint32_t>(f);
}

} // namespace gf
