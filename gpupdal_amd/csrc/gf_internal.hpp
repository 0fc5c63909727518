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
inline uint32_t bits_for(uint64_t v) {
    uint32_t b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
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
