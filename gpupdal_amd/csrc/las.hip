// LAS point records on either side of the path.
//
// In: the point records of an uncompressed LAS file image are copied to HBM
// once, as bytes, and X, Y, Z decoded there into the cloud's FP64 arrays as
// int32 * scale + offset (io/LasReader.cpp:926-932 for point formats 0-5,
// 1008-1014 for 6-10; each product rounded before the sum, as PDAL's FP64
// expression without contraction).  Records are of any length (20-38 bytes
// of base record plus extra bytes, so not 4-byte aligned): a workgroup
// copies its run of records into LDS with aligned dword loads, then each
// thread assembles its record's three int32 from LDS bytes.
//
// Out: double columns packed into point records at a byte offset, the extra
// bytes writers.las appends after the base record of each point.
//
// The header rules (field order, point count, format bits, supported
// formats, validation messages) follow io/private/las/Header.cpp:15-144,
// Header.hpp:75-138 and Utils.hpp:93-107.
#include "gf_internal.hpp"
#include "gf_kernels.hpp"

#include <cstring>

namespace gf {

namespace {

constexpr int kLasThreads = 256;
constexpr uint32_t kLasStageBytes = 32768;  // LDS per workgroup for its records

__device__ __forceinline__ int32_t lds_i32(const uint8_t* p)
{
    return static_cast<int32_t>(static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) |
                                (static_cast<uint32_t>(p[2]) << 16) |
                                (static_cast<uint32_t>(p[3]) << 24));
}

// Records [r0, r0 + cnt) of `rec` (record length L): the workgroup's run,
// cnt * L <= kLasStageBytes - 8.  rec is padded by at least 4 bytes past the
// last record.
__global__ __launch_bounds__(kLasThreads) void k_las_xyz(const uint32_t* __restrict__ rec,
                                                         uint64_t n, uint32_t L, uint32_t per,
                                                         LasXform xf, double* __restrict__ x,
                                                         double* __restrict__ y,
                                                         double* __restrict__ z)
{
    extern __shared__ uint32_t s_rec[];
    const uint64_t r0 = static_cast<uint64_t>(blockIdx.x) * per;
    if (r0 >= n) return;
    const uint32_t cnt = static_cast<uint32_t>(n - r0 < per ? n - r0 : per);
    const uint64_t b0 = r0 * L;                 // first byte of the run
    const uint64_t w0 = b0 >> 2;                // its dword
    const uint32_t lead = static_cast<uint32_t>(b0 & 3u);
    const uint32_t words = (lead + cnt * L + 3u) >> 2;
    for (uint32_t i = threadIdx.x; i < words; i += kLasThreads) s_rec[i] = rec[w0 + i];
    __syncthreads();
    const uint8_t* bytes = reinterpret_cast<const uint8_t*>(s_rec) + lead;
    for (uint32_t r = threadIdx.x; r < cnt; r += kLasThreads) {
        const uint8_t* p = bytes + static_cast<size_t>(r) * L;
        const int32_t xi = lds_i32(p), yi = lds_i32(p + 4), zi = lds_i32(p + 8);
        const uint64_t o = r0 + r;
        x[o] = __dadd_rn(__dmul_rn(static_cast<double>(xi), xf.scale[0]), xf.offset[0]);
        y[o] = __dadd_rn(__dmul_rn(static_cast<double>(yi), xf.scale[1]), xf.offset[1]);
        z[o] = __dadd_rn(__dmul_rn(static_cast<double>(zi), xf.scale[2]), xf.offset[2]);
    }
}

// Row i's ncols doubles to out + i * stride + offset, little-endian; 8-byte
// stores when every destination is 8-byte aligned, byte stores otherwise.
__global__ __launch_bounds__(kLasThreads) void k_las_pack(LasColumns cols, int ncols, uint64_t n,
                                                          uint8_t* __restrict__ out,
                                                          uint32_t stride, uint32_t offset,
                                                          int aligned)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kLasThreads) + threadIdx.x;
    if (i >= n) return;
    uint8_t* dst = out + i * stride + offset;
    for (int c = 0; c < ncols; ++c) {
        const double v = cols.col[c][i];
        if (aligned) {
            reinterpret_cast<double*>(dst)[c] = v;
        } else {
            uint64_t bits;
            __builtin_memcpy(&bits, &v, 8);
            for (int k = 0; k < 8; ++k) dst[8 * c + k] = static_cast<uint8_t>(bits >> (8 * k));
        }
    }
}

uint16_t rd16(const uint8_t* p) { return static_cast<uint16_t>(p[0] | (p[1] << 8)); }
uint32_t rd32(const uint8_t* p)
{
    return static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 8) |
           (static_cast<uint32_t>(p[2]) << 16) | (static_cast<uint32_t>(p[3]) << 24);
}
uint64_t rd64(const uint8_t* p)
{
    return static_cast<uint64_t>(rd32(p)) | (static_cast<uint64_t>(rd32(p + 4)) << 32);
}
double rdf64(const uint8_t* p)
{
    const uint64_t b = rd64(p);
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}

} // namespace

// Base record size of a supported point format, 0 otherwise
// (Header.cpp:15-36 with Utils.hpp:93-107).
uint32_t las_base_size(int format)
{
    switch (format) {
    case 0: return 20;
    case 1: return 28;
    case 2: return 26;
    case 3: return 34;
    case 6: return 30;
    case 7: return 36;
    case 8: return 38;
    default: return 0;
    }
}

// Header fields (Header.cpp:38-75); the first of PDAL's validation errors
// (Header.cpp:118-144) goes to `err` and the return value is false.
bool las_parse(const uint8_t* b, uint64_t nbytes, gf_las_info& h, std::string& err)
{
    std::memset(&h, 0, sizeof(h));
    if (nbytes < 227) {
        err = "Invalid LAS header: file shorter than the header.";
        return false;
    }
    const bool magic = std::memcmp(b, "LASF", 4) == 0;
    h.version_major = b[24];
    h.version_minor = b[25];
    h.header_size = rd16(b + 94);
    h.point_offset = rd32(b + 96);
    const uint8_t bits = b[104];
    h.point_format = bits & 0x0F;
    h.compressed = (bits & 0x80) ? 1 : 0;
    h.point_size = rd16(b + 105);
    h.point_count = rd32(b + 107);
    for (int d = 0; d < 3; ++d) {
        h.scale[d] = rdf64(b + 131 + 8 * d);
        h.offset[d] = rdf64(b + 155 + 8 * d);
        h.max[d] = rdf64(b + 179 + 16 * d);
        h.min[d] = rdf64(b + 187 + 16 * d);
    }
    if (h.version_minor >= 3) {
        if (nbytes < (h.version_minor >= 4 ? 375u : 235u)) {
            err = "Invalid LAS header: file shorter than the header.";
            return false;
        }
        if (h.version_minor >= 4) h.point_count = rd64(b + 247);
    }
    if (!magic) {
        err = "Invalid file signature. Was expecting 'LASF', Check the first four  bytes of "
              "the file.";
        return false;
    }
    if (!h.compressed && h.point_offset > nbytes) {
        err = "Invalid point offset - exceeds file size.";
        return false;
    }
    // count * size > bytes after the offset, without overflow.
    if (!h.compressed && h.point_size &&
        h.point_count > (nbytes - h.point_offset) / h.point_size) {
        err = "Invalid point count: " + std::to_string(h.point_count) +
              ". Number of points too large for file size.";
        return false;
    }
    if (h.header_size > nbytes) {
        err = "Invalid VLR offset - exceeds file size.";
        return false;
    }
    const uint32_t base = las_base_size(h.point_format);
    if (!base) {
        err = "Unsupported LAS input point format: " + std::to_string(h.point_format) + ".";
        return false;
    }
    if (h.point_size < base) {
        err = "Invalid point size of " + std::to_string(h.point_size) +
              ". Less than minimum required for point format " +
              std::to_string(h.point_format) + ".";
        return false;
    }
    if (h.version_major != 1 || h.version_minor > 4) {
        err = "Unsupported LAS version '" + std::to_string(h.version_major) + "." +
              std::to_string(h.version_minor) + "'.";
        return false;
    }
    return true;
}

// At most one record per thread (measured: four per thread, 34% of peak
// against 49%; the short kernel wants every workgroup it can get).
uint32_t las_records_per_block(uint32_t L)
{
    const uint32_t per = (kLasStageBytes - 8) / L;
    return per < 1 ? 1 : (per > kLasThreads ? kLasThreads : per);
}

void launch_las_xyz(const uint8_t* rec, uint64_t n, uint32_t L, const LasXform& xf, double* x,
                    double* y, double* z, hipStream_t s)
{
    if (!n) return;
    const uint32_t per = las_records_per_block(L);
    const size_t lds = (static_cast<size_t>(per) * L + 8 + 3) & ~static_cast<size_t>(3);
    hipLaunchKernelGGL(k_las_xyz, dim3(static_cast<uint32_t>(div_up(n, per))),
                       dim3(kLasThreads), lds, s, reinterpret_cast<const uint32_t*>(rec), n, L,
                       per, xf, x, y, z);
}

namespace {

// A double as LAS Extra Bytes data type t (1..10: u8, i8, u16, i16, u32,
// i32, u64, i64, f32, f64): integer types take the value rounded toward
// zero and held to the type's range (NaN: 0); returns the bits and size.
__device__ uint64_t las_typed_bits(double v, int t, int& size)
{
    auto clampd = [](double x, double lo, double hi) {
        return x != x ? 0.0 : (x < lo ? lo : (x > hi ? hi : x));
    };
    switch (t) {
    case 1: size = 1; return static_cast<uint8_t>(clampd(v, 0.0, 255.0));
    case 2: size = 1; return static_cast<uint8_t>(static_cast<int8_t>(clampd(v, -128.0, 127.0)));
    case 3: size = 2; return static_cast<uint16_t>(clampd(v, 0.0, 65535.0));
    case 4:
        size = 2;
        return static_cast<uint16_t>(static_cast<int16_t>(clampd(v, -32768.0, 32767.0)));
    case 5: size = 4; return static_cast<uint32_t>(clampd(v, 0.0, 4294967295.0));
    case 6:
        size = 4;
        return static_cast<uint32_t>(static_cast<int32_t>(clampd(v, -2147483648.0, 2147483647.0)));
    case 7: {
        size = 8;
        const double c = clampd(v, 0.0, 18446744073709549568.0);  // largest double < 2^64
        return static_cast<uint64_t>(c);
    }
    case 8: {
        size = 8;
        const double c = clampd(v, -9223372036854775808.0, 9223372036854774784.0);
        return static_cast<uint64_t>(static_cast<int64_t>(c));
    }
    case 9: {
        size = 4;
        const float f = static_cast<float>(v);
        uint32_t b;
        __builtin_memcpy(&b, &f, 4);
        return b;
    }
    default: {
        size = 8;
        uint64_t b;
        __builtin_memcpy(&b, &v, 8);
        return b;
    }
    }
}

__global__ __launch_bounds__(kLasThreads) void k_las_pack_typed(LasColumns cols, LasTypes types,
                                                                int ncols, uint64_t n,
                                                                uint8_t* __restrict__ out,
                                                                uint32_t stride,
                                                                uint32_t offset)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(kLasThreads) + threadIdx.x;
    if (i >= n) return;
    uint8_t* dst = out + i * stride + offset;
    for (int c = 0; c < ncols; ++c) {
        int size = 8;
        const uint64_t bits = las_typed_bits(cols.col[c][i], types.type[c], size);
        for (int k = 0; k < size; ++k) dst[k] = static_cast<uint8_t>(bits >> (8 * k));
        dst += size;
    }
}

} // namespace

uint32_t las_type_size(int t)
{
    static const uint32_t size[11] = {0, 1, 1, 2, 2, 4, 4, 8, 8, 4, 8};
    return (t >= 1 && t <= 10) ? size[t] : 0;
}

void launch_las_pack_typed(const LasColumns& cols, const LasTypes& types, int ncols, uint64_t n,
                           uint8_t* out, uint32_t stride, uint32_t offset, hipStream_t s)
{
    if (!n || !ncols) return;
    hipLaunchKernelGGL(k_las_pack_typed, dim3(static_cast<uint32_t>(div_up(n, kLasThreads))),
                       dim3(kLasThreads), 0, s, cols, types, ncols, n, out, stride, offset);
}

void launch_las_pack(const LasColumns& cols, int ncols, uint64_t n, uint8_t* out,
                     uint32_t stride, uint32_t offset, hipStream_t s)
{
    if (!n || !ncols) return;
    const int aligned = (reinterpret_cast<uintptr_t>(out) % 8 == 0 && stride % 8 == 0 &&
                         offset % 8 == 0)
                            ? 1
                            : 0;
    hipLaunchKernelGGL(k_las_pack, dim3(static_cast<uint32_t>(div_up(n, kLasThreads))),
                       dim3(kLasThreads), 0, s, cols, ncols, n, out, stride, offset, aligned);
}

} // namespace gf
