// las_write.cpp -- writers.las extra_dims: a LAS file image from a source
// image, with columns (FP64 values, each written as its LAS Extra Bytes
// data type) appended to every point record and described in the Extra
// Bytes VLR (io/LasWriter.cpp:583-596, io/private/las/Utils.cpp:244-267;
// LAS 1.4 R15 2.6 "Extra Bytes").
//
// The header and VLRs are rewritten on the host; the point records are
// widened and packed in HBM: one pitched copy spreads the source records to
// the wider stride, the pack kernel writes the columns little-endian after
// each record's source bytes, and the records come back in one copy.
#include "gf_kernels.hpp"
#include "gf_runtime.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace gf {

namespace {

constexpr uint32_t kVlrHeader = 54;
constexpr uint32_t kEbDescriptor = 192;
constexpr uint16_t kEbRecordId = 4;
constexpr uint8_t kEbDouble = 10;

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
void wr16(uint8_t* p, uint16_t v)
{
    p[0] = static_cast<uint8_t>(v);
    p[1] = static_cast<uint8_t>(v >> 8);
}
void wr32(uint8_t* p, uint32_t v)
{
    for (int i = 0; i < 4; ++i) p[i] = static_cast<uint8_t>(v >> (8 * i));
}
void wr64(uint8_t* p, uint64_t v)
{
    for (int i = 0; i < 8; ++i) p[i] = static_cast<uint8_t>(v >> (8 * i));
}

// Bytes one Extra Bytes descriptor describes: data_type 0 is `options`
// undocumented bytes; 1..10 one value of u8, i8, u16, i16, u32, i32, u64,
// i64, float, double; 11..30 two or three of them.
uint32_t eb_bytes(const uint8_t* d)
{
    static const uint32_t size[11] = {0, 1, 1, 2, 2, 4, 4, 8, 8, 4, 8};
    const uint8_t type = d[2];
    if (type == 0) return d[3];
    if (type > 30) return 0;
    const uint32_t fields = (type - 1) / 10 + 1;
    return fields * size[(type - 1) % 10 + 1];
}

// A descriptor (io/private/las/Utils.cpp:244-267): one typed value with the
// scale and offset options clear, no-data/min/max zero, name and
// description zero-padded to 32 bytes (longer ones truncated).
void put_descriptor(std::vector<uint8_t>& out, uint8_t type, uint8_t options,
                    const std::string& name, const std::string& description)
{
    uint8_t d[kEbDescriptor];
    std::memset(d, 0, sizeof(d));
    d[2] = type;
    d[3] = options;
    std::memcpy(d + 4, name.data(), std::min<size_t>(name.size(), 32));
    std::memcpy(d + 160, description.data(), std::min<size_t>(description.size(), 32));
    out.insert(out.end(), d, d + kEbDescriptor);
}

struct VlrSpan {
    uint64_t begin, size;  // header + data bytes
    bool extra_bytes;
};

} // namespace

void Context::las_write_extra(const uint8_t* src, uint64_t src_nbytes, int ncols,
                              const char* const* names, const char* const* descriptions,
                              const uint8_t* types, const double* const* cols, int where,
                              uint8_t* dst, uint64_t dst_capacity, uint64_t* dst_nbytes)
{
    LasTypes lt{};
    uint32_t extra = 0;
    for (int c = 0; c < ncols; ++c) {
        lt.type[c] = types ? types[c] : kEbDouble;
        const uint32_t sz = las_type_size(lt.type[c]);
        if (!sz) throw Error(GF_ERR_ARG, "extra dimension types are 1..10 (u8 .. f64)");
        extra += sz;
    }
    gf_las_info h;
    std::string err;
    if (!las_parse(src, src_nbytes, h, err)) throw Error(GF_ERR_ARG, err);
    if (h.compressed) throw Error(GF_ERR_UNSUPPORTED, "LAZ-compressed data is not supported");
    const uint32_t nvlr = rd32(src + 100);
    const uint32_t S = h.point_size;
    const uint64_t n = h.point_count;
    const uint32_t base = las_base_size(h.point_format);

    // The source VLRs, and its Extra Bytes VLR if any.
    std::vector<VlrSpan> vlrs;
    uint64_t at = h.header_size;
    int eb = -1;
    for (uint32_t v = 0; v < nvlr; ++v) {
        if (at + kVlrHeader > h.point_offset)
            throw Error(GF_ERR_ARG, "VLR " + std::to_string(v) + " runs past the point data");
        const uint64_t len = kVlrHeader + rd16(src + at + 20);
        if (at + len > h.point_offset)
            throw Error(GF_ERR_ARG, "VLR " + std::to_string(v) + " runs past the point data");
        char user[17] = {0};
        std::memcpy(user, src + at + 2, 16);
        const bool is_eb =
            std::strcmp(user, "LASF_Spec") == 0 && rd16(src + at + 18) == kEbRecordId;
        if (is_eb && eb < 0) eb = static_cast<int>(vlrs.size());
        vlrs.push_back({at, len, is_eb});
        at += len;
    }
    const uint64_t vlr_end = at;

    // Descriptors to add: bytes of the source records past the base format
    // and past what an existing Extra Bytes VLR describes are declared as
    // undocumented first, so the new columns are described at their offset.
    uint64_t described = base;
    if (eb >= 0) {
        const uint8_t* d = src + vlrs[eb].begin + kVlrHeader;
        const uint64_t cnt = (vlrs[eb].size - kVlrHeader) / kEbDescriptor;
        for (uint64_t i = 0; i < cnt; ++i) described += eb_bytes(d + i * kEbDescriptor);
    }
    if (described > S)
        throw Error(GF_ERR_ARG, "the Extra Bytes VLR describes more bytes than the records hold");
    std::vector<uint8_t> add;
    for (uint64_t gap = S - described; gap > 0;) {
        const uint8_t g = static_cast<uint8_t>(std::min<uint64_t>(gap, 255));
        put_descriptor(add, 0, g, "", "");
        gap -= g;
    }
    for (int c = 0; c < ncols; ++c)
        put_descriptor(add, static_cast<uint8_t>(lt.type[c]), 0,
                       names && names[c] ? names[c] : "",
                       descriptions && descriptions[c] ? descriptions[c] : "");
    const uint64_t grow = add.size() + (eb >= 0 ? 0 : kVlrHeader);
    if (eb >= 0 && vlrs[eb].size - kVlrHeader + add.size() > 65535)
        throw Error(GF_ERR_UNSUPPORTED, "the Extra Bytes VLR would exceed 65535 bytes");
    if (S + static_cast<uint64_t>(extra) > 65535)
        throw Error(GF_ERR_UNSUPPORTED, "the point record would exceed 65535 bytes");

    const uint32_t S2 = S + extra;
    const uint64_t new_offset = h.point_offset + grow;
    if (new_offset > 0xFFFFFFFFull) throw Error(GF_ERR_UNSUPPORTED, "point offset above 4 GiB");
    const uint64_t src_points_end = h.point_offset + n * S;
    const uint64_t tail = src_nbytes - src_points_end;  // EVLRs, waveform data
    const uint64_t total = new_offset + n * S2 + tail;
    *dst_nbytes = total;
    if (!dst) return;
    if (dst_capacity < total) throw Error(GF_ERR_ARG, "destination buffer too small");

    // Header.
    std::memcpy(dst, src, h.header_size);
    wr32(dst + 96, static_cast<uint32_t>(new_offset));
    wr32(dst + 100, nvlr + (eb >= 0 ? 0u : 1u));
    wr16(dst + 105, static_cast<uint16_t>(S2));
    const uint64_t after = new_offset + n * S2 - src_points_end;  // shift of what follows
    if (h.version_minor >= 3 && h.header_size >= 235) {
        const uint64_t wave = rd64(src + 227);
        if (wave) wr64(dst + 227, wave + after);
    }
    if (h.version_minor >= 4 && h.header_size >= 243) {
        const uint64_t evlr = rd64(src + 235);
        if (evlr) wr64(dst + 235, evlr + after);
    }
    // VLRs, the grown Extra Bytes VLR in place, or a new one after the last.
    uint8_t* o = dst + h.header_size;
    for (size_t v = 0; v < vlrs.size(); ++v) {
        std::memcpy(o, src + vlrs[v].begin, vlrs[v].size);
        if (static_cast<int>(v) == eb) {
            wr16(o + 20, static_cast<uint16_t>(vlrs[v].size - kVlrHeader + add.size()));
            std::memcpy(o + vlrs[v].size, add.data(), add.size());
            o += add.size();
        }
        o += vlrs[v].size;
    }
    if (eb < 0) {
        std::memset(o, 0, kVlrHeader);
        std::memcpy(o + 2, "LASF_Spec", 9);
        wr16(o + 18, kEbRecordId);
        wr16(o + 20, static_cast<uint16_t>(add.size()));
        std::memcpy(o + 22, "Extra Bytes Record", 18);
        std::memcpy(o + kVlrHeader, add.data(), add.size());
        o += kVlrHeader + add.size();
    }
    // Bytes between the VLRs and the points (kept as they were).
    std::memcpy(o, src + vlr_end, h.point_offset - vlr_end);
    // EVLRs and anything else after the points.
    std::memcpy(dst + new_offset + n * S2, src + src_points_end, tail);
    if (!n) return;

    // Point records: widened and packed in HBM.
    const uint64_t in_bytes = n * S, out_bytes = n * S2;
    lasbuf.ensure(in_bytes + out_bytes + 16);
    uint8_t* din = lasbuf.as<uint8_t>();
    uint8_t* dout = din + ((in_bytes + 15) & ~static_cast<uint64_t>(15));
    GF_HIP(hipMemcpyAsync(din, src + h.point_offset, in_bytes, hipMemcpyHostToDevice, stream));
    GF_HIP(hipMemcpy2DAsync(dout, S2, din, S, S, n, hipMemcpyDeviceToDevice, stream));
    LasColumns lc{};
    if (where == GF_WHERE_DEVICE) {
        for (int c = 0; c < ncols; ++c) lc.col[c] = cols[c];
    } else {
        outbuf.ensure(n * 8 * static_cast<uint64_t>(ncols) + 8);
        for (int c = 0; c < ncols; ++c) {
            double* d = outbuf.as<double>() + static_cast<uint64_t>(c) * n;
            GF_HIP(hipMemcpyAsync(d, cols[c], n * 8, hipMemcpyHostToDevice, stream));
            lc.col[c] = d;
        }
    }
    launch_las_pack_typed(lc, lt, ncols, n, dout, S2, S, stream);
    GF_HIP(hipMemcpyAsync(dst + new_offset, dout, out_bytes, hipMemcpyDeviceToHost, stream));
    GF_HIP(hipStreamSynchronize(stream));
}

} // namespace gf
