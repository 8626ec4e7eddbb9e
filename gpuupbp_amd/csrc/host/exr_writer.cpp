// Minimal OpenEXR writer: single-part scanline image, NO_COMPRESSION, one
// scanline per block, three HALF channels stored in the order B, G, R.
//
// Produces the same file layout that tinyexr_wrapper::save_exr produces in the
// reference (gpuupbp/misc/tinyexr_wrapper.cpp:34-95: channels B,G,R, requested
// pixel type HALF, zero-initialised header => compression NONE), written from
// the OpenEXR file-layout specification rather than with tinyexr.
#include "../../../include/upbp_capi.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

// IEEE binary32 -> binary16, round to nearest even; overflow to +-inf, NaN kept.
uint16_t to_half(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) return static_cast<uint16_t>(sign | 0x7c00u | (man ? 0x200u : 0u));
    const int e = static_cast<int>(exp) - 127 + 15;
    if (e >= 31) return static_cast<uint16_t>(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return static_cast<uint16_t>(sign);
        man |= 0x800000u;
        const int shift = 14 - e; // 14..24
        uint32_t h = man >> shift;
        const uint32_t rem = man & ((1u << shift) - 1u);
        const uint32_t half_point = 1u << (shift - 1);
        if (rem > half_point || (rem == half_point && (h & 1u))) h++;
        return static_cast<uint16_t>(sign | h);
    }
    uint32_t h = (static_cast<uint32_t>(e) << 10) | (man >> 13);
    const uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++; // may carry into the exponent: correct
    return static_cast<uint16_t>(sign | h);
}

struct Buf {
    std::vector<unsigned char> b;
    void bytes(const void* p, size_t n)
    {
        const auto* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
    }
    void str(const char* s) { bytes(s, strlen(s) + 1); }
    void i32(int32_t v) { bytes(&v, 4); } // little-endian host
    void u64(uint64_t v) { bytes(&v, 8); }
    void f32(float v) { bytes(&v, 4); }
    void u8(uint8_t v) { bytes(&v, 1); }
    void attr(const char* name, const char* type, int32_t size)
    {
        str(name);
        str(type);
        i32(size);
    }
};

} // namespace

extern "C" int upbp_save_exr(const char* path, const float* rgb, uint32_t width, uint32_t height)
{
    if (!path || !rgb || width == 0 || height == 0) return UPBP_E_ARG;
    Buf h;
    const unsigned char magic[8] = {0x76, 0x2f, 0x31, 0x01, 0x02, 0x00, 0x00, 0x00};
    h.bytes(magic, 8);

    static const char* chans[3] = {"B", "G", "R"};
    h.attr("channels", "chlist", 3 * (2 + 16) + 1);
    for (const char* c : chans) {
        h.str(c);
        h.i32(1); // HALF
        h.u8(0);  // pLinear
        h.u8(0);
        h.u8(0);
        h.u8(0);
        h.i32(1); // xSampling
        h.i32(1); // ySampling
    }
    h.u8(0);
    h.attr("compression", "compression", 1);
    h.u8(0); // NO_COMPRESSION
    h.attr("dataWindow", "box2i", 16);
    h.i32(0);
    h.i32(0);
    h.i32(static_cast<int32_t>(width) - 1);
    h.i32(static_cast<int32_t>(height) - 1);
    h.attr("displayWindow", "box2i", 16);
    h.i32(0);
    h.i32(0);
    h.i32(static_cast<int32_t>(width) - 1);
    h.i32(static_cast<int32_t>(height) - 1);
    h.attr("lineOrder", "lineOrder", 1);
    h.u8(0); // INCREASING_Y
    h.attr("pixelAspectRatio", "float", 4);
    h.f32(1.f);
    h.attr("screenWindowCenter", "v2f", 8);
    h.f32(0.f);
    h.f32(0.f);
    h.attr("screenWindowWidth", "float", 4);
    h.f32(1.f);
    h.u8(0); // end of header

    const uint64_t line_bytes = 3ull * width * 2ull;
    const uint64_t block_bytes = 8ull + line_bytes;
    uint64_t offset = h.b.size() + 8ull * height;
    for (uint32_t y = 0; y < height; y++) {
        h.u64(offset);
        offset += block_bytes;
    }
    std::vector<uint16_t> line(3ull * width);
    for (uint32_t y = 0; y < height; y++) {
        h.i32(static_cast<int32_t>(y));
        h.i32(static_cast<int32_t>(line_bytes));
        const float* row = rgb + 3ull * width * y;
        for (int c = 0; c < 3; c++) {
            const int src = 2 - c; // B, G, R
            for (uint32_t x = 0; x < width; x++) line[c * width + x] = to_half(row[3ull * x + src]);
        }
        h.bytes(line.data(), line.size() * 2);
    }

    FILE* f = fopen(path, "wb");
    if (!f) return UPBP_E_IO;
    const size_t n = fwrite(h.b.data(), 1, h.b.size(), f);
    const int rc = fclose(f);
    return (n == h.b.size() && rc == 0) ? UPBP_OK : UPBP_E_IO;
}
