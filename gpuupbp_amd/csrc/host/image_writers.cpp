// BMP and Radiance HDR image output and the extension dispatch of the reference's
// save() (nolmoonen/gpuupbp gpuupbp/host/framebuffer.cpp:114-233).
//
// BMP: 24 bits per pixel, uncompressed, rows bottom-up in B,G,R order after gamma
// correction and clamping to [0, 255]. Rows are padded to four bytes as the format requires
// (the reference writes them unpadded, which is identical when width * 3 is a multiple of 4).
// HDR: "#?RADIANCE" header, FORMAT=32-bit_rle_rgbe, "-Y h +X w", flat (not run-length
// encoded) scanlines of shared-exponent RGBE pixels (Ward, "Real pixels", Graphics Gems II).
#include "upbp_capi.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

void put_u32(std::vector<unsigned char>& b, uint32_t v)
{
    for (int i = 0; i < 4; i++) b.push_back(static_cast<unsigned char>((v >> (8 * i)) & 255u));
}
void put_u16(std::vector<unsigned char>& b, uint16_t v)
{
    b.push_back(static_cast<unsigned char>(v & 255u));
    b.push_back(static_cast<unsigned char>(v >> 8));
}

int write_all(const char* path, const std::vector<unsigned char>& data)
{
    FILE* f = fopen(path, "wb");
    if (!f) return UPBP_E_IO;
    const size_t n = fwrite(data.data(), 1, data.size(), f);
    const int rc = fclose(f);
    return (n == data.size() && rc == 0) ? UPBP_OK : UPBP_E_IO;
}

} // namespace

extern "C" int upbp_save_bmp(const char* path, const float* rgb, uint32_t width, uint32_t height, float gamma)
{
    if (!path || !rgb || width == 0 || height == 0 || !(gamma > 0.f)) return UPBP_E_ARG;
    const uint32_t row = (width * 3u + 3u) & ~3u;
    const uint32_t data = row * height;
    std::vector<unsigned char> b;
    b.reserve(54 + data);
    b.push_back('B');
    b.push_back('M');
    put_u32(b, 54 + data); // file size
    put_u32(b, 0);         // reserved
    put_u32(b, 54);        // pixel data offset
    put_u32(b, 40);        // BITMAPINFOHEADER size
    put_u32(b, width);
    put_u32(b, height);
    put_u16(b, 1);  // planes
    put_u16(b, 24); // bits per pixel
    put_u32(b, 0);  // BI_RGB
    put_u32(b, data);
    put_u32(b, 2953); // 75 dpi
    put_u32(b, 2953);
    put_u32(b, 0);
    put_u32(b, 0);
    const float inv_gamma = 1.f / gamma;
    auto byte_of = [&](float v) {
        const float g = std::pow(v, inv_gamma) * 255.f;
        return static_cast<unsigned char>(std::min(255.f, std::max(0.f, g)));
    };
    for (uint32_t y = 0; y < height; y++) { // bottom row first
        const float* src = rgb + size_t(height - 1 - y) * width * 3;
        for (uint32_t x = 0; x < width; x++) {
            b.push_back(byte_of(src[3 * x + 2]));
            b.push_back(byte_of(src[3 * x + 1]));
            b.push_back(byte_of(src[3 * x]));
        }
        for (uint32_t p = width * 3; p < row; p++) b.push_back(0);
    }
    return write_all(path, b);
}

extern "C" int upbp_save_hdr(const char* path, const float* rgb, uint32_t width, uint32_t height)
{
    if (!path || !rgb || width == 0 || height == 0) return UPBP_E_ARG;
    std::string head = "#?RADIANCE\n# gpuupbp_amd\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(height) + " +X " +
                       std::to_string(width) + "\n";
    std::vector<unsigned char> b(head.begin(), head.end());
    b.reserve(b.size() + size_t(width) * height * 4);
    for (size_t i = 0; i < size_t(width) * height; i++) {
        const float r = rgb[3 * i], g = rgb[3 * i + 1], bl = rgb[3 * i + 2];
        const float v = std::max(r, std::max(g, bl));
        unsigned char e[4] = {0, 0, 0, 0};
        if (v >= 1e-32f) {
            int ex;
            const float scale = std::frexp(v, &ex) * 256.f / v;
            e[0] = static_cast<unsigned char>(r * scale);
            e[1] = static_cast<unsigned char>(g * scale);
            e[2] = static_cast<unsigned char>(bl * scale);
            e[3] = static_cast<unsigned char>(ex + 128);
        }
        b.insert(b.end(), e, e + 4);
    }
    return write_all(path, b);
}

extern "C" int upbp_save_image(const char* path, const float* rgb, uint32_t width, uint32_t height)
{
    if (!path) return UPBP_E_ARG;
    const size_t n = strlen(path);
    const std::string ext = n >= 3 ? std::string(path + n - 3) : std::string();
    if (ext == "exr") return upbp_save_exr(path, rgb, width, height);
    if (ext == "bmp") return upbp_save_bmp(path, rgb, width, height, 2.2f);
    if (ext == "hdr") return upbp_save_hdr(path, rgb, width, height);
    return UPBP_E_ARG;
}
