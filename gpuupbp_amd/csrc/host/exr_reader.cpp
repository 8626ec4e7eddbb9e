// OpenEXR reader for environment maps: replaces tinyexr_wrapper::load_image
// (nolmoonen/gpuupbp gpuupbp/misc/tinyexr_wrapper.cpp:97-136, which calls tinyexr's LoadEXR).
//
// Written from the published OpenEXR file layout ("Reading and Writing OpenEXR Image Files",
// "OpenEXR File Layout"): magic 20000630, version 2 single-part scanline files; a header of
// (name, type, size, value) attributes ending in an empty name; a table of chunk offsets; chunks
// of (y, byte count, data) holding lines_per_chunk scanlines, each scanline storing the channels
// in header order (alphabetical), each channel as width values of its pixel type.
// Compressions: NONE (1 line per chunk), RLE (1), ZIPS (1), ZIP (16). RLE and ZIP data are
// byte-wise delta coded and split into even / odd byte halves before compression; both steps
// are undone here. PIZ, PXR24 and the B44 / DWA family are rejected with a message.
// Channels R, G, B (A is ignored) give the colour; a file with one channel gives grey.
#include "upbp_capi.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

float half_to_float(uint16_t h)
{
    const uint32_t s = uint32_t(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else { // subnormal: m * 2^-24
            const float v = std::ldexp(float(m), -24);
            return (h >> 15) ? -v : v;
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

struct Reader {
    const std::vector<unsigned char>& b;
    size_t p = 0;
    bool ok = true;
    template <typename T>
    T get()
    {
        T v{};
        // written as a subtraction so that a huge p (a crafted offset) cannot wrap around
        if (p > b.size() || sizeof(T) > b.size() - p) {
            ok = false;
            return v;
        }
        memcpy(&v, &b[p], sizeof(T));
        p += sizeof(T);
        return v;
    }
    std::string str()
    {
        std::string s;
        while (p < b.size() && b[p]) s += char(b[p++]);
        if (p >= b.size()) ok = false;
        p++;
        return s;
    }
};

struct Channel {
    std::string name;
    int type; // 0 UINT, 1 HALF, 2 FLOAT
    int xs, ys;
};

// undo the predictor and the even / odd byte split of RLE and ZIP data
void unpredict(std::vector<unsigned char>& t, unsigned char* out)
{
    const size_t n = t.size();
    for (size_t i = 1; i < n; i++) t[i] = static_cast<unsigned char>(int(t[i - 1]) + int(t[i]) - 128);
    const size_t half = (n + 1) / 2;
    for (size_t i = 0; i < n; i++) out[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
}

bool rle_decode(const unsigned char* in, size_t n, std::vector<unsigned char>& out, size_t expect)
{
    size_t i = 0;
    while (i < n) {
        const int c = static_cast<signed char>(in[i++]);
        if (c < 0) {
            if (i + size_t(-c) > n) return false;
            out.insert(out.end(), in + i, in + i - c);
            i += size_t(-c);
        } else {
            if (i >= n) return false;
            out.insert(out.end(), size_t(c) + 1, in[i++]);
        }
        if (out.size() > expect) return false;
    }
    return out.size() == expect;
}

} // namespace

// Reads an OpenEXR file into RGB floats, rows from the top. Returns an empty string on success.
std::string upbp_read_exr_rgb(const std::string& path, std::vector<float>& rgb, uint32_t& width, uint32_t& height)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return "could not open " + path;
    std::vector<unsigned char> b;
    unsigned char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + n);
    fclose(f);
    Reader r{b};
    if (r.get<uint32_t>() != 20000630u) return path + ": not an OpenEXR file";
    const uint32_t version = r.get<uint32_t>();
    if ((version & 255u) != 2u || (version & (0x200u | 0x800u | 0x1000u)))
        return path + ": only single-part scanline OpenEXR files are supported";
    std::vector<Channel> ch;
    int comp = -1, xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    for (;;) {
        const std::string name = r.str();
        if (!r.ok) return path + ": truncated header";
        if (name.empty()) break;
        const std::string type = r.str();
        const int32_t size = r.get<int32_t>();
        if (!r.ok || size < 0 || size_t(size) > b.size() - r.p) return path + ": truncated header";
        const size_t end = r.p + size_t(size);
        if (name == "channels" && type == "chlist") {
            while (r.p < end) {
                const std::string cn = r.str();
                if (cn.empty()) break;
                Channel c;
                c.name = cn;
                c.type = r.get<int32_t>();
                r.get<uint32_t>(); // pLinear + reserved
                c.xs = r.get<int32_t>();
                c.ys = r.get<int32_t>();
                ch.push_back(c);
            }
        } else if (name == "compression") {
            if (size < 1) return path + ": empty compression attribute";
            comp = b[r.p];
        } else if (name == "dataWindow" && size == 16) {
            xmin = r.get<int32_t>();
            ymin = r.get<int32_t>();
            xmax = r.get<int32_t>();
            ymax = r.get<int32_t>();
        }
        r.p = end;
    }
    if (ch.empty() || xmax < xmin || ymax < ymin) return path + ": missing channels or data window";
    if (int64_t(xmax) - xmin >= (1 << 16) || int64_t(ymax) - ymin >= (1 << 16)) return path + ": image too large";
    for (const Channel& c : ch)
        if (c.xs != 1 || c.ys != 1 || c.type < 0 || c.type > 2) return path + ": sub-sampled channels are not supported";
    int lines;
    if (comp == 0 || comp == 1 || comp == 2) lines = 1;
    else if (comp == 3) lines = 16;
    else return path + ": unsupported OpenEXR compression " + std::to_string(comp) + " (NONE, RLE, ZIPS and ZIP are read)";
    width = uint32_t(xmax - xmin + 1);
    height = uint32_t(ymax - ymin + 1);
    size_t pixel_bytes = 0;
    for (const Channel& c : ch) pixel_bytes += c.type == 1 ? 2 : 4;
    const size_t line_bytes = pixel_bytes * width;
    // deflate expands at most ~1032:1 and RLE 64:1, so a data window beyond that ratio of the file
    // is corrupt; checked before the pixels are allocated
    if (double(line_bytes) * height > 1100.0 * double(b.size()) + double(1 << 20))
        return path + ": data window larger than the file can hold";
    const uint32_t chunks = (height + uint32_t(lines) - 1) / uint32_t(lines);
    std::vector<uint64_t> offs(chunks);
    for (auto& o : offs) o = r.get<uint64_t>();
    if (!r.ok) return path + ": truncated offset table";
    for (const uint64_t o : offs) // each chunk starts with its 8-byte (y, size) prefix
        if (b.size() < 8 || o > uint64_t(b.size() - 8)) return path + ": chunk offset outside the file";

    int ci[3] = {-1, -1, -1};
    for (size_t k = 0; k < ch.size(); k++) {
        if (ch[k].name == "R") ci[0] = int(k);
        if (ch[k].name == "G") ci[1] = int(k);
        if (ch[k].name == "B") ci[2] = int(k);
    }
    if (ci[0] < 0 || ci[1] < 0 || ci[2] < 0) {
        if (ch.size() != 1) return path + ": no R, G, B channels";
        ci[0] = ci[1] = ci[2] = 0;
    }
    std::vector<size_t> ch_off(ch.size());
    for (size_t k = 0, o = 0; k < ch.size(); k++) {
        ch_off[k] = o;
        o += (ch[k].type == 1 ? 2 : 4) * size_t(width);
    }
    rgb.assign(size_t(width) * height * 3, 0.f);
    std::vector<unsigned char> raw, tmp;
    for (uint32_t c = 0; c < chunks; c++) {
        Reader cr{b};
        cr.p = size_t(offs[c]);
        const int32_t y = cr.get<int32_t>();
        const int32_t size = cr.get<int32_t>();
        if (!cr.ok || size < 0 || size_t(size) > b.size() - cr.p) return path + ": truncated chunk";
        if (y < ymin || y > ymax) return path + ": chunk outside the data window";
        const uint32_t y0 = uint32_t(y - ymin);
        const uint32_t nl = std::min<uint32_t>(uint32_t(lines), height - y0);
        const size_t expect = line_bytes * nl;
        const unsigned char* src = &b[cr.p];
        raw.resize(expect);
        if (size_t(size) == expect) { // stored as is (also when compression did not pay)
            memcpy(raw.data(), src, expect);
        } else if (comp == 1) {
            tmp.clear();
            if (!rle_decode(src, size_t(size), tmp, expect)) return path + ": corrupt RLE chunk";
            unpredict(tmp, raw.data());
        } else if (comp == 2 || comp == 3) {
            tmp.resize(expect);
            uLongf dl = uLongf(expect);
            if (uncompress(tmp.data(), &dl, src, uLong(size)) != Z_OK || dl != expect)
                return path + ": corrupt ZIP chunk";
            unpredict(tmp, raw.data());
        } else {
            return path + ": chunk size does not match an uncompressed scanline";
        }
        for (uint32_t l = 0; l < nl; l++) {
            const unsigned char* line = raw.data() + line_bytes * l;
            float* dst = rgb.data() + size_t(y0 + l) * width * 3;
            for (int k = 0; k < 3; k++) {
                const Channel& C = ch[size_t(ci[k])];
                const unsigned char* s = line + ch_off[size_t(ci[k])];
                for (uint32_t x = 0; x < width; x++) {
                    float v;
                    if (C.type == 1) {
                        uint16_t h;
                        memcpy(&h, s + 2 * x, 2);
                        v = half_to_float(h);
                    } else if (C.type == 2) {
                        memcpy(&v, s + 4 * x, 4);
                    } else {
                        uint32_t u;
                        memcpy(&u, s + 4 * x, 4);
                        v = float(u);
                    }
                    dst[3 * x + k] = v;
                }
            }
        }
    }
    return std::string();
}

extern "C" int upbp_read_exr(const char* path, float* rgb, uint64_t capacity_floats, uint32_t* width, uint32_t* height)
{
    if (!path || !width || !height) return UPBP_E_ARG;
    std::vector<float> img;
    uint32_t w = 0, h = 0;
    std::string err;
    try { // the C ABI lets no C++ exception through (an image too large for host memory)
        err = upbp_read_exr_rgb(path, img, w, h);
    } catch (const std::exception& e) {
        err = std::string(path) + ": " + e.what();
    }
    if (!err.empty()) {
        fprintf(stderr, "Error: %s\n", err.c_str());
        return err.compare(0, 14, "could not open") == 0 ? UPBP_E_IO : UPBP_E_PARSE;
    }
    *width = w;
    *height = h;
    if (!rgb) return UPBP_OK;
    if (capacity_floats < img.size()) return UPBP_E_CAPACITY;
    memcpy(rgb, img.data(), img.size() * sizeof(float));
    return UPBP_OK;
}
