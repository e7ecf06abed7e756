// png.cpp -- see png.hpp.  Chunk parsing, zlib inflate, the five PNG scanline
// filters (PNG specification, section 9), channel expansion.
#include "png.hpp"

#include <zlib.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace sol {

namespace {

uint32_t be32(const uint8_t *p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

}  // namespace

ImageRgba8 load_png_rgba8_flipv(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<uint8_t> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (raw.size() < 8 || std::memcmp(raw.data(), sig, 8) != 0) throw std::runtime_error("not a PNG: " + path);
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat;
    size_t off = 8;
    while (off + 12 <= raw.size()) {
        const uint32_t len = be32(&raw[off]);
        const char *type = (const char *)&raw[off + 4];
        if (off + 12 + (size_t)len > raw.size()) throw std::runtime_error("truncated PNG chunk");
        const uint8_t *data = &raw[off + 8];
        if (std::memcmp(type, "IHDR", 4) == 0) {
            W = be32(data);
            H = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), data, data + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            break;
        }
        off += 12 + (size_t)len;
    }
    if (interlace) throw std::runtime_error("interlaced PNG not supported");
    int channels;
    switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: throw std::runtime_error("unsupported PNG colour type");
    }
    if (depth != 8 && depth != 16) throw std::runtime_error("unsupported PNG bit depth");
    const size_t bpp = (size_t)channels * (size_t)(depth / 8);
    const size_t stride = (size_t)W * bpp;
    std::vector<uint8_t> inflated((stride + 1) * H);
    uLongf dst_len = (uLongf)inflated.size();
    if (uncompress(inflated.data(), &dst_len, idat.data(), (uLong)idat.size()) != Z_OK || dst_len != inflated.size())
        throw std::runtime_error("PNG inflate failed");
    std::vector<uint8_t> img(stride * H);
    for (uint32_t y = 0; y < H; ++y) {
        const uint8_t filter = inflated[y * (stride + 1)];
        const uint8_t *src = &inflated[y * (stride + 1) + 1];
        uint8_t *row = &img[y * stride];
        const uint8_t *prev = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? row[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int v = src[i];
            switch (filter) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default: throw std::runtime_error("bad PNG filter");
            }
            row[i] = (uint8_t)v;
        }
    }
    ImageRgba8 out;
    out.width = W;
    out.height = H;
    out.rgba.resize((size_t)W * H * 4);
    for (uint32_t y = 0; y < H; ++y) {
        const uint8_t *row = &img[(size_t)(H - 1 - y) * stride];   // flipv
        for (uint32_t x = 0; x < W; ++x) {
            uint32_t ch[4] = {0, 0, 0, 255};
            for (int k = 0; k < channels; ++k) {
                uint32_t v;
                if (depth == 16) {
                    const uint32_t v16 = ((uint32_t)row[x * bpp + 2 * k] << 8) | row[x * bpp + 2 * k + 1];
                    v = (v16 + 128u) / 257u;
                } else {
                    v = row[x * bpp + k];
                }
                ch[k] = v;
            }
            uint8_t *o = &out.rgba[((size_t)y * W + x) * 4];
            if (channels == 1 || channels == 2) {
                o[0] = o[1] = o[2] = (uint8_t)ch[0];
                o[3] = channels == 2 ? (uint8_t)ch[1] : 255;
            } else {
                o[0] = (uint8_t)ch[0];
                o[1] = (uint8_t)ch[1];
                o[2] = (uint8_t)ch[2];
                o[3] = channels == 4 ? (uint8_t)ch[3] : 255;
            }
        }
    }
    return out;
}

}  // namespace sol
