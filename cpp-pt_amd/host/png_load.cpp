// PNG reader for image textures (the reference reads them with stb_image,
// src/image/rgb_image.h:161-197; `stbi_load(..., 0)` semantics): returns 8-bit
// samples with the file's own channel count (gray 1, gray+alpha 2, RGB 3,
// RGBA 4; palette images expand to RGB, or RGBA with a tRNS chunk), 16-bit
// samples reduced to their high byte, 1/2/4-bit gray scaled to 0..255.
// Non-interlaced images only.  Inflate is zlib's.
#include <zlib.h>

#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cpt_host.h"

namespace cpppt {

namespace {
uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}
}  // namespace

void load_png(const std::string& path, int* w, int* h, int* n, std::vector<uint8_t>* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot open image file: " + path);
    std::vector<uint8_t> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (buf.size() < 8 || !std::equal(sig, sig + 8, buf.begin()))
        throw std::runtime_error("Unsupported image format (PNG only): " + path);
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    size_t pos = 8;
    while (pos + 12 <= buf.size()) {
        uint32_t len = be32(&buf[pos]);
        if (pos + 12 + (size_t)len > buf.size()) throw std::runtime_error("Truncated PNG: " + path);
        std::string type(reinterpret_cast<const char*>(&buf[pos + 4]), 4);
        const uint8_t* d = &buf[pos + 8];
        if (type == "IHDR") {
            if (len < 13) throw std::runtime_error("Bad PNG header: " + path);
            W = be32(d);
            H = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (type == "PLTE") {
            plte.assign(d, d + len);
        } else if (type == "tRNS") {
            trns.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (W == 0 || H == 0 || W > (1u << 15) || H > (1u << 15)) throw std::runtime_error("Bad PNG size: " + path);
    if (interlace) throw std::runtime_error("Interlaced PNG not supported: " + path);
    int ch;
    switch (ctype) {
        case 0: ch = 1; break;
        case 2: ch = 3; break;
        case 3: ch = 1; break;
        case 4: ch = 2; break;
        case 6: ch = 4; break;
        default: throw std::runtime_error("Bad PNG colour type: " + path);
    }
    const bool depth_ok = depth == 8 || depth == 16 || ((ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4));
    if (!depth_ok) throw std::runtime_error("Unsupported PNG bit depth: " + path);
    const size_t bits_px = (size_t)ch * depth;
    const size_t stride = (W * bits_px + 7) / 8;
    const size_t bpp = std::max<size_t>(1, bits_px / 8);  // filter unit
    std::vector<uint8_t> raw((stride + 1) * H);
    uLongf raw_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK || raw_len != raw.size())
        throw std::runtime_error("Corrupt PNG data: " + path);
    // undo the per-row filters
    std::vector<uint8_t> img(stride * H);
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t* s = &raw[y * (stride + 1) + 1];
        uint8_t* r = &img[y * stride];
        const uint8_t* p = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t x = 0; x < stride; x++) {
            int a = x >= bpp ? r[x - bpp] : 0, b = p ? p[x] : 0, c = (p && x >= bpp) ? p[x - bpp] : 0;
            int v;
            switch (ft) {
                case 0: v = s[x]; break;
                case 1: v = s[x] + a; break;
                case 2: v = s[x] + b; break;
                case 3: v = s[x] + ((a + b) >> 1); break;
                case 4: v = s[x] + paeth(a, b, c); break;
                default: throw std::runtime_error("Bad PNG filter: " + path);
            }
            r[x] = (uint8_t)v;
        }
    }
    // samples -> 8 bit, palette expanded
    const int out_ch = ctype == 3 ? (trns.empty() ? 3 : 4) : ch;
    out->assign((size_t)W * H * out_ch, 0);
    for (uint32_t y = 0; y < H; y++) {
        const uint8_t* r = &img[y * stride];
        for (uint32_t x = 0; x < W; x++) {
            uint8_t* o = &(*out)[((size_t)y * W + x) * out_ch];
            if (depth == 8) {
                if (ctype == 3) {
                    const size_t k = r[x];
                    if (3 * k + 2 >= plte.size()) throw std::runtime_error("Bad PNG palette index: " + path);
                    o[0] = plte[3 * k];
                    o[1] = plte[3 * k + 1];
                    o[2] = plte[3 * k + 2];
                    if (out_ch == 4) o[3] = k < trns.size() ? trns[k] : 255;
                } else {
                    for (int c = 0; c < ch; c++) o[c] = r[(size_t)x * ch + c];
                }
            } else if (depth == 16) {
                for (int c = 0; c < ch; c++) o[c] = r[((size_t)x * ch + c) * 2];  // high byte
            } else {  // 1, 2, 4 bit gray or palette
                const int per = 8 / depth, shift = (per - 1 - (int)(x % per)) * depth;
                const int v = (r[x / per] >> shift) & ((1 << depth) - 1);
                if (ctype == 3) {
                    if (3 * (size_t)v + 2 >= plte.size()) throw std::runtime_error("Bad PNG palette index: " + path);
                    o[0] = plte[3 * v];
                    o[1] = plte[3 * v + 1];
                    o[2] = plte[3 * v + 2];
                    if (out_ch == 4) o[3] = (size_t)v < trns.size() ? trns[v] : 255;
                } else {
                    o[0] = (uint8_t)(v * (255 / ((1 << depth) - 1)));
                }
            }
        }
    }
    *w = (int)W;
    *h = (int)H;
    *n = out_ch;
}

}  // namespace cpppt
