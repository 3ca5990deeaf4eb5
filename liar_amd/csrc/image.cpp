// image.cpp -- image writers for tone-mapped films (host code).
//
// LiAR picks an image codec by file extension (src/liar/kernel/image_codec.cpp:38-45,
// src/liar/codecs/lodepng.cpp:248-267, src/liar/codecs/openexr.cpp:95-112, 196-226).
// The formats are written here from their public specifications:
//   PNG  (ISO/IEC 15948): 8-bit RGBA, filter 0, zlib deflate
//   EXR  (OpenEXR file layout): scanline, NO_COMPRESSION, half A/B/G/R
//   Radiance RGBE (Ward, Graphics Gems II): flat scanlines
//   PFM  (Portable Float Map): little-endian float RGB, bottom row first
//   TGA  (Truevision TGA 2.0): uncompressed 32-bit, top-left origin
#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/liar_amd.h"
#include "lr_util.h"

namespace {

using Bytes = std::vector<unsigned char>;

void put32be(Bytes& b, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
}
void put32le(Bytes& b, uint32_t v) {
    for (int s = 0; s < 32; s += 8) b.push_back((unsigned char)(v >> s));
}
void put16le(Bytes& b, uint16_t v) {
    b.push_back((unsigned char)v);
    b.push_back((unsigned char)(v >> 8));
}
void put64le(Bytes& b, uint64_t v) {
    for (int s = 0; s < 64; s += 8) b.push_back((unsigned char)(v >> s));
}
void putf32le(Bytes& b, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put32le(b, u);
}
void putstr(Bytes& b, const char* s) {
    while (*s) b.push_back((unsigned char)*s++);
    b.push_back(0);
}

// The 8-bit quantisation of LiAR's PNG writer (codecs/lodepng.cpp:255-258):
// clamp(255 x, 0, 255), truncated.  NaN maps to 0.
unsigned char q8(float v) {
    float c = 255.0f * v;
    if (!(c > 0.0f)) return 0;
    if (c > 255.0f) c = 255.0f;
    return (unsigned char)c;
}

// IEEE binary16 with round-to-nearest-even (overflow to infinity, NaN kept).
uint16_t to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7FFFFFFFu;
    if (absx > 0x7F800000u) return (uint16_t)(sign | 0x7E00u);  // NaN
    if (absx >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to >= 65520: infinity
    if (absx < 0x38800000u) {  // subnormal half (or zero): value / 2^-24, rounded to even
        if (absx < 0x33000000u) return (uint16_t)sign;  // < 2^-25: rounds to 0
        const uint32_t e = absx >> 23;
        const uint32_t m = (absx & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126 - e;  // value = m 2^(e - 150) = h 2^-24  ->  h = m 2^(e - 126)
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((absx - 0x38000000u) >> 13);  // rebias exponent 127 -> 15
    const uint32_t rem = absx & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
    return (uint16_t)(sign | h);
}

void crc_chunk(Bytes& out, const char* type, const Bytes& data) {
    put32be(out, (uint32_t)data.size());
    const size_t start = out.size();
    for (int k = 0; k < 4; ++k) out.push_back((unsigned char)type[k]);
    out.insert(out.end(), data.begin(), data.end());
    const uLong crc = crc32(0L, out.data() + start, (uInt)(out.size() - start));
    put32be(out, (uint32_t)crc);
}

int encode_png(const float* rgba, uint32_t w, uint32_t h, Bytes& out) {
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    out.assign(sig, sig + 8);
    Bytes ihdr;
    put32be(ihdr, w);
    put32be(ihdr, h);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(6);  // colour type RGBA
    ihdr.push_back(0);  // deflate
    ihdr.push_back(0);  // filter method 0
    ihdr.push_back(0);  // no interlace
    crc_chunk(out, "IHDR", ihdr);
    Bytes raw;
    raw.reserve((size_t)h * (1 + 4 * (size_t)w));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);  // filter type None
        const float* row = rgba + (size_t)y * w * 4;
        for (uint32_t k = 0; k < 4 * w; ++k) raw.push_back(q8(row[k]));
    }
    uLongf zlen = compressBound((uLong)raw.size());
    Bytes z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return lr::fail(LR_E_NOMEM, "png: zlib compression failed");
    z.resize(zlen);
    crc_chunk(out, "IDAT", z);
    crc_chunk(out, "IEND", Bytes());
    return LR_OK;
}

void exr_attr(Bytes& b, const char* name, const char* type, const Bytes& value) {
    putstr(b, name);
    putstr(b, type);
    put32le(b, (uint32_t)value.size());
    b.insert(b.end(), value.begin(), value.end());
}

int encode_exr(const float* rgba, uint32_t w, uint32_t h, const lr_rgb_space& sp, Bytes& out) {
    out.clear();
    put32le(out, 20000630u);  // magic
    put32le(out, 2u);         // version 2, single-part scanline
    {
        Bytes ch;  // chlist, sorted by name: A, B, G, R; HALF = 1
        for (const char* name : {"A", "B", "G", "R"}) {
            putstr(ch, name);
            put32le(ch, 1u);            // pixel type HALF
            ch.push_back(0);            // pLinear
            ch.push_back(0);
            ch.push_back(0);
            ch.push_back(0);            // reserved
            put32le(ch, 1u);            // xSampling
            put32le(ch, 1u);            // ySampling
        }
        ch.push_back(0);
        exr_attr(out, "channels", "chlist", ch);
    }
    {
        Bytes c;
        for (float v : {sp.red[0], sp.red[1], sp.green[0], sp.green[1], sp.blue[0], sp.blue[1], sp.white[0],
                        sp.white[1]})
            putf32le(c, v);
        exr_attr(out, "chromaticities", "chromaticities", c);
    }
    exr_attr(out, "compression", "compression", Bytes{0});  // NO_COMPRESSION
    {
        Bytes box;
        put32le(box, 0);
        put32le(box, 0);
        put32le(box, w - 1);
        put32le(box, h - 1);
        exr_attr(out, "dataWindow", "box2i", box);
        exr_attr(out, "displayWindow", "box2i", box);
    }
    exr_attr(out, "lineOrder", "lineOrder", Bytes{0});  // INCREASING_Y
    {
        Bytes f;
        putf32le(f, 1.0f);
        exr_attr(out, "pixelAspectRatio", "float", f);
    }
    {
        Bytes v;
        putf32le(v, 0.0f);
        putf32le(v, 0.0f);
        exr_attr(out, "screenWindowCenter", "v2f", v);
    }
    {
        Bytes f;
        putf32le(f, 1.0f);
        exr_attr(out, "screenWindowWidth", "float", f);
    }
    out.push_back(0);  // end of header
    const uint64_t line_bytes = 8 + 4 * 2 * (uint64_t)w;
    const uint64_t table = out.size();
    for (uint32_t y = 0; y < h; ++y) put64le(out, table + 8 * (uint64_t)h + y * line_bytes);
    for (uint32_t y = 0; y < h; ++y) {
        put32le(out, y);
        put32le(out, (uint32_t)(4 * 2 * (uint64_t)w));
        const float* row = rgba + (size_t)y * w * 4;
        for (int c : {3, 2, 1, 0})  // A, B, G, R
            for (uint32_t x = 0; x < w; ++x) put16le(out, to_half(row[4 * x + c]));
    }
    return LR_OK;
}

// Ward's RGBE: shared exponent of the largest component.
void rgbe(float r, float g, float b, unsigned char* e4) {
    float v = std::fmax(r, std::fmax(g, b));
    if (!(v >= 1e-32f) || !std::isfinite(v)) {  // black (also for NaN / infinite input)
        e4[0] = e4[1] = e4[2] = e4[3] = 0;
        return;
    }
    int e;
    const float m = std::frexp(v, &e);
    const float s = m * 256.0f / v;
    e4[0] = (unsigned char)std::fmax(0.0f, r * s);
    e4[1] = (unsigned char)std::fmax(0.0f, g * s);
    e4[2] = (unsigned char)std::fmax(0.0f, b * s);
    e4[3] = (unsigned char)(e + 128);
}

int encode_hdr(const float* rgba, uint32_t w, uint32_t h, Bytes& out) {
    char head[128];
    std::snprintf(head, sizeof(head), "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
    out.assign(head, head + std::strlen(head));
    const size_t base = out.size();
    out.resize(base + 4 * (size_t)w * h);
    for (size_t k = 0; k < (size_t)w * h; ++k) rgbe(rgba[4 * k], rgba[4 * k + 1], rgba[4 * k + 2], &out[base + 4 * k]);
    return LR_OK;
}

int encode_pfm(const float* rgba, uint32_t w, uint32_t h, Bytes& out) {
    char head[64];
    std::snprintf(head, sizeof(head), "PF\n%u %u\n-1.0\n", w, h);
    out.assign(head, head + std::strlen(head));
    for (uint32_t y = h; y-- > 0;)
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) putf32le(out, rgba[4 * ((size_t)y * w + x) + c]);
    return LR_OK;
}

int encode_tga(const float* rgba, uint32_t w, uint32_t h, Bytes& out) {
    if (w > 65535 || h > 65535) return lr::fail(LR_E_INVALID, "tga: image larger than 65535 pixels");
    out.assign(18, 0);
    out[2] = 2;  // uncompressed true colour
    out[12] = (unsigned char)w;
    out[13] = (unsigned char)(w >> 8);
    out[14] = (unsigned char)h;
    out[15] = (unsigned char)(h >> 8);
    out[16] = 32;    // bits per pixel
    out[17] = 0x28;  // 8 alpha bits, top-left origin
    out.reserve(18 + 4 * (size_t)w * h);
    for (size_t k = 0; k < (size_t)w * h; ++k) {
        out.push_back(q8(rgba[4 * k + 2]));
        out.push_back(q8(rgba[4 * k + 1]));
        out.push_back(q8(rgba[4 * k]));
        out.push_back(q8(rgba[4 * k + 3]));
    }
    return LR_OK;
}

std::string lower_ext(const char* path) {
    std::string p(path);
    const size_t dot = p.find_last_of('.');
    const size_t slash = p.find_last_of("/\\");
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = p.substr(dot);
    for (char& c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

}  // namespace

extern "C" int lr_image_write(const char* path, const float* rgba, uint32_t width, uint32_t height,
                              const lr_rgb_space* space) {
    if (!path || !rgba) return lr::fail(LR_E_INVALID, "NULL path or pixels");
    if (!width || !height) return lr::fail(LR_E_INVALID, "empty image");
    const lr_rgb_space srgb{{0.64f, 0.33f}, {0.30f, 0.60f}, {0.15f, 0.06f}, {0.3127f, 0.3290f}, 2.2f};
    const lr_rgb_space& sp = space ? *space : srgb;
    const std::string ext = lower_ext(path);
    Bytes out;
    int rc;
    if (ext == ".png") rc = encode_png(rgba, width, height, out);
    else if (ext == ".exr") rc = encode_exr(rgba, width, height, sp, out);
    else if (ext == ".hdr" || ext == ".rgbe" || ext == ".pic") rc = encode_hdr(rgba, width, height, out);
    else if (ext == ".pfm") rc = encode_pfm(rgba, width, height, out);
    else if (ext == ".tga" || ext == ".targa") rc = encode_tga(rgba, width, height, out);
    else return lr::fail(LR_E_UNSUPPORTED, "no image codec registered for the extension '" + ext + "'");
    if (rc) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return lr::fail(LR_E_INVALID, std::string("cannot open ") + path + " for writing");
    const size_t n = std::fwrite(out.data(), 1, out.size(), f);
    const int closed = std::fclose(f);
    if (n != out.size() || closed != 0) return lr::fail(LR_E_INVALID, std::string("short write to ") + path);
    return LR_OK;
}
