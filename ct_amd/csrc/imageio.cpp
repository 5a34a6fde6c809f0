// Projection image decoding (PGM, TIFF, PNG); see imageio.h.
#include "imageio.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

namespace ct {

namespace {

bool read_file(const std::string& path, std::vector<uint8_t>& buf, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    if (buf.empty()) {
        err = "empty file " + path;
        return false;
    }
    return true;
}

// ---------------------------------------------------------------------- PGM

bool decode_pgm(const std::vector<uint8_t>& b, Image& out, std::string& err) {
    size_t pos = 2;
    const bool ascii = b[1] == '2';
    auto skip_ws = [&]() {
        while (pos < b.size()) {
            if (b[pos] == '#') {
                while (pos < b.size() && b[pos] != '\n') ++pos;
            } else if (b[pos] == ' ' || b[pos] == '\t' || b[pos] == '\r' || b[pos] == '\n') {
                ++pos;
            } else {
                break;
            }
        }
    };
    auto read_uint = [&](long& v) -> bool {
        skip_ws();
        if (pos >= b.size() || b[pos] < '0' || b[pos] > '9') return false;
        v = 0;
        while (pos < b.size() && b[pos] >= '0' && b[pos] <= '9') v = v * 10 + (b[pos++] - '0');
        return true;
    };
    long w, h, maxval;
    if (!read_uint(w) || !read_uint(h) || !read_uint(maxval) || w <= 0 || h <= 0 || maxval <= 0 || maxval > 65535) {
        err = "malformed PGM header";
        return false;
    }
    out.width = (int)w;
    out.height = (int)h;
    out.depth = maxval < 256 ? PixelDepth::U8 : PixelDepth::U16;
    const size_t n = (size_t)w * h;
    out.data.assign(n * out.bytesPerPixel(), 0);
    if (ascii) {
        for (size_t i = 0; i < n; ++i) {
            long v;
            if (!read_uint(v)) {
                err = "truncated PGM data";
                return false;
            }
            if (out.depth == PixelDepth::U8) out.data[i] = (uint8_t)v;
            else reinterpret_cast<uint16_t*>(out.data.data())[i] = (uint16_t)v;
        }
        return true;
    }
    ++pos;  // exactly one whitespace byte after maxval
    const size_t need = n * out.bytesPerPixel();
    if (pos + need > b.size()) {
        err = "truncated PGM data";
        return false;
    }
    if (out.depth == PixelDepth::U8) {
        memcpy(out.data.data(), b.data() + pos, need);
    } else {
        uint16_t* d = reinterpret_cast<uint16_t*>(out.data.data());
        for (size_t i = 0; i < n; ++i) d[i] = (uint16_t)((b[pos + 2 * i] << 8) | b[pos + 2 * i + 1]);
    }
    return true;
}

// --------------------------------------------------------------------- TIFF

struct TiffReader {
    const std::vector<uint8_t>& b;
    bool le;
    uint16_t u16(size_t o) const {
        return le ? (uint16_t)(b[o] | (b[o + 1] << 8)) : (uint16_t)((b[o] << 8) | b[o + 1]);
    }
    uint32_t u32(size_t o) const {
        return le ? (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8) | ((uint32_t)b[o + 2] << 16) | ((uint32_t)b[o + 3] << 24)
                  : ((uint32_t)b[o] << 24) | ((uint32_t)b[o + 1] << 16) | ((uint32_t)b[o + 2] << 8) | (uint32_t)b[o + 3];
    }
};

bool packbits_decode(const uint8_t* src, size_t n, std::vector<uint8_t>& dst, size_t expect) {
    size_t i = 0;
    while (i < n && dst.size() < expect) {
        int8_t c = (int8_t)src[i++];
        if (c >= 0) {
            size_t cnt = (size_t)c + 1;
            if (i + cnt > n) return false;
            dst.insert(dst.end(), src + i, src + i + cnt);
            i += cnt;
        } else if (c != -128) {
            if (i >= n) return false;
            dst.insert(dst.end(), (size_t)(1 - c), src[i++]);
        }
    }
    return dst.size() >= expect;
}

// TIFF LZW (MSB-first codes, 9..12 bits, early change).
bool lzw_decode(const uint8_t* src, size_t n, std::vector<uint8_t>& dst, size_t expect) {
    std::vector<std::vector<uint8_t>> table;
    auto reset = [&]() {
        table.assign(258, {});
        for (int i = 0; i < 256; ++i) table[i] = {(uint8_t)i};
    };
    reset();
    size_t bitpos = 0;
    int width = 9;
    auto next = [&]() -> int {
        if (bitpos + width > n * 8) return 257;
        uint32_t v = 0;
        for (int k = 0; k < width; ++k) {
            size_t bp = bitpos + k;
            v = (v << 1) | ((src[bp >> 3] >> (7 - (bp & 7))) & 1);
        }
        bitpos += width;
        return (int)v;
    };
    int old = -1;
    while (dst.size() < expect) {
        int code = next();
        if (code == 257) break;
        if (code == 256) {
            reset();
            width = 9;
            code = next();
            if (code == 257) break;
            if (code >= (int)table.size()) return false;
            dst.insert(dst.end(), table[code].begin(), table[code].end());
            old = code;
            continue;
        }
        std::vector<uint8_t> entry;
        if (code < (int)table.size()) {
            entry = table[code];
            if (old >= 0) {
                std::vector<uint8_t> ne = table[old];
                ne.push_back(entry[0]);
                table.push_back(std::move(ne));
            }
        } else if (old >= 0 && code == (int)table.size()) {
            entry = table[old];
            entry.push_back(table[old][0]);
            table.push_back(entry);
        } else {
            return false;
        }
        dst.insert(dst.end(), entry.begin(), entry.end());
        old = code;
        const size_t ts = table.size();
        if (ts + 1 >= 4096) width = 12;
        else if (ts + 1 >= 2048) width = 12;
        else if (ts + 1 >= 1024) width = 11;
        else if (ts + 1 >= 512) width = 10;
    }
    return dst.size() >= expect;
}

bool inflate_all(const uint8_t* src, size_t n, std::vector<uint8_t>& dst, size_t expect) {
    dst.resize(expect);
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef*>(src);
    zs.avail_in = (uInt)n;
    zs.next_out = dst.data();
    zs.avail_out = (uInt)expect;
    int rc = inflate(&zs, Z_FINISH);
    size_t got = expect - zs.avail_out;
    inflateEnd(&zs);
    if (rc != Z_STREAM_END && rc != Z_OK && rc != Z_BUF_ERROR) return false;
    return got == expect;
}

bool decode_tiff(const std::vector<uint8_t>& b, Image& out, std::string& err) {
    TiffReader r{b, b[0] == 'I'};
    if (b.size() < 8 || r.u16(2) != 42) {
        err = "not a classic TIFF";
        return false;
    }
    size_t ifd = r.u32(4);
    if (ifd + 2 > b.size()) {
        err = "bad TIFF IFD offset";
        return false;
    }
    uint32_t width = 0, height = 0, bps = 1, comp = 1, spp = 1, rps = 0xffffffffu, fmt = 1, pred = 1, planar = 1;
    std::vector<uint32_t> offs, counts;
    const uint16_t nent = r.u16(ifd);
    auto values = [&](size_t e, std::vector<uint32_t>& v) {
        const uint16_t type = r.u16(e + 2);
        const uint32_t cnt = r.u32(e + 4);
        const size_t sz = type == 3 ? 2 : 4;
        size_t base = (sz * cnt <= 4) ? e + 8 : r.u32(e + 8);
        v.resize(cnt);
        for (uint32_t i = 0; i < cnt; ++i) {
            if (base + sz * (i + 1) > b.size()) { v.clear(); return; }
            v[i] = type == 3 ? r.u16(base + 2 * i) : r.u32(base + 4 * i);
        }
    };
    for (uint16_t i = 0; i < nent; ++i) {
        const size_t e = ifd + 2 + 12 * (size_t)i;
        if (e + 12 > b.size()) break;
        const uint16_t tag = r.u16(e);
        std::vector<uint32_t> v;
        values(e, v);
        if (v.empty()) continue;
        switch (tag) {
            case 256: width = v[0]; break;
            case 257: height = v[0]; break;
            case 258: bps = v[0]; break;
            case 259: comp = v[0]; break;
            case 273: offs = v; break;
            case 277: spp = v[0]; break;
            case 278: rps = v[0]; break;
            case 279: counts = v; break;
            case 284: planar = v[0]; break;
            case 317: pred = v[0]; break;
            case 339: fmt = v[0]; break;
            default: break;
        }
    }
    if (width == 0 || height == 0 || offs.empty() || offs.size() != counts.size()) {
        err = "unsupported TIFF layout (needs strips)";
        return false;
    }
    if (spp != 1 || planar != 1) {
        err = "only single-channel TIFF is supported";
        return false;
    }
    if (bps == 8 && fmt == 1) out.depth = PixelDepth::U8;
    else if (bps == 16 && fmt == 1) out.depth = PixelDepth::U16;
    else if (bps == 32 && fmt == 3) out.depth = PixelDepth::F32;
    else {
        err = "TIFF depth must be 8/16-bit unsigned or 32-bit float";
        return false;
    }
    out.width = (int)width;
    out.height = (int)height;
    const size_t bpp = out.bytesPerPixel();
    const size_t row = (size_t)width * bpp;
    if (rps > height) rps = height;
    std::vector<uint8_t> raw;
    raw.reserve(row * height);
    for (size_t s = 0; s < offs.size(); ++s) {
        const size_t rows = std::min<size_t>(rps, height - std::min<size_t>(height, s * (size_t)rps));
        const size_t expect = rows * row;
        if ((size_t)offs[s] + counts[s] > b.size()) {
            err = "TIFF strip outside the file";
            return false;
        }
        const uint8_t* src = b.data() + offs[s];
        std::vector<uint8_t> strip;
        bool ok = true;
        if (comp == 1) strip.assign(src, src + std::min<size_t>(counts[s], expect));
        else if (comp == 32773) ok = packbits_decode(src, counts[s], strip, expect);
        else if (comp == 5) ok = lzw_decode(src, counts[s], strip, expect);
        else if (comp == 8 || comp == 32946) ok = inflate_all(src, counts[s], strip, expect);
        else {
            err = "unsupported TIFF compression " + std::to_string(comp);
            return false;
        }
        if (!ok || strip.size() < expect) {
            err = "corrupt TIFF strip";
            return false;
        }
        strip.resize(expect);
        raw.insert(raw.end(), strip.begin(), strip.end());
    }
    if (raw.size() < row * height) {
        err = "truncated TIFF data";
        return false;
    }
    out.data.assign(row * height, 0);
    const bool swap = (!r.le) && bpp > 1;
    for (size_t i = 0; i < (size_t)width * height; ++i) {
        for (size_t k = 0; k < bpp; ++k) out.data[i * bpp + k] = raw[i * bpp + (swap ? bpp - 1 - k : k)];
    }
    if (pred == 2 && out.depth != PixelDepth::F32) {
        for (uint32_t y = 0; y < height; ++y) {
            if (out.depth == PixelDepth::U8) {
                uint8_t* p = out.data.data() + (size_t)y * width;
                for (uint32_t x = 1; x < width; ++x) p[x] = (uint8_t)(p[x] + p[x - 1]);
            } else {
                uint16_t* p = reinterpret_cast<uint16_t*>(out.data.data()) + (size_t)y * width;
                for (uint32_t x = 1; x < width; ++x) p[x] = (uint16_t)(p[x] + p[x - 1]);
            }
        }
    }
    return true;
}

// ---------------------------------------------------------------------- PNG

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

bool decode_png(const std::vector<uint8_t>& b, Image& out, std::string& err) {
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat;
    while (pos + 8 <= b.size()) {
        const uint32_t len = be32(&b[pos]);
        const char* type = reinterpret_cast<const char*>(&b[pos + 4]);
        if (pos + 12 + (size_t)len > b.size()) {
            err = "truncated PNG chunk";
            return false;
        }
        const uint8_t* d = &b[pos + 8];
        if (!memcmp(type, "IHDR", 4)) {
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        }
        pos += 12 + len;
    }
    if (w == 0 || h == 0 || (depth != 8 && depth != 16)) {
        err = "PNG must be 8- or 16-bit";
        return false;
    }
    if (interlace) {
        err = "interlaced PNG is not supported";
        return false;
    }
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 4: channels = 2; break;
        case 2: channels = 3; break;
        case 6: channels = 4; break;
        default: err = "palette PNG is not supported"; return false;
    }
    const size_t bps = depth / 8, bpp = bps * channels, stride = (size_t)w * bpp;
    std::vector<uint8_t> raw;
    if (!inflate_all(idat.data(), idat.size(), raw, (stride + 1) * h)) {
        err = "corrupt PNG data";
        return false;
    }
    std::vector<uint8_t> img(stride * h), prev(stride, 0);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t f = raw[y * (stride + 1)];
        const uint8_t* s = &raw[y * (stride + 1) + 1];
        uint8_t* o = &img[y * stride];
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= bpp ? o[x - bpp] : 0, up = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
            int v = s[x];
            switch (f) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += up; break;
                case 3: v += (a + up) / 2; break;
                case 4: {
                    const int p = a + up - c, pa = std::abs(p - a), pb = std::abs(p - up), pc = std::abs(p - c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? up : c);
                    break;
                }
                default: err = "bad PNG filter"; return false;
            }
            o[x] = (uint8_t)v;
        }
        memcpy(prev.data(), o, stride);
    }
    out.width = (int)w;
    out.height = (int)h;
    out.depth = depth == 8 ? PixelDepth::U8 : PixelDepth::U16;
    out.data.assign((size_t)w * h * bps, 0);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        double g;
        auto sample = [&](int ch) -> double {
            const uint8_t* p = &img[i * bpp + ch * bps];
            return bps == 1 ? p[0] : (double)((p[0] << 8) | p[1]);
        };
        if (channels <= 2) g = sample(0);
        else g = 0.299 * sample(0) + 0.587 * sample(1) + 0.114 * sample(2);
        const long v = std::lround(g);
        if (bps == 1) out.data[i] = (uint8_t)v;
        else reinterpret_cast<uint16_t*>(out.data.data())[i] = (uint16_t)v;
    }
    return true;
}

}  // namespace

bool readImage(const std::string& path, Image& out, std::string& err) {
    std::vector<uint8_t> b;
    if (!read_file(path, b, err)) return false;
    out = Image();
    if (b.size() >= 2 && b[0] == 'P' && (b[1] == '5' || b[1] == '2')) return decode_pgm(b, out, err);
    if (b.size() >= 4 && ((b[0] == 'I' && b[1] == 'I') || (b[0] == 'M' && b[1] == 'M'))) return decode_tiff(b, out, err);
    static const uint8_t png_sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    if (b.size() >= 8 && !memcmp(b.data(), png_sig, 8)) return decode_png(b, out, err);
    err = "unsupported image format: " + path;
    return false;
}

bool writePgm16(const std::string& path, const uint16_t* px, int w, int h, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    std::fprintf(f, "P5\n%d %d\n65535\n", w, h);
    std::vector<uint8_t> row(2 * (size_t)w);
    bool ok = true;
    for (int y = 0; y < h && ok; ++y) {
        for (int x = 0; x < w; ++x) {
            const uint16_t v = px[(size_t)y * w + x];
            row[2 * x] = (uint8_t)(v >> 8);
            row[2 * x + 1] = (uint8_t)(v & 0xff);
        }
        ok = std::fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) err = "write failed: " + path;
    return ok;
}

}  // namespace ct
