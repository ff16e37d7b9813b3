// PNG files of the reference's harness (src/main.cpp:126-140,59-67):
// decoding to 8-bit RGBA exactly as libpng 1.6's simplified read API
// delivers PNG_FORMAT_RGBA, and the float conversions of
// includes/utils.hpp, on zlib only (no libpng headers are needed).
//
// Decoding covers every PNG colour type and bit depth, Adam7 interlace,
// tRNS, sBIT, gAMA and sRGB.  The conversion to 8-bit RGBA follows the
// published behaviour of the simplified API for an RGBA output without a
// background:
//   - gray is replicated to RGB; 1/2/4-bit gray is expanded by replication
//     of its bits (v * 255 / (2^d - 1)); palette entries give RGB and the
//     tRNS table gives alpha; a tRNS colour key gives alpha 0, else 255;
//   - the output is sRGB (libpng's screen gamma 2.2).  The file gamma is
//     the sRGB chunk's (0.45455), else gAMA's, else the default: sRGB for
//     8-bit data and linear (1.0) for 16-bit data;
//   - 8-bit samples pass through unless the correction 1/(file gamma * 2.2)
//     is significant (outside 1 +/- 0.05); then each becomes
//     floor(255 * (v/255)^correction + 0.5);
//   - 16-bit samples go through libpng's 16-to-8 gamma table when the
//     correction file gamma * 2.2 is significant: the table is indexed by
//     the input's top 16 - shift bits (shift = max(5, 16 - sBIT), at most
//     8) and steps from output i to i + 1 at the input that the exponent
//     maps the midpoint i*257 + 128 back to (build_16to8 below); otherwise
//     they are scaled by (v*255 + 32895) >> 16.  16-bit alpha is always
//     scaled that way.
// Gamma is applied to colour samples only, never to alpha (libpng's
// PNG_ALPHA_PNG mode).  tests/test_png.py pins every rule against the
// system libpng where it is installed.
//
// Encoding writes 8-bit RGBA, non-interlaced, one IDAT.

#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pfs.h"
#include "pfs_internal.h"

namespace pfs {
namespace {

constexpr uint8_t kSignature[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
constexpr int64_t kMaxPixels = (int64_t)1 << 28;  // 256 Mpx: 1 GiB of RGBA8

uint32_t be32(const uint8_t *p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

bool read_file(const char *path, std::vector<uint8_t> *out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out->insert(out->end(), buf, buf + n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

struct Header {
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    int channels() const {
        switch (ctype) {
            case 0: return 1;  // gray
            case 2: return 3;  // RGB
            case 3: return 1;  // palette index
            case 4: return 2;  // gray + alpha
            case 6: return 4;  // RGBA
        }
        return 0;
    }
};

struct Decoded {
    Header hd;
    std::vector<uint8_t> plte;         // 3 bytes per entry
    std::vector<uint8_t> trns_alpha;   // palette images
    bool has_key = false;              // tRNS colour key (gray / RGB)
    uint16_t key[3] = {0, 0, 0};
    bool has_gama = false, has_srgb = false;
    uint32_t gama = 0;
    int sbit = 0;                      // max significant bits of the colour channels, 0: none
    std::vector<uint8_t> idat;
};

int fail(std::string *err, const std::string &msg) {
    if (err) *err = msg;
    return PFS_ERR_INVALID;
}

bool valid_depth(int ctype, int depth) {
    switch (ctype) {
        case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
        case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
        case 2: case 4: case 6: return depth == 8 || depth == 16;
    }
    return false;
}

int parse(const std::vector<uint8_t> &f, Decoded *d, std::string *err) {
    if (f.size() < 8 || std::memcmp(f.data(), kSignature, 8) != 0)
        return fail(err, "not a PNG file (bad signature)");
    size_t pos = 8;
    bool seen_ihdr = false, seen_iend = false;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        if (len > 0x7fffffffu || pos + 12 + (size_t)len > f.size())
            return fail(err, "truncated PNG chunk");
        const uint8_t *type = &f[pos + 4], *data = &f[pos + 8];
        const uint32_t crc = be32(&f[pos + 8 + len]);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc)
            return fail(err, std::string("CRC error in chunk ") + std::string((const char *)type, 4));
        const std::string t((const char *)type, 4);
        if (!seen_ihdr && t != "IHDR") return fail(err, "IHDR is not the first chunk");
        if (t == "IHDR") {
            if (len != 13 || seen_ihdr) return fail(err, "bad IHDR");
            seen_ihdr = true;
            d->hd.w = be32(data);
            d->hd.h = be32(data + 4);
            d->hd.depth = data[8];
            d->hd.ctype = data[9];
            d->hd.interlace = data[12];
            if (d->hd.w == 0 || d->hd.h == 0 || d->hd.w > 0x7fffffffu || d->hd.h > 0x7fffffffu)
                return fail(err, "bad image size");
            if ((int64_t)d->hd.w * d->hd.h > kMaxPixels) return fail(err, "image too large");
            if (!valid_depth(d->hd.ctype, d->hd.depth)) return fail(err, "bad colour type / depth");
            if (data[10] != 0 || data[11] != 0 || d->hd.interlace > 1)
                return fail(err, "unknown compression, filter or interlace method");
        } else if (t == "PLTE") {
            if (len % 3 != 0 || len == 0 || len > 768) return fail(err, "bad PLTE");
            d->plte.assign(data, data + len);
        } else if (t == "tRNS") {
            if (d->hd.ctype == 3) {
                d->trns_alpha.assign(data, data + len);
            } else if (d->hd.ctype == 0 && len == 2) {
                d->has_key = true;
                d->key[0] = (uint16_t)((data[0] << 8) | data[1]);
            } else if (d->hd.ctype == 2 && len == 6) {
                d->has_key = true;
                for (int c = 0; c < 3; ++c) d->key[c] = (uint16_t)((data[2 * c] << 8) | data[2 * c + 1]);
            }  // other tRNS forms are invalid; libpng ignores them
        } else if (t == "gAMA") {
            if (len == 4 && be32(data) != 0) {
                d->has_gama = true;
                d->gama = be32(data);
            }
        } else if (t == "sRGB") {
            d->has_srgb = true;
        } else if (t == "sBIT") {
            int m = 0;
            const int nc = d->hd.ctype == 3 ? 3 : (d->hd.ctype & 2 ? 3 : 1);
            for (uint32_t k = 0; k < len && (int)k < nc; ++k) m = data[k] > m ? data[k] : m;
            d->sbit = m;
        } else if (t == "IDAT") {
            d->idat.insert(d->idat.end(), data, data + len);
        } else if (t == "IEND") {
            seen_iend = true;
            break;
        } else if (!(type[0] & 0x20)) {
            return fail(err, "unknown critical chunk " + t);
        }
        pos += 12 + (size_t)len;
    }
    if (!seen_ihdr || !seen_iend) return fail(err, "missing IHDR or IEND");
    if (d->hd.ctype == 3 && d->plte.empty()) return fail(err, "palette image without PLTE");
    if (d->idat.empty()) return fail(err, "no image data");
    return PFS_OK;
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint8_t)((pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c));
}

// Reverses the per-row filters of a w x h (sub)image in place; raw holds h
// rows of 1 + rowbytes bytes.  Returns false on an unknown filter type.
bool unfilter(uint8_t *raw, size_t rowbytes, uint32_t h, int bpp) {
    std::vector<uint8_t> zero(rowbytes, 0);
    const uint8_t *prev = zero.data();
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t *row = raw + (size_t)y * (rowbytes + 1);
        const int ft = row[0];
        uint8_t *r = row + 1;
        for (size_t i = 0; i < rowbytes; ++i) {
            const int a = i >= (size_t)bpp ? r[i - bpp] : 0, b = prev[i];
            const int c = i >= (size_t)bpp ? prev[i - bpp] : 0;
            switch (ft) {
                case 0: break;
                case 1: r[i] = (uint8_t)(r[i] + a); break;
                case 2: r[i] = (uint8_t)(r[i] + b); break;
                case 3: r[i] = (uint8_t)(r[i] + ((a + b) >> 1)); break;
                case 4: r[i] = (uint8_t)(r[i] + paeth(a, b, c)); break;
                default: return false;
            }
        }
        prev = r;
    }
    return true;
}

// Samples (one uint16 per channel, full resolution) of the decoded image.
int samples(const Decoded &d, std::vector<uint16_t> *out, std::string *err) {
    const Header &hd = d.hd;
    const int ch = hd.channels(), depth = hd.depth;
    const int bits = ch * depth, bpp = bits >= 8 ? bits / 8 : 1;
    struct Pass {
        int x0, y0, dx, dy;
    };
    static const Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const Pass kWhole = {0, 0, 1, 1};
    const Pass *passes = hd.interlace ? kAdam7 : &kWhole;
    const int npass = hd.interlace ? 7 : 1;
    size_t total = 0;
    for (int p = 0; p < npass; ++p) {
        const uint32_t pw = hd.w > (uint32_t)passes[p].x0
                                ? (hd.w - passes[p].x0 + passes[p].dx - 1) / passes[p].dx : 0;
        const uint32_t ph = hd.h > (uint32_t)passes[p].y0
                                ? (hd.h - passes[p].y0 + passes[p].dy - 1) / passes[p].dy : 0;
        if (pw && ph) total += (size_t)ph * (1 + ((size_t)pw * bits + 7) / 8);
    }
    std::vector<uint8_t> raw(total);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) return fail(err, "zlib init failed");
    zs.next_in = const_cast<Bytef *>(d.idat.data());
    zs.avail_in = (uInt)d.idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR) || got != total)
        return fail(err, "corrupt or short image data");
    out->assign((size_t)hd.w * hd.h * ch, 0);
    size_t off = 0;
    for (int p = 0; p < npass; ++p) {
        const Pass &ps = passes[p];
        const uint32_t pw = hd.w > (uint32_t)ps.x0 ? (hd.w - ps.x0 + ps.dx - 1) / ps.dx : 0;
        const uint32_t ph = hd.h > (uint32_t)ps.y0 ? (hd.h - ps.y0 + ps.dy - 1) / ps.dy : 0;
        if (!pw || !ph) continue;
        const size_t rowbytes = ((size_t)pw * bits + 7) / 8;
        uint8_t *base = raw.data() + off;
        if (!unfilter(base, rowbytes, ph, bpp)) return fail(err, "unknown row filter");
        for (uint32_t y = 0; y < ph; ++y) {
            const uint8_t *r = base + (size_t)y * (rowbytes + 1) + 1;
            const size_t oy = (size_t)(ps.y0 + y * ps.dy);
            for (uint32_t x = 0; x < pw; ++x) {
                const size_t ox = (size_t)(ps.x0 + x * ps.dx);
                uint16_t *dst = out->data() + (oy * hd.w + ox) * ch;
                for (int c = 0; c < ch; ++c) {
                    uint16_t v;
                    if (depth == 16) {
                        const uint8_t *s = r + ((size_t)x * ch + c) * 2;
                        v = (uint16_t)((s[0] << 8) | s[1]);
                    } else if (depth == 8) {
                        v = r[(size_t)x * ch + c];
                    } else {  // 1, 2, 4 bits, one channel, MSB first
                        const size_t bit = (size_t)x * depth;
                        v = (uint16_t)((r[bit / 8] >> (8 - depth - bit % 8)) & ((1 << depth) - 1));
                    }
                    dst[c] = v;
                }
            }
        }
        off += (size_t)ph * (rowbytes + 1);
    }
    return PFS_OK;
}

// libpng's fixed-point gamma products (1.0 = 100000), rounded as libpng does.
int64_t fp_product(int64_t a, int64_t b) { return (int64_t)std::floor((double)a * 1e-5 * b + 0.5); }
int64_t fp_reciprocal(int64_t a, int64_t b) { return (int64_t)std::floor(1e15 / a / b + 0.5); }
bool significant(int64_t g) { return g < 100000 - 5000 || g > 100000 + 5000; }

// libpng's 16-to-8 gamma table: indexed by the input's top (16 - shift)
// bits, each entry the 8-bit output whose boundaries (outputs i + 1/2, as
// 16-bit i*257 + 128) map back through the exponent to the input range
// [0, 2^(16-shift) - 1].
std::vector<uint8_t> build_16to8(int shift, int64_t gamma_fp) {
    const uint32_t n = 1u << (16 - shift), max = n - 1;
    std::vector<uint8_t> t(n);
    const double g = (double)gamma_fp * 1e-5;
    uint32_t last = 0;
    for (uint32_t i = 0; i < 255; ++i) {
        const double out = (double)(i * 257u + 128u);
        uint32_t bound = (uint32_t)std::floor(65535.0 * std::pow(out / 65535.0, g) + 0.5);
        bound = (uint32_t)(((uint64_t)bound * max + 32768u) / 65535u + 1u);
        for (; last < bound && last < n; ++last) t[last] = (uint8_t)i;
    }
    for (; last < n; ++last) t[last] = 255;
    return t;
}

int to_rgba8(const Decoded &d, const std::vector<uint16_t> &s, std::vector<uint8_t> *out,
             std::string *err) {
    const Header &hd = d.hd;
    const size_t n = (size_t)hd.w * hd.h;
    const int ch = hd.channels();
    const bool wide = hd.depth == 16;
    // File gamma (sRGB chunk first), default sRGB for 8-bit data and
    // linear for 16-bit data; output (screen) gamma 2.2.
    const int64_t file_gamma = d.has_srgb ? 45455 : d.has_gama ? (int64_t)d.gama
                                                                  : (wide ? 100000 : 45455);
    const int64_t screen = 220000;
    std::vector<uint8_t> lut8, lut16;
    int shift = 0;
    if (!wide) {
        const int64_t g = fp_reciprocal(file_gamma, screen);
        if (significant(g)) {
            lut8.resize(256);
            for (int i = 0; i < 256; ++i)
                lut8[i] = (uint8_t)std::floor(255.0 * std::pow(i / 255.0, (double)g * 1e-5) + 0.5);
        }
    } else {
        const int64_t g = fp_product(file_gamma, screen);
        if (significant(g)) {
            shift = (d.sbit > 0 && d.sbit < 16) ? 16 - d.sbit : 0;
            if (shift < 5) shift = 5;
            if (shift > 8) shift = 8;
            lut16 = build_16to8(shift, g);
        }
    }
    auto colour = [&](uint16_t v) -> uint8_t {
        if (wide) return lut16.empty() ? (uint8_t)(((uint32_t)v * 255u + 32895u) >> 16)
                                       : lut16[v >> shift];
        return lut8.empty() ? (uint8_t)v : lut8[v];
    };
    auto alpha16 = [](uint16_t v) { return (uint8_t)(((uint32_t)v * 255u + 32895u) >> 16); };
    // Low bit depths expand to 8 bits by replication (gray only; palette
    // samples are indices).
    const int depth = hd.depth;
    auto expand = [&](uint16_t v) -> uint16_t {
        if (depth >= 8) return v;
        return (uint16_t)(v * (255 / ((1 << depth) - 1)));
    };
    out->resize(n * 4);
    uint8_t *o = out->data();
    const size_t npal = d.plte.size() / 3;
    for (size_t i = 0; i < n; ++i, o += 4) {
        const uint16_t *p = s.data() + i * ch;
        switch (hd.ctype) {
            case 0: {
                const uint8_t g = colour(expand(p[0]));
                o[0] = o[1] = o[2] = g;
                o[3] = (d.has_key && p[0] == d.key[0]) ? 0 : 255;
                break;
            }
            case 2:
                for (int c = 0; c < 3; ++c) o[c] = colour(p[c]);
                o[3] = (d.has_key && p[0] == d.key[0] && p[1] == d.key[1] && p[2] == d.key[2])
                           ? 0 : 255;
                break;
            case 3: {
                if (p[0] >= npal) return fail(err, "palette index out of range");
                for (int c = 0; c < 3; ++c) o[c] = colour(d.plte[3 * p[0] + c]);
                o[3] = p[0] < d.trns_alpha.size() ? d.trns_alpha[p[0]] : 255;
                break;
            }
            case 4: {
                const uint8_t g = colour(p[0]);
                o[0] = o[1] = o[2] = g;
                o[3] = wide ? alpha16(p[1]) : (uint8_t)p[1];
                break;
            }
            case 6:
                for (int c = 0; c < 3; ++c) o[c] = colour(p[c]);
                o[3] = wide ? alpha16(p[3]) : (uint8_t)p[3];
                break;
        }
    }
    return PFS_OK;
}

int decode(const char *path, Header *hd, std::vector<uint8_t> *rgba, std::string *err) {
    std::vector<uint8_t> f;
    if (!path || !read_file(path, &f)) return fail(err, std::string("cannot read ") + (path ? path : "(null)"));
    Decoded d;
    int rc = parse(f, &d, err);
    if (rc != PFS_OK) return rc;
    *hd = d.hd;
    if (!rgba) return PFS_OK;
    std::vector<uint16_t> s;
    rc = samples(d, &s, err);
    if (rc != PFS_OK) return rc;
    return to_rgba8(d, s, rgba, err);
}

void put32(std::vector<uint8_t> *v, uint32_t x) {
    for (int k = 3; k >= 0; --k) v->push_back((uint8_t)(x >> (8 * k)));
}

void put_chunk(std::vector<uint8_t> *f, const char *type, const uint8_t *data, size_t len) {
    put32(f, (uint32_t)len);
    const size_t start = f->size();
    f->insert(f->end(), type, type + 4);
    if (len) f->insert(f->end(), data, data + len);
    put32(f, (uint32_t)crc32(crc32(0L, Z_NULL, 0), f->data() + start, (uInt)(len + 4)));
}

int encode(const char *path, const uint8_t *rgba, int w, int h, std::string *err) {
    if (!path || !rgba || w < 1 || h < 1 || (int64_t)w * h > kMaxPixels)
        return fail(err, "bad arguments to the PNG writer");
    const size_t rowbytes = (size_t)w * 4;
    // Filter each row with Sub (the difference to the pixel on its left).
    std::vector<uint8_t> raw((size_t)h * (rowbytes + 1));
    for (int y = 0; y < h; ++y) {
        uint8_t *r = raw.data() + (size_t)y * (rowbytes + 1);
        const uint8_t *s = rgba + (size_t)y * rowbytes;
        r[0] = 1;
        for (size_t i = 0; i < rowbytes; ++i) r[1 + i] = (uint8_t)(s[i] - (i >= 4 ? s[i - 4] : 0));
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return fail(err, "zlib compression failed");
    std::vector<uint8_t> f(kSignature, kSignature + 8);
    uint8_t ihdr[13];
    for (int k = 0; k < 4; ++k) {
        ihdr[k] = (uint8_t)((uint32_t)w >> (24 - 8 * k));
        ihdr[4 + k] = (uint8_t)((uint32_t)h >> (24 - 8 * k));
    }
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 6;   // RGBA
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    put_chunk(&f, "IHDR", ihdr, 13);
    put_chunk(&f, "IDAT", z.data(), zlen);
    put_chunk(&f, "IEND", nullptr, 0);
    FILE *o = std::fopen(path, "wb");
    if (!o) return fail(err, std::string("cannot write ") + path);
    const bool ok = std::fwrite(f.data(), 1, f.size(), o) == f.size();
    if (std::fclose(o) != 0 || !ok) return fail(err, std::string("cannot write ") + path);
    return PFS_OK;
}

// (png_byte)(x * 255.0) of includes/utils.hpp:129 as x86-64 evaluates it:
// the double is truncated to a 32-bit integer (out of range and NaN give
// INT32_MIN) and its low byte is stored.  In range [0, 1] this is the
// truncation of x * 255.
uint8_t float_to_byte(float x) {
    const double d = (double)x * 255.0;
    const int32_t t = (d > -2147483649.0 && d < 2147483648.0) ? (int32_t)d : INT32_MIN;
    return (uint8_t)(uint32_t)t;
}

}  // namespace
}  // namespace pfs

using namespace pfs;

extern "C" {

int pfs_png_size(const char *path, int *width, int *height) {
    if (!width || !height) return set_error(PFS_ERR_INVALID, "width/height NULL");
    Header hd;
    std::string err;
    const int rc = decode(path, &hd, nullptr, &err);
    if (rc != PFS_OK) return set_error(rc, err);
    *width = (int)hd.w;
    *height = (int)hd.h;
    return PFS_OK;
}

int pfs_png_read_rgba8(const char *path, uint8_t *out, size_t n_bytes) {
    Header hd;
    std::vector<uint8_t> rgba;
    std::string err;
    const int rc = decode(path, &hd, &rgba, &err);
    if (rc != PFS_OK) return set_error(rc, err);
    if (!out || n_bytes < rgba.size()) return set_error(PFS_ERR_INVALID, "output buffer too small");
    std::memcpy(out, rgba.data(), rgba.size());
    return PFS_OK;
}

int pfs_png_read(const char *path, float *out, size_t n_floats) {
    Header hd;
    std::vector<uint8_t> rgba;
    std::string err;
    const int rc = decode(path, &hd, &rgba, &err);
    if (rc != PFS_OK) return set_error(rc, err);
    if (!out || n_floats < rgba.size()) return set_error(PFS_ERR_INVALID, "output buffer too small");
    for (size_t i = 0; i < rgba.size(); ++i) out[i] = (float)((double)(float)rgba[i] / 255.0);
    return PFS_OK;
}

int pfs_png_write_rgba8(const char *path, const uint8_t *rgba, int width, int height) {
    std::string err;
    const int rc = encode(path, rgba, width, height, &err);
    return rc == PFS_OK ? PFS_OK : set_error(rc, err);
}

int pfs_png_write(const char *path, const float *data, int width, int height) {
    if (!data || width < 1 || height < 1 || (int64_t)width * height > kMaxPixels)
        return set_error(PFS_ERR_INVALID, "bad arguments to the PNG writer");
    std::vector<uint8_t> rgba((size_t)width * height * 4);
    for (size_t i = 0; i < rgba.size(); ++i) rgba[i] = float_to_byte(data[i]);
    return pfs_png_write_rgba8(path, rgba.data(), width, height);
}

}  // extern "C"
