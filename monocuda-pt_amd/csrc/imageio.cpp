// imageio.cpp — the image files a scene reads, host code:
//   * binary PPM textures (loadPpmAsRgba8, src/path_tracer.cu:236-294):
//     "P6", then width, height and maxval as decimal integers separated by
//     whitespace, '#' starting a comment that runs to the end of its line,
//     maxval 255 only, exactly one separator byte after the maxval, then
//     width*height RGB triples, top row first; returned as RGBA8, alpha 255.
//   * PFM environment maps (loadPfmAsRgba32f, :335-368): "PF" (three
//     channels), width, height and scale read as by fscanf("%2s %d %d %f"),
//     one byte after the scale, little-endian FP32 only (a positive scale
//     means big-endian and is refused), rows stored bottom first; returned as
//     RGBA FP32, top row first, alpha 1.
//   * Radiance RGBE .hdr environment maps, which the reference decodes with
//     stb_image (github.com/nothings/stb at 9600522, stbi_loadf) and converts
//     with scripts/hdr2pfm.py:14-61: a "#?" signature, header lines up to an
//     empty line (a FORMAT line other than 32-bit_rle_rgbe is refused), the
//     resolution line "-Y h +X w" only, then per row either the new-style
//     run-length form (2, 2, width high byte, width low byte, then the four
//     byte planes as runs: a count above 128 repeats the next byte
//     count - 128 times, otherwise count literal bytes follow) when
//     8 <= width < 32768, or width flat RGBE quadruples; each texel is
//     byte * 2^(e - 136) for e > 0 and 0 for e = 0, alpha 1.
//   * OpenEXR environment maps, which the reference decodes with tinyexr
//     (github.com/syoyo/tinyexr at e48a52f, LoadEXR) into RGBA FP32. Read
//     here from the published OpenEXR file layout: single-part scanline
//     files; channels of type UINT, HALF or FLOAT; compression NONE, RLE,
//     ZIPS or ZIP (zlib, then the byte predictor and the two-half
//     interleave undone); pixels placed by their data-window row, so either
//     line order reads the same. R, G, B and A are taken by name (A = 1 when
//     absent); a file with one channel repeats it in R, G and B. Tiled,
//     deep and multi-part files and the PIZ, PXR24, B44 and DWA
//     compressions are refused.
// Every size is checked against the bytes the file holds before anything
// of that size is allocated.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <zlib.h>

#include "host_common.hpp"

namespace mcpt {

namespace {

// pixels of one image at most (4 channels each)
constexpr long long kMaxPixels = 1LL << 28;

struct File {
    std::FILE* f = nullptr;
    explicit File(const char* path) : f(std::fopen(path, "rb")) {}
    ~File() {
        if (f) std::fclose(f);
    }
    long long remaining() const {  // bytes from the current position to the end
        const long here = std::ftell(f);
        if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
        const long end = std::ftell(f);
        std::fseek(f, here, SEEK_SET);
        return end < 0 ? -1 : (long long)(end - here);
    }
};

// readNextPpmInteger (:236-256): skips whitespace and '#' comment lines,
// reads decimal digits and consumes the one separator after them
bool ppm_int(std::FILE* f, int& value) {
    if (std::feof(f) || std::ferror(f)) return false;
    int ch = std::fgetc(f);
    while (ch != EOF) {
        if (ch == '#') {
            while (ch != '\n' && ch != EOF) ch = std::fgetc(f);
        } else if (std::isspace(ch)) {
            ch = std::fgetc(f);
        } else {
            break;
        }
    }
    if (ch == EOF || !std::isdigit(ch)) return false;
    long long v = 0;
    while (ch != EOF && std::isdigit(ch)) {
        v = v * 10 + (ch - '0');
        if (v > (1LL << 30)) return false;  // no image has a side this long
        ch = std::fgetc(f);
    }
    value = (int)v;
    return true;
}

}  // namespace

int read_ppm(const char* path, int& w, int& h, std::vector<uint8_t>* rgba) {
    File file(path);
    std::FILE* f = file.f;
    if (!f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
    int maxv = 0;
    const bool ok = std::fgetc(f) == 'P' && std::fgetc(f) == '6' && ppm_int(f, w) && ppm_int(f, h) &&
                    ppm_int(f, maxv) && maxv == 255 && w > 0 && h > 0 && !std::feof(f) && !std::ferror(f);
    if (!ok) return set_error(MCPT_ERR_INVALID, "'%s' is not a binary PPM with maxval 255", path);
    const long long n = (long long)w * h;
    if (n > kMaxPixels) return set_error(MCPT_ERR_INVALID, "'%s': %d x %d pixels is too large", path, w, h);
    if (file.remaining() < 3 * n)
        return set_error(MCPT_ERR_INVALID, "'%s': the file holds fewer than %lld pixel bytes", path, 3 * n);
    if (!rgba) return MCPT_OK;
    std::vector<uint8_t> rgb((size_t)n * 3);
    if (std::fread(rgb.data(), 1, rgb.size(), f) != rgb.size())
        return set_error(MCPT_ERR_INVALID, "'%s': short read of the pixels", path);
    rgba->resize((size_t)n * 4);
    uint8_t* o = rgba->data();
    for (long long i = 0; i < n; i++) {
        o[4 * i] = rgb[3 * i];
        o[4 * i + 1] = rgb[3 * i + 1];
        o[4 * i + 2] = rgb[3 * i + 2];
        o[4 * i + 3] = 255;
    }
    return MCPT_OK;
}

int read_pfm(const char* path, int& w, int& h, std::vector<float>* rgba) {
    File file(path);
    std::FILE* f = file.f;
    if (!f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
    char magic[3] = {0, 0, 0};
    float scale = 0.f;
    if (std::fscanf(f, "%2s %d %d %f", magic, &w, &h, &scale) != 4 || std::strcmp(magic, "PF") != 0 || w <= 0 ||
        h <= 0)
        return set_error(MCPT_ERR_INVALID, "'%s' is not a three-channel PFM", path);
    if (std::fgetc(f) == EOF) return set_error(MCPT_ERR_INVALID, "'%s': no pixels after the header", path);
    if (scale > 0.f) return set_error(MCPT_ERR_INVALID, "'%s': big-endian PFM files are not supported", path);
    const long long n = (long long)w * h;
    if (n > kMaxPixels) return set_error(MCPT_ERR_INVALID, "'%s': %d x %d pixels is too large", path, w, h);
    if (file.remaining() < 12 * n)
        return set_error(MCPT_ERR_INVALID, "'%s': the file holds fewer than %lld pixel bytes", path, 12 * n);
    if (!rgba) return MCPT_OK;
    std::vector<float> rgb((size_t)n * 3);
    if (std::fread(rgb.data(), sizeof(float), rgb.size(), f) != rgb.size())
        return set_error(MCPT_ERR_INVALID, "'%s': short read of the pixels", path);
    rgba->resize((size_t)n * 4);
    for (int row = 0; row < h; row++) {
        const float* src = &rgb[(size_t)(h - 1 - row) * w * 3];  // bottom row first in the file
        float* dst = &(*rgba)[(size_t)row * w * 4];
        for (int x = 0; x < w; x++) {
            dst[4 * x] = src[3 * x];
            dst[4 * x + 1] = src[3 * x + 1];
            dst[4 * x + 2] = src[3 * x + 2];
            dst[4 * x + 3] = 1.0f;
        }
    }
    return MCPT_OK;
}

int read_hdr(const char* path, int& w, int& h, std::vector<float>* rgba) {
    File file(path);
    std::FILE* f = file.f;
    if (!f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
    const long long size = file.remaining();
    if (size < 0) return set_error(MCPT_ERR_INVALID, "'%s': cannot size the file", path);
    std::vector<uint8_t> data((size_t)size);
    if (size && std::fread(data.data(), 1, data.size(), f) != data.size())
        return set_error(MCPT_ERR_INVALID, "'%s': short read", path);
    size_t at = 0;
    auto line = [&](std::string& out) {  // the next '\n'-terminated line
        const void* nl = at < data.size() ? std::memchr(&data[at], '\n', data.size() - at) : nullptr;
        if (!nl) return false;
        const size_t end = (size_t)((const uint8_t*)nl - data.data());
        out.assign((const char*)&data[at], end - at);
        at = end + 1;
        return true;
    };
    if (data.size() < 2 || data[0] != '#' || data[1] != '?')
        return set_error(MCPT_ERR_INVALID, "'%s' is not a Radiance HDR file", path);
    std::string text;
    for (;;) {
        if (!line(text)) return set_error(MCPT_ERR_INVALID, "'%s': the header does not end", path);
        if (text.empty()) break;
        if (text.compare(0, 7, "FORMAT=") == 0 && text != "FORMAT=32-bit_rle_rgbe")
            return set_error(MCPT_ERR_INVALID, "'%s': %s is not supported", path, text.c_str());
    }
    if (!line(text)) return set_error(MCPT_ERR_INVALID, "'%s': no resolution line", path);
    char ya[8] = {0}, xa[8] = {0};
    long long hh = 0, ww = 0;
    char tail = 0;
    if (std::sscanf(text.c_str(), "%7s %lld %7s %lld %c", ya, &hh, xa, &ww, &tail) != 4 ||
        std::strcmp(ya, "-Y") != 0 || std::strcmp(xa, "+X") != 0)
        return set_error(MCPT_ERR_INVALID, "'%s': only the -Y h +X w orientation is supported", path);
    if (hh <= 0 || ww <= 0 || hh * ww > kMaxPixels)
        return set_error(MCPT_ERR_INVALID, "'%s': bad size %lld x %lld", path, ww, hh);
    w = (int)ww;
    h = (int)hh;
    if (!rgba) return MCPT_OK;
    std::vector<uint8_t> row((size_t)w * 4);
    rgba->resize((size_t)w * h * 4);
    auto bad = [&](int y) { return set_error(MCPT_ERR_INVALID, "'%s': bad pixel data in row %d", path, y); };
    for (int y = 0; y < h; y++) {
        const bool rle = w >= 8 && w < 32768 && at + 4 <= data.size() && data[at] == 2 && data[at + 1] == 2 &&
                         ((int)data[at + 2] << 8 | (int)data[at + 3]) == w;
        if (rle) {
            at += 4;
            for (int ch = 0; ch < 4; ch++) {
                for (int x = 0; x < w;) {
                    if (at >= data.size()) return bad(y);
                    int count = data[at++];
                    if (count > 128) {
                        count -= 128;
                        if (x + count > w || at >= data.size()) return bad(y);
                        const uint8_t v = data[at++];
                        for (int k = 0; k < count; k++) row[(size_t)4 * (x + k) + ch] = v;
                    } else {
                        if (count == 0 || x + count > w || at + count > data.size()) return bad(y);
                        for (int k = 0; k < count; k++) row[(size_t)4 * (x + k) + ch] = data[at + k];
                        at += count;
                    }
                    x += count;
                }
            }
        } else {
            if (at + row.size() > data.size()) return bad(y);
            std::memcpy(row.data(), &data[at], row.size());
            at += row.size();
        }
        float* dst = &(*rgba)[(size_t)y * w * 4];
        for (int x = 0; x < w; x++) {
            const uint8_t* q = &row[(size_t)4 * x];
            const float s = q[3] ? std::ldexp(1.0f, (int)q[3] - 136) : 0.0f;
            dst[4 * x] = (float)q[0] * s;
            dst[4 * x + 1] = (float)q[1] * s;
            dst[4 * x + 2] = (float)q[2] * s;
            dst[4 * x + 3] = 1.0f;
        }
    }
    return MCPT_OK;
}

namespace {

float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else {  // subnormal: m * 2^-24
            const float v = (float)m * (1.0f / 16777216.0f);
            uint32_t b;
            std::memcpy(&b, &v, 4);
            bits = sign | b;
        }
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (m << 13);
    } else {
        bits = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// undo the ZIP/RLE byte predictor, then interleave the two halves
void exr_unpredict(std::vector<uint8_t>& t) {
    for (size_t i = 1; i < t.size(); i++) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
    std::vector<uint8_t> out(t.size());
    const size_t half = (t.size() + 1) / 2;
    for (size_t i = 0; i < t.size(); i++) out[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
    t.swap(out);
}

}  // namespace

int read_exr(const char* path, int& w, int& h, std::vector<float>* rgba) {
    File file(path);
    if (!file.f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
    const long long size = file.remaining();
    if (size < 8) return set_error(MCPT_ERR_INVALID, "'%s' is not an OpenEXR file", path);
    std::vector<uint8_t> d((size_t)size);
    if (std::fread(d.data(), 1, d.size(), file.f) != d.size()) return set_error(MCPT_ERR_INVALID, "'%s': short read", path);
    auto u32 = [&](size_t at) {
        uint32_t v;
        std::memcpy(&v, &d[at], 4);
        return v;
    };
    if (u32(0) != 20000630u) return set_error(MCPT_ERR_INVALID, "'%s' is not an OpenEXR file", path);
    const uint32_t ver = u32(4);
    if ((ver & 0xff) != 2 || (ver & (0x200u | 0x800u | 0x1000u)))
        return set_error(MCPT_ERR_INVALID, "'%s': tiled, deep or multi-part OpenEXR files are not read", path);
    struct Chan {
        std::string name;
        int type, xs, ys;
    };
    std::vector<Chan> ch;
    int comp = -1, xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    size_t at = 8;
    auto cstr = [&](std::string& out) {
        const void* z = at < d.size() ? std::memchr(&d[at], 0, d.size() - at) : nullptr;
        if (!z) return false;
        const size_t e = (size_t)((const uint8_t*)z - d.data());
        out.assign((const char*)&d[at], e - at);
        at = e + 1;
        return true;
    };
    for (;;) {
        std::string name, type;
        if (!cstr(name)) return set_error(MCPT_ERR_INVALID, "'%s': the header does not end", path);
        if (name.empty()) break;
        if (!cstr(type) || at + 4 > d.size()) return set_error(MCPT_ERR_INVALID, "'%s': bad header", path);
        const uint32_t len = u32(at);
        at += 4;
        if (at + len > d.size()) return set_error(MCPT_ERR_INVALID, "'%s': bad header", path);
        const size_t end = at + len;
        if (name == "channels" && type == "chlist") {
            while (at < end && d[at] != 0) {
                Chan c;
                if (!cstr(c.name) || at + 16 > end) return set_error(MCPT_ERR_INVALID, "'%s': bad channel list", path);
                c.type = (int)u32(at);
                c.xs = (int)u32(at + 8);
                c.ys = (int)u32(at + 12);
                at += 16;
                if (c.type < 0 || c.type > 2 || c.xs != 1 || c.ys != 1)
                    return set_error(MCPT_ERR_INVALID, "'%s': channel %s is subsampled or of unknown type", path,
                                     c.name.c_str());
                ch.push_back(c);
            }
        } else if (name == "compression" && len >= 1) {
            comp = d[at];
        } else if (name == "dataWindow" && len >= 16) {
            xmin = (int)u32(at), ymin = (int)u32(at + 4), xmax = (int)u32(at + 8), ymax = (int)u32(at + 12);
        }
        at = end;
    }
    const long long W = (long long)xmax - xmin + 1, H = (long long)ymax - ymin + 1;
    if (ch.empty() || W < 1 || H < 1 || W * H > kMaxPixels)
        return set_error(MCPT_ERR_INVALID, "'%s': no channels or a bad data window", path);
    const int lines = comp == 0 || comp == 1 || comp == 2 ? 1 : comp == 3 ? 16 : 0;
    if (lines == 0) return set_error(MCPT_ERR_INVALID, "'%s': compression %d is not read (NONE, RLE, ZIPS, ZIP are)", path, comp);
    w = (int)W;
    h = (int)H;
    if (!rgba) return MCPT_OK;
    // channels are stored in name order; each line holds every channel in turn
    std::vector<int> order(ch.size());
    for (size_t i = 0; i < ch.size(); i++) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return ch[a].name < ch[b].name; });
    int slot[4] = {-1, -1, -1, -1};  // channel feeding R, G, B, A
    const char* want[4] = {"R", "G", "B", "A"};
    for (int k = 0; k < 4; k++)
        for (size_t i = 0; i < ch.size(); i++)
            if (ch[i].name == want[k]) slot[k] = (int)i;
    if (ch.size() == 1 && slot[0] < 0) slot[0] = slot[1] = slot[2] = 0;
    size_t line_bytes = 0;
    for (const Chan& c : ch) line_bytes += (size_t)W * (c.type == 1 ? 2 : 4);
    const long long chunks = (H + lines - 1) / lines;
    if (at + (size_t)chunks * 8 > d.size()) return set_error(MCPT_ERR_INVALID, "'%s': truncated offset table", path);
    rgba->assign((size_t)W * H * 4, 0.0f);
    for (long long p = 0; p < W * H; p++) (*rgba)[p * 4 + 3] = 1.0f;
    std::vector<uint8_t> buf;
    for (long long k = 0; k < chunks; k++) {
        uint64_t off;
        std::memcpy(&off, &d[at + k * 8], 8);
        // compare without adding to the file's value, which may be near 2^64
        if (off > d.size() || d.size() - off < 8)
            return set_error(MCPT_ERR_INVALID, "'%s': bad chunk offset", path);
        const int y0 = (int)u32(off);
        const uint32_t n = u32(off + 4);
        if (n > d.size() - off - 8) return set_error(MCPT_ERR_INVALID, "'%s': truncated chunk", path);
        const long long first = (long long)y0 - ymin;
        if (first < 0 || first >= H) return set_error(MCPT_ERR_INVALID, "'%s': chunk outside the data window", path);
        const long long nl = std::min<long long>(lines, H - first);
        const size_t raw = (size_t)nl * line_bytes;
        const uint8_t* src = &d[off + 8];
        if (n == raw) {  // stored as is (also the fallback of a block that would not shrink)
            buf.assign(src, src + raw);
        } else if (comp == 2 || comp == 3) {
            buf.resize(raw);
            uLongf got = (uLongf)raw;
            if (uncompress(buf.data(), &got, src, n) != Z_OK || got != raw)
                return set_error(MCPT_ERR_INVALID, "'%s': bad ZIP block", path);
            exr_unpredict(buf);
        } else if (comp == 1) {
            buf.clear();
            for (uint32_t i = 0; i < n;) {
                const int c = (int8_t)src[i++];
                if (c < 0) {
                    if (i + (uint32_t)(-c) > n) return set_error(MCPT_ERR_INVALID, "'%s': bad RLE block", path);
                    buf.insert(buf.end(), src + i, src + i - c);
                    i += (uint32_t)(-c);
                } else {
                    if (i >= n) return set_error(MCPT_ERR_INVALID, "'%s': bad RLE block", path);
                    buf.insert(buf.end(), (size_t)c + 1, src[i++]);
                }
                if (buf.size() > raw) return set_error(MCPT_ERR_INVALID, "'%s': bad RLE block", path);
            }
            if (buf.size() != raw) return set_error(MCPT_ERR_INVALID, "'%s': bad RLE block", path);
            exr_unpredict(buf);
        } else {
            return set_error(MCPT_ERR_INVALID, "'%s': chunk size %u, expected %zu", path, n, raw);
        }
        size_t q = 0;
        for (long long l = 0; l < nl; l++) {
            const long long row = first + l;
            for (int oi : order) {
                const Chan& c = ch[oi];
                const int sz = c.type == 1 ? 2 : 4;
                for (long long x = 0; x < W; x++, q += sz) {
                    float v;
                    if (c.type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, &buf[q], 2);
                        v = half_to_float(hv);
                    } else if (c.type == 2) {
                        std::memcpy(&v, &buf[q], 4);
                    } else {
                        uint32_t uv;
                        std::memcpy(&uv, &buf[q], 4);
                        v = (float)uv;
                    }
                    for (int k2 = 0; k2 < 4; k2++)
                        if (slot[k2] == oi) (*rgba)[((size_t)row * W + x) * 4 + k2] = v;
                }
            }
        }
    }
    return MCPT_OK;
}


// ---------------------------------------------------------------- PNG
// PNG textures, which the reference reads with stb_image
// (loadImageRgba8, src/image_io.cpp:30-39, the fallback of
// getOrLoadObjTexture at src/path_tracer.cu:2602). Read here from the PNG
// specification (ISO/IEC 15948): chunks with their CRC checked, IHDR first,
// IDAT concatenated and inflated, unknown critical chunks refused; every
// colour type and bit depth the format allows, filters None/Sub/Up/
// Average/Paeth, Adam7 interlacing; 16-bit samples keep their high byte
// and 1/2/4-bit grey is scaled to 0..255, as stb_image returns them; a tRNS
// chunk gives palette alpha, or alpha 0 to the one grey or RGB value it
// names; returned as RGBA8, top row first.
namespace {

const uint8_t kPngSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};

uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// reverses one row's filter in place; prev is the previous row of the same
// pass after its own reversal (zeros above a pass's first row)
bool png_unfilter(int type, uint8_t* row, const uint8_t* prev, size_t n, size_t bpp) {
    switch (type) {
        case 0:
            return true;
        case 1:
            for (size_t i = bpp; i < n; i++) row[i] = (uint8_t)(row[i] + row[i - bpp]);
            return true;
        case 2:
            for (size_t i = 0; i < n; i++) row[i] = (uint8_t)(row[i] + prev[i]);
            return true;
        case 3:
            for (size_t i = 0; i < n; i++) row[i] = (uint8_t)(row[i] + (((i >= bpp ? row[i - bpp] : 0) + prev[i]) >> 1));
            return true;
        case 4:
            for (size_t i = 0; i < n; i++) {
                const int a = i >= bpp ? row[i - bpp] : 0, c = i >= bpp ? prev[i - bpp] : 0;
                row[i] = (uint8_t)(row[i] + paeth(a, prev[i], c));
            }
            return true;
    }
    return false;
}

}  // namespace

int read_png(const char* path, int& w, int& h, std::vector<uint8_t>* rgba) {
    File file(path);
    if (!file.f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
    const long long size = file.remaining();
    if (size < 8 + 25) return set_error(MCPT_ERR_INVALID, "'%s' is not a PNG file", path);
    std::vector<uint8_t> d((size_t)size);
    if (std::fread(d.data(), 1, d.size(), file.f) != d.size()) return set_error(MCPT_ERR_INVALID, "'%s': short read", path);
    if (std::memcmp(d.data(), kPngSig, 8) != 0) return set_error(MCPT_ERR_INVALID, "'%s' is not a PNG file", path);
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> plte, trns, idat;
    bool ihdr = false, iend = false;
    size_t at = 8;
    while (!iend) {
        if (d.size() - at < 12) return set_error(MCPT_ERR_INVALID, "'%s': truncated PNG chunk", path);
        const uint32_t len = be32(&d[at]);
        if (len > 0x7fffffffu || (size_t)len > d.size() - at - 12)
            return set_error(MCPT_ERR_INVALID, "'%s': PNG chunk runs past the end of the file", path);
        const uint8_t* type = &d[at + 4];
        const uint8_t* data = &d[at + 8];
        const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4);
        if (crc != be32(data + len)) return set_error(MCPT_ERR_INVALID, "'%s': PNG chunk CRC mismatch", path);
        const std::string t((const char*)type, 4);
        if (!ihdr && t != "IHDR") return set_error(MCPT_ERR_INVALID, "'%s': PNG does not start with IHDR", path);
        if (t == "IHDR") {
            if (ihdr || len != 13) return set_error(MCPT_ERR_INVALID, "'%s': bad PNG IHDR", path);
            ihdr = true;
            const uint32_t pw = be32(data), ph = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            const bool ok_depth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                                  (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                                  ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
            if (pw == 0 || ph == 0 || pw > 0x7fffffffu || ph > 0x7fffffffu || !ok_depth || data[10] != 0 ||
                data[11] != 0 || interlace > 1)
                return set_error(MCPT_ERR_INVALID, "'%s': unsupported or invalid PNG header", path);
            if ((long long)pw * (long long)ph > kMaxPixels)
                return set_error(MCPT_ERR_INVALID, "'%s': PNG image too large", path);
            w = (int)pw;
            h = (int)ph;
            if (!rgba) return MCPT_OK;
        } else if (t == "PLTE") {
            if (len % 3 != 0 || len == 0 || len > 768) return set_error(MCPT_ERR_INVALID, "'%s': bad PNG palette", path);
            plte.assign(data, data + len);
        } else if (t == "tRNS") {
            trns.assign(data, data + len);
        } else if (t == "IDAT") {
            idat.insert(idat.end(), data, data + len);
        } else if (t == "IEND") {
            iend = true;
        } else if (!(type[0] & 0x20)) {
            return set_error(MCPT_ERR_INVALID, "'%s': unknown critical PNG chunk %s", path, t.c_str());
        }
        at += 12 + (size_t)len;
    }
    if (ctype == 3 && plte.empty()) return set_error(MCPT_ERR_INVALID, "'%s': palette PNG without PLTE", path);
    const int chans = ctype == 0 || ctype == 3 ? 1 : ctype == 2 ? 3 : ctype == 4 ? 2 : 4;
    const size_t bits = (size_t)chans * depth;           // per pixel
    const size_t bpp = std::max<size_t>(1, bits / 8);    // filter distance in bytes
    // the Adam7 passes (x0, y0, dx, dy), or the whole image
    static const int kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                     {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const int kWhole[1][4] = {{0, 0, 1, 1}};
    const int (*passes)[4] = interlace ? kAdam7 : kWhole;
    const int npass = interlace ? 7 : 1;
    size_t raw_bytes = 0;
    for (int p = 0; p < npass; p++) {
        const size_t pw = (size_t)(w > passes[p][0] ? (w - passes[p][0] + passes[p][2] - 1) / passes[p][2] : 0);
        const size_t ph = (size_t)(h > passes[p][1] ? (h - passes[p][1] + passes[p][3] - 1) / passes[p][3] : 0);
        if (pw && ph) raw_bytes += ph * (1 + (pw * bits + 7) / 8);
    }
    // deflate cannot expand its input more than 1032-fold: a longer image
    // cannot be in this file, so nothing of its size is allocated
    if (raw_bytes / 1032 > idat.size() + 64) return set_error(MCPT_ERR_INVALID, "'%s': PNG image data too short", path);
    std::vector<uint8_t> raw(raw_bytes);
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return set_error(MCPT_ERR_INVALID, "'%s': zlib init failed", path);
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if (zr != Z_STREAM_END || got != raw.size())
        return set_error(MCPT_ERR_INVALID, "'%s': PNG image data does not inflate to the image size", path);
    std::vector<uint8_t>& out = *rgba;
    out.assign((size_t)w * h * 4, 0);
    const int maxv = (1 << depth) - 1;
    size_t off = 0;
    for (int p = 0; p < npass; p++) {
        const int x0 = passes[p][0], y0 = passes[p][1], dx = passes[p][2], dy = passes[p][3];
        const size_t pw = (size_t)(w > x0 ? (w - x0 + dx - 1) / dx : 0);
        const size_t ph = (size_t)(h > y0 ? (h - y0 + dy - 1) / dy : 0);
        if (!pw || !ph) continue;
        const size_t rb = (pw * bits + 7) / 8;
        std::vector<uint8_t> zero(rb, 0);
        const uint8_t* prev = zero.data();
        for (size_t r = 0; r < ph; r++) {
            uint8_t* row = &raw[off + 1];
            if (!png_unfilter(raw[off], row, prev, rb, bpp))
                return set_error(MCPT_ERR_INVALID, "'%s': bad PNG filter type %d", path, raw[off]);
            for (size_t x = 0; x < pw; x++) {
                // the pixel's samples at full depth
                uint32_t v[4] = {0, 0, 0, 0};
                for (int c = 0; c < chans; c++) {
                    if (depth < 8) {
                        const size_t bit = x * depth;
                        v[c] = (uint32_t)(row[bit / 8] >> (8 - depth - (int)(bit % 8))) & (uint32_t)maxv;
                    } else if (depth == 8) {
                        v[c] = row[x * chans + c];
                    } else {
                        v[c] = (uint32_t)row[(x * chans + c) * 2] << 8 | row[(x * chans + c) * 2 + 1];
                    }
                }
                auto to8 = [&](uint32_t s) -> uint8_t {
                    return depth == 16 ? (uint8_t)(s >> 8) : depth == 8 ? (uint8_t)s : (uint8_t)(s * (255 / maxv));
                };
                uint8_t px[4];
                switch (ctype) {
                    case 0: {
                        px[0] = px[1] = px[2] = to8(v[0]);
                        const bool key = trns.size() >= 2 && v[0] == (uint32_t)(trns[0] << 8 | trns[1]);
                        px[3] = key ? 0 : 255;
                        break;
                    }
                    case 2: {
                        for (int c = 0; c < 3; c++) px[c] = to8(v[c]);
                        const bool key = trns.size() >= 6 && v[0] == (uint32_t)(trns[0] << 8 | trns[1]) &&
                                         v[1] == (uint32_t)(trns[2] << 8 | trns[3]) &&
                                         v[2] == (uint32_t)(trns[4] << 8 | trns[5]);
                        px[3] = key ? 0 : 255;
                        break;
                    }
                    case 3: {
                        if ((size_t)v[0] * 3 + 2 >= plte.size())
                            return set_error(MCPT_ERR_INVALID, "'%s': PNG palette index %u out of range", path, v[0]);
                        for (int c = 0; c < 3; c++) px[c] = plte[v[0] * 3 + c];
                        px[3] = v[0] < trns.size() ? trns[v[0]] : 255;
                        break;
                    }
                    case 4:
                        px[0] = px[1] = px[2] = to8(v[0]);
                        px[3] = to8(v[1]);
                        break;
                    default:
                        for (int c = 0; c < 4; c++) px[c] = to8(v[c]);
                        break;
                }
                const size_t X = (size_t)x0 + x * dx, Y = (size_t)y0 + r * dy;
                std::memcpy(&out[(Y * w + X) * 4], px, 4);
            }
            prev = row;
            off += 1 + rb;
        }
    }
    return MCPT_OK;
}

// ---------------------------------------------------------------- TGA
// Truevision TGA textures (another of stb_image's formats, the reference's
// texture fallback): image types 1/9 (colour-mapped, 8-bit index), 2/10
// (true colour, 15/16/24/32 bits) and 3/11 (grey, 8 bits), raw or
// run-length packets (a header byte c: (c & 127) + 1 pixels, one pixel
// repeated when c & 128, else that many literal pixels); BGR(A) order;
// 15/16-bit colour expanded as v * 255 / 31 with alpha 255; rows bottom
// first unless descriptor bit 5 is set; returned as RGBA8, top row first.
int read_tga(const char* path, int& w, int& h, std::vector<uint8_t>* rgba) {
    File file(path);
    if (!file.f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
    const long long size = file.remaining();
    if (size < 18) return set_error(MCPT_ERR_INVALID, "'%s' is not a TGA file", path);
    std::vector<uint8_t> d((size_t)size);
    if (std::fread(d.data(), 1, d.size(), file.f) != d.size()) return set_error(MCPT_ERR_INVALID, "'%s': short read", path);
    auto le16 = [&](size_t at) { return (int)d[at] | (int)d[at + 1] << 8; };
    const int idlen = d[0], cmtype = d[1], itype = d[2], cmfirst = le16(3), cmlen = le16(5), cmbits = d[7];
    const int tw = le16(12), th = le16(14), bits = d[16], desc = d[17];
    const int base = itype & 7;
    const bool rle = (itype & 8) != 0;
    const bool ok = (itype & ~11) == 0 && base >= 1 && base <= 3 && tw > 0 && th > 0 &&
                    ((base == 1 && cmtype == 1 && bits == 8 && cmlen > 0 &&
                      (cmbits == 15 || cmbits == 16 || cmbits == 24 || cmbits == 32)) ||
                     (base == 2 && (bits == 15 || bits == 16 || bits == 24 || bits == 32)) || (base == 3 && bits == 8));
    if (!ok) return set_error(MCPT_ERR_INVALID, "'%s': unsupported or invalid TGA header", path);
    w = tw;
    h = th;
    if (!rgba) return MCPT_OK;
    size_t at = 18 + (size_t)idlen;
    const size_t cbp = (size_t)(cmbits + 7) / 8, bp = (size_t)(bits + 7) / 8;
    const size_t cm_bytes = cmtype == 1 ? (size_t)cmlen * cbp : 0;
    if (at + cm_bytes > d.size()) return set_error(MCPT_ERR_INVALID, "'%s': TGA file too short", path);
    const uint8_t* cmap = &d[at];
    at += cm_bytes;
    const size_t npx = (size_t)tw * th, left = d.size() - at;
    // every pixel needs its bytes, or a run packet's share of them
    if (rle ? npx > 128 * (left / (1 + bp) + 1) : npx * bp > left)
        return set_error(MCPT_ERR_INVALID, "'%s': TGA file too short", path);
    auto colour = [](const uint8_t* p, int nbits, uint8_t* o) {
        if (nbits == 15 || nbits == 16) {
            const int v = (int)p[0] | (int)p[1] << 8;
            o[0] = (uint8_t)(((v >> 10) & 31) * 255 / 31);
            o[1] = (uint8_t)(((v >> 5) & 31) * 255 / 31);
            o[2] = (uint8_t)((v & 31) * 255 / 31);
            o[3] = 255;
        } else {
            o[0] = p[2];
            o[1] = p[1];
            o[2] = p[0];
            o[3] = nbits == 32 ? p[3] : 255;
        }
    };
    std::vector<uint8_t>& out = *rgba;
    out.assign(npx * 4, 0);
    uint8_t px[4];
    auto decode = [&](const uint8_t* p) -> bool {
        if (base == 3) {
            px[0] = px[1] = px[2] = p[0];
            px[3] = 255;
        } else if (base == 1) {
            const int k = (int)p[0] - cmfirst;
            if (k < 0 || k >= cmlen) return false;
            colour(cmap + (size_t)k * cbp, cmbits, px);
        } else {
            colour(p, bits, px);
        }
        return true;
    };
    const bool top_first = (desc & 0x20) != 0;
    auto put = [&](size_t i) {
        const size_t r = i / tw, c = i % tw;
        const size_t row = top_first ? r : (size_t)th - 1 - r;
        std::memcpy(&out[(row * tw + c) * 4], px, 4);
    };
    size_t i = 0;
    while (i < npx) {
        size_t count = 1;
        bool repeat = false;
        if (rle) {
            if (at >= d.size()) return set_error(MCPT_ERR_INVALID, "'%s': TGA run data ends early", path);
            const uint8_t c = d[at++];
            count = (size_t)(c & 127) + 1;
            repeat = (c & 128) != 0;
            if (i + count > npx) return set_error(MCPT_ERR_INVALID, "'%s': TGA run passes the last pixel", path);
        }
        for (size_t k = 0; k < count; k++) {
            if (k == 0 || !repeat) {
                if (at + bp > d.size()) return set_error(MCPT_ERR_INVALID, "'%s': TGA pixel data ends early", path);
                if (!decode(&d[at])) return set_error(MCPT_ERR_INVALID, "'%s': TGA colour index out of range", path);
                at += bp;
            }
            put(i++);
        }
    }
    return MCPT_OK;
}

int read_image8(const char* path, int& w, int& h, std::vector<uint8_t>* rgba) {
    const size_t n = std::strlen(path);
    if (n >= 4) {
        std::string ext(path + n - 4);
        for (char& ch : ext) ch = (char)std::tolower((unsigned char)ch);
        if (ext == ".tga") return read_tga(path, w, h, rgba);
    }
    uint8_t sig[8] = {0};
    {
        File file(path);
        if (!file.f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
        if (std::fread(sig, 1, 8, file.f) != 8) return read_ppm(path, w, h, rgba);
    }
    if (sig[0] == 0xFF && sig[1] == 0xD8 && sig[2] == 0xFF) return read_jpeg(path, w, h, rgba);
    return std::memcmp(sig, kPngSig, 8) == 0 ? read_png(path, w, h, rgba) : read_ppm(path, w, h, rgba);
}
}  // namespace mcpt
