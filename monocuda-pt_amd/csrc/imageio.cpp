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
// Every size is checked against the bytes the file holds before anything
// of that size is allocated.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

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

}  // namespace mcpt
