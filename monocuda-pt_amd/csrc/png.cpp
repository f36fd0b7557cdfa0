// png.cpp — PNG file of RGB8 pixels (saveImageRgb8Png, src/image_io.cpp:62-65,
// which hands them to a third-party encoder). PNG is lossless, so any valid
// encoding of the same pixels is the same image: this one writes the rows
// with filter 0 inside zlib "stored" deflate blocks, with the CRC-32 of each
// chunk and the Adler-32 of the zlib stream as the format defines them.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "host_common.hpp"
#include "scene.hpp"

namespace mcpt {

namespace {

uint32_t crc_table[256];
bool crc_ready = false;

void crc_init() {
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
    crc_ready = true;
}

uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c;
}

void put32(std::vector<uint8_t>& o, uint32_t v) {
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data) {
    put32(o, (uint32_t)data.size());
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put32(o, crc32(o.data() + at, 4 + data.size()) ^ 0xFFFFFFFFu);
}

}  // namespace

int write_png(const char* path, int w, int h, const uint8_t* rgb8) {
    if (!path || !rgb8 || w < 1 || h < 1) return set_error(MCPT_ERR_INVALID, "bad PNG arguments");
    if (!crc_ready) crc_init();
    // raw scanlines: filter byte 0, then the row
    const size_t row = (size_t)w * 3;
    std::vector<uint8_t> raw;
    raw.reserve((row + 1) * h);
    for (int y = 0; y < h; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb8 + (size_t)y * row, rgb8 + (size_t)(y + 1) * row);
    }
    // zlib: header, stored blocks of at most 65535 bytes, Adler-32
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t at = 0; at < raw.size() || at == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - at);
        const bool last = at + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)len);
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len);
        z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + at, raw.begin() + at + len);
        at += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB, deflate, filter 0, no interlace
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(MCPT_ERR_INVALID, "cannot write '%s'", path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return set_error(MCPT_ERR_INVALID, "write to '%s' failed", path);
    return MCPT_OK;
}

}  // namespace mcpt
