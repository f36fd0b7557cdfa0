// jpeg.cpp — baseline JPEG textures (host code). The reference reads JPEG
// image textures through stb_image (github.com/nothings/stb at 9600522,
// stbi_load, src/image_io.cpp:1-40); this reader is written from ITU-T T.81
// (ISO/IEC 10918-1) and the JFIF colour convention:
//   * SOF0 / SOF1 frames (sequential DCT, Huffman coding), 8-bit samples,
//     one (grey) or three (YCbCr) components, sampling factors 1..4 per
//     component, interleaved and non-interleaved scans, restart intervals
//     (DRI + RSTn); progressive, lossless, hierarchical and arithmetic-coded
//     files and 12-bit samples are refused;
//   * Huffman decoding by the canonical code lengths (T.81 Annex C / F.2.2.3),
//     DC prediction per component, coefficients dequantized in zig-zag order;
//   * the inverse DCT of T.81 A.3.3 in floating point, level shift +128,
//     rounded to nearest and clamped to 0..255;
//   * chroma planes with half the luma resolution brought up with the
//     triangle filter (weights 3/4 and 1/4 on the nearest and the next sample
//     per axis, chroma sited between luma samples), other ratios by
//     replication; YCbCr -> RGB by the JFIF equations.
// Returned as RGBA8, top row first, alpha 255. Every length is checked
// against the bytes the file holds before it is used.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_common.hpp"

namespace mcpt {
namespace {

constexpr long long kMaxPixels = 1LL << 28;

// natural (row-major) index of the k-th coefficient in zig-zag order
constexpr int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    bool defined = false;
    int mincode[17], maxcode[18], valptr[17];
    uint8_t vals[256];
    // canonical codes from the 16 code-length counts (T.81 C.1, C.2, F.16)
    bool build(const uint8_t* counts, const uint8_t* v, int nv) {
        int code = 0, k = 0;
        for (int l = 1; l <= 16; l++) {
            valptr[l] = k;
            mincode[l] = code;
            code += counts[l - 1];
            k += counts[l - 1];
            maxcode[l] = counts[l - 1] ? code - 1 : -1;
            if (code > (1 << l)) return false;  // more codes than l bits hold
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        if (k != nv || nv > 256) return false;
        std::memcpy(vals, v, nv);
        defined = true;
        return true;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;    // Huffman tables of the current scan
    int bw = 0, bh = 0;    // blocks per row / column in the padded plane
    std::vector<uint8_t> plane;  // bw*8 x bh*8 samples
    int pred = 0;
};

struct Reader {
    const std::vector<uint8_t>& d;
    size_t pos;
    uint32_t acc = 0;
    int cnt = 0;
    bool past_data = false;  // a marker ended the entropy-coded segment
    explicit Reader(const std::vector<uint8_t>& data, size_t at) : d(data), pos(at) {}
    void fill() {
        while (cnt <= 24) {
            uint32_t b = 0;
            if (!past_data && pos < d.size()) {
                b = d[pos];
                if (b == 0xFF) {
                    const uint32_t nxt = pos + 1 < d.size() ? d[pos + 1] : 0xD9;
                    if (nxt == 0x00) {
                        pos += 2;  // stuffed zero byte
                    } else {
                        past_data = true;  // a marker: feed zeros from here
                        b = 0;
                    }
                } else {
                    pos++;
                }
            }
            acc |= b << (24 - cnt);
            cnt += 8;
        }
    }
    int bit() {
        if (cnt < 1) fill();
        const int b = (int)(acc >> 31);
        acc <<= 1;
        cnt--;
        return b;
    }
    int bits(int n) {
        if (n == 0) return 0;
        if (cnt < n) fill();
        const int v = (int)(acc >> (32 - n));
        acc <<= n;
        cnt -= n;
        return v;
    }
    // the next RSTn marker: drop the rest of the current byte, then expect it
    bool restart(int expect) {
        acc = 0;
        cnt = 0;
        past_data = false;
        // the interval's data is decoded: skip its padding and fill bytes up to the marker
        while (pos + 1 < d.size() && !(d[pos] == 0xFF && d[pos + 1] != 0x00 && d[pos + 1] != 0xFF)) pos++;
        if (pos + 1 >= d.size() || d[pos] != 0xFF || d[pos + 1] != 0xD0 + expect) return false;
        pos += 2;
        return true;
    }
};

int decode_symbol(Reader& r, const Huffman& t) {
    int code = 0;
    for (int l = 1; l <= 16; l++) {
        code = (code << 1) | r.bit();
        if (t.maxcode[l] >= 0 && code <= t.maxcode[l]) return t.vals[t.valptr[l] + code - t.mincode[l]];
    }
    return -1;  // no code of 16 bits or fewer
}

// EXTEND (T.81 F.12): s magnitude bits to a signed value
int extend(int v, int s) { return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v; }

struct Idct {
    float c[8][8];  // c[x][u] = C(u)/2 cos((2x+1) u pi / 16)
    Idct() {
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++)
                c[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) / 2.0 * std::cos((2 * x + 1) * u * M_PI / 16.0));
    }
    // T.81 A.3.3 on dequantized coefficients (natural order), level shift,
    // rounding and clamping into dst (stride in samples)
    void run(const int* coef, uint8_t* dst, int stride) const {
        float tmp[64];
        for (int v = 0; v < 8; v++)  // rows of coefficients -> spatial columns
            for (int x = 0; x < 8; x++) {
                float s = 0.0f;
                for (int u = 0; u < 8; u++) s += c[x][u] * (float)coef[v * 8 + u];
                tmp[v * 8 + x] = s;
            }
        for (int x = 0; x < 8; x++)
            for (int y = 0; y < 8; y++) {
                float s = 0.0f;
                for (int v = 0; v < 8; v++) s += c[y][v] * tmp[v * 8 + x];
                const int q = (int)std::lround(s + 128.0f);
                dst[y * stride + x] = (uint8_t)std::min(255, std::max(0, q));
            }
    }
};

struct Decoder {
    const std::vector<uint8_t>& d;
    const char* path;
    uint16_t qt[4][64];
    bool qdef[4] = {false, false, false, false};
    Huffman dc[4], ac[4];
    std::vector<Component> comp;
    int W = 0, H = 0, hmax = 1, vmax = 1, restart = 0;
    bool frame = false;
    Idct idct;
    explicit Decoder(const std::vector<uint8_t>& data, const char* p) : d(data), path(p) {}

    int err(const char* what) const { return set_error(MCPT_ERR_INVALID, "'%s': %s", path, what); }

    int dqt(size_t at, size_t len) {
        size_t i = 0;
        while (i < len) {
            const int pq = d[at + i] >> 4, tq = d[at + i] & 15;
            if (tq > 3) return err("bad quantization table id");
            const size_t n = pq ? 128 : 64;
            if (len - i - 1 < n) return err("truncated quantization table");
            for (int k = 0; k < 64; k++)
                qt[tq][k] = pq ? (uint16_t)((d[at + i + 1 + 2 * k] << 8) | d[at + i + 2 + 2 * k]) : d[at + i + 1 + k];
            qdef[tq] = true;
            i += 1 + n;
        }
        return MCPT_OK;
    }

    int dht(size_t at, size_t len) {
        size_t i = 0;
        while (i < len) {
            if (len - i < 17) return err("truncated Huffman table");
            const int tc = d[at + i] >> 4, th = d[at + i] & 15;
            if (tc > 1 || th > 3) return err("bad Huffman table class or id");
            int nv = 0;
            for (int l = 0; l < 16; l++) nv += d[at + i + 1 + l];
            if (len - i - 17 < (size_t)nv) return err("truncated Huffman table");
            Huffman& t = tc ? ac[th] : dc[th];
            if (!t.build(&d[at + i + 1], &d[at + i + 17], nv)) return err("bad Huffman code lengths");
            i += 17 + nv;
        }
        return MCPT_OK;
    }

    int sof(size_t at, size_t len) {
        if (frame) return err("more than one frame");
        if (len < 6) return err("truncated frame header");
        if (d[at] != 8) return err("only 8-bit samples are read");
        H = (d[at + 1] << 8) | d[at + 2];
        W = (d[at + 3] << 8) | d[at + 4];
        const int nc = d[at + 5];
        if (W < 1 || H < 1) return err("zero image size (DNL is not read)");
        if ((long long)W * H > kMaxPixels) return err("image too large");
        if (nc != 1 && nc != 3) return err("only grey and three-component images are read");
        if (len < 6 + 3 * (size_t)nc) return err("truncated frame header");
        comp.resize(nc);
        for (int i = 0; i < nc; i++) {
            Component& c = comp[i];
            c.id = d[at + 6 + 3 * i];
            c.h = d[at + 7 + 3 * i] >> 4;
            c.v = d[at + 7 + 3 * i] & 15;
            c.tq = d[at + 8 + 3 * i];
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return err("bad component parameters");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        const int mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
        for (Component& c : comp) {
            if (hmax % c.h || vmax % c.v) return err("sampling factors that do not divide the largest");
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
        }
        frame = true;
        return MCPT_OK;
    }

    int block(Reader& r, Component& c, int bx, int by) {
        const Huffman& hd = dc[c.td];
        const Huffman& ha = ac[c.ta];
        const uint16_t* q = qt[c.tq];
        int coef[64] = {0};
        const int t = decode_symbol(r, hd);
        if (t < 0 || t > 11) return err("bad DC code");
        const int diff = t ? extend(r.bits(t), t) : 0;
        c.pred += diff;
        if (c.pred < -(1 << 20) || c.pred > (1 << 20)) return err("DC value out of range");
        coef[0] = c.pred * q[0];
        for (int k = 1; k < 64;) {
            const int rs = decode_symbol(r, ha);
            if (rs < 0) return err("bad AC code");
            const int run = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (run != 15) break;  // end of block
                k += 16;
                continue;
            }
            k += run;
            if (k > 63) return err("AC coefficients past the block");
            coef[kZigzag[k]] = extend(r.bits(s), s) * q[k];
            k++;
        }
        const int stride = c.bw * 8;
        idct.run(coef, &c.plane[((size_t)by * 8) * stride + (size_t)bx * 8], stride);
        return MCPT_OK;
    }

    int sos(size_t at, size_t len, size_t& end) {
        if (!frame) return err("scan before the frame header");
        if (len < 1) return err("truncated scan header");
        const int ns = d[at];
        if (ns < 1 || ns > 4 || len < 4 + 2 * (size_t)ns) return err("bad scan header");
        std::vector<Component*> sc;
        for (int i = 0; i < ns; i++) {
            const int cid = d[at + 1 + 2 * i];
            Component* found = nullptr;
            for (Component& c : comp)
                if (c.id == cid) found = &c;
            if (!found) return err("scan names an unknown component");
            found->td = d[at + 2 + 2 * i] >> 4;
            found->ta = d[at + 2 + 2 * i] & 15;
            if (found->td > 3 || found->ta > 3 || !dc[found->td].defined || !ac[found->ta].defined)
                return err("scan uses an undefined Huffman table");
            if (!qdef[found->tq]) return err("component uses an undefined quantization table");
            found->pred = 0;
            sc.push_back(found);
        }
        const int ss = d[at + 1 + 2 * ns], se = d[at + 2 + 2 * ns], a = d[at + 3 + 2 * ns];
        if (ss != 0 || se != 63 || a != 0) return err("not a sequential scan");
        Reader r(d, at + len);
        int units_x, units_y;
        if (ns == 1) {  // non-interleaved: the component's own blocks in raster order
            const Component& c = *sc[0];
            units_x = (int)(((long long)W * c.h + 8 * hmax - 1) / (8 * hmax));
            units_y = (int)(((long long)H * c.v + 8 * vmax - 1) / (8 * vmax));
        } else {
            units_x = (W + 8 * hmax - 1) / (8 * hmax);
            units_y = (H + 8 * vmax - 1) / (8 * vmax);
        }
        const long long units = (long long)units_x * units_y;
        int rst = 0;
        for (long long u = 0; u < units; u++) {
            if (restart && u > 0 && u % restart == 0) {
                if (!r.restart(rst)) return err("missing restart marker");
                rst = (rst + 1) & 7;
                for (Component* c : sc) c->pred = 0;
            }
            const int ux = (int)(u % units_x), uy = (int)(u / units_x);
            if (ns == 1) {
                MCPT_TRY(block(r, *sc[0], ux, uy));
            } else {
                for (Component* c : sc)
                    for (int v = 0; v < c->v; v++)
                        for (int h = 0; h < c->h; h++) MCPT_TRY(block(r, *c, ux * c->h + h, uy * c->v + v));
            }
        }
        // continue after the entropy-coded data: the next marker
        size_t p = r.pos;
        while (p + 1 < d.size() && !(d[p] == 0xFF && d[p + 1] != 0x00 && !(d[p + 1] >= 0xD0 && d[p + 1] <= 0xD7)))
            p++;
        end = p;
        return MCPT_OK;
    }

    // sample of component c at full-resolution pixel (x, y)
    void upsample(const Component& c, std::vector<uint8_t>& out) const {
        const int stride = c.bw * 8;
        out.resize((size_t)W * H);
        const int cw = (int)(((long long)W * c.h + hmax - 1) / hmax), ch = (int)(((long long)H * c.v + vmax - 1) / vmax);
        auto at = [&](int x, int y) {
            x = std::min(std::max(x, 0), cw - 1);
            y = std::min(std::max(y, 0), ch - 1);
            return (int)c.plane[(size_t)y * stride + x];
        };
        const int fx = hmax / c.h, fy = vmax / c.v;
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                int val;
                if (fx == 1 && fy == 1) {
                    val = at(x, y);
                } else if (fx <= 2 && fy <= 2) {
                    // triangle filter per halved axis: 3/4 of the nearest
                    // sample, 1/4 of the next one on the output's side
                    const int sx = x / fx, sy = y / fy;
                    const int nx = fx == 2 ? sx + ((x & 1) ? 1 : -1) : sx;
                    const int ny = fy == 2 ? sy + ((y & 1) ? 1 : -1) : sy;
                    const int wx = fx == 2 ? 3 : 4, wy = fy == 2 ? 3 : 4;
                    const int sum = wy * (wx * at(sx, sy) + (4 - wx) * at(nx, sy)) +
                                    (4 - wy) * (wx * at(sx, ny) + (4 - wx) * at(nx, ny));
                    val = (sum + 8) >> 4;
                } else {
                    val = at(x / fx, y / fy);
                }
                out[(size_t)y * W + x] = (uint8_t)val;
            }
    }

    int run(std::vector<uint8_t>* rgba) {
        if (d.size() < 4 || d[0] != 0xFF || d[1] != 0xD8) return err("not a JPEG file");
        size_t p = 2;
        while (true) {
            while (p < d.size() && d[p] == 0xFF && p + 1 < d.size() && d[p + 1] == 0xFF) p++;  // fill bytes
            if (p + 1 >= d.size() || d[p] != 0xFF) return err("truncated file or bad marker");
            const int m = d[p + 1];
            p += 2;
            if (m == 0xD9) break;  // EOI
            if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01) continue;
            if (p + 2 > d.size()) return err("truncated marker segment");
            const size_t len = ((size_t)d[p] << 8) | d[p + 1];
            if (len < 2 || len > d.size() - p) return err("truncated marker segment");
            const size_t at = p + 2, n = len - 2;
            if (m == 0xC0 || m == 0xC1) {
                MCPT_TRY(sof(at, n));
                if (!rgba) return MCPT_OK;  // the size is all a header read needs
            } else if ((m >= 0xC2 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                return err("progressive, lossless, hierarchical or arithmetic-coded JPEG is not read");
            } else if (m == 0xC4) {
                MCPT_TRY(dht(at, n));
            } else if (m == 0xDB) {
                MCPT_TRY(dqt(at, n));
            } else if (m == 0xDD) {
                if (n < 2) return err("truncated restart interval");
                restart = (d[at] << 8) | d[at + 1];
            } else if (m == 0xDA) {
                size_t end = 0;
                MCPT_TRY(sos(at, n, end));
                p = end;
                continue;
            }
            p += len;
        }
        if (!frame) return err("no frame header");
        std::vector<uint8_t> planes[3];
        for (size_t i = 0; i < comp.size(); i++) upsample(comp[i], planes[i]);
        rgba->assign((size_t)W * H * 4, 255);
        for (size_t i = 0; i < (size_t)W * H; i++) {
            uint8_t* o = &(*rgba)[4 * i];
            if (comp.size() == 1) {
                o[0] = o[1] = o[2] = planes[0][i];
                continue;
            }
            // JFIF: R = Y + 1.402 Cr', G = Y - 0.344136 Cb' - 0.714136 Cr', B = Y + 1.772 Cb'
            const float y = planes[0][i], cb = planes[1][i] - 128.0f, cr = planes[2][i] - 128.0f;
            const float rgb[3] = {y + 1.402f * cr, y - 0.344136f * cb - 0.714136f * cr, y + 1.772f * cb};
            for (int k = 0; k < 3; k++) o[k] = (uint8_t)std::min(255.0f, std::max(0.0f, std::floor(rgb[k] + 0.5f)));
        }
        return MCPT_OK;
    }
};

}  // namespace

int read_jpeg(const char* path, int& w, int& h, std::vector<uint8_t>* rgba) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(MCPT_ERR_INVALID, "cannot open '%s'", path);
    std::vector<uint8_t> d;
    uint8_t buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) d.insert(d.end(), buf, buf + n);
    std::fclose(f);
    Decoder dec(d, path);
    MCPT_TRY(dec.run(rgba));
    if (!dec.frame) return set_error(MCPT_ERR_INVALID, "'%s': no frame header", path);
    w = dec.W;
    h = dec.H;
    return MCPT_OK;
}

}  // namespace mcpt
