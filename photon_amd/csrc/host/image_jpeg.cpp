// Baseline JPEG decoding (ITU-T T.81 sequential Huffman, 8-bit samples) for glTF textures: the
// reference decodes them with stb_image (src/scene/image.cpp:23-37). Supported: SOF0/SOF1, 1 (grey) or 3
// (YCbCr, or RGB with an Adobe transform of 0) components with any sampling factors, interleaved and
// single-component scans, restart intervals, 8- and 16-bit quantisation tables. Progressive, lossless,
// arithmetic-coded, 12-bit and CMYK files fail with a message.
//
// The inverse DCT is the separable float transform of the definition (T.81 A.3.3); chroma planes sampled
// at half resolution are upsampled with the triangular ("fancy") filter the common decoders use, so
// results agree with them within a few code values per channel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "image_io.hpp"

namespace phh {

namespace {

constexpr int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct JpegError {
    std::string msg;
};

[[noreturn]] void fail(const std::string& m) { throw JpegError{"JPEG: " + m}; }

// Canonical Huffman table (T.81 C): codes by length, decoded bit by bit against the per-length ranges.
struct Huffman {
    int maxcode[18];  // largest code of each length, -1 if none
    int valptr[17];   // index into vals of the first code of each length
    int mincode[17];
    std::vector<uint8_t> vals;
    bool present = false;

    void build(const uint8_t* counts, const uint8_t* symbols, int total) {
        vals.assign(symbols, symbols + total);
        int code = 0, k = 0;
        for (int len = 1; len <= 16; len++) {
            valptr[len] = k;
            mincode[len] = code;
            code += counts[len - 1];
            k += counts[len - 1];
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        present = true;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;
    int bw = 0, bh = 0;  // blocks across and down (padded to whole MCUs)
    std::vector<uint8_t> plane;  // bw*8 x bh*8 samples
    int dc_pred = 0;
};

class BitReader {
public:
    BitReader(const uint8_t* d, size_t n, size_t pos) : d_(d), n_(n), pos_(pos) {}

    int bit() {
        if (bits_ == 0) fill();
        bits_--;
        return int((acc_ >> bits_) & 1u);
    }
    int bits(int n) {
        int v = 0;
        for (int i = 0; i < n; i++) v = (v << 1) | bit();
        return v;
    }
    // Skips to the next restart marker (after the current byte-aligned position).
    void restart() {
        bits_ = 0;
        marker_ = false;
        while (pos_ + 1 < n_ && !(d_[pos_] == 0xFF && d_[pos_ + 1] >= 0xD0 && d_[pos_ + 1] <= 0xD7)) pos_++;
        if (pos_ + 1 >= n_) fail("missing restart marker");
        pos_ += 2;
    }
    size_t pos() const { return pos_; }

private:
    const uint8_t* d_;
    size_t n_, pos_;
    uint32_t acc_ = 0;
    int bits_ = 0;
    bool marker_ = false;

    void fill() {
        // Past a marker (or the end) the stream reads as 1-bits, as decoders conventionally do.
        uint8_t b = 0xFF;
        if (!marker_ && pos_ < n_) {
            b = d_[pos_];
            if (b == 0xFF) {
                const uint8_t nx = pos_ + 1 < n_ ? d_[pos_ + 1] : 0xD9;
                if (nx == 0x00) {
                    pos_ += 2;
                } else {
                    marker_ = true;  // a marker: leave it for the caller
                }
            } else {
                pos_++;
            }
        }
        acc_ = (acc_ << 8) | b;
        bits_ = 8;
    }
};

int decode_symbol(BitReader& br, const Huffman& h) {
    int code = br.bit();
    int len = 1;
    while (len <= 16 && code > h.maxcode[len]) {
        code = (code << 1) | br.bit();
        len++;
    }
    if (len > 16) fail("bad Huffman code");
    const size_t idx = size_t(h.valptr[len] + code - h.mincode[len]);
    if (idx >= h.vals.size()) fail("bad Huffman code");
    return h.vals[idx];
}

int extend(int v, int t) { return t == 0 ? 0 : (v < (1 << (t - 1)) ? v - (1 << t) + 1 : v); }

struct Idct {
    float c[8][8];  // c[x][u] = C(u)/2 cos((2x+1) u pi / 16)
    Idct() {
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++) c[x][u] = float((u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * M_PI / 16.0));
    }
    void run(const int* coef, uint8_t* out, int stride) const {
        float tmp[8][8];
        for (int v = 0; v < 8; v++)
            for (int x = 0; x < 8; x++) {
                float s = 0;
                for (int u = 0; u < 8; u++) s += c[x][u] * float(coef[v * 8 + u]);
                tmp[v][x] = s;
            }
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) {
                float s = 0;
                for (int v = 0; v < 8; v++) s += c[y][v] * tmp[v][x];
                const int p = int(std::lround(s + 128.0f));
                out[y * stride + x] = uint8_t(p < 0 ? 0 : p > 255 ? 255 : p);
            }
    }
};

uint8_t clamp8(int v) { return uint8_t(v < 0 ? 0 : v > 255 ? 255 : v); }

// Full-resolution plane of a component sampled at (hs, vs) of the maximum factors (hmax, vmax).
std::vector<uint8_t> upsample(const Component& c, int hmax, int vmax, int W, int H) {
    const int cw = c.bw * 8;
    std::vector<uint8_t> out(size_t(W) * size_t(H));
    const int fx = hmax / c.h, fy = vmax / c.v;
    const int pw = (W * c.h + hmax - 1) / hmax, ph = (H * c.v + vmax - 1) / vmax;  // the component's own size
    auto at = [&](int x, int y) { return int(c.plane[size_t(std::min(std::max(y, 0), ph - 1)) * size_t(cw) + size_t(std::min(std::max(x, 0), pw - 1))]); };
    if (fx == 1 && fy == 1) {
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) out[size_t(y) * W + x] = uint8_t(at(x, y));
    } else if (fx == 2 && fy == 1) {
        // Triangular filter: each output sample is 3/4 of its source sample and 1/4 of the nearer neighbour.
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const int i = x >> 1, near = (x & 1) ? i + 1 : i - 1;
                out[size_t(y) * W + x] = clamp8((3 * at(i, y) + at(near, y) + ((x & 1) ? 2 : 1)) >> 2);
            }
    } else if (fx == 2 && fy == 2) {
        // The same filter vertically, then horizontally on the column sums (weights 9/3/3/1 over 16).
        for (int y = 0; y < H; y++) {
            const int j = y >> 1, nj = (y & 1) ? j + 1 : j - 1;
            for (int x = 0; x < W; x++) {
                const int i = x >> 1, ni = (x & 1) ? i + 1 : i - 1;
                const int col = 3 * at(i, j) + at(i, nj), coln = 3 * at(ni, j) + at(ni, nj);
                out[size_t(y) * W + x] = clamp8((3 * col + coln + ((x & 1) ? 7 : 8)) >> 4);
            }
        }
    } else {
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) out[size_t(y) * W + x] = uint8_t(at(x / fx, y / fy));
    }
    return out;
}

void decode(const uint8_t* d, size_t n, Image8& out) {
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) fail("missing SOI marker");
    int qt[4][64] = {};
    bool qt_set[4] = {};
    Huffman dc[4], ac[4];
    std::vector<Component> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, restart = 0, adobe = -1;
    bool frame = false, scanned = false;
    const Idct idct;
    size_t pos = 2;
    while (pos + 4 <= n) {
        if (d[pos] != 0xFF) fail("marker expected");
        const uint8_t m = d[pos + 1];
        if (m == 0xFF) {  // fill byte
            pos++;
            continue;
        }
        pos += 2;
        if (m == 0xD9) break;  // EOI
        if (m >= 0xD0 && m <= 0xD7) continue;
        const size_t len = (size_t(d[pos]) << 8) | d[pos + 1];
        if (len < 2 || pos + len > n) fail("truncated segment");
        const uint8_t* s = d + pos + 2;
        const size_t sl = len - 2;
        switch (m) {
            case 0xC0:
            case 0xC1: {
                if (sl < 6 || s[0] != 8) fail("only 8-bit samples are supported");
                H = (s[1] << 8) | s[2];
                W = (s[3] << 8) | s[4];
                const int nc = s[5];
                if (W <= 0 || H <= 0 || int64_t(W) * H > (int64_t(1) << 28)) fail("image size out of range");
                if ((nc != 1 && nc != 3) || sl < size_t(6 + 3 * nc)) fail("only 1- and 3-component images are supported");
                comps.resize(size_t(nc));
                for (int k = 0; k < nc; k++) {
                    comps[size_t(k)].id = s[6 + 3 * k];
                    comps[size_t(k)].h = s[7 + 3 * k] >> 4;
                    comps[size_t(k)].v = s[7 + 3 * k] & 15;
                    comps[size_t(k)].tq = s[8 + 3 * k];
                    if (comps[size_t(k)].tq > 3) fail("bad quantisation table selector");
                    if (comps[size_t(k)].h < 1 || comps[size_t(k)].h > 4 || comps[size_t(k)].v < 1 || comps[size_t(k)].v > 4) fail("bad sampling factors");
                    hmax = std::max(hmax, comps[size_t(k)].h);
                    vmax = std::max(vmax, comps[size_t(k)].v);
                }
                const int mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
                for (Component& c : comps) {
                    c.bw = mcux * c.h;
                    c.bh = mcuy * c.v;
                    c.plane.assign(size_t(c.bw) * 8 * size_t(c.bh) * 8, 0);
                }
                frame = true;
                break;
            }
            case 0xC2: fail("progressive JPEG is not supported");
            case 0xC3:
            case 0xC5:
            case 0xC6:
            case 0xC7:
            case 0xC9:
            case 0xCA:
            case 0xCB:
            case 0xCD:
            case 0xCE:
            case 0xCF: fail("lossless, hierarchical and arithmetic-coded JPEG are not supported");
            case 0xDB: {
                size_t p = 0;
                while (p < sl) {
                    const int pq = s[p] >> 4, tq = s[p] & 15;
                    if (pq > 1) fail("bad quantisation table precision");
                    if (tq > 3) fail("bad quantisation table selector");
                    p++;
                    if (p + size_t(pq ? 128 : 64) > sl) fail("truncated quantisation table");
                    for (int k = 0; k < 64; k++) qt[tq][kZigzag[k]] = pq ? (s[p + 2 * size_t(k)] << 8 | s[p + 2 * size_t(k) + 1]) : s[p + size_t(k)];
                    qt_set[tq] = true;
                    p += pq ? 128 : 64;
                }
                break;
            }
            case 0xC4: {
                size_t p = 0;
                while (p + 17 <= sl) {
                    const int tc = s[p] >> 4, th = s[p] & 15;
                    if (tc > 1 || th > 3) fail("bad Huffman table selector");
                    int total = 0;
                    for (int k = 0; k < 16; k++) total += s[p + 1 + size_t(k)];
                    if (p + 17 + size_t(total) > sl || total > 256) fail("truncated Huffman table");
                    (tc ? ac : dc)[th].build(s + p + 1, s + p + 17, total);
                    p += 17 + size_t(total);
                }
                break;
            }
            case 0xDD:
                if (sl < 2) fail("bad DRI segment");
                restart = (s[0] << 8) | s[1];
                break;
            case 0xEE:
                if (sl >= 12 && std::memcmp(s, "Adobe", 5) == 0) adobe = s[11];
                break;
            case 0xDA: {
                if (!frame) fail("scan before frame header");
                const int ns = s[0];
                if (ns < 1 || ns > 4 || sl < size_t(1 + 2 * ns + 3)) fail("bad scan header");
                std::vector<Component*> sc;
                for (int k = 0; k < ns; k++) {
                    Component* c = nullptr;
                    for (Component& cc : comps)
                        if (cc.id == s[1 + 2 * k]) c = &cc;
                    if (!c) fail("scan names an unknown component");
                    c->td = s[2 + 2 * k] >> 4;
                    c->ta = s[2 + 2 * k] & 15;
                    if (c->td > 3 || c->ta > 3) fail("bad Huffman table selector");
                    if (!dc[c->td].present || !ac[c->ta].present || !qt_set[c->tq]) fail("missing table");
                    sc.push_back(c);
                }
                BitReader br(d, n, pos + len);
                for (Component* c : sc) c->dc_pred = 0;
                // Units: MCUs of an interleaved scan, or single blocks of a one-component scan (over the
                // component's own extent, without the MCU padding).
                int ux, uy;
                if (ns == 1) {
                    const Component& c = *sc[0];
                    ux = (((W * c.h + hmax - 1) / hmax) + 7) / 8;
                    uy = (((H * c.v + vmax - 1) / vmax) + 7) / 8;
                } else {
                    ux = (W + 8 * hmax - 1) / (8 * hmax);
                    uy = (H + 8 * vmax - 1) / (8 * vmax);
                }
                int coef[64];
                int zz[64];
                auto block = [&](Component& c, int bx, int by) {
                    std::memset(zz, 0, sizeof(zz));
                    const int t = decode_symbol(br, dc[c.td]);
                    if (t > 11) fail("bad DC magnitude");
                    c.dc_pred += extend(br.bits(t), t);
                    zz[0] = c.dc_pred;
                    for (int k = 1; k < 64;) {
                        const int rs = decode_symbol(br, ac[c.ta]);
                        const int r = rs >> 4, sz = rs & 15;
                        if (sz == 0) {
                            if (r != 15) break;  // EOB
                            k += 16;
                            continue;
                        }
                        k += r;
                        if (k > 63) fail("AC coefficient index out of range");
                        zz[k++] = extend(br.bits(sz), sz);
                    }
                    for (int k = 0; k < 64; k++) coef[kZigzag[k]] = zz[k] * qt[c.tq][kZigzag[k]];
                    if (bx < c.bw && by < c.bh) idct.run(coef, c.plane.data() + size_t(by) * 8 * size_t(c.bw) * 8 + size_t(bx) * 8, c.bw * 8);
                };
                int units = 0;
                for (int my = 0; my < uy; my++)
                    for (int mx = 0; mx < ux; mx++) {
                        if (restart && units > 0 && units % restart == 0) {
                            br.restart();
                            for (Component* c : sc) c->dc_pred = 0;
                        }
                        if (ns == 1) {
                            block(*sc[0], mx, my);
                        } else {
                            for (Component* c : sc)
                                for (int v = 0; v < c->v; v++)
                                    for (int h = 0; h < c->h; h++) block(*c, mx * c->h + h, my * c->v + v);
                        }
                        units++;
                    }
                // Continue after the entropy-coded data: find the next marker that is not a restart.
                size_t p = br.pos();
                while (p + 1 < n && !(d[p] == 0xFF && d[p + 1] != 0x00 && !(d[p + 1] >= 0xD0 && d[p + 1] <= 0xD7))) p++;
                pos = p;
                scanned = true;
                continue;
            }
            default: break;  // APPn, COM and others
        }
        pos += len;
    }
    if (!frame || !scanned) fail("no image data");
    out.width = W;
    out.height = H;
    out.px.assign(size_t(W) * size_t(H), 0);
    if (comps.size() == 1) {
        const std::vector<uint8_t> g = upsample(comps[0], hmax, vmax, W, H);
        for (size_t i = 0; i < g.size(); i++) out.px[i] = uint32_t(g[i]) * 0x010101u | 0xFF000000u;
        return;
    }
    const std::vector<uint8_t> a = upsample(comps[0], hmax, vmax, W, H), b = upsample(comps[1], hmax, vmax, W, H), c = upsample(comps[2], hmax, vmax, W, H);
    const bool rgb = adobe == 0;
    for (size_t i = 0; i < a.size(); i++) {
        int R, G, B;
        if (rgb) {
            R = a[i];
            G = b[i];
            B = c[i];
        } else {
            // JFIF YCbCr to RGB (T.871).
            const float y = a[i], cb = float(b[i]) - 128.0f, cr = float(c[i]) - 128.0f;
            R = int(std::lround(y + 1.402f * cr));
            G = int(std::lround(y - 0.344136f * cb - 0.714136f * cr));
            B = int(std::lround(y + 1.772f * cb));
        }
        out.px[i] = uint32_t(clamp8(R)) | uint32_t(clamp8(G)) << 8 | uint32_t(clamp8(B)) << 16 | 0xFF000000u;
    }
}

}  // namespace

bool decode_jpeg(const uint8_t* data, size_t size, Image8& out, std::string& error) {
    try {
        decode(data, size, out);
        return true;
    } catch (const JpegError& e) {
        error = e.msg;
        return false;
    }
}

}  // namespace phh
