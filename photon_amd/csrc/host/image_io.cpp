// Image codecs of the host library (image_io.hpp): PNG over zlib, Radiance RGBE, PFM (baseline JPEG is in
// image_jpeg.cpp).
//
// Behaviour kept from the reference (src/scene/image.cpp:23-116, which calls stb_image):
//   * decoded textures are RGBA8 with row 0 at the top; 16-bit PNG samples keep their high byte;
//   * an 8-bit environment image is linearised with srgb_to_linear, an HDR one is taken as is
//     (mantissa x 2^(exponent - 136)), and non-finite or negative texels are cleared to 0;
//   * PFM files are written "PF\n<w> <h>\n-1.0\n" with rows bottom to top.
#include "image_io.hpp"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>

namespace phh {

namespace {

bool read_all(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamoff n = f.tellg();
    if (n < 0) return false;
    out.resize(size_t(n));
    f.seekg(0);
    return bool(f.read(reinterpret_cast<char*>(out.data()), n)) || n == 0;
}

uint32_t be32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]); }

uint32_t rgba(uint32_t r, uint32_t g, uint32_t b, uint32_t a) { return r | (g << 8) | (b << 16) | (a << 24); }

// ---- PNG decoding ---------------------------------------------------------------------------------------

struct PngHeader {
    uint32_t width = 0, height = 0;
    int depth = 0, colour = 0, interlace = 0;
    int channels() const { return colour == 0 ? 1 : colour == 2 ? 3 : colour == 3 ? 1 : colour == 4 ? 2 : 4; }
    int bits_per_pixel() const { return channels() * depth; }
    size_t row_bytes(uint32_t w) const { return (size_t(w) * size_t(bits_per_pixel()) + 7) / 8; }
};

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Reverses the per-row filters of one (sub)image in place: `data` holds h rows of 1 + row bytes.
bool unfilter(uint8_t* data, uint32_t h, size_t row, int bpp_bytes, std::vector<uint8_t>& out, std::string& error) {
    out.assign(size_t(h) * row, 0);
    const int step = std::max(1, bpp_bytes);
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t type = data[size_t(y) * (row + 1)];
        const uint8_t* in = data + size_t(y) * (row + 1) + 1;
        uint8_t* cur = out.data() + size_t(y) * row;
        const uint8_t* up = y ? cur - row : nullptr;
        for (size_t i = 0; i < row; i++) {
            const int a = i >= size_t(step) ? cur[i - size_t(step)] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= size_t(step)) ? up[i - size_t(step)] : 0;
            int pred;
            switch (type) {
                case 0: pred = 0; break;
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: pred = paeth(a, b, c); break;
                default: error = "PNG: invalid filter type " + std::to_string(type); return false;
            }
            cur[i] = uint8_t(in[i] + pred);
        }
    }
    return true;
}

// Sample k (0-based) of a row of `depth`-bit samples.
uint32_t sample_at(const uint8_t* row, size_t k, int depth) {
    if (depth == 8) return row[k];
    if (depth == 16) return (uint32_t(row[2 * k]) << 8) | row[2 * k + 1];
    const size_t bit = k * size_t(depth);
    const int shift = 8 - depth - int(bit % 8);
    return (row[bit / 8] >> shift) & ((1u << depth) - 1u);
}

bool decode_png(const uint8_t* data, size_t size, Image8& out, std::string& error) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (size < 8 || std::memcmp(data, sig, 8) != 0) {
        error = "not a PNG file";
        return false;
    }
    PngHeader hd;
    std::vector<uint8_t> idat, palette, trns;
    bool have_header = false, ended = false;
    size_t off = 8;
    while (off + 12 <= size && !ended) {
        const uint32_t len = be32(data + off);
        const uint8_t* type = data + off + 4;
        if (size_t(len) > size - off - 12) {
            error = "PNG: truncated chunk";
            return false;
        }
        const uint8_t* body = data + off + 8;
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len < 13) {
                error = "PNG: short IHDR";
                return false;
            }
            hd.width = be32(body);
            hd.height = be32(body + 4);
            hd.depth = body[8];
            hd.colour = body[9];
            hd.interlace = body[12];
            if (body[10] != 0 || body[11] != 0 || hd.interlace > 1) {
                error = "PNG: unsupported compression, filter or interlace method";
                return false;
            }
            have_header = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            palette.assign(body, body + len);
        } else if (std::memcmp(type, "tRNS", 4) == 0) {
            trns.assign(body, body + len);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), body, body + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            ended = true;
        }
        off += size_t(len) + 12;
    }
    if (!have_header || idat.empty()) {
        error = "PNG: missing IHDR or IDAT";
        return false;
    }
    const int d = hd.depth, ct = hd.colour;
    const bool depth_ok = (ct == 0 && (d == 1 || d == 2 || d == 4 || d == 8 || d == 16)) || (ct == 3 && (d == 1 || d == 2 || d == 4 || d == 8)) ||
                          ((ct == 2 || ct == 4 || ct == 6) && (d == 8 || d == 16));
    if (!depth_ok) {
        error = "PNG: invalid colour type / bit depth " + std::to_string(ct) + "/" + std::to_string(d);
        return false;
    }
    if (hd.width == 0 || hd.height == 0 || hd.width > (1u << 24) || hd.height > (1u << 24) || uint64_t(hd.width) * hd.height > (uint64_t(1) << 30)) {
        error = "PNG: image size out of range";
        return false;
    }
    if (ct == 3 && (palette.empty() || palette.size() % 3)) {
        error = "PNG: palette image without a valid PLTE";
        return false;
    }
    // Pass geometry (one pass, or the seven Adam7 passes).
    struct Pass {
        uint32_t x0, y0, dx, dy, w, h;
    };
    std::vector<Pass> passes;
    if (hd.interlace == 0) {
        passes.push_back({0, 0, 1, 1, hd.width, hd.height});
    } else {
        static const uint32_t a7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
        for (const auto& p : a7) {
            const uint32_t w = hd.width > p[0] ? (hd.width - p[0] + p[2] - 1) / p[2] : 0;
            const uint32_t h = hd.height > p[1] ? (hd.height - p[1] + p[3] - 1) / p[3] : 0;
            passes.push_back({p[0], p[1], p[2], p[3], w, h});
        }
    }
    size_t raw_size = 0;
    for (const Pass& p : passes)
        if (p.w && p.h) raw_size += size_t(p.h) * (hd.row_bytes(p.w) + 1);
    std::vector<uint8_t> raw(raw_size);
    uLongf got = uLongf(raw_size);
    const int zr = uncompress(raw.data(), &got, idat.data(), uLong(idat.size()));
    if (zr != Z_OK || got != raw_size) {
        error = zr == Z_OK ? "PNG: image data shorter than the header says" : "PNG: corrupt image data";
        return false;
    }
    // tRNS: palette alphas, or the one grey / RGB value that is transparent.
    const uint32_t sample_max = (1u << d) - 1u;
    int key[3] = {-1, -1, -1};
    if (!trns.empty() && ct == 0 && trns.size() >= 2) key[0] = int((uint32_t(trns[0]) << 8 | trns[1]) & sample_max);
    if (!trns.empty() && ct == 2 && trns.size() >= 6)
        for (int k = 0; k < 3; k++) key[k] = int((uint32_t(trns[2 * k]) << 8 | trns[2 * k + 1]) & sample_max);
    auto to8 = [&](uint32_t v) -> uint32_t {
        if (d == 16) return v >> 8;
        if (d == 8) return v;
        return v * (255u / sample_max);  // 1, 2 and 4-bit grey scale to the full range
    };
    out.width = int(hd.width);
    out.height = int(hd.height);
    out.px.assign(size_t(hd.width) * hd.height, 0);
    const int bpp_bytes = (hd.bits_per_pixel() + 7) / 8;
    size_t at = 0;
    std::vector<uint8_t> rows;
    for (const Pass& p : passes) {
        if (!p.w || !p.h) continue;
        const size_t row = hd.row_bytes(p.w);
        if (!unfilter(raw.data() + at, p.h, row, bpp_bytes, rows, error)) return false;
        at += size_t(p.h) * (row + 1);
        const int nc = hd.channels();
        for (uint32_t y = 0; y < p.h; y++) {
            const uint8_t* r = rows.data() + size_t(y) * row;
            for (uint32_t x = 0; x < p.w; x++) {
                uint32_t s[4] = {0, 0, 0, 0};
                for (int c = 0; c < nc; c++) s[c] = sample_at(r, size_t(x) * size_t(nc) + size_t(c), d);
                uint32_t px;
                if (ct == 3) {
                    const size_t i = s[0];
                    if (3 * i + 2 >= palette.size()) {
                        error = "PNG: palette index out of range";
                        return false;
                    }
                    const uint32_t a = i < trns.size() ? trns[i] : 255u;
                    px = rgba(palette[3 * i], palette[3 * i + 1], palette[3 * i + 2], a);
                } else if (ct == 0) {
                    const uint32_t g = to8(s[0]);
                    px = rgba(g, g, g, key[0] >= 0 && int(s[0]) == key[0] ? 0u : 255u);
                } else if (ct == 4) {
                    const uint32_t g = to8(s[0]);
                    px = rgba(g, g, g, to8(s[1]));
                } else if (ct == 2) {
                    const bool clear = key[0] >= 0 && int(s[0]) == key[0] && int(s[1]) == key[1] && int(s[2]) == key[2];
                    px = rgba(to8(s[0]), to8(s[1]), to8(s[2]), clear ? 0u : 255u);
                } else {
                    px = rgba(to8(s[0]), to8(s[1]), to8(s[2]), to8(s[3]));
                }
                const size_t X = size_t(p.x0) + size_t(x) * p.dx, Y = size_t(p.y0) + size_t(y) * p.dy;
                out.px[Y * hd.width + X] = px;
            }
        }
    }
    return true;
}

// ---- Radiance RGBE ------------------------------------------------------------------------------------

bool decode_hdr(const std::vector<uint8_t>& f, Backdrop& env, std::string& error) {
    size_t pos = 0;
    auto line = [&](std::string& s) {
        s.clear();
        while (pos < f.size() && f[pos] != '\n') s.push_back(char(f[pos++]));
        if (pos >= f.size()) return false;
        pos++;
        return true;
    };
    std::string s;
    if (!line(s) || (s != "#?RADIANCE" && s != "#?RGBE")) {
        error = "not a Radiance HDR file";
        return false;
    }
    bool format_ok = false;
    while (true) {
        if (!line(s)) {
            error = "HDR: truncated header";
            return false;
        }
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") format_ok = true;
    }
    if (!format_ok) {
        error = "HDR: unsupported format (32-bit_rle_rgbe only)";
        return false;
    }
    if (!line(s)) {
        error = "HDR: missing resolution";
        return false;
    }
    int w = 0, h = 0;
    if (std::sscanf(s.c_str(), "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0 || int64_t(w) * h > (int64_t(1) << 30)) {
        error = "HDR: unsupported resolution line (-Y h +X w only)";
        return false;
    }
    std::vector<uint8_t> rgbe(size_t(w) * size_t(h) * 4);
    std::vector<uint8_t> scan(size_t(w) * 4);
    for (int y = 0; y < h; y++) {
        uint8_t* dst = rgbe.data() + size_t(y) * size_t(w) * 4;
        const bool rle = w >= 8 && w < 32768 && pos + 4 <= f.size() && f[pos] == 2 && f[pos + 1] == 2 && (f[pos + 2] & 0x80) == 0;
        if (!rle) {
            if (pos + size_t(w) * 4 > f.size()) {
                error = "HDR: truncated pixel data";
                return false;
            }
            std::memcpy(dst, &f[pos], size_t(w) * 4);
            pos += size_t(w) * 4;
            continue;
        }
        if ((int(f[pos + 2]) << 8 | f[pos + 3]) != w) {
            error = "HDR: scanline width mismatch";
            return false;
        }
        pos += 4;
        // Four planes (R, G, B, E), each as runs (count > 128: repeat) and literals.
        for (int c = 0; c < 4; c++) {
            int x = 0;
            while (x < w) {
                if (pos >= f.size()) {
                    error = "HDR: truncated run";
                    return false;
                }
                int n = f[pos++];
                if (n > 128) {
                    n -= 128;
                    if (pos >= f.size() || x + n > w) {
                        error = "HDR: bad run";
                        return false;
                    }
                    const uint8_t v = f[pos++];
                    for (int k = 0; k < n; k++) scan[size_t(x++) * 4 + size_t(c)] = v;
                } else {
                    if (n == 0 || x + n > w || pos + size_t(n) > f.size()) {
                        error = "HDR: bad literal";
                        return false;
                    }
                    for (int k = 0; k < n; k++) scan[size_t(x++) * 4 + size_t(c)] = f[pos++];
                }
            }
        }
        std::memcpy(dst, scan.data(), scan.size());
    }
    env.width = w;
    env.height = h;
    env.rgb.resize(size_t(w) * size_t(h));
    for (size_t i = 0; i < env.rgb.size(); i++) {
        const uint8_t* p = &rgbe[4 * i];
        if (p[3] == 0) {
            env.rgb[i] = vec3(0);
        } else {
            const float scale = std::ldexp(1.0f, int(p[3]) - 136);
            env.rgb[i] = vec3(float(p[0]) * scale, float(p[1]) * scale, float(p[2]) * scale);
        }
    }
    return true;
}

// ---- PNG encoding -------------------------------------------------------------------------------------

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back(uint8_t(v >> s));
}

void put_chunk(std::vector<uint8_t>& o, const char* type, const uint8_t* body, size_t n) {
    put_be32(o, uint32_t(n));
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    if (n) o.insert(o.end(), body, body + n);
    put_be32(o, uint32_t(crc32(0, o.data() + start, uInt(o.size() - start))));
}

}  // namespace

bool decode_image(const uint8_t* data, size_t size, Image8& out, std::string& error) {
    if (size >= 8 && data[0] == 0x89 && data[1] == 'P') return decode_png(data, size, out, error);
    if (size >= 3 && data[0] == 0xFF && data[1] == 0xD8 && data[2] == 0xFF) return decode_jpeg(data, size, out, error);
    error = "unsupported image format (PNG and baseline JPEG only)";
    return false;
}

bool load_image(const std::string& path, Image8& out, std::string& error) {
    std::vector<uint8_t> data;
    if (!read_all(path, data)) {
        error = "cannot open " + path;
        return false;
    }
    if (!decode_image(data.data(), data.size(), out, error)) {
        error = path + ": " + error;
        return false;
    }
    return true;
}

bool load_environment(const std::string& path, Backdrop& env, std::string& error) {
    std::vector<uint8_t> data;
    if (!read_all(path, data)) {
        error = "cannot open " + path;
        return false;
    }
    if (data.size() >= 2 && data[0] == '#' && data[1] == '?') {
        if (!decode_hdr(data, env, error)) {
            error = path + ": " + error;
            return false;
        }
    } else if (data.size() >= 2 && data[0] == 'P' && (data[1] == 'F' || data[1] == 'f')) {
        std::vector<float> rgb;
        int w = 0, h = 0;
        if (!load_pfm(path, w, h, rgb)) {
            error = path + ": invalid PFM file";
            return false;
        }
        env.width = w;
        env.height = h;
        env.rgb.resize(size_t(w) * size_t(h));
        for (size_t i = 0; i < env.rgb.size(); i++) env.rgb[i] = vec3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    } else {
        Image8 t;
        if (!decode_image(data.data(), data.size(), t, error)) {
            error = path + ": " + error;
            return false;
        }
        env.width = t.width;
        env.height = t.height;
        env.rgb.resize(t.px.size());
        for (size_t i = 0; i < t.px.size(); i++) {
            const uint32_t p = t.px[i];
            env.rgb[i] = vec3(srgb_to_linear(float(p & 0xFF) / 255.0f), srgb_to_linear(float((p >> 8) & 0xFF) / 255.0f),
                                 srgb_to_linear(float((p >> 16) & 0xFF) / 255.0f));
        }
    }
    for (vec3& t : env.rgb)
        if (!std::isfinite(t.x) || !std::isfinite(t.y) || !std::isfinite(t.z) || t.x < 0 || t.y < 0 || t.z < 0) t = vec3(0);
    return true;
}

bool save_png(const std::string& path, int width, int height, const uint32_t* px) {
    if (width <= 0 || height <= 0 || !px) return false;
    const size_t row = size_t(width) * 4;
    // Each row takes the filter (none, sub, up, average, Paeth) whose output has the smallest sum of
    // magnitudes as signed bytes.
    std::vector<uint8_t> filtered(size_t(height) * (row + 1));
    std::vector<uint8_t> cand(row);
    const uint8_t* bytes = reinterpret_cast<const uint8_t*>(px);
    for (int y = 0; y < height; y++) {
        const uint8_t* cur = bytes + size_t(y) * row;
        const uint8_t* up = y ? cur - row : nullptr;
        uint64_t best_cost = ~uint64_t(0);
        uint8_t* dst = filtered.data() + size_t(y) * (row + 1);
        for (int type = 0; type < 5; type++) {
            uint64_t cost = 0;
            for (size_t i = 0; i < row; i++) {
                const int a = i >= 4 ? cur[i - 4] : 0, b = up ? up[i] : 0, c = (up && i >= 4) ? up[i - 4] : 0;
                const int pred = type == 0 ? 0 : type == 1 ? a : type == 2 ? b : type == 3 ? (a + b) >> 1 : paeth(a, b, c);
                cand[i] = uint8_t(cur[i] - pred);
                cost += uint64_t(std::abs(int(int8_t(cand[i]))));
            }
            if (cost < best_cost) {
                best_cost = cost;
                dst[0] = uint8_t(type);
                std::memcpy(dst + 1, cand.data(), row);
            }
        }
    }
    uLongf zn = compressBound(uLong(filtered.size()));
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, filtered.data(), uLong(filtered.size()), 6) != Z_OK) return false;
    std::vector<uint8_t> o = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    uint8_t ihdr[13];
    for (int k = 0; k < 4; k++) {
        ihdr[k] = uint8_t(uint32_t(width) >> (24 - 8 * k));
        ihdr[4 + k] = uint8_t(uint32_t(height) >> (24 - 8 * k));
    }
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 6;   // RGBA
    ihdr[10] = ihdr[11] = ihdr[12] = 0;
    put_chunk(o, "IHDR", ihdr, 13);
    put_chunk(o, "IDAT", z.data(), zn);
    put_chunk(o, "IEND", nullptr, 0);
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char*>(o.data()), std::streamsize(o.size()));
    return bool(f);
}

bool save_pfm(const std::string& path, int width, int height, const float* rgb) {
    if (width <= 0 || height <= 0 || !rgb) return false;
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "PF\n%d %d\n-1.0\n", width, height);
    bool ok = true;
    for (int y = height - 1; y >= 0 && ok; y--) ok = std::fwrite(rgb + size_t(y) * size_t(width) * 3, sizeof(float) * 3, size_t(width), f) == size_t(width);
    ok = std::fclose(f) == 0 && ok;
    return ok;
}

bool load_pfm(const std::string& path, int& width, int& height, std::vector<float>& rgb) {
    std::vector<uint8_t> f;
    if (!read_all(path, f)) return false;
    // Header: "PF" (RGB) or "Pf" (grey), width, height and scale as white-space separated tokens, then
    // exactly one white-space byte before the rows.
    size_t pos = 0;
    auto token = [&](std::string& t) {
        t.clear();
        while (pos < f.size() && std::isspace(f[pos])) pos++;
        while (pos < f.size() && !std::isspace(f[pos])) t.push_back(char(f[pos++]));
        return !t.empty();
    };
    std::string magic, ws, hs, ss;
    if (!token(magic) || (magic != "PF" && magic != "Pf") || !token(ws) || !token(hs) || !token(ss) || pos >= f.size()) return false;
    pos++;
    char* end = nullptr;
    const long w = std::strtol(ws.c_str(), &end, 10);
    if (*end) return false;
    const long h = std::strtol(hs.c_str(), &end, 10);
    if (*end) return false;
    const double scale = std::strtod(ss.c_str(), &end);
    if (*end || scale == 0.0 || w <= 0 || h <= 0 || w * h > (long(1) << 30)) return false;
    const int nc = magic == "PF" ? 3 : 1;
    const size_t n = size_t(w) * size_t(h) * size_t(nc);
    if (f.size() - pos < n * 4) return false;
    std::vector<float> vals(n);
    std::memcpy(vals.data(), &f[pos], n * 4);
    if (scale > 0)  // big endian
        for (float& v : vals) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            u = __builtin_bswap32(u);
            std::memcpy(&v, &u, 4);
        }
    width = int(w);
    height = int(h);
    rgb.assign(size_t(w) * size_t(h) * 3, 0.0f);
    for (long y = 0; y < h; y++)
        for (long x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) rgb[(size_t(y) * size_t(w) + size_t(x)) * 3 + size_t(c)] = vals[((size_t(h - 1 - y)) * size_t(w) + size_t(x)) * size_t(nc) + size_t(nc == 3 ? c : 0)];
    return true;
}

}  // namespace phh
