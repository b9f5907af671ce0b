// glTF 2.0 import and export for the host scene (gltf.hpp).
//
// Import follows the reference loader's rules (src/scene/gltf.cpp:72-487) so that a file yields the
// same scene arrays:
//   * accessors convert to float with the spec's normalisation (unsigned x / (2^n - 1), signed
//     max(x / (2^(n-1) - 1), -1)); an accessor without a buffer view reads as zeros;
//   * strips and fans become lists with the strip's alternating winding; points and lines are dropped;
//   * a node's transform is its matrix (column major in the file) or T * R * S; normals go through the
//     inverse transpose; a mirroring transform (negative determinant) swaps each triangle's last two
//     indices and negates tangent handedness;
//   * a primitive without a material gets one shared default (grey 0.8, dielectric, roughness 0.6);
//   * textures are decoded once per image and added once per (texture, colour space); one that
//     cannot be decoded is skipped with a warning and the material keeps no texture;
//   * cameras and KHR_lights_punctual lights take the node's origin and its -z axis;
//   * hierarchies deeper than 64 levels fail (they are taken as cyclic).
// Where the reference differs from its own comment (with no "scenes" list it loads every node, children
// included, so children are loaded twice), this loader loads the root nodes only.
#include "gltf.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <string_view>
#include <unordered_map>

#include "image_io.hpp"
#include "json.hpp"

namespace phh {

namespace {

namespace fs = std::filesystem;
using json::Node;

// Row-major 4x4 transform.
struct Xform {
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
};

Xform operator*(const Xform& a, const Xform& b) {
    Xform r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}

vec3 apply_point(const Xform& x, vec3 p) {
    const auto& m = x.m;
    const float px = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    const float py = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    const float pz = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    const float pw = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (pw != 0.0f && pw != 1.0f) return vec3(px / pw, py / pw, pz / pw);
    return vec3(px, py, pz);
}

vec3 apply_dir(const Xform& x, vec3 v) {
    const auto& m = x.m;
    return vec3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}

float det3(const Xform& x) {
    const auto& m = x.m;
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

// Inverse transpose of the linear part, as cofactors over the determinant (the translation column of
// an affine transform does not enter a normal's transform).
struct NormalXform {
    float c[3][3];
};

NormalXform normal_xform(const Xform& x) {
    const auto& m = x.m;
    NormalXform n;
    n.c[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    n.c[0][1] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    n.c[0][2] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    n.c[1][0] = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    n.c[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
    n.c[1][2] = m[0][1] * m[2][0] - m[0][0] * m[2][1];
    n.c[2][0] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    n.c[2][1] = m[0][2] * m[1][0] - m[0][0] * m[1][2];
    n.c[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
    const float inv = 1.0f / det3(x);
    for (auto& row : n.c)
        for (float& v : row) v *= inv;
    return n;
}

vec3 apply_normal(const NormalXform& n, vec3 v) {
    return normalize(vec3(n.c[0][0] * v.x + n.c[0][1] * v.y + n.c[0][2] * v.z, n.c[1][0] * v.x + n.c[1][1] * v.y + n.c[1][2] * v.z,
                          n.c[2][0] * v.x + n.c[2][1] * v.y + n.c[2][2] * v.z));
}

// T * R * S of a glTF node (quaternion x, y, z, w).
Xform trs(vec3 t, vec4 q, vec3 s) {
    const float x = q.x, y = q.y, z = q.z, w = q.w;
    const float r[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                           {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                           {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
    const float sc[3] = {s.x, s.y, s.z}, tr[3] = {t.x, t.y, t.z};
    Xform out;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) out.m[i][j] = r[i][j] * sc[j];
        out.m[i][3] = tr[i];
    }
    return out;
}

fs::path utf8_path(const std::string& s) { return fs::path(std::u8string(s.begin(), s.end())); }

std::string percent_decode(const std::string& uri) {
    std::string out;
    for (size_t i = 0; i < uri.size(); i++) {
        if (uri[i] == '%' && i + 2 < uri.size() && std::isxdigit(uint8_t(uri[i + 1])) && std::isxdigit(uint8_t(uri[i + 2]))) {
            out.push_back(char(std::stoi(uri.substr(i + 1, 2), nullptr, 16)));
            i += 2;
        } else {
            out.push_back(uri[i]);
        }
    }
    return out;
}

// Sextet of each base64 digit; 64 for the white space a data URI may carry, 65 for '=', 66 otherwise.
struct Base64Table {
    uint8_t v[256];
    constexpr Base64Table() : v() {
        for (int c = 0; c < 256; c++) v[c] = 66;
        for (int c = 0; c < 26; c++) {
            v['A' + c] = uint8_t(c);
            v['a' + c] = uint8_t(26 + c);
        }
        for (int c = 0; c < 10; c++) v['0' + c] = uint8_t(52 + c);
        v[int('+')] = 62;
        v[int('/')] = 63;
        v[int('\n')] = v[int('\r')] = v[int(' ')] = 64;
        v[int('=')] = 65;
    }
};
constexpr Base64Table kBase64;

// Decodes until the first '=' (or the end), skipping white space; false on any other character.
bool base64_decode(std::string_view in, std::vector<uint8_t>& out) {
    const size_t start = out.size();
    out.resize(start + in.size() / 4 * 3 + 3);
    uint8_t* dst = out.data() + start;
    const uint8_t* p = reinterpret_cast<const uint8_t*>(in.data());
    const uint8_t* const end = p + in.size();
    // Whole quads of digits four characters at a time; anything else drops to the general loop.
    while (end - p >= 4) {
        const uint32_t a = kBase64.v[p[0]], b = kBase64.v[p[1]], c = kBase64.v[p[2]], d = kBase64.v[p[3]];
        if ((a | b | c | d) >= 64) break;
        const uint32_t q = (a << 18) | (b << 12) | (c << 6) | d;
        dst[0] = uint8_t(q >> 16);
        dst[1] = uint8_t(q >> 8);
        dst[2] = uint8_t(q);
        dst += 3;
        p += 4;
    }
    uint32_t acc = 0;
    int bits = 0;
    for (; p < end; p++) {
        const uint32_t v = kBase64.v[*p];
        if (v == 65) break;
        if (v == 64) continue;
        if (v > 65) {
            out.resize(size_t(dst - out.data()));
            return false;
        }
        acc = (acc << 6) | v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            *dst++ = uint8_t(acc >> bits);
        }
    }
    out.resize(size_t(dst - out.data()));
    return true;
}

std::string base64_encode(const std::vector<uint8_t>& bin) {
    static const char* digits = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve((bin.size() + 2) / 3 * 4);
    for (size_t i = 0; i < bin.size(); i += 3) {
        const size_t left = bin.size() - i;
        uint32_t v = uint32_t(bin[i]) << 16;
        if (left > 1) v |= uint32_t(bin[i + 1]) << 8;
        if (left > 2) v |= uint32_t(bin[i + 2]);
        out.push_back(digits[(v >> 18) & 63]);
        out.push_back(digits[(v >> 12) & 63]);
        out.push_back(left > 1 ? digits[(v >> 6) & 63] : '=');
        out.push_back(left > 2 ? digits[v & 63] : '=');
    }
    return out;
}

bool read_file(const fs::path& p, std::vector<uint8_t>& out) {
    std::ifstream f(p, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamoff n = f.tellg();
    if (n < 0) return false;
    out.resize(size_t(n));
    f.seekg(0);
    return bool(f.read(reinterpret_cast<char*>(out.data()), n)) || n == 0;
}

struct Failure {
    std::string msg;
};

[[noreturn]] void fail(const std::string& msg) { throw Failure{msg}; }

// A byte offset, length or stride of the file: a whole number in [0, 2^40), so that sums and products
// of them cannot wrap.
size_t byte_size(Node v, double fallback, const char* what) {
    const double d = v.real(fallback);
    if (!(d >= 0.0) || d >= 1099511627776.0 || d != std::floor(d)) fail(std::string("invalid ") + what);
    return size_t(d);
}

// The stride between elements of an accessor: the view's byteStride, which glTF 2.0 bounds to [4, 252]
// and which cannot be shorter than one element, or the element size when the view has none.
size_t element_stride(Node view, size_t elem) {
    if (!view.holds("byteStride")) return elem;
    const size_t stride = byte_size(view.member("byteStride"), 0, "byteStride");
    if (stride < 4 || stride > 252 || stride < elem) fail("invalid byteStride");
    return stride;
}

// True when `count` elements of `elem` bytes, `stride` apart from `base`, all lie inside `size` bytes.
// The arithmetic is checked, so no product or sum can wrap.
bool span_fits(size_t base, size_t stride, int count, size_t elem, size_t size) {
    if (count <= 0) return true;
    size_t last = 0, end = 0;
    if (__builtin_mul_overflow(stride, size_t(count - 1), &last)) return false;
    if (__builtin_add_overflow(base, last, &end) || __builtin_add_overflow(end, elem, &end)) return false;
    return end <= size;
}

// Component types of accessors (glTF 2.0, 3.6.2.2).
enum Component : int { kI8 = 5120, kU8 = 5121, kI16 = 5122, kU16 = 5123, kU32 = 5125, kF32 = 5126 };

int component_bytes(int ct) { return (ct == kF32 || ct == kU32) ? 4 : (ct == kI16 || ct == kU16) ? 2 : 1; }

// One component as float: integers normalised as the specification says when `norm` is set
// (unsigned x / (2^n - 1), signed max(x / (2^(n-1) - 1), -1)), converted as they are otherwise.
float component_value(const uint8_t* q, int ct, bool norm) {
    switch (ct) {
        case kF32: {
            float v;
            std::memcpy(&v, q, 4);
            return v;
        }
        case kU8: return norm ? float(q[0]) / 255.0f : float(q[0]);
        case kI8: return norm ? std::max(float(int8_t(q[0])) / 127.0f, -1.0f) : float(int8_t(q[0]));
        case kU16: {
            uint16_t u;
            std::memcpy(&u, q, 2);
            return norm ? float(u) / 65535.0f : float(u);
        }
        case kI16: {
            int16_t v;
            std::memcpy(&v, q, 2);
            return norm ? std::max(float(v) / 32767.0f, -1.0f) : float(v);
        }
        case kU32: {
            uint32_t u;
            std::memcpy(&u, q, 4);
            return float(u);
        }
        default: fail("unsupported component type");
    }
}

int components_of(std::string_view type) {
    static const std::pair<std::string_view, int> known[] = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4}, {"MAT4", 16}};
    for (const auto& [name, n] : known)
        if (name == type) return n;
    fail("unsupported accessor type");
}

vec3 vec3_at(Node a, float otherwise) {
    return vec3(float(a.element(0).real(otherwise)), float(a.element(1).real(otherwise)), float(a.element(2).real(otherwise)));
}

// A node's local transform: its matrix (column major in the file) or T * R * S.
Xform node_local(Node nd) {
    if (nd.holds("matrix")) {
        const Node a = nd.member("matrix");
        Xform x;
        for (int col = 0; col < 4; col++)
            for (int row = 0; row < 4; row++) x.m[row][col] = float(a.element(size_t(col * 4 + row)).real(row == col ? 1.0 : 0.0));
        return x;
    }
    const vec3 t = nd.holds("translation") ? vec3_at(nd.member("translation"), 0.0f) : vec3(0.0f);
    const vec3 sc = nd.holds("scale") ? vec3_at(nd.member("scale"), 1.0f) : vec3(1.0f);
    vec4 q(0, 0, 0, 1);
    if (nd.holds("rotation")) {
        const Node r = nd.member("rotation");
        q = vec4(float(r.element(0).real(0)), float(r.element(1).real(0)), float(r.element(2).real(0)), float(r.element(3).real(1)));
    }
    return trs(t, q, sc);
}

// Material extensions that carry one scalar each, and the field it sets.
struct ScalarExtension {
    const char* extension;
    const char* key;
    float ph_material::*field;
    double otherwise;
};
constexpr ScalarExtension kScalarExtensions[] = {
    {"KHR_materials_transmission", "transmissionFactor", &ph_material::transmission, 0.0},
    {"KHR_materials_ior", "ior", &ph_material::ior, 1.5},
};

int alpha_mode_of(std::string_view mode) {
    if (mode == "MASK") return PH_ALPHA_MASK;
    if (mode == "BLEND") return PH_ALPHA_BLEND;
    return PH_ALPHA_OPAQUE;
}

class GltfImport {
public:
    GltfImport(Scene& s, const std::string& path) : scene_(s), path_(path) {}

    void run() {
        std::vector<uint8_t> file;
        if (!read_file(utf8_path(path_), file)) fail("cannot open " + path_);
        dir_ = utf8_path(path_).parent_path();
        std::string glb_text;
        const bool glb = file.size() >= 12 && std::memcmp(file.data(), "glTF", 4) == 0;
        if (glb) split_glb(file, glb_text);
        const std::string_view text = glb ? std::string_view(glb_text) : std::string_view(reinterpret_cast<const char*>(file.data()), file.size());
        try {
            doc_ = json::Document::read(text);
        } catch (const json::SyntaxError& e) {
            fail(e.what());
        }
        root_ = doc_.root();
        if (root_.member("asset").member("version").text().substr(0, 1) != "2") fail("only glTF 2.0 is supported");
        load_buffers(glb);
        decoded_.resize(root_.member("images").length());
        decode_state_.assign(decoded_.size(), 0);
        const Node mats = root_.member("materials");
        material_base_ = int(scene_.materials.size());
        for (size_t i = 0; i < mats.length(); i++) scene_.materials.push_back(read_material(mats.element(i)));
        for (const int n : root_nodes()) visit(n, Xform(), 0);
    }

private:
    Scene& scene_;
    std::string path_;
    fs::path dir_;
    json::Document doc_;
    Node root_;
    std::vector<uint8_t> glb_bin_;
    std::vector<std::vector<uint8_t>> buffers_;
    std::vector<Image8> decoded_;
    std::vector<int> decode_state_;  // 0 not tried, 1 decoded, -1 failed
    std::map<std::pair<int, bool>, int> texture_ids_;
    int material_base_ = 0;
    int fallback_ = -1;

    Node at(const char* array, int i) const { return i < 0 ? Node() : root_.member(array).element(size_t(i)); }

    // The default scene's nodes; with no scenes list, every node that is no other node's child.
    std::vector<int> root_nodes() const {
        const Node sc = at("scenes", root_.member("scene").whole(0));
        std::vector<int> out;
        if (!sc.null()) {
            const Node list = sc.member("nodes");
            for (size_t i = 0; i < list.length(); i++) out.push_back(list.element(i).whole());
            return out;
        }
        const Node nodes = root_.member("nodes");
        std::vector<char> child(nodes.length(), 0);
        for (size_t i = 0; i < nodes.length(); i++) {
            const Node kids = nodes.element(i).member("children");
            for (size_t k = 0; k < kids.length(); k++) {
                const int c = kids.element(k).whole();
                if (c >= 0 && size_t(c) < child.size()) child[size_t(c)] = 1;
            }
        }
        for (size_t i = 0; i < child.size(); i++)
            if (!child[i]) out.push_back(int(i));
        return out;
    }

    void split_glb(const std::vector<uint8_t>& file, std::string& text) {
        // Header (magic, version, length), then chunks of (length, type, data padded to 4 bytes).
        for (size_t off = 12; off + 8 <= file.size();) {
            uint32_t len, type;
            std::memcpy(&len, &file[off], 4);
            std::memcpy(&type, &file[off + 4], 4);
            off += 8;
            if (size_t(len) > file.size() - off) fail("truncated GLB chunk");
            if (type == 0x4E4F534Au)  // "JSON"
                text.assign(reinterpret_cast<const char*>(&file[off]), len);
            else if (type == 0x004E4942u)  // "BIN\0"
                glb_bin_.assign(file.begin() + long(off), file.begin() + long(off + len));
            off += (size_t(len) + 3) & ~size_t(3);
        }
    }

    void load_buffers(bool glb) {
        const Node bufs = root_.member("buffers");
        for (size_t i = 0; i < bufs.length(); i++) {
            const Node b = bufs.element(i);
            std::vector<uint8_t> data;
            if (!b.holds("uri")) {
                if (!glb) fail("buffer " + std::to_string(i) + " has no data");
                data = glb_bin_;
            } else {
                const std::string_view uri = b.member("uri").text();
                if (uri.substr(0, 5) == "data:") {
                    const size_t comma = uri.find(',');
                    if (comma == std::string_view::npos || !base64_decode(uri.substr(comma + 1), data)) fail("invalid data URI");
                } else if (!read_file(dir_ / utf8_path(percent_decode(std::string(uri))), data)) {
                    fail("cannot read buffer " + std::string(uri));
                }
            }
            buffers_.push_back(std::move(data));
        }
    }

    const std::vector<uint8_t>& buffer_of(Node view) const {
        const int b = view.member("buffer").whole(0);
        if (b < 0 || size_t(b) >= buffers_.size()) fail("buffer view refers to a missing buffer");
        return buffers_[size_t(b)];
    }

    static int count_of(Node a) {
        const size_t n = byte_size(a.member("count"), 0, "accessor count");
        if (n > size_t(1) << 30) fail("accessor count out of range");
        return int(n);
    }

    // Where an accessor's elements lie: the buffer, the first byte and the stride, range-checked.
    struct Span {
        const std::vector<uint8_t>* buf = nullptr;
        size_t base = 0, stride = 0;
    };
    Span span_of(Node a, int count, size_t element, const char* what) const {
        const Node view = at("bufferViews", a.member("bufferView").whole());
        Span s;
        s.buf = &buffer_of(view);
        s.base = byte_size(view.member("byteOffset"), 0, "byteOffset") + byte_size(a.member("byteOffset"), 0, "byteOffset");
        s.stride = element_stride(view, element);
        if (!span_fits(s.base, s.stride, count, element, s.buf->size())) fail(what);
        return s;
    }

    // Accessor -> count x `want` floats (components the accessor lacks stay 0; an accessor without a
    // buffer view reads as zeros).
    std::vector<float> floats(int accessor, int want) const {
        const Node a = at("accessors", accessor);
        if (a.null()) fail("missing accessor");
        const int count = count_of(a);
        const int nc = components_of(a.member("type").text());
        const int ct = a.member("componentType").whole();
        const bool norm = a.member("normalized").truth();
        const int cb = component_bytes(ct);
        std::vector<float> out(size_t(count) * size_t(want), 0.0f);
        if (!a.holds("bufferView")) return out;
        const Span s = span_of(a, count, size_t(nc * cb), "accessor out of the buffer's range");
        const int take = std::min(nc, want);
        for (int i = 0; i < count; i++) {
            const uint8_t* elem = s.buf->data() + s.base + s.stride * size_t(i);
            for (int c = 0; c < take; c++) out[size_t(i) * size_t(want) + size_t(c)] = component_value(elem + size_t(c * cb), ct, norm);
        }
        return out;
    }

    std::vector<int> indices(int accessor) const {
        const Node a = at("accessors", accessor);
        const int count = count_of(a);
        const int ct = a.member("componentType").whole();
        const int cb = ct == kU32 ? 4 : ct == kU16 ? 2 : 1;
        if (!a.holds("bufferView")) fail("index accessor without data");
        const Span s = span_of(a, count, size_t(cb), "index accessor out of range");
        std::vector<int> out(static_cast<size_t>(count));
        for (int i = 0; i < count; i++) {
            const uint8_t* p = s.buf->data() + s.base + s.stride * size_t(i);
            uint32_t v = p[0];
            if (cb == 2) {
                uint16_t u;
                std::memcpy(&u, p, 2);
                v = u;
            } else if (cb == 4) {
                std::memcpy(&v, p, 4);
            }
            out[size_t(i)] = int(v);
        }
        return out;
    }

    // Decodes glTF image `i` into decoded_[i]; false (with the reason) when it cannot.
    bool decode(int i, std::string& why) {
        const Node img = at("images", i);
        Image8& out = decoded_[size_t(i)];
        if (img.holds("uri")) {
            const std::string uri(img.member("uri").text());
            if (uri.rfind("data:", 0) != 0) return load_image((dir_ / utf8_path(percent_decode(uri))).string(), out, why);
            std::vector<uint8_t> data;
            const size_t comma = uri.find(',');
            if (comma == std::string::npos || !base64_decode(uri.substr(comma + 1), data)) {
                why = "invalid image data URI";
                return false;
            }
            return decode_image(data.data(), data.size(), out, why);
        }
        if (!img.holds("bufferView")) {
            why = "image without data";
            return false;
        }
        const Node view = at("bufferViews", img.member("bufferView").whole());
        const int b = view.member("buffer").whole(0);
        if (b < 0 || size_t(b) >= buffers_.size()) {
            why = "image buffer view refers to a missing buffer";
            return false;
        }
        const std::vector<uint8_t>& buf = buffers_[size_t(b)];
        const double off = view.member("byteOffset").real(0), len = view.member("byteLength").real(0);
        if (!(off >= 0 && len >= 0 && off + len <= double(buf.size()))) {
            why = "image buffer view out of range";
            return false;
        }
        return decode_image(buf.data() + size_t(off), size_t(len), out, why);
    }

    // The scene image for a material's texture reference in a colour space (each glTF texture is added
    // once per colour space it is used in); -1 when absent or undecodable.
    int texture(Node ref, bool srgb) {
        if (!ref.is_object() || !ref.holds("index")) return -1;
        const int ti = ref.member("index").whole();
        const auto key = std::make_pair(ti, srgb);
        if (const auto hit = texture_ids_.find(key); hit != texture_ids_.end()) return hit->second;
        const Node t = at("textures", ti);
        int source = t.member("source").whole(-1);
        if (source < 0) source = t.member("extensions").member("EXT_texture_webp").member("source").whole(-1);
        int id = -1;
        if (source >= 0 && size_t(source) < decoded_.size()) {
            int& state = decode_state_[size_t(source)];
            if (state == 0) {
                std::string why;
                state = decode(source, why) ? 1 : -1;
                if (state < 0) std::fprintf(stderr, "warning: texture skipped: %s\n", why.c_str());
            }
            if (state > 0) {
                Image8 img = decoded_[size_t(source)];
                img.srgb = srgb;
                // CLAMP_TO_EDGE (33071) clamps; the repeating wraps (and no sampler) repeat.
                img.tile = at("samplers", t.member("sampler").whole(-1)).member("wrapS").whole(10497) != 33071;
                id = scene_.add_image(std::move(img));
            }
        }
        texture_ids_[key] = id;
        return id;
    }

    ph_material read_material(Node m) {
        ph_material out = default_material();
        const Node pbr = m.member("pbrMetallicRoughness");
        if (pbr.holds("baseColorFactor")) {
            const Node f = pbr.member("baseColorFactor");
            for (int k = 0; k < 4; k++) out.base_color[k] = float(f.element(size_t(k)).real(1.0));
        }
        out.metallic = float(pbr.member("metallicFactor").real(1.0));
        out.roughness = float(pbr.member("roughnessFactor").real(1.0));
        // Colour textures are sRGB-encoded, data textures linear.
        out.base_color_tex = texture(pbr.member("baseColorTexture"), true);
        out.metallic_roughness_tex = texture(pbr.member("metallicRoughnessTexture"), false);
        out.normal_tex = texture(m.member("normalTexture"), false);
        out.normal_scale = float(m.member("normalTexture").member("scale").real(1.0));
        out.emissive_tex = texture(m.member("emissiveTexture"), true);
        const Node ext = m.member("extensions");
        const float strength = float(ext.member("KHR_materials_emissive_strength").member("emissiveStrength").real(1.0));
        const Node emissive = m.member("emissiveFactor");
        for (int k = 0; k < 3; k++) out.emission[k] = (emissive.null() ? 0.0f : float(emissive.element(size_t(k)).real(0.0))) * strength;
        for (const ScalarExtension& e : kScalarExtensions) out.*e.field = float(ext.member(e.extension).member(e.key).real(e.otherwise));
        out.double_sided = m.member("doubleSided").truth() ? 1 : 0;
        out.alpha_mode = alpha_mode_of(m.member("alphaMode").text());
        out.alpha_cutoff = out.alpha_mode == PH_ALPHA_BLEND ? 0.5f : float(m.member("alphaCutoff").real(0.5));
        return out;
    }

    // The material of primitives that name none: shared, grey 0.8, dielectric, roughness 0.6.
    int fallback_material() {
        if (fallback_ < 0) {
            ph_material m = default_material();
            std::fill(m.base_color, m.base_color + 3, 0.8f);
            m.metallic = 0.0f;
            m.roughness = 0.6f;
            fallback_ = scene_.add_material(m);
        }
        return fallback_;
    }

    // Triangle list of a primitive: its indices (or 0..n-1), strips and fans unrolled with the strip's
    // alternating winding.
    std::vector<int> triangle_list(Node prim, int mode, size_t nv) const {
        std::vector<int> idx;
        if (prim.holds("indices")) {
            idx = indices(prim.member("indices").whole());
        } else {
            idx.resize(nv);
            for (size_t i = 0; i < nv; i++) idx[i] = int(i);
        }
        if (mode == 4) return idx;
        std::vector<int> list;
        for (size_t i = 2; i < idx.size(); i++) {
            if (mode == 6)
                list.insert(list.end(), {idx[0], idx[i - 1], idx[i]});
            else if (i % 2 == 0)
                list.insert(list.end(), {idx[i - 2], idx[i - 1], idx[i]});
            else
                list.insert(list.end(), {idx[i - 1], idx[i - 2], idx[i]});
        }
        return list;
    }

    void primitive(Node prim, const Xform& world) {
        const int mode = prim.member("mode").whole(4);
        if (mode != 4 && mode != 5 && mode != 6) return;  // points and lines carry no surface
        const Node attrs = prim.member("attributes");
        if (!attrs.holds("POSITION")) return;
        auto stream = [&](const char* name, int width) {
            return attrs.holds(name) ? floats(attrs.member(name).whole(), width) : std::vector<float>{};
        };
        const std::vector<float> pos = floats(attrs.member("POSITION").whole(), 3);
        const std::vector<float> nrm = stream("NORMAL", 3), uv = stream("TEXCOORD_0", 2), tan = stream("TANGENT", 4);
        const size_t nv = pos.size() / 3;
        std::vector<int> idx = triangle_list(prim, mode, nv);
        for (const int i : idx)
            if (i < 0 || size_t(i) >= nv) fail("vertex index out of range");
        // A mirroring transform turns the winding: swap each triangle's last two corners.
        const bool mirrored = det3(world) < 0.0f;
        if (mirrored)
            for (size_t i = 0; i + 2 < idx.size(); i += 3) std::swap(idx[i + 1], idx[i + 2]);
        std::vector<vec3> p(nv), n;
        std::vector<vec2> t;
        for (size_t i = 0; i < nv; i++) p[i] = apply_point(world, vec3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
        if (!nrm.empty()) {
            const NormalXform nx = normal_xform(world);
            n.resize(nv);
            for (size_t i = 0; i < nv; i++) n[i] = apply_normal(nx, vec3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
        }
        if (!uv.empty()) {
            t.resize(nv);
            for (size_t i = 0; i < nv; i++) t[i] = vec2(uv[2 * i], uv[2 * i + 1]);
        }
        const int material = prim.holds("material") ? material_base_ + prim.member("material").whole() : fallback_material();
        Mesh& mesh = scene_.mesh;
        const size_t first = mesh.pos.size();
        append_mesh(mesh, p, idx, material, n, t);
        // Tangents: the file's, transformed (handedness flipped under mirroring), once any primitive has
        // them; other primitives' vertices get +x.
        if (!tan.empty()) {
            mesh.tan.resize(first, vec4(1, 0, 0, 1));
            for (size_t i = 0; i < nv; i++) {
                const vec3 dir = normalize(apply_dir(world, vec3(tan[4 * i], tan[4 * i + 1], tan[4 * i + 2])));
                mesh.tan.push_back(vec4(dir, tan[4 * i + 3] * (mirrored ? -1.0f : 1.0f)));
            }
        } else if (!mesh.tan.empty()) {
            mesh.tan.resize(mesh.pos.size(), vec4(1, 0, 0, 1));
        }
    }

    void camera(Node cam, const Xform& world) {
        if (cam.member("type").text() != "perspective") return;
        Viewpoint v;
        v.eye = apply_point(world, vec3(0.0f));
        v.aim = v.eye + normalize(apply_dir(world, vec3(0, 0, -1)));
        v.sky = normalize(apply_dir(world, vec3(0, 1, 0)));
        v.fov_deg = float(cam.member("perspective").member("yfov").real(0.8)) * 180.0f / kPi;
        scene_.file_views.push_back(v);
    }

    // A KHR_lights_punctual light at the node's origin, shining down its -z axis.
    void lamp(Node l, const Xform& world) {
        const vec3 colour = l.holds("color") ? vec3_at(l.member("color"), 1.0f) : vec3(1.0f);
        const vec3 power = colour * float(l.member("intensity").real(1.0));
        const vec3 at_pos = apply_point(world, vec3(0.0f));
        const vec3 dir = normalize(apply_dir(world, vec3(0, 0, -1)));
        const std::string_view type = l.member("type").text();
        ph_punctual_light pl{};
        pl.type = type == "directional" ? PH_LIGHT_DIRECTIONAL : type == "spot" ? PH_LIGHT_SPOT : PH_LIGHT_POINT;
        for (int k = 0; k < 3; k++) {
            pl.position[k] = at_pos[k];
            pl.direction[k] = dir[k];
            pl.intensity[k] = power[k];
        }
        const Node spot = l.member("spot");
        pl.cos_inner = std::cos(float(spot.member("innerConeAngle").real(0.0)));
        pl.cos_outer = std::cos(float(spot.member("outerConeAngle").real(double(kPi / 4))));
        scene_.lamps.push_back(pl);
    }

    void visit(int index, const Xform& parent, int depth) {
        if (depth > 64) fail("node hierarchy too deep (or cyclic)");
        const Node nd = at("nodes", index);
        const Xform world = parent * node_local(nd);
        if (nd.holds("mesh")) {
            const Node prims = at("meshes", nd.member("mesh").whole()).member("primitives");
            for (size_t i = 0; i < prims.length(); i++) primitive(prims.element(i), world);
        }
        if (nd.holds("camera")) camera(at("cameras", nd.member("camera").whole()), world);
        const Node light_ref = nd.member("extensions").member("KHR_lights_punctual");
        if (light_ref.holds("light"))
            lamp(root_.member("extensions").member("KHR_lights_punctual").member("lights").element(size_t(light_ref.member("light").whole())), world);
        const Node kids = nd.member("children");
        for (size_t i = 0; i < kids.length(); i++) visit(kids.element(i).whole(), world, depth + 1);
    }
};

// ---- export ----------------------------------------------------------------------------------------

// The binary payload of an export: arrays appended back to back, each padded to 4 bytes.
struct Payload {
    std::vector<uint8_t> bytes;
    size_t add(const void* p, size_t n) {
        const size_t off = bytes.size();
        const uint8_t* b = static_cast<const uint8_t*>(p);
        bytes.insert(bytes.end(), b, b + n);
        bytes.resize((bytes.size() + 3) & ~size_t(3), 0);
        return off;
    }
};

// One primitive per material: the vertices its triangles use (scene ids, in order of first use within
// the primitive) and its triangles' corners as indices into that list.
struct Part {
    std::vector<uint32_t> verts;
    std::vector<uint32_t> corners;
    size_t first = 0;  // its first vertex in the written streams
};

std::map<int, Part> parts_by_material(const Mesh& mesh) {
    std::map<int, Part> parts;
    std::map<int, std::unordered_map<uint32_t, uint32_t>> local;  // per material: scene vertex -> index in the part
    for (int t = 0; t < mesh.triangles(); t++) {
        const int m = mesh.mat[size_t(t)];
        Part& p = parts[m];
        auto& ids = local[m];
        for (int k = 0; k < 3; k++) {
            const uint32_t v = uint32_t(mesh.corner[size_t(3 * t + k)]);
            const auto [it, added] = ids.emplace(v, uint32_t(p.verts.size()));
            if (added) p.verts.push_back(v);
            p.corners.push_back(it->second);
        }
    }
    return parts;
}

// A column-major node matrix with axes x, y, z and origin t.
void node_matrix(json::Writer& w, vec3 x, vec3 y, vec3 z, vec3 t) {
    const float m[16] = {x.x, x.y, x.z, 0, y.x, y.y, y.z, 0, z.x, z.y, z.z, 0, t.x, t.y, t.z, 1};
    w.name("matrix").numbers(m, 16);
}

std::string gltf_text(const Scene& scene) {
    const Mesh& mesh = scene.mesh;
    const size_t nv = mesh.pos.size();
    const bool with_normals = nv > 0 && mesh.nrm.size() == nv;
    const bool with_uvs = nv > 0 && mesh.uv.size() == nv;
    std::map<int, Part> parts = parts_by_material(mesh);

    // Vertex streams, part after part, then each part's index list.
    std::vector<vec3> pos, nrm;
    std::vector<vec2> uv;
    for (auto& [m, p] : parts) {
        p.first = pos.size();
        for (const uint32_t v : p.verts) {
            pos.push_back(mesh.pos[v]);
            if (with_normals) nrm.push_back(mesh.nrm[v]);
            if (with_uvs) uv.push_back(mesh.uv[v]);
        }
    }
    const size_t n = pos.size();
    Payload bin;
    struct View {
        size_t offset, length;
    };
    std::vector<View> views;
    views.push_back({bin.add(pos.data(), n * 12), n * 12});
    const int normal_view = with_normals ? int(views.size()) : -1;
    if (with_normals) views.push_back({bin.add(nrm.data(), n * 12), n * 12});
    const int uv_view = with_uvs ? int(views.size()) : -1;
    if (with_uvs) views.push_back({bin.add(uv.data(), n * 8), n * 8});

    struct Accessor {
        int view;
        size_t offset, count;
        int component;
        const char* type;
        bool bounds;
        vec3 lo, hi;
    };
    std::vector<Accessor> accessors;
    struct Prim {
        int material, position, normal, uv, index;
    };
    std::vector<Prim> prims;
    for (const auto& [m, p] : parts) {
        vec3 lo(kInf), hi(-kInf);
        for (const uint32_t v : p.verts) {
            lo = vmin(lo, mesh.pos[v]);
            hi = vmax(hi, mesh.pos[v]);
        }
        Prim pr{m, int(accessors.size()), -1, -1, -1};
        accessors.push_back({0, p.first * 12, p.verts.size(), kF32, "VEC3", true, lo, hi});
        if (with_normals) {
            pr.normal = int(accessors.size());
            accessors.push_back({normal_view, p.first * 12, p.verts.size(), kF32, "VEC3", false, {}, {}});
        }
        if (with_uvs) {
            pr.uv = int(accessors.size());
            accessors.push_back({uv_view, p.first * 8, p.verts.size(), kF32, "VEC2", false, {}, {}});
        }
        views.push_back({bin.add(p.corners.data(), p.corners.size() * 4), p.corners.size() * 4});
        pr.index = int(accessors.size());
        accessors.push_back({int(views.size()) - 1, 0, p.corners.size(), kU32, "SCALAR", false, {}, {}});
        prims.push_back(pr);
    }

    json::Writer w;
    w.open_object();
    w.name("asset").open_object().field("version", "2.0").field("generator", "photon_amd").close_object();
    w.name("extensionsUsed").open_array();
    for (const char* e : {"KHR_materials_emissive_strength", "KHR_materials_transmission", "KHR_materials_ior", "KHR_lights_punctual"}) w.string(e);
    w.close_array();
    w.name("buffers").open_array().open_object().field("byteLength", bin.bytes.size());
    w.field("uri", "data:application/octet-stream;base64," + base64_encode(bin.bytes)).close_object().close_array();
    w.name("bufferViews").open_array();
    for (const View& v : views) w.open_object().field("buffer", 0).field("byteOffset", v.offset).field("byteLength", v.length).close_object();
    w.close_array();
    w.name("accessors").open_array();
    for (const Accessor& a : accessors) {
        w.open_object().field("bufferView", a.view).field("byteOffset", a.offset).field("componentType", a.component);
        w.field("count", a.count).field("type", a.type);
        if (a.bounds) {
            const float lo[3] = {a.lo.x, a.lo.y, a.lo.z}, hi[3] = {a.hi.x, a.hi.y, a.hi.z};
            w.name("min").numbers(lo, 3);
            w.name("max").numbers(hi, 3);
        }
        w.close_object();
    }
    w.close_array();

    // Emission is written as a factor of at most 1 times KHR_materials_emissive_strength.
    w.name("materials").open_array();
    for (const ph_material& m : scene.materials) {
        const float strength = std::max({m.emission[0], m.emission[1], m.emission[2], 1.0f});
        const float factor[3] = {m.emission[0] / strength, m.emission[1] / strength, m.emission[2] / strength};
        w.open_object();
        w.name("pbrMetallicRoughness").open_object();
        w.name("baseColorFactor").numbers(m.base_color, 4);
        w.field("metallicFactor", m.metallic).field("roughnessFactor", m.roughness).close_object();
        w.name("emissiveFactor").numbers(factor, 3);
        w.field("doubleSided", m.double_sided != 0);
        w.name("extensions").open_object();
        w.name("KHR_materials_emissive_strength").open_object().field("emissiveStrength", strength).close_object();
        w.name("KHR_materials_transmission").open_object().field("transmissionFactor", m.transmission).close_object();
        w.name("KHR_materials_ior").open_object().field("ior", m.ior).close_object();
        w.close_object().close_object();
    }
    w.close_array();

    w.name("meshes").open_array().open_object().name("primitives").open_array();
    for (const Prim& p : prims) {
        w.open_object().name("attributes").open_object().field("POSITION", p.position);
        if (p.normal >= 0) w.field("NORMAL", p.normal);
        if (p.uv >= 0) w.field("TEXCOORD_0", p.uv);
        w.close_object().field("indices", p.index).field("material", p.material).close_object();
    }
    w.close_array().close_object().close_array();

    const Viewpoint& v = scene.view;
    w.name("cameras").open_array().open_object().field("type", "perspective");
    w.name("perspective").open_object().field("yfov", v.fov_deg * kPi / 180.0f).field("znear", 0.01).close_object();
    w.close_object().close_array();

    // Lights: colour of largest component 1, times an intensity.
    w.name("extensions").open_object().name("KHR_lights_punctual").open_object().name("lights").open_array();
    for (const ph_punctual_light& l : scene.lamps) {
        const float power = std::max({l.intensity[0], l.intensity[1], l.intensity[2], 1e-6f});
        const float colour[3] = {l.intensity[0] / power, l.intensity[1] / power, l.intensity[2] / power};
        w.open_object().field("type", l.type == PH_LIGHT_DIRECTIONAL ? "directional" : l.type == PH_LIGHT_SPOT ? "spot" : "point");
        w.name("color").numbers(colour, 3);
        w.field("intensity", power);
        if (l.type == PH_LIGHT_SPOT)
            w.name("spot").open_object().field("innerConeAngle", std::acos(l.cos_inner)).field("outerConeAngle", std::acos(l.cos_outer)).close_object();
        w.close_object();
    }
    w.close_array().close_object().close_object();

    // Nodes: the mesh, the camera (looking down its -z), then one node per light (its -z the light's
    // direction).
    w.name("nodes").open_array();
    w.open_object().field("mesh", 0).close_object();
    const vec3 f = normalize(v.aim - v.eye), r = normalize(cross(f, v.sky)), u = cross(r, f);
    w.open_object().field("camera", 0);
    node_matrix(w, r, u, -f, v.eye);
    w.close_object();
    for (size_t i = 0; i < scene.lamps.size(); i++) {
        const ph_punctual_light& l = scene.lamps[i];
        const vec3 z = -normalize(vec3(l.direction[0], l.direction[1], l.direction[2]));
        const vec3 x = normalize(cross(std::fabs(z.y) > 0.99f ? vec3(1, 0, 0) : vec3(0, 1, 0), z));
        w.open_object();
        node_matrix(w, x, cross(z, x), z, vec3(l.position[0], l.position[1], l.position[2]));
        w.name("extensions").open_object().name("KHR_lights_punctual").open_object().field("light", i).close_object().close_object();
        w.close_object();
    }
    w.close_array();
    w.name("scenes").open_array().open_object().name("nodes").open_array();
    for (size_t i = 0; i < 2 + scene.lamps.size(); i++) w.integer(int64_t(i));
    w.close_array().close_object().close_array();
    w.field("scene", 0);
    w.close_object();
    return w.str();
}

}  // namespace

bool load_gltf(const std::string& path, Scene& scene, std::string& error) {
    try {
        GltfImport(scene, path).run();
    } catch (const Failure& f) {
        error = f.msg;
        return false;
    }
    if (scene.label.empty()) scene.label = utf8_path(path).stem().string();
    if (!scene.file_views.empty()) scene.view = scene.file_views.front();
    return true;
}

bool save_gltf(const std::string& path, const Scene& scene, std::string& error) {
    const std::string text = gltf_text(scene);
    std::ofstream out(utf8_path(path), std::ios::binary);
    if (out) out.write(text.data(), std::streamsize(text.size()));
    if (!out) {
        error = "cannot write " + path;
        return false;
    }
    return true;
}

}  // namespace phh
