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

#include "image_io.hpp"
#include "json.hpp"

namespace phh {

namespace {

namespace fs = std::filesystem;
using json::Value;

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

bool base64_decode(const std::string& in, std::vector<uint8_t>& out) {
    out.reserve(out.size() + in.size() / 4 * 3 + 3);
    uint32_t acc = 0;
    int bits = 0;
    for (const char ch : in) {
        if (ch == '=') break;
        int v;
        if (ch >= 'A' && ch <= 'Z')
            v = ch - 'A';
        else if (ch >= 'a' && ch <= 'z')
            v = ch - 'a' + 26;
        else if (ch >= '0' && ch <= '9')
            v = ch - '0' + 52;
        else if (ch == '+')
            v = 62;
        else if (ch == '/')
            v = 63;
        else if (ch == '\n' || ch == '\r' || ch == ' ')
            continue;
        else
            return false;
        acc = (acc << 6) | uint32_t(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(uint8_t(acc >> bits));
        }
    }
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
size_t byte_size(const json::Value& v, double fallback, const char* what) {
    const double d = v.number(fallback);
    if (!(d >= 0.0) || d >= 1099511627776.0 || d != std::floor(d)) fail(std::string("invalid ") + what);
    return size_t(d);
}

// The stride between elements of an accessor: the view's byteStride, which glTF 2.0 bounds to [4, 252]
// and which cannot be shorter than one element, or the element size when the view has none.
size_t element_stride(const json::Value& view, size_t elem) {
    if (!view.has("byteStride")) return elem;
    const size_t stride = byte_size(view["byteStride"], 0, "byteStride");
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

class GltfImport {
public:
    GltfImport(Scene& s, const std::string& path) : scene_(s), path_(path) {}

    void run() {
        std::vector<uint8_t> file;
        if (!read_file(utf8_path(path_), file)) fail("cannot open " + path_);
        dir_ = utf8_path(path_).parent_path();
        std::string text;
        bool glb = false;
        if (file.size() >= 12 && std::memcmp(file.data(), "glTF", 4) == 0) {
            glb = true;
            split_glb(file, text);
        } else {
            text.assign(file.begin(), file.end());
        }
        try {
            doc_ = json::parse(text);
        } catch (const json::ParseError& e) {
            fail(e.what());
        }
        if (doc_["asset"]["version"].string().rfind('2', 0) != 0) fail("only glTF 2.0 is supported");
        load_buffers(glb);
        images_.resize(doc_["images"].size());
        image_state_.assign(images_.size(), 0);
        load_materials();
        const Value& sc = doc_["scenes"][size_t(doc_["scene"].integer(0))];
        const Xform root;
        if (sc.is_null()) {
            std::vector<bool> is_child(doc_["nodes"].size(), false);
            for (size_t i = 0; i < doc_["nodes"].size(); i++) {
                const Value& kids = doc_["nodes"][i]["children"];
                for (size_t k = 0; k < kids.size(); k++) {
                    const int c = kids[k].integer();
                    if (c >= 0 && size_t(c) < is_child.size()) is_child[size_t(c)] = true;
                }
            }
            for (size_t i = 0; i < is_child.size(); i++)
                if (!is_child[i]) node(int(i), root, 0);
        } else {
            const Value& roots = sc["nodes"];
            for (size_t i = 0; i < roots.size(); i++) node(roots[i].integer(), root, 0);
        }
    }

private:
    Scene& scene_;
    std::string path_;
    fs::path dir_;
    Value doc_;
    std::vector<uint8_t> glb_bin_;
    std::vector<std::vector<uint8_t>> buffers_;
    std::vector<Texture> images_;
    std::vector<int> image_state_;  // 0 not tried, 1 decoded, -1 failed
    std::map<std::pair<int, bool>, int> textures_;
    int material_base_ = 0;
    int default_material_ = -1;

    void split_glb(const std::vector<uint8_t>& file, std::string& text) {
        // Header (magic, version, length), then chunks of (length, type, data padded to 4 bytes).
        size_t off = 12;
        while (off + 8 <= file.size()) {
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
        const Value& bufs = doc_["buffers"];
        for (size_t i = 0; i < bufs.size(); i++) {
            const Value& b = bufs[i];
            std::vector<uint8_t> data;
            if (!b.has("uri")) {
                if (!glb) fail("buffer " + std::to_string(i) + " has no data");
                data = glb_bin_;
            } else {
                const std::string& uri = b["uri"].string();
                if (uri.rfind("data:", 0) == 0) {
                    const size_t comma = uri.find(',');
                    if (comma == std::string::npos || !base64_decode(uri.substr(comma + 1), data)) fail("invalid data URI");
                } else if (!read_file(dir_ / utf8_path(percent_decode(uri)), data)) {
                    fail("cannot read buffer " + uri);
                }
            }
            buffers_.push_back(std::move(data));
        }
    }

    const std::vector<uint8_t>& buffer_of(const Value& view) {
        const int b = view["buffer"].integer(0);
        if (b < 0 || size_t(b) >= buffers_.size()) fail("buffer view refers to a missing buffer");
        return buffers_[size_t(b)];
    }

    static int accessor_count(const Value& a) {
        const size_t n = byte_size(a["count"], 0, "accessor count");
        if (n > size_t(1) << 30) fail("accessor count out of range");
        return int(n);
    }

    static int component_size(int ct) { return (ct == 5126 || ct == 5125) ? 4 : (ct == 5122 || ct == 5123) ? 2 : 1; }

    // Accessor -> count x `want` floats (missing components stay 0).
    std::vector<float> floats(int accessor, int want) {
        const Value& a = doc_["accessors"][size_t(accessor)];
        if (a.is_null()) fail("missing accessor");
        const int count = accessor_count(a);
        static const std::map<std::string, int> widths = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4}, {"MAT4", 16}};
        const auto wit = widths.find(a["type"].string());
        if (wit == widths.end()) fail("unsupported accessor type");
        const int nc = wit->second;
        const int ct = a["componentType"].integer();
        const bool norm = a["normalized"].boolean();
        const int cs = component_size(ct);
        std::vector<float> out(size_t(std::max(count, 0)) * size_t(want), 0.0f);
        if (!a.has("bufferView")) return out;
        const Value& view = doc_["bufferViews"][size_t(a["bufferView"].integer())];
        const std::vector<uint8_t>& buf = buffer_of(view);
        const size_t base = byte_size(view["byteOffset"], 0, "byteOffset") + byte_size(a["byteOffset"], 0, "byteOffset");
        const size_t stride = element_stride(view, size_t(nc * cs));
        if (!span_fits(base, stride, count, size_t(nc * cs), buf.size())) fail("accessor out of the buffer's range");
        const int take = std::min(nc, want);
        for (int i = 0; i < count; i++) {
            const uint8_t* p = buf.data() + base + stride * size_t(i);
            for (int c = 0; c < take; c++) {
                const uint8_t* q = p + size_t(c * cs);
                float v;
                switch (ct) {
                    case 5126: std::memcpy(&v, q, 4); break;
                    case 5121: v = norm ? float(q[0]) / 255.0f : float(q[0]); break;
                    case 5120: v = norm ? std::max(float(int8_t(q[0])) / 127.0f, -1.0f) : float(int8_t(q[0])); break;
                    case 5123: {
                        uint16_t u;
                        std::memcpy(&u, q, 2);
                        v = norm ? float(u) / 65535.0f : float(u);
                        break;
                    }
                    case 5122: {
                        int16_t s;
                        std::memcpy(&s, q, 2);
                        v = norm ? std::max(float(s) / 32767.0f, -1.0f) : float(s);
                        break;
                    }
                    case 5125: {
                        uint32_t u;
                        std::memcpy(&u, q, 4);
                        v = float(u);
                        break;
                    }
                    default: fail("unsupported component type");
                }
                out[size_t(i) * size_t(want) + size_t(c)] = v;
            }
        }
        return out;
    }

    std::vector<int> indices(int accessor) {
        const Value& a = doc_["accessors"][size_t(accessor)];
        const int count = accessor_count(a);
        const int cs = a["componentType"].integer() == 5125 ? 4 : a["componentType"].integer() == 5123 ? 2 : 1;
        if (!a.has("bufferView")) fail("index accessor without data");
        const Value& view = doc_["bufferViews"][size_t(a["bufferView"].integer())];
        const std::vector<uint8_t>& buf = buffer_of(view);
        const size_t base = byte_size(view["byteOffset"], 0, "byteOffset") + byte_size(a["byteOffset"], 0, "byteOffset");
        const size_t stride = element_stride(view, size_t(cs));
        if (!span_fits(base, stride, count, size_t(cs), buf.size())) fail("index accessor out of range");
        std::vector<int> out(size_t(std::max(count, 0)));
        for (int i = 0; i < count; i++) {
            const uint8_t* p = buf.data() + base + stride * size_t(i);
            uint32_t v = p[0];
            if (cs == 2) {
                uint16_t u;
                std::memcpy(&u, p, 2);
                v = u;
            } else if (cs == 4) {
                std::memcpy(&v, p, 4);
            }
            out[size_t(i)] = int(v);
        }
        return out;
    }

    // Decodes glTF image `i` into images_[i]; false (with the reason) when it cannot.
    bool decode(int i, std::string& why) {
        const Value& img = doc_["images"][size_t(i)];
        Texture& tex = images_[size_t(i)];
        if (img.has("uri")) {
            const std::string& uri = img["uri"].string();
            if (uri.rfind("data:", 0) == 0) {
                std::vector<uint8_t> data;
                const size_t comma = uri.find(',');
                if (comma == std::string::npos || !base64_decode(uri.substr(comma + 1), data)) {
                    why = "invalid image data URI";
                    return false;
                }
                return decode_image(data.data(), data.size(), tex, why);
            }
            return load_image((dir_ / utf8_path(percent_decode(uri))).string(), tex, why);
        }
        if (img.has("bufferView")) {
            const Value& view = doc_["bufferViews"][size_t(img["bufferView"].integer())];
            const int b = view["buffer"].integer(0);
            if (b < 0 || size_t(b) >= buffers_.size()) {
                why = "image buffer view refers to a missing buffer";
                return false;
            }
            const std::vector<uint8_t>& buf = buffers_[size_t(b)];
            const double od = view["byteOffset"].number(0), ld = view["byteLength"].number(0);
            if (!(od >= 0 && ld >= 0 && od + ld <= double(buf.size()))) {
                why = "image buffer view out of range";
                return false;
            }
            return decode_image(buf.data() + size_t(od), size_t(ld), tex, why);
        }
        why = "image without data";
        return false;
    }

    // Scene texture of a material's texture reference in a colour space; -1 when absent or unusable.
    int texture(const Value& ref, bool srgb) {
        if (!ref.is_object() || !ref.has("index")) return -1;
        const int ti = ref["index"].integer();
        const auto key = std::make_pair(ti, srgb);
        if (const auto it = textures_.find(key); it != textures_.end()) return it->second;
        const Value& t = doc_["textures"][size_t(ti)];
        int source = t["source"].integer(-1);
        if (source < 0 && t["extensions"].has("EXT_texture_webp")) source = t["extensions"]["EXT_texture_webp"]["source"].integer(-1);
        if (source < 0 || size_t(source) >= images_.size()) return -1;
        if (image_state_[size_t(source)] == 0) {
            std::string why;
            image_state_[size_t(source)] = decode(source, why) ? 1 : -1;
            if (image_state_[size_t(source)] < 0) std::fprintf(stderr, "warning: texture skipped: %s\n", why.c_str());
        }
        if (image_state_[size_t(source)] < 0) {
            textures_[key] = -1;
            return -1;
        }
        Texture tex = images_[size_t(source)];
        tex.srgb = srgb;
        const Value& sampler = doc_["samplers"][size_t(t["sampler"].integer(-1))];
        tex.repeat = sampler["wrapS"].integer(10497) != 33071;  // CLAMP_TO_EDGE clamps, the others repeat
        const int id = int(scene_.textures.size());
        scene_.textures.push_back(std::move(tex));
        textures_[key] = id;
        return id;
    }

    void load_materials() {
        const Value& mats = doc_["materials"];
        material_base_ = int(scene_.materials.size());
        for (size_t i = 0; i < mats.size(); i++) {
            const Value& m = mats[i];
            const Value& pbr = m["pbrMetallicRoughness"];
            const Value& ext = m["extensions"];
            ph_material mat = default_material();
            if (pbr.has("baseColorFactor"))
                for (int k = 0; k < 4; k++) mat.base_color[k] = float(pbr["baseColorFactor"][size_t(k)].number(1.0));
            mat.metallic = float(pbr["metallicFactor"].number(1.0));
            mat.roughness = float(pbr["roughnessFactor"].number(1.0));
            mat.base_color_tex = texture(pbr["baseColorTexture"], true);
            mat.metallic_roughness_tex = texture(pbr["metallicRoughnessTexture"], false);
            mat.normal_tex = texture(m["normalTexture"], false);
            mat.normal_scale = float(m["normalTexture"]["scale"].number(1.0));
            mat.emissive_tex = texture(m["emissiveTexture"], true);
            const float strength = float(ext["KHR_materials_emissive_strength"]["emissiveStrength"].number(1.0));
            for (int k = 0; k < 3; k++) mat.emission[k] = (m.has("emissiveFactor") ? float(m["emissiveFactor"][size_t(k)].number(0.0)) : 0.0f) * strength;
            mat.transmission = float(ext["KHR_materials_transmission"]["transmissionFactor"].number(0.0));
            mat.ior = float(ext["KHR_materials_ior"]["ior"].number(1.5));
            mat.double_sided = m["doubleSided"].boolean() ? 1 : 0;
            const std::string& mode = m["alphaMode"].string();
            mat.alpha_mode = mode == "MASK" ? PH_ALPHA_MASK : mode == "BLEND" ? PH_ALPHA_BLEND : PH_ALPHA_OPAQUE;
            mat.alpha_cutoff = mode == "BLEND" ? 0.5f : float(m["alphaCutoff"].number(0.5));
            scene_.materials.push_back(mat);
        }
    }

    int fallback_material() {
        if (default_material_ < 0) {
            ph_material m = default_material();
            m.base_color[0] = m.base_color[1] = m.base_color[2] = 0.8f;
            m.metallic = 0.0f;
            m.roughness = 0.6f;
            default_material_ = scene_.add_material(m);
        }
        return default_material_;
    }

    static vec3 vec3_of(const Value& a, float fallback) {
        return vec3(float(a[0].number(fallback)), float(a[1].number(fallback)), float(a[2].number(fallback)));
    }

    static Xform local_xform(const Value& nd) {
        if (nd.has("matrix")) {
            Xform x;
            const Value& a = nd["matrix"];
            for (int c = 0; c < 4; c++)
                for (int r = 0; r < 4; r++) x.m[r][c] = float(a[size_t(c * 4 + r)].number(r == c ? 1.0 : 0.0));
            return x;
        }
        const vec3 t = nd.has("translation") ? vec3_of(nd["translation"], 0.0f) : vec3(0.0f);
        const vec3 s = nd.has("scale") ? vec3_of(nd["scale"], 1.0f) : vec3(1.0f);
        vec4 q(0, 0, 0, 1);
        if (nd.has("rotation")) {
            const Value& r = nd["rotation"];
            q = vec4(float(r[0].number(0)), float(r[1].number(0)), float(r[2].number(0)), float(r[3].number(1)));
        }
        return trs(t, q, s);
    }

    void primitive(const Value& prim, const Xform& world) {
        const int mode = prim["mode"].integer(4);
        if (mode != 4 && mode != 5 && mode != 6) return;
        const Value& attrs = prim["attributes"];
        if (!attrs.has("POSITION")) return;
        const std::vector<float> pos = floats(attrs["POSITION"].integer(), 3);
        const std::vector<float> nrm = attrs.has("NORMAL") ? floats(attrs["NORMAL"].integer(), 3) : std::vector<float>{};
        const std::vector<float> uv = attrs.has("TEXCOORD_0") ? floats(attrs["TEXCOORD_0"].integer(), 2) : std::vector<float>{};
        const std::vector<float> tan = attrs.has("TANGENT") ? floats(attrs["TANGENT"].integer(), 4) : std::vector<float>{};
        const size_t nv = pos.size() / 3;
        std::vector<int> idx;
        if (prim.has("indices")) {
            idx = indices(prim["indices"].integer());
        } else {
            idx.resize(nv);
            for (size_t i = 0; i < nv; i++) idx[i] = int(i);
        }
        if (mode != 4) {
            std::vector<int> list;
            for (size_t i = 2; i < idx.size(); i++) {
                if (mode == 6)
                    list.insert(list.end(), {idx[0], idx[i - 1], idx[i]});
                else if (i % 2 == 0)
                    list.insert(list.end(), {idx[i - 2], idx[i - 1], idx[i]});
                else
                    list.insert(list.end(), {idx[i - 1], idx[i - 2], idx[i]});
            }
            idx.swap(list);
        }
        for (const int i : idx)
            if (i < 0 || size_t(i) >= nv) fail("vertex index out of range");
        const bool mirrored = det3(world) < 0.0f;
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
        if (mirrored)
            for (size_t i = 0; i + 2 < idx.size(); i += 3) std::swap(idx[i + 1], idx[i + 2]);
        const int material = prim.has("material") ? material_base_ + prim["material"].integer() : fallback_material();
        const size_t vbase = scene_.positions.size();
        scene_.add_mesh(p, idx, material, n, t);
        if (!tan.empty()) {
            scene_.tangents.resize(vbase, vec4(1, 0, 0, 1));
            for (size_t i = 0; i < nv; i++) {
                const vec3 tv = normalize(apply_dir(world, vec3(tan[4 * i], tan[4 * i + 1], tan[4 * i + 2])));
                scene_.tangents.push_back(vec4(tv, tan[4 * i + 3] * (mirrored ? -1.0f : 1.0f)));
            }
        } else if (!scene_.tangents.empty()) {
            scene_.tangents.resize(scene_.positions.size(), vec4(1, 0, 0, 1));
        }
    }

    void node(int index, const Xform& parent, int depth) {
        if (depth > 64) fail("node hierarchy too deep (or cyclic)");
        const Value& nd = doc_["nodes"][size_t(index)];
        const Xform world = parent * local_xform(nd);
        if (nd.has("mesh")) {
            const Value& prims = doc_["meshes"][size_t(nd["mesh"].integer())]["primitives"];
            for (size_t i = 0; i < prims.size(); i++) primitive(prims[i], world);
        }
        if (nd.has("camera")) {
            const Value& cam = doc_["cameras"][size_t(nd["camera"].integer())];
            if (cam["type"].string() == "perspective") {
                Camera c;
                c.position = apply_point(world, vec3(0.0f));
                c.target = c.position + normalize(apply_dir(world, vec3(0, 0, -1)));
                c.up = normalize(apply_dir(world, vec3(0, 1, 0)));
                c.vfov = float(cam["perspective"]["yfov"].number(0.8)) * 180.0f / kPi;
                scene_.cameras.push_back(c);
            }
        }
        const Value& lref = nd["extensions"]["KHR_lights_punctual"];
        if (lref.has("light")) {
            const Value& l = doc_["extensions"]["KHR_lights_punctual"]["lights"][size_t(lref["light"].integer())];
            const vec3 colour = l.has("color") ? vec3_of(l["color"], 1.0f) : vec3(1.0f);
            const vec3 power = colour * float(l["intensity"].number(1.0));
            const vec3 at = apply_point(world, vec3(0.0f));
            const vec3 dir = normalize(apply_dir(world, vec3(0, 0, -1)));
            const std::string& type = l["type"].string();
            ph_punctual_light pl{};
            pl.type = type == "directional" ? PH_LIGHT_DIRECTIONAL : type == "spot" ? PH_LIGHT_SPOT : PH_LIGHT_POINT;
            const float p3[3] = {at.x, at.y, at.z}, d3[3] = {dir.x, dir.y, dir.z}, i3[3] = {power.x, power.y, power.z};
            for (int k = 0; k < 3; k++) {
                pl.position[k] = p3[k];
                pl.direction[k] = d3[k];
                pl.intensity[k] = i3[k];
            }
            pl.cos_inner = std::cos(float(l["spot"]["innerConeAngle"].number(0.0)));
            pl.cos_outer = std::cos(float(l["spot"]["outerConeAngle"].number(double(kPi / 4))));
            scene_.punctual.push_back(pl);
        }
        const Value& kids = nd["children"];
        for (size_t i = 0; i < kids.size(); i++) node(kids[i].integer(), world, depth + 1);
    }
};

// Appends raw bytes to the export buffer, padded to 4 bytes; returns their offset.
size_t put(std::vector<uint8_t>& bin, const void* p, size_t n) {
    const size_t off = bin.size();
    const uint8_t* b = static_cast<const uint8_t*>(p);
    bin.insert(bin.end(), b, b + n);
    while (bin.size() % 4) bin.push_back(0);
    return off;
}

}  // namespace

bool load_gltf(const std::string& path, Scene& scene, std::string& error) {
    try {
        GltfImport(scene, path).run();
    } catch (const Failure& f) {
        error = f.msg;
        return false;
    }
    if (scene.name.empty()) scene.name = utf8_path(path).stem().string();
    if (!scene.cameras.empty()) scene.camera = scene.cameras.front();
    return true;
}

bool save_gltf(const std::string& path, const Scene& scene, std::string& error) {
    const size_t nv = scene.positions.size();
    const bool has_n = scene.normals.size() == nv && nv > 0;
    const bool has_uv = scene.uvs.size() == nv && nv > 0;
    // One primitive per material, in material order. Each carries the vertices it references, in order
    // of first use, so that a file of many materials does not repeat the whole vertex array per
    // primitive (a vertex shared by several materials is written once for each).
    struct Group {
        std::vector<uint32_t> vertices;  // scene vertex ids
        std::vector<uint32_t> indices;   // into `vertices`
        size_t first = 0;                // first of its vertices in the written arrays
    };
    std::map<int, Group> groups;
    std::map<int, std::map<uint32_t, uint32_t>> shared;  // local ids of vertices used by a second group
    std::vector<int> owner(nv, -1);
    std::vector<uint32_t> local(nv, 0);
    for (int t = 0; t < scene.num_tris(); t++) {
        const int m = scene.tri_material[size_t(t)];
        Group& g = groups[m];
        for (int k = 0; k < 3; k++) {
            const uint32_t v = uint32_t(scene.indices[size_t(t) * 3 + size_t(k)]);
            if (owner[v] < 0) {
                owner[v] = m;
                local[v] = uint32_t(g.vertices.size());
                g.vertices.push_back(v);
            }
            if (owner[v] == m) {
                g.indices.push_back(local[v]);
                continue;
            }
            auto& sh = shared[m];
            auto it = sh.find(v);
            if (it == sh.end()) {
                it = sh.emplace(v, uint32_t(g.vertices.size())).first;
                g.vertices.push_back(v);
            }
            g.indices.push_back(it->second);
        }
    }
    std::vector<vec3> pos, nrm;
    std::vector<vec2> uvs;
    for (auto& [m, g] : groups) {
        g.first = pos.size();
        for (const uint32_t v : g.vertices) {
            pos.push_back(scene.positions[v]);
            if (has_n) nrm.push_back(scene.normals[v]);
            if (has_uv) uvs.push_back(scene.uvs[v]);
        }
    }
    const size_t nw = pos.size();
    std::vector<uint8_t> bin;
    const size_t pos_off = put(bin, pos.data(), nw * sizeof(vec3));
    const size_t nrm_off = has_n ? put(bin, nrm.data(), nw * sizeof(vec3)) : 0;
    const size_t uv_off = has_uv ? put(bin, uvs.data(), nw * sizeof(vec2)) : 0;
    std::vector<std::pair<size_t, size_t>> index_views;  // (offset, count)
    for (const auto& [m, g] : groups) index_views.push_back({put(bin, g.indices.data(), g.indices.size() * 4), g.indices.size()});

    std::ostringstream o;
    o.precision(9);
    auto v3 = [&](vec3 v) {
        std::ostringstream s;
        s.precision(9);
        s << "[" << v.x << ", " << v.y << ", " << v.z << "]";
        return s.str();
    };
    o << "{\n\"asset\": {\"version\": \"2.0\", \"generator\": \"photon_amd\"},\n";
    o << "\"extensionsUsed\": [\"KHR_materials_emissive_strength\", \"KHR_materials_transmission\", \"KHR_materials_ior\", \"KHR_lights_punctual\"],\n";
    o << "\"buffers\": [{\"byteLength\": " << bin.size() << ", \"uri\": \"data:application/octet-stream;base64," << base64_encode(bin) << "\"}],\n";
    // Views: positions, normals and uvs when present, then one index list per group. Accessors: per
    // group its ranges of the vertex views (positions with their bounds), then its indices.
    std::vector<std::string> views, accessors;
    auto view = [&](size_t off, size_t len) {
        views.push_back("{\"buffer\": 0, \"byteOffset\": " + std::to_string(off) + ", \"byteLength\": " + std::to_string(len) + "}");
        return int(views.size()) - 1;
    };
    auto accessor = [&](int v, size_t byte_off, int ct, size_t count, const char* type, const std::string& extra) {
        accessors.push_back("{\"bufferView\": " + std::to_string(v) + ", \"byteOffset\": " + std::to_string(byte_off) + ", \"componentType\": " +
                            std::to_string(ct) + ", \"count\": " + std::to_string(count) + ", \"type\": \"" + type + "\"" + extra + "}");
        return int(accessors.size()) - 1;
    };
    const int pos_view = view(pos_off, nw * 12);
    const int nrm_view = has_n ? view(nrm_off, nw * 12) : -1;
    const int uv_view = has_uv ? view(uv_off, nw * 8) : -1;
    struct Prim {
        int material, pos, nrm, uv, idx;
    };
    std::vector<Prim> prims;
    size_t gi = 0;
    for (const auto& [m, g] : groups) {
        vec3 lo(kInf), hi(-kInf);
        for (const uint32_t v : g.vertices) {
            lo = vmin(lo, scene.positions[v]);
            hi = vmax(hi, scene.positions[v]);
        }
        const size_t n = g.vertices.size();
        Prim pr{m, accessor(pos_view, g.first * 12, 5126, n, "VEC3", ", \"min\": " + v3(lo) + ", \"max\": " + v3(hi)), -1, -1, -1};
        if (has_n) pr.nrm = accessor(nrm_view, g.first * 12, 5126, n, "VEC3", "");
        if (has_uv) pr.uv = accessor(uv_view, g.first * 8, 5126, n, "VEC2", "");
        pr.idx = accessor(view(index_views[gi].first, index_views[gi].second * 4), 0, 5125, index_views[gi].second, "SCALAR", "");
        prims.push_back(pr);
        gi++;
    }
    auto list = [&](const std::vector<std::string>& items) {
        std::string s;
        for (size_t i = 0; i < items.size(); i++) s += (i ? ", " : "") + items[i];
        return s;
    };
    o << "\"bufferViews\": [" << list(views) << "],\n\"accessors\": [" << list(accessors) << "],\n\"materials\": [";
    for (size_t i = 0; i < scene.materials.size(); i++) {
        const ph_material& m = scene.materials[i];
        // Emission splits into a factor of at most 1 and a strength.
        const float strength = std::max({m.emission[0], m.emission[1], m.emission[2], 1.0f});
        const vec3 factor = vec3(m.emission[0], m.emission[1], m.emission[2]) / strength;
        o << (i ? ", " : "") << "{\"pbrMetallicRoughness\": {\"baseColorFactor\": [" << m.base_color[0] << ", " << m.base_color[1] << ", " << m.base_color[2]
          << ", " << m.base_color[3] << "], \"metallicFactor\": " << m.metallic << ", \"roughnessFactor\": " << m.roughness << "}, \"emissiveFactor\": " << v3(factor)
          << ", \"doubleSided\": " << (m.double_sided ? "true" : "false") << ", \"extensions\": {\"KHR_materials_emissive_strength\": {\"emissiveStrength\": "
          << strength << "}, \"KHR_materials_transmission\": {\"transmissionFactor\": " << m.transmission << "}, \"KHR_materials_ior\": {\"ior\": " << m.ior
          << "}}}";
    }
    o << "],\n\"meshes\": [{\"primitives\": [";
    for (size_t i = 0; i < prims.size(); i++) {
        const Prim& pr = prims[i];
        o << (i ? ", " : "") << "{\"attributes\": {\"POSITION\": " << pr.pos;
        if (pr.nrm >= 0) o << ", \"NORMAL\": " << pr.nrm;
        if (pr.uv >= 0) o << ", \"TEXCOORD_0\": " << pr.uv;
        o << "}, \"indices\": " << pr.idx << ", \"material\": " << pr.material << "}";
    }
    o << "]}],\n";
    const Camera& c = scene.camera;
    const vec3 f = normalize(c.target - c.position), r = normalize(cross(f, c.up)), u = cross(r, f);
    o << "\"cameras\": [{\"type\": \"perspective\", \"perspective\": {\"yfov\": " << c.vfov * kPi / 180.0f << ", \"znear\": 0.01}}],\n";
    o << "\"extensions\": {\"KHR_lights_punctual\": {\"lights\": [";
    for (size_t i = 0; i < scene.punctual.size(); i++) {
        const ph_punctual_light& l = scene.punctual[i];
        const float power = std::max({l.intensity[0], l.intensity[1], l.intensity[2], 1e-6f});
        const vec3 colour = vec3(l.intensity[0], l.intensity[1], l.intensity[2]) / power;
        const char* type = l.type == PH_LIGHT_DIRECTIONAL ? "directional" : l.type == PH_LIGHT_SPOT ? "spot" : "point";
        o << (i ? ", " : "") << "{\"type\": \"" << type << "\", \"color\": " << v3(colour) << ", \"intensity\": " << power;
        if (l.type == PH_LIGHT_SPOT) o << ", \"spot\": {\"innerConeAngle\": " << std::acos(l.cos_inner) << ", \"outerConeAngle\": " << std::acos(l.cos_outer) << "}";
        o << "}";
    }
    o << "]}},\n";
    // Node matrices, column major: x (right), y (up), z (backwards), translation.
    auto matrix = [&](vec3 x, vec3 y, vec3 z, vec3 t) {
        std::ostringstream s;
        s.precision(9);
        s << "[" << x.x << ", " << x.y << ", " << x.z << ", 0, " << y.x << ", " << y.y << ", " << y.z << ", 0, " << z.x << ", " << z.y << ", " << z.z << ", 0, " << t.x
          << ", " << t.y << ", " << t.z << ", 1]";
        return s.str();
    };
    o << "\"nodes\": [{\"mesh\": 0}, {\"camera\": 0, \"matrix\": " << matrix(r, u, -f, c.position) << "}";
    for (size_t i = 0; i < scene.punctual.size(); i++) {
        const ph_punctual_light& l = scene.punctual[i];
        const vec3 z = -normalize(vec3(l.direction[0], l.direction[1], l.direction[2]));
        const vec3 x = normalize(cross(std::fabs(z.y) > 0.99f ? vec3(1, 0, 0) : vec3(0, 1, 0), z));
        o << ", {\"matrix\": " << matrix(x, cross(z, x), z, vec3(l.position[0], l.position[1], l.position[2])) << ", \"extensions\": {\"KHR_lights_punctual\": {\"light\": " << i
          << "}}}";
    }
    o << "],\n\"scenes\": [{\"nodes\": [0, 1";
    for (size_t i = 0; i < scene.punctual.size(); i++) o << ", " << i + 2;
    o << "]}],\n\"scene\": 0\n}\n";
    std::ofstream out(utf8_path(path), std::ios::binary);
    if (!out) {
        error = "cannot write " + path;
        return false;
    }
    out << o.str();
    if (!out) {
        error = "cannot write " + path;
        return false;
    }
    return true;
}

}  // namespace phh
