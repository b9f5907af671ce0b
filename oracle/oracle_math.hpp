// ORACLE — test infrastructure only. Nothing in the product (photon_amd/) links, loads or calls this.
//
// Vector and sampling arithmetic of the CPU oracle, restating the reference's shared headers
// operation by operation so that results are bit-identical to the reference's own CPU renderer built
// with g++ (no FMA contraction, glibc's float libm):
//   vec ops, normalize, luminance, reflect ........ src/core/math.h:52-124
//   ONB (Duff et al. 2017) ......................... src/core/math.h:127-137, 148-149
//   PCG32 seeding/stepping, hash, warps, MIS, equirect ... src/core/sampling.h:9-99
//   sRGB / ACES / tone map / pack .................. src/core/color.h:8-41
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace orc {

// Threads of the oracle's parallel loops when the caller asks for the default: OMP_NUM_THREADS when set
// (the share of host cores a job was given), else every hardware thread.
inline int default_threads() {
    const char* e = std::getenv("OMP_NUM_THREADS");
    const int n = e ? std::atoi(e) : 0;
    return n > 0 ? n : int(std::thread::hardware_concurrency() ? std::thread::hardware_concurrency() : 1);
}

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInvTwoPi = 0.15915494309189533577f;
constexpr float kInf = 1e30f;

struct V2 {
    float x = 0, y = 0;
};
inline V2 v2(float x, float y) { return V2{x, y}; }
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }

struct V3 {
    float x = 0, y = 0, z = 0;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 v3s(float s) { return V3{s, s, s}; }
inline V3 v3p(const float* p) { return V3{p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
// Division by a scalar multiplies by its reciprocal (math.h:65-68).
inline V3 operator/(V3 a, float s) {
    float inv = 1.0f / s;
    return {a.x * inv, a.y * inv, a.z * inv};
}
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len2(V3 a) { return dot(a, a); }
inline V3 unit(V3 a) {
    float l2 = dot(a, a);
    return l2 > 0 ? a * (1.0f / sqrtf(l2)) : v3(0, 0, 1);
}
inline V3 vmin3(V3 a, V3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
inline V3 vmax3(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
inline float maxc(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
inline float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
inline V3 clamp3(V3 v, float lo, float hi) { return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)}; }
inline float sat(float v) { return clampf(v, 0.0f, 1.0f); }
inline V3 mirror(V3 v, V3 n) { return n * (2.0f * dot(v, n)) - v; }
inline float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
inline bool black(V3 c) { return c.x <= 0 && c.y <= 0 && c.z <= 0; }
inline bool finite3(V3 c) { return std::isfinite(c.x) && std::isfinite(c.y) && std::isfinite(c.z); }
inline float sq(float x) { return x * x; }
inline float ssqrt(float x) { return sqrtf(fmaxf(x, 0.0f)); }
inline V3 lerp3(V3 a, V3 b, float t) { return a + (b - a) * t; }

struct V4 {
    float x = 0, y = 0, z = 0, w = 0;
    V3 xyz() const { return {x, y, z}; }
};
inline V4 v4(float x, float y, float z, float w) { return V4{x, y, z, w}; }
inline V4 v4(V3 a, float w) { return V4{a.x, a.y, a.z, w}; }
inline V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

// Orthonormal basis around a unit normal.
struct Onb {
    V3 t, b, n;
    explicit Onb(V3 nn) : n(nn) {
        float sign = copysignf(1.0f, n.z);
        float a = -1.0f / (sign + n.z);
        float bb = n.x * n.y * a;
        t = v3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
        b = v3(bb, sign + n.y * n.y * a, -n.y);
    }
    V3 local(V3 v) const { return {dot(v, t), dot(v, b), dot(v, n)}; }
    V3 world(V3 v) const { return t * v.x + b * v.y + n * v.z; }
};

struct Pcg {
    uint64_t state, inc;
    Pcg(uint64_t seed, uint64_t stream) {
        state = 0;
        inc = (stream << 1u) | 1u;
        u32();
        state += seed;
        u32();
    }
    uint32_t u32() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    float uni() { return fminf(float(u32() >> 8) * (1.0f / 16777216.0f), 0.99999994f); }
    V2 uni2() {
        float a = uni();
        float b = uni();
        return {a, b};
    }
};

inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

inline V2 disk_concentric(V2 u) {
    float a = 2 * u.x - 1, b = 2 * u.y - 1;
    if (a == 0 && b == 0) return {0, 0};
    float r, phi;
    if (fabsf(a) > fabsf(b)) {
        r = a;
        phi = (kPi / 4) * (b / a);
    } else {
        r = b;
        phi = (kPi / 2) - (kPi / 4) * (a / b);
    }
    return {r * cosf(phi), r * sinf(phi)};
}

inline V3 hemi_cosine(V2 u) {
    V2 d = disk_concentric(u);
    float z = ssqrt(1 - d.x * d.x - d.y * d.y);
    return {d.x, d.y, z};
}

inline float hemi_cosine_pdf(float c) { return fmaxf(c, 0.0f) * kInvPi; }

inline V3 sphere_uniform(V2 u) {
    float z = 1 - 2 * u.x;
    float r = ssqrt(1 - z * z);
    float phi = kTwoPi * u.y;
    return {r * cosf(phi), r * sinf(phi), z};
}

inline V2 tri_uniform(V2 u) {
    float su = sqrtf(u.x);
    return {1 - su, u.y * su};
}

inline float mis_power(float pa, float pb) {
    float a = pa * pa, b = pb * pb;
    return a + b > 0 ? a / (a + b) : 0.0f;
}

inline V2 dir_to_latlong(V3 d) {
    float u = atan2f(d.x, -d.z) * kInvTwoPi + 0.5f;
    float v = acosf(clampf(d.y, -1.0f, 1.0f)) * kInvPi;
    return {u - floorf(u), v};
}

inline V3 latlong_to_dir(V2 uv) {
    float phi = (uv.x - 0.5f) * kTwoPi;
    float theta = uv.y * kPi;
    float s = sinf(theta);
    return {s * sinf(phi), cosf(theta), -s * cosf(phi)};
}

inline float srgb_decode(float c) { return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f); }
inline float srgb_encode(float c) {
    c = sat(c);
    return c <= 0.0031308f ? c * 12.92f : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}

inline V3 aces(V3 c) {
    V3 v = v3(0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z, 0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z,
              0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z);
    V3 a = v * (v + v3s(0.0245786f)) - v3s(0.000090537f);
    V3 b = v * (v * 0.983729f + v3s(0.4329510f)) + v3s(0.238081f);
    v = a / b;
    return clamp3(v3(1.60475f * v.x - 0.53108f * v.y - 0.07367f * v.z, -0.10208f * v.x + 1.10813f * v.y - 0.00605f * v.z,
                     -0.00327f * v.x - 0.07276f * v.y + 1.07602f * v.z),
                  0.0f, 1.0f);
}

inline V3 tonemap(V3 c, float exposure, int op) {
    c = c * exposure;
    if (op == 1) return aces(c);
    if (op == 2) return c / (v3s(1.0f) + c);
    return clamp3(c, 0.0f, 1.0f);
}

inline uint32_t pack8(V3 c) {
    auto q = [](float v) { return uint32_t(sat(v) * 255.0f + 0.5f); };
    return q(srgb_encode(c.x)) | (q(srgb_encode(c.y)) << 8) | (q(srgb_encode(c.z)) << 16) | 0xFF000000u;
}

inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace orc
