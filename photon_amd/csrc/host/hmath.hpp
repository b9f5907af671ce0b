// Host-side vector math for the scene producers (BVH build, light tables, generators).
//
// The arithmetic follows the reference's core/math.h and core/sampling.h operation by operation
// (src/core/math.h:52-150, src/core/sampling.h:9-99): the BVH's SAH decisions and the light tables are
// float computations whose last bits decide the tree shape and therefore the BVH-order triangle IDs,
// so every sum and product here is evaluated in the same order with no contraction
// (built with -ffp-contract=off).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace phh {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInvTwoPi = 0.15915494309189533577f;
constexpr float kInf = 1e30f;

struct vec2 {
    float x = 0, y = 0;
    vec2() = default;
    constexpr vec2(float a, float b) : x(a), y(b) {}
};

struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    constexpr vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit constexpr vec3(float s) : x(s), y(s), z(s) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct vec4 {
    float x = 0, y = 0, z = 0, w = 0;
    vec4() = default;
    constexpr vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    constexpr vec4(vec3 v, float d) : x(v.x), y(v.y), z(v.z), w(d) {}
    vec3 xyz() const { return {x, y, z}; }
};

inline vec2 operator+(vec2 a, vec2 b) { return {a.x + b.x, a.y + b.y}; }
inline vec2 operator-(vec2 a, vec2 b) { return {a.x - b.x, a.y - b.y}; }
inline vec2 operator*(vec2 a, float s) { return {a.x * s, a.y * s}; }

inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator/(vec3 a, float s) {
    float r = 1.0f / s;
    return {a.x * r, a.y * r, a.z * r};
}
inline vec3& operator+=(vec3& a, vec3 b) {
    a = a + b;
    return a;
}
inline vec3& operator*=(vec3& a, float s) {
    a = a * s;
    return a;
}

inline float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(vec3 a) { return std::sqrt(dot(a, a)); }
inline float length_squared(vec3 a) { return dot(a, a); }
inline vec3 normalize(vec3 a) {
    float l2 = dot(a, a);
    return l2 > 0 ? a * (1.0f / std::sqrt(l2)) : vec3(0, 0, 1);
}
inline vec3 vmin(vec3 a, vec3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline vec3 vmax(vec3 a, vec3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline vec3 lerp(vec3 a, vec3 b, float t) { return a + (b - a) * t; }
inline float clampf(float v, float lo, float hi) { return std::fmin(std::fmax(v, lo), hi); }
inline float luminance(vec3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

inline vec4 operator+(vec4 a, vec4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline vec4 operator*(vec4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

// Tangent of an orthonormal basis around a unit normal (Duff et al. 2017), used for fallback tangents.
inline vec3 basis_tangent(vec3 n) {
    float sign = std::copysign(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    return vec3(1.0f + sign * n.x * n.x * a, sign * (n.x * n.y * a), -sign * n.x);
}

// PCG32 with the reference's seeding (src/core/sampling.h:9-34); the generators draw from it.
struct Pcg32 {
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;
    Pcg32() = default;
    Pcg32(uint64_t seed, uint64_t stream) {
        state = 0;
        inc = (stream << 1u) | 1u;
        next_u32();
        state += seed;
        next_u32();
    }
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    float next() { return std::fmin(float(next_u32() >> 8) * (1.0f / 16777216.0f), 0.99999994f); }
};

inline float srgb_to_linear(float c) { return c <= 0.04045f ? c / 12.92f : std::pow((c + 0.055f) / 1.055f, 2.4f); }

// Equirectangular (u, v) in [0,1)^2 -> direction, +y up (src/core/sampling.h:94-99).
inline vec3 equirect_to_direction(vec2 uv) {
    float phi = (uv.x - 0.5f) * kTwoPi;
    float theta = uv.y * kPi;
    float s = std::sin(theta);
    return {s * std::sin(phi), std::cos(theta), -s * std::cos(phi)};
}

}  // namespace phh
