// Device arithmetic for gfx950: vectors, the random sequences and the warps of the integrator.
//
// Bit-level agreement with the CPU build is a requirement of this library (primary-hit triangle IDs
// identical, radiance within 0.1% on 99.9% of pixels), so the rules are:
//   * every float expression is evaluated in the order the reference writes it
//     (src/core/math.h:52-124, src/core/sampling.h:9-99), built with -ffp-contract=off;
//   * division and square root are IEEE correctly rounded (-fhip-fp32-correctly-rounded-divide-sqrt);
//   * sin, cos, atan2, acos, exp and pow are evaluated in fp64 and rounded once to fp32, which gives
//     the correctly rounded fp32 value except when the fp64 result lies within one fp64 ulp of an fp32
//     rounding boundary (probability ~2^-29 per call). glibc's fp32 functions used by the CPU build are
//     correctly rounded on all but a tiny fraction of inputs, so the two agree almost everywhere; fp64
//     runs at half the fp32 rate on CDNA4, and these calls are a few per path vertex.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define PH_DEV __device__ __forceinline__

namespace phd {

// A load through a pointer known to address device (global) memory. Pointers read from memory are
// generic to the compiler, which would otherwise issue flat loads for them. Structs are read word by
// word.
template <class T>
PH_DEV T ldg(const T* p) {
    if constexpr (std::is_arithmetic_v<T>) {
        return *(const T __attribute__((address_space(1)))*)p;
    } else {
        static_assert(sizeof(T) % 4 == 0, "ldg reads whole words");
        T v;
        const uint32_t __attribute__((address_space(1)))* src = (const uint32_t __attribute__((address_space(1)))*)p;
        uint32_t* dst = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
        for (int i = 0; i < int(sizeof(T) / 4); i++) dst[i] = src[i];
        return v;
    }
}

PH_DEV float4 ldg(const float4* p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 v = *(const v4 __attribute__((address_space(1)))*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

// A pointer whose value is the same in every lane, moved to scalar registers.
template <class T>
PH_DEV const T* uniform_ptr(const T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v)), hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<const T*>((uint64_t(hi) << 32) | lo);
}


constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInvTwoPi = 0.15915494309189533577f;
constexpr float kInf = 1e30f;

// ---- correctly rounded transcendentals -----------------------------------------------------------
PH_DEV float sin_cr(float x) { return float(::sin(double(x))); }
PH_DEV float cos_cr(float x) { return float(::cos(double(x))); }
PH_DEV void sincos_cr(float x, float* s, float* c) {
    double sd, cd;
    ::sincos(double(x), &sd, &cd);
    *s = float(sd);
    *c = float(cd);
}
PH_DEV float atan2_cr(float y, float x) { return float(::atan2(double(y), double(x))); }
PH_DEV float acos_cr(float x) { return float(::acos(double(x))); }
PH_DEV float exp_cr(float x) { return float(::exp(double(x))); }
PH_DEV float pow_cr(float x, float y) { return float(::pow(double(x), double(y))); }

// ---- vectors ------------------------------------------------------------------------------------
struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};
struct f4 {
    float x, y, z, w;
};

PH_DEV f2 mk2(float x, float y) { return {x, y}; }
PH_DEV f3 mk3(float x, float y, float z) { return {x, y, z}; }
PH_DEV f3 mk3(float s) { return {s, s, s}; }
PH_DEV f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
PH_DEV f3 xyz(float4 v) { return {v.x, v.y, v.z}; }
PH_DEV f4 mk4(f3 a, float w) { return {a.x, a.y, a.z, w}; }

PH_DEV f2 operator+(f2 a, f2 b) { return {a.x + b.x, a.y + b.y}; }
PH_DEV f2 operator*(f2 a, float s) { return {a.x * s, a.y * s}; }

PH_DEV f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PH_DEV f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PH_DEV f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
PH_DEV f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
PH_DEV f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PH_DEV f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
PH_DEV f3 operator/(f3 a, float s) {  // reciprocal, then products (math.h:65-68)
    float r = 1.0f / s;
    return {a.x * r, a.y * r, a.z * r};
}
PH_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PH_DEV f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PH_DEV float len2(f3 a) { return dot(a, a); }
PH_DEV f3 unit(f3 a) {
    float l2 = dot(a, a);
    return l2 > 0 ? a * (1.0f / __builtin_sqrtf(l2)) : mk3(0, 0, 1);
}
PH_DEV f3 vmin3(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
PH_DEV f3 vmax3(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
PH_DEV float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
PH_DEV float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
PH_DEV f3 clamp3(f3 v, float lo, float hi) { return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)}; }
PH_DEV float sat(float v) { return clampf(v, 0.0f, 1.0f); }
PH_DEV f3 mirror(f3 v, f3 n) { return n * (2.0f * dot(v, n)) - v; }
PH_DEV float lum(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
PH_DEV bool black(f3 c) { return c.x <= 0 && c.y <= 0 && c.z <= 0; }
PH_DEV bool finite3(f3 c) { return __builtin_isfinite(c.x) && __builtin_isfinite(c.y) && __builtin_isfinite(c.z); }
PH_DEV float sq(float x) { return x * x; }
PH_DEV float ssqrt(float x) { return __builtin_sqrtf(fmaxf(x, 0.0f)); }
PH_DEV f3 lerp3(f3 a, f3 b, float t) { return a + (b - a) * t; }

PH_DEV f4 operator+(f4 a, f4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
PH_DEV f4 operator*(f4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

// Orthonormal basis around a unit normal (Duff et al. 2017; math.h:127-137).
struct Onb {
    f3 t, b, n;
};
PH_DEV Onb onb(f3 n) {
    Onb f;
    f.n = n;
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float bb = n.x * n.y * a;
    f.t = mk3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
    f.b = mk3(bb, sign + n.y * n.y * a, -n.y);
    return f;
}
PH_DEV f3 to_local(const Onb& f, f3 v) { return {dot(v, f.t), dot(v, f.b), dot(v, f.n)}; }
PH_DEV f3 to_world(const Onb& f, f3 v) { return f.t * v.x + f.b * v.y + f.n * v.z; }

// ---- PCG32 and the seeding hash (sampling.h:9-44) ----------------------------------------------------
struct Pcg {
    uint64_t state, inc;
};
PH_DEV uint32_t pcg_next(Pcg& r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31));
}
PH_DEV Pcg pcg_seed(uint64_t seed, uint64_t stream) {
    Pcg r;
    r.state = 0;
    r.inc = (stream << 1u) | 1u;
    pcg_next(r);
    r.state += seed;
    pcg_next(r);
    return r;
}
PH_DEV float pcg_uniform(Pcg& r) { return fminf(float(pcg_next(r) >> 8) * (1.0f / 16777216.0f), 0.99999994f); }
PH_DEV f2 pcg_uniform2(Pcg& r) {
    float a = pcg_uniform(r);
    float b = pcg_uniform(r);
    return {a, b};
}
PH_DEV uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
// The generator of one (pixel, sample) pair (pixel.h:20-21).
PH_DEV Pcg pixel_rng(uint32_t pixel, uint32_t frame, uint32_t seed) {
    return pcg_seed(mix32(pixel ^ mix32(seed)) | (uint64_t(mix32(frame + 0x9e3779b9u)) << 32), pixel);
}

// ---- warps ------------------------------------------------------------------------------------------
PH_DEV f2 disk_concentric(f2 u) {
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
    float s, c;
    sincos_cr(phi, &s, &c);
    return {r * c, r * s};
}
PH_DEV f3 hemi_cosine(f2 u) {
    f2 d = disk_concentric(u);
    float z = ssqrt(1 - d.x * d.x - d.y * d.y);
    return {d.x, d.y, z};
}
PH_DEV float hemi_cosine_pdf(float c) { return fmaxf(c, 0.0f) * kInvPi; }
PH_DEV f3 sphere_uniform(f2 u) {
    float z = 1 - 2 * u.x;
    float r = ssqrt(1 - z * z);
    float phi = kTwoPi * u.y;
    float s, c;
    sincos_cr(phi, &s, &c);
    return {r * c, r * s, z};
}
PH_DEV f2 tri_uniform(f2 u) {
    float su = __builtin_sqrtf(u.x);
    return {1 - su, u.y * su};
}
PH_DEV float mis_power(float pa, float pb) {
    float a = pa * pa, b = pb * pb;
    return a + b > 0 ? a / (a + b) : 0.0f;
}
PH_DEV f2 dir_to_latlong(f3 d) {
    float u = atan2_cr(d.x, -d.z) * kInvTwoPi + 0.5f;
    float v = acos_cr(clampf(d.y, -1.0f, 1.0f)) * kInvPi;
    return {u - floorf(u), v};
}
PH_DEV f3 latlong_to_dir(f2 uv) {
    float phi = (uv.x - 0.5f) * kTwoPi;
    float theta = uv.y * kPi;
    float st, ct, sp, cp;
    sincos_cr(theta, &st, &ct);
    sincos_cr(phi, &sp, &cp);
    return {st * sp, ct, -st * cp};
}

// ---- colour (color.h:8-41) ---------------------------------------------------------------------------
PH_DEV float srgb_decode(float c) { return c <= 0.04045f ? c / 12.92f : pow_cr((c + 0.055f) / 1.055f, 2.4f); }
PH_DEV float srgb_encode(float c) {
    c = sat(c);
    return c <= 0.0031308f ? c * 12.92f : 1.055f * pow_cr(c, 1.0f / 2.4f) - 0.055f;
}
PH_DEV f3 aces(f3 c) {
    f3 v = mk3(0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z, 0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z,
               0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z);
    f3 a = v * (v + mk3(0.0245786f)) - mk3(0.000090537f);
    f3 b = v * (v * 0.983729f + mk3(0.4329510f)) + mk3(0.238081f);
    v = a / b;
    return clamp3(mk3(1.60475f * v.x - 0.53108f * v.y - 0.07367f * v.z, -0.10208f * v.x + 1.10813f * v.y - 0.00605f * v.z,
                      -0.00327f * v.x - 0.07276f * v.y + 1.07602f * v.z),
                  0.0f, 1.0f);
}
PH_DEV f3 tonemap(f3 c, float exposure, int op) {
    c = c * exposure;
    if (op == 1) return aces(c);
    if (op == 2) return c / (mk3(1.0f) + c);
    return clamp3(c, 0.0f, 1.0f);
}
PH_DEV uint32_t quant8(float v) { return uint32_t(sat(v) * 255.0f + 0.5f); }
PH_DEV uint32_t pack8(f3 c) {
    return quant8(srgb_encode(c.x)) | (quant8(srgb_encode(c.y)) << 8) | (quant8(srgb_encode(c.z)) << 16) | 0xFF000000u;
}

}  // namespace phd
