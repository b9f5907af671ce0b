// The compact G-buffer of the real-time workload (photon_hip.h, ph_compact_gbuffer): the shading normal
// as two fp16 octahedral coordinates in one uint32, and hit positions rebuilt from the primary hit
// distance along each pixel's centre ray.
//
// The fp16 conversions are integer code with round-to-nearest-even (subnormals kept), so that the
// device and any CPU restatement of the same steps produce the same bits without relying on a
// conversion instruction's rounding or denormal modes.
#pragma once

#include "dmath.h"
#include "photon_types.h"

namespace phd {

PH_DEV uint32_t f32_to_f16_bits(float f) {
    const uint32_t x = __float_as_uint(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t e = (x >> 23) & 0xffu;
    uint32_t m = x & 0x7fffffu;
    if (e == 0xffu) return sign | 0x7c00u | (m ? 0x200u : 0u);  // infinity, NaN
    const int exp = int(e) - 127 + 15;
    if (exp >= 31) return sign | 0x7c00u;  // overflow to infinity
    if (exp <= 0) {
        // Subnormal half: the 24-bit significand times 2^(exp - 14).
        if (exp < -10) return sign;
        m |= 0x800000u;
        const int shift = 14 - exp;
        uint32_t h = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (h & 1u))) h++;
        return sign | h;
    }
    uint32_t h = sign | (uint32_t(exp) << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // a carry moves into the exponent: still correct
    return h;
}

PH_DEV float f16_bits_to_f32(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {
            int k = 0;
            while (!(m & 0x400u)) {
                m <<= 1;
                k++;
            }
            x = sign | (uint32_t(113 - k) << 23) | ((m & 0x3ffu) << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    return __uint_as_float(x);
}

// Octahedral map of a unit vector to the square [-1, 1]^2 (a zero vector maps to the origin).
PH_DEV uint32_t oct_encode(f3 n) {
    const float s = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    if (!(s > 0.0f)) return 0u;
    float x = n.x / s, y = n.y / s;
    if (n.z < 0.0f) {
        const float ox = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        const float oy = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = ox;
        y = oy;
    }
    return f32_to_f16_bits(x) | (f32_to_f16_bits(y) << 16);
}

PH_DEV f3 oct_decode(uint32_t e) {
    float x = f16_bits_to_f32(e & 0xffffu), y = f16_bits_to_f32(e >> 16);
    const float z = 1.0f - fabsf(x) - fabsf(y);
    if (z < 0.0f) {
        const float ox = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        const float oy = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = ox;
        y = oy;
    }
    // |x| + |y| + |z| = 1 here, so the length is at least 1/sqrt(3).
    const float len = __builtin_sqrtf(x * x + y * y + z * z);
    return mk3(x / len, y / len, z / len);
}

// The same unit vector normalised with the hardware reciprocal square root (about 1 ulp from
// oct_decode's per component), for the denoiser's continuous edge-stopping weights, where a tap
// decodes the normal of every neighbour it weighs.
PH_DEV f3 oct_decode_approx(uint32_t e) {
    // The hardware's half-to-single conversion: exact for every half (subnormals too, under the default
    // mode that keeps half denormals; flushed, they would move the unit vector by less than 1e-4).
    float x = float(__builtin_bit_cast(_Float16, uint16_t(e & 0xffffu))), y = float(__builtin_bit_cast(_Float16, uint16_t(e >> 16)));
    const float z = 1.0f - fabsf(x) - fabsf(y);
    if (z < 0.0f) {
        const float ox = (1.0f - fabsf(y)) * (x >= 0.0f ? 1.0f : -1.0f);
        const float oy = (1.0f - fabsf(x)) * (y >= 0.0f ? 1.0f : -1.0f);
        x = ox;
        y = oy;
    }
    const float r = __builtin_amdgcn_rsqf(x * x + y * y + z * z);
    return mk3(x * r, y * r, z * r);
}

// Position at distance `depth` along the centre ray of pixel (x, y) of a pinhole view of `c` (the
// camera's own origin; the lens sample and the pixel jitter of the traced ray are not stored).
PH_DEV f3 centre_position(const ph_camera_view& c, int x, int y, int width, int height, float depth) {
    const float px = float(x) + 0.5f, py = float(y) + 0.5f;
    const float sx = (2.0f * px / float(width) - 1.0f) * c.tan_half_fov * c.aspect;
    const float sy = (1.0f - 2.0f * py / float(height)) * c.tan_half_fov;
    const f3 d = unit(ld3(c.forward) + ld3(c.right) * sx + ld3(c.up) * sy);
    return ld3(c.position) + d * depth;
}

}  // namespace phd
