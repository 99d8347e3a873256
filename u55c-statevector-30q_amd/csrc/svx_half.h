// IEEE binary16 conversions for host code (matrix splitting for the split-FP16 matrix-core path).
// Plain integer arithmetic, so the result does not depend on the host CPU's F16C support.
#ifndef SVX_HALF_H
#define SVX_HALF_H
#include <cstdint>
#include <cstring>

namespace svx {

// float -> binary16, round to nearest, ties to even; overflow -> inf, NaN stays NaN
inline uint16_t f32_to_f16_bits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);   // >= 65520 rounds to inf
    if (ax <= 0x33000000u) return (uint16_t)sign;               // <= 2^-25 rounds to 0
    const int e = (int)(ax >> 23) - 127;
    const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    if (e < -14) {   // binary16 subnormal: units of 2^-24
        const int shift = 13 + (-14 - e);                        // 14 .. 24
        uint32_t r = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (r & 1u))) ++r;
        return (uint16_t)(sign | r);
    }
    uint32_t r = ((uint32_t)(e + 15) << 10) | ((mant >> 13) & 0x3ffu);
    const uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;     // a carry moves into the exponent
    return (uint16_t)(sign | r);
}

inline float f16_bits_to_f32(uint16_t hbits) {
    const uint32_t sign = (uint32_t)(hbits & 0x8000u) << 16;
    const uint32_t e = (hbits >> 10) & 0x1fu, m = hbits & 0x3ffu;
    uint32_t x;
    if (e == 0x1fu) {
        x = sign | 0x7f800000u | (m << 13);
    } else if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {   // subnormal: m * 2^-24, exact in float
            float v = (float)m * (1.0f / 16777216.0f);
            std::memcpy(&x, &v, 4);
            x |= sign;
        }
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float out;
    std::memcpy(&out, &x, 4);
    return out;
}

}  // namespace svx
#endif
