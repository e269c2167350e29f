// Scalar single-precision transcendental functions that give bit-identical
// results on the host (x86-64, g++/clang) and on gfx950.
//
// The integrator's sampling decisions (lobe selection, Russian-roulette-like
// early outs, hit/miss at triangle edges) depend on sin/cos/log/exp results.
// libm on the host and ocml on the device differ in the last ulp, which would
// let a CPU restatement and the GPU kernels take different branches on the
// same random numbers.  Every function here is built only from IEEE-754
// correctly rounded operations (+, -, *, /, sqrt) and integer bit
// manipulation, so it evaluates identically on both sides as long as the
// translation unit is compiled with -ffp-contract=off (no FMA contraction).
//
// Methods (all public, published):
//   sin/cos : Cody & Waite (1980) three-part reduction by pi/4 and the
//             single-precision minimax polynomials of S. L. Moshier's Cephes
//             library (sinf.c / cosf.c).  Accurate for |x| <= 8192.
//   exp     : Cody-Waite reduction by ln 2 and the Cephes expf polynomial.
//   log     : frexp-style split at sqrt(1/2) and the Cephes logf polynomial.
//   asin/acos: Cephes asinf polynomial with the half-angle identity for
//             |x| > 0.5; acos(x) from asin via pi/2 - asin(x) and
//             2*asin(sqrt((1-x)/2)).
//   pow     : exp(y * log(x)) for x > 0 (only used with small integral y).
// Errors against glibc's correctly rounded double evaluations are checked by
// the unit tests (a few ulp at most over the argument ranges the integrator
// uses).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define TT_HD __host__ __device__ __forceinline__
#else
#define TT_HD static inline
#endif

namespace ttm {

TT_HD uint32_t float_bits(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}

TT_HD float bits_float(uint32_t u) {
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}

TT_HD float sqrt(float x) { return __builtin_sqrtf(x); }
TT_HD float abs(float x) { return bits_float(float_bits(x) & 0x7fffffffu); }
TT_HD bool isnan(float x) { return (float_bits(x) & 0x7fffffffu) > 0x7f800000u; }
TT_HD bool isinf(float x) { return (float_bits(x) & 0x7fffffffu) == 0x7f800000u; }

// 2^n for integer n in [-126, 127] exactly.
TT_HD float exp2i(int n) { return bits_float(uint32_t(n + 127) << 23); }

// x * 2^n, exact when the result is a normal number.  Splits large shifts so
// that intermediate powers stay representable.
TT_HD float ldexp(float x, int n) {
    if (n > 127) {
        x = x * exp2i(127);
        n -= 127;
        if (n > 127) n = 127;
    } else if (n < -126) {
        x = x * exp2i(-126);
        n += 126;
        if (n < -126) n = -126;
    }
    return x * exp2i(n);
}

// sin and cos share the reduction.  j is the octant index (even), y the
// reduced argument in [-pi/4, pi/4].
TT_HD void reduce_quarter_pi(float x, int& j, float& y) {
    const float FOUR_OVER_PI = 1.27323954473516f;
    const float DP1 = 0.78515625f;
    const float DP2 = 2.4187564849853515625e-4f;
    const float DP3 = 3.77489497744594108e-8f;
    float ax = abs(x);
    j = int(ax * FOUR_OVER_PI);
    j = (j + 1) & ~1;
    float fj = float(j);
    y = ((ax - fj * DP1) - fj * DP2) - fj * DP3;
}

TT_HD float sin_poly(float y) {
    float z = y * y;
    return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * y + y;
}

TT_HD float cos_poly(float y) {
    float z = y * y;
    return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z *
            z -
        0.5f * z + 1.0f;
}

// floor(x) for |x| < 2^31.
TT_HD float floor_small(float x) {
    float t = float(int(x));
    return t > x ? t - 1.0f : t;
}

// sin(pi v) and cos(pi v) for |v| < 2^22.  v = h / 2 + r with h the nearest
// integer to 2v and |r| <= 1/4; v - h / 2 is exact (Sterbenz), so only the
// product pi r rounds, whatever the size of v.  Same polynomials as sin/cos.
TT_HD void sincospi(float v, float& s, float& c) {
    float h = floor_small(2.0f * v + 0.5f);
    float r = v - 0.5f * h;
    int q = int(h) & 3;
    float y = r * 3.14159265358979323846f;
    float sp = sin_poly(y), cp = cos_poly(y);
    s = q == 0 ? sp : (q == 1 ? cp : (q == 2 ? -sp : -cp));
    c = q == 0 ? cp : (q == 1 ? -sp : (q == 2 ? -cp : sp));
}

TT_HD float sin(float x) {
    if (isnan(x) || isinf(x)) return bits_float(0x7fc00000u);
    bool negative = x < 0.0f;
    int j;
    float y;
    reduce_quarter_pi(x, j, y);
    int octant = j & 7;
    if (octant > 3) {
        negative = !negative;
        octant -= 4;
    }
    float r = (octant == 2) ? cos_poly(y) : sin_poly(y);
    return negative ? -r : r;
}

TT_HD float cos(float x) {
    if (isnan(x) || isinf(x)) return bits_float(0x7fc00000u);
    int j;
    float y;
    reduce_quarter_pi(x, j, y);
    int octant = j & 7;
    bool negative = false;
    if (octant > 3) {
        negative = true;
        octant -= 4;
    }
    if (octant > 1) negative = !negative;
    float r = (octant == 2) ? sin_poly(y) : cos_poly(y);
    return negative ? -r : r;
}

TT_HD float exp(float x) {
    if (isnan(x)) return x;
    if (x > 88.7228391f) return bits_float(0x7f800000u);
    if (x < -103.278929903f) return 0.0f;
    const float LOG2E = 1.44269504088896341f;
    const float C1 = 0.693359375f;
    const float C2 = -2.12194440e-4f;
    // n = round(x / ln 2) with ties away from zero, computed without
    // library rounding functions.
    float t = x * LOG2E;
    int n = int(t + (t >= 0.0f ? 0.5f : -0.5f));
    float fn = float(n);
    float r = (x - fn * C1) - fn * C2;
    float z = r * r;
    float p = ((((1.9875691500e-4f * r + 1.3981999507e-3f) * r + 8.3334519073e-3f) * r +
                4.1665795894e-2f) *
                   r +
               1.6666665459e-1f) *
            r +
        5.0000001201e-1f;
    float e = p * z + r + 1.0f;
    return ldexp(e, n);
}

TT_HD float log(float x) {
    if (isnan(x)) return x;
    if (x < 0.0f) return bits_float(0x7fc00000u);
    if (x == 0.0f) return bits_float(0xff800000u);
    if (isinf(x)) return x;
    uint32_t bits = float_bits(x);
    int e = 0;
    if ((bits & 0x7f800000u) == 0) {
        // Subnormal: scale into the normal range first.
        x = x * 8388608.0f;
        bits = float_bits(x);
        e = -23;
    }
    e += int((bits >> 23) & 0xffu) - 126;
    // Mantissa in [0.5, 1).
    float m = bits_float((bits & 0x007fffffu) | 0x3f000000u);
    const float SQRTHF = 0.707106781186547524f;
    float xr;
    if (m < SQRTHF) {
        e -= 1;
        xr = m + m - 1.0f;
    } else {
        xr = m - 1.0f;
    }
    float z = xr * xr;
    float y = ((((((((7.0376836292e-2f * xr - 1.1514610310e-1f) * xr + 1.1676998740e-1f) * xr -
                   1.2420140846e-1f) *
                      xr +
                  1.4249322787e-1f) *
                     xr -
                 1.6668057665e-1f) *
                    xr +
                2.0000714765e-1f) *
                   xr -
               2.4999993993e-1f) *
                  xr +
              3.3333331174e-1f) *
        xr * z;
    float fe = float(e);
    y = y + -2.12194440e-4f * fe;
    y = y + -0.5f * z;
    float r = xr + y;
    r = r + 0.693359375f * fe;
    return r;
}

TT_HD float log2(float x) {
    const float LOG2E = 1.44269504088896341f;
    return log(x) * LOG2E;
}

// asin for |x| <= 1.
TT_HD float asin(float x) {
    if (isnan(x)) return x;
    float a = abs(x);
    if (a > 1.0f) return bits_float(0x7fc00000u);
    const float PIO2 = 1.5707963267948966192f;
    float z, s;
    bool flag = a > 0.5f;
    if (flag) {
        z = 0.5f * (1.0f - a);
        s = sqrt(z);
    } else {
        z = a * a;
        s = a;
    }
    float r = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z +
                7.4953002686e-2f) *
                   z +
               1.6666752422e-1f) *
            z * s +
        s;
    if (flag) {
        r = r + r;
        r = PIO2 - r;
    }
    return x < 0.0f ? -r : r;
}

TT_HD float acos(float x) {
    if (isnan(x)) return x;
    if (x < -1.0f || x > 1.0f) return bits_float(0x7fc00000u);
    const float PI = 3.14159265358979323846f;
    const float PIO2 = 1.5707963267948966192f;
    if (x < -0.5f) return PI - 2.0f * asin(sqrt(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * asin(sqrt(0.5f * (1.0f - x)));
    return PIO2 - asin(x);
}

// IEEE binary32 -> binary16 bits, round to nearest even (overflow to inf).
TT_HD uint16_t float_to_half(float f) {
    uint32_t x = float_bits(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t e = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (e == 0xffu) return uint16_t(sign | 0x7c00u | (man ? 0x200u : 0u));
    int he = int(e) - 127 + 15;
    if (he >= 0x1f) return uint16_t(sign | 0x7c00u);
    if (he <= 0) {
        if (he < -10) return uint16_t(sign);
        man |= 0x800000u;
        int shift = 14 - he;
        uint32_t hm = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
        return uint16_t(sign | hm);
    }
    uint32_t hm = man >> 13;
    uint32_t rem = man & 0x1fffu;
    uint32_t h = sign | (uint32_t(he) << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h++;
    return uint16_t(h);
}

// IEEE binary16 bits -> binary32 (exact).
TT_HD float half_to_float(uint16_t h) {
    uint32_t sign = uint32_t(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t man = h & 0x3ffu;
    if (e == 0) {
        if (man == 0) return bits_float(sign);
        // subnormal half: man * 2^-24
        float v = float(man) * 5.9604644775390625e-8f;
        return sign ? -v : v;
    }
    if (e == 31) return bits_float(sign | 0x7f800000u | (man << 13));
    return bits_float(sign | ((e + 112u) << 23) | (man << 13));
}

// Power for x >= 0.  GLSL leaves pow(x<0, y) undefined; this returns NaN.
TT_HD float pow(float x, float y) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : bits_float(0x7f800000u));
    if (x == 1.0f || y == 0.0f) return 1.0f;
    return exp(y * log(x));
}

}  // namespace ttm
