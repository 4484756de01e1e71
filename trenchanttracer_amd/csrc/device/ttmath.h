// Float math of the render stages, written so that host restatements compute
// the same bits: only IEEE +, -, *, /, sqrt, rint and bit casts, unfused.
//   sincos: Cody–Waite reduction by pi/2 (three-part constant 12 + 12 + 24 bits),
//           Taylor polynomials of sin to r^9 and cos to r^10 on |r| <= pi/4
//   log2:   x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)) to f^11
//   exp2:   2^n e^(g ln 2), |g| <= 1/2, Taylor to t^8
//   atan2:  octant reduction a = min(|x|,|y|) / max(|x|,|y|); above tan(pi/8)
//           atan a = pi/4 + atan((a-1)/(a+1)); Taylor series of atan to t^19 on
//           |t| <= tan(pi/8); pi/4, pi/2, pi as a float plus a float correction
//   acos:   half-angle identity acos x = 2 atan2(sqrt(1-x), sqrt(1+x))
// Coefficients are 1/n! and 1/(2k+1) rounded to float.
#pragma once

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace tt {

__device__ __forceinline__ float tm_sincos(float x, float *c_out) {
    float k = rintf(x * 0x1.45f306p-1f);
    float r = x - k * 0x1.92p+0f;
    r = r - k * 0x1.fb4p-12f;
    r = r - k * 0x1.4442d2p-24f;
    float r2 = r * r;
    float ps = -0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f));
    float s = r + r * r2 * ps;
    float pc = 0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f));
    float c = 1.0f - 0.5f * r2 + r2 * r2 * pc;
    int q = ((int)k) & 3;
    float sn, cs;
    if (q == 0) {
        sn = s;
        cs = c;
    } else if (q == 1) {
        sn = c;
        cs = -s;
    } else if (q == 2) {
        sn = -s;
        cs = -c;
    } else {
        sn = -c;
        cs = s;
    }
    *c_out = cs;
    return sn;
}

__device__ __forceinline__ float tm_log2(float x) {
    int e = 0;
    if (x < 0x1p-126f) {
        x = x * 0x1p24f;
        e = -24;
    }
    unsigned b = __float_as_uint(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float f = (m - 1.0f) / (m + 1.0f);
    float f2 = f * f;
    float p = f2 * (0x1.555556p-2f +
                    f2 * (0x1.99999ap-3f + f2 * (0x1.24924ap-3f + f2 * (0x1.c71c72p-4f + f2 * 0x1.745d18p-4f))));
    float ln_m = (f + f * p) * 2.0f;
    return (float)e + ln_m * 0x1.715476p+0f;
}

__device__ __forceinline__ float tm_exp2(float z) {
    if (z < -150.0f) return 0.0f;
    if (z >= 128.0f) return __builtin_inff();
    float n = rintf(z);
    float g = z - n;
    float t = g * 0x1.62e430p-1f;
    float p = 1.0f + t * (1.0f + t * (0.5f + t * (0x1.555556p-3f + t * (0x1.555556p-5f + t * (0x1.111112p-7f +
              t * (0x1.6c16c2p-10f + t * (0x1.a01a02p-13f + t * 0x1.a01a02p-16f)))))));
    int ni = (int)n;
    if (ni < -126) {
        p = p * 0x1p-64f;
        ni += 64;
    }
    return p * __uint_as_float((unsigned)(ni + 127) << 23);
}

// x in [0, 1], y > 0
__device__ __forceinline__ float tm_pow(float x, float y) {
    if (x == 0.0f) return 0.0f;
    return tm_exp2(y * tm_log2(x));
}

// C99 atan2 special values for zeros (atan2(+-0, +0) = +-0, atan2(+-0, -0) = +-pi);
// NaN propagates. Within 2.5 ulp of the exact value.
__device__ __forceinline__ float tm_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const bool swap = ay > ax;
    const float num = swap ? ax : ay, den = swap ? ay : ax;
    float r;
    if (den == 0.0f) {
        r = 0.0f;
    } else {
        const float a = num / den;
        float hi = 0.0f, lo = 0.0f, t = a;
        if (a > 0x1.a8279ap-2f) {
            t = (a - 1.0f) / (a + 1.0f);
            hi = 0x1.921fb6p-1f;
            lo = -0x1.777a5cp-26f;
        }
        const float t2 = t * t;
        const float p = t2 * (-0x1.555556p-2f + t2 * (0x1.99999ap-3f + t2 * (-0x1.24924ap-3f + t2 * (0x1.c71c72p-4f +
                        t2 * (-0x1.745d18p-4f + t2 * (0x1.3b13b2p-4f + t2 * (-0x1.111112p-4f + t2 * (0x1.e1e1e2p-5f +
                        t2 * -0x1.af286cp-5f))))))));
        r = hi + (lo + (t + t * p));
    }
    if (swap) r = 0x1.921fb6p+0f + (-0x1.777a5cp-25f - r);
    if (x < 0.0f || (den == 0.0f && (__float_as_uint(x) >> 31))) r = 0x1.921fb6p+1f + (-0x1.777a5cp-24f - r);
    return (__float_as_uint(y) >> 31) ? -r : r;
}

// Within 3.5 ulp of the exact value on [-1, 1]; NaN outside.
__device__ __forceinline__ float tm_acos(float x) { return 2.0f * tm_atan2(sqrtf(1.0f - x), sqrtf(1.0f + x)); }

}  // namespace tt
