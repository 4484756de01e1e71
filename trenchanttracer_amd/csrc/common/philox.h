// Philox-4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw: "Parallel
// random numbers: as easy as 1, 2, 3", SC'11). One call maps a 128-bit counter and
// a 64-bit key to 128 random bits; streams are addressed by counter, so the
// GPU stages and the host code draw identical numbers without carrying state.
// Usable from host C++ and HIP device code.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define TT_HD __host__ __device__ inline
#else
#define TT_HD inline
#endif

namespace tt {

struct u32x4 {
    uint32_t x, y, z, w;
};

TT_HD void philox_mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

TT_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        philox_mulhilo(M0, c.x, hi0, lo0);
        philox_mulhilo(M1, c.z, hi1, lo1);
        u32x4 n = {hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// 24 high bits -> float in [0, 1).
TT_HD float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
// 53 bits from two words -> double in [0, 1).
TT_HD double u01d(uint32_t a, uint32_t b) {
    uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)v * (1.0 / 9007199254740992.0);
}

}  // namespace tt
