// sp_rng.h -- counter-based noise for the ensemble SDE engine (host + device).
//
// Replaces the reference's NoiseField (SDEFramework/NoiseField.inl:11-44): one
// pcg64 + Boost normal generator per noise component, seeded from
// std::random_device.  Here every normal is a pure function of
// (seed, particle, index), so a particle's noise does not depend on the lane,
// work-group, batch or chunk it runs in.
//
// Stream definition (shared by the FP64 and the mixed-precision paths):
//   normal index k = step * dim_W + j     (step counted from 0, j < dim_W)
//   Philox4x32-10 call c = k / 4 with counter (c_lo, c_hi, particle, tag) and
//   key (seed_lo, seed_hi) gives w0..w3; u_i = (w_i + 1/2) 2^-32 in (0,1);
//   z_{4c}   = r0 cos(2 pi u1), z_{4c+1} = r0 sin(2 pi u1), r0 = sqrt(-2 ln u0)
//   z_{4c+2} = r1 cos(2 pi u3), z_{4c+3} = r1 sin(2 pi u3), r1 = sqrt(-2 ln u2)
//   dW_j = sqrt(dt) z_k  (the reference draws N(0, sqrt(dt)), NoiseField.inl:19).
//
// Philox4x32-10 is the generator of Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): round multipliers 0xD2511F53 and
// 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85, ten rounds.
//
// The FP64 ln u and sin/cos(2 pi u) are written for this engine so that host
// and device produce bit-identical normals: they only use IEEE operations that
// are correctly rounded on x86-64 and on gfx950 (add, mul, fma, sqrt, ldexp,
// int<->double conversion) and the constant tables of sp_rng_tables.h,
// and they work on the Philox word w itself, so the reduction is exact:
//   ln u: v = w + 1/2 = m 2^e with m in [3/4, 3/2); j = rint(128 (m - 1)),
//     c_j = 1 + j/128, r = (m - c_j) / c_j (|r| <= 1/192) and
//     ln u = (e - 32) ln 2 + ln c_j + ln(1 + r), ln(1 + r) by its series to r^7;
//   2 pi u = t_j + x with t_j = (j + 1/2) 2 pi / 128 for the top 7 bits j of w
//     and |x| <= 2 pi / 256; sin x and cos x - 1 by their series (to x^7, x^6),
//     combined with the tabulated sin t_j, cos t_j by the angle-sum formulas.
#pragma once
#include <stdint.h>
#include <math.h>

#include "sp_rng_tables.h"

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define SP_HD __host__ __device__ __forceinline__
#else
#define SP_HD inline
#endif

namespace sp {

struct U4 { uint32_t x, y, z, w; };

// a ^ b ^ c; on the device one gfx950 three-input logic instruction (truth table 0x96)
SP_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

SP_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{xor3(hi1, c.y, k0), lo1, xor3(hi0, c.w, k1), lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// one Philox block of a particle's noise stream
SP_HD U4 noise_block(uint64_t seed, uint64_t particle, uint64_t call, uint32_t tag) {
    return philox4x32_10(U4{(uint32_t)call, (uint32_t)(call >> 32), (uint32_t)particle, tag},
                         (uint32_t)seed, (uint32_t)(seed >> 32));
}

SP_HD double u01_f64(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }  // (w+1/2) 2^-32

// ---- FP64 ln u, u = (w + 1/2) 2^-32 ----------------------------------------------
// The reduction is read off the bits of v = w + 1/2: adding half an exponent unit to the high word
// carries into the exponent exactly when the mantissa is >= 3/2, which gives the exponent e of
// v = m 2^e with m in [3/4, 3/2); j = rint(128 (m - 1)) comes from the round-to-integer constant
// 1.5 * 2^52, whose sum with 128 (m - 1) holds j in its low word.
// LT: the ln table (SP_LOG_TAB, or a copy of it in LDS on the device)
SP_HD double log_u01_f64(uint32_t w, const double (*LT)[2] = SP_LOG_TAB) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    constexpr double LN2_HI = 0x1.62e42fee00000p-1;   // ln 2 to 32 significant bits: E * LN2_HI is exact
    constexpr double LN2_LO = 0x1.a39ef35793c76p-33;  // ln 2 - LN2_HI
    constexpr double RND = 0x1.8p52;                  // x + RND - RND = rint(x) for |x| < 2^51
    const double v = (double)w + 0.5;                 // exact
    const uint32_t be = ((uint32_t)(__builtin_bit_cast(uint64_t, v) >> 32) + 0x80000u) >> 20;   // e + 1023
    const double u = ldexp(v, 1030 - (int)be) - 128.0;   // 128 (m - 1), exact, in [-32, 64)
    const double t = u + RND;
    const double jd = t - RND;                        // j = rint(128 (m - 1)), exact
    const uint32_t row = (uint32_t)__builtin_bit_cast(uint64_t, t) + 32u;   // j + 32 from the low word
    const double* T = LT[row];                        // ln c_j, 2^-7 / c_j
    const double r = (u - jd) * T[1];                 // (m - c_j) / c_j: 128 (m - c_j) is exact
    double p = 0.14285714285714285;                   // 1/7
    p = fma(p, r, -0.16666666666666666);
    p = fma(p, r, 0.2);
    p = fma(p, r, -0.25);
    p = fma(p, r, 0.3333333333333333);
    p = fma(p, r, -0.5);
    const double l1p = fma(p, r * r, r);              // ln(1 + r)
    const double E = (double)((int)be - 1055);        // e - 32
    return fma(E, LN2_HI, T[0] + fma(E, LN2_LO, l1p));
}

// sqrt(x) for the Box-Muller radius, x = -2 ln u in [4e-10, 45]: on the device the correctly rounded
// square root without the rescaling of tiny arguments or the zero/infinity checks (v_rsq_f64 and two
// Newton corrections of the root and the half reciprocal), on the host the C library's.
SP_HD double sqrt_radius(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma clang fp contract(off)
    double h = __builtin_amdgcn_rsq(x);
    double s = x * h;
    h = h * 0.5;
    const double e = fma(-h, s, 0.5);
    s = fma(s, e, s);
    h = fma(h, e, h);
    double d = fma(-s, s, x);
    s = fma(d, h, s);
    d = fma(-s, s, x);
    return fma(d, h, s);
#else
    return sqrt(x);
#endif
}

// ---- FP64 sin and cos of 2 pi u, u = (w + 1/2) 2^-32 ------------------------------
// ST: the sin/cos table (SP_SINCOS_TAB, or a copy of it in LDS on the device)
SP_HD void sincos2pi_u01_f64(uint32_t w, double* s_out, double* c_out, const double (*ST)[2] = SP_SINCOS_TAB) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    constexpr double TWO_PI_2M32 = 0x1.921fb54442d18p-30;   // (nearest double to 2 pi) 2^-32
    const double* T = ST[w >> 25];                         // sin t_j, cos t_j
    const int d = (int)(w & 0x1FFFFFFu) - 0x1000000;       // u = (j + 1/2)/128 + (d + 1/2) 2^-32
    const double x = ((double)d + 0.5) * TWO_PI_2M32;      // |x| <= 2 pi / 256
    const double x2 = x * x;
    double sp = -0.0001984126984126984;                    // -1/5040
    sp = fma(sp, x2, 0.008333333333333333);
    sp = fma(sp, x2, -0.16666666666666666);
    const double sx = fma(sp * x2, x, x);                  // sin x
    double cp = -0.001388888888888889;                     // -1/720
    cp = fma(cp, x2, 0.041666666666666664);
    cp = fma(cp, x2, -0.5);
    const double cm1 = cp * x2;                            // cos x - 1
    *s_out = T[0] + fma(T[0], cm1, T[1] * sx);
    *c_out = T[1] + fma(T[1], cm1, -(T[0] * sx));
}

// four standard normals (FP64) from one Philox block; the tables may be given (LDS copies on the device)
SP_HD void box_muller4_f64(U4 w, double z[4], const double (*LT)[2] = SP_LOG_TAB, const double (*ST)[2] = SP_SINCOS_TAB) {
    const double r0 = sqrt_radius(-2.0 * log_u01_f64(w.x, LT));
    const double r1 = sqrt_radius(-2.0 * log_u01_f64(w.z, LT));
    double s0, c0, s1, c1;
    sincos2pi_u01_f64(w.y, &s0, &c0, ST);
    sincos2pi_u01_f64(w.w, &s1, &c1, ST);
    z[0] = r0 * c0;
    z[1] = r0 * s0;
    z[2] = r1 * c1;
    z[3] = r1 * s1;
}

#if defined(__HIPCC__) || defined(__HIP__)
// four standard normals (FP32) from the same Philox block: gfx950 hardware
// transcendentals.  v_log_f32 gives log2, v_sin_f32/v_cos_f32 take the angle in
// revolutions, so sin(2 pi u) is one instruction.
__device__ __forceinline__ void box_muller4_f32(U4 w, float z[4]) {
    constexpr float SC = 2.3283064365386963e-10f, HALF = 1.1641532182693481e-10f;
    const float u0 = fmaf((float)w.x, SC, HALF), u1 = fmaf((float)w.y, SC, HALF);
    const float u2 = fmaf((float)w.z, SC, HALF), u3 = fmaf((float)w.w, SC, HALF);
    // -2 ln u = -2 ln2 log2 u.  u can round to 1 (w >= 2^32 - 128); log2 is clamped at 0 so that an
    // approximation error of the hardware log at u -> 1 can never make the radicand negative.
    const float r0 = __builtin_amdgcn_sqrtf(-1.3862943611198906f * fminf(__builtin_amdgcn_logf(u0), 0.0f));
    const float r1 = __builtin_amdgcn_sqrtf(-1.3862943611198906f * fminf(__builtin_amdgcn_logf(u2), 0.0f));
    z[0] = r0 * __builtin_amdgcn_cosf(u1);
    z[1] = r0 * __builtin_amdgcn_sinf(u1);
    z[2] = r1 * __builtin_amdgcn_cosf(u3);
    z[3] = r1 * __builtin_amdgcn_sinf(u3);
}
#endif

}  // namespace sp
