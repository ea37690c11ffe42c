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
// The FP64 logarithm and sin/cos(2 pi u) below are written for this engine so
// that host and device produce bit-identical normals: both only use IEEE
// operations that are correctly rounded on x86-64 and on gfx950 (add, mul, fma,
// div, sqrt, frexp, rint).  log uses m = 2^-e u in [sqrt(1/2), sqrt(2)),
// s = (m-1)/(m+1) and log m = 2 atanh(s) summed as a series in s^2 with the
// exact coefficients 1/(2k+1).  sin/cos reduce 4u exactly to a quadrant q and a
// remainder r in [-1/2, 1/2] and sum the Taylor series of sin(pi r/2) and
// cos(pi r/2); the coefficients are (pi/2)^n / n! rounded to double.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define SP_HD __host__ __device__ __forceinline__
#else
#define SP_HD inline
#endif

namespace sp {

struct U4 { uint32_t x, y, z, w; };

SP_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
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

// ---- FP64 log(u), u in (0, 1] ---------------------------------------------------
SP_HD double log_f64(double u) {
    int e;
    double m = frexp(u, &e);                 // u = m 2^e, m in [1/2, 1)
    if (m < 0.7071067811865476) { m = m * 2.0; e -= 1; }
    const double f = m - 1.0;                // exact (Sterbenz)
    const double s = f / (2.0 + f);
    const double t = s * s;
    double P = 0.043478260869565216;         // 1/23
    P = fma(P, t, 0.047619047619047616);     // 1/21
    P = fma(P, t, 0.05263157894736842);      // 1/19
    P = fma(P, t, 0.058823529411764705);     // 1/17
    P = fma(P, t, 0.06666666666666667);      // 1/15
    P = fma(P, t, 0.07692307692307693);      // 1/13
    P = fma(P, t, 0.09090909090909091);      // 1/11
    P = fma(P, t, 0.1111111111111111);       // 1/9
    P = fma(P, t, 0.14285714285714285);      // 1/7
    P = fma(P, t, 0.2);                      // 1/5
    P = fma(P, t, 0.3333333333333333);       // 1/3
    P = P * t;                               // sum_{k>=1} t^k/(2k+1)
    // log m = 2s(1+P) = f - s (f - 2P)   (using 2s = f - s f)
    const double lm = fma(-s, fma(-2.0, P, f), f);
    const double de = (double)e;
    return fma(de, 0.6931471805599453, fma(de, 2.3190468138462996e-17, lm));
}

// ---- FP64 sin and cos of 2 pi u, u in [0, 1) -----------------------------------
SP_HD void sincos2pi_f64(double u, double* s_out, double* c_out) {
    const double x = 4.0 * u;                // exact
    const double q = rint(x);                // quadrant, no ties possible for u = (w+1/2)2^-32
    const double r = x - q;                  // exact, |r| <= 1/2
    const double r2 = r * r;
    double S = -4.377065467313742e-14;
    S = fma(S, r2, 6.0669357311061955e-12);
    S = fma(S, r2, -6.688035109811468e-10);
    S = fma(S, r2, 5.692172921967927e-08);
    S = fma(S, r2, -3.598843235212085e-06);
    S = fma(S, r2, 0.00016044118478735983);
    S = fma(S, r2, -0.004681754135318688);
    S = fma(S, r2, 0.07969262624616705);
    S = fma(S, r2, -0.6459640975062463);
    S = fma(S, r2, 1.5707963267948966);
    S = S * r;                               // sin(pi r / 2)
    double C = 3.437739179098607e-15;
    C = fma(C, r2, -5.294400200734623e-13);
    C = fma(C, r2, 6.565963114979473e-11);
    C = fma(C, r2, -6.386603083791852e-09);
    C = fma(C, r2, 4.710874778818172e-07);
    C = fma(C, r2, -2.5202042373060607e-05);
    C = fma(C, r2, 0.0009192602748394266);
    C = fma(C, r2, -0.02086348076335296);
    C = fma(C, r2, 0.25366950790104803);
    C = fma(C, r2, -1.2337005501361697);
    C = fma(C, r2, 1.0);                     // cos(pi r / 2)
    const int iq = ((int)q) & 3;
    double s, c;
    if (iq == 0)      { s = S;  c = C;  }
    else if (iq == 1) { s = C;  c = -S; }
    else if (iq == 2) { s = -S; c = -C; }
    else              { s = -C; c = S;  }
    *s_out = s;
    *c_out = c;
}

// four standard normals (FP64) from one Philox block
SP_HD void box_muller4_f64(U4 w, double z[4]) {
    const double r0 = sqrt(-2.0 * log_f64(u01_f64(w.x)));
    const double r1 = sqrt(-2.0 * log_f64(u01_f64(w.z)));
    double s0, c0, s1, c1;
    sincos2pi_f64(u01_f64(w.y), &s0, &c0);
    sincos2pi_f64(u01_f64(w.w), &s1, &c1);
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
