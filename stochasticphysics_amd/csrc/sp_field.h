// sp_field.h -- the applied fields, evaluated identically on the host (time/field axis) and on the
// device (once per step per wave, batched 64 steps at a time).
//
// Restated from the reference's field classes:
//   zero         Fields/ZeroField.h:39
//   constant     Fields/ConstantField.h:43                  H = offset
//   sinusoidal   Fields/SinusoidalField.h:39,47-51          H = A sin(2 pi f t + phi) + offset
//   lissajous    Fields/LissajousField.h:50-64,98-101       H_i = offset_i + A_i sin(2 pi f_i t + phi_i)
//   triangular   Fields/TriangularField.h:51-67
//   rectangular  Fields/RectangularField.h:26-63 (exponential edges with time constant Tau, optionally
//                alternating; the derived constants are computed once, on the host, in field_params)
//   sinc         Fields/SincField.h:34-45                   H = A sinc(pi a x)^2 + offset, a = 11 / (P/2)
#pragma once
#include <math.h>
#include <float.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define SP_FHD __host__ __device__ __forceinline__
#else
#define SP_FHD inline
#endif

namespace sp {

enum : int { FIELD_ZERO = 1, FIELD_CONSTANT = 2, FIELD_SINUSOIDAL = 3, FIELD_LISSAJOUS = 4,
             FIELD_TRIANGULAR = 5, FIELD_RECTANGULAR = 6, FIELD_SINC = 7 };

// All doubles, so that the device can copy it with plain (volatile) loads.
struct FieldParams {
    double type;
    double off[3], amp[3], w[3], ph[3];    // offset, amplitude, angular frequency 2 pi f, phase
    double period, half, toff;             // Periode, Periode / 2, TimeOffset
    double tau_inv;                        // rectangular: 1 / Tau (DBL_MAX when |Tau| <= DBL_MIN)
    double ramp[3], rmax[3], roff[3];      // rectangular: amplitude (x2 if alternating), maxField, offset
    double rtoff;                          // rectangular: time offset (shifted if alternating)
    double sinc_a;                         // sinc: 11 / (Periode / 2)
    double pad;
};
static_assert(sizeof(FieldParams) % 8 == 0, "FieldParams is read as doubles");

// RectangularField's constructor (RectangularField.h:26-34): derived constants of the edge shape.
inline void field_rectangular_constants(FieldParams& f, double tau, bool alternating) {
    f.tau_inv = fabs(tau) <= DBL_MIN ? DBL_MAX : 1.0 / tau;
    double na2 = 0.0, mx2 = 0.0;
    for (int i = 0; i < 3; ++i) {
        f.ramp[i] = alternating ? 2.0 * f.amp[i] : f.amp[i];
        f.rmax[i] = f.ramp[i] * (-expm1(-f.half * f.tau_inv));
        na2 += f.ramp[i] * f.ramp[i];
        mx2 += f.rmax[i] * f.rmax[i];
    }
    f.rtoff = alternating ? f.toff - log(1.0 - sqrt(mx2) / (2.0 * sqrt(na2))) * tau : f.toff;
    for (int i = 0; i < 3; ++i) f.roff[i] = alternating ? f.ramp[i] * expm1(-f.rtoff * f.tau_inv) + f.off[i] : f.off[i];
}

SP_FHD void field_eval(const FieldParams& f, double t, double h[3]) {
    const int type = (int)f.type;
    if (type == FIELD_SINUSOIDAL) {
        const double s = sin(f.w[0] * t + f.ph[0]);
        for (int i = 0; i < 3; ++i) h[i] = f.amp[i] * s + f.off[i];
    } else if (type == FIELD_LISSAJOUS) {
        for (int i = 0; i < 3; ++i) h[i] = f.off[i] + f.amp[i] * sin(f.w[i] * t + f.ph[i]);
    } else if (type == FIELD_CONSTANT) {
        for (int i = 0; i < 3; ++i) h[i] = f.off[i];
    } else if (type == FIELD_TRIANGULAR) {
        // as written in the reference: 3A at the start of a period, A at half period, 3A again at its end
        const double nt = t + f.toff;
        const double pos = nt < 0.0 ? f.period + fmod(nt, f.period) : fmod(nt, f.period);
        const double s = pos > f.half ? 2.0 * pos / f.half - 1.0 : -2.0 * pos / f.half + 3.0;
        for (int i = 0; i < 3; ++i) h[i] = s * f.amp[i] + f.off[i];
    } else if (type == FIELD_RECTANGULAR) {
        const double nt = t + f.rtoff;
        const double pos = nt < 0.0 ? fmod(nt + f.period, f.period) : fmod(nt, f.period);
        if (pos <= f.half) {
            const double e = -expm1(-pos * f.tau_inv);
            for (int i = 0; i < 3; ++i) h[i] = f.ramp[i] * e + f.roff[i];
        } else {
            const double e = exp(-(pos - f.half) * f.tau_inv);
            for (int i = 0; i < 3; ++i) h[i] = f.ramp[i] * e + f.roff[i] - (f.ramp[i] - f.rmax[i]);
        }
    } else if (type == FIELD_SINC) {
        const double pos = t < 0.0 ? fmod(t + f.period, f.period) - f.half : fmod(t, f.period) - f.half;
        const double x = 3.14159265358979323846 * f.sinc_a * pos;
        const double sc = x == 0.0 ? 1.0 : sin(x) / x;        // boost::math::sinc_pi
        for (int i = 0; i < 3; ++i) h[i] = f.amp[i] * (sc * sc) + f.off[i];
    } else {
        h[0] = h[1] = h[2] = 0.0;
    }
}

}  // namespace sp
