// sp_params.cpp -- per-particle parameter calculators and request shapes (host).
//
// Evaluates, once per particle, the reference's parameter calculators (particle_params of sp_provider.h,
// which the device provider kernel runs too) into the param-major block the kernels load.
#include "sp_host.h"
#include "sp_provider.h"

#include <cmath>
#include <cstring>

namespace sp {

static constexpr double kPi = 3.14159265358979323846;

int request_dims(int problem, uint32_t* dy, uint32_t* dw, uint32_t* dout) {
    switch (problem) {
        case SP_PROBLEM_NEEL: *dy = 3; *dw = 3; *dout = 3; return 0;
        case SP_PROBLEM_NEEL_SPHERICAL: *dy = 2; *dw = 3; *dout = 3; return 0;
        case SP_PROBLEM_BROWN_AND_NEEL: *dy = 6; *dw = 6; *dout = 6; return 0;
        // (phi, theta, psi, theta_N, phi_N), 6 noise components, output (x axis, y axis, m)
        case SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL: *dy = 5; *dw = 6; *dout = 9; return 0;
    }
    return -1;
}

void compute_params(const sp_request& rq, const sp_particles& pt, double* out, uint64_t stride) {
    parallel_ranges(rq.n_particles, [&](uint64_t p0, uint64_t p1) { compute_params_range(rq, pt, out, stride, p0, p1); });
}

void compute_params_range(const sp_request& rq, const sp_particles& pt, double* out, uint64_t stride, uint64_t p0,
                          uint64_t p1) {
    ParticleProps q{};
    double v[NPARAM];
    for (uint64_t p = p0; p < p1; ++p) {
        q.radius = pt.magnetic_radius[p]; q.Ms = pt.saturation_magnetisation[p]; q.alpha = pt.damping[p];
        q.gamma = pt.gyromagnetic_ratio[p]; q.T = pt.temperature[p]; q.K = pt.anisotropy_constant[p];
        q.r_h = pt.hydrodynamic_radius ? pt.hydrodynamic_radius[p] : 0.0;
        q.visc = pt.viscosity ? pt.viscosity[p] : 0.0;
        for (int k = 0; k < 9; ++k) q.axes[k] = pt.axes ? pt.axes[p * 9 + k] : ((k == 0 || k == 4 || k == 8) ? 1.0 : 0.0);
        if (rq.anisotropy == SP_ANISO_UNIAXIALCUBIC) {
            q.K2 = pt.anisotropy_constant_cubic[p];
            for (int k = 0; k < 3; ++k) q.urot[k] = pt.uniaxial_rotation[3 * p + k];
        }
        particle_params(rq.problem, rq.anisotropy, q, v);
        for (int k = 0; k < NPARAM; ++k) out[k * stride + p] = v[k];
    }
}

FieldParams field_params(const sp_request& rq) {
    FieldParams f{};
    f.type = (double)rq.field;
    for (int i = 0; i < 3; ++i) {
        f.off[i] = rq.field_offset[i];
        f.amp[i] = rq.field_amplitude[i];
        f.ph[i] = rq.field_phase[i];
        f.w[i] = 2.0 * kPi * rq.field_freq[i];   // SinusoidalField.h:39, LissajousField.h:50-64
    }
    if (rq.field == SP_FIELD_SINUSOIDAL)
        for (int i = 1; i < 3; ++i) { f.w[i] = f.w[0]; f.ph[i] = f.ph[0]; }
    f.period = rq.field_period;
    f.half = rq.field_period * 0.5;                                         // Periode / 2 (exact)
    f.toff = rq.field_time_offset;
    f.sinc_a = rq.field == SP_FIELD_SINC ? 11.0 / f.half : 0.0;            // SincField.h:36
    if (rq.field == SP_FIELD_RECTANGULAR) field_rectangular_constants(f, rq.field_tau, rq.field_alternating != 0);
    return f;
}

// SingleParticleSimulator::calculateTimeAndField (SingleParticleSimulator.h:236-273)
void time_and_field(const sp_request& rq, double* times, double* fields) {
    const uint64_t points = rq.n_steps / rq.oversampling;
    if (points == 0) return;
    const uint64_t OS = rq.oversampling;
    const FieldParams f = field_params(rq);
    double start = 0.0, end = 0.0;
    const double timeshift = rq.dt * ((double)OS) * 0.5;
    times[0] = start;
    field_eval(f, start, fields);
    uint64_t counter = 0;
    for (uint64_t i = 1; i < points; ++i) {
        double tmp[3] = {0.0, 0.0, 0.0}, h[3];
        start = end + rq.dt;
        for (uint64_t l = OS; l--;) {
            field_eval(f, rq.dt * ((double)++counter), h);
            for (int c = 0; c < 3; ++c) tmp[c] += h[c];
        }
        end = rq.dt * ((double)counter) + start;
        for (int c = 0; c < 3; ++c) fields[i * 3 + c] = tmp[c] / (double)OS;
        times[i] = (end - start) - timeshift;
    }
}

}  // namespace sp
