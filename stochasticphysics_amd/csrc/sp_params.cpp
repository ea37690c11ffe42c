// sp_params.cpp -- per-particle parameter calculators and request shapes (host).
//
// Evaluates, once per particle, the reference's parameter calculators
// (Problems/Helpers/ParameterCalculatorNeel.h:44-92, ParameterCalculatorBrown.h:39-58,
// ParameterCalculatorBrownAndNeel.h:67-76, UniaxialAnisotropy.h:71-85,
// CubicAnisotropy.h:92-94, UniaxialCubicAnisotropy.hpp:92-96) into the param-major block the kernels load.
#include "sp_host.h"

#include <cmath>
#include <cstring>

namespace sp {

static constexpr double kBoltzmann = 1.3806485279E-23;      // ParameterCalculatorNeel.h:43
static constexpr double kPi = 3.14159265358979323846;

// MyCEL sphere volume 4/3 pi r^3 (MagneticProperties.h:159, HydrodynamicProperties.h:46)
static inline double sphere_volume(double r) { return 4.0 / 3.0 * kPi * r * r * r; }

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
    const uint64_t N = rq.n_particles;
    const bool euler = rq.problem == SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL;
    const bool bn = rq.problem == SP_PROBLEM_BROWN_AND_NEEL || euler;
    for (uint64_t p = 0; p < N; ++p) {
        const double r = pt.magnetic_radius[p], Ms = pt.saturation_magnetisation[p];
        const double alpha = pt.damping[p], gamma = pt.gyromagnetic_ratio[p], T = pt.temperature[p];
        const double K = pt.anisotropy_constant[p];
        const double V = sphere_volume(r), mu = Ms * V;
        // NeelCalculator::calcNeelParams
        double diffsquare = 2.0 * alpha * kBoltzmann * T / (gamma * mu);
        if (std::isinf(diffsquare) || std::isnan(diffsquare)) diffsquare = 0.0;
        double a, b;
        if (std::isinf(alpha)) { a = 0.0; b = 0.0; }
        else { a = -gamma / (1.0 + std::pow(alpha, 2)); b = std::abs(a) * alpha; }
        const double D = (diffsquare == 0.0) ? 0.0 : -1.0 * std::pow(a, 2) * (1.0 + std::pow(alpha, 2)) * diffsquare;
        const double sig = std::sqrt(diffsquare);
        out[P_A * stride + p] = a;
        out[P_B * stride + p] = b;
        out[P_SIG * stride + p] = sig;
        out[P_D * stride + p] = D;
        out[P_MU * stride + p] = mu;
        out[P_K1 * stride + p] = -2.0 * K / Ms;
        out[P_PT * stride + p] = -2.0 * K * V;
        double c = 0.0, cF = 0.0, d = 0.0;
        if (bn) {   // BrownianRotationCalculator::calcBrownRotationParams
            const double Vh = sphere_volume(pt.hydrodynamic_radius[p]), eta = pt.viscosity[p];
            c = 1.0 / (6.0 * eta * Vh);
            double diff = std::sqrt(2.0 * kBoltzmann * T * (6.0 * eta * Vh));
            if (std::isinf(diff) || std::isnan(diff)) diff = 0.0;
            if (std::isnan(c)) c = 0.0;
            cF = c * diff;
            d = c * mu;   // BrownAndNeelRelaxation.inl:47-48
        }
        out[P_C * stride + p] = c;
        out[P_CF * stride + p] = cF;
        out[P_DD * stride + p] = d;
        double ax[9];
        for (int k = 0; k < 9; ++k) {
            ax[k] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
            if (pt.axes) ax[k] = pt.axes[p * 9 + k];
            out[(P_AX + k) * stride + p] = ax[k];
        }
        double ku = 0.0, u[3] = {0.0, 0.0, 0.0};
        if (rq.anisotropy == SP_ANISO_UNIAXIALCUBIC) {
            // the cubic part takes P_K1 / P_PT (CubicAnisotropy.h:92-94): K is K_cubic here
            const double Kc = pt.anisotropy_constant_cubic[p];
            out[P_K1 * stride + p] = -2.0 * Kc / Ms;
            out[P_PT * stride + p] = -2.0 * Kc * V;
            ku = -2.0 * K / Ms;                              // UniaxialAnisotropy.h:71-77 with K_uniaxial
            double R[3][3];
            uniaxial_rotation_matrix(pt.uniaxial_rotation + 3 * p, R);
            if (euler) {   // the body axes move with the state: keep R itself (axis rows) and -2K_u V (torque)
                for (int k = 0; k < 9; ++k) out[(P_AX + k) * stride + p] = R[k / 3][k % 3];
                u[0] = -2.0 * K * V;
            } else {
                for (int i = 0; i < 3; ++i) u[i] = R[0][i] * ax[6] + R[1][i] * ax[7] + R[2][i] * ax[8];   // R^T z
            }
        }
        out[P_KU * stride + p] = ku;
        for (int i = 0; i < 3; ++i) out[(P_U + i) * stride + p] = u[i];
    }
}

// UniaxialCubicAnisotropy::calcRotation (UniaxialCubicAnisotropy.hpp:61-90): the matrix built from the
// Euler angles (phi, theta, psi) of Uniaxial_Rotation
void uniaxial_rotation_matrix(const double* e, double R[3][3]) {
    const double cphi = std::cos(e[0]), ctheta = std::cos(e[1]), cpsi = std::cos(e[2]);
    const double sphi = std::sin(e[0]), stheta = std::sin(e[1]), spsi = std::sin(e[2]);
    const double cphicpsi = cphi * cpsi, sphicpsi = sphi * cpsi, cphispsi = cphi * spsi, sphispsi = sphi * spsi;
    R[0][0] = cphicpsi - ctheta * sphispsi;
    R[0][1] = -sphicpsi - ctheta * cphispsi;
    R[0][2] = stheta * spsi;
    R[1][0] = ctheta * sphicpsi + cphispsi;
    R[1][1] = ctheta * cphicpsi - sphispsi;
    R[1][2] = -stheta * cpsi;
    R[2][0] = stheta * sphi;
    R[2][1] = stheta * cphi;
    R[2][2] = ctheta;
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
