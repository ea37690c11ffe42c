// sp_provider.h -- one particle of an ensemble from its particle types (host + device).
//
// Replaces the reference's ParticleProvider::getProvidedObject (Provider/ParticleProvider.h:230-295),
// ParticleSimulationSettings::applySettingsToParticleProperties (Settings/ParticleSimulationSettings.h:51-152)
// and the start-state initialisers (ParticleStateInitializer.h:42-144, BrownAndNeelRelaxation.inl:277-324,
// BrownAndNeelRelaxationEulerSpherical.h:741-753), followed by the per-particle parameter calculators
// (ParameterCalculatorNeel.h:44-92, ParameterCalculatorBrown.h:39-58, ParameterCalculatorBrownAndNeel.h:67-76,
// UniaxialAnisotropy.h:71-85, CubicAnisotropy.h:92-94, UniaxialCubicAnisotropy.hpp:61-114).
//
// Particle i is a pure function of (seed, i) and the type table: its draws come from Philox stream tag 1 with
// counter (k, particle i), one Philox block per draw (uniform: u = (w0 + 1/2) 2^-32; normal: the first
// Box-Muller normal of the block).  The host materialisation of a parameter file and the device provider
// kernel run this same code, so they select the same type for every particle and draw the same words; the
// floating-point results agree to the last bit except where exp, sin, cos, acos or atan2 enter (the device
// math library and the C library may round those differently by an ulp).
//
// Choices where the reference draws from std::random_device (unreproducible): the type selection, the
// random orientations and magnetisations use the Philox draws above; radius, shell and anisotropy
// distributions are delta, normal or lognormal with the given mean and width (lognormal: the normal
// parameters that give that mean and standard deviation).
#pragma once
#include <stdint.h>
#include <math.h>

#include "../../include/stophys_gpu.h"
#include "sp_rng.h"
#include "sp_types.h"

namespace sp {

enum : int32_t { PD_DELTA = 0, PD_NORMAL = 1, PD_LOGNORMAL = 2 };

// a distribution of one property, reduced on the host to what a draw needs
struct ProvDist {
    int32_t type;         // PD_*; a zero width is PD_DELTA (no draw)
    int32_t pad;
    double mean;
    double sd;            // normal: absolute standard deviation
    double sigma, half_s2;   // lognormal: sqrt(s2), s2 / 2 with s2 = ln(1 + sd^2 / mean^2)
};

// one particle file of the Particle_List
struct ProvType {
    ProvDist rad, shell, K, K2;       // magnetic radius, hydrodynamic shell, K (K_uniaxial), K_cubic
    double Ms, alpha, gamma, T, visc;
    double euler[3], mag[3], urot[3];
    double cum;                       // running sum of the type counts up to this type (discrete selection)
    uint64_t cum_count;               // the same as an integer (sequential selection)
    uint32_t rand_orient, rand_mag, rand_urot, pad;
};

// what the provider kernel needs besides the type table
struct ProvDesc {
    uint64_t seed, N;
    int32_t problem, aniso;
    uint32_t n_types, discrete;
    double total_count;               // sum of the type counts (discrete selection)
    uint64_t n_check;                 // particles of the mixed-precision check (leading ones)
    uint32_t dy, pad;
    const ProvType* types;
    double* params;                   // [NPARAM][N]
    double* y0;                       // [dy][N]
    double* sums;                     // [block][PROV_NSUM]
};

// per-block sums the provider kernel writes: the Mean_Properties inputs T, radius, Ms, alpha, gamma, K,
// r_h, viscosity, then mu and mu over the check particles
constexpr int PROV_NSUM = 10;
constexpr uint32_t PROV_BLOCK = 256;   // particles (threads) per work-group of the provider kernel

// one particle's properties, body axes and start state
struct ParticleProps {
    uint32_t type;
    double radius, r_h, Ms, alpha, gamma, T, K, K2, visc;
    double axes[9], urot[3];
    double y0[6];
};

struct ProvDraw {
    uint64_t seed, p, k;
    SP_HD double uniform() { return u01_f64(noise_block(seed, p, k++, 1u).x); }
    SP_HD double normal() {             // z[0] of box_muller4_f64, without the three normals it discards
        const U4 w = noise_block(seed, p, k++, 1u);
        double s, c;
        sincos2pi_u01_f64(w.y, &s, &c);
        return sqrt_radius(-2.0 * log_u01_f64(w.x)) * c;
    }
};

SP_HD double prov_draw(const ProvDist& d, ProvDraw& rng) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    if (d.type == PD_NORMAL) return d.mean + d.sd * rng.normal();
    if (d.type == PD_LOGNORMAL) return d.mean * exp(d.sigma * rng.normal() - d.half_s2);
    return d.mean;
}

// particle i of the ensemble (settings.cpp's host materialisation and k_provide)
SP_HD void provide_particle(const ProvDesc& d, const ProvType* types, uint64_t i, ParticleProps& o) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    constexpr double PI = 3.141592653589793, TWO_PI = 6.283185307179586;
    ProvDraw rng{d.seed, i, 0};
    // Provider::getProvidedObject: a discrete draw weighted by the counts, or the types in list order,
    // each for its count of particles (the last type takes any particles beyond the counts)
    uint32_t sel = 0;
    if (d.discrete && d.total_count > 0.0) {
        const double u = rng.uniform() * d.total_count;
        while (sel + 1 < d.n_types && !(u < types[sel].cum)) ++sel;
    } else {
        while (sel + 1 < d.n_types && types[sel].cum_count <= i) ++sel;
    }
    const ProvType& t = types[sel];
    o.type = sel;
    const double rad = prov_draw(t.rad, rng);
    const double shell = prov_draw(t.shell, rng);
    o.radius = rad; o.Ms = t.Ms; o.alpha = t.alpha; o.gamma = t.gamma; o.T = t.T; o.visc = t.visc;
    o.K = prov_draw(t.K, rng);
    o.r_h = shell + rad;
    o.K2 = 0.0;
    o.urot[0] = o.urot[1] = o.urot[2] = 0.0;
    if (d.aniso == SP_ANISO_UNIAXIALCUBIC) {     // UniaxialCubic_Distribution::applyDistribution (:63-76)
        o.K2 = prov_draw(t.K2, rng);
        for (int k = 0; k < 3; ++k) o.urot[k] = t.rand_urot ? rng.uniform() * (k == 1 ? PI : TWO_PI) : t.urot[k];
    }
    // axes from Euler angles (NeelRelaxation.h:223-261); random: each angle U(0, 2 pi), or for
    // BrownAndNeelEulerSpherical phi, psi U(0, 2 pi) and theta U(0, pi) (ParticleStateInitializer.h:42-62)
    const bool euler_state = d.problem == SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL;
    double e[3] = {t.euler[0], t.euler[1], t.euler[2]};
    if (t.rand_orient)
        for (int k = 0; k < 3; ++k) e[k] = rng.uniform() * (euler_state && k == 1 ? PI : TWO_PI);
    const double cphi = cos(e[0]), cth = cos(e[1]), cpsi = cos(e[2]);
    const double sphi = sin(e[0]), sth = sin(e[1]), spsi = sin(e[2]);
    const double cc = cphi * cpsi, sc = sphi * cpsi, cs = cphi * spsi, ss = sphi * spsi;
    o.axes[0] = cc - cth * ss; o.axes[1] = -sc - cth * cs; o.axes[2] = sth * spsi;      // x axis
    o.axes[3] = cth * sc + cs; o.axes[4] = cth * cc - ss; o.axes[5] = -sth * cpsi;      // y axis
    o.axes[6] = sth * sphi; o.axes[7] = sth * cphi; o.axes[8] = cth;                    // z axis
    double m[3] = {t.mag[0], t.mag[1], t.mag[2]};
    if (t.rand_mag && euler_state) {     // ParticleStateInitializer.h:107-130: phi U(0, 2 pi), theta U(0, pi)
        const double ph = rng.uniform() * TWO_PI, th = rng.uniform() * PI;
        m[0] = cos(ph) * sin(th); m[1] = sin(ph) * sin(th); m[2] = cos(th);
    } else if (t.rand_mag) {
        for (int k = 0; k < 3; ++k) m[k] = rng.normal();
    }
    const double mn = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    for (int k = 0; k < 3; ++k) m[k] /= mn;
    for (int k = 0; k < 6; ++k) o.y0[k] = 0.0;
    if (euler_state) {                   // getStart (BrownAndNeelRelaxationEulerSpherical.h:741-753)
        for (int k = 0; k < 3; ++k) o.y0[k] = e[k];
        o.y0[3] = acos(m[2]);
        o.y0[4] = atan2(m[1], m[0]);
    } else if (d.problem == SP_PROBLEM_NEEL) {
        for (int k = 0; k < 3; ++k) o.y0[k] = m[k];
    } else if (d.problem == SP_PROBLEM_NEEL_SPHERICAL) {     // ParticleStateInitializer.h:132-144
        o.y0[0] = acos(m[2]);
        o.y0[1] = atan2(m[1], m[0]);
    } else {                                                 // BrownAndNeelRelaxation.inl:277-324
        double n[3];
        if (t.rand_orient) {
            for (int k = 0; k < 3; ++k) n[k] = rng.normal();
        } else {
            n[0] = cos(e[0]) * cos(e[2]) - sin(e[0]) * cos(e[1]) * sin(e[2]);
            n[1] = -cos(e[0]) * sin(e[2]) - sin(e[0]) * cos(e[1]) * cos(e[2]);
            n[2] = sin(e[0]) * sin(e[1]);
        }
        const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) { o.y0[k] = n[k] / nn; o.y0[3 + k] = m[k]; }
    }
}

// UniaxialCubicAnisotropy::calcRotation (UniaxialCubicAnisotropy.hpp:61-90): the matrix built from the
// Euler angles (phi, theta, psi) of Uniaxial_Rotation
SP_HD void uniaxial_rotation(const double* e, double R[3][3]) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double cphi = cos(e[0]), ctheta = cos(e[1]), cpsi = cos(e[2]);
    const double sphi = sin(e[0]), stheta = sin(e[1]), spsi = sin(e[2]);
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

// the kernels' parameter block of one particle (the param-major rows of sp_types.h); props.axes are the body
// axes (identity when the caller has none)
SP_HD void particle_params(int problem, int aniso, const ParticleProps& q, double out[NPARAM]) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    constexpr double kBoltzmann = 1.3806485279E-23;      // ParameterCalculatorNeel.h:43
    constexpr double kPi = 3.14159265358979323846;
    const bool euler = problem == SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL;
    const bool bn = problem == SP_PROBLEM_BROWN_AND_NEEL || euler;
    const double r = q.radius, Ms = q.Ms, alpha = q.alpha, gamma = q.gamma, T = q.T, K = q.K;
    const double V = 4.0 / 3.0 * kPi * r * r * r, mu = Ms * V;   // MagneticProperties.h:159-166
    // NeelCalculator::calcNeelParams
    double diffsquare = 2.0 * alpha * kBoltzmann * T / (gamma * mu);
    if (isinf(diffsquare) || isnan(diffsquare)) diffsquare = 0.0;
    double a, b;
    if (isinf(alpha)) { a = 0.0; b = 0.0; }
    else { a = -gamma / (1.0 + alpha * alpha); b = fabs(a) * alpha; }
    const double D = (diffsquare == 0.0) ? 0.0 : -1.0 * (a * a) * (1.0 + alpha * alpha) * diffsquare;
    out[P_A] = a;
    out[P_B] = b;
    out[P_SIG] = sqrt(diffsquare);
    out[P_D] = D;
    out[P_MU] = mu;
    out[P_K1] = -2.0 * K / Ms;
    out[P_PT] = -2.0 * K * V;
    double c = 0.0, cF = 0.0, dd = 0.0;
    if (bn) {   // BrownianRotationCalculator::calcBrownRotationParams
        const double rh = q.r_h, eta = q.visc;
        const double Vh = 4.0 / 3.0 * kPi * rh * rh * rh;      // HydrodynamicProperties.h:46
        c = 1.0 / (6.0 * eta * Vh);
        double diff = sqrt(2.0 * kBoltzmann * T * (6.0 * eta * Vh));
        if (isinf(diff) || isnan(diff)) diff = 0.0;
        if (isnan(c)) c = 0.0;
        cF = c * diff;
        dd = c * mu;   // BrownAndNeelRelaxation.inl:47-48
    }
    out[P_C] = c;
    out[P_CF] = cF;
    out[P_DD] = dd;
    for (int k = 0; k < 9; ++k) out[P_AX + k] = q.axes[k];
    double ku = 0.0, u[3] = {0.0, 0.0, 0.0};
    if (aniso == SP_ANISO_UNIAXIALCUBIC) {
        // the cubic part takes P_K1 / P_PT (CubicAnisotropy.h:92-94): K2 is K_cubic here
        out[P_K1] = -2.0 * q.K2 / Ms;
        out[P_PT] = -2.0 * q.K2 * V;
        ku = -2.0 * K / Ms;                              // UniaxialAnisotropy.h:71-77 with K_uniaxial
        double R[3][3];
        uniaxial_rotation(q.urot, R);
        if (euler) {   // the body axes move with the state: keep R itself (axis rows) and -2K_u V (torque)
            for (int k = 0; k < 9; ++k) out[P_AX + k] = R[k / 3][k % 3];
            u[0] = -2.0 * K * V;
        } else {
            const double* ax = q.axes;
            for (int i = 0; i < 3; ++i) u[i] = R[0][i] * ax[6] + R[1][i] * ax[7] + R[2][i] * ax[8];   // R^T z
        }
    }
    out[P_KU] = ku;
    for (int i = 0; i < 3; ++i) out[P_U + i] = u[i];
}

}  // namespace sp
