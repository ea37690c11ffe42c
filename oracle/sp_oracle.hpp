// sp_oracle.hpp -- CPU oracle for the StochasticPhysics particle-simulation hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a scalar, Eigen-free restatement of the
// reference's algorithm, used by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py as the checker.  The product (the HIP library in
// stochasticphysics_amd/) never includes, links or calls it.
//
// Every function follows the reference file:line named beside it, in the
// reference's order of operations where that is visible (Eigen expression
// order is approximated by left-to-right evaluation).  Parity pins, one per
// function, live in tests/test_oracle_*.py:
//   - golden vectors of the reference's own gtests (tests/golden/reference_vectors.json):
//     Neel det/S/drift/JacDet/JacSto, NeelSpherical det/S/drift/finish,
//     BrownAndNeel det/S(full, simple)/drift(full, simple), uniaxial Jacobian,
//     Neel parameter closed forms, Newton roots of Test_Function_1..3;
//   - closed forms / invariants where the reference has no vector: cubic field
//     = -(1/Ms) grad of the cubic energy (finite differences), Box-Muller log
//     and sin/cos against libm, Philox against the published KAT vectors,
//     Dual-number Jacobians against finite differences, solver one-step
//     formulas against a direct evaluation, equilibrium statistics.
#pragma once
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <type_traits>

namespace oracle {

// ------------------------------------------------------------------ small algebra
template <class S, int N> struct Vec {
    S v[N];
    S& operator[](int i) { return v[i]; }
    const S& operator[](int i) const { return v[i]; }
    static Vec zero() { Vec r; for (int i = 0; i < N; ++i) r.v[i] = S(0.0); return r; }
};
template <class S, int R, int C> struct Mat {
    S m[R][C];
    S& operator()(int i, int j) { return m[i][j]; }
    const S& operator()(int i, int j) const { return m[i][j]; }
    static Mat zero() { Mat r; for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) r.m[i][j] = S(0.0); return r; }
    static Mat identity() { Mat r = zero(); for (int i = 0; i < R && i < C; ++i) r.m[i][i] = S(1.0); return r; }
};
template <class S, int N> Vec<S, N> operator+(const Vec<S, N>& a, const Vec<S, N>& b) { Vec<S, N> r; for (int i = 0; i < N; ++i) r[i] = a[i] + b[i]; return r; }
template <class S, int N> Vec<S, N> operator-(const Vec<S, N>& a, const Vec<S, N>& b) { Vec<S, N> r; for (int i = 0; i < N; ++i) r[i] = a[i] - b[i]; return r; }
template <class S, class T, int N> Vec<S, N> operator*(const T& s, const Vec<S, N>& a) { Vec<S, N> r; for (int i = 0; i < N; ++i) r[i] = s * a[i]; return r; }
template <class S, int N> Vec<S, N>& operator+=(Vec<S, N>& a, const Vec<S, N>& b) { for (int i = 0; i < N; ++i) a[i] = a[i] + b[i]; return a; }
template <class S, int N> S dot(const Vec<S, N>& a, const Vec<S, N>& b) { S r = a[0] * b[0]; for (int i = 1; i < N; ++i) r = r + a[i] * b[i]; return r; }
template <class S> Vec<S, 3> cross(const Vec<S, 3>& a, const Vec<S, 3>& b) {
    Vec<S, 3> r;
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
    return r;
}
template <class S, int R, int C, int K> Mat<S, R, C> matmul(const Mat<S, R, K>& a, const Mat<S, K, C>& b) {
    Mat<S, R, C> r;
    for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) { S s = a(i, 0) * b(0, j); for (int k = 1; k < K; ++k) s = s + a(i, k) * b(k, j); r(i, j) = s; }
    return r;
}
template <class S, class T, int R, int C> Vec<S, R> matvec(const Mat<S, R, C>& a, const Vec<T, C>& x) {
    Vec<S, R> r;
    for (int i = 0; i < R; ++i) { S s = a(i, 0) * x[0]; for (int j = 1; j < C; ++j) s = s + a(i, j) * x[j]; r[i] = s; }
    return r;
}
template <class S> Vec<S, 3> v3(S a, S b, S c) { Vec<S, 3> r; r[0] = a; r[1] = b; r[2] = c; return r; }
template <class S, class T> Vec<S, 3> lift3(const Vec<T, 3>& a) { return v3<S>(S(a[0]), S(a[1]), S(a[2])); }

// ------------------------------------------------------------------ forward-mode duals
// Used only to form exact Newton Jacobians for the ImplicitMidpoint oracle.
template <int N> struct Dual {
    double v; double d[N];
    Dual() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    static Dual var(double x, int k) { Dual r(x); r.d[k] = 1.0; return r; }
};
template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a) { Dual<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v / b.v; for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v); return r; }
template <int N> Dual<N> operator*(double a, const Dual<N>& b) { return Dual<N>(a) * b; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b) { return a * Dual<N>(b); }
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> Dual<N> chain(const Dual<N>& a, double f, double df) { Dual<N> r; r.v = f; for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i]; return r; }
template <int N> Dual<N> sin(const Dual<N>& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
template <int N> Dual<N> cos(const Dual<N>& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
template <int N> Dual<N> sqrt(const Dual<N>& a) { const double s = std::sqrt(a.v); return chain(a, s, 0.5 / s); }
template <int N> Dual<N> acos(const Dual<N>& a) { return chain(a, std::acos(a.v), -1.0 / std::sqrt(1.0 - a.v * a.v)); }
template <int N> Dual<N> atan2(const Dual<N>& y, const Dual<N>& x) {
    Dual<N> r; r.v = std::atan2(y.v, x.v);
    const double den = x.v * x.v + y.v * y.v;
    for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / den;
    return r;
}
template <int N> Dual<N> fmod(const Dual<N>& a, double m) { Dual<N> r = a; r.v = std::fmod(a.v, m); return r; }
inline double val(double x) { return x; }
template <int N> double val(const Dual<N>& x) { return x.v; }
using std::sin; using std::cos; using std::sqrt; using std::acos; using std::atan2; using std::fmod;

// ------------------------------------------------------------------ constants
constexpr double kB = 1.3806485279E-23;      // ParameterCalculatorNeel.h:43
constexpr double PI = 3.14159265358979323846;
constexpr double TWO_PI = 6.28318530717958647692;

// MyCEL sphere volume, 4/3 pi r^3; pinned by the golden vectors (V enters mu)
inline double sphere_volume(double r) { return 4.0 / 3.0 * PI * r * r * r; }

struct MagneticProperties {          // Properties/MagneticProperties.h:100-267
    double radius, Ms, alpha, gamma, K;   // K: K_uniaxial or K_cubic
    int aniso;                            // 1 uniaxial, 2 cubic, 4 uniaxialcubic
    double K2 = 0.0;                      // uniaxialcubic: K_cubic (K is K_uniaxial)
    double urot[3] = {0.0, 0.0, 0.0};     // uniaxialcubic: Uniaxial_Rotation (UniaxialCubic.hpp:26)
    double volume() const { return sphere_volume(radius); }           // :159
    double moment() const { return Ms * volume(); }                   // :166
};

struct NeelParams { double NeelFactor1, NeelFactor2, DriftPrefactor, NoisePrefactor, Damping; };

// ParameterCalculatorNeel.h:44-92
inline NeelParams calcNeelParams(const MagneticProperties& mp, double T) {
    NeelParams p{};
    const double alpha = mp.alpha;
    double diffsquare = 2.0 * alpha * kB * T / (mp.gamma * mp.moment());
    if (std::isinf(diffsquare) || std::isnan(diffsquare)) diffsquare = 0.0;
    if (std::isinf(alpha)) { p.NeelFactor1 = 0.0; p.NeelFactor2 = 0.0; }
    else {
        p.NeelFactor1 = -mp.gamma / (1.0 + std::pow(alpha, 2));
        p.NeelFactor2 = std::abs(p.NeelFactor1) * alpha;
    }
    if (diffsquare == 0.0) p.DriftPrefactor = 0.0;
    else p.DriftPrefactor = -1.0 * std::pow(p.NeelFactor1, 2) * (1.0 + std::pow(alpha, 2)) * diffsquare;
    p.NoisePrefactor = std::sqrt(diffsquare);
    p.Damping = alpha;
    return p;
}

struct BrownParams { double BrownPrefactor, BrownDiffusion, Brown_F_Noise; };

// ParameterCalculatorBrown.h:39-58 (hydrodynamic volume 4/3 pi r_h^3, HydrodynamicProperties.h:46)
inline BrownParams calcBrownRotationParams(double r_h, double viscosity, double T) {
    BrownParams p{};
    const double Vh = sphere_volume(r_h);
    p.BrownPrefactor = 1.0 / (6.0 * viscosity * Vh);
    p.BrownDiffusion = std::sqrt(2.0 * kB * T * (6.0 * viscosity * Vh));
    if (std::isinf(p.BrownDiffusion) || std::isnan(p.BrownDiffusion)) p.BrownDiffusion = 0.0;
    if (std::isnan(p.BrownPrefactor)) p.BrownPrefactor = 0.0;
    p.Brown_F_Noise = p.BrownPrefactor * p.BrownDiffusion;
    return p;
}

// ------------------------------------------------------------------ anisotropies
struct Axes { Vec<double, 3> x, y, z; };

// UniaxialAnisotropy.h:68-150
struct Uniaxial {
    double prefactorField, prefactorTorque;   // -2K/Ms, -2K V
    explicit Uniaxial(const MagneticProperties& mp)
        : prefactorField(-2.0 * mp.K / mp.Ms), prefactorTorque(-2.0 * mp.K * mp.volume()) {}
    template <class S> void prepareField(const Vec<S, 3>&, const Axes&) const {}
    template <class S> Vec<S, 3> field(const Vec<S, 3>& m, const Axes& ax) const { return field_axis(m, lift3<S>(ax.z)); }
    template <class S> Vec<S, 3> field_axis(const Vec<S, 3>& m, const Vec<S, 3>& n) const {   // :141-144
        return (prefactorField * dot(m, n)) * n;
    }
    template <class S> Vec<S, 3> torque_axis(const Vec<S, 3>& m, const Vec<S, 3>& n) const {  // :146-150
        return cross(n, (prefactorTorque * dot(m, n)) * m);
    }
    // getEffTorque with body axes (:108-118): the easy axis is the body z axis
    Vec<double, 3> torqueBody(const Vec<double, 3>& m, const Axes& ax, const double*, const double*) const {
        return torque_axis<double>(m, ax.z);
    }
    Mat<double, 3, 3> jacobian(const Axes& ax) const {                                       // :130-137
        Mat<double, 3, 3> J;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) J(i, j) = (prefactorField * ax.z[i]) * ax.z[j];
        return J;
    }
};

// CubicAnisotropy.h:92-236; the field uses the cache prepareField fills (:109-121)
struct Cubic {
    double K1_MS, K1VM;
    explicit Cubic(const MagneticProperties& mp) : K1_MS(-2.0 * mp.K / mp.Ms), K1VM(-2.0 * mp.K * mp.volume()) {}
    template <class S> Vec<S, 3> field(const Vec<S, 3>& m, const Axes& ax) const {
        const Vec<S, 3> xa = lift3<S>(ax.x), ya = lift3<S>(ax.y), za = lift3<S>(ax.z);
        const S c1m = dot(m, xa), c2m = dot(m, ya), c3m = dot(m, za);
        const S c1m_2 = c1m * c1m, c2m_2 = c2m * c2m, c3m_2 = c3m * c3m;
        const Vec<S, 3> t = ((c2m_2 + c3m_2) * c1m) * xa + ((c1m_2 + c3m_2) * c2m) * ya + ((c1m_2 + c2m_2) * c3m) * za;
        return K1_MS * t;                                                                     // :139
    }
    // getEffTorque (:161-193) from the prepareField cache (:109-121) and the sines / cosines of the body's Euler
    // angles (phi, theta, psi); division by sin(theta) as written (0/0 when theta = 0)
    Vec<double, 3> torqueBody(const Vec<double, 3>& m, const Axes& ax, const double* sn, const double* cs) const {
        const double c1m = dot(m, ax.x), c2m = dot(m, ax.y), c3m = dot(m, ax.z);
        const double c1m_2 = c1m * c1m, c2m_2 = c2m * c2m;
        const double ctheta = cs[1], cpsi = cs[2], stheta = sn[1], spsi = sn[2];
        const Vec<double, 3> theta = v3(ctheta * spsi, -ctheta * cpsi, -stheta);
        Vec<double, 3> symetric;
        for (int i = 0; i < 3; ++i) symetric[i] = (m[i] * c3m) * (c1m_2 + c2m_2);
        const Vec<double, 3> symetricresult = cross(ax.z, symetric);
        Vec<double, 3> asymetricdir;
        for (int i = 0; i < 3; ++i) asymetricdir[i] = ((theta[i] * ctheta) * 1.0) / stheta + ax.z[i];
        const double asymetricterm = c1m * c2m * (c2m_2 - c1m_2);
        return K1VM * (symetricresult + asymetricterm * asymetricdir);
    }
    // Jacobian of the field (the reference returns Identity here, :228-236 "TODO");
    // analytic form d/dm of the expression above, used for documentation tests.
    Mat<double, 3, 3> jacobian(const Vec<double, 3>& m, const Axes& ax) const {
        const Vec<double, 3> a[3] = {ax.x, ax.y, ax.z};
        double c[3]; for (int i = 0; i < 3; ++i) c[i] = dot(m, a[i]);
        Mat<double, 3, 3> J = Mat<double, 3, 3>::zero();
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1) % 3, k = (i + 2) % 3;
            // term_i = c_i (c_j^2 + c_k^2) a_i ; grad term_i = (c_j^2+c_k^2) a_i a_i^T + 2 c_i c_j a_i a_j^T + 2 c_i c_k a_i a_k^T
            for (int r = 0; r < 3; ++r) for (int s = 0; s < 3; ++s)
                J(r, s) = J(r, s) + K1_MS * (a[i][r] * ((c[j] * c[j] + c[k] * c[k]) * a[i][s] + 2.0 * c[i] * c[j] * a[j][s] + 2.0 * c[i] * c[k] * a[k][s]));
        }
        return J;
    }
};

// UniaxialCubicAnisotropy.hpp:40-146: a uniaxial part about u = R^T z (R from the Euler angles of
// Uniaxial_Rotation, calcRotation :61-90) plus a cubic part about the particle's axes (:114).  Only the
// field is restated: the problems this oracle runs with it (Neel, NeelSpherical) use no torque.
struct UniaxialCubic {
    double prefactorField;       // -2 K_uniaxial / Ms  (UniaxialAnisotropy.h:71-77)
    Cubic cubic;                 // -2 K_cubic / Ms
    Mat<double, 3, 3> R;
    static MagneticProperties cubicPart(const MagneticProperties& mp) { MagneticProperties c = mp; c.K = mp.K2; return c; }
    double uniTorque;            // -2 K_uniaxial V (UniaxialAnisotropy.h:79-85)
    explicit UniaxialCubic(const MagneticProperties& mp)
        : prefactorField(-2.0 * mp.K / mp.Ms), cubic(cubicPart(mp)), R(rotation(mp.urot)),
          uniTorque(-2.0 * mp.K * mp.volume()) {}
    static Mat<double, 3, 3> rotation(const double* e) {                     // :61-90
        const Vec<double, 3> ang = v3(e[0], e[1], e[2]);
        const double cphi = std::cos(ang[0]), ctheta = std::cos(ang[1]), cpsi = std::cos(ang[2]);
        const double sphi = std::sin(ang[0]), stheta = std::sin(ang[1]), spsi = std::sin(ang[2]);
        const double cphicpsi = cphi * cpsi, sphicpsi = sphi * cpsi, cphispsi = cphi * spsi, sphispsi = sphi * spsi;
        Mat<double, 3, 3> r;
        r(0, 0) = cphicpsi - ctheta * sphispsi;  r(0, 1) = -sphicpsi - ctheta * cphispsi; r(0, 2) = stheta * spsi;
        r(1, 0) = ctheta * sphicpsi + cphispsi;  r(1, 1) = ctheta * cphicpsi - sphispsi;  r(1, 2) = -stheta * cpsi;
        r(2, 0) = stheta * sphi;                 r(2, 1) = stheta * cphi;                 r(2, 2) = ctheta;
        return r;
    }
    Vec<double, 3> easyAxis(const Axes& ax) const {                          // uniaxial_rotation.transpose() * zi
        Vec<double, 3> u;
        for (int i = 0; i < 3; ++i) u[i] = R(0, i) * ax.z[0] + R(1, i) * ax.z[1] + R(2, i) * ax.z[2];
        return u;
    }
    template <class S> Vec<S, 3> field(const Vec<S, 3>& m, const Axes& ax) const {
        const Vec<S, 3> u = lift3<S>(easyAxis(ax));
        return (prefactorField * dot(m, u)) * u + cubic.field(m, ax);
    }
    // getEffTorque (:120-127): the uniaxial part's torque takes R m (not the field's R^T z), plus the cubic torque
    Vec<double, 3> torqueBody(const Vec<double, 3>& m, const Axes& ax, const double* sn, const double* cs) const {
        const Vec<double, 3> Rm = matvec(R, m);
        const Vec<double, 3> tu = cross(ax.z, (uniTorque * dot(Rm, ax.z)) * Rm);
        return tu + cubic.torqueBody(m, ax, sn, cs);
    }
};

// Euler angles -> body axes (NeelRelaxation.h:223-261)
inline Axes axesFromEuler(double phi, double theta, double psi) {
    const double cphi = std::cos(phi), ctheta = std::cos(theta), cpsi = std::cos(psi);
    const double sphi = std::sin(phi), stheta = std::sin(theta), spsi = std::sin(psi);
    const double cphicpsi = cphi * cpsi, sphicpsi = sphi * cpsi, cphispsi = cphi * spsi, sphispsi = sphi * spsi;
    Axes a;
    a.x = v3(cphicpsi - ctheta * sphispsi, -sphicpsi - ctheta * cphispsi, stheta * spsi);
    a.y = v3(ctheta * sphicpsi + cphispsi, ctheta * cphicpsi - sphispsi, -stheta * cpsi);
    a.z = v3(stheta * sphi, stheta * cphi, ctheta);
    return a;
}

template <class Aniso, class S> Vec<S, 3> aniso_field(const Aniso& an, const Vec<S, 3>& m, const Axes& ax) { return an.field(m, ax); }

// Eigen's MatrixBase::normalize() on a 3-block: v /= sqrt(v.v) when v.v > 0.
template <int N> void normalize3(Vec<double, N>& y, int o) {
    const double q = (y[o] * y[o] + y[o + 1] * y[o + 1]) + y[o + 2] * y[o + 2];
    if (q > 0.0) {
        const double r = std::sqrt(q);
        y[o] = y[o] / r; y[o + 1] = y[o + 1] / r; y[o + 2] = y[o + 2] / r;
    }
}

// ------------------------------------------------------------------ Neel problem
// Problems/NeelRelaxation.h:27-309.  Stratonovich (NeelRelaxation_Definitions.h:105).
template <class Aniso> struct NeelProblem {
    static constexpr int DY = 3, DW = 3, DOUT = 3;
    NeelParams mParams; Aniso mAnisotropy; Axes ax; double mu;
    NeelProblem(const MagneticProperties& mp, double T, const Axes& axes)
        : mParams(calcNeelParams(mp, T)), mAnisotropy(mp), ax(axes), mu(mp.moment()) {}

    template <class S> void prepare(Vec<S, 3>&) {}                         // :195-196
    template <class S> void finish(Vec<S, 3>&) const {}                     // :201-202
    void normalize(Vec<double, 3>& y) const { normalize3(y, 0); }           // :204-208
    template <class S> Mat<S, 3, 3> sto(const Vec<S, 3>& yi) const {        // :91-109
        const double pre2 = mParams.NeelFactor2 * mParams.NoisePrefactor;
        const double pre1 = mParams.NeelFactor1 * mParams.NoisePrefactor;
        Mat<S, 3, 3> M;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M(i, j) = S(i == j ? pre2 : 0.0) - (pre2 * yi[i]) * yi[j];
        const Vec<S, 3> yi2 = pre1 * yi;
        M(0, 1) = M(0, 1) - yi2[2]; M(0, 2) = M(0, 2) + yi2[1];
        M(1, 0) = M(1, 0) + yi2[2]; M(1, 2) = M(1, 2) - yi2[0];
        M(2, 0) = M(2, 0) - yi2[1]; M(2, 1) = M(2, 1) + yi2[0];
        return M;
    }
    template <class S> Vec<S, 3> drift(const Vec<S, 3>& yi) const { return mParams.DriftPrefactor * yi; }   // :111-115
    template <class S> Vec<S, 3> det(const Vec<S, 3>& yi, const Vec<double, 3>& xi) const {                 // :117-128
        const Vec<S, 3> Heff = aniso_field(mAnisotropy, yi, ax) + lift3<S>(xi);
        return mParams.NeelFactor1 * cross(yi, Heff) - mParams.NeelFactor2 * cross(yi, cross(yi, Heff));
    }
    // reference Jacobians (:134-190), kept for the golden-vector pin only
    Mat<double, 3, 3> jacDetReference(const Vec<double, 3>& m, const Vec<double, 3>& xi, double dt) const {
        const Mat<double, 3, 3> HJ = mAnisotropy.jacobian(ax);
        Mat<double, 3, 3> mp = Mat<double, 3, 3>::zero();
        mp(0, 1) = -m[2]; mp(0, 2) = m[1]; mp(1, 0) = m[2]; mp(1, 2) = -m[0]; mp(2, 0) = -m[1]; mp(2, 1) = m[0];
        const Vec<double, 3> H = aniso_field(mAnisotropy, m, ax) + xi;
        Mat<double, 3, 3> J = matmul(mp, HJ);
        J(0, 1) += H[2]; J(0, 2) -= H[1]; J(1, 0) -= H[2]; J(1, 2) += H[0]; J(2, 0) += H[1]; J(2, 1) -= H[0];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) J(i, j) = mParams.NeelFactor1 * J(i, j) - 2.0 * mParams.Damping / dt * mp(i, j);
        return J;
    }
    Mat<double, 3, 3> jacStoReference(const Vec<double, 3>& dW) const {
        Mat<double, 3, 3> J = Mat<double, 3, 3>::zero();
        const double pre1 = mParams.NeelFactor1 * mParams.NoisePrefactor;
        const Vec<double, 3> dw2 = (-pre1) * dW;
        J(0, 1) -= dw2[2]; J(0, 2) += dw2[1]; J(1, 0) += dw2[2]; J(1, 2) -= dw2[0]; J(2, 0) -= dw2[1]; J(2, 1) += dw2[0];
        return J;
    }
    Vec<double, 3> output(const Vec<double, 3>& y) const { return y; }          // :210-214
    Vec<double, 3> weight() const { return v3(mu, mu, mu); }                   // :287-291
};

// ------------------------------------------------------------------ NeelSpherical problem
// Problems/NeelRelaxationSpherical.h:32-647; state (theta, phi), noise R^3.
template <class Aniso, class S> struct SphericalCache {
    bool isRotated = false;
    S one_div_sin_t, Mx, My, Mz;
    Vec<S, 3> MagDir, e_theta, e_phi;
    Mat<S, 2, 3> P;
};

template <class Aniso> struct NeelSphericalProblem {
    static constexpr int DY = 2, DW = 3, DOUT = 3;
    NeelParams mParams; Aniso mAnisotropy; Axes ax; double mu;
    bool rotate; double minAngle, maxAngle;
    NeelSphericalProblem(const MagneticProperties& mp, double T, const Axes& axes, bool useRot, double minA)
        : mParams(calcNeelParams(mp, T)), mAnisotropy(mp), ax(axes), mu(mp.moment()),
          rotate(useRot), minAngle(minA), maxAngle(PI - minA) {}   // :109-116

    template <class S> bool needsRotation(const Vec<S, 2>& y) const {            // :546-559
        if (rotate) { const double th = val(y[0]); if (th < minAngle || th > maxAngle) return true; }
        return false;
    }
    template <class S> static Vec<S, 2> rotateFwd(const Vec<S, 2>& y) {             // :568-580
        const S sin_t = sin(y[0]);
        Vec<S, 2> r; r[0] = acos(-cos(y[1]) * sin_t); r[1] = atan2(sin(y[1]) * sin_t, cos(y[0])); return r;
    }
    template <class S> static Vec<S, 2> rotateBack(const Vec<S, 2>& y) {            // :589-602
        const S sin_t = sin(y[0]);
        Vec<S, 2> r; r[0] = acos(cos(y[1]) * sin_t); r[1] = atan2(sin(y[1]) * sin_t, -cos(y[0])); return r;
    }
    // prepareCalculations (:124-237): may rotate y in place and fills the chart cache
    template <class S> void prepare(Vec<S, 2>& y, SphericalCache<Aniso, S>& c) const {
        if (needsRotation(y)) { y = rotateFwd(y); c.isRotated = true; } else c.isRotated = false;
        const S cos_t = cos(y[0]), cos_p = cos(y[1]), sin_t = sin(y[0]), sin_p = sin(y[1]);
        S ods = S(1.0) / sin_t;
        if (std::isinf(val(ods))) ods = S(0.0);
        c.one_div_sin_t = ods;
        const double a = mParams.NeelFactor1, b = mParams.NeelFactor2;
        if (!c.isRotated) {
            c.MagDir = v3(sin_t * cos_p, sin_t * sin_p, cos_t);
            c.e_theta = v3(cos_t * cos_p, cos_t * sin_p, -sin_t);
            c.e_phi = v3(-sin_p, cos_p, S(0.0));
        } else {
            c.MagDir = v3(-cos_t, sin_t * sin_p, sin_t * cos_p);
            c.e_theta = v3(sin_t, cos_t * sin_p, cos_t * cos_p);
            c.e_phi = v3(S(0.0), cos_p, -sin_p);
        }
        for (int j = 0; j < 3; ++j) {
            c.P(0, j) = -a * c.e_phi[j] + b * c.e_theta[j];
            c.P(1, j) = ods * (a * c.e_theta[j] + b * c.e_phi[j]);
        }
        if (val(ods) == 0.0) { if (!c.isRotated) c.P(1, 2) = S(a); else c.P(1, 0) = S(-a); }
    }
    template <class S> Vec<S, 2> det(const SphericalCache<Aniso, S>& c, const Vec<double, 3>& xi) const {   // :239-265
        const Vec<S, 3> Heff = aniso_field(mAnisotropy, c.MagDir, ax) + lift3<S>(xi);
        Vec<S, 2> r;
        for (int i = 0; i < 2; ++i) r[i] = c.P(i, 0) * Heff[0] + c.P(i, 1) * Heff[1] + c.P(i, 2) * Heff[2];
        return r;
    }
    template <class S> Mat<S, 2, 3> sto(const SphericalCache<Aniso, S>& c) const {                            // :267-293
        Mat<S, 2, 3> r;
        for (int i = 0; i < 2; ++i) for (int j = 0; j < 3; ++j) r(i, j) = mParams.NoisePrefactor * c.P(i, j);
        return r;
    }
    template <class S> Vec<S, 2> drift(const SphericalCache<Aniso, S>& c) const {                             // :295-309
        const S cos_t = c.isRotated ? -c.MagDir[0] : c.MagDir[2];
        Vec<S, 2> r; r[0] = -0.5 * mParams.DriftPrefactor * cos_t * c.one_div_sin_t; r[1] = S(0.0);
        return r;
    }
    void normalize(Vec<double, 2>&) const {}                                                                 // :534-535
    template <class S> void finish(Vec<S, 2>& y, const SphericalCache<Aniso, S>& c) const {                  // :440-471
        if (c.isRotated) y = rotateBack(y);
        else { y[0] = fmod(y[0], TWO_PI); y[1] = fmod(y[1], TWO_PI); }
    }
    Vec<double, 3> output(const Vec<double, 2>& y) const {                                                   // :499-516
        const double st = std::sin(y[0]), ct = std::cos(y[0]), sp = std::sin(y[1]), cp = std::cos(y[1]);
        return v3(st * cp, st * sp, ct);
    }
    Vec<double, 3> weight() const { return v3(mu, mu, mu); }                                                 // :527-531
};

// ------------------------------------------------------------------ BrownAndNeel problem
// Problems/BrownAndNeelRelaxation.h:71-193, .inl:16-324; y = (n, m), noise R^6.
template <class Aniso, bool Simple> struct BrownAndNeelProblem {
    static constexpr int DY = 6, DW = 6, DOUT = 6;
    NeelParams mNeel; BrownParams mBrown; Aniso mAnisotropy; double MagneticMoment; bool isCubic;
    BrownAndNeelProblem(const MagneticProperties& mp, double T, double r_h, double visc)
        : mNeel(calcNeelParams(mp, T)), mBrown(calcBrownRotationParams(r_h, visc, T)), mAnisotropy(mp),
          MagneticMoment(mp.moment()), isCubic(mp.aniso == 2) {}

    template <class S> void prepare(Vec<S, 6>&) {}
    template <class S> void finish(Vec<S, 6>&) const {}
    void normalize(Vec<double, 6>& y) const { normalize3(y, 0); normalize3(y, 3); }   // BrownAndNeelRelaxation.h:168-172
    template <class S> static Vec<S, 3> head(const Vec<S, 6>& y) { return v3(y[0], y[1], y[2]); }
    template <class S> static Vec<S, 3> tail(const Vec<S, 6>& y) { return v3(y[3], y[4], y[5]); }
    template <class S> static Vec<S, 6> join(const Vec<S, 3>& a, const Vec<S, 3>& b) { Vec<S, 6> r; for (int i = 0; i < 3; ++i) { r[i] = a[i]; r[i + 3] = b[i]; } return r; }

    // getStochasticMatrixFull (.inl:39-111) / Simplified (:114-170)
    Mat<double, 6, 6> sto(const Vec<double, 6>& yi) const {
        const Vec<double, 3> ni = head(yi), ei = tail(yi);
        const double a = mNeel.NeelFactor1, b = mNeel.NeelFactor2, c = mBrown.BrownPrefactor, d = c * MagneticMoment;
        const double F = mBrown.Brown_F_Noise;
        Mat<double, 6, 6> M = Mat<double, 6, 6>::zero();
        auto skewT = [&](const Vec<double, 3>& v, int r0, int c0) {   // Brown_F / Neel_F pattern
            M(r0 + 1, c0 + 0) = -v[2]; M(r0 + 2, c0 + 0) = v[1]; M(r0 + 0, c0 + 1) = v[2];
            M(r0 + 2, c0 + 1) = -v[0]; M(r0 + 0, c0 + 2) = -v[1]; M(r0 + 1, c0 + 2) = v[0];
        };
        if (!Simple) {
            const Vec<double, 3> dni = (d * mNeel.NoisePrefactor) * ni;
            const double s = dot(dni, ei);
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M(i, 3 + j) = (i == j ? s : 0.0) - ei[i] * dni[j];
            skewT(F * ei, 3, 0);
        }
        skewT(F * ni, 0, 0);
        const double pre = (Simple ? b : (b + d)) * mNeel.NoisePrefactor;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M(3 + i, 3 + j) = pre * ((i == j ? 1.0 : 0.0) - ei[i] * ei[j]);
        const Vec<double, 3> aei = (a * mNeel.NoisePrefactor) * ei;
        M(4, 3) += aei[2]; M(5, 3) -= aei[1]; M(3, 4) -= aei[2]; M(5, 4) += aei[0]; M(3, 5) += aei[1]; M(4, 5) -= aei[0];
        return M;
    }
    // getStratonovichtoItoFull (.inl:198-219) / Simplified (:179-196)
    Vec<double, 6> drift(const Vec<double, 6>& yi) const {
        const Vec<double, 3> ni = head(yi), ei = tail(yi);
        const double a = mNeel.NeelFactor1, b = mNeel.NeelFactor2, c = mBrown.BrownPrefactor, d = c * MagneticMoment;
        const double cF_2 = mBrown.Brown_F_Noise * mBrown.Brown_F_Noise;
        const double PreH_2 = mNeel.NoisePrefactor * mNeel.NoisePrefactor;
        if (Simple) return join((-cF_2) * ni, (-((a * a + b * b) * PreH_2)) * ei);
        const double bpd = b + d;
        const Vec<double, 3> nie = dot(ei, ei) * ni;          // (n e^T) e
        const Vec<double, 3> h = (-cF_2) * ni + PreH_2 * ((d * a) * cross(ni, ei) - (0.5 * d * d) * (nie - ni));
        return join(h, (-(cF_2 + (a * a + bpd * bpd) * PreH_2)) * ei);
    }
    // getDeterministicVector (.inl:222-271); returns zero for cubic (:225-229)
    template <class S> Vec<S, 6> det(const Vec<S, 6>& yi, const Vec<double, 3>& xi) const {
        if constexpr (!std::is_same<Aniso, Uniaxial>::value) { return Vec<S, 6>::zero(); } else {
        const Vec<S, 3> ni = head(yi), ei = tail(yi);
        const Vec<S, 3> Heff = mAnisotropy.template field_axis<S>(ei, ni) + lift3<S>(xi);
        const Vec<S, 3> Teff = mAnisotropy.template torque_axis<S>(ei, ni);
        const double a = mNeel.NeelFactor1, b = mNeel.NeelFactor2, c = mBrown.BrownPrefactor, d = c * MagneticMoment;
        const Vec<S, 3> mxHeff = cross(ei, Heff);
        const Vec<S, 3> omegabrown = c * Teff + d * mxHeff;
        const Vec<S, 3> omeganeel = (-a) * Heff + (b + d) * mxHeff + c * Teff;
        return join(cross(omegabrown, ni), cross(omeganeel, ei));
        }
    }
    Vec<double, 6> output(const Vec<double, 6>& y) const { return y; }
    Vec<double, 6> weight() const { Vec<double, 6> w; for (int i = 0; i < 3; ++i) { w[i] = 1.0; w[i + 3] = MagneticMoment; } return w; }  // .h:152-159
};

// ------------------------------------------------------------------ BrownAndNeelEulerSpherical problem
// Problems/BrownAndNeelRelaxationEulerSpherical.h:40-927.  y = (phi, theta, psi, theta_N, phi_N): the Euler angles
// (R313) of the particle body and the spherical angles of the magnetisation; dim_W = 6 (torque noise, field noise);
// output (x axis, y axis, m), all weighted by mu (:674-760).  Stratonovich.
struct BNESCache {
    bool brownRot = false, neelRot = false;
    double sn[5], cs[5];                 // sines / cosines of the prepared state (StateSines, StateCosines)
    Mat<double, 3, 3> R;                 // EulerRotationMatrix: columns are the body axes in world coordinates
    Mat<double, 3, 3> E;                 // EulerProjectionMatrix: omega -> Euler angle rates
    double csctheta = 0.0;
    Vec<double, 3> M;                    // MagnetisationDir
    double ods = 0.0;                    // one_div_sin_t
    Mat<double, 2, 3> PS;                // SphericalProjectionMatrix
};

template <class Aniso> struct BrownAndNeelEulerSphericalProblem {
    static constexpr int DY = 5, DW = 6, DOUT = 9;
    NeelParams mNeel; BrownParams mBrown; Aniso mAnisotropy; double MagneticMoment;
    bool brownRotate, neelRotate; double brownMin, neelMin;
    BrownAndNeelEulerSphericalProblem(const MagneticProperties& mp, double T, double r_h, double visc, bool useEuler,
                                      double brownMinAngle, bool useSpherical, double neelMinAngle)    // :124-133
        : mNeel(calcNeelParams(mp, T)), mBrown(calcBrownRotationParams(r_h, visc, T)), mAnisotropy(mp),
          MagneticMoment(mp.moment()), brownRotate(useEuler), neelRotate(useSpherical),
          brownMin(brownMinAngle), neelMin(neelMinAngle) {}

    // needsBrownCoordRotation (:796-814), needsNeelCoordRotation (:770-788): |theta| against the window
    bool needsBrown(const Vec<double, 5>& y) const {
        const double t = std::abs(y[1]);
        return brownRotate && (t < brownMin || t > PI - brownMin);
    }
    bool needsNeel(const Vec<double, 5>& y) const {
        const double t = std::abs(y[3]);
        return neelRotate && (t < neelMin || t > PI - neelMin);
    }
    // EulertoEulerRotated (:859-877), EulerRotatedtoEuler (:880-898)
    static void eulerToRotated(Vec<double, 5>& y) {
        const double S0 = std::sin(y[0]), S1 = std::sin(y[1]), S2 = std::sin(y[2]);
        const double C0 = std::cos(y[0]), C1 = std::cos(y[1]), C2 = std::cos(y[2]);
        const double phi = std::atan2(-C2 * C0 + C1 * S0 * S2, S0 * C2 + C1 * C0 * S2);
        const double theta = std::acos(-S1 * S2);
        const double psi = std::atan2(C1, C2 * S1);
        y[0] = phi; y[1] = theta; y[2] = psi;
    }
    static void rotatedToEuler(Vec<double, 5>& y) {
        const double S0 = std::sin(y[0]), S1 = std::sin(y[1]), S2 = std::sin(y[2]);
        const double C0 = std::cos(y[0]), C1 = std::cos(y[1]), C2 = std::cos(y[2]);
        const double phi = std::atan2(C0 * C2 - C1 * S0 * S2, -C2 * S0 - C1 * C0 * S2);
        const double theta = std::acos(S1 * S2);
        const double psi = std::atan2(-C1, C2 * S1);
        y[0] = phi; y[1] = theta; y[2] = psi;
    }
    // Rotate2DSphericalCoordinate90DegreeAroundYAxis (:823-834) and its inverse (:844-855) on (theta_N, phi_N)
    static void neelToRotated(Vec<double, 5>& y) {
        const double sin_t = std::sin(y[3]);
        const double t = std::acos(-std::cos(y[4]) * sin_t), p = std::atan2(std::sin(y[4]) * sin_t, std::cos(y[3]));
        y[3] = t; y[4] = p;
    }
    static void rotatedToNeel(Vec<double, 5>& y) {
        const double sin_t = std::sin(y[3]);
        const double t = std::acos(std::cos(y[4]) * sin_t), p = std::atan2(std::sin(y[4]) * sin_t, -std::cos(y[3]));
        y[3] = t; y[4] = p;
    }
    // the part of prepareCalculations after the rotation decision (:170-361): caches of state y in the given charts
    void chartAt(const Vec<double, 5>& y, bool brownRot, bool neelRot, BNESCache& c) const {
        c.brownRot = brownRot; c.neelRot = neelRot;
        for (int i = 0; i < 5; ++i) { c.sn[i] = std::sin(y[i]); c.cs[i] = std::cos(y[i]); }
        const double cphi = c.cs[0], ctheta = c.cs[1], cpsi = c.cs[2], sphi = c.sn[0], stheta = c.sn[1], spsi = c.sn[2];
        const double cphicpsi = cphi * cpsi, sphicpsi = sphi * cpsi, cphispsi = cphi * spsi, sphispsi = sphi * spsi;
        const double cthetacpsi = ctheta * cpsi, cthetaspsi = ctheta * spsi;
        c.csctheta = 1.0 / stheta;
        if (std::isinf(c.csctheta)) c.csctheta = 0.0;
        const double csc = c.csctheta;
        if (!brownRot) {                                                             // :183-238
            c.R(0, 0) = cphicpsi - ctheta * sphispsi; c.R(0, 1) = -sphicpsi - ctheta * cphispsi; c.R(0, 2) = stheta * spsi;
            c.R(1, 0) = ctheta * sphicpsi + cphispsi; c.R(1, 1) = ctheta * cphicpsi - sphispsi; c.R(1, 2) = -stheta * cpsi;
            c.R(2, 0) = stheta * sphi;                c.R(2, 1) = stheta * cphi;                c.R(2, 2) = ctheta;
            c.E(0, 0) = spsi * csc;        c.E(0, 1) = -cpsi * csc;       c.E(0, 2) = 0.0;
            c.E(1, 0) = cpsi;              c.E(1, 1) = spsi;              c.E(1, 2) = 0.0;
            c.E(2, 0) = -cthetaspsi * csc; c.E(2, 1) = cthetacpsi * csc;  c.E(2, 2) = 1.0;
        } else {                                                                     // :239-303
            c.R(0, 0) = -stheta * sphi;               c.R(0, 1) = -stheta * cphi;               c.R(0, 2) = -ctheta;
            c.R(1, 0) = ctheta * sphicpsi + cphispsi; c.R(1, 1) = ctheta * cphicpsi - sphispsi; c.R(1, 2) = -stheta * cpsi;
            c.R(2, 0) = cphicpsi - ctheta * sphispsi; c.R(2, 1) = -sphicpsi - ctheta * cphispsi; c.R(2, 2) = stheta * spsi;
            c.E(0, 0) = 0.0;   c.E(0, 1) = -cpsi * csc;       c.E(0, 2) = spsi * csc;
            c.E(1, 0) = 0.0;   c.E(1, 1) = spsi;              c.E(1, 2) = cpsi;
            c.E(2, 0) = -1.0;  c.E(2, 1) = cthetacpsi * csc;  c.E(2, 2) = -cthetaspsi * csc;
        }
        const double cos_t = c.cs[3], cos_p = c.cs[4], sin_t = c.sn[3], sin_p = c.sn[4];
        c.ods = 1.0 / sin_t;
        if (std::isinf(c.ods)) c.ods = 0.0;
        if (!neelRot) {                                                              // :307-330
            c.M = v3(sin_t * cos_p, sin_t * sin_p, cos_t);
            c.PS(0, 0) = -sin_p; c.PS(1, 0) = -cos_p * cos_t * c.ods;
            c.PS(0, 1) = cos_p;  c.PS(1, 1) = -sin_p * cos_t * c.ods;
            c.PS(0, 2) = 0.0;    c.PS(1, 2) = 1.0;
        } else {                                                                     // :331-360
            c.M = v3(-cos_t, sin_t * sin_p, sin_t * cos_p);
            c.PS(0, 0) = 0.0;    c.PS(1, 0) = -1.0;
            c.PS(0, 1) = cos_p;  c.PS(1, 1) = -sin_p * cos_t * c.ods;
            c.PS(0, 2) = -sin_p; c.PS(1, 2) = -cos_p * cos_t * c.ods;
        }
    }
    // prepareCalculations (:142-361)
    void prepare(Vec<double, 5>& y, BNESCache& c) const {
        const bool br = needsBrown(y);
        if (br) eulerToRotated(y);
        const bool nr = needsNeel(y);
        if (nr) neelToRotated(y);
        chartAt(y, br, nr, c);
    }
    Axes axes(const BNESCache& c) const {
        Axes ax;
        ax.x = v3(c.R(0, 0), c.R(1, 0), c.R(2, 0)); ax.y = v3(c.R(0, 1), c.R(1, 1), c.R(2, 1)); ax.z = v3(c.R(0, 2), c.R(1, 2), c.R(2, 2));
        return ax;
    }
    // getDeterministicVector (:364-425)
    Vec<double, 5> det(const BNESCache& c, const Vec<double, 3>& xi) const {
        const Axes ax = axes(c);
        const Vec<double, 3> Heff = mAnisotropy.field(c.M, ax) + xi;
        const Vec<double, 3> Teff = mAnisotropy.torqueBody(c.M, ax, c.sn, c.cs);
        const double a = mNeel.NeelFactor1, b = mNeel.NeelFactor2, cc = mBrown.BrownPrefactor, d = cc * MagneticMoment;
        const Vec<double, 3> mxHeff = cross(c.M, Heff);
        const Vec<double, 3> omegabrown = cc * Teff + d * mxHeff;
        const Vec<double, 3> omeganeel = (-a) * Heff + (b + d) * mxHeff + cc * Teff;
        const Vec<double, 3> br = matvec(c.E, omegabrown);
        const Vec<double, 2> ne = matvec(c.PS, omeganeel);
        Vec<double, 5> r; r[0] = br[0]; r[1] = br[1]; r[2] = br[2]; r[3] = ne[0]; r[4] = ne[1];
        return r;
    }
    // getStochasticMatrix (:428-497)
    Mat<double, 5, 6> sto(const BNESCache& c) const {
        const double a = mNeel.NeelFactor1, b = mNeel.NeelFactor2, cc = mBrown.BrownPrefactor, d = cc * MagneticMoment;
        const double F = mBrown.Brown_F_Noise, sig = mNeel.NoisePrefactor;
        auto skew = [](const Vec<double, 3>& t) {     // the matrices built at :449-457 and :473-481: Mat v = t x v
            Mat<double, 3, 3> K = Mat<double, 3, 3>::zero();
            K(1, 0) = t[2]; K(2, 0) = -t[1]; K(0, 1) = -t[2]; K(2, 1) = t[0]; K(0, 2) = t[1]; K(1, 2) = -t[0];
            return K;
        };
        const Mat<double, 3, 3> BF = matmul(c.E, skew((d * sig) * c.M));
        Mat<double, 3, 3> NF = skew(((b + d) * sig) * c.M);
        for (int i = 0; i < 3; ++i) NF(i, i) = NF(i, i) - a * sig;
        const Mat<double, 2, 3> NH = matmul(c.PS, NF);
        Mat<double, 5, 6> S = Mat<double, 5, 6>::zero();
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) { S(i, j) = F * c.E(i, j); S(i, 3 + j) = BF(i, j); }
        for (int i = 0; i < 2; ++i) for (int j = 0; j < 3; ++j) { S(3 + i, j) = c.PS(i, j) * F; S(3 + i, 3 + j) = NH(i, j); }
        return S;
    }
    // getDrift (:500-601); y is the prepared state
    Vec<double, 5> drift(const Vec<double, 5>& y, const BNESCache& c) const {
        const double a = mNeel.NeelFactor1, b = mNeel.NeelFactor2, cc = mBrown.BrownPrefactor, d = cc * MagneticMoment;
        const double DN = mNeel.NoisePrefactor, DB = mBrown.BrownDiffusion;
        const double ctheta = c.cs[1], cpsi = c.cs[2], spsi = c.sn[2];
        const double cos_t = c.cs[3], cos_p = c.cs[4], sin_t = c.sn[3], sin_p = c.sn[4];
        const double csctheta = c.csctheta;
        const double DN_2 = DN * DN, DB_2 = DB * DB, c_2_half = 0.5 * cc * cc, d_2_fourth = 0.25 * d * d, a_d = a * d;
        const double bpd = b + d;
        Vec<double, 5> r;
        if ((!c.brownRot && !c.neelRot) || (c.brownRot && c.neelRot)) {             // :542-558
            const double cotb = csctheta * ctheta;
            const double cos_2t = cos_t * cos_t - sin_t * sin_t, sin_2t = 2.0 * cos_t * sin_t, sin_t_2 = sin_t * sin_t;
            const double cos_pm = cos_p * cpsi + spsi * sin_p, sin_pm = cpsi * sin_p - cos_p * spsi;
            const double two = 2.0 * (y[4] - y[2]);
            const double cos_2pm = std::cos(two), sin_2pm = std::sin(two);
            r[0] = -DN_2 * csctheta * sin_t * (-a * d * sin_pm + 2.0 * d_2_fourth * (cos_t * cos_pm + cotb * sin_t * sin_2pm));
            r[1] = c_2_half * DB_2 * cotb - DN_2 * a_d * (cos_pm * sin_t) +
                   d_2_fourth * DN_2 * (0.5 * (3.0 + cos_2t + 2.0 * cos_2pm * sin_t_2) * cotb - sin_2t * sin_pm);
            r[2] = -DN_2 * (a_d * (cos_t + sin_t * cotb * sin_pm) -
                            d_2_fourth * sin_t * (2.0 * cos_t * cotb * cos_pm + (-1.0 + 2.0 * csctheta * csctheta) * sin_t * sin_2pm));
        } else {
            const double sin_2t = 2.0 * cos_t * sin_t, sin_t_2 = sin_t * sin_t;
            const double cpsi_2 = cpsi * cpsi, spsi_2 = spsi * spsi, c2psi = cpsi_2 - spsi_2, s2psi = 2.0 * cpsi * spsi;
            const double sin_2p = 2.0 * cos_p * sin_p, cotb = csctheta * ctheta;
            const double cos_t_2 = cos_t * cos_t, sin_p_2 = sin_p * sin_p, cos_p_2 = cos_p * cos_p;
            if (c.brownRot) {                                                        // :559-576
                r[0] = DN_2 * csctheta * (a_d * (-spsi * cos_t + sin_p * cpsi * sin_t) +
                                          2.0 * d_2_fourth * (-c2psi * cotb * sin_2t * sin_p + cos_p * sin_t * (cos_t * cpsi + sin_t * sin_p * spsi) +
                                                              cotb * s2psi * (cos_t_2 - sin_t_2 * sin_p_2)));
                r[1] = c_2_half * DB_2 * cotb + DN_2 * (a_d * (-cos_t * cpsi - spsi * sin_t * sin_p) +
                                                        d_2_fourth * (2.0 * cpsi_2 * cotb * (cos_t_2 + cos_p_2 * sin_t_2) + cpsi * sin_t_2 * sin_2p +
                                                                      2.0 * spsi * (-sin_t * cos_t * cos_p + cotb * sin_t_2 * spsi) +
                                                                      cotb * sin_2t * sin_p * s2psi));
                r[2] = DN_2 * (a_d * (sin_t * cos_p - cpsi * cotb * sin_t * sin_p + cotb * cos_t * spsi) -
                               d_2_fourth * (cos_t * cpsi + sin_t * sin_p * spsi) *
                                   (2.0 * cos_p * cotb * sin_t - (3.0 + c2psi) * csctheta * csctheta * (cpsi * sin_t * sin_p - cos_t * spsi)));
            } else {                                                                 // :577-594
                r[0] = DN_2 * csctheta * (a_d * (spsi * cos_t + sin_p * cpsi * sin_t) +
                                          2.0 * d_2_fourth * (c2psi * cotb * sin_2t * sin_p + cos_p * sin_t * (cos_t * cpsi - sin_t * sin_p * spsi) +
                                                              cotb * s2psi * (cos_t_2 - sin_t_2 * sin_p_2)));
                r[1] = c_2_half * DB_2 * cotb + DN_2 * (a_d * (+cos_t * cpsi - spsi * sin_t * sin_p) +
                                                        d_2_fourth * (2.0 * cpsi_2 * cotb * (cos_t_2 + cos_p_2 * sin_t_2) - cpsi * sin_t_2 * sin_2p +
                                                                      2.0 * spsi * (-sin_t * cos_t * cos_p + cotb * sin_t_2 * spsi) -
                                                                      cotb * sin_2t * sin_p * s2psi));
                r[2] = DN_2 * (a_d * (-sin_t * cos_p - cpsi * cotb * sin_t * sin_p - cotb * cos_t * spsi) -
                               d_2_fourth * (cos_t * cpsi - sin_t * sin_p * spsi) *
                                   (2.0 * cos_p * cotb * sin_t + (3.0 + c2psi) * csctheta * csctheta * (cpsi * sin_t * sin_p + cos_t * spsi)));
            }
        }
        r[3] = (c_2_half * DB_2 + 0.5 * (a * a + bpd * bpd) * DN_2) * c.ods * cos_t;   // :596-598
        r[4] = 0.0;
        return r;
    }
    // finishCalculations (:635-665)
    void finish(Vec<double, 5>& y, const BNESCache& c) const {
        if (c.brownRot) rotatedToEuler(y);
        else { y[0] = std::fmod(y[0], TWO_PI); y[1] = std::fmod(y[1], TWO_PI); y[2] = std::fmod(y[2], TWO_PI); }
        if (c.neelRot) rotatedToNeel(y);
        else { y[3] = std::fmod(y[3], TWO_PI); y[4] = std::fmod(y[4], TWO_PI); }
    }
    void normalize(Vec<double, 5>&) const {}                                             // :738-739
    // calculateOutputResult (:674-735): body x axis, body y axis, magnetisation direction
    Vec<double, 9> output(const Vec<double, 5>& y) const {
        double sn[5], cs[5];
        for (int i = 0; i < 5; ++i) { sn[i] = std::sin(y[i]); cs[i] = std::cos(y[i]); }
        const double cphi = cs[0], ctheta = cs[1], cpsi = cs[2], sphi = sn[0], stheta = sn[1], spsi = sn[2];
        const double cphicpsi = cphi * cpsi, sphicpsi = sphi * cpsi, cphispsi = cphi * spsi, sphispsi = sphi * spsi;
        Vec<double, 9> o;
        o[0] = cphicpsi - ctheta * sphispsi; o[1] = ctheta * sphicpsi + cphispsi; o[2] = stheta * sphi;
        o[3] = -sphicpsi - ctheta * cphispsi; o[4] = ctheta * cphicpsi - sphispsi; o[5] = stheta * cphi;
        o[6] = sn[3] * cs[4]; o[7] = sn[3] * sn[4]; o[8] = cs[3];
        return o;
    }
    Vec<double, 9> weight() const { Vec<double, 9> w; for (int i = 0; i < 9; ++i) w[i] = MagneticMoment; return w; }   // :755-759
};

// ------------------------------------------------------------------ fields
// Fields/ZeroField.h:39, ConstantField.h:43, SinusoidalField.h:38-51, LissajousField.h:50-101,
// TriangularField.h:42-67, RectangularField.h:26-63, SincField.h:34-45
struct FieldSpec {
    int type = 1;     // 1 zero, 2 constant, 3 sinusoidal, 4 lissajous, 5 triangular, 6 rectangular, 7 sinc
    double offset[3] = {0, 0, 0}, amp[3] = {0, 0, 0}, freq[3] = {0, 0, 0}, phase[3] = {0, 0, 0};
    double periode = 0, timeOffset = 0, tau = 0;
    bool alternating = false;
    // RectangularField constructor: derived constants (RectangularField.h:26-34)
    double halfP = 0, tauInv = 0, newAmp[3] = {0, 0, 0}, maxField[3] = {0, 0, 0}, newTimeOffset = 0, newOffset[3] = {0, 0, 0};
    void prepare() {
        halfP = periode / 2.0;
        tauInv = std::abs(tau) <= std::numeric_limits<double>::min() ? std::numeric_limits<double>::max() : 1.0 / tau;
        double na = 0, mf = 0;
        for (int i = 0; i < 3; ++i) {
            newAmp[i] = alternating ? 2 * amp[i] : amp[i];
            maxField[i] = newAmp[i] * (-std::expm1(-halfP * tauInv));
            na += newAmp[i] * newAmp[i]; mf += maxField[i] * maxField[i];
        }
        newTimeOffset = alternating ? timeOffset - std::log(1 - std::sqrt(mf) / (2 * std::sqrt(na))) * tau : timeOffset;
        for (int i = 0; i < 3; ++i) newOffset[i] = alternating ? newAmp[i] * std::expm1(-newTimeOffset * tauInv) + offset[i] : offset[i];
    }
    Vec<double, 3> at(double t) const {
        if (type == 2) return v3(offset[0], offset[1], offset[2]);
        if (type == 3) {
            const double w = TWO_PI * freq[0];
            const double s = std::sin(w * t + phase[0]);
            return v3(amp[0] * s + offset[0], amp[1] * s + offset[1], amp[2] * s + offset[2]);
        }
        if (type == 4) {
            Vec<double, 3> r;
            for (int i = 0; i < 3; ++i) r[i] = offset[i] + amp[i] * std::sin(TWO_PI * freq[i] * t + phase[i]);
            return r;
        }
        if (type == 5) {
            const double nt = t + timeOffset, half = periode / 2.0;
            const double pos = nt < 0 ? periode + std::fmod(nt, periode) : std::fmod(nt, periode);
            const double f = pos > half ? (2.0 * pos / half - 1.0) : (-2.0 * pos / half + 3.0);
            return v3(f * amp[0] + offset[0], f * amp[1] + offset[1], f * amp[2] + offset[2]);
        }
        if (type == 6) {
            const double nt = t + newTimeOffset;
            const double pos = nt < 0 ? std::fmod(nt + periode, periode) : std::fmod(nt, periode);
            Vec<double, 3> r;
            if (pos <= halfP) { const double e = -std::expm1(-pos * tauInv); for (int i = 0; i < 3; ++i) r[i] = newAmp[i] * e + newOffset[i]; }
            else { const double e = std::exp(-(pos - halfP) * tauInv); for (int i = 0; i < 3; ++i) r[i] = newAmp[i] * e + newOffset[i] - (newAmp[i] - maxField[i]); }
            return r;
        }
        if (type == 7) {
            const double half = periode * 0.5, a = 11.0 / half;
            const double pos = t < 0 ? std::fmod(t + periode, periode) - half : std::fmod(t, periode) - half;
            const double x = 3.14159265358979323846 * a * pos;
            const double sc = x == 0.0 ? 1.0 : std::sin(x) / x;
            return v3(amp[0] * sc * sc + offset[0], amp[1] * sc * sc + offset[1], amp[2] * sc * sc + offset[2]);
        }
        return Vec<double, 3>::zero();
    }
};

// ------------------------------------------------------------------ noise
// Philox4x32-10 (Salmon et al. SC'11) and the engine's Box-Muller stream; an
// independent copy of the generator the product uses (same definition).
inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}
inline double u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }
// The engine's FP64 Box-Muller (its own stream definition; the reference draws
// pcg64 + Boost normals instead): ln u by a 97-entry table of ln(1 + j/128)
// and 1/(1 + j/128) plus the series of ln(1 + r) to r^7; sin/cos(2 pi u) by a
// 128-entry table of sin/cos((j + 1/2) 2 pi/128) and the angle-sum formulas
// with the series of sin x (to x^7) and cos x - 1 (to x^6).  The tables are
// computed here with the C library from the same double arguments.
struct BoxMullerTables {
    double sc[128][2], lg[97][2];
    BoxMullerTables() {
        const double step = TWO_PI / 128.0;
        for (int j = 0; j < 128; ++j) { const double t = (j + 0.5) * step; sc[j][0] = std::sin(t); sc[j][1] = std::cos(t); }
        for (int j = -32; j <= 64; ++j) { const double c = 1.0 + j / 128.0; lg[j + 32][0] = std::log(c); lg[j + 32][1] = 1.0 / c; }
    }
};
inline const BoxMullerTables& bmTables() { static const BoxMullerTables t; return t; }
inline double bm_log(uint32_t w) {
    const double LN2_HI = 0.6931471803691238, LN2_LO = 1.9082149292705877e-10;
    const double v = (double)w + 0.5;
    int e; double m = std::frexp(v, &e) * 2.0; e -= 1;
    if (m >= 1.5) { m = m * 0.5; e += 1; }
    const double jd = std::rint((m - 1.0) * 128.0);
    const double* T = bmTables().lg[(int)jd + 32];
    const double c = std::fma(jd, 0.0078125, 1.0);
    const double r = (m - c) * T[1];
    static const double A[6] = {0.14285714285714285, -0.16666666666666666, 0.2, -0.25, 0.3333333333333333, -0.5};
    double p = A[0];
    for (int i = 1; i < 6; ++i) p = std::fma(p, r, A[i]);
    const double l1p = std::fma(p, r * r, r);
    const double E = (double)(e - 32);
    return std::fma(E, LN2_HI, T[0] + std::fma(E, LN2_LO, l1p));
}
inline void bm_sincos2pi(uint32_t w, double* s, double* c) {
    const double* T = bmTables().sc[w >> 25];
    const int d = (int)(w & 0x1FFFFFFu) - 0x1000000;
    const double x = ((double)d + 0.5) * 1.4629180792671596e-09;
    const double x2 = x * x;
    double sp = std::fma(std::fma(-0.0001984126984126984, x2, 0.008333333333333333), x2, -0.16666666666666666);
    const double sx = std::fma(sp * x2, x, x);
    const double cp = std::fma(std::fma(-0.001388888888888889, x2, 0.041666666666666664), x2, -0.5);
    const double cm1 = cp * x2;
    *s = T[0] + std::fma(T[0], cm1, T[1] * sx);
    *c = T[1] + std::fma(T[1], cm1, -(T[0] * sx));
}
inline void box_muller4(const uint32_t c[4], double z[4]) {
    const double r0 = std::sqrt(-2.0 * bm_log(c[0])), r1 = std::sqrt(-2.0 * bm_log(c[2]));
    double s0, c0, s1, c1;
    bm_sincos2pi(c[1], &s0, &c0); bm_sincos2pi(c[3], &s1, &c1);
    z[0] = r0 * c0; z[1] = r0 * s0; z[2] = r1 * c1; z[3] = r1 * s1;
}
inline void normals4(uint64_t seed, uint64_t particle, uint64_t call, uint32_t tag, double z[4], bool libm = false) {
    uint32_t c[4] = {(uint32_t)call, (uint32_t)(call >> 32), (uint32_t)particle, tag};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    if (libm) {   // textbook Box-Muller with libm, for the tolerance check of bm_log / bm_sincos2pi
        const double r0 = std::sqrt(-2.0 * std::log(u01(c[0]))), r1 = std::sqrt(-2.0 * std::log(u01(c[2])));
        z[0] = r0 * std::cos(TWO_PI * u01(c[1])); z[1] = r0 * std::sin(TWO_PI * u01(c[1]));
        z[2] = r1 * std::cos(TWO_PI * u01(c[3])); z[3] = r1 * std::sin(TWO_PI * u01(c[3]));
        return;
    }
    box_muller4(c, z);
}

// per-particle standard-normal source: Philox stream or replayed buffer
struct NoiseSource {
    uint64_t seed = 0, particle = 0, n_particles = 1;
    const double* replay = nullptr; uint64_t replay_count = 0;
    uint64_t k = 0;                  // next normal index
    double buf[4]; uint64_t buf_call = ~0ull;
    double next() {
        double z;
        if (replay) z = (k < replay_count) ? replay[k * n_particles + particle] : 0.0;
        else {
            const uint64_t call = k >> 2;
            if (call != buf_call) { normals4(seed, particle, call, 0u, buf); buf_call = call; }
            z = buf[k & 3];
        }
        ++k;
        return z;
    }
};

// ------------------------------------------------------------------ small linear solve
template <int N> bool solveLinear(double A[N][N], double b[N], double x[N]) {   // Gaussian elimination, partial pivoting
    for (int k = 0; k < N; ++k) {
        int p = k; double best = std::abs(A[k][k]);
        for (int i = k + 1; i < N; ++i) if (std::abs(A[i][k]) > best) { best = std::abs(A[i][k]); p = i; }
        if (best == 0.0) return false;
        if (p != k) { for (int j = 0; j < N; ++j) std::swap(A[k][j], A[p][j]); std::swap(b[k], b[p]); }
        for (int i = k + 1; i < N; ++i) {
            const double f = A[i][k] / A[k][k];
            for (int j = k; j < N; ++j) A[i][j] -= f * A[k][j];
            b[i] -= f * b[k];
        }
    }
    for (int i = N - 1; i >= 0; --i) { double s = b[i]; for (int j = i + 1; j < N; ++j) s -= A[i][j] * x[j]; x[i] = s / A[i][i]; }
    return true;
}

// Newton iteration with the stopping rule of Implicit_Midpoint.inl:107-180
// (dx = J^-1 (-F); y += dx; stop when |dx| <= tol) and at most maxIter steps.
// fdf(y, F, J) evaluates residual and Jacobian.  Returns iterations used, or -1.
template <int N, class FDF> int newton(FDF&& fdf, double y[N], double tol, int maxIter) {
    for (int it = 1; it <= maxIter; ++it) {
        double F[N], J[N][N], dx[N];
        fdf(y, F, J);
        for (int i = 0; i < N; ++i) F[i] = -F[i];
        if (!solveLinear<N>(J, F, dx)) return -1;
        double nrm = 0.0;
        for (int i = 0; i < N; ++i) { y[i] += dx[i]; nrm += dx[i] * dx[i]; }
        if (std::sqrt(nrm) <= tol) return it;
    }
    return -1;
}

}  // namespace oracle
