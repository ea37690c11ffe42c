// oracle_capi.cpp -- extern "C" surface of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Loaded with ctypes by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg.  The product library never links this file.
#include "sp_oracle_sim.hpp"
#include <chrono>

using namespace oracle;

namespace {
MagneticProperties magFrom(const double* pr, int aniso) {
    // pr = [radius, Ms, alpha, gamma, T, K, r_h, viscosity] (+ [K_cubic, rotation(3)] for uniaxialcubic)
    MagneticProperties mp{pr[0], pr[1], pr[2], pr[3], pr[5], aniso};
    if (aniso == 4) { mp.K2 = pr[8]; mp.urot[0] = pr[9]; mp.urot[1] = pr[10]; mp.urot[2] = pr[11]; }
    return mp;
}
int propsStride(int aniso) { return aniso == 4 ? 12 : 8; }
Axes axesFrom(const double* a) {
    Axes ax;
    if (!a) { ax.x = v3(1.0, 0.0, 0.0); ax.y = v3(0.0, 1.0, 0.0); ax.z = v3(0.0, 0.0, 1.0); return ax; }
    ax.x = v3(a[0], a[1], a[2]); ax.y = v3(a[3], a[4], a[5]); ax.z = v3(a[6], a[7], a[8]);
    return ax;
}
}  // namespace

extern "C" {

void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

void or_normals(uint64_t seed, uint64_t particle, uint64_t first, uint64_t count, int libm, double* out) {
    double z[4]; uint64_t have = ~0ull;
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t k = first + i;
        if ((k >> 2) != have) { normals4(seed, particle, k >> 2, 0u, z, libm != 0); have = k >> 2; }
        out[i] = z[k & 3];
    }
}

double or_bm_log(uint32_t w) { return bm_log(w); }
void or_bm_sincos2pi(uint32_t w, double* s, double* c) { bm_sincos2pi(w, s, c); }
void or_box_muller(const uint32_t* words, uint64_t n_blocks, double* out) {
    for (uint64_t i = 0; i < n_blocks; ++i) box_muller4(words + 4 * i, out + 4 * i);
}

void or_neel_params(double r, double Ms, double alpha, double gamma, double T, double* out) {
    MagneticProperties mp{r, Ms, alpha, gamma, 0.0, 1};
    const NeelParams p = calcNeelParams(mp, T);
    out[0] = p.NeelFactor1; out[1] = p.NeelFactor2; out[2] = p.DriftPrefactor; out[3] = p.NoisePrefactor; out[4] = p.Damping;
}
void or_brown_params(double rh, double visc, double T, double* out) {
    const BrownParams p = calcBrownRotationParams(rh, visc, T);
    out[0] = p.BrownPrefactor; out[1] = p.BrownDiffusion; out[2] = p.Brown_F_Noise;
}

// anisotropy field (and Jacobian) at m: aniso 1 uniaxial (easy axis = axes z), 2 cubic
void or_aniso_field(int aniso, const double* props, const double* m, const double* axes, double* out, double* jac) {
    const MagneticProperties mp = magFrom(props, aniso);
    const Axes ax = axesFrom(axes);
    const Vec<double, 3> mv = v3(m[0], m[1], m[2]);
    Vec<double, 3> H;
    Mat<double, 3, 3> J;
    if (aniso == 1) { Uniaxial u(mp); H = u.field(mv, ax); J = u.jacobian(ax); }
    else if (aniso == 4) { UniaxialCubic uc(mp); H = uc.field(mv, ax); J = Mat<double, 3, 3>::zero(); }
    else { Cubic c(mp); H = c.field(mv, ax); J = c.jacobian(mv, ax); }
    for (int i = 0; i < 3; ++i) out[i] = H[i];
    if (jac) for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) jac[i * 3 + j] = J(i, j);
}

// dspec[3] type, [4..15] offset, amplitude, frequency, phase, [16] Periode, [17] TimeOffset, [18] Tau, [19] Alternating
static FieldSpec fieldFrom(const double* dspec) {
    FieldSpec f;
    f.type = (int)dspec[3];
    for (int i = 0; i < 3; ++i) { f.offset[i] = dspec[4 + i]; f.amp[i] = dspec[7 + i]; f.freq[i] = dspec[10 + i]; f.phase[i] = dspec[13 + i]; }
    f.periode = dspec[16]; f.timeOffset = dspec[17]; f.tau = dspec[18]; f.alternating = dspec[19] != 0.0;
    f.prepare();
    return f;
}

void or_field(const double* dspec, double t, double* out) {
    const Vec<double, 3> h = fieldFrom(dspec).at(t);
    for (int i = 0; i < 3; ++i) out[i] = h[i];
}

void or_euler_axes(double phi, double theta, double psi, double* out) {
    const Axes a = axesFromEuler(phi, theta, psi);
    for (int i = 0; i < 3; ++i) { out[i] = a.x[i]; out[3 + i] = a.y[i]; out[6 + i] = a.z[i]; }
}

// The particles of an ensemble from its particle types, restated from the engine's definition of its
// provider (the reference draws them from std::random_device with MyCEL distributions, which cannot be
// reproduced): draw k of particle i is Philox block (k, i) of stream tag 1 under the seed, a uniform is
// u(w0) and a normal the first Box-Muller normal of the block.  Per particle, in this order: the type
// (discrete: u * total count against the running counts; else the list order, each type for its count),
// the magnetic radius, the hydrodynamic shell (r_h = radius + shell) and K from their distributions
// (delta; normal mean + sd z; lognormal mean exp(s z - s^2 / 2) with s^2 = ln(1 + sd^2 / mean^2); sd is
// width * mean when the width is relative); uniaxialcubic: K_cubic, then the three rotation angles
// (U(0, 2 pi), U(0, pi), U(0, 2 pi)) when random; the Euler angles when the orientation is random
// (U(0, 2 pi) each; theta U(0, pi) for BrownAndNeelEulerSpherical, ParticleStateInitializer.h:42-62); the
// magnetisation when random (BrownAndNeelEulerSpherical: phi U(0, 2 pi), theta U(0, pi),
// ParticleStateInitializer.h:107-130; otherwise three normals, normalised); BrownAndNeel with a random
// orientation: three more normals for the axis n (BrownAndNeelRelaxation.inl:277-324).
// types: [n_types][34] = count, T, radius, Ms, alpha, gamma, K, r_h, viscosity, K_cubic, euler(3), mag(3),
// urot(3), random orientation, random magnetisation, random rotation, then four distributions (radius,
// shell, K, K_cubic) of (type 0 delta / 1 normal / 2 lognormal, width, relative).
// Outputs: props [N][12] (the layout of or_simulate), axes [N][9], y0 [N][dim_y], type [N].
void or_provide(int problem, int aniso, int n_types, int discrete, uint64_t seed, uint64_t N, const double* types,
                double* props, double* axes_out, double* y0, int64_t* type_out) {
    const int TS = 34;
    auto draw = [&](uint64_t i, uint64_t& k, bool normal) {
        double z[4];
        uint32_t c[4] = {(uint32_t)k, (uint32_t)(k >> 32), (uint32_t)i, 1u};
        ++k;
        if (normal) { normals4(seed, i, k - 1, 1u, z); return z[0]; }
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        return u01(c[0]);
    };
    auto sample = [&](const double* dist, double mean, uint64_t i, uint64_t& k) {
        const int kind = (int)dist[0];
        const double width = dist[1];
        if (kind == 0 || width == 0.0) return mean;
        const double sd = dist[2] != 0.0 ? width * mean : width;
        if (kind == 1) return mean + sd * draw(i, k, true);
        const double s2 = std::log1p(sd * sd / (mean * mean));
        return mean * std::exp(std::sqrt(s2) * draw(i, k, true) - 0.5 * s2);
    };
    const bool bnes = problem == PROB_BNES;
    const int dy = problem == PROB_NEEL ? 3 : (problem == PROB_SPHERICAL ? 2 : (bnes ? 5 : 6));
    double total = 0.0;
    for (int t = 0; t < n_types; ++t) total += types[t * TS];
    for (uint64_t i = 0; i < N; ++i) {
        uint64_t k = 0;
        int sel = 0;
        if (discrete && total > 0.0) {
            const double u = draw(i, k, false) * total;
            double run = 0.0;
            for (sel = 0; sel < n_types - 1; ++sel) { run += types[sel * TS]; if (u < run) break; }
        } else {
            uint64_t run = 0;
            for (sel = 0; sel < n_types - 1; ++sel) { run += (uint64_t)types[sel * TS]; if (i < run) break; }
        }
        const double* t = types + sel * TS;
        type_out[i] = sel;
        double* pr = props + 12 * i;
        const double radius = sample(t + 22, t[2], i, k);
        const double shell = sample(t + 25, t[7] - t[2], i, k);
        pr[0] = radius; pr[1] = t[3]; pr[2] = t[4]; pr[3] = t[5]; pr[4] = t[1];
        pr[5] = sample(t + 28, t[6], i, k);
        pr[6] = radius + shell; pr[7] = t[8];
        pr[8] = 0.0; pr[9] = pr[10] = pr[11] = 0.0;
        if (aniso == 4) {
            pr[8] = sample(t + 31, t[9], i, k);
            for (int a = 0; a < 3; ++a) pr[9 + a] = t[21] != 0.0 ? draw(i, k, false) * (a == 1 ? PI : TWO_PI) : t[16 + a];
        }
        double e[3] = {t[10], t[11], t[12]};
        if (t[19] != 0.0)
            for (int a = 0; a < 3; ++a) e[a] = draw(i, k, false) * (bnes && a == 1 ? PI : TWO_PI);
        or_euler_axes(e[0], e[1], e[2], axes_out + 9 * i);
        double m[3] = {t[13], t[14], t[15]};
        if (t[20] != 0.0 && bnes) {
            const double ph = draw(i, k, false) * TWO_PI, th = draw(i, k, false) * PI;
            m[0] = std::cos(ph) * std::sin(th); m[1] = std::sin(ph) * std::sin(th); m[2] = std::cos(th);
        } else if (t[20] != 0.0) {
            for (double& x : m) x = draw(i, k, true);
        }
        const double mn = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (double& x : m) x /= mn;
        double* y = y0 + (uint64_t)dy * i;
        if (bnes) {
            y[0] = e[0]; y[1] = e[1]; y[2] = e[2]; y[3] = std::acos(m[2]); y[4] = std::atan2(m[1], m[0]);
        } else if (problem == PROB_NEEL) {
            y[0] = m[0]; y[1] = m[1]; y[2] = m[2];
        } else if (problem == PROB_SPHERICAL) {
            y[0] = std::acos(m[2]); y[1] = std::atan2(m[1], m[0]);
        } else {
            double n[3];
            if (t[19] != 0.0) {
                for (double& x : n) x = draw(i, k, true);
            } else {   // BrownAndNeelRelaxation.inl:297-307: the axis from the Euler angles
                n[0] = std::cos(e[0]) * std::cos(e[2]) - std::sin(e[0]) * std::cos(e[1]) * std::sin(e[2]);
                n[1] = -std::cos(e[0]) * std::sin(e[2]) - std::sin(e[0]) * std::cos(e[1]) * std::cos(e[2]);
                n[2] = std::sin(e[0]) * std::sin(e[1]);
            }
            const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int a = 0; a < 3; ++a) { y[a] = n[a] / nn; y[3 + a] = m[a]; }
        }
    }
}

// Evaluate one problem term at state y with field h (golden-vector checks).
// what: 0 det, 1 drift, 2 sto (row-major), 3 Neel reference JacDet (dt), 4 Neel reference
// JacSto (h = dW), 5 NeelSpherical prepare+finish round trip, 6 NeelSpherical isRotated
// flags: bit0 BN simple model, bit1 coordinate transformation (NeelSpherical; Neel part of
// BrownAndNeelEulerSpherical), bit3 Euler coordinate transformation (its Brown part, minimum angle passed as dt);
// what 7: output vector
int or_eval(int problem, int aniso, const double* props, const double* axes, int flags, double minAngle,
            const double* y, const double* h, double dt, int what, double* out) {
    const MagneticProperties mp = magFrom(props, aniso);
    const double T = props[4];
    const Axes ax = axesFrom(axes);
    const Vec<double, 3> hv = v3(h ? h[0] : 0.0, h ? h[1] : 0.0, h ? h[2] : 0.0);
    auto neel = [&](auto an) -> int {
        using A = decltype(an);
        NeelProblem<A> P(mp, T, ax);
        const Vec<double, 3> yv = v3(y[0], y[1], y[2]);
        if (what == 0) { auto r = P.det(yv, hv); for (int i = 0; i < 3; ++i) out[i] = r[i]; return 0; }
        if (what == 1) { auto r = P.drift(yv); for (int i = 0; i < 3; ++i) out[i] = r[i]; return 0; }
        if (what == 2) { auto M = P.sto(yv); for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[i * 3 + j] = M(i, j); return 0; }
        if constexpr (std::is_same<A, Uniaxial>::value) {
            if (what == 3) { auto M = P.jacDetReference(yv, hv, dt); for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[i * 3 + j] = M(i, j); return 0; }
            if (what == 4) { auto M = P.jacStoReference(hv); for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) out[i * 3 + j] = M(i, j); return 0; }
        }
        return 1;
    };
    auto sph = [&](auto an) -> int {
        using A = decltype(an);
        NeelSphericalProblem<A> P(mp, T, ax, (flags & 2) != 0, minAngle);
        Vec<double, 2> yv; yv[0] = y[0]; yv[1] = y[1];
        SphericalCache<A, double> c;
        P.prepare(yv, c);
        if (what == 0) { auto r = P.det(c, hv); out[0] = r[0]; out[1] = r[1]; return 0; }
        if (what == 1) { auto r = P.drift(c); out[0] = r[0]; out[1] = r[1]; return 0; }
        if (what == 2) { auto M = P.sto(c); for (int i = 0; i < 2; ++i) for (int j = 0; j < 3; ++j) out[i * 3 + j] = M(i, j); return 0; }
        if (what == 5) { P.finish(yv, c); out[0] = yv[0]; out[1] = yv[1]; return 0; }
        if (what == 6) { out[0] = c.isRotated ? 1.0 : 0.0; return 0; }
        return 1;
    };
    auto bnes = [&](auto an) -> int {   // flags bit1 spherical, bit3 Euler transformation; dt carries the Brown angle
        using A = decltype(an);
        BrownAndNeelEulerSphericalProblem<A> P(mp, T, props[6], props[7], (flags & 8) != 0, dt, (flags & 2) != 0, minAngle);
        Vec<double, 5> yv; for (int i = 0; i < 5; ++i) yv[i] = y[i];
        BNESCache c;
        P.prepare(yv, c);
        if (what == 0) { auto r = P.det(c, hv); for (int i = 0; i < 5; ++i) out[i] = r[i]; return 0; }
        if (what == 1) { auto r = P.drift(yv, c); for (int i = 0; i < 5; ++i) out[i] = r[i]; return 0; }
        if (what == 2) { auto M = P.sto(c); for (int i = 0; i < 5; ++i) for (int j = 0; j < 6; ++j) out[i * 6 + j] = M(i, j); return 0; }
        if (what == 5) { for (int i = 0; i < 5; ++i) out[i] = yv[i]; P.finish(yv, c); for (int i = 0; i < 5; ++i) out[5 + i] = yv[i]; return 0; }
        if (what == 6) { out[0] = c.brownRot ? 1.0 : 0.0; out[1] = c.neelRot ? 1.0 : 0.0; return 0; }
        if (what == 7) { auto o = P.output(yv); for (int i = 0; i < 9; ++i) out[i] = o[i]; return 0; }
        return 1;
    };
    auto bn = [&](auto an, auto simple) -> int {
        using A = decltype(an);
        BrownAndNeelProblem<A, decltype(simple)::value> P(mp, T, props[6], props[7]);
        Vec<double, 6> yv; for (int i = 0; i < 6; ++i) yv[i] = y[i];
        if (what == 0) { auto r = P.det(yv, hv); for (int i = 0; i < 6; ++i) out[i] = r[i]; return 0; }
        if (what == 1) { auto r = P.drift(yv); for (int i = 0; i < 6; ++i) out[i] = r[i]; return 0; }
        if (what == 2) { auto M = P.sto(yv); for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) out[i * 6 + j] = M(i, j); return 0; }
        return 1;
    };
    if (problem == PROB_NEEL) return aniso == 1 ? neel(Uniaxial(mp)) : (aniso == 4 ? neel(UniaxialCubic(mp)) : neel(Cubic(mp)));
    if (problem == PROB_SPHERICAL) return aniso == 1 ? sph(Uniaxial(mp)) : (aniso == 4 ? sph(UniaxialCubic(mp)) : sph(Cubic(mp)));
    if (problem == PROB_BNES) return aniso == 1 ? bnes(Uniaxial(mp)) : (aniso == 4 ? bnes(UniaxialCubic(mp)) : bnes(Cubic(mp)));
    if (problem == PROB_BN) {
        const bool simple = flags & 1;
        if (aniso == 1) return simple ? bn(Uniaxial(mp), std::true_type{}) : bn(Uniaxial(mp), std::false_type{});
        return simple ? bn(Cubic(mp), std::true_type{}) : bn(Cubic(mp), std::false_type{});
    }
    return 2;
}

// Newton on the reference's Implicit_Solver_Test functions (Test_Function_1..3.h)
int or_newton_testfn(int fn, const double* x0, double tol, int maxit, double* xout) {
    if (fn == 1 || fn == 2) {
        double y[2] = {x0[0], x0[1]};
        auto fdf = [&](const double* x, double* F, double (*J)[2]) {
            if (fn == 1) {
                F[0] = x[0] + 2.0 * x[1] - 9.0; F[1] = 3.0 * x[0] + x[1] + 6 * x[0] * x[1] - 64;
                J[0][0] = 1; J[0][1] = 2; J[1][0] = 3 + 6 * x[1]; J[1][1] = 1 + 6 * x[0];
            } else {
                F[0] = 16.0 * x[0] + 2.0 * std::pow(x[1], 3) - 9.0; F[1] = 3 * x[0] * x[0] + std::pow(x[1], 4) + 6 * x[0] * x[1] - 64;
                J[0][0] = 16.0; J[0][1] = 6.0 * x[1] * x[1]; J[1][0] = 6.0 * x[0] + 6.0 * x[1]; J[1][1] = 6.0 * x[0] + 4.0 * std::pow(x[1], 3);
            }
        };
        const int it = newton<2>(fdf, y, tol, maxit);
        xout[0] = y[0]; xout[1] = y[1];
        return it;
    }
    double y[3] = {x0[0], x0[1], x0[2]};
    auto fdf = [&](const double* x, double* F, double (*J)[3]) {
        F[0] = x[0] * x[0] - 3.0 * x[2] * x[2] - 9; F[1] = 16.0 * x[1] + x[1] * x[1] * x[2] + 64; F[2] = x[0] * x[1] * x[1] - x[2] * x[2] + 16;
        J[0][0] = 2.0 * x[0]; J[0][1] = 0.0; J[0][2] = -6.0 * x[2];
        J[1][0] = 0.0; J[1][1] = 16.0 + 2.0 * x[1] * x[2]; J[1][2] = x[1] * x[1];
        J[2][0] = x[1] * x[1]; J[2][1] = 2.0 * x[0] * x[1]; J[2][2] = -2.0 * x[2];
    };
    const int it = newton<3>(fdf, y, tol, maxit);
    xout[0] = y[0]; xout[1] = y[1]; xout[2] = y[2];
    return it;
}

// Full ensemble run.  ispec = [problem, solver, aniso, flags, newton_max_iter, nthreads];
// uspec = [n_particles, n_steps, oversampling, seed, replay_count];
// dspec = [dt, min_angle, newton_tol, field_type, offset(3), amp(3), freq(3), phase(3), Periode, TimeOffset, Tau,
// Alternating, brown_min_angle]; ispec[3] flags: 1 BN simple, 2 spherical chart rotation, 4 renormalize, 8 Euler chart
// rotation (BrownAndNeelEulerSpherical)
// props [N][8], axes [N][9] (or NULL), y0 [N][dim_y].
// Outputs: mean [n_out*dout], weight [dout], times [n_out], fields [n_out*3],
// final_state [N*dim_y] (or NULL), per_particle [N*n_out*dout] weighted outputs (or NULL).
// Returns wall seconds in stats[0], newton failures in stats[1].
int or_simulate(const int64_t* ispec, const uint64_t* uspec, const double* dspec,
                const double* props, const double* axes, const double* y0, const double* replay,
                double* mean, double* weight, double* times, double* fields, double* final_state,
                double* per_particle, double* stats) {
    RunSpec rs{};
    rs.problem = (int)ispec[0]; rs.solver = (int)ispec[1]; rs.aniso = (int)ispec[2];
    rs.bnSimple = (ispec[3] & 1) != 0; rs.coordTransform = (ispec[3] & 2) != 0; rs.renormalize = (ispec[3] & 4) != 0;
    rs.eulerTransform = (ispec[3] & 8) != 0;
    rs.newtonMaxIter = (int)ispec[4];
    const int nThreads = (int)ispec[5];
    if (rs.solver != SOLVER_EM && rs.solver != SOLVER_EM_NORMALIZED && rs.solver != SOLVER_IM && rs.solver != SOLVER_MILSTEIN &&
        rs.solver != SOLVER_HEUN && rs.solver != SOLVER_HNC)
        return 3;
    rs.nParticles = uspec[0]; rs.nSteps = uspec[1]; rs.oversampling = uspec[2]; rs.seed = uspec[3];
    rs.replayCount = uspec[4]; rs.replay = replay;
    rs.dt = dspec[0]; rs.minAngle = dspec[1]; rs.newtonTol = dspec[2]; rs.brownMinAngle = dspec[20];
    rs.field.type = (int)dspec[3];
    rs.field = fieldFrom(dspec);
    const uint64_t N = rs.nParticles;
    const int dy = rs.problem == PROB_NEEL ? 3 : (rs.problem == PROB_SPHERICAL ? 2 : (rs.problem == PROB_BNES ? 5 : 6));
    if (rs.problem == PROB_BNES && rs.solver == SOLVER_IM) return 3;
    std::vector<ParticleIn> parts(N);
    const int ps = propsStride(rs.aniso);
    for (uint64_t p = 0; p < N; ++p) {
        parts[p].mp = magFrom(props + p * ps, rs.aniso);
        parts[p].T = props[p * ps + 4]; parts[p].r_h = props[p * ps + 6]; parts[p].visc = props[p * ps + 7];
        parts[p].ax = axesFrom(axes ? axes + p * 9 : nullptr);
        for (int i = 0; i < dy; ++i) parts[p].y0[i] = y0[p * dy + i];
    }
    RunResult rr;
    std::vector<double> pp;
    const auto t0 = std::chrono::steady_clock::now();
    const bool keep = final_state != nullptr;
    std::vector<double>* ppo = per_particle ? &pp : nullptr;
#define NEEL_RUN(A) runEnsemble<NeelProblem<A>, CartesianStepper<NeelProblem<A>>>(rs, parts.data(), [&](const ParticleIn& q) { return NeelProblem<A>(q.mp, q.T, q.ax); }, rr, nThreads, keep, ppo)
#define SPH_RUN(A) runEnsemble<NeelSphericalProblem<A>, SphericalStepper<A>>(rs, parts.data(), [&](const ParticleIn& q) { return NeelSphericalProblem<A>(q.mp, q.T, q.ax, rs.coordTransform, rs.minAngle); }, rr, nThreads, keep, ppo)
    if (rs.problem == PROB_NEEL) {
        if (rs.aniso == 1) NEEL_RUN(Uniaxial); else if (rs.aniso == 4) NEEL_RUN(UniaxialCubic); else NEEL_RUN(Cubic);
    } else if (rs.problem == PROB_SPHERICAL) {
        if (rs.aniso == 1) SPH_RUN(Uniaxial); else if (rs.aniso == 4) SPH_RUN(UniaxialCubic); else SPH_RUN(Cubic);
    } else if (rs.problem == PROB_BNES) {
#define BNES_RUN(A) runEnsemble<BrownAndNeelEulerSphericalProblem<A>, BNESStepper<A>>(rs, parts.data(), [&](const ParticleIn& q) { return BrownAndNeelEulerSphericalProblem<A>(q.mp, q.T, q.r_h, q.visc, rs.eulerTransform, rs.brownMinAngle, rs.coordTransform, rs.minAngle); }, rr, nThreads, keep, ppo)
        if (rs.aniso == 1) BNES_RUN(Uniaxial); else if (rs.aniso == 4) BNES_RUN(UniaxialCubic); else BNES_RUN(Cubic);
#undef BNES_RUN
    } else if (rs.problem == PROB_BN) {
        if (rs.aniso == 4) return 2;   // the reference's torque is NaN for this combination
#define BN_RUN(A, S) runEnsemble<BrownAndNeelProblem<A, S>, CartesianStepper<BrownAndNeelProblem<A, S>>>(rs, parts.data(), [&](const ParticleIn& q) { return BrownAndNeelProblem<A, S>(q.mp, q.T, q.r_h, q.visc); }, rr, nThreads, keep, ppo)
        if (rs.aniso == 1) { if (rs.bnSimple) BN_RUN(Uniaxial, true); else BN_RUN(Uniaxial, false); }
        else { if (rs.bnSimple) BN_RUN(Cubic, true); else BN_RUN(Cubic, false); }
#undef BN_RUN
    } else return 2;
#undef NEEL_RUN
#undef SPH_RUN
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::copy(rr.mean.begin(), rr.mean.end(), mean);
    std::copy(rr.weight.begin(), rr.weight.end(), weight);
    if (times) std::copy(rr.times.begin(), rr.times.end(), times);
    if (fields) std::copy(rr.fields.begin(), rr.fields.end(), fields);
    if (keep) std::copy(rr.finalState.begin(), rr.finalState.end(), final_state);
    if (per_particle) std::copy(pp.begin(), pp.end(), per_particle);
    stats[0] = secs; stats[1] = (double)rr.newtonFail;
    return 0;
}

}  // extern "C"
