// sp_oracle_sim.hpp -- solvers, simulator loop and ensemble mean of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (see sp_oracle.hpp).  Follows
//   SDEFramework/Solver/EulerMaruyama.inl:41-72, EulerMaruyama_Normalized.inl:40-55, Heun_Strong.inl:36-47,
//   Millstein.inl:34-44, Implicit_Midpoint.inl:38-188,
//   Simulator/SingleParticleSimulator.h:129-159 and :236-273,
//   Results/MeanSimulationResult.h:63-119.
// Engine decisions where the reference has no working path:
//   * Heun_Strong and Millstein take xi = field(t_{n+1}), like EulerMaruyama.
//   * NeelSpherical under Heun_Strong / Millstein: the chart (rotated or not) is
//     chosen once from y_i, every stage of the step is evaluated in that chart,
//     and the result is mapped back (finishCalculations) at the end of the step.
//   * ImplicitMidpoint uses the exact Jacobian of its residual (forward-mode
//     duals); the reference's Jacobians are approximate (NeelRelaxation.h:168,
//     CubicAnisotropy.h:228-236, BrownAndNeelRelaxation.h:178-187) but the
//     Newton fixed point is the same.
#pragma once
#include "sp_oracle.hpp"
#include <thread>
#include <atomic>

namespace oracle {

enum { SOLVER_EM = 1, SOLVER_EM_NORMALIZED = 2, SOLVER_IM = 3, SOLVER_MILSTEIN = 6, SOLVER_HEUN = 7, SOLVER_HNC = 8 };
enum { PROB_NEEL = 1, PROB_SPHERICAL = 2, PROB_BN = 3, PROB_BNES = 4 };

struct RunSpec {
    int problem, solver, aniso;
    bool bnSimple, coordTransform;
    bool eulerTransform = false;   // BrownAndNeelEulerSpherical: Use_euler_coordinate_transformation
    double brownMinAngle = 0.0;    // BrownAndNeelEulerSpherical: Brown_Min_angle_before_transformation
    bool renormalize;          // problem.normalize() after every step, any solver (engine option)
    double dt, minAngle, newtonTol;
    int newtonMaxIter;
    uint64_t nSteps, oversampling, seed, nParticles;
    FieldSpec field;
    const double* replay; uint64_t replayCount;
};

struct ParticleIn {
    MagneticProperties mp; double T, r_h, visc; Axes ax; double y0[6];
};

// ---------------------------------------------------------------- Cartesian step functions
template <class Prob, int DY>
Vec<double, DY> noiseVec(NoiseSource& ns, double sqdt) {
    Vec<double, DY> dW; for (int i = 0; i < DY; ++i) dW[i] = sqdt * ns.next(); return dW;
}

// stochastic matrix with a generic scalar (for the Newton residual)
template <class S, class Aniso> Mat<S, 3, 3> neelSto(const NeelProblem<Aniso>& P, const Vec<S, 3>& yi) { return P.template sto<S>(yi); }
template <class S, class Aniso, bool Simple> Mat<S, 6, 6> bnSto(const BrownAndNeelProblem<Aniso, Simple>& P, const Vec<S, 6>& yi) {
    // generic-scalar restatement of BrownAndNeelRelaxation.inl:39-170 (same entries as sto())
    const Vec<S, 3> ni = v3(yi[0], yi[1], yi[2]), ei = v3(yi[3], yi[4], yi[5]);
    const double a = P.mNeel.NeelFactor1, b = P.mNeel.NeelFactor2, c = P.mBrown.BrownPrefactor, d = c * P.MagneticMoment;
    const double F = P.mBrown.Brown_F_Noise, sig = P.mNeel.NoisePrefactor;
    Mat<S, 6, 6> M = Mat<S, 6, 6>::zero();
    auto skewT = [&](const Vec<S, 3>& v, int r0, int c0) {
        M(r0 + 1, c0 + 0) = -v[2]; M(r0 + 2, c0 + 0) = v[1]; M(r0 + 0, c0 + 1) = v[2];
        M(r0 + 2, c0 + 1) = -v[0]; M(r0 + 0, c0 + 2) = -v[1]; M(r0 + 1, c0 + 2) = v[0];
    };
    if (!Simple) {
        const Vec<S, 3> dni = (d * sig) * ni;
        const S s = dot(dni, ei);
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M(i, 3 + j) = (i == j ? s : S(0.0)) - ei[i] * dni[j];
        skewT(F * ei, 3, 0);
    }
    skewT(F * ni, 0, 0);
    const double pre = (Simple ? b : (b + d)) * sig;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) M(3 + i, 3 + j) = pre * (S(i == j ? 1.0 : 0.0) - ei[i] * ei[j]);
    const Vec<S, 3> aei = (a * sig) * ei;
    M(4, 3) = M(4, 3) + aei[2]; M(5, 3) = M(5, 3) - aei[1]; M(3, 4) = M(3, 4) - aei[2];
    M(5, 4) = M(5, 4) + aei[0]; M(3, 5) = M(3, 5) + aei[1]; M(4, 5) = M(4, 5) - aei[0];
    return M;
}
template <class S, class A> Mat<S, 3, 3> stoAny(const NeelProblem<A>& P, const Vec<S, 3>& y) { return neelSto<S>(P, y); }
template <class S, class A, bool Sm> Mat<S, 6, 6> stoAny(const BrownAndNeelProblem<A, Sm>& P, const Vec<S, 6>& y) { return bnSto<S>(P, y); }

template <class Prob> struct CartesianStepper {
    static constexpr int DY = Prob::DY, DW = Prob::DW;
    using Y = Vec<double, DY>;
    Prob& P; const RunSpec& rs; uint64_t newtonFail = 0;
    CartesianStepper(Prob& p, const RunSpec& r) : P(p), rs(r) {}

    Y em(const Y& yi, double time, NoiseSource& ns) {                                  // EulerMaruyama.inl:41-72
        Y res = yi; P.prepare(res);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, DW> dW = noiseVec<Prob, DW>(ns, std::sqrt(dt));
        const Y a_ = P.det(res, xi), bb = P.drift(res);
        const auto b = P.sto(res);
        const Y bdW = matvec(b, dW);
        for (int i = 0; i < DY; ++i) res[i] = res[i] + ((a_[i] + bb[i]) * dt + bdW[i]);
        P.finish(res);
        return res;
    }
    Y heun(const Y& yi, double time, NoiseSource& ns) {                                // Heun_Strong.inl:36-47
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, DW> dW = noiseVec<Prob, DW>(ns, std::sqrt(dt));
        const Y ayi = P.det(yi, xi) + P.drift(yi);
        const auto byi = P.sto(yi);
        const Y bdW = matvec(byi, dW);
        Y support; for (int i = 0; i < DY; ++i) support[i] = yi[i] + ayi[i] * dt + bdW[i];
        const Y a2 = P.det(support, xi) + P.drift(support);
        Y out; for (int i = 0; i < DY; ++i) out[i] = yi[i] + 0.5 * (a2[i] + ayi[i]) * dt + bdW[i];
        return out;
    }
    // Heun_NotConsistent.inl:34-45: dW and dW2 drawn in this order; S at y_i and at the support on dW2
    Y heunNC(const Y& yi, double time, NoiseSource& ns) {
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, DW> dW = noiseVec<Prob, DW>(ns, std::sqrt(dt));
        const Vec<double, DW> dW2 = noiseVec<Prob, DW>(ns, std::sqrt(dt));
        const Y ayi = P.det(yi, xi) + P.drift(yi);
        const auto byi = P.sto(yi);
        const Y bdW = matvec(byi, dW);
        Y support; for (int i = 0; i < DY; ++i) support[i] = yi[i] + ayi[i] * dt + bdW[i];
        const Y a2 = P.det(support, xi) + P.drift(support);
        const Y b2 = matvec(P.sto(support), dW2), b1 = matvec(byi, dW2);
        Y out; for (int i = 0; i < DY; ++i) out[i] = yi[i] + 0.5 * (a2[i] + ayi[i]) * dt + 0.5 * (b2[i] + b1[i]);
        return out;
    }
    Y milstein(const Y& yi, double time, NoiseSource& ns) {                            // Millstein.inl:34-44
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, DW> dW = noiseVec<Prob, DW>(ns, std::sqrt(dt));
        const Y a_ = P.det(yi, xi), bb = P.drift(yi);
        const Y bdW = matvec(P.sto(yi), dW);
        const double dWdW = dot(dW, dW);
        Y out; for (int i = 0; i < DY; ++i) out[i] = yi[i] + a_[i] * dt + bdW[i] + bb[i] * dWdW;
        return out;
    }
    // ImplicitMidpoint.inl:38-188.  Residual F(y) = y_i + a(ybar) dt + b(ybar) dW - y.
    template <class S> Vec<S, DY> residual(const Vec<S, DY>& y, const Y& yi, const Vec<double, 3>& xj, const Vec<double, DW>& dW) {
        Vec<S, DY> res; for (int i = 0; i < DY; ++i) res[i] = (y[i] + S(yi[i])) * 0.5;
        P.prepare(res);
        const Vec<S, DY> a = P.det(res, xj);
        const auto b = stoAny<S>(P, res);
        const Vec<S, DY> bdW = matvec(b, dW);
        for (int i = 0; i < DY; ++i) res[i] = res[i] + (a[i] * rs.dt + bdW[i]);
        P.finish(res);
        for (int i = 0; i < DY; ++i) res[i] = res[i] + (0.5 * yi[i] - 1.5 * y[i]);
        return res;
    }
    Y implicitMidpoint(const Y& yi, double time, NoiseSource& ns) {
        const double dt = rs.dt;
        const Vec<double, DW> dW = noiseVec<Prob, DW>(ns, std::sqrt(dt));
        const Vec<double, 3> xj = rs.field.at(time + 0.5 * dt);
        double y[DY]; for (int i = 0; i < DY; ++i) y[i] = yi[i];
        auto fdf = [&](const double* yv, double* F, double (*J)[DY]) {
            Vec<Dual<DY>, DY> yd; for (int i = 0; i < DY; ++i) yd[i] = Dual<DY>::var(yv[i], i);
            const Vec<Dual<DY>, DY> r = residual<Dual<DY>>(yd, yi, xj, dW);
            for (int i = 0; i < DY; ++i) { F[i] = r[i].v; for (int j = 0; j < DY; ++j) J[i][j] = r[i].d[j]; }
        };
        if (newton<DY>(fdf, y, rs.newtonTol, rs.newtonMaxIter) < 0) ++newtonFail;
        Y out; for (int i = 0; i < DY; ++i) out[i] = y[i];
        return out;
    }
};

// ---------------------------------------------------------------- spherical step functions
template <class Aniso> struct SphericalStepper {
    using Prob = NeelSphericalProblem<Aniso>;
    using Y = Vec<double, 2>;
    Prob& P; const RunSpec& rs; uint64_t newtonFail = 0;
    SphericalStepper(Prob& p, const RunSpec& r) : P(p), rs(r) {}

    // chart quantities at y in a given chart (the part of prepareCalculations after the rotation decision)
    template <class S> void chartAt(const Vec<S, 2>& y, bool rotated, SphericalCache<Aniso, S>& c) {
        // prepare() decides the rotation from theta; emulate a fixed decision
        const bool saved = P.rotate;
        P.rotate = false;
        Vec<S, 2> tmp = y;
        P.prepare(tmp, c);           // unrotated formulas on the given coordinates
        P.rotate = saved;
        if (rotated) {               // same coordinates, rotated-chart formulas (:197-236)
            const S cos_t = cos(y[0]), cos_p = cos(y[1]), sin_t = sin(y[0]), sin_p = sin(y[1]);
            const double a = P.mParams.NeelFactor1, b = P.mParams.NeelFactor2;
            c.isRotated = true;
            c.MagDir = v3(-cos_t, sin_t * sin_p, sin_t * cos_p);
            c.e_theta = v3(sin_t, cos_t * sin_p, cos_t * cos_p);
            c.e_phi = v3(S(0.0), cos_p, -sin_p);
            for (int j = 0; j < 3; ++j) {
                c.P(0, j) = -a * c.e_phi[j] + b * c.e_theta[j];
                c.P(1, j) = c.one_div_sin_t * (a * c.e_theta[j] + b * c.e_phi[j]);
            }
            if (val(c.one_div_sin_t) == 0.0) c.P(1, 0) = S(-a);
        }
    }
    Y em(const Y& yi, double time, NoiseSource& ns) {
        Y res = yi; SphericalCache<Aniso, double> c;
        P.prepare(res, c);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 3> dW = noiseVec<Prob, 3>(ns, std::sqrt(dt));
        const Y a_ = P.det(c, xi), bb = P.drift(c);
        const Y bdW = matvec(P.sto(c), dW);
        for (int i = 0; i < 2; ++i) res[i] = res[i] + ((a_[i] + bb[i]) * dt + bdW[i]);
        P.finish(res, c);
        return res;
    }
    Y heun(const Y& yi, double time, NoiseSource& ns) {
        Y res = yi; SphericalCache<Aniso, double> c0, c1;
        P.prepare(res, c0);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 3> dW = noiseVec<Prob, 3>(ns, std::sqrt(dt));
        const Y ayi = P.det(c0, xi) + P.drift(c0);
        const Y bdW = matvec(P.sto(c0), dW);
        Y support; for (int i = 0; i < 2; ++i) support[i] = res[i] + ayi[i] * dt + bdW[i];
        chartAt(support, c0.isRotated, c1);
        const Y a2 = P.det(c1, xi) + P.drift(c1);
        Y out; for (int i = 0; i < 2; ++i) out[i] = res[i] + 0.5 * (a2[i] + ayi[i]) * dt + bdW[i];
        P.finish(out, c0);
        return out;
    }
    Y heunNC(const Y& yi, double time, NoiseSource& ns) {       // Heun_NotConsistent.inl:34-45, chart of y_i
        Y res = yi; SphericalCache<Aniso, double> c0, c1;
        P.prepare(res, c0);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 3> dW = noiseVec<Prob, 3>(ns, std::sqrt(dt));
        const Vec<double, 3> dW2 = noiseVec<Prob, 3>(ns, std::sqrt(dt));
        const Y ayi = P.det(c0, xi) + P.drift(c0);
        const Y bdW = matvec(P.sto(c0), dW);
        Y support; for (int i = 0; i < 2; ++i) support[i] = res[i] + ayi[i] * dt + bdW[i];
        chartAt(support, c0.isRotated, c1);
        const Y a2 = P.det(c1, xi) + P.drift(c1);
        const Y b2 = matvec(P.sto(c1), dW2), b1 = matvec(P.sto(c0), dW2);
        Y out; for (int i = 0; i < 2; ++i) out[i] = res[i] + 0.5 * (a2[i] + ayi[i]) * dt + 0.5 * (b2[i] + b1[i]);
        P.finish(out, c0);
        return out;
    }
    Y milstein(const Y& yi, double time, NoiseSource& ns) {
        Y res = yi; SphericalCache<Aniso, double> c;
        P.prepare(res, c);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 3> dW = noiseVec<Prob, 3>(ns, std::sqrt(dt));
        const Y a_ = P.det(c, xi), bb = P.drift(c);
        const Y bdW = matvec(P.sto(c), dW);
        const double dWdW = dot(dW, dW);
        for (int i = 0; i < 2; ++i) res[i] = res[i] + a_[i] * dt + bdW[i] + bb[i] * dWdW;
        P.finish(res, c);
        return res;
    }
    template <class S> Vec<S, 2> residual(const Vec<S, 2>& y, const Y& yi, const Vec<double, 3>& xj, const Vec<double, 3>& dW) {
        Vec<S, 2> res; for (int i = 0; i < 2; ++i) res[i] = (y[i] + S(yi[i])) * 0.5;
        SphericalCache<Aniso, S> c;
        P.prepare(res, c);
        const Vec<S, 2> a = P.det(c, xj);
        const Vec<S, 2> bdW = matvec(P.sto(c), dW);
        for (int i = 0; i < 2; ++i) res[i] = res[i] + (a[i] * rs.dt + bdW[i]);
        P.finish(res, c);
        for (int i = 0; i < 2; ++i) res[i] = res[i] + (0.5 * yi[i] - 1.5 * y[i]);
        return res;
    }
    Y implicitMidpoint(const Y& yi, double time, NoiseSource& ns) {
        const double dt = rs.dt;
        const Vec<double, 3> dW = noiseVec<Prob, 3>(ns, std::sqrt(dt));
        const Vec<double, 3> xj = rs.field.at(time + 0.5 * dt);
        double y[2] = {yi[0], yi[1]};
        auto fdf = [&](const double* yv, double* F, double (*J)[2]) {
            Vec<Dual<2>, 2> yd; yd[0] = Dual<2>::var(yv[0], 0); yd[1] = Dual<2>::var(yv[1], 1);
            const Vec<Dual<2>, 2> r = residual<Dual<2>>(yd, yi, xj, dW);
            for (int i = 0; i < 2; ++i) { F[i] = r[i].v; J[i][0] = r[i].d[0]; J[i][1] = r[i].d[1]; }
        };
        if (newton<2>(fdf, y, rs.newtonTol, rs.newtonMaxIter) < 0) ++newtonFail;
        Y out; out[0] = y[0]; out[1] = y[1];
        return out;
    }
};

// ---------------------------------------------------------------- BrownAndNeelEulerSpherical step functions
// The same chart convention as SphericalStepper: the charts (Brown and Neel rotation) chosen from y_i hold for
// every stage of a step, and the result is mapped back at the end of the step.  Implicit_Midpoint has no
// counterpart: the reference's Jacobians of this problem are not implemented (assert(false),
// BrownAndNeelRelaxationEulerSpherical.h:606-671).
template <class Aniso> struct BNESStepper {
    using Prob = BrownAndNeelEulerSphericalProblem<Aniso>;
    using Y = Vec<double, 5>;
    Prob& P; const RunSpec& rs; uint64_t newtonFail = 0;
    BNESStepper(Prob& p, const RunSpec& r) : P(p), rs(r) {}
    Y em(const Y& yi, double time, NoiseSource& ns) {                                  // EulerMaruyama.inl:41-72
        Y res = yi; BNESCache c;
        P.prepare(res, c);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 6> dW = noiseVec<Prob, 6>(ns, std::sqrt(dt));
        const Y a_ = P.det(c, xi), bb = P.drift(res, c);
        const Y bdW = matvec(P.sto(c), dW);
        for (int i = 0; i < 5; ++i) res[i] = res[i] + ((a_[i] + bb[i]) * dt + bdW[i]);
        P.finish(res, c);
        return res;
    }
    Y heun(const Y& yi, double time, NoiseSource& ns) {                                // Heun_Strong.inl:36-47
        Y res = yi; BNESCache c0, c1;
        P.prepare(res, c0);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 6> dW = noiseVec<Prob, 6>(ns, std::sqrt(dt));
        const Y ayi = P.det(c0, xi) + P.drift(res, c0);
        const Y bdW = matvec(P.sto(c0), dW);
        Y support; for (int i = 0; i < 5; ++i) support[i] = res[i] + ayi[i] * dt + bdW[i];
        P.chartAt(support, c0.brownRot, c0.neelRot, c1);
        const Y a2 = P.det(c1, xi) + P.drift(support, c1);
        Y out; for (int i = 0; i < 5; ++i) out[i] = res[i] + 0.5 * (a2[i] + ayi[i]) * dt + bdW[i];
        P.finish(out, c0);
        return out;
    }
    Y heunNC(const Y& yi, double time, NoiseSource& ns) {                              // Heun_NotConsistent.inl:34-45
        Y res = yi; BNESCache c0, c1;
        P.prepare(res, c0);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 6> dW = noiseVec<Prob, 6>(ns, std::sqrt(dt));
        const Vec<double, 6> dW2 = noiseVec<Prob, 6>(ns, std::sqrt(dt));
        const Y ayi = P.det(c0, xi) + P.drift(res, c0);
        const Y bdW = matvec(P.sto(c0), dW);
        Y support; for (int i = 0; i < 5; ++i) support[i] = res[i] + ayi[i] * dt + bdW[i];
        P.chartAt(support, c0.brownRot, c0.neelRot, c1);
        const Y a2 = P.det(c1, xi) + P.drift(support, c1);
        const Y b2 = matvec(P.sto(c1), dW2), b1 = matvec(P.sto(c0), dW2);
        Y out; for (int i = 0; i < 5; ++i) out[i] = res[i] + 0.5 * (a2[i] + ayi[i]) * dt + 0.5 * (b2[i] + b1[i]);
        P.finish(out, c0);
        return out;
    }
    Y milstein(const Y& yi, double time, NoiseSource& ns) {                            // Millstein.inl:34-44
        Y res = yi; BNESCache c;
        P.prepare(res, c);
        const Vec<double, 3> xi = rs.field.at(time);
        const double dt = rs.dt;
        const Vec<double, 6> dW = noiseVec<Prob, 6>(ns, std::sqrt(dt));
        const Y a_ = P.det(c, xi), bb = P.drift(res, c);
        const Y bdW = matvec(P.sto(c), dW);
        const double dWdW = dot(dW, dW);
        for (int i = 0; i < 5; ++i) res[i] = res[i] + a_[i] * dt + bdW[i] + bb[i] * dWdW;
        P.finish(res, c);
        return res;
    }
    Y implicitMidpoint(const Y& yi, double, NoiseSource&) { ++newtonFail; return yi; }   // not offered (see above)
};

// ---------------------------------------------------------------- simulator
// One particle: SingleParticleSimulator::startSimulation (:129-159).  Writes
// nOut rows of DOUT to out (already multiplied by the weight if weighted).
template <class Prob, class Stepper>
void simulateParticle(Prob& prob, Stepper& st, const RunSpec& rs, uint64_t p, const double* y0,
                      double* out /* nOut*DOUT, weighted */, double* finalState) {
    constexpr int DY = Prob::DY, DOUT = Prob::DOUT;
    const uint64_t nOut = rs.nSteps / rs.oversampling;
    Vec<double, DY> yi; for (int i = 0; i < DY; ++i) yi[i] = y0[i];
    NoiseSource ns; ns.seed = rs.seed; ns.particle = p; ns.n_particles = rs.nParticles;
    ns.replay = rs.replay; ns.replay_count = rs.replayCount;
    const Vec<double, DOUT> w = prob.weight();
    {
        const Vec<double, DOUT> o = prob.output(yi);
        for (int c = 0; c < DOUT; ++c) out[c] = o[c] * w[c];
    }
    uint64_t counter = 0;
    for (uint64_t k = 1; k < nOut; ++k) {
        Vec<double, DOUT> sum = Vec<double, DOUT>::zero(), comp = Vec<double, DOUT>::zero();
        for (uint64_t l = rs.oversampling; l--;) {
            const double time = rs.dt * ((double)++counter);
            switch (rs.solver) {
                case SOLVER_EM: yi = st.em(yi, time, ns); break;
                case SOLVER_EM_NORMALIZED:                    // EulerMaruyama_Normalized.inl:40-55
                    yi = st.em(yi, time, ns);
                    prob.normalize(yi);
                    break;
                case SOLVER_HEUN: yi = st.heun(yi, time, ns); break;
                case SOLVER_HNC: yi = st.heunNC(yi, time, ns); break;
                case SOLVER_MILSTEIN: yi = st.milstein(yi, time, ns); break;
                default: yi = st.implicitMidpoint(yi, time, ns); break;
            }
            if (rs.renormalize) prob.normalize(yi);
            const Vec<double, DOUT> o = prob.output(yi);
            for (int c = 0; c < DOUT; ++c) {                  // Kahan summation (:148-153)
                const double y = o[c] - comp[c];
                const double t = sum[c] + y;
                comp[c] = (t - sum[c]) - y;
                sum[c] = t;
            }
        }
        for (int c = 0; c < DOUT; ++c) out[k * DOUT + c] = (sum[c] / (double)rs.oversampling) * w[c];
    }
    if (finalState) for (int i = 0; i < DY; ++i) finalState[i] = yi[i];
}

// calculateTimeAndField (SingleParticleSimulator.h:236-273)
inline void timeAndField(const RunSpec& rs, double* times, double* fields) {
    const uint64_t points = rs.nSteps / rs.oversampling;
    if (points == 0) return;
    double start = 0.0, end = 0.0;
    const double timeshift = rs.dt * ((double)rs.oversampling) * 0.5;
    times[0] = start;
    { const Vec<double, 3> f = rs.field.at(start); for (int c = 0; c < 3; ++c) fields[c] = f[c]; }
    uint64_t counter = 0;
    for (uint64_t i = 1; i < points; ++i) {
        Vec<double, 3> tmp = Vec<double, 3>::zero();
        start = end + rs.dt;
        for (uint64_t l = rs.oversampling; l--;) tmp += rs.field.at(rs.dt * ((double)++counter));
        end = rs.dt * ((double)counter) + start;
        for (int c = 0; c < 3; ++c) fields[i * 3 + c] = tmp[c] / (double)rs.oversampling;
        times[i] = (end - start) - timeshift;
    }
}

struct RunResult {
    std::vector<double> mean, weight, times, fields, finalState;
    uint64_t newtonFail = 0;
    int dout = 0;
};

// Build the problem for particle p and run it.
template <class Prob, class Stepper, class Make>
void runEnsemble(const RunSpec& rs, const ParticleIn* parts, Make&& make, RunResult& rr, int nThreads, bool keepFinal,
                 std::vector<double>* perParticleOut = nullptr) {
    constexpr int DY = Prob::DY, DOUT = Prob::DOUT;
    const uint64_t nOut = rs.nSteps / rs.oversampling, N = rs.nParticles;
    rr.dout = DOUT;
    rr.mean.assign(nOut * DOUT, 0.0); rr.weight.assign(DOUT, 0.0);
    rr.times.assign(nOut, 0.0); rr.fields.assign(nOut * 3, 0.0);
    if (keepFinal) rr.finalState.assign(N * DY, 0.0);
    timeAndField(rs, rr.times.data(), rr.fields.data());
    std::vector<double> all(N * nOut * DOUT);
    std::vector<uint64_t> fails(N, 0);
    std::atomic<uint64_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const uint64_t p = next.fetch_add(1);
            if (p >= N) break;
            Prob prob = make(parts[p]);
            Stepper st(prob, rs);
            simulateParticle(prob, st, rs, p, parts[p].y0, &all[p * nOut * DOUT], keepFinal ? &rr.finalState[p * DY] : nullptr);
            fails[p] = st.newtonFail;
        }
    };
    if (nThreads <= 1) worker();
    else { std::vector<std::thread> th; for (int i = 0; i < nThreads; ++i) th.emplace_back(worker); for (auto& t : th) t.join(); }
    // MeanSimulationResult::operator+= in particle order, then normalize() (:63-75)
    for (uint64_t p = 0; p < N; ++p) {
        Prob prob = make(parts[p]);
        const Vec<double, DOUT> w = prob.weight();
        for (int c = 0; c < DOUT; ++c) rr.weight[c] += w[c];
        for (uint64_t k = 0; k < nOut * DOUT; ++k) rr.mean[k] += all[p * nOut * DOUT + k];
        rr.newtonFail += fails[p];
    }
    bool zero = true; for (int c = 0; c < DOUT; ++c) if (rr.weight[c] != 0.0) zero = false;
    if (!zero) for (uint64_t k = 0; k < nOut; ++k) for (int c = 0; c < DOUT; ++c) rr.mean[k * DOUT + c] /= rr.weight[c];
    if (perParticleOut) *perParticleOut = std::move(all);
}

}  // namespace oracle
