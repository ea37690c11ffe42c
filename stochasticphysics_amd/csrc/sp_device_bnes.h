// sp_device_bnes.h -- BrownAndNeelEulerSpherical on gfx950: the body's R313 Euler angles and the
// magnetisation's spherical angles, y = (phi, theta, psi, theta_N, phi_N), dim_W = 6, output (body x axis,
// body y axis, m), all weighted by mu.  Restates Problems/BrownAndNeelRelaxationEulerSpherical.h; each piece
// names the lines it follows.  One particle per lane as for the other problems (sp_device.h).
#pragma once
#include "sp_device.h"

namespace sp {

template <class R> struct Trig5 { R s[5], c[5]; };

// the sines and cosines of the angles [I0, I1) in double whatever the precision of the bulk (near a chart switch)
template <int I0, int I1, class R> __device__ __forceinline__ void trig_exact(const double* y, Trig5<R>& T) {
#pragma unroll
    for (int i = I0; i < I1; ++i) {
        double sd, cd;
        sincos_fast(y[i], &sd, &cd);
        T.s[i] = (R)sd; T.c[i] = (R)cd;
    }
}
// the same as the step takes them in the standard charts: the table-driven FP64 sin/cos, or in the mixed mode the
// hardware FP32 ones after a reduction to one turn in double (they take revolutions)
template <int I0, int I1> __device__ __forceinline__ void trig_std(const double* y, Trig5<double>& T) { trig_exact<I0, I1>(y, T); }
template <int I0, int I1> __device__ __forceinline__ void trig_std(const double* y, Trig5<float>& T) {
#pragma unroll
    for (int i = I0; i < I1; ++i) {
        double r = y[i] * INV_TWO_PI;
        r -= floor(r);
        const float f = (float)r;
        T.s[i] = __builtin_amdgcn_sinf(f);
        T.c[i] = __builtin_amdgcn_cosf(f);
    }
}
template <class R> __device__ __forceinline__ void trig5(const double* y, Trig5<R>& T) { trig_std<0, 5>(y, T); }
template <class R> __device__ __forceinline__ void trig5_exact(const double* y, Trig5<R>& T) { trig_exact<0, 5>(y, T); }

template <int ANISO, class R> struct BNESLane {
    R a, bpd, c, d;              // NeelFactor1, b + d, BrownPrefactor, d = c mu
    R cFs, dss, bpdss, ass;      // cF, d sigma, (b + d) sigma, a sigma, each times sqrt(dt)
    R DN2, hcF2, ad, d2f, nd;    // sigma^2, cF^2 / 2, a d, d^2 / 4, cF^2 / 2 + (a^2 + (b + d)^2) sigma^2 / 2
    R dt;
    R k1, kv;                    // anisotropy (cubic part for uniaxialcubic): -2K/Ms, -2K V
    R ku, kuv;                   // uniaxialcubic: -2K_u/Ms, -2K_u V
    R Ru[9];                     // uniaxialcubic: rotation of the uniaxial part, row-major
    double w;
};

template <int ANISO, class R>
__device__ __forceinline__ void bnes_load(const DevRequest& rq, uint64_t p, BNESLane<ANISO, R>& L) {
    const double a = ld(rq, P_A, p), b = ld(rq, P_B, p), s = ld(rq, P_SIG, p);
    const double c = ld(rq, P_C, p), cF = ld(rq, P_CF, p), d = ld(rq, P_DD, p);
    const double sq = rq.sqdt, bpd = b + d;
    L.a = (R)a; L.bpd = (R)bpd; L.c = (R)c; L.d = (R)d;
    L.cFs = (R)(cF * sq); L.dss = (R)(d * s * sq); L.bpdss = (R)(bpd * s * sq); L.ass = (R)(a * s * sq);
    // the reference's c^2 D_B^2 / 2 is cF^2 / 2 (cF = c D_B, ParameterCalculatorBrown.h:45-56)
    L.DN2 = (R)(s * s); L.hcF2 = (R)(0.5 * cF * cF); L.ad = (R)(a * d); L.d2f = (R)(0.25 * d * d);
    L.nd = (R)(0.5 * cF * cF + 0.5 * (a * a + bpd * bpd) * (s * s));
    L.dt = (R)rq.dt;
    L.k1 = (R)ld(rq, P_K1, p); L.kv = (R)ld(rq, P_PT, p);
    if (ANISO == 4) {
        L.ku = (R)ld(rq, P_KU, p); L.kuv = (R)ld(rq, P_U, p);
#pragma unroll
        for (int k = 0; k < 9; ++k) L.Ru[k] = (R)ld(rq, P_AX + k, p);
    } else {
        L.ku = L.kuv = (R)0;
#pragma unroll
        for (int k = 0; k < 9; ++k) L.Ru[k] = (R)0;
    }
    L.w = ld(rq, P_MU, p);
}

// chart quantities of prepareCalculations after the rotation decision (:183-360): body axes (columns of the
// EulerRotationMatrix), EulerProjectionMatrix E, MagnetisationDir M and the SphericalProjectionMatrix rows
template <class R> struct BChart {
    V3<R> ex, ey, ez;
    R E[3][3];
    R csc;
    V3<R> M;
    R ods;
    V3<R> P0, P1;
};
template <class R>
__device__ __forceinline__ void bnes_chart(const Trig5<R>& T, bool brot, bool nrot, BChart<R>& C) {
    const R cphi = T.c[0], ctheta = T.c[1], cpsi = T.c[2], sphi = T.s[0], stheta = T.s[1], spsi = T.s[2];
    const R cphicpsi = cphi * cpsi, sphicpsi = sphi * cpsi, cphispsi = cphi * spsi, sphispsi = sphi * spsi;
    const R cthetacpsi = ctheta * cpsi, cthetaspsi = ctheta * spsi;
    R csc = (R)1 / stheta;
    if (isinf(csc)) csc = (R)0;
    C.csc = csc;
    const V3<R> bx = {cphicpsi - ctheta * sphispsi, ctheta * sphicpsi + cphispsi, stheta * sphi};
    const V3<R> by = {-sphicpsi - ctheta * cphispsi, ctheta * cphicpsi - sphispsi, stheta * cphi};
    const V3<R> bz = {stheta * spsi, -stheta * cpsi, ctheta};
    if (!brot) {                                                    // :183-238
        C.ex = bx; C.ey = by; C.ez = bz;
        C.E[0][0] = spsi * csc;        C.E[0][1] = -cpsi * csc;       C.E[0][2] = (R)0;
        C.E[1][0] = cpsi;              C.E[1][1] = spsi;              C.E[1][2] = (R)0;
        C.E[2][0] = -cthetaspsi * csc; C.E[2][1] = cthetacpsi * csc;  C.E[2][2] = (R)1;
    } else {                                                        // :239-303: rows cycled back to world axes
        C.ex = {-bx.z, bx.y, bx.x};
        C.ey = {-by.z, by.y, by.x};
        C.ez = {-bz.z, bz.y, bz.x};
        C.E[0][0] = (R)0;  C.E[0][1] = -cpsi * csc;       C.E[0][2] = spsi * csc;
        C.E[1][0] = (R)0;  C.E[1][1] = spsi;              C.E[1][2] = cpsi;
        C.E[2][0] = (R)-1; C.E[2][1] = cthetacpsi * csc;  C.E[2][2] = -cthetaspsi * csc;
    }
    const R cos_t = T.c[3], cos_p = T.c[4], sin_t = T.s[3], sin_p = T.s[4];
    R ods = (R)1 / sin_t;
    if (isinf(ods)) ods = (R)0;
    C.ods = ods;
    if (!nrot) {                                                    // :307-330
        C.M = {sin_t * cos_p, sin_t * sin_p, cos_t};
        C.P0 = {-sin_p, cos_p, (R)0};
        C.P1 = {-cos_p * cos_t * ods, -sin_p * cos_t * ods, (R)1};
    } else {                                                        // :331-360
        C.M = {-cos_t, sin_t * sin_p, sin_t * cos_p};
        C.P0 = {(R)0, cos_p, -sin_p};
        C.P1 = {(R)-1, -sin_p * cos_t * ods, -cos_p * cos_t * ods};
    }
}

template <class R> __device__ __forceinline__ V3<R> mat3v(const R (&E)[3][3], V3<R> v) {
    return {E[0][0] * v.x + E[0][1] * v.y + E[0][2] * v.z, E[1][0] * v.x + E[1][1] * v.y + E[1][2] * v.z,
            E[2][0] * v.x + E[2][1] * v.y + E[2][2] * v.z};
}

// cubic part: field (CubicAnisotropy.h:109-146) and torque (:161-193) with the Euler sines / cosines
template <class R>
__device__ __forceinline__ void bnes_cubic(R k1, R kv, const BChart<R>& C, const Trig5<R>& T, V3<R>& H, V3<R>& Tq) {
    const R c1 = dot(C.M, C.ex), c2 = dot(C.M, C.ey), c3 = dot(C.M, C.ez);
    const R c1s = c1 * c1, c2s = c2 * c2, c3s = c3 * c3;
    H = k1 * (((c2s + c3s) * c1) * C.ex + ((c1s + c3s) * c2) * C.ey + ((c1s + c2s) * c3) * C.ez);
    const R ctheta = T.c[1], stheta = T.s[1], cpsi = T.c[2], spsi = T.s[2];
    const V3<R> sym = cross(C.ez, (c3 * (c1s + c2s)) * C.M);
    const V3<R> th = {ctheta * spsi, -ctheta * cpsi, -stheta};
    const V3<R> asd = {(th.x * ctheta) / stheta + C.ez.x, (th.y * ctheta) / stheta + C.ez.y,
                       (th.z * ctheta) / stheta + C.ez.z};
    Tq = kv * (sym + (c1 * c2 * (c2s - c1s)) * asd);
}

// anisotropy field + applied field, and the effective torque (:380-382)
template <int ANISO, class R>
__device__ __forceinline__ void bnes_field(const BNESLane<ANISO, R>& L, const BChart<R>& C, const Trig5<R>& T, V3<R> h,
                                           V3<R>& H, V3<R>& Tq) {
    if (ANISO == 1) {                                               // UniaxialAnisotropy.h:141-150
        const R mz = dot(C.M, C.ez);
        H = (L.k1 * mz) * C.ez + h;
        Tq = (L.kv * mz) * cross(C.ez, C.M);
        return;
    }
    V3<R> Hc, Tc;
    bnes_cubic(L.k1, L.kv, C, T, Hc, Tc);
    if (ANISO == 2) {
        H = Hc + h;
        Tq = Tc;
        return;
    }
    // UniaxialCubicAnisotropy.hpp:114, :126: field about u = R^T z, torque of the uniaxial part at R m
    const V3<R> u = {L.Ru[0] * C.ez.x + L.Ru[3] * C.ez.y + L.Ru[6] * C.ez.z,
                     L.Ru[1] * C.ez.x + L.Ru[4] * C.ez.y + L.Ru[7] * C.ez.z,
                     L.Ru[2] * C.ez.x + L.Ru[5] * C.ez.y + L.Ru[8] * C.ez.z};
    const V3<R> Rm = {L.Ru[0] * C.M.x + L.Ru[1] * C.M.y + L.Ru[2] * C.M.z,
                      L.Ru[3] * C.M.x + L.Ru[4] * C.M.y + L.Ru[5] * C.M.z,
                      L.Ru[6] * C.M.x + L.Ru[7] * C.M.y + L.Ru[8] * C.M.z};
    H = ((L.ku * dot(C.M, u)) * u + Hc) + h;
    Tq = cross(C.ez, (L.kuv * dot(Rm, C.ez)) * Rm) + Tc;
}

// getDrift (:500-601), in the chart of the prepared state
template <int ANISO, class R>
__device__ __forceinline__ void bnes_drift(const BNESLane<ANISO, R>& L, const BChart<R>& C, const Trig5<R>& T, bool brot,
                                           bool nrot, R dr[4]) {
    const R ctheta = T.c[1], cpsi = T.c[2], spsi = T.s[2];
    const R cos_t = T.c[3], cos_p = T.c[4], sin_t = T.s[3], sin_p = T.s[4];
    const R csc = C.csc, cotb = csc * ctheta;
    const R DN2 = L.DN2, ad = L.ad, d2f = L.d2f;
    const R sin_2t = (R)2 * cos_t * sin_t, sin_t_2 = sin_t * sin_t;
    if (brot == nrot) {                                             // :542-558
        const R cos_2t = cos_t * cos_t - sin_t * sin_t;
        const R cos_pm = cos_p * cpsi + spsi * sin_p, sin_pm = cpsi * sin_p - cos_p * spsi;
        const R cos_2pm = cos_pm * cos_pm - sin_pm * sin_pm, sin_2pm = (R)2 * sin_pm * cos_pm;   // of 2 (phi_N - psi)
        dr[0] = -DN2 * csc * sin_t * (-ad * sin_pm + (R)2 * d2f * (cos_t * cos_pm + cotb * sin_t * sin_2pm));
        dr[1] = L.hcF2 * cotb - DN2 * ad * (cos_pm * sin_t) +
                d2f * DN2 * ((R)0.5 * ((R)3 + cos_2t + (R)2 * cos_2pm * sin_t_2) * cotb - sin_2t * sin_pm);
        dr[2] = -DN2 * (ad * (cos_t + sin_t * cotb * sin_pm) -
                        d2f * sin_t * ((R)2 * cos_t * cotb * cos_pm + ((R)-1 + (R)2 * csc * csc) * sin_t * sin_2pm));
    } else {
        const R cpsi_2 = cpsi * cpsi, spsi_2 = spsi * spsi, c2psi = cpsi_2 - spsi_2, s2psi = (R)2 * cpsi * spsi;
        const R sin_2p = (R)2 * cos_p * sin_p, cos_t_2 = cos_t * cos_t, sin_p_2 = sin_p * sin_p, cos_p_2 = cos_p * cos_p;
        // the Brown-rotated (:559-576) and the Neel-rotated (:577-594) formulas differ only in signs: sg = +1 / -1
        const R sg = brot ? (R)1 : (R)-1;
        dr[0] = DN2 * csc * (ad * (-sg * spsi * cos_t + sin_p * cpsi * sin_t) +
                             (R)2 * d2f * (-sg * c2psi * cotb * sin_2t * sin_p + cos_p * sin_t * (cos_t * cpsi + sg * sin_t * sin_p * spsi) +
                                           cotb * s2psi * (cos_t_2 - sin_t_2 * sin_p_2)));
        dr[1] = L.hcF2 * cotb + DN2 * (ad * (-sg * cos_t * cpsi - spsi * sin_t * sin_p) +
                                       d2f * ((R)2 * cpsi_2 * cotb * (cos_t_2 + cos_p_2 * sin_t_2) + sg * cpsi * sin_t_2 * sin_2p +
                                              (R)2 * spsi * (-sin_t * cos_t * cos_p + cotb * sin_t_2 * spsi) +
                                              sg * cotb * sin_2t * sin_p * s2psi));
        dr[2] = DN2 * (ad * (sg * sin_t * cos_p - cpsi * cotb * sin_t * sin_p + sg * cotb * cos_t * spsi) -
                       d2f * (cos_t * cpsi + sg * sin_t * sin_p * spsi) *
                           ((R)2 * cos_p * cotb * sin_t - sg * ((R)3 + c2psi) * csc * csc * (cpsi * sin_t * sin_p - sg * cos_t * spsi)));
    }
    dr[3] = L.nd * C.ods * cos_t;                                   // :597
}

// deterministic rates (:364-425) plus the drift (:500-601); f[0..2] Euler angles, f[3..4] Neel angles
template <int ANISO, class R>
__device__ __forceinline__ void bnes_rates(const BNESLane<ANISO, R>& L, const BChart<R>& C, const Trig5<R>& T, bool brot,
                                           bool nrot, V3<R> h, bool with_drift, R f[5]) {
    V3<R> H, Tq;
    bnes_field<ANISO>(L, C, T, h, H, Tq);
    const V3<R> mxH = cross(C.M, H);
    const V3<R> wB = L.c * Tq + L.d * mxH;
    const V3<R> wN = (L.bpd * mxH + L.c * Tq) - L.a * H;
    const V3<R> eb = mat3v(C.E, wB);
    f[0] = eb.x; f[1] = eb.y; f[2] = eb.z;
    f[3] = dot(C.P0, wN); f[4] = dot(C.P1, wN);
    if (!with_drift) return;
    R dr[4];
    bnes_drift(L, C, T, brot, nrot, dr);
#pragma unroll
    for (int i = 0; i < 4; ++i) f[i] += dr[i];
}

// S dW (:428-497): E (cF dW_F + d sigma (M x dW_H)) and P (cF dW_F + (b + d) sigma (M x dW_H) - a sigma dW_H)
template <int ANISO, class R>
__device__ __forceinline__ void bnes_noise(const BNESLane<ANISO, R>& L, const BChart<R>& C, V3<R> zF, V3<R> zH, R s[5]) {
    const V3<R> wF = L.cFs * zF;
    const V3<R> u = cross(C.M, zH);
    const V3<R> eb = mat3v(C.E, wF + L.dss * u);
    const V3<R> wn = (wF + L.bpdss * u) - L.ass * zH;
    s[0] = eb.x; s[1] = eb.y; s[2] = eb.z;
    s[3] = dot(C.P0, wn); s[4] = dot(C.P1, wn);
}

// EulertoEulerRotated (:859-877), EulerRotatedtoEuler (:880-898), the 90 degree chart maps of the Neel angles
// (:823-834, :844-855); all in double, with the table-driven atan2 and acos.  The rotated charts are those of the
// body and of the moment turned by 90 degrees about y: Euler angles (phi', theta', psi') of the rotated chart give
// the rotation matrix Ry(90) A(phi, theta, psi), Neel angles of the rotated chart the moment Ry(90) m.
__device__ __forceinline__ void euler_to_rotated(double S0, double C0, double S1, double C1, double S2, double C2, double* r) {
    r[0] = atan2_fast(-C2 * C0 + C1 * S0 * S2, S0 * C2 + C1 * C0 * S2);
    r[1] = acos_fast(-S1 * S2);
    r[2] = atan2_fast(C1, C2 * S1);
}
__device__ __forceinline__ void euler_to_rotated(const double* y, double* r) {
    double S0, C0, S1, C1, S2, C2;
    sincos_fast(y[0], &S0, &C0); sincos_fast(y[1], &S1, &C1); sincos_fast(y[2], &S2, &C2);
    euler_to_rotated(S0, C0, S1, C1, S2, C2, r);
}
__device__ __forceinline__ void rotated_to_euler(double S0, double C0, double S1, double C1, double S2, double C2, double* r) {
    r[0] = atan2_fast(C0 * C2 - C1 * S0 * S2, -C2 * S0 - C1 * C0 * S2);
    r[1] = acos_fast(S1 * S2);
    r[2] = atan2_fast(-C1, C2 * S1);
}
__device__ __forceinline__ void rotated_to_euler(const double* y, double* r) {
    double S0, C0, S1, C1, S2, C2;
    sincos_fast(y[0], &S0, &C0); sincos_fast(y[1], &S1, &C1); sincos_fast(y[2], &S2, &C2);
    rotated_to_euler(S0, C0, S1, C1, S2, C2, r);
}
__device__ __forceinline__ void neel_to_rotated(double st, double ct, double sp, double cp, double* r) {
    r[0] = acos_fast(-cp * st);
    r[1] = atan2_fast(sp * st, ct);
}
__device__ __forceinline__ void neel_to_rotated(const double* y, double* r) {
    double st, ct, sp, cp;
    sincos_fast(y[0], &st, &ct); sincos_fast(y[1], &sp, &cp);
    neel_to_rotated(st, ct, sp, cp, r);
}
__device__ __forceinline__ void rotated_to_neel(double st, double ct, double sp, double cp, double* r) {
    r[0] = acos_fast(cp * st);
    r[1] = atan2_fast(sp * st, -ct);
}
__device__ __forceinline__ void rotated_to_neel(const double* y, double* r) {
    double st, ct, sp, cp;
    sincos_fast(y[0], &st, &ct); sincos_fast(y[1], &sp, &cp);
    rotated_to_neel(st, ct, sp, cp, r);
}
template <int ANISO, int SOLVER, class R> struct BNESCfg {
    static constexpr int DY = 5, DW = 6, DOUT = 9, NORM = 0;
    static constexpr int NZ = noise_per_step<SOLVER, DW>(), G = 12 / NZ;
    using Lane = BNESLane<ANISO, R>;
    // the trig of the state in its charts; brot, nrot: the Euler angles, the Neel angles are in their rotated charts.
    // A block is kept in its rotated chart from step to step while the reference would rotate it at every step
    // (prepareCalculations maps it in, finishCalculations maps it back, :142-169, :635-665); those two maps cancel
    // up to rounding, so they are applied only where a chart changes and at the end of every output window.
    struct Cache { Trig5<R> T; bool brot, nrot; };
    __device__ static void load(const DevRequest& rq, uint64_t p, Lane& L) { bnes_load<ANISO, R>(rq, p, L); }
    __device__ static void init_cache(const Lane&, const double* y, Cache& c) { trig5(y, c.T); c.brot = c.nrot = false; }
    __device__ static void weights(const Lane& L, double w[9]) {
#pragma unroll
        for (int i = 0; i < 9; ++i) w[i] = L.w;                     // getWeighting (:755-759)
    }
    // calculateOutputResult (:674-735): body x axis, body y axis, magnetisation direction, in the lab frame from
    // either chart (a vector v' of the rotated chart is Ry(90) v: v = (-v'_z, v'_y, v'_x))
    __device__ static void output(const Lane&, const double*, const Cache& c, double o[9]) {
        const Trig5<R>& T = c.T;
        const R cphi = T.c[0], ctheta = T.c[1], cpsi = T.c[2], sphi = T.s[0], stheta = T.s[1], spsi = T.s[2];
        const R cphicpsi = cphi * cpsi, sphicpsi = sphi * cpsi, cphispsi = cphi * spsi, sphispsi = sphi * spsi;
        const R ex[3] = {cphicpsi - ctheta * sphispsi, ctheta * sphicpsi + cphispsi, stheta * sphi};
        const R ey[3] = {-sphicpsi - ctheta * cphispsi, ctheta * cphicpsi - sphispsi, stheta * cphi};
        const R m[3] = {T.s[3] * T.c[4], T.s[3] * T.s[4], T.c[3]};
        if (!c.brot) {
            o[0] = (double)ex[0]; o[1] = (double)ex[1]; o[2] = (double)ex[2];
            o[3] = (double)ey[0]; o[4] = (double)ey[1]; o[5] = (double)ey[2];
        } else {
            o[0] = (double)(-ex[2]); o[1] = (double)ex[1]; o[2] = (double)ex[0];
            o[3] = (double)(-ey[2]); o[4] = (double)ey[1]; o[5] = (double)ey[0];
        }
        if (!c.nrot) { o[6] = (double)m[0]; o[7] = (double)m[1]; o[8] = (double)m[2]; }
        else { o[6] = (double)(-m[2]); o[7] = (double)m[1]; o[8] = (double)m[0]; }
    }
    // the charts of this step, as prepareCalculations chooses them (:142-169, windows :770-814): a block is rotated
    // when |theta| of its standard chart leaves its window.  A block in its rotated chart has the standard theta of
    // the map back, acos(sin theta' sin psi') for the body (:880-898) and acos(cos phi' sin theta') for the moment
    // (:844-855), compared here through its cosine.  In FP64 the cache holds exactly the sines and cosines of y
    // the chart maps need; the mixed mode evaluates them again in FP64.
    __device__ __forceinline__ static void to_chart(const DevRequest& rq, double* y, Cache& c) {
        bool bneed, nneed;
        if (!c.brot) bneed = fabs(y[1]) < rq.bmin || fabs(y[1]) > rq.bmax;
        else {
            const double x = (double)c.T.s[1] * (double)c.T.s[2];
            bneed = x > rq.bcmin || x < rq.bcmax;
        }
        if (!c.nrot) nneed = fabs(y[3]) < rq.tmin || fabs(y[3]) > rq.tmax;
        else {
            const double x = (double)c.T.c[4] * (double)c.T.s[3];
            nneed = x > rq.ctmin || x < rq.ctmax;
        }
        if (bneed != c.brot) {
            double Y[3];
            if constexpr (sizeof(R) == 8) {
                if (bneed) euler_to_rotated(c.T.s[0], c.T.c[0], c.T.s[1], c.T.c[1], c.T.s[2], c.T.c[2], Y);
                else rotated_to_euler(c.T.s[0], c.T.c[0], c.T.s[1], c.T.c[1], c.T.s[2], c.T.c[2], Y);
            } else {
                if (bneed) euler_to_rotated(y, Y);
                else rotated_to_euler(y, Y);
            }
            y[0] = Y[0]; y[1] = Y[1]; y[2] = Y[2];
            if (bneed) trig_exact<0, 3>(y, c.T);
            else trig_std<0, 3>(y, c.T);
            c.brot = bneed;
        }
        if (nneed != c.nrot) {
            double Y[2];
            if constexpr (sizeof(R) == 8) {
                if (nneed) neel_to_rotated(c.T.s[3], c.T.c[3], c.T.s[4], c.T.c[4], Y);
                else rotated_to_neel(c.T.s[3], c.T.c[3], c.T.s[4], c.T.c[4], Y);
            } else {
                if (nneed) neel_to_rotated(y + 3, Y);
                else rotated_to_neel(y + 3, Y);
            }
            y[3] = Y[0]; y[4] = Y[1];
            if (nneed) trig_exact<3, 5>(y, c.T);
            else trig_std<3, 5>(y, c.T);
            c.nrot = nneed;
        }
    }
    // end of an output window: a block in its rotated chart is mapped back (finishCalculations, :635-665) and the
    // trig of the standard state is evaluated from it, as a launch that starts at this window does (init_cache), so
    // a run gives the same bits however its steps are split into launches (launches start and end on window ends)
    __device__ __forceinline__ static void window_end(const Lane&, double* y, Cache& c) {
        if (!c.brot && !c.nrot) return;
        if (c.brot) {
            double Y[3];
            if constexpr (sizeof(R) == 8) rotated_to_euler(c.T.s[0], c.T.c[0], c.T.s[1], c.T.c[1], c.T.s[2], c.T.c[2], Y);
            else rotated_to_euler(y, Y);
            y[0] = Y[0]; y[1] = Y[1]; y[2] = Y[2];
        }
        if (c.nrot) {
            double Y[2];
            if constexpr (sizeof(R) == 8) rotated_to_neel(c.T.s[3], c.T.c[3], c.T.s[4], c.T.c[4], Y);
            else rotated_to_neel(y + 3, Y);
            y[3] = Y[0]; y[4] = Y[1];
        }
        trig5(y, c.T);
        c.brot = c.nrot = false;
    }
    __device__ __forceinline__ static void step(const DevRequest& rq, const Lane& L, double* y, Cache& c, const R* z, V3<double> hd) {
        const V3<R> h = {(R)hd.x, (R)hd.y, (R)hd.z};
        const V3<R> zF = {z[0], z[1], z[2]}, zH = {z[3], z[4], z[5]};
        to_chart(rq, y, c);
        const bool brot = c.brot, nrot = c.nrot;
        double Y[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) Y[i] = y[i];
        const Trig5<R>& T0 = c.T;
        BChart<R> c0;
        bnes_chart(T0, brot, nrot, c0);
        R f0[5], s0[5];
        bnes_noise(L, c0, zF, zH, s0);
        double r[5];
        if (SOLVER == S_EM) {                                       // EulerMaruyama.inl:41-72
            bnes_rates(L, c0, T0, brot, nrot, h, true, f0);
#pragma unroll
            for (int i = 0; i < 5; ++i) r[i] = Y[i] + (double)(f0[i] * L.dt + s0[i]);
        } else if (SOLVER == S_HEUN || SOLVER == S_HNC) {           // charts of y_i kept for both stages
            bnes_rates(L, c0, T0, brot, nrot, h, true, f0);
            double sup[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) { f0[i] = f0[i] * L.dt; sup[i] = Y[i] + (double)(f0[i] + s0[i]); }
            Trig5<R> T1;
            if (brot || nrot) trig5_exact(sup, T1);
            else trig5(sup, T1);
            BChart<R> c1;
            bnes_chart(T1, brot, nrot, c1);
            R f1[5];
            bnes_rates(L, c1, T1, brot, nrot, h, true, f1);
            if (SOLVER == S_HEUN) {
#pragma unroll
                for (int i = 0; i < 5; ++i) r[i] = Y[i] + (double)((R)0.5 * (f1[i] * L.dt + f0[i]) + s0[i]);
            } else {                                                // Heun_NotConsistent.inl:34-45: S at both points on dW2
                const V3<R> zF2 = {z[6], z[7], z[8]}, zH2 = {z[9], z[10], z[11]};
                R t1[5], t0[5];
                bnes_noise(L, c1, zF2, zH2, t1);
                bnes_noise(L, c0, zF2, zH2, t0);
#pragma unroll
                for (int i = 0; i < 5; ++i) r[i] = Y[i] + (double)((R)0.5 * (f1[i] * L.dt + f0[i]) + (R)0.5 * (t1[i] + t0[i]));
            }
        } else {                                                    // Millstein: y + det dt + S dW + drift (dW.dW)
            bnes_rates(L, c0, T0, brot, nrot, h, false, f0);
            R dr[4];
            bnes_drift(L, c0, T0, brot, nrot, dr);
            const R dWdW = L.dt * (dot(zF, zF) + dot(zH, zH));
#pragma unroll
            for (int i = 0; i < 4; ++i) r[i] = Y[i] + (double)(f0[i] * L.dt + s0[i] + dr[i] * dWdW);
            r[4] = Y[4] + (double)(f0[4] * L.dt + s0[4]);
        }
        // the new state in the charts of the step (a rotated block's trig in FP64, as the step's stages take it)
#pragma unroll
        for (int i = 0; i < 5; ++i) y[i] = fmod_2pi(r[i]);
        if (brot) trig_exact<0, 3>(y, c.T);
        else trig_std<0, 3>(y, c.T);
        if (nrot) trig_exact<3, 5>(y, c.T);
        else trig_std<3, 5>(y, c.T);
    }
};

}  // namespace sp
