// sp_device.h -- gfx950 device code of the ensemble SDE engine.
//
// One particle per lane.  A wave owns a 64-particle tile of one request and
// keeps the particle state, its parameters and the window accumulators in
// VGPRs for a whole step range; noise is Philox4x32-10 + Box-Muller in
// registers; the field is evaluated once per step per wave (each lane computes
// one of the next 64 steps, v_readlane broadcasts it); at the end of every
// output window the moment-weighted window means are reduced across the wave
// and lane 0 writes DOUT doubles.  Nothing else touches HBM inside the loop.
//
// Physics follows the reference problem classes; each function names the
// file:line it restates.  R is the arithmetic type of drift, diffusion and
// solver (double, or float in the mixed-precision mode); state accumulation,
// ensemble accumulation, chart rotation and Newton stay in double.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "sp_rng.h"
#include "sp_types.h"

namespace sp {

constexpr double TWO_PI = 6.283185307179586476925286766559;
constexpr double INV_TWO_PI = 0.15915494309189533576888376337251;

template <class R> struct V3 { R x, y, z; };
template <class R> __device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> __device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> __device__ __forceinline__ V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> __device__ __forceinline__ R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> __device__ __forceinline__ V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const unsigned long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ unsigned wave_sum_u(unsigned v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// ---------------------------------------------------------------- field
// Field values for 64 consecutive steps, one step per lane; step n (0-based) is
// evaluated at t = dt (n+1) (+ dt/2 for ImplicitMidpoint), the reference's
// time = _timestep*(++counter) (SingleParticleSimulator.h:145) and
// Implicit_Midpoint.inl:59; the field itself is field_eval (sp_field.h).  The
// refill runs once per 64 steps and is kept out of line, so that the code of every
// field type stays out of the simulation loop; it copies the parameters with
// volatile loads so that they are not kept in scalar registers across the loop.
static __device__ __attribute__((noinline)) V3<double> field_refill(const DevRequest* rq, uint64_t n, int midpoint) {
    FieldParams f;
    const volatile double* src = reinterpret_cast<const volatile double*>(&rq->field);
    double* dst = reinterpret_cast<double*>(&f);
    for (int i = 0; i < (int)(sizeof(FieldParams) / 8); ++i) dst[i] = src[i];
    const int lane = threadIdx.x & 63;
    double t = rq->dt * (double)(n + lane + 1);
    if (midpoint) t = t + 0.5 * rq->dt;
    double h[3];
    field_eval(f, t, h);
    return {h[0], h[1], h[2]};
}

struct FieldCache {
    uint64_t base;
    double hx, hy, hz;
    bool active;
    __device__ __forceinline__ void init(const DevRequest& rq) {
        base = ~0ull; hx = hy = hz = 0.0;
        active = __builtin_amdgcn_readfirstlane((int)rq.field.type) != FIELD_ZERO;   // uniform
    }
    // make steps [n, n + count) available (count <= 64, 64 % count == 0, n aligned to count)
    template <bool MIDPOINT>
    __device__ __forceinline__ void ensure(const DevRequest& rq, uint64_t n, uint32_t count) {
        if (!active) return;
        if (base != ~0ull && n - base + count <= 64) return;
        base = n;
        const V3<double> h = field_refill(&rq, n, MIDPOINT ? 1 : 0);
        hx = h.x; hy = h.y; hz = h.z;
    }
    // branch-free: an inactive (zero) field keeps hx = hy = hz = 0 in every lane
    __device__ __forceinline__ V3<double> at(uint64_t n) const {
        const int k = (int)(n - base) & 63;
        return {readlane_d(hx, k), readlane_d(hy, k), readlane_d(hz, k)};
    }
};

// ---------------------------------------------------------------- noise
// 12 standard normals = 3 Philox blocks, starting at block `call`.
// The FP64 Box-Muller tables in LDS: the simulation kernels read a table row per log and per sin/cos, and an
// LDS read returns in a fraction of the latency of a vector-memory load that hits the cache.  Every FP64
// simulation kernel fills its work-group's copy once at its start (bm_lds_fill); the values are those of
// SP_LOG_TAB and SP_SINCOS_TAB, so the normals stay bit-identical to the host's.
// The same LDS block holds the table of sin and cos at the multiples of pi/64 that sincos_fast reads; every
// simulation kernel fills it.
struct BMTables {
    double log_tab[97][2];
    double sc_tab[128][2];
    double trig_tab[128][4];     // (sin, cos)(k pi / 64) as hi and lo parts, SP_TRIG64_TAB
    double atan_tab[33][2];      // atan(k / 32) as hi and lo parts, SP_ATAN32_TAB
};
static __shared__ BMTables sp_bm_lds;

__device__ __forceinline__ void bm_lds_fill() {
    constexpr int NL = 97 * 2, NS = 128 * 2, NT = 128 * 4, NA = 33 * 2;
    double* dst = &sp_bm_lds.log_tab[0][0];          // the four tables are contiguous
    const double* lt = &SP_LOG_TAB[0][0];
    const double* st = &SP_SINCOS_TAB[0][0];
    const double* tt = &SP_TRIG64_TAB[0][0];
    const double* at = &SP_ATAN32_TAB[0][0];
    for (int i = threadIdx.x; i < NL + NS + NT + NA; i += blockDim.x)
        dst[i] = i < NL ? lt[i] : (i < NL + NS ? st[i - NL] : (i < NL + NS + NT ? tt[i - NL - NS] : at[i - NL - NS - NT]));
    __syncthreads();
}

template <class R> struct NormalsT;
template <> struct NormalsT<double> {
    __device__ __forceinline__ static void block(uint64_t seed, uint64_t p, uint64_t call, double z[4]) {
        box_muller4_f64(noise_block(seed, p, call, 0u), z, sp_bm_lds.log_tab, sp_bm_lds.sc_tab);
    }
    __device__ __forceinline__ static void philox(uint64_t seed, uint64_t p, uint64_t call, double z[12]) {
#pragma unroll
        for (int b = 0; b < 3; ++b) block(seed, p, call + b, z + 4 * b);
    }
};
template <> struct NormalsT<float> {
    __device__ __forceinline__ static void block(uint64_t seed, uint64_t p, uint64_t call, float z[4]) {
        box_muller4_f32(noise_block(seed, p, call, 0u), z);
    }
    __device__ __forceinline__ static void philox(uint64_t seed, uint64_t p, uint64_t call, float z[12]) {
#pragma unroll
        for (int b = 0; b < 3; ++b) block(seed, p, call + b, z + 4 * b);
    }
};
// Whether a kernel class computes the next step group's normals beside the current group's physics (PIPE_NOISE,
// the default) or each Philox block just before the step that first uses it, which keeps 12 fewer normals live.
template <class Cfg, class = void> struct pipe_noise { static constexpr bool value = true; };
template <class Cfg> struct pipe_noise<Cfg, std::void_t<decltype(Cfg::PIPE_NOISE)>> {
    static constexpr bool value = Cfg::PIPE_NOISE;
};
template <class R>
__device__ __forceinline__ void normals12(const DevRequest& rq, uint64_t p, uint64_t k0, R z[12]) {
    if (rq.flags & RQ_REPLAY) {
        const uint64_t N = rq.n_particles;
#pragma unroll
        for (int j = 0; j < 12; ++j) {
            const uint64_t k = k0 + j;
            z[j] = (R)(k < rq.replay_count ? rq.replay[k * N + p] : 0.0);
        }
    } else {
        NormalsT<R>::philox(rq.seed, p, k0 >> 2, z);
    }
}

__device__ __forceinline__ double ld(const DevRequest& rq, int k, uint64_t p) { return rq.params[(uint64_t)k * rq.n_particles + p]; }

// ---------------------------------------------------------------- anisotropy
// ANISO is the reference's IAnisotropy value: 1 uniaxial, 2 cubic, 4 uniaxialcubic
template <class R> struct AnisoLane { R k1; V3<R> ex, ey, ez; R ku; V3<R> eu; };
template <int ANISO, class R>
__device__ __forceinline__ void load_aniso(const DevRequest& rq, uint64_t p, AnisoLane<R>& A) {
    A.k1 = (R)ld(rq, P_K1, p);
    A.ez = {(R)ld(rq, P_AX + 6, p), (R)ld(rq, P_AX + 7, p), (R)ld(rq, P_AX + 8, p)};
    if (ANISO != 1) {
        A.ex = {(R)ld(rq, P_AX + 0, p), (R)ld(rq, P_AX + 1, p), (R)ld(rq, P_AX + 2, p)};
        A.ey = {(R)ld(rq, P_AX + 3, p), (R)ld(rq, P_AX + 4, p), (R)ld(rq, P_AX + 5, p)};
    } else {
        A.ex = A.ey = {(R)0, (R)0, (R)0};
    }
    if (ANISO == 4) {
        A.ku = (R)ld(rq, P_KU, p);
        A.eu = {(R)ld(rq, P_U + 0, p), (R)ld(rq, P_U + 1, p), (R)ld(rq, P_U + 2, p)};
    } else {
        A.ku = (R)0;
        A.eu = {(R)0, (R)0, (R)0};
    }
}
// UniaxialAnisotropy.h:141-144: H = (-2K/Ms)(m.z) z
// CubicAnisotropy.h:109-146:   H = K1_MS sum_i c_i (c_j^2 + c_k^2) axis_i, c_i = m.axis_i
// UniaxialCubicAnisotropy.hpp:114: H = uniaxial field about u = R^T z (-2K_u/Ms) + cubic field (-2K_c/Ms)
template <int ANISO, class R>
__device__ __forceinline__ V3<R> aniso_field(const AnisoLane<R>& A, V3<R> m) {
    if (ANISO == 1) return (A.k1 * dot(m, A.ez)) * A.ez;
    const R c1 = dot(m, A.ex), c2 = dot(m, A.ey), c3 = dot(m, A.ez);
    const R c1s = c1 * c1, c2s = c2 * c2, c3s = c3 * c3;
    const V3<R> hc = A.k1 * (((c2s + c3s) * c1) * A.ex + ((c1s + c3s) * c2) * A.ey + ((c1s + c2s) * c3) * A.ez);
    if (ANISO == 2) return hc;
    return (A.ku * dot(m, A.eu)) * A.eu + hc;
}
// the derivative of aniso_field at m along v, H'(m) v: d_i = v.axis_i enters as
// d/dm [c_i (c_j^2 + c_k^2)] . v = d_i (c_j^2 + c_k^2) + 2 c_i (c_j d_j + c_k d_k)
template <int ANISO>
__device__ __forceinline__ V3<double> aniso_dir(const AnisoLane<double>& A, V3<double> m, V3<double> v) {
    if (ANISO == 1) return (A.k1 * dot(v, A.ez)) * A.ez;
    const double c1 = dot(m, A.ex), c2 = dot(m, A.ey), c3 = dot(m, A.ez);
    const double d1 = dot(v, A.ex), d2 = dot(v, A.ey), d3 = dot(v, A.ez);
    const double g1 = d1 * (c2 * c2 + c3 * c3) + 2.0 * c1 * (c2 * d2 + c3 * d3);
    const double g2 = d2 * (c1 * c1 + c3 * c3) + 2.0 * c2 * (c1 * d1 + c3 * d3);
    const double g3 = d3 * (c1 * c1 + c2 * c2) + 2.0 * c3 * (c1 * d1 + c2 * d2);
    const V3<double> hc = A.k1 * (g1 * A.ex + g2 * A.ey + g3 * A.ez);
    if (ANISO == 2) return hc;
    return (A.ku * dot(v, A.eu)) * A.eu + hc;
}

enum { S_EM = 1, S_IM = 3, S_MIL = 6, S_HEUN = 7, S_HNC = 8 };
// normals per step: dim_W, twice that for Heun_NotConsistent (dW and dW2, Heun_NotConsistent.inl:37-38)
template <int SOLVER, int DW> constexpr int noise_per_step() { return SOLVER == S_HNC ? 2 * DW : DW; }

// ================================================================ Neel (Cartesian)
// Problems/NeelRelaxation.h: det = a m x H - b m x (m x H) (:117-128);
// S dW = b sigma (dW - m (m.dW)) + a sigma (m x dW) (:91-109); drift D m (:111-115).
template <int ANISO, class R> struct NeelLane {
    R a, b, d, as, bs;   // a dt, b dt, D dt, a sigma sqrt(dt), b sigma sqrt(dt)
    AnisoLane<R> A;
    double w;
};
template <int ANISO, class R>
__device__ __forceinline__ void neel_load(const DevRequest& rq, uint64_t p, NeelLane<ANISO, R>& L) {
    const double a = ld(rq, P_A, p), b = ld(rq, P_B, p), s = ld(rq, P_SIG, p), D = ld(rq, P_D, p);
    L.a = (R)(a * rq.dt); L.b = (R)(b * rq.dt); L.d = (R)(D * rq.dt);
    L.as = (R)(a * s * rq.sqdt); L.bs = (R)(b * s * rq.sqdt);
    load_aniso<ANISO, R>(rq, p, L.A);
    L.w = ld(rq, P_MU, p);
}
// (det + drift) dt
template <int ANISO, class R>
__device__ __forceinline__ V3<R> neel_f(const NeelLane<ANISO, R>& L, V3<R> m, V3<R> h) {
    const V3<R> H = aniso_field<ANISO>(L.A, m) + h;
    const V3<R> mxH = cross(m, H);
    const V3<R> mxmxH = cross(m, mxH);
    return (L.a * mxH - L.b * mxmxH) + L.d * m;
}
template <int ANISO, class R>
__device__ __forceinline__ V3<R> neel_det(const NeelLane<ANISO, R>& L, V3<R> m, V3<R> h) {
    const V3<R> H = aniso_field<ANISO>(L.A, m) + h;
    const V3<R> mxH = cross(m, H);
    return L.a * mxH - L.b * cross(m, mxH);
}
template <int ANISO, class R>
__device__ __forceinline__ V3<R> neel_s(const NeelLane<ANISO, R>& L, V3<R> m, V3<R> z) {
    return L.bs * (z - dot(m, z) * m) + L.as * cross(m, z);
}

template <int ANISO, int SOLVER, class R> struct NeelCfg {
    static constexpr int DY = 3, DW = 3, DOUT = 3, NORM = 1;   // NORM: 3-blocks that normalize() rescales
    // FP64: the compiler's register count (208-235, two waves per SIMD); FP32 physics: four waves (128 registers)
    static constexpr int WAVES_PER_SIMD = sizeof(R) == 8 ? 1 : 4;
    static constexpr int NZ = noise_per_step<SOLVER, DW>(), G = 12 / NZ;
    using Lane = NeelLane<ANISO, R>;
    struct Cache {};
    __device__ static void load(const DevRequest& rq, uint64_t p, Lane& L) { neel_load<ANISO, R>(rq, p, L); }
    __device__ static void init_cache(const Lane&, const double*, Cache&) {}
    __device__ static void weights(const Lane& L, double w[3]) { w[0] = w[1] = w[2] = L.w; }
    __device__ static void output(const Lane&, const double* y, const Cache&, double o[3]) { o[0] = y[0]; o[1] = y[1]; o[2] = y[2]; }
    __device__ __forceinline__ static void step(const DevRequest&, const Lane& L, double* y, Cache&, const R* z, V3<double> hd) {
        const V3<R> m = {(R)y[0], (R)y[1], (R)y[2]};
        const V3<R> h = {(R)hd.x, (R)hd.y, (R)hd.z};
        const V3<R> zz = {z[0], z[1], z[2]};
        V3<R> inc;
        if (SOLVER == S_EM) {                 // EulerMaruyama.inl:41-72
            inc = neel_f<ANISO>(L, m, h) + neel_s<ANISO>(L, m, zz);
        } else if (SOLVER == S_HEUN) {        // Heun_Strong.inl:36-47
            const V3<R> f0 = neel_f<ANISO>(L, m, h);
            const V3<R> s0 = neel_s<ANISO>(L, m, zz);
            const V3<R> sup = m + f0 + s0;
            const V3<R> f1 = neel_f<ANISO>(L, sup, h);
            inc = (R)0.5 * (f1 + f0) + s0;
        } else if (SOLVER == S_HNC) {         // Heun_NotConsistent.inl:34-45: second stage with S at both points on dW2
            const V3<R> z2 = {z[3], z[4], z[5]};
            const V3<R> f0 = neel_f<ANISO>(L, m, h);
            const V3<R> sup = m + f0 + neel_s<ANISO>(L, m, zz);
            const V3<R> f1 = neel_f<ANISO>(L, sup, h);
            inc = (R)0.5 * (f1 + f0) + (R)0.5 * (neel_s<ANISO>(L, sup, z2) + neel_s<ANISO>(L, m, z2));
        } else {                              // Millstein.inl:34-44: y + a dt + S dW + drift (dW.dW)
            const R zz2 = dot(zz, zz);
            inc = neel_det<ANISO>(L, m, h) + neel_s<ANISO>(L, m, zz) + (L.d * zz2) * m;
        }
        y[0] += (double)inc.x; y[1] += (double)inc.y; y[2] += (double)inc.z;
    }
};

// ================================================================ BrownAndNeel
// Problems/BrownAndNeelRelaxation.inl: det (:222-271), S full/simple (:39-170),
// Stratonovich->Ito drift full/simple (:179-219).  y = (n, m), dW = (dW_F, dW_H).
template <int ANISO, class R> struct BNLane {
    R pf, pt;                 // -2K/Ms, -2K V
    R a, bpd, c, d;           // a dt, (b+d) dt, c dt, d dt
    R cFs, dsig, bpds, as;    // cF sqrt(dt), d sigma sqrt(dt), (b+d|b) sigma sqrt(dt), a sigma sqrt(dt)
    R cFm;                    // cF sqrt(dt) in the full model, 0 in the simple one
    R kn1, kn2, kn3, km;      // drift coefficients times dt
    double w;
};
template <int ANISO, class R>
__device__ __forceinline__ void bn_load(const DevRequest& rq, uint64_t p, BNLane<ANISO, R>& L) {
    const double a = ld(rq, P_A, p), b = ld(rq, P_B, p), s = ld(rq, P_SIG, p);
    const double c = ld(rq, P_C, p), cF = ld(rq, P_CF, p), d = ld(rq, P_DD, p);
    const bool simple = (rq.flags & RQ_BN_SIMPLE) != 0;
    const double dt = rq.dt, sq = rq.sqdt, s2 = s * s, cF2 = cF * cF, bpd = b + d;
    L.pf = (R)ld(rq, P_K1, p); L.pt = (R)ld(rq, P_PT, p);
    L.a = (R)(a * dt); L.bpd = (R)(bpd * dt); L.c = (R)(c * dt); L.d = (R)(d * dt);
    // the simple model drops the coupling terms (.inl:114-170, :179-198); its coefficients are zero here so
    // that the step has no branch: the full model's extra terms then add exact zeros
    L.cFs = (R)(cF * sq); L.dsig = (R)(simple ? 0.0 : d * s * sq); L.as = (R)(a * s * sq);
    L.cFm = (R)(simple ? 0.0 : cF * sq);
    L.bpds = (R)((simple ? b : bpd) * s * sq);
    L.kn1 = (R)(-cF2 * dt);
    L.kn2 = (R)(simple ? 0.0 : s2 * d * a * dt);
    L.kn3 = (R)(simple ? 0.0 : -0.5 * d * d * s2 * dt);
    L.km = (R)(simple ? -((a * a + b * b) * s2) * dt : -(cF2 + (a * a + bpd * bpd) * s2) * dt);
    L.w = ld(rq, P_MU, p);
}
// (det + drift) dt for both blocks
template <int ANISO, class R>
__device__ __forceinline__ void bn_f(const BNLane<ANISO, R>& L, V3<R> n, V3<R> m, V3<R> h, V3<R>& fn, V3<R>& fm, bool with_drift) {
    if (ANISO == 1) {
        const R mn = dot(m, n);
        const V3<R> H = (L.pf * mn) * n + h;
        const V3<R> Teff = (L.pt * mn) * cross(n, m);      // n x (pt (m.n) m), UniaxialAnisotropy.h:146-150
        const V3<R> mxH = cross(m, H);
        const V3<R> wB = L.c * Teff + L.d * mxH;
        const V3<R> wN = (L.bpd * mxH + L.c * Teff) - L.a * H;
        fn = cross(wB, n);
        fm = cross(wN, m);
    } else {                                               // cubic: zero deterministic vector (.inl:225-229)
        fn = {(R)0, (R)0, (R)0};
        fm = {(R)0, (R)0, (R)0};
    }
    if (with_drift) {
        fn = fn + L.kn1 * n + L.kn2 * cross(n, m) + (L.kn3 * (dot(m, m) - (R)1)) * n;
        fm = fm + L.km * m;
    }
}
template <int ANISO, class R>
__device__ __forceinline__ void bn_s(const BNLane<ANISO, R>& L, V3<R> n, V3<R> m, V3<R> zF, V3<R> zH, V3<R>& sn, V3<R>& sm) {
    sn = L.cFs * cross(zF, n) + L.dsig * (dot(n, m) * zH - dot(n, zH) * m);
    sm = L.bpds * (zH - dot(m, zH) * m) + L.as * cross(m, zH) + L.cFm * cross(zF, m);
}

template <int ANISO, int SOLVER, class R> struct BNCfg {
    static constexpr int DY = 6, DW = 6, DOUT = 6, NORM = 2;
    // FP64 at two waves per SIMD (with a one-wave build for small launches), FP32 physics at three (168 registers)
    static constexpr int WAVES_PER_SIMD = sizeof(R) == 8 ? 2 : 3;
    static constexpr int NZ = noise_per_step<SOLVER, DW>(), G = 12 / NZ;
    using Lane = BNLane<ANISO, R>;
    struct Cache {};
    __device__ static void load(const DevRequest& rq, uint64_t p, Lane& L) { bn_load<ANISO, R>(rq, p, L); }
    __device__ static void init_cache(const Lane&, const double*, Cache&) {}
    __device__ static void weights(const Lane& L, double w[6]) { w[0] = w[1] = w[2] = 1.0; w[3] = w[4] = w[5] = L.w; }
    __device__ static void output(const Lane&, const double* y, const Cache&, double o[6]) {
        for (int i = 0; i < 6; ++i) o[i] = y[i];
    }
    __device__ __forceinline__ static void step(const DevRequest&, const Lane& L, double* y, Cache&, const R* z, V3<double> hd) {
        const V3<R> n = {(R)y[0], (R)y[1], (R)y[2]}, m = {(R)y[3], (R)y[4], (R)y[5]};
        const V3<R> h = {(R)hd.x, (R)hd.y, (R)hd.z};
        const V3<R> zF = {z[0], z[1], z[2]}, zH = {z[3], z[4], z[5]};
        V3<R> in, im;
        if (SOLVER == S_EM) {
            V3<R> fn, fm, sn, sm;
            bn_f<ANISO>(L, n, m, h, fn, fm, true);
            bn_s<ANISO>(L, n, m, zF, zH, sn, sm);
            in = fn + sn; im = fm + sm;
        } else if (SOLVER == S_HEUN) {
            V3<R> fn0, fm0, sn, sm, fn1, fm1;
            bn_f<ANISO>(L, n, m, h, fn0, fm0, true);
            bn_s<ANISO>(L, n, m, zF, zH, sn, sm);
            bn_f<ANISO>(L, n + fn0 + sn, m + fm0 + sm, h, fn1, fm1, true);
            in = (R)0.5 * (fn1 + fn0) + sn;
            im = (R)0.5 * (fm1 + fm0) + sm;
        } else if (SOLVER == S_HNC) {         // Heun_NotConsistent.inl:34-45
            const V3<R> zF2 = {z[6], z[7], z[8]}, zH2 = {z[9], z[10], z[11]};
            V3<R> fn0, fm0, sn, sm, fn1, fm1, sn1, sm1, sn0, sm0;
            bn_f<ANISO>(L, n, m, h, fn0, fm0, true);
            bn_s<ANISO>(L, n, m, zF, zH, sn, sm);
            const V3<R> ns = n + fn0 + sn, ms = m + fm0 + sm;
            bn_f<ANISO>(L, ns, ms, h, fn1, fm1, true);
            bn_s<ANISO>(L, ns, ms, zF2, zH2, sn1, sm1);
            bn_s<ANISO>(L, n, m, zF2, zH2, sn0, sm0);
            in = (R)0.5 * (fn1 + fn0) + (R)0.5 * (sn1 + sn0);
            im = (R)0.5 * (fm1 + fm0) + (R)0.5 * (sm1 + sm0);
        } else {   // Millstein: y + det dt + S dW + drift (dW.dW)
            V3<R> fn, fm, sn, sm, dn, dm;
            bn_f<ANISO>(L, n, m, h, fn, fm, false);
            bn_s<ANISO>(L, n, m, zF, zH, sn, sm);
            // drift (dW.dW) = drift dt (z.z) and the drift coefficients already carry dt
            const R zz = dot(zF, zF) + dot(zH, zH);
            dn = L.kn1 * n + L.kn2 * cross(n, m) + (L.kn3 * (dot(m, m) - (R)1)) * n;
            dm = L.km * m;
            in = fn + sn + zz * dn;
            im = fm + sm + zz * dm;
        }
        y[0] += (double)in.x; y[1] += (double)in.y; y[2] += (double)in.z;
        y[3] += (double)im.x; y[4] += (double)im.y; y[5] += (double)im.z;
    }
};

// ================================================================ NeelSpherical
// Problems/NeelRelaxationSpherical.h: chart rotation (:124-145, :546-602), chart
// vectors and projection P (:162-236), det = P H (:239-265), S = sigma P
// (:267-293), drift (-D/2 cos(theta)/sin(theta), 0) (:295-309), finish (:440-471),
// output (sin t cos p, sin t sin p, cos t) (:499-516).
// sin and cos of a double, |x| < 16384, within about half an ulp: x = k pi/64 + r with pi/64 in three parts
// (k * part exact for |k| < 2^20) and |r| <= pi/128; sin r by its series to r^7 and cos r - 1 to r^6 (the next
// terms are below 4e-19 relative and 4e-18 absolute), combined with (sin, cos)(k pi/64) from the LDS table --
// each value as a double and the double nearest to its remainder -- by the angle-sum formulas.  The table covers
// the whole circle, so k mod 128 picks the row and no quadrant logic is needed.  Larger arguments (never reached
// by angles that stay near [0, 2 pi]) take the library routine.
struct SinCos { double s, c; };
// The library routines for the arguments the inline paths leave out, kept out of line: inlined, their long
// reductions would be copied into every call site of the simulation loops and crowd the instruction cache.
static __device__ __attribute__((noinline)) SinCos sincos_library(double x) {
    SinCos r;
    sincos(x, &r.s, &r.c);
    return r;
}
static __device__ __attribute__((noinline)) double fmod_library(double x, double y) { return fmod(x, y); }

__device__ __forceinline__ void sincos_fast(double x, double* s_out, double* c_out) {
    if (!(fabs(x) < 16384.0)) {
        const SinCos r = sincos_library(x);
        *s_out = r.s;
        *c_out = r.c;
        return;
    }
    const double kd = rint(x * 0x1.45f306dc9c883p+4);         // x 64/pi
    double r = fma(-kd, 0x1.921fb54400000p-5, x);             // pi/64 = P1 + P2 + P3
    r = fma(-kd, 0x1.0b4611a600000p-39, r);
    r = fma(-kd, 0x1.3198a2e037073p-74, r);
    const double* T = sp_bm_lds.trig_tab[(int)kd & 127];      // sin, cos of k pi/64: hi, hi, lo, lo
    const double r2 = r * r;
    const double sr = fma(r * r2, fma(fma(-1.984126984126984e-4, r2, 8.333333333333333e-3), r2, -0.16666666666666666), r);
    const double cm1 = r2 * fma(fma(-1.388888888888889e-3, r2, 4.1666666666666664e-2), r2, -0.5);
    *s_out = T[0] + fma(T[1], sr, fma(T[0], cm1, T[2]));
    *c_out = T[1] + fma(-T[0], sr, fma(T[1], cm1, T[3]));
}

// 1 / x for |x| >= 2^-1020: the hardware reciprocal and two Newton corrections (within an ulp)
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
static __device__ __attribute__((noinline)) double inv_tiny(double x) {
    const double r = 1.0 / x;
    return isinf(r) ? 0.0 : r;
}
static __device__ __attribute__((noinline)) double atan2_library(double y, double x) { return atan2(y, x); }

// a + b = s + e exactly (Knuth's two-sum)
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// atan2(y, x) within about 1.8 ulp (the C library's on 96% of arguments): t = min(|x|, |y|) / max(|x|, |y|) with its
// rounding error tl, t = c + (t - c) for the nearest c = k/32, atan t = atan c + atan u with u = (t - c)/(1 + t c),
// |u| <= 1/64, by its series to u^9, plus tl/(1 + t^2); atan c from the LDS table as hi and lo parts; pi/2 - a and
// pi - a for the other octants by exact two-sums.  Zeros, infinities, NaNs and arguments beyond 2^+-500 take the
// library routine.
__device__ __forceinline__ double atan2_fast(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    const bool swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    if (!(den >= 0x1p-500 && den <= 0x1p500)) return atan2_library(y, x);
    const double r = rcp_nr(den);
    double t = num * r;
    t = fma(fma(-den, t, num), r, t);                          // num / den, within half an ulp
    const double tl = fma(-den, t, num) * r;                   // num / den - t
    const double kd = rint(t * 32.0);
    const double c = kd * 0.03125;
    const double q = rcp_nr(fma(t, c, 1.0));
    const double u = (t - c) * q;                              // t - c is exact
    const double u2 = u * u;
    const double p = fma(fma(fma(0.1111111111111111, u2, -0.14285714285714285), u2, 0.2), u2, -0.3333333333333333);
    const double au = fma(u * u2, p, u);
    const double* T = sp_bm_lds.atan_tab[(int)kd];
    double hi = T[0], lo = T[1] + fma(tl, q * q * fma(c, c, 1.0), au);
    if (swap) {
        double s, e;
        two_sum(0x1.921fb54442d18p+0, -hi, s, e);              // pi/2 (hi and lo parts) - a
        hi = s;
        lo = (0x1.1a62633145c07p-54 - lo) + e;
    }
    if (signbit(x)) {
        double s, e;
        two_sum(0x1.921fb54442d18p+1, -hi, s, e);              // pi - a
        hi = s;
        lo = (0x1.1a62633145c07p-53 - lo) + e;
    }
    return copysign(hi + lo, y);
}

// acos(c) for |c| <= 1 as atan2(sqrt((1 - c)(1 + c)), c)
__device__ __forceinline__ double acos_fast(double c) { return atan2_fast(sqrt((1.0 - c) * (1.0 + c)), c); }

// 1 / sin(theta) with an infinite value replaced by 0 (NeelRelaxationSpherical.h:156-160)
__device__ __forceinline__ double inv_or_zero(double x) { return fabs(x) >= 0x1p-1020 ? rcp_nr(x) : inv_tiny(x); }
__device__ __forceinline__ float inv_or_zero(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return isinf(r) ? 0.0f : r;
}

// std::fmod(x, 2 pi): x itself when |x| < 2 pi; x -+ 2 pi when 2 pi <= |x| < 4 pi (exact by Sterbenz's lemma,
// and exactly what fmod returns there); the library routine otherwise
__device__ __forceinline__ double fmod_2pi(double x) {
    const double a = fabs(x);
    if (a < TWO_PI) return x;
    if (a < 2.0 * TWO_PI) return x > 0.0 ? x - TWO_PI : x + TWO_PI;
    return fmod_library(x, TWO_PI);
}

template <class R> struct Trig { R st, ct, sp, cp; };
// A chart map of the spherical coordinates (NeelRelaxationSpherical.h:568-602): the angles
// (acos(c), atan2(ys, xs)) of a point on the unit sphere (c^2 + xs^2 + ys^2 = 1), and their sines and cosines
// without evaluating them: acos(c) = atan2(sqrt((1 - c)(1 + c)), c), whose sine is that root and whose cosine is c,
// and (xs, ys) / hypot(xs, ys) are the cosine and sine of atan2(ys, xs); where hypot(xs, ys) = 0, atan2(+-0, x) is
// +-0 for x = +0 and +-pi for x = -0.  FAST takes the table-driven atan2 (fewer registers: the FP32-physics kernels
// gain occupancy from it), otherwise the library's acos and atan2 (as many instructions, and faster in the FP64
// kernels, whose occupancy is fixed by their own register count)
template <bool FAST>
__device__ __forceinline__ void chart_map(double c, double ys, double xs, double Y[2], Trig<double>& T) {
    const double s = sqrt((1.0 - c) * (1.0 + c));
    if constexpr (FAST) {
        Y[0] = atan2_fast(s, c);
        Y[1] = atan2_fast(ys, xs);
    } else {
        Y[0] = acos(c);
        Y[1] = atan2(ys, xs);
    }
    T.st = s;
    T.ct = c;
    const double h = sqrt(fma(xs, xs, ys * ys));
    if (h > 0.0) {
        const double ih = rcp_nr(h);
        T.cp = xs * ih;
        T.sp = ys * ih;
    } else {
        T.cp = signbit(xs) ? -1.0 : 1.0;
        T.sp = ys;
    }
}
__device__ __forceinline__ void trig_of(double th, double ph, Trig<double>& T) {
    sincos_fast(th, &T.st, &T.ct);
    sincos_fast(ph, &T.sp, &T.cp);
}
__device__ __forceinline__ void trig_of(double th, double ph, Trig<float>& T) {
    // hardware sin/cos take revolutions; the reduction to one turn is done in double
    double rt = th * INV_TWO_PI, rp = ph * INV_TWO_PI;
    rt -= floor(rt); rp -= floor(rp);
    const float ft = (float)rt, fp = (float)rp;
    T.st = __builtin_amdgcn_sinf(ft); T.ct = __builtin_amdgcn_cosf(ft);
    T.sp = __builtin_amdgcn_sinf(fp); T.cp = __builtin_amdgcn_cosf(fp);
}

template <int ANISO, class R> struct SphLane {
    R a, b, dr, sg, dt;       // a, b, -D/2, sigma sqrt(dt), dt
    AnisoLane<R> A;
    double w;
};
template <class R> struct Chart { V3<R> P0, P1; R drift0; };

// chart quantities from the (possibly rotated) coordinates' trig values
template <int ANISO, class R>
__device__ __forceinline__ void make_chart(const SphLane<ANISO, R>& L, const Trig<R>& T, bool rot, V3<R> h, Chart<R>& c, V3<R>& det) {
    const R ods = inv_or_zero(T.st);
    V3<R> M, et, ep;
    if (!rot) {
        M = {T.st * T.cp, T.st * T.sp, T.ct};
        et = {T.ct * T.cp, T.ct * T.sp, -T.st};
        ep = {-T.sp, T.cp, (R)0};
    } else {
        M = {-T.ct, T.st * T.sp, T.st * T.cp};
        et = {T.st, T.ct * T.sp, T.ct * T.cp};
        ep = {(R)0, T.cp, -T.sp};
    }
    c.P0 = L.b * et - L.a * ep;
    c.P1 = ods * (L.a * et + L.b * ep);
    if (ods == (R)0) {
        if (!rot) c.P1.z = L.a; else c.P1.x = -L.a;
    }
    c.drift0 = L.dr * T.ct * ods;     // cos_t = rot ? -M(0) : M(2) = ct in both charts
    const V3<R> H = aniso_field<ANISO>(L.A, M) + h;
    det = {dot(c.P0, H), dot(c.P1, H), (R)0};
}

template <int ANISO, int SOLVER, class R> struct SphCfg {
    static constexpr int DY = 2, DW = 3, DOUT = 3, NORM = 0;
    static constexpr bool PIPE_NOISE = false;
    // uniaxial and cubic: FP64 at three waves per SIMD (their translation unit is built so that it fits 168
    // registers without spilling), FP32 physics at four; uniaxialcubic at two
    static constexpr int WAVES_PER_SIMD = ANISO == 4 ? 2 : (sizeof(R) == 8 ? 3 : 4);
    static constexpr int NZ = noise_per_step<SOLVER, DW>(), G = 12 / NZ;
    using Lane = SphLane<ANISO, R>;
    // the trig of the state in its chart; rot: the state is in the rotated chart.  A state is kept in the rotated
    // chart from step to step while the reference would rotate it at every step (prepareCalculations maps it in,
    // finishCalculations maps it back, NeelRelaxationSpherical.h:440-471, 568-602); those two maps cancel up to
    // rounding, so they are applied only where the chart changes and at the end of every output window.
    struct Cache { Trig<R> T; bool rot; };
    __device__ static void load(const DevRequest& rq, uint64_t p, Lane& L) {
        L.a = (R)ld(rq, P_A, p); L.b = (R)ld(rq, P_B, p);
        L.dr = (R)(-0.5 * ld(rq, P_D, p));
        L.sg = (R)(ld(rq, P_SIG, p) * rq.sqdt);
        L.dt = (R)rq.dt;
        load_aniso<ANISO, R>(rq, p, L.A);
        L.w = ld(rq, P_MU, p);
    }
    __device__ static void init_cache(const Lane&, const double* y, Cache& c) { trig_of(y[0], y[1], c.T); c.rot = false; }
    __device__ static void weights(const Lane& L, double w[3]) { w[0] = w[1] = w[2] = L.w; }
    // m in the lab frame from either chart (NeelRelaxationSpherical.h:499-516 after the map back)
    __device__ static void output(const Lane&, const double*, const Cache& c, double o[3]) {
        const Trig<R>& T = c.T;
        if (!c.rot) { o[0] = (double)(T.st * T.cp); o[1] = (double)(T.st * T.sp); o[2] = (double)T.ct; }
        else { o[0] = (double)(-T.ct); o[1] = (double)(T.st * T.sp); o[2] = (double)(T.st * T.cp); }
    }
    // the FP64 sines and cosines of the state (the cached ones in FP64, evaluated again from y in the mixed mode)
    __device__ __forceinline__ static void trig64(const double* y, const Cache& c, double& st, double& ct, double& sp, double& cp) {
        if constexpr (sizeof(R) == 8) { st = c.T.st; ct = c.T.ct; sp = c.T.sp; cp = c.T.cp; }
        else { sincos_fast(y[0], &st, &ct); sincos_fast(y[1], &sp, &cp); }
    }
    // the chart of this step, as prepareCalculations chooses it (:568-580): rotated (90 degrees about y) when the
    // polar angle of the state leaves [tmin, tmax].  A state in the rotated chart has the polar angle
    // acos(cos phi' sin theta') of the map back (:440-471), compared here through its cosine.
    __device__ __forceinline__ static void to_chart(const DevRequest& rq, double* y, Cache& c) {
        bool need;
        if (!c.rot) {
            need = (y[0] < rq.tmin) || (y[0] > rq.tmax);
        } else {
            const double cth = (double)c.T.cp * (double)c.T.st;
            need = cth > rq.ctmin || cth < rq.ctmax;
        }
        if (need == c.rot) return;
        double st, ct, sp, cp;
        trig64(y, c, st, ct, sp, cp);
        Trig<double> Td;
        if (need) chart_map<sizeof(R) == 4>(-cp * st, sp * st, ct, y, Td);    // (acos(-cp st), atan2(sp st, ct))
        else chart_map<sizeof(R) == 4>(cp * st, sp * st, -ct, y, Td);         // (acos(cp st), atan2(sp st, -ct))
        c.T = {(R)Td.st, (R)Td.ct, (R)Td.sp, (R)Td.cp};
        c.rot = need;
    }
    // end of an output window: a state in the rotated chart is mapped back (finishCalculations, :440-471) and the
    // trig of the standard state is evaluated from it, as a launch that starts at this window does (init_cache), so
    // a run gives the same bits however its steps are split into launches (launches start and end on window ends)
    __device__ __forceinline__ static void window_end(const Lane&, double* y, Cache& c) {
        if (!c.rot) return;
        double st, ct, sp, cp;
        trig64(y, c, st, ct, sp, cp);
        Trig<double> Td;
        chart_map<sizeof(R) == 4>(cp * st, sp * st, -ct, y, Td);
        trig_of(y[0], y[1], c.T);
        c.rot = false;
    }
    __device__ __forceinline__ static void step(const DevRequest& rq, const Lane& L, double* y, Cache& c, const R* z, V3<double> hd) {
        const V3<R> h = {(R)hd.x, (R)hd.y, (R)hd.z};
        const V3<R> zz = {z[0], z[1], z[2]};
        to_chart(rq, y, c);
        const bool rot = c.rot;
        const double Y[2] = {y[0], y[1]};
        Chart<R> c0; V3<R> det0;
        make_chart<ANISO>(L, c.T, rot, h, c0, det0);
        const R s0 = L.sg * dot(c0.P0, zz), s1 = L.sg * dot(c0.P1, zz);
        double r[2];
        if (SOLVER == S_EM) {                         // EulerMaruyama.inl:41-72
            r[0] = Y[0] + (double)((det0.x + c0.drift0) * L.dt + s0);
            r[1] = Y[1] + (double)(det0.y * L.dt + s1);
        } else if (SOLVER == S_HEUN || SOLVER == S_HNC) {   // chart of y_i kept for both stages
            const R f00 = (det0.x + c0.drift0) * L.dt, f01 = det0.y * L.dt;
            const double sup0 = Y[0] + (double)(f00 + s0), sup1 = Y[1] + (double)(f01 + s1);
            Trig<R> T1;
            if (rot) { Trig<double> Td; trig_of(sup0, sup1, Td); T1 = {(R)Td.st, (R)Td.ct, (R)Td.sp, (R)Td.cp}; }
            else trig_of(sup0, sup1, T1);
            Chart<R> c1; V3<R> det1;
            make_chart<ANISO>(L, T1, rot, h, c1, det1);
            const R f10 = (det1.x + c1.drift0) * L.dt, f11 = det1.y * L.dt;
            if (SOLVER == S_HEUN) {
                r[0] = Y[0] + (double)((R)0.5 * (f10 + f00) + s0);
                r[1] = Y[1] + (double)((R)0.5 * (f11 + f01) + s1);
            } else {                                  // Heun_NotConsistent.inl:34-45: S at both points on dW2
                const V3<R> z2 = {z[3], z[4], z[5]};
                const R t0 = L.sg * (dot(c1.P0, z2) + dot(c0.P0, z2)), t1 = L.sg * (dot(c1.P1, z2) + dot(c0.P1, z2));
                r[0] = Y[0] + (double)((R)0.5 * (f10 + f00) + (R)0.5 * t0);
                r[1] = Y[1] + (double)((R)0.5 * (f11 + f01) + (R)0.5 * t1);
            }
        } else {                                      // Millstein: y + det dt + S dW + drift (dW.dW)
            const R dWdW = L.dt * dot(zz, zz);
            r[0] = Y[0] + (double)(det0.x * L.dt + s0 + c0.drift0 * dWdW);
            r[1] = Y[1] + (double)(det0.y * L.dt + s1);
        }
        // the new state in the chart of the step (the rotated one in FP64 trig, as the step's stages)
        y[0] = fmod_2pi(r[0]);
        y[1] = fmod_2pi(r[1]);
        if (rot) { Trig<double> Td; trig_of(y[0], y[1], Td); c.T = {(R)Td.st, (R)Td.ct, (R)Td.sp, (R)Td.cp}; }
        else trig_of(y[0], y[1], c.T);
    }
};

// ================================================================ ImplicitMidpoint (FP64 Newton)
// Implicit_Midpoint.inl:38-188: F(y) = y_i + a(ybar) dt + b(ybar) dW - y with
// ybar = (y + y_i)/2, prepare/finish around the stage, field at t_{n+1} + dt/2.
// The Jacobian is exact (forward-mode duals through the whole residual);
// Newton stops when |dx| <= Accuracy_Goal or after Max_Iterations.
template <int N> struct Dual {
    double v, d[N];
    __device__ Dual() {}
    __device__ Dual(double x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
};
template <int N> __device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) { Dual<N> r; r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> __device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r; r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> __device__ __forceinline__ Dual<N> operator*(double s, const Dual<N>& b) { Dual<N> r; r.v = s * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = s * b.d[i]; return r; }
template <int N> __device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v / b.v; const double ib = 1.0 / b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N> __device__ __forceinline__ Dual<N> chain(const Dual<N>& a, double f, double df) { Dual<N> r; r.v = f;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i]; return r; }
template <int N> __device__ __forceinline__ void dsincos(const Dual<N>& a, Dual<N>& s, Dual<N>& c) {
    double sv, cv; sincos_fast(a.v, &sv, &cv); s = chain(a, sv, cv); c = chain(a, cv, -sv);
}
template <int N> __device__ __forceinline__ Dual<N> dacos(const Dual<N>& a) { return chain(a, acos(a.v), -1.0 / sqrt(1.0 - a.v * a.v)); }
template <int N> __device__ __forceinline__ Dual<N> datan2(const Dual<N>& y, const Dual<N>& x) {
    Dual<N> r; r.v = atan2(y.v, x.v); const double den = 1.0 / (x.v * x.v + y.v * y.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * den;
    return r;
}
template <int N> __device__ __forceinline__ Dual<N> dfmod_2pi(const Dual<N>& a) { Dual<N> r = a; r.v = fmod_2pi(a.v); return r; }

template <int N> struct DV3 { Dual<N> x, y, z; };
template <int N> __device__ __forceinline__ DV3<N> operator+(const DV3<N>& a, const DV3<N>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <int N> __device__ __forceinline__ DV3<N> operator-(const DV3<N>& a, const DV3<N>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <int N> __device__ __forceinline__ DV3<N> operator*(const Dual<N>& s, const DV3<N>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <int N> __device__ __forceinline__ DV3<N> operator*(double s, const DV3<N>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <int N> __device__ __forceinline__ Dual<N> ddot(const DV3<N>& a, const DV3<N>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <int N> __device__ __forceinline__ DV3<N> dcross(const DV3<N>& a, const DV3<N>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <int N> __device__ __forceinline__ DV3<N> dconst(V3<double> v) { return {Dual<N>(v.x), Dual<N>(v.y), Dual<N>(v.z)}; }
template <int N, int ANISO> __device__ __forceinline__ DV3<N> daniso(const AnisoLane<double>& A, const DV3<N>& m) {
    const DV3<N> ez = dconst<N>(A.ez);
    if (ANISO == 1) return (A.k1 * ddot(m, ez)) * ez;
    const DV3<N> ex = dconst<N>(A.ex), ey = dconst<N>(A.ey);
    const Dual<N> c1 = ddot(m, ex), c2 = ddot(m, ey), c3 = ddot(m, ez);
    const Dual<N> c1s = c1 * c1, c2s = c2 * c2, c3s = c3 * c3;
    const DV3<N> hc = A.k1 * (((c2s + c3s) * c1) * ex + ((c1s + c3s) * c2) * ey + ((c1s + c2s) * c3) * ez);
    if (ANISO == 2) return hc;
    const DV3<N> eu = dconst<N>(A.eu);
    return (A.ku * ddot(m, eu)) * eu + hc;
}

// linear solve A x = b (DY <= 6), Gaussian elimination with partial pivoting
template <int N> __device__ __forceinline__ bool solve_lin(double A[N][N], double b[N], double x[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int p = k; double best = fabs(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < N; ++i) { const double v = fabs(A[i][k]); if (v > best) { best = v; p = i; } }
        if (best == 0.0) return false;
#pragma unroll
        for (int i = k + 1; i < N; ++i) {      // swap rows k and p without dynamic indexing
            if (i == p) {
#pragma unroll
                for (int j = 0; j < N; ++j) { const double t = A[k][j]; A[k][j] = A[i][j]; A[i][j] = t; }
                const double t = b[k]; b[k] = b[i]; b[i] = t;
            }
        }
        const double inv = 1.0 / A[k][k];
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const double f = A[i][k] * inv;
#pragma unroll
            for (int j = k; j < N; ++j) A[i][j] -= f * A[k][j];
            b[i] -= f * b[k];
        }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double s = b[i];
#pragma unroll
        for (int j = i + 1; j < N; ++j) s -= A[i][j] * x[j];
        x[i] = s / A[i][i];
    }
    return true;
}

// residuals per problem -------------------------------------------------------------
template <int ANISO> struct IMNeel {
    static constexpr int DY = 3, DW = 3, DOUT = 3, G = 4, NORM = 1;
    double a, b, as, bs, dt; AnisoLane<double> A; double w;
    __device__ void load(const DevRequest& rq, uint64_t p) {
        a = ld(rq, P_A, p); b = ld(rq, P_B, p);
        const double s = ld(rq, P_SIG, p);
        as = a * s * rq.sqdt; bs = b * s * rq.sqdt; dt = rq.dt;
        load_aniso<ANISO, double>(rq, p, A); w = ld(rq, P_MU, p);
    }
    __device__ void weights(double wv[3]) const { wv[0] = wv[1] = wv[2] = w; }
    __device__ void output(const double* y, double o[3]) const { o[0] = y[0]; o[1] = y[1]; o[2] = y[2]; }
    __device__ void residual(const double* yv, const double* yi, V3<double> h, const double* z, double F[3], double J[3][3]) const {
        DV3<3> y, ybar;
        Dual<3>* yp = &y.x;
        for (int i = 0; i < 3; ++i) { yp[i] = Dual<3>(yv[i]); yp[i].d[i] = 1.0; }
        const DV3<3> yid = dconst<3>({yi[0], yi[1], yi[2]});
        ybar = 0.5 * (y + yid);
        const DV3<3> H = daniso<3, ANISO>(A, ybar) + dconst<3>(h);
        const DV3<3> mxH = dcross(ybar, H);
        const DV3<3> det = (a * dt) * mxH - (b * dt) * dcross(ybar, mxH);
        const DV3<3> zz = dconst<3>({z[0], z[1], z[2]});
        const DV3<3> sdw = bs * (zz - ddot(ybar, zz) * ybar) + as * dcross(ybar, zz);
        const DV3<3> r = ybar + det + sdw + (0.5 * yid - 1.5 * y);
        const Dual<3>* rp = &r.x;
        for (int i = 0; i < 3; ++i) { F[i] = rp[i].v; for (int j = 0; j < 3; ++j) J[i][j] = rp[i].d[j]; }
    }
};
template <int ANISO> struct IMBN {
    static constexpr int DY = 6, DW = 6, DOUT = 6, G = 2, NORM = 2;
    double pf, pt, a, b, c, d, cF, s, dt, sq; bool simple; double w;
    __device__ void load(const DevRequest& rq, uint64_t p) {
        a = ld(rq, P_A, p); b = ld(rq, P_B, p); s = ld(rq, P_SIG, p);
        c = ld(rq, P_C, p); cF = ld(rq, P_CF, p); d = ld(rq, P_DD, p);
        pf = ld(rq, P_K1, p); pt = ld(rq, P_PT, p); dt = rq.dt; sq = rq.sqdt;
        simple = (rq.flags & RQ_BN_SIMPLE) != 0; w = ld(rq, P_MU, p);
    }
    __device__ void weights(double wv[6]) const { wv[0] = wv[1] = wv[2] = 1.0; wv[3] = wv[4] = wv[5] = w; }
    __device__ void output(const double* y, double o[6]) const { for (int i = 0; i < 6; ++i) o[i] = y[i]; }
    __device__ void residual(const double* yv, const double* yi, V3<double> h, const double* z, double F[6], double J[6][6]) const {
        Dual<6> yd[6], yb[6];
        for (int i = 0; i < 6; ++i) { yd[i] = Dual<6>(yv[i]); yd[i].d[i] = 1.0; yb[i] = 0.5 * (yd[i] + Dual<6>(yi[i])); }
        const DV3<6> n = {yb[0], yb[1], yb[2]}, m = {yb[3], yb[4], yb[5]};
        DV3<6> fn = dconst<6>({0, 0, 0}), fm = dconst<6>({0, 0, 0});
        if (ANISO == 1) {
            const Dual<6> mn = ddot(m, n);
            const DV3<6> H = (pf * mn) * n + dconst<6>(h);
            const DV3<6> Teff = (pt * mn) * dcross(n, m);
            const DV3<6> mxH = dcross(m, H);
            const DV3<6> wB = c * Teff + d * mxH;
            const DV3<6> wN = ((b + d) * mxH + c * Teff) - a * H;
            fn = dt * dcross(wB, n);
            fm = dt * dcross(wN, m);
        }
        const DV3<6> zF = dconst<6>({z[0], z[1], z[2]}), zH = dconst<6>({z[3], z[4], z[5]});
        DV3<6> sn = (cF * sq) * dcross(zF, n);
        DV3<6> sm = ((simple ? b : b + d) * s * sq) * (zH - ddot(m, zH) * m) + (a * s * sq) * dcross(m, zH);
        if (!simple) {
            sn = sn + (d * s * sq) * (ddot(n, m) * zH - ddot(n, zH) * m);
            sm = sm + (cF * sq) * dcross(zF, m);
        }
        const Dual<6>* np = &n.x; const Dual<6>* mp = &m.x;
        const Dual<6>* fnp = &fn.x; const Dual<6>* fmp = &fm.x; const Dual<6>* snp = &sn.x; const Dual<6>* smp = &sm.x;
        for (int i = 0; i < 3; ++i) {
            const Dual<6> r0 = np[i] + fnp[i] + snp[i] + Dual<6>(0.5 * yi[i]) - 1.5 * yd[i];
            const Dual<6> r1 = mp[i] + fmp[i] + smp[i] + Dual<6>(0.5 * yi[3 + i]) - 1.5 * yd[3 + i];
            F[i] = r0.v; F[3 + i] = r1.v;
            for (int j = 0; j < 6; ++j) { J[i][j] = r0.d[j]; J[3 + i][j] = r1.d[j]; }
        }
    }
};
template <int ANISO> struct IMSph {
    static constexpr int DY = 2, DW = 3, DOUT = 3, G = 4, NORM = 0;
    static constexpr int WAVES_PER_SIMD = 2;     // the chain-rule Jacobian fits 256 registers
    double a, b, sg, dt, tmin, tmax; AnisoLane<double> A; double w;
    __device__ void load(const DevRequest& rq, uint64_t p) {
        a = ld(rq, P_A, p); b = ld(rq, P_B, p); sg = ld(rq, P_SIG, p) * rq.sqdt; dt = rq.dt;
        tmin = rq.tmin; tmax = rq.tmax;
        load_aniso<ANISO, double>(rq, p, A); w = ld(rq, P_MU, p);
    }
    __device__ void weights(double wv[3]) const { wv[0] = wv[1] = wv[2] = w; }
    __device__ void output(const double* y, double o[3]) const {
        double st, ct, sp, cp; sincos_fast(y[0], &st, &ct); sincos_fast(y[1], &sp, &cp);
        o[0] = st * cp; o[1] = st * sp; o[2] = ct;
    }
    // The residual and its exact Jacobian by the chain rule written out: in either chart, with M(theta, phi) the
    // moment, e_t = dM/dtheta, e_p = (dM/dphi) / sin(theta) and e_r = -de_p/dphi,
    //   de_t/dtheta = -M, de_t/dphi = cos(theta) e_p, de_p/dtheta = 0,
    // so P0 = b e_t - a e_p and P1 = (a e_t + b e_p) / sin(theta) have
    //   dP0/dtheta = -b M, dP0/dphi = b cos(theta) e_p + a e_r,
    //   dP1/dtheta = -cos(theta)/sin(theta) P1 - a M / sin(theta), dP1/dphi = (a cos(theta) e_p - b e_r) / sin(theta)
    // (zero where 1/sin(theta) is replaced by 0), and q_r = u_r + dt P_r.H(M) + sigma P_r.dW has
    //   dq_r/du_k = delta_rk + dP_r/du_k . (dt H + sigma dW) + dt P_r . (H'(M) dM/du_k);
    // the chart maps (acos, atan2) contribute their 2x2 Jacobians before and after.
    __device__ void residual(const double* yv, const double* yi, V3<double> h, const double* z, double F[2], double J[2][2]) const {
        const double r0 = 0.5 * (yv[0] + yi[0]), r1 = 0.5 * (yv[1] + yi[1]);
        const bool rot = (r0 < tmin) || (r0 > tmax);
        double u0 = r0, u1 = r1;
        double JA[2][2] = {{1.0, 0.0}, {0.0, 1.0}};          // d(chart coordinates) / d(midpoint)
        if (rot) {                                           // prepareCalculations (:568-580)
            double st, ct, sp, cp;
            sincos_fast(r0, &st, &ct); sincos_fast(r1, &sp, &cp);
            const double x = -(cp * st), Y = sp * st;
            u0 = acos(x); u1 = atan2(Y, ct);
            const double k = -1.0 / sqrt(1.0 - x * x), den = 1.0 / (ct * ct + Y * Y);
            JA[0][0] = k * -(cp * ct); JA[0][1] = k * (sp * st);
            JA[1][0] = (ct * (sp * ct) + Y * st) * den; JA[1][1] = (ct * (cp * st)) * den;
        }
        double st, ct, sp, cp;
        sincos_fast(u0, &st, &ct); sincos_fast(u1, &sp, &cp);
        double ods = 1.0 / st;
        const bool odz = isinf(ods);
        if (odz) ods = 0.0;
        V3<double> M, et, ep, er;
        if (!rot) { M = {st * cp, st * sp, ct}; et = {ct * cp, ct * sp, -st}; ep = {-sp, cp, 0.0}; er = {cp, sp, 0.0}; }
        else { M = {-ct, st * sp, st * cp}; et = {st, ct * sp, ct * cp}; ep = {0.0, cp, -sp}; er = {0.0, sp, cp}; }
        const V3<double> P0 = b * et - a * ep;
        V3<double> P1 = ods * (a * et + b * ep);
        if (odz) { if (!rot) P1.z = a; else P1.x = -a; }
        const V3<double> H = aniso_field<ANISO, double>(A, M) + h;
        const V3<double> zz = {z[0], z[1], z[2]};
        double q0 = u0 + dt * dot(P0, H) + sg * dot(P0, zz);
        double q1 = u1 + dt * dot(P1, H) + sg * dot(P1, zz);
        // d q / d u
        const V3<double> W = dt * H + sg * zz;
        const V3<double> Gt = aniso_dir<ANISO>(A, M, et), Gp = st * aniso_dir<ANISO>(A, M, ep);
        const V3<double> dP0t = -b * M, dP0p = (b * ct) * ep + a * er;
        V3<double> dP1t = {0.0, 0.0, 0.0}, dP1p = {0.0, 0.0, 0.0};
        if (!odz) {
            dP1t = (-ct * ods) * P1 - (a * ods) * M;
            dP1p = ods * ((a * ct) * ep - b * er);
        }
        double JQ[2][2] = {{1.0 + dot(dP0t, W) + dt * dot(P0, Gt), dot(dP0p, W) + dt * dot(P0, Gp)},
                           {dot(dP1t, W) + dt * dot(P1, Gt), 1.0 + dot(dP1p, W) + dt * dot(P1, Gp)}};
        if (rot) {                                           // finishCalculations (:440-471)
            double s0, c0, s1, c1;
            sincos_fast(q0, &s0, &c0); sincos_fast(q1, &s1, &c1);
            const double x = c1 * s0, Y = s1 * s0;
            q0 = acos(x); q1 = atan2(Y, -c0);
            const double k = -1.0 / sqrt(1.0 - x * x), den = 1.0 / (c0 * c0 + Y * Y);
            const double JB[2][2] = {{k * (c1 * c0), k * -(s1 * s0)}, {(-c0 * (s1 * c0) - Y * s0) * den, (-c0 * (c1 * s0)) * den}};
            double T[2][2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) T[i][j] = JB[i][0] * JQ[0][j] + JB[i][1] * JQ[1][j];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) JQ[i][j] = T[i][0] * JA[0][j] + T[i][1] * JA[1][j];
        } else {
            q0 = fmod_2pi(q0); q1 = fmod_2pi(q1);
        }
        F[0] = (q0 + 0.5 * yi[0]) - 1.5 * yv[0];
        F[1] = (q1 + 0.5 * yi[1]) - 1.5 * yv[1];
        J[0][0] = 0.5 * JQ[0][0] - 1.5; J[0][1] = 0.5 * JQ[0][1];
        J[1][0] = 0.5 * JQ[1][0]; J[1][1] = 0.5 * JQ[1][1] - 1.5;
    }
};

template <class Prob, class = void> struct im_waves { static constexpr int value = 1; };
template <class Prob> struct im_waves<Prob, std::void_t<decltype(Prob::WAVES_PER_SIMD)>> {
    static constexpr int value = Prob::WAVES_PER_SIMD;
};
template <class Prob> struct IMCfg {
    static constexpr int DY = Prob::DY, DW = Prob::DW, DOUT = Prob::DOUT, G = Prob::G, NORM = Prob::NORM, NZ = DW;
    static constexpr int WAVES_PER_SIMD = im_waves<Prob>::value;
    // each Philox block is drawn just before the step that first uses it when the class runs at two waves per SIMD
    // (12 fewer normals live through the Newton iterations)
    static constexpr bool PIPE_NOISE = WAVES_PER_SIMD == 1;
    using Lane = Prob;
    struct Cache { unsigned fails; };
    __device__ static void load(const DevRequest& rq, uint64_t p, Lane& L) { L.load(rq, p); }
    __device__ static void init_cache(const Lane&, const double*, Cache& c) { c.fails = 0; }
    __device__ static void weights(const Lane& L, double w[DOUT]) { L.weights(w); }
    __device__ static void output(const Lane& L, const double* y, const Cache&, double o[DOUT]) { L.output(y, o); }
    __device__ static void step(const DevRequest& rq, const Lane& L, double* y, Cache& c, const double* z, V3<double> h) {
        double yi[DY], yj[DY];
#pragma unroll
        for (int i = 0; i < DY; ++i) { yi[i] = y[i]; yj[i] = y[i]; }
        bool conv = false;
        for (uint32_t it = 0; it < rq.newton_max_iter && !conv; ++it) {
            double F[DY], J[DY][DY], dx[DY];
            L.residual(yj, yi, h, z, F, J);
#pragma unroll
            for (int i = 0; i < DY; ++i) F[i] = -F[i];
            if (!solve_lin<DY>(J, F, dx)) break;
            double nrm = 0.0;
#pragma unroll
            for (int i = 0; i < DY; ++i) { yj[i] += dx[i]; nrm += dx[i] * dx[i]; }
            conv = sqrt(nrm) <= rq.newton_tol;
        }
        if (!conv) ++c.fails;
#pragma unroll
        for (int i = 0; i < DY; ++i) y[i] = yj[i];
    }
};

// Eigen's normalize() on each 3-block of the state, in FP64: v / |v| when |v| > 0
// (NeelRelaxation.h:204-208, BrownAndNeelRelaxation.h:168-172; a no-op for
// NeelSpherical, NeelRelaxationSpherical.h:534-535).
// 1 / sqrt(q) for a normal q > 0: the hardware reciprocal square root and two Newton corrections (within an ulp)
__device__ __forceinline__ double rsqrt_nr(double q) {
    double r = __builtin_amdgcn_rsq(q);
    const double h = 0.5 * q;
    r = r * fma(-h * r, r, 1.5);
    return r * fma(-h * r, r, 1.5);
}
static __device__ __attribute__((noinline)) double rsqrt_library(double q) { return 1.0 / sqrt(q); }

template <int NORM>
__device__ __forceinline__ void normalize_blocks(double* y) {
#pragma unroll
    for (int b = 0; b < NORM; ++b) {
        double* v = y + 3 * b;
        const double q = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
        if (q > 0.0) {                       // v * (1 / |v|), within an ulp of v / |v|
            const double inv = (q >= 0x1p-1000 && q <= 0x1p1000) ? rsqrt_nr(q) : rsqrt_library(q);
            v[0] = v[0] * inv; v[1] = v[1] * inv; v[2] = v[2] * inv;
        }
    }
}

template <bool B> struct BoolTag { static constexpr bool value = B; };
// whether a kernel class has work to do at the end of every output window (Cfg::window_end)
template <class Cfg, class = void> struct has_window_end : std::false_type {};
template <class Cfg> struct has_window_end<Cfg, std::void_t<decltype(&Cfg::window_end)>> : std::true_type {};

// ================================================================ split noise
// A tile of a small launch (one that leaves most SIMDs idle) runs as two waves of one work-group
// (k_plain_split): a noise wave computes the Philox + Box-Muller normals of the tile's 64 particles into an LDS
// ring, the integrator wave reads them and runs the physics, so the noise work moves to an otherwise idle SIMD.
// The ring holds two buffers of SPLIT_NG step groups (12 normals per lane and group, packed 16 bytes per lane,
// consecutive lanes at consecutive addresses); iteration i of both waves ends at a barrier, the noise wave fills
// buffer i & 1 with batch i and the integrator consumes batch i - 1 from the other buffer, so a buffer is written
// only after the barrier that follows its last read.  Both waves run the same number of iterations, decided from
// the same request fields.
template <class R> struct alignas(16) NoisePack { R v[16 / sizeof(R)]; };
template <class R> constexpr int noise_packs() { return 12 / (int)(16 / sizeof(R)); }
constexpr int SPLIT_NG = 2;       // step groups per ring buffer and noise wave
// ring of a split tile with NW noise waves: two buffers of NW * SPLIT_NG step groups
template <class R, int NW> constexpr int split_ring_packs() { return 2 * NW * SPLIT_NG * noise_packs<R>() * 64; }

// whether a tile takes the fast path of run_tile (Philox noise, windows ending on group boundaries); uniform
template <class Cfg>
__device__ __forceinline__ bool fast_path(const DevRequest& rq) {
    const uint32_t flags = __builtin_amdgcn_readfirstlane(rq.flags);
    const uint32_t OS = __builtin_amdgcn_readfirstlane(rq.oversampling);
    return OS % Cfg::G == 0 && (flags & RQ_REPLAY) == 0;
}

// noise wave k of the NW of a split tile: groups [k SPLIT_NG, (k + 1) SPLIT_NG) of every batch of NW SPLIT_NG
// groups of the steps [tw.step_begin, tw.step_end)
template <class Cfg, class R, int NW>
__device__ __forceinline__ void produce_noise(const DevRequest& rq, const TileWork& tw, NoisePack<R>* ring, int k) {
    constexpr int G = Cfg::G, NZ = Cfg::NZ, PK = 16 / (int)sizeof(R), NP = noise_packs<R>(), NB = NW * SPLIT_NG;
    const int lane = threadIdx.x & 63;
    const uint64_t N = rq.n_particles;
    const uint64_t pr = (uint64_t)tw.tile * 64 + lane;
    const uint64_t p = pr < N ? pr : N - 1;
    const uint64_t seed = rq.seed;
    const uint64_t ng = (tw.step_end - tw.step_begin) / G;
    const uint64_t nb = (ng + NB - 1) / NB;
    for (uint64_t i = 0; i <= nb; ++i) {
        if (i < nb) {
            NoisePack<R>* buf = ring + (i & 1) * (NB * NP * 64);
#pragma unroll
            for (int g = k * SPLIT_NG; g < (k + 1) * SPLIT_NG; ++g) {
                const uint64_t gi = i * NB + g;
                if (gi < ng) {
                    R z[12];
                    NormalsT<R>::philox(seed, p, ((tw.step_begin + gi * G) * NZ) >> 2, z);
#pragma unroll
                    for (int k = 0; k < NP; ++k) {
                        NoisePack<R> pk;
#pragma unroll
                        for (int e = 0; e < PK; ++e) pk.v[e] = z[k * PK + e];
                        buf[(g * NP + k) * 64 + lane] = pk;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ================================================================ tile driver
// Runs steps [tw.step_begin, tw.step_end) of one 64-particle tile
// (SingleParticleSimulator::startSimulation, :129-159, for 64 particles at
// once) and writes one moment-weighted partial sum per window and component
// (MeanSimulationResult::operator+=, :87-119, for the wave's 64 particles).
// NW > 0: the integrator wave of a split tile, whose fast path reads its normals from the ring of its NW noise waves.
template <class Cfg, class R, bool KAHAN, bool MIDPOINT, int NW = 0>
__device__ __forceinline__ void run_tile(const DevRequest& rq, const TileWork& tw, const NoisePack<R>* ring = nullptr) {
    constexpr int DY = Cfg::DY, NZ = Cfg::NZ, DOUT = Cfg::DOUT, G = Cfg::G;   // NZ normals per step, 12 per group
    const int lane = threadIdx.x & 63;
    const uint64_t N = rq.n_particles;
    const uint64_t pr = (uint64_t)tw.tile * 64 + lane;
    const bool active = pr < N;
    const uint64_t p = active ? pr : N - 1;

    typename Cfg::Lane L;
    Cfg::load(rq, p, L);
    double y[DY];
#pragma unroll
    for (int i = 0; i < DY; ++i) y[i] = rq.state[(uint64_t)i * N + p];
    typename Cfg::Cache cache;
    Cfg::init_cache(L, y, cache);
    double w[DOUT];
    Cfg::weights(L, w);
    if (!active) {
#pragma unroll
        for (int c = 0; c < DOUT; ++c) w[c] = 0.0;
    }
    const uint32_t OS = rq.oversampling;
    uint32_t row = 0;
    // single-simulation outputs (SimulationResultManager.h:115-149): particle p is kept when
    // (p + 1) % save_interval == 0, as slot (p + 1) / save_interval - 1; window rows are global
    double* const traj = rq.traj;
    const uint64_t n_saved = rq.n_saved;
    const uint64_t first_window = tw.step_begin == 0 ? 0 : tw.step_begin / OS + 1;
    uint64_t slot = ~0ull;
    if (traj != nullptr && active && (pr + 1) % rq.save_interval == 0) slot = (pr + 1) / rq.save_interval - 1;
    auto flush = [&](const double* v) {
        double sums[DOUT];
#pragma unroll
        for (int c = 0; c < DOUT; ++c) sums[c] = wave_sum(v[c] * w[c]);
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < DOUT; ++c) tw.partial[((uint64_t)row * DOUT + c) * tw.pstride + tw.pslot] = sums[c];
        }
        if (slot != ~0ull) {
#pragma unroll
            for (int c = 0; c < DOUT; ++c) traj[((first_window + row) * DOUT + c) * n_saved + slot] = v[c];
        }
        ++row;
    };
    if (tw.step_begin == 0) {                       // window 0: the start state (:135-136)
        double o[DOUT];
        Cfg::output(L, y, cache, o);
        flush(o);
    }
    FieldCache fc;
    fc.init(rq);
    const uint32_t flags = __builtin_amdgcn_readfirstlane(rq.flags);    // uniform: scalar branches only
    const bool renorm = Cfg::NORM > 0 && (flags & RQ_NORMALIZE) != 0;
    const bool replay = (flags & RQ_REPLAY) != 0;
    double acc[DOUT], comp[DOUT];
#pragma unroll
    for (int c = 0; c < DOUT; ++c) { acc[c] = 0.0; comp[c] = 0.0; }
    uint32_t l = 0;
    auto accumulate = [&](const double* o) {
#pragma unroll
        for (int c = 0; c < DOUT; ++c) {
            if (KAHAN) {                      // Kahan summation (:148-153)
                const double yk = o[c] - comp[c];
                const double t = acc[c] + yk;
                comp[c] = (t - acc[c]) - yk;
                acc[c] = t;
            } else {
                acc[c] += o[c];
            }
        }
    };
    auto close_window = [&]() {
        double mean[DOUT];
#pragma unroll
        for (int c = 0; c < DOUT; ++c) { mean[c] = acc[c] / (double)OS; acc[c] = 0.0; comp[c] = 0.0; }
        flush(mean);
        if constexpr (has_window_end<Cfg>::value) Cfg::window_end(L, y, cache);
    };
    // Fast path (Philox noise, windows ending on group boundaries): each group is straight-line code, the
    // next group's Philox + Box-Muller is issued alongside this group's physics, and the per-step
    // normalisation is decided outside the loop.
    auto fast_unpiped = [&](auto renorm_tag) {
        constexpr bool RN = decltype(renorm_tag)::value;
        const uint64_t seed = rq.seed;
        for (uint64_t n0 = tw.step_begin; n0 < tw.step_end; n0 += G) {
            fc.ensure<MIDPOINT>(rq, n0, G);
            const uint64_t call0 = (n0 * NZ) >> 2;
            R z[12];
            double gsum[DOUT];
#pragma unroll
            for (int s = 0; s < G; ++s) {
                // block b is first used by step (4 b) / NZ of the group
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    if (s == (4 * b) / NZ) NormalsT<R>::block(seed, p, call0 + b, z + 4 * b);
                Cfg::step(rq, L, y, cache, z + s * NZ, fc.at(n0 + s));
                if constexpr (RN) normalize_blocks<Cfg::NORM>(y);
                double o[DOUT];
                Cfg::output(L, y, cache, o);
#pragma unroll
                for (int c = 0; c < DOUT; ++c) gsum[c] = s == 0 ? o[c] : gsum[c] + o[c];
            }
            accumulate(gsum);
            l += G;
            if (l == OS) { close_window(); l = 0; }
        }
    };
    // one step group from its 12 normals; the G outputs of the group are added first and enter the window's
    // compensated sum as one term (the reference compensates every step, SingleParticleSimulator.h:148-153; the
    // two differ by the rounding of G - 1 plain additions)
    auto group = [&](uint64_t n0, const R* zc, auto renorm_tag) {
        constexpr bool RN = decltype(renorm_tag)::value;
        double gsum[DOUT];
#pragma unroll
        for (int s = 0; s < G; ++s) {
            Cfg::step(rq, L, y, cache, zc + s * NZ, fc.at(n0 + s));
            if constexpr (RN) normalize_blocks<Cfg::NORM>(y);
            double o[DOUT];
            Cfg::output(L, y, cache, o);
#pragma unroll
            for (int c = 0; c < DOUT; ++c) gsum[c] = s == 0 ? o[c] : gsum[c] + o[c];
        }
        accumulate(gsum);
        l += G;
        if (l == OS) { close_window(); l = 0; }
    };
    auto fast = [&](auto renorm_tag) {
        if constexpr (!pipe_noise<Cfg>::value) {
            fast_unpiped(renorm_tag);
            return;
        }
        const uint64_t seed = rq.seed;
        R zc[12], zn[12];
        if (tw.step_begin < tw.step_end) NormalsT<R>::philox(seed, p, (tw.step_begin * NZ) >> 2, zc);
        for (uint64_t n0 = tw.step_begin; n0 < tw.step_end; n0 += G) {
            fc.ensure<MIDPOINT>(rq, n0, G);
            NormalsT<R>::philox(seed, p, ((n0 + G) * NZ) >> 2, zn);
            group(n0, zc, renorm_tag);
#pragma unroll
            for (int j = 0; j < 12; ++j) zc[j] = zn[j];
        }
    };
    // the integrator of a split tile: batch i - 1 from the ring, then the barrier of iteration i (produce_noise)
    auto fast_split = [&](auto renorm_tag) {
        constexpr int PK = 16 / (int)sizeof(R), NP = noise_packs<R>(), NB = (NW > 0 ? NW : 1) * SPLIT_NG;
        const uint64_t ng = (tw.step_end - tw.step_begin) / G;
        const uint64_t nb = (ng + NB - 1) / NB;
        for (uint64_t i = 0; i <= nb; ++i) {
            if (i > 0) {
                const NoisePack<R>* buf = ring + ((i - 1) & 1) * (NB * NP * 64);
#pragma unroll
                for (int g = 0; g < NB; ++g) {
                    const uint64_t gi = (i - 1) * NB + g;
                    if (gi < ng) {
                        const uint64_t n0 = tw.step_begin + gi * G;
                        R zc[12];
#pragma unroll
                        for (int k = 0; k < NP; ++k) {
                            const NoisePack<R> pk = buf[(g * NP + k) * 64 + lane];
#pragma unroll
                            for (int e = 0; e < PK; ++e) zc[k * PK + e] = pk.v[e];
                        }
                        fc.ensure<MIDPOINT>(rq, n0, G);
                        group(n0, zc, renorm_tag);
                    }
                }
            }
            __syncthreads();
        }
    };
    if (NW > 0 && fast_path<Cfg>(rq)) {
        if constexpr (NW > 0) {
            if (renorm) fast_split(BoolTag<true>{});
            else fast_split(BoolTag<false>{});
        }
    } else if (OS % G == 0 && !replay) {
        if (renorm) fast(BoolTag<true>{});
        else fast(BoolTag<false>{});
    } else {
        for (uint64_t n0 = tw.step_begin; n0 < tw.step_end; n0 += G) {
            R z[12];
            normals12<R>(rq, p, n0 * NZ, z);
            fc.ensure<MIDPOINT>(rq, n0, G);
#pragma unroll
            for (int s = 0; s < G; ++s) {
                const uint64_t n = n0 + s;
                if (n < tw.step_end) {
                    Cfg::step(rq, L, y, cache, z + s * NZ, fc.at(n));
                    if (renorm) normalize_blocks<Cfg::NORM>(y);
                    double o[DOUT];
                    Cfg::output(L, y, cache, o);
                    accumulate(o);
                    if (++l == OS) { close_window(); l = 0; }
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < DY; ++i)
        if (active) rq.state[(uint64_t)i * N + p] = y[i];
    if constexpr (MIDPOINT) {
        const unsigned f = wave_sum_u(active ? cache.fails : 0u);
        if (lane == 0 && f) atomicAdd(rq.newton_fail, (unsigned long long)f);
    }
}

}  // namespace sp
