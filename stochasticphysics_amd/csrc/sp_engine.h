// sp_engine.h -- internal state shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "sp_host.h"
#include "sp_kernels.h"
#include "sp_provider.h"
#include "sp_rng.h"

namespace sp {

#define SP_HIP(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return set_error(SP_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

using Clock = std::chrono::steady_clock;
inline double ms_since(Clock::time_point t0, Clock::time_point t1) {
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// ------------------------------------------------------------------ validated request
// particles drawn on the device by the provider kernel instead of uploaded from host arrays
struct ProvSpec {
    ProvDesc base;                 // seed, N, problem, anisotropy and selection of the ensemble (no pointers)
    std::vector<ProvType> types;
};

struct HostRequest {
    sp_request rq{};
    uint32_t dy = 0, dw = 0, dout = 0, kind = 0;
    bool mixed = false;
    uint64_t N = 0, n_out = 0, total_steps = 0;
    uint32_t n_tiles = 0;
    std::vector<double> params;   // [NPARAM][N]
    std::vector<double> y0;       // [dy][N]
    std::vector<double> replay;   // [K][N]
    std::shared_ptr<const ProvSpec> prov;   // set: params and y0 are drawn on the device (both vectors empty)
    double wsum[9] = {0}, wsum_check[9] = {0};
    double prop_sum[8] = {0};     // sums of T, radius, Ms, alpha, gamma, K, r_h, viscosity over the particles
    uint32_t n_check_tiles = 0;
    uint64_t n_saved = 0;         // particles with single-simulation outputs (save_interval)
};

inline uint64_t gcd64(uint64_t a, uint64_t b) { while (b) { const uint64_t t = a % b; a = b; b = t; } return a; }
inline uint64_t lcm64(uint64_t a, uint64_t b) { return a / gcd64(a, b) * b; }

// normals drawn per step: dim_W, or dW and dW2 for Heun_NotConsistent (Heun_NotConsistent.inl:37-38)
inline uint32_t normals_per_step(const sp_request& rq, uint32_t dw) {
    return rq.solver == SP_SOLVER_HEUN_NOT_CONSISTENT ? 2 * dw : dw;
}

// pt == nullptr: the particles are drawn on the device (only the request is checked)
inline int validate(const sp_request* rq, const sp_particles* pt, bool provided = false) {
    if (!rq || (!pt && !provided)) return set_error(SP_ERR_INVALID_ARG, "request or particles is NULL");
    uint32_t dy, dw, dout;
    if (request_dims(rq->problem, &dy, &dw, &dout)) return set_error(SP_ERR_UNSUPPORTED, "unknown problem type");
    if (rq->solver != SP_SOLVER_EULER_MARUYAMA && rq->solver != SP_SOLVER_EULER_MARUYAMA_NORMALIZED &&
        rq->solver != SP_SOLVER_HEUN_STRONG && rq->solver != SP_SOLVER_MILLSTEIN && rq->solver != SP_SOLVER_IMPLICIT_MIDPOINT &&
        rq->solver != SP_SOLVER_HEUN_NOT_CONSISTENT)
        return set_error(SP_ERR_UNSUPPORTED, "solver not available on the GPU backend");
    if (rq->field < SP_FIELD_ZERO || rq->field > SP_FIELD_SINC)
        return set_error(SP_ERR_UNSUPPORTED, "field type not available on the GPU backend");
    if ((rq->field == SP_FIELD_TRIANGULAR || rq->field == SP_FIELD_RECTANGULAR || rq->field == SP_FIELD_SINC) &&
        !(rq->field_period > 0.0))
        return set_error(SP_ERR_INVALID_ARG, "the field's Periode must be > 0");
    if (rq->anisotropy != SP_ANISO_UNIAXIAL && rq->anisotropy != SP_ANISO_CUBIC && rq->anisotropy != SP_ANISO_UNIAXIALCUBIC)
        return set_error(SP_ERR_UNSUPPORTED, "anisotropy type not available on the GPU backend");
    if (rq->anisotropy == SP_ANISO_UNIAXIALCUBIC) {
        // BrownAndNeel passes zero x/y axes and zero Euler sines/cosines to the anisotropy
        // (BrownAndNeelRelaxation.inl:236-242); the cubic part's torque then divides 0 by sin(theta) = 0
        // (CubicAnisotropy.h:183-192) and the reference's deterministic vector is NaN for every state
        if (rq->problem == SP_PROBLEM_BROWN_AND_NEEL)
            return set_error(SP_ERR_UNSUPPORTED, "BrownAndNeel with uniaxialcubic anisotropy: the reference's effective "
                                                 "torque is 0/0 (NaN) for this combination");
    }
    if (rq->problem == SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL && rq->solver == SP_SOLVER_IMPLICIT_MIDPOINT) {
        // the reference's Jacobians of this problem are not implemented (assert(false),
        // BrownAndNeelRelaxationEulerSpherical.h:606-671)
        return set_error(SP_ERR_UNSUPPORTED, "Implicit_Midpoint is not available for BrownAndNeelEulerSpherical");
    }
    if (rq->anisotropy == SP_ANISO_UNIAXIALCUBIC && pt) {
        if (!pt->anisotropy_constant_cubic || !pt->uniaxial_rotation)
            return set_error(SP_ERR_INVALID_ARG, "uniaxialcubic needs anisotropy_constant_cubic and uniaxial_rotation");
    }
    if (rq->precision != SP_PREC_FP64 && rq->precision != SP_PREC_MIXED) return set_error(SP_ERR_INVALID_ARG, "bad precision");
    if (rq->noise != SP_NOISE_PHILOX && rq->noise != SP_NOISE_REPLAY) return set_error(SP_ERR_INVALID_ARG, "bad noise source");
    if (rq->n_particles == 0 || rq->n_particles > 0xFFFFFFFFull) return set_error(SP_ERR_INVALID_ARG, "n_particles must be in [1, 2^32)");
    if (rq->oversampling == 0) return set_error(SP_ERR_INVALID_ARG, "oversampling must be positive");
    if (rq->n_steps < rq->oversampling) return set_error(SP_ERR_INVALID_ARG, "n_steps must be >= oversampling");
    if (!(rq->dt > 0.0) || !std::isfinite(rq->dt)) return set_error(SP_ERR_INVALID_ARG, "dt must be positive");
    if (pt && (!pt->magnetic_radius || !pt->saturation_magnetisation || !pt->damping || !pt->gyromagnetic_ratio ||
               !pt->temperature || !pt->anisotropy_constant || !pt->initial_state))
        return set_error(SP_ERR_INVALID_ARG, "missing particle property array");
    if ((rq->problem == SP_PROBLEM_BROWN_AND_NEEL || rq->problem == SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL) && pt &&
        (!pt->hydrodynamic_radius || !pt->viscosity))
        return set_error(SP_ERR_INVALID_ARG, "BrownAndNeel problems need hydrodynamic_radius and viscosity");
    if (rq->solver == SP_SOLVER_IMPLICIT_MIDPOINT && (rq->newton_max_iter == 0 || !(rq->newton_tol > 0.0)))
        return set_error(SP_ERR_INVALID_ARG, "Implicit_Midpoint needs Max_Iterations > 0 and Accuracy_Goal > 0");
    const uint64_t total = (rq->n_steps / rq->oversampling - 1) * rq->oversampling;
    if (rq->noise == SP_NOISE_REPLAY && (!rq->replay_normals || rq->replay_count < total * normals_per_step(*rq, dw)))
        return set_error(SP_ERR_INVALID_ARG, "replay needs n_particles*steps*dim_W normals");
    return SP_OK;
}

// kernel class of a request; EulerMaruyama_Normalized runs the EulerMaruyama kernels with the
// per-step normalize switched on
inline uint32_t kind_for(const sp_request& rq, bool mixed) {
    const int ks = rq.solver == SP_SOLVER_EULER_MARUYAMA_NORMALIZED ? SP_SOLVER_EULER_MARUYAMA : rq.solver;
    return kernel_kind(rq.problem, ks, rq.anisotropy, mixed ? 1 : 0);
}

// 64-particle tiles covered by the mixed-precision check: the request's first 1/8, at least 8 and
// at most 64 tiles (512..4096 particles), never more than the request has
inline uint32_t check_tiles_for(uint32_t n_tiles) {
    return std::min<uint32_t>(n_tiles, std::max<uint32_t>(8u, std::min<uint32_t>(64u, n_tiles / 8)));
}

// the request-level fields of a HostRequest (shapes, kernel class, replayed noise, check tiles)
inline void prepare_shape(const sp_request* rq, HostRequest& hr) {
    hr.rq = *rq;
    request_dims(rq->problem, &hr.dy, &hr.dw, &hr.dout);
    hr.N = rq->n_particles;
    hr.n_out = rq->n_steps / rq->oversampling;
    hr.total_steps = (hr.n_out - 1) * rq->oversampling;
    hr.n_tiles = (uint32_t)((hr.N + 63) / 64);
    hr.mixed = rq->precision == SP_PREC_MIXED && rq->solver != SP_SOLVER_IMPLICIT_MIDPOINT &&
               rq->solver != SP_SOLVER_HEUN_NOT_CONSISTENT;     // both run in FP64 only
    hr.kind = kind_for(*rq, hr.mixed);
    if (rq->noise == SP_NOISE_REPLAY) {
        const uint64_t K = hr.total_steps * normals_per_step(*rq, hr.dw);
        hr.replay.assign(rq->replay_normals, rq->replay_normals + K * hr.N);
        hr.rq.replay_count = K;
    }
    hr.rq.replay_normals = nullptr;
    hr.n_saved = rq->save_interval ? hr.N / rq->save_interval : 0;
    hr.n_check_tiles = check_tiles_for(hr.n_tiles);
}

inline int prepare_request(const sp_request* rq, const sp_particles* pt, HostRequest& hr) {
    if (int e = validate(rq, pt)) return e;
    prepare_shape(rq, hr);
    hr.params.assign((size_t)NPARAM * hr.N, 0.0);
    compute_params(*rq, *pt, hr.params.data(), hr.N);
    hr.y0.resize((size_t)hr.dy * hr.N);
    parallel_ranges(hr.N, [&](uint64_t p0, uint64_t p1) {
        for (uint64_t p = p0; p < p1; ++p)
            for (uint32_t i = 0; i < hr.dy; ++i) hr.y0[i * hr.N + p] = pt->initial_state[p * hr.dy + i];
    });
    const uint64_t n_check = std::min<uint64_t>(hr.N, 64ull * hr.n_check_tiles);
    const double* mu = &hr.params[P_MU * hr.N];
    const double mu_sum = weight_sum_fixed(mu, hr.N), mu_sum_check = weight_sum_fixed(mu, n_check);
    for (uint32_t c = 0; c < hr.dout; ++c) {
        const bool unit = hr.dout == 6 && c < 3;    // BrownAndNeelRelaxation.h:152-159: weight 1; all mu otherwise
        hr.wsum[c] = unit ? (double)hr.N : mu_sum;
        hr.wsum_check[c] = unit ? (double)n_check : mu_sum_check;
    }
    const double* props[8] = {pt->temperature, pt->magnetic_radius, pt->saturation_magnetisation, pt->damping,
                              pt->gyromagnetic_ratio, pt->anisotropy_constant, pt->hydrodynamic_radius, pt->viscosity};
    for (int k = 0; k < 8; ++k) {
        hr.prop_sum[k] = 0.0;
        if (props[k]) for (uint64_t p = 0; p < hr.N; ++p) hr.prop_sum[k] += props[k][p];
    }
    return SP_OK;
}

// a request whose particles the provider kernel draws on the device when its buffers are allocated
inline int prepare_provided(const sp_request* rq, std::shared_ptr<const ProvSpec> spec, HostRequest& hr) {
    if (int e = validate(rq, nullptr, true)) return e;
    if (!spec || spec->types.empty()) return set_error(SP_ERR_INVALID_ARG, "no particle types");
    if (spec->base.N != rq->n_particles || spec->base.seed != rq->seed || spec->base.problem != rq->problem ||
        spec->base.aniso != rq->anisotropy)
        return set_error(SP_ERR_INVALID_ARG, "particle types do not match the request");
    prepare_shape(rq, hr);
    hr.prov = std::move(spec);
    return SP_OK;
}

// ------------------------------------------------------------------ device side of one request
// All buffers of a request are carved from one device block (256-byte aligned pieces).
struct ReqDev {
    double* params = nullptr;
    double* state = nullptr;
    double* replay = nullptr;
    double* out = nullptr;        // [n_out*dout] weighted sums
    double* check = nullptr;      // [n_out*dout][n_check_tiles] weighted sums of each of the first tiles
    unsigned long long* newton_fail = nullptr;
    DevRequest* dreq = nullptr;   // device copy of the request descriptor (plain path)
    double* traj = nullptr;       // [n_out][dout][n_saved] single-simulation outputs
    size_t bytes = 0;             // the request's footprint (admission accounting)
    void* block = nullptr;        // the device block holding all of the above
    size_t block_bytes = 0;
};

// Device blocks of finished served requests, kept for the requests that follow: hipFree waits for the work of every
// stream of the device (an iteration in flight on the server's other lane included) and each hipMalloc costs tens
// of microseconds, so the server worker takes blocks from here and gives them back.  A block is given back only
// when no work on any stream uses it.  At most CAP bytes are kept; used by one thread (the worker).
struct BlockCache {
    static constexpr size_t CAP = size_t(1) << 30;
    std::multimap<size_t, void*> blocks;   // by size
    size_t bytes = 0;
    // a cached block of at least `need` bytes and at most twice that, or null
    void* take(size_t need, size_t& got) {
        auto it = blocks.lower_bound(need);
        if (it == blocks.end() || it->first > 2 * need) return nullptr;
        void* p = it->second;
        got = it->first;
        bytes -= it->first;
        blocks.erase(it);
        return p;
    }
    void give(void* p, size_t n) {
        if (bytes + n > CAP) { (void)hipFree(p); return; }
        blocks.emplace(n, p);
        bytes += n;
    }
    void clear() {
        for (auto& b : blocks) (void)hipFree(b.second);
        blocks.clear();
        bytes = 0;
    }
};

inline size_t traj_bytes(const HostRequest& hr) { return (size_t)hr.n_out * hr.dout * hr.n_saved * 8; }

inline size_t request_bytes(const HostRequest& hr) {
    return (size_t)NPARAM * hr.N * 8 + (size_t)hr.dy * hr.N * 8 + hr.replay.size() * 8 +
           (1 + (size_t)std::max<uint32_t>(1, hr.n_check_tiles)) * hr.n_out * hr.dout * 8 + 64 + sizeof(DevRequest) +
           traj_bytes(hr);
}

inline size_t check_bytes(const HostRequest& hr) {
    return (size_t)hr.n_out * hr.dout * std::max<uint32_t>(1, hr.n_check_tiles) * 8;
}

// the request's block back to `cache`, or freed (which waits for the device) without one
inline void free_dev(ReqDev& d, BlockCache* cache = nullptr) {
    if (d.block) {
        if (cache) cache->give(d.block, d.block_bytes);
        else (void)hipFree(d.block);
    }
    d = ReqDev{};
}

int run_provider(HostRequest& hr, ReqDev& d, hipStream_t s);   // sp_runtime.cpp

// device buffers of a request; its parameters and initial state are uploaded, or drawn on the device by
// the provider kernel (which also fills the weight and property sums of hr)
// (from `cache` when it holds a block of about the size; a failed allocation empties the cache and tries again)
inline int alloc_dev(HostRequest& hr, ReqDev& d, hipStream_t s, BlockCache* cache = nullptr) {
    size_t off = 0;
    auto piece = [&](size_t n) { const size_t o = off; off += (n + 255) / 256 * 256; return o; };
    const size_t o_params = piece((size_t)NPARAM * hr.N * 8), o_state = piece((size_t)hr.dy * hr.N * 8);
    const size_t o_replay = piece(hr.replay.size() * 8), o_out = piece(hr.n_out * hr.dout * 8);
    const size_t o_check = piece(check_bytes(hr)), o_nf = piece(8), o_dreq = piece(sizeof(DevRequest));
    const size_t o_traj = piece(traj_bytes(hr));
    size_t got = 0;
    void* blk = cache ? cache->take(off, got) : nullptr;
    if (!blk) {
        hipError_t e = hipMalloc(&blk, off);
        if (e != hipSuccess && cache && !cache->blocks.empty()) {
            (void)hipGetLastError();
            cache->clear();
            e = hipMalloc(&blk, off);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();      // clear the failed allocation's error
            d = ReqDev{};
            return set_error(e == hipErrorOutOfMemory ? SP_ERR_OUT_OF_MEMORY : SP_ERR_DEVICE,
                             std::string("allocating request buffers: ") + hipGetErrorString(e));
        }
        got = off;
    }
    char* b = (char*)blk;
    d.block = blk;
    d.block_bytes = got;
    d.params = (double*)(b + o_params);
    d.state = (double*)(b + o_state);
    d.replay = hr.replay.empty() ? nullptr : (double*)(b + o_replay);
    d.out = (double*)(b + o_out);
    d.check = (double*)(b + o_check);
    d.newton_fail = (unsigned long long*)(b + o_nf);
    d.dreq = (DevRequest*)(b + o_dreq);
    d.traj = traj_bytes(hr) ? (double*)(b + o_traj) : nullptr;
    d.bytes = request_bytes(hr);
    if (hr.prov) {
        if (int e2 = run_provider(hr, d, s)) { free_dev(d); return e2; }
    } else {
        SP_HIP(hipMemcpyAsync(d.params, hr.params.data(), hr.params.size() * 8, hipMemcpyHostToDevice, s));
        SP_HIP(hipMemcpyAsync(d.state, hr.y0.data(), hr.y0.size() * 8, hipMemcpyHostToDevice, s));
    }
    if (!hr.replay.empty()) SP_HIP(hipMemcpyAsync(d.replay, hr.replay.data(), hr.replay.size() * 8, hipMemcpyHostToDevice, s));
    SP_HIP(hipMemsetAsync(d.out, 0, hr.n_out * hr.dout * 8, s));
    SP_HIP(hipMemsetAsync(d.check, 0, check_bytes(hr), s));
    SP_HIP(hipMemsetAsync(d.newton_fail, 0, 8, s));
    return SP_OK;
}

inline DevRequest make_dev_request(const HostRequest& hr, const ReqDev& d) {
    const sp_request& r = hr.rq;
    DevRequest dr{};
    dr.dt = r.dt;
    dr.sqdt = std::sqrt(r.dt);
    const bool rot = (r.flags & SP_FLAG_COORD_TRANSFORM) != 0;
    dr.tmin = rot ? r.min_angle : -1e300;   // no chart rotation: theta may leave [0, pi] freely
    dr.tmax = rot ? (3.14159265358979323846 - r.min_angle) : 1e300;
    dr.ctmin = rot ? std::cos(dr.tmin) : 2.0;     // with no rotation no state is ever in the rotated chart
    dr.ctmax = rot ? std::cos(dr.tmax) : -2.0;
    const bool erot = (r.flags & SP_FLAG_EULER_TRANSFORM) != 0;     // BrownAndNeelEulerSpherical Brown chart
    dr.bmin = erot ? r.brown_min_angle : -1e300;
    dr.bmax = erot ? (3.14159265358979323846 - r.brown_min_angle) : 1e300;
    dr.bcmin = erot ? std::cos(dr.bmin) : 2.0;    // with no rotation no body is ever in the rotated chart
    dr.bcmax = erot ? std::cos(dr.bmax) : -2.0;
    dr.newton_tol = r.newton_tol;
    dr.field = field_params(r);
    dr.n_particles = hr.N;
    dr.seed = r.seed;
    dr.replay_count = hr.replay.empty() ? 0 : r.replay_count;
    dr.oversampling = (uint32_t)r.oversampling;
    const bool norm = (r.flags & SP_FLAG_RENORMALIZE) || r.solver == SP_SOLVER_EULER_MARUYAMA_NORMALIZED;
    dr.flags = ((r.flags & SP_FLAG_BN_SIMPLE_MODEL) ? RQ_BN_SIMPLE : 0u) | (hr.replay.empty() ? 0u : RQ_REPLAY) |
               (norm ? RQ_NORMALIZE : 0u);
    dr.newton_max_iter = r.newton_max_iter;
    dr.n_tiles = hr.n_tiles;
    dr.kind = hr.kind;
    dr.params = d.params;
    dr.state = d.state;
    dr.replay = d.replay;
    dr.newton_fail = d.newton_fail;
    dr.traj = d.traj;
    dr.n_saved = d.traj ? hr.n_saved : 0;
    dr.save_interval = d.traj ? r.save_interval : 0;
    return dr;
}

// ------------------------------------------------------------------ results
struct Outputs {
    std::vector<double> time, field, mean, weight, final_state, traj;
    std::vector<double> check_tiles;   // [n_out*dout][n_check_tiles] raw weighted tile sums
    double device_ms = 0.0, latency_ms = 0.0, particle_steps = 0.0, check_error = -1.0, check_noise = 0.0;
    double check_bias = -1.0, check_bias_se = 0.0;
    uint32_t used_fallback = 0, precision_used = SP_PREC_FP64;
    uint64_t newton_failures = 0;
    bool check_failed = false;
};

// axes: the request's time and field axes when the caller has them already (a prepared request executed again)
inline int download_outputs(const HostRequest& hr, const ReqDev& d, hipStream_t s, bool keep_final, bool want_check, Outputs& o,
                            const Outputs* axes = nullptr) {
    const uint64_t M = hr.n_out * hr.dout;
    std::vector<double> sums(M);
    SP_HIP(hipMemcpyAsync(sums.data(), d.out, M * 8, hipMemcpyDeviceToHost, s));
    if (want_check) {
        o.check_tiles.resize(check_bytes(hr) / 8);
        SP_HIP(hipMemcpyAsync(o.check_tiles.data(), d.check, check_bytes(hr), hipMemcpyDeviceToHost, s));
    }
    unsigned long long nf = 0;
    SP_HIP(hipMemcpyAsync(&nf, d.newton_fail, 8, hipMemcpyDeviceToHost, s));
    if (d.traj) {
        o.traj.resize(traj_bytes(hr) / 8);
        SP_HIP(hipMemcpyAsync(o.traj.data(), d.traj, traj_bytes(hr), hipMemcpyDeviceToHost, s));
    }
    if (keep_final) {
        o.final_state.resize(hr.N * hr.dy);
        std::vector<double> soa(hr.N * hr.dy);
        SP_HIP(hipMemcpyAsync(soa.data(), d.state, soa.size() * 8, hipMemcpyDeviceToHost, s));
        SP_HIP(hipStreamSynchronize(s));
        for (uint64_t p = 0; p < hr.N; ++p)
            for (uint32_t i = 0; i < hr.dy; ++i) o.final_state[p * hr.dy + i] = soa[i * hr.N + p];
    }
    SP_HIP(hipStreamSynchronize(s));
    o.newton_failures = nf;
    // MeanSimulationResult::normalize (MeanSimulationResult.h:63-75)
    bool zero = true;
    for (uint32_t c = 0; c < hr.dout; ++c) if (hr.wsum[c] != 0.0) zero = false;
    o.mean.assign(M, 0.0);
    for (uint64_t k = 0; k < hr.n_out; ++k)
        for (uint32_t c = 0; c < hr.dout; ++c) o.mean[k * hr.dout + c] = zero ? sums[k * hr.dout + c] : sums[k * hr.dout + c] / hr.wsum[c];
    o.weight.assign(hr.wsum, hr.wsum + hr.dout);
    if (axes && axes->time.size() == hr.n_out) {
        o.time = axes->time;
        o.field = axes->field;
    } else {
        o.time.resize(hr.n_out);
        o.field.resize(hr.n_out * 3);
        time_and_field(hr.rq, o.time.data(), o.field.data());
    }
    o.particle_steps = (double)hr.N * (double)hr.total_steps;
    o.precision_used = hr.mixed ? SP_PREC_MIXED : SP_PREC_FP64;
    return SP_OK;
}

inline void copy_result(const HostRequest& hr, const Outputs& o, sp_result* r) {
    r->n_out = hr.n_out;
    r->dim_out = hr.dout;
    r->dim_y = hr.dy;
    if (r->time) std::copy(o.time.begin(), o.time.end(), r->time);
    if (r->field) std::copy(o.field.begin(), o.field.end(), r->field);
    if (r->mean) std::copy(o.mean.begin(), o.mean.end(), r->mean);
    if (r->weight) std::copy(o.weight.begin(), o.weight.end(), r->weight);
    if (r->final_state && !o.final_state.empty()) std::copy(o.final_state.begin(), o.final_state.end(), r->final_state);
    r->n_saved = o.traj.empty() ? 0 : hr.n_saved;
    if (r->trajectories && !o.traj.empty()) std::copy(o.traj.begin(), o.traj.end(), r->trajectories);
    r->device_ms = o.device_ms;
    r->latency_ms = o.latency_ms;
    r->particle_steps = o.particle_steps;
    r->check_error = o.check_error;
    r->check_noise = o.check_noise;
    r->check_bias = o.check_bias;
    r->check_bias_se = o.check_bias_se;
    r->used_fallback = o.used_fallback;
    r->precision_used = o.precision_used;
    r->newton_failures = o.newton_failures;
    for (int k = 0; k < 8; ++k) r->particle_means[k] = hr.prop_sum[k] / (double)hr.N;
}

// ------------------------------------------------------------------ context
// Device scratch and streams of one sequence of iterations: the persistent launches of an iteration's kernel classes
// run side by side (class k > 0 on side[(k - 1) % n_side], forked from and joined back into `stream` with events, all
// streams at the default priority; with n_side = 0 they follow each other on `stream`); tables, partial sums and work
// counters are reused from iteration to iteration.  The synchronous paths and the server's first lane use the
// context's set (two side streams); the server's Implicit_Midpoint lane has its own stream and scratch, so that its
// iterations overlap the first lane's.  A context creates four streams in all: the runtime maps streams onto four
// hardware queues per process by default, and two streams sharing a queue would run their kernels one after the other.
struct IterRes {
    static constexpr int N_SIDE = 2;
    static constexpr int N_CLASS_EV = 32;   // classes timed per iteration (further classes are not timed)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int n_side = 0;                         // side streams created (0..N_SIDE)
    hipStream_t side[N_SIDE] = {nullptr, nullptr};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[N_SIDE] = {nullptr, nullptr};
    double* partial = nullptr;
    size_t partial_bytes = 0;
    void* table = nullptr;                  // device copy of request / work / reduce tables
    size_t table_bytes = 0;
    void* pinned = nullptr;                 // pinned host staging of the tables
    size_t pinned_bytes = 0;
    uint32_t* counters = nullptr;           // per-launch work counters
    size_t n_counters = 0;
    hipEvent_t class_ev[2 * N_CLASS_EV] = {};   // per-class timing (the server's iterations)
    bool class_timing = false;
};
// streams and events of r: `stream` is used as r's own when given, else one is created; n_side side streams
inline int init_res(IterRes& r, hipStream_t stream, int n_side) {
    if (stream) r.stream = stream;
    else { SP_HIP(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking)); r.own_stream = true; }
    for (int i = 0; i < std::min(n_side, IterRes::N_SIDE); ++i) {
        SP_HIP(hipStreamCreateWithFlags(&r.side[i], hipStreamNonBlocking));
        SP_HIP(hipEventCreateWithFlags(&r.join_ev[i], hipEventDisableTiming));
        r.n_side = i + 1;
    }
    SP_HIP(hipEventCreateWithFlags(&r.fork_ev, hipEventDisableTiming));
    return SP_OK;
}
// the per-class timing events (created on first use by the server; without them the iteration budget is off)
inline void init_class_timing(IterRes& r) {
    if (r.class_timing) return;
    r.class_timing = true;
    for (int i = 0; i < 2 * IterRes::N_CLASS_EV && r.class_timing; ++i)
        if (!r.class_ev[i] && hipEventCreate(&r.class_ev[i]) != hipSuccess) {
            (void)hipGetLastError();
            r.class_timing = false;
        }
}
inline void destroy_res(IterRes& r) {
    if (r.stream) hipStreamSynchronize(r.stream);
    for (int i = 0; i < IterRes::N_SIDE; ++i)
        if (r.side[i]) hipStreamSynchronize(r.side[i]);
    if (r.partial) hipFree(r.partial);
    if (r.table) hipFree(r.table);
    if (r.pinned) hipHostFree(r.pinned);
    if (r.counters) hipFree(r.counters);
    for (int i = 0; i < IterRes::N_SIDE; ++i) {
        if (r.side[i]) hipStreamDestroy(r.side[i]);
        if (r.join_ev[i]) hipEventDestroy(r.join_ev[i]);
    }
    if (r.fork_ev) hipEventDestroy(r.fork_ev);
    for (int i = 0; i < 2 * IterRes::N_CLASS_EV; ++i)
        if (r.class_ev[i]) hipEventDestroy(r.class_ev[i]);
    if (r.own_stream && r.stream) hipStreamDestroy(r.stream);
    r = IterRes{};
}

struct Job {
    uint64_t id = 0;
    HostRequest hr;
    Outputs out;
    int status = SP_ERR_NOT_READY;       // SP_ERR_NOT_READY while pending/running
    bool cancel = false;
    bool keep_final = false;
    Clock::time_point t_submit;
    double device_ms_acc = 0.0;
    uint64_t steps_done = 0;              // steps of the request's particles completed (chunk granularity)
};

struct Slot {                 // an admitted request (or the FP64 shadow of a mixed one)
    std::shared_ptr<Job> job;
    HostRequest* hr = nullptr;
    HostRequest shadow_req;   // storage when this slot is a shadow
    ReqDev dev;
    DevRequest dr{};
    uint64_t next_step = 0;
    uint64_t chunk = 0;       // steps per iteration
    uint64_t tuned_chunk = 0; // the class's tuned chunk (the upper bound of `chunk`)
    bool first = true;
    bool done = false;
    bool is_shadow = false;
    Slot* shadow = nullptr;   // FP64 check slot of a mixed request
    Slot* parent = nullptr;
    double work_share = 0.0;  // particle-steps in the current iteration
};

struct sp_ctx_impl;
}  // namespace sp

struct sp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    std::mutex mu;                          // protects everything below
    std::map<uint64_t, sp_tuning> tuning;   // by (problem, solver, precision)
    uint64_t mem_budget = 0;
    uint32_t max_active = 512;              // admitted requests at once (512 small requests fill the device)
    // server
    std::thread worker;
    bool worker_started = false;
    bool stop = false;
    std::condition_variable cv;             // wakes the worker
    std::condition_variable done_cv;        // wakes waiters
    std::deque<std::shared_ptr<sp::Job>> pending;
    std::map<uint64_t, std::shared_ptr<sp::Job>> jobs;
    uint64_t next_id = 1;
    uint64_t mem_used = 0;
    // statistics
    double stat_active = 0, stat_completed = 0, stat_iterations = 0, stat_busy_ms = 0, stat_particle_steps = 0, stat_fallbacks = 0;
    sp::Clock::time_point t_start = sp::Clock::now();
    // scratch owned by the worker (or by a synchronous call holding run_mu); res.stream is `stream`
    std::mutex run_mu;                      // serialises device use between sp_run and the worker
    std::atomic<int> run_waiters{0};        // synchronous calls waiting for run_mu (sp::RunLock)
    sp::IterRes res;
    sp::BlockCache blocks;                  // request blocks for reuse (the server worker's own)
    // plain path: largest partial-sum buffer for reducing a whole pass at once (larger passes are reduced
    // chunk by chunk)
    uint64_t whole_pass_limit = 4ull << 30;
    // server: the next n iterations fail before their launches (a host-side hook for the error paths)
    uint32_t inject_failures = 0;
    // server: each kernel class's chunk is sized so that its launch takes about this long (ms), at most its tuned
    // chunk; 0 keeps the tuned chunk.  The worker thread owns the measured costs and the timing events.
    double iteration_budget_ms = 5.0;
    std::map<uint32_t, double> class_ns_per_ps;   // kernel class -> smoothed device ns per particle-step
    // last synchronous run statistics
    double last_stats[8] = {0};
};

namespace sp {

// run_mu held by a synchronous call.  The server's worker keeps run_mu while iterations are in flight; the count of
// callers waiting for it makes the worker stop issuing iterations and let them in.
struct RunLock {
    sp_ctx* c;
    explicit RunLock(sp_ctx* ctx) : c(ctx) {
        c->run_waiters.fetch_add(1);
        c->run_mu.lock();
        c->run_waiters.fetch_sub(1);
    }
    ~RunLock() { c->run_mu.unlock(); }
    RunLock(const RunLock&) = delete;
    RunLock& operator=(const RunLock&) = delete;
};

inline uint64_t tuning_key(int problem, int solver, int precision) { return ((uint64_t)problem << 16) | ((uint64_t)solver << 8) | (uint64_t)precision; }

inline sp_tuning default_tuning(const sp_ctx* ctx, const HostRequest& hr) {
    sp_tuning t{};
    t.block_size = 256;
    t.chunk_steps = 4000;
    t.use_graph = 1;
    t.persistent_waves_per_simd = 2;
    (void)ctx; (void)hr;
    return t;
}

// work-group size of a launch of `tiles` 64-particle tiles: the tuned size, or one wave per work-group when the launch
// has fewer tiles than the device has SIMDs, so that its waves spread over every compute unit instead of sharing the
// front end of a few (16,384-particle NeelSpherical FP64: +5%; served iterations: Implicit_Midpoint p50 -9%)
inline uint32_t launch_block(uint32_t tuned, uint64_t tiles) {
    if (tiles < device_simds()) return 64;
    return std::max<uint32_t>(64, std::min<uint32_t>(MAX_BLOCK, tuned ? tuned / 64 * 64 : MAX_BLOCK));
}

inline sp_tuning get_tuning(sp_ctx* ctx, const HostRequest& hr) {
    auto it = ctx->tuning.find(tuning_key(hr.rq.problem, hr.rq.solver, hr.mixed ? SP_PREC_MIXED : SP_PREC_FP64));
    return it != ctx->tuning.end() ? it->second : default_tuning(ctx, hr);
}

inline int ensure_partial(IterRes& r, size_t bytes) {
    if (bytes <= r.partial_bytes) return SP_OK;
    if (r.partial) hipFree(r.partial);
    r.partial = nullptr;
    r.partial_bytes = 0;
    hipError_t e = hipMalloc(&r.partial, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();      // a failed allocation must not be reported by a later launch check
        r.partial = nullptr;
        return set_error(SP_ERR_OUT_OF_MEMORY, std::string("partial buffer: ") + hipGetErrorString(e));
    }
    r.partial_bytes = bytes;
    return SP_OK;
}
inline int ensure_table(IterRes& r, size_t bytes) {
    if (bytes > r.table_bytes) {
        if (r.table) hipFree(r.table);
        r.table = nullptr;
        r.table_bytes = 0;
        SP_HIP(hipMalloc(&r.table, bytes));
        r.table_bytes = bytes;
    }
    if (bytes > r.pinned_bytes) {
        if (r.pinned) hipHostFree(r.pinned);
        r.pinned = nullptr;
        r.pinned_bytes = 0;
        SP_HIP(hipHostMalloc(&r.pinned, bytes, hipHostMallocDefault));
        r.pinned_bytes = bytes;
    }
    return SP_OK;
}
inline int ensure_counters(IterRes& r, size_t n) {
    if (n <= r.n_counters) return SP_OK;
    if (r.counters) hipFree(r.counters);
    r.counters = nullptr;
    r.n_counters = 0;
    SP_HIP(hipMalloc(&r.counters, n * sizeof(uint32_t)));
    r.n_counters = n;
    return SP_OK;
}
inline int ensure_partial(sp_ctx* ctx, size_t bytes) { return ensure_partial(ctx->res, bytes); }
inline int ensure_table(sp_ctx* ctx, size_t bytes) { return ensure_table(ctx->res, bytes); }
inline int ensure_counters(sp_ctx* ctx, size_t n) { return ensure_counters(ctx->res, n); }

// defined in sp_runtime.cpp
struct Prepared {
    HostRequest hr;
    ReqDev d;
    double* state0 = nullptr;     // device copy of the initial state
    bool keep_final = false;
    Outputs axes;                 // time and field axes, computed at the first execution
    std::vector<hipEvent_t> events;   // per-chunk timing events, created at the first execution
};
void free_prepared(Prepared& pr);
int prepare_plain(sp_ctx* ctx, Prepared& pr);
int execute_plain(sp_ctx* ctx, Prepared& pr, Outputs& o, const sp_tuning& tun);
int run_plain(sp_ctx* ctx, HostRequest& hr, Outputs& o, bool keep_final, const sp_tuning& tun);
// defined in sp_server.cpp
void stop_worker(sp_ctx* ctx);
// queues a validated request on the server (sp_submit after its checks)
int submit_request(sp_ctx* ctx, HostRequest&& hr, bool keep_final, uint64_t* id);
}  // namespace sp

struct sp_prepared {
    sp::Prepared pr;
    sp_ctx* ctx = nullptr;
};
