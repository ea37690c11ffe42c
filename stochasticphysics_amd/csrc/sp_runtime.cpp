// sp_runtime.cpp -- host runtime of the ensemble SDE engine: requests, device
// memory, the plain one-request path, the persistent batched path with
// hipGraph-captured chunks, and the continuous-batching request server.
//
// Replaces the orchestration of the reference's SimulationManager
// (Simulator/SimulationManager.h:125-221, :545-599): instead of one host thread
// task per particle and a mutex-protected merge of per-particle trajectories,
// one host thread drives a HIP stream, particles run 64 to a wave, and the
// ensemble mean is reduced on the device.
#include "sp_engine.h"

namespace sp {

static thread_local std::string g_err;
int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}


// ------------------------------------------------------------------ plain path
// One request, one kernel launch per chunk over all its tiles, then the reduction.
// A Prepared request keeps its parameters and initial state resident in HBM, so
// it can be executed repeatedly without host-to-device traffic.
void free_prepared(Prepared& pr) {
    for (auto e : pr.events) hipEventDestroy(e);
    pr.events.clear();
    free_dev(pr.d);
    if (pr.state0) hipFree(pr.state0);
    pr.state0 = nullptr;
}

int prepare_plain(sp_ctx* ctx, Prepared& pr) {
    SP_HIP(hipSetDevice(ctx->device));
    if (int e = alloc_dev(pr.hr, pr.d, ctx->stream)) return e;
    const size_t state_bytes = (size_t)pr.hr.dy * pr.hr.N * 8;
    hipError_t e = hipMalloc(&pr.state0, state_bytes);
    if (e != hipSuccess) { free_dev(pr.d); return set_error(SP_ERR_OUT_OF_MEMORY, "initial state copy"); }
    SP_HIP(hipMemcpyAsync(pr.state0, pr.d.state, state_bytes, hipMemcpyDeviceToDevice, ctx->stream));
    SP_HIP(hipStreamSynchronize(ctx->stream));
    return SP_OK;
}

int execute_plain(sp_ctx* ctx, Prepared& pr, Outputs& o, const sp_tuning& tun) {
    RunLock lk(ctx);
    HostRequest& hr = pr.hr;
    ReqDev& d = pr.d;
    hipStream_t s = ctx->stream;
    SP_HIP(hipSetDevice(ctx->device));
    const auto t0 = Clock::now();
    SP_HIP(hipMemcpyAsync(d.state, pr.state0, (size_t)hr.dy * hr.N * 8, hipMemcpyDeviceToDevice, s));
    SP_HIP(hipMemsetAsync(d.out, 0, hr.n_out * hr.dout * 8, s));
    SP_HIP(hipMemsetAsync(d.check, 0, check_bytes(hr), s));
    SP_HIP(hipMemsetAsync(d.newton_fail, 0, 8, s));
    const DevRequest dr = make_dev_request(hr, d);
    SP_HIP(hipMemcpyAsync(d.dreq, &dr, sizeof(DevRequest), hipMemcpyHostToDevice, s));
    const uint64_t OS = hr.rq.oversampling;
    const uint64_t align = lcm64(OS, 4);
    // chunk so that a launch's partial sums fit a 256 MiB buffer
    const uint64_t row_bytes = (uint64_t)hr.dout * hr.n_tiles * 8;
    uint64_t max_rows = std::max<uint64_t>(2, (256ull << 20) / row_bytes);
    uint64_t C = std::max<uint64_t>(align, ((max_rows - 1) * OS) / align * align);
    if (tun.chunk_steps && tun.chunk_steps < hr.total_steps) C = std::min<uint64_t>(C, std::max<uint64_t>(align, tun.chunk_steps / align * align));
    C = std::min<uint64_t>(C, std::max<uint64_t>(hr.total_steps, align));
    const uint64_t rows_max = C / OS + 1;
    // when the partial sums of the whole pass fit the context's reduction limit (4 GiB by default) and a
    // quarter of the free device memory, every chunk writes its own rows and one reduction runs at the end
    // (one kernel boundary per chunk instead of two); otherwise, or when that buffer cannot be allocated,
    // each chunk is reduced before the next one reuses a buffer of one chunk's rows
    bool whole = hr.n_out * row_bytes <= ctx->whole_pass_limit && hr.n_out * hr.dout < (1ull << 32);
    if (whole && hr.n_out * row_bytes > ctx->res.partial_bytes) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || hr.n_out * row_bytes > free_b / 4) whole = false;
    }
    if (whole && ensure_partial(ctx, hr.n_out * row_bytes) != SP_OK) whole = false;
    if (!whole) {
        if (int e = ensure_partial(ctx, rows_max * row_bytes)) return e;
    }
    const uint64_t n_chunks = std::max<uint64_t>(1, (hr.total_steps + C - 1) / C);
    if (int e = ensure_table(ctx, n_chunks * sizeof(ReduceWork) + 256)) return e;
    ReduceWork* rw_host = (ReduceWork*)ctx->res.pinned;
    ReduceWork* rw_dev = (ReduceWork*)ctx->res.table;
    const uint32_t block = launch_block(tun.block_size, hr.n_tiles);
    const uint32_t waves_per_block = block / 64;
    const dim3 grid((hr.n_tiles + waves_per_block - 1) / waves_per_block);
    // timing events, kept with the prepared request so that repeated executions do not create them again
    if (pr.events.size() != 2 * n_chunks + 2) {
        for (auto e : pr.events) hipEventDestroy(e);
        pr.events.clear();
        for (uint64_t k = 0; k < 2 * n_chunks + 2; ++k) {
            hipEvent_t e;
            SP_HIP(hipEventCreate(&e));
            pr.events.push_back(e);
        }
    }
    const std::vector<hipEvent_t>& ev = pr.events;
    // the reduction work goes to the device in one copy, ahead of the launches
    std::vector<uint32_t> chunk_rows(n_chunks);
    std::vector<double*> chunk_partial(n_chunks, ctx->res.partial);
    {
        uint64_t sb = 0;
        for (uint64_t k = 0; k < n_chunks; ++k) {
            const uint64_t se = std::min(sb + C, hr.total_steps);
            const uint64_t first_window = sb == 0 ? 0 : sb / OS + 1;
            const uint64_t rows = (se - sb) / OS + (sb == 0 ? 1 : 0);
            chunk_rows[k] = (uint32_t)(rows * hr.dout);
            if (whole) chunk_partial[k] = ctx->res.partial + first_window * hr.dout * hr.n_tiles;
            else rw_host[k] = ReduceWork{ctx->res.partial, hr.n_tiles, hr.n_check_tiles, chunk_rows[k],
                                         (uint32_t)(first_window * hr.dout), d.out, d.check};
            sb = se;
        }
        if (whole) rw_host[0] = ReduceWork{ctx->res.partial, hr.n_tiles, hr.n_check_tiles, (uint32_t)(hr.n_out * hr.dout), 0u,
                                           d.out, d.check};
    }
    SP_HIP(hipMemcpyAsync(rw_dev, rw_host, (whole ? 1 : n_chunks) * sizeof(ReduceWork), hipMemcpyHostToDevice, s));
    SP_HIP(hipEventRecord(ev[2 * n_chunks], s));
    uint64_t sb = 0;
    for (uint64_t k = 0; k < n_chunks; ++k) {
        const uint64_t se = std::min(sb + C, hr.total_steps);
        SP_HIP(hipEventRecord(ev[2 * k], s));
        if (!launch_plain(hr.kind, grid, dim3(block), s, d.dreq, sb, se, chunk_partial[k], hr.n_tiles))
            return set_error(SP_ERR_UNSUPPORTED, "no kernel for this problem/solver/anisotropy/precision");
        SP_HIP(hipGetLastError());
        SP_HIP(hipEventRecord(ev[2 * k + 1], s));
        if (!whole) {
            launch_reduce(s, rw_dev + k, 1, chunk_rows[k]);
            SP_HIP(hipGetLastError());
        }
        sb = se;
    }
    if (whole) {
        launch_reduce(s, rw_dev, 1, (uint32_t)(hr.n_out * hr.dout));
        SP_HIP(hipGetLastError());
    }
    SP_HIP(hipEventRecord(ev[2 * n_chunks + 1], s));
    SP_HIP(hipStreamSynchronize(s));
    double kern_ms = 0.0;
    for (uint64_t k = 0; k < n_chunks; ++k) {
        float m = 0.f;
        SP_HIP(hipEventElapsedTime(&m, ev[2 * k], ev[2 * k + 1]));
        kern_ms += m;
    }
    float total_ms = 0.f;
    SP_HIP(hipEventElapsedTime(&total_ms, ev[2 * n_chunks], ev[2 * n_chunks + 1]));
    if (int e = download_outputs(hr, d, s, pr.keep_final, false, o, &pr.axes)) return e;
    if (pr.axes.time.size() != hr.n_out) { pr.axes.time = o.time; pr.axes.field = o.field; }
    o.device_ms = total_ms;
    o.latency_ms = ms_since(t0, Clock::now());
    KernelInfo ki{};
    kernel_info(hr.kind, &ki);
    ctx->last_stats[0] = (double)n_chunks;
    ctx->last_stats[1] = kern_ms;
    ctx->last_stats[2] = total_ms - kern_ms;
    ctx->last_stats[3] = o.latency_ms;
    ctx->last_stats[4] = o.particle_steps;
    ctx->last_stats[5] = ki.vgprs;
    ctx->last_stats[6] = (double)grid.x;
    ctx->last_stats[7] = (double)block;
    return SP_OK;
}

// ------------------------------------------------------------------ device-side particle provider
// One thread per particle draws the particle, evaluates its parameter block and start state into the
// request's buffers, and each work-group leaves the sums of its particles' properties and weights; the host
// adds the per-group sums in group order, so the totals do not depend on scheduling.
int run_provider(HostRequest& hr, ReqDev& d, hipStream_t s) {
    const ProvSpec& spec = *hr.prov;
    const uint32_t blocks = (uint32_t)((hr.N + PROV_BLOCK - 1) / PROV_BLOCK);
    ProvType* d_types = nullptr;
    double* d_sums = nullptr;
    struct Guard {
        ProvType** t; double** s;
        ~Guard() { if (*t) hipFree(*t); if (*s) hipFree(*s); }
    } guard{&d_types, &d_sums};
    SP_HIP(hipMalloc(&d_types, spec.types.size() * sizeof(ProvType)));
    SP_HIP(hipMalloc(&d_sums, (size_t)blocks * PROV_NSUM * 8));
    SP_HIP(hipMemcpyAsync(d_types, spec.types.data(), spec.types.size() * sizeof(ProvType), hipMemcpyHostToDevice, s));
    ProvDesc desc = spec.base;
    desc.N = hr.N;
    desc.n_types = (uint32_t)spec.types.size();
    desc.n_check = std::min<uint64_t>(hr.N, 64ull * hr.n_check_tiles);
    desc.dy = hr.dy;
    desc.types = d_types;
    desc.params = d.params;
    desc.y0 = d.state;
    desc.sums = d_sums;
    launch_provide(s, blocks, desc);
    SP_HIP(hipGetLastError());
    std::vector<double> sums((size_t)blocks * PROV_NSUM);
    SP_HIP(hipMemcpyAsync(sums.data(), d_sums, sums.size() * 8, hipMemcpyDeviceToHost, s));
    SP_HIP(hipStreamSynchronize(s));
    double acc[PROV_NSUM] = {0};
    for (uint32_t b = 0; b < blocks; ++b)
        for (int k = 0; k < PROV_NSUM; ++k) acc[k] += sums[(size_t)b * PROV_NSUM + k];
    for (uint32_t c = 0; c < hr.dout; ++c) {
        const bool unit = hr.dout == 6 && c < 3;      // BrownAndNeelRelaxation.h:152-159: weight 1 for the axis
        hr.wsum[c] = unit ? (double)hr.N : acc[8];
        hr.wsum_check[c] = unit ? (double)desc.n_check : acc[9];
    }
    for (int k = 0; k < 8; ++k) hr.prop_sum[k] = acc[k];
    return SP_OK;
}

int run_plain(sp_ctx* ctx, HostRequest& hr, Outputs& o, bool keep_final, const sp_tuning& tun) {
    Prepared pr;
    pr.hr = std::move(hr);
    pr.keep_final = keep_final;
    if (int e = prepare_plain(ctx, pr)) return e;
    const int rc = execute_plain(ctx, pr, o, tun);
    free_prepared(pr);
    hr = std::move(pr.hr);
    return rc;
}

}  // namespace sp

// ================================================================== C ABI (part 1)
using namespace sp;

extern "C" {

const char* sp_last_error(void) { return g_err.c_str(); }
int sp_abi_version(void) { return SP_ABI_VERSION; }

int sp_request_shape(const sp_request* req, uint64_t* n_out, uint32_t* dim_y, uint32_t* dim_w, uint32_t* dim_out) {
    if (!req) return set_error(SP_ERR_INVALID_ARG, "request is NULL");
    uint32_t dy, dw, dout;
    if (request_dims(req->problem, &dy, &dw, &dout)) return set_error(SP_ERR_UNSUPPORTED, "unknown problem type");
    if (req->oversampling == 0) return set_error(SP_ERR_INVALID_ARG, "oversampling must be positive");
    if (n_out) *n_out = req->n_steps / req->oversampling;
    if (dim_y) *dim_y = dy;
    if (dim_w) *dim_w = dw;
    if (dim_out) *dim_out = dout;
    return SP_OK;
}

int sp_compute_particle_params(const sp_request* req, const sp_particles* parts, double* out, uint32_t* n_params) {
    if (n_params) *n_params = NPARAM;
    if (!out) return SP_OK;
    HostRequest hr;
    if (int e = validate(req, parts)) return e;
    compute_params(*req, *parts, out, req->n_particles);
    return SP_OK;
}

int sp_host_normals(uint64_t seed, uint64_t particle, uint64_t first, uint64_t count, double* out) {
    if (!out && count) return set_error(SP_ERR_INVALID_ARG, "out is NULL");
    double z[4];
    uint64_t have = ~0ull;
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t k = first + i;
        if ((k >> 2) != have) { box_muller4_f64(noise_block(seed, particle, k >> 2, 0u), z); have = k >> 2; }
        out[i] = z[k & 3];
    }
    return SP_OK;
}

int sp_ctx_create(int device, sp_ctx** out) {
    if (!out) return set_error(SP_ERR_INVALID_ARG, "out is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return set_error(SP_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return set_error(SP_ERR_INVALID_ARG, "device index out of range");
    std::unique_ptr<sp_ctx> c(new sp_ctx());
    c->device = device;
    SP_HIP(hipSetDevice(device));
    SP_HIP(hipGetDeviceProperties(&c->prop, device));
    SP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    if (int e = init_res(c->res, c->stream, IterRes::N_SIDE)) {
        destroy_res(c->res);
        hipStreamDestroy(c->stream);
        return e;
    }
    size_t free_b = 0, total_b = 0;
    SP_HIP(hipMemGetInfo(&free_b, &total_b));
    c->mem_budget = (uint64_t)(0.5 * (double)free_b);
    *out = c.release();
    return SP_OK;
}

int sp_ctx_set_tuning(sp_ctx* ctx, int problem, int solver, int precision, const sp_tuning* t) {
    if (!ctx || !t) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    if (t->block_size < 64 || t->block_size > MAX_BLOCK || t->block_size % 64)
        return set_error(SP_ERR_INVALID_ARG, "block_size must be 64, 128, 192 or 256");
    if (t->chunk_steps == 0 || t->persistent_waves_per_simd == 0 || t->persistent_waves_per_simd > 8)
        return set_error(SP_ERR_INVALID_ARG, "chunk_steps must be > 0 and persistent_waves_per_simd in 1..8");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->tuning[tuning_key(problem, solver, precision)] = *t;
    return SP_OK;
}
int sp_ctx_get_tuning(sp_ctx* ctx, int problem, int solver, int precision, sp_tuning* t) {
    if (!ctx || !t) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tuning.find(tuning_key(problem, solver, precision));
    if (it == ctx->tuning.end()) return set_error(SP_ERR_NOT_FOUND, "no tuning stored");
    *t = it->second;
    return SP_OK;
}
int sp_ctx_set_memory_budget(sp_ctx* ctx, uint64_t bytes, uint32_t max_active) {
    if (!ctx) return set_error(SP_ERR_INVALID_ARG, "NULL context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->mem_budget = bytes;
    ctx->max_active = max_active ? max_active : 1;
    return SP_OK;
}

int sp_ctx_set_reduction_limit(sp_ctx* ctx, uint64_t bytes) {
    if (!ctx) return set_error(SP_ERR_INVALID_ARG, "NULL context");
    RunLock lk(ctx);
    ctx->whole_pass_limit = bytes;
    return SP_OK;
}

int sp_ctx_set_iteration_budget(sp_ctx* ctx, double ms) {
    if (!ctx) return set_error(SP_ERR_INVALID_ARG, "NULL context");
    if (!(ms >= 0.0) || !std::isfinite(ms)) return set_error(SP_ERR_INVALID_ARG, "iteration budget must be finite and >= 0");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->iteration_budget_ms = ms;
    return SP_OK;
}

int sp_ctx_inject_failures(sp_ctx* ctx, uint32_t n_iterations) {
    if (!ctx) return set_error(SP_ERR_INVALID_ARG, "NULL context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->inject_failures = n_iterations;
    return SP_OK;
}

int sp_ctx_last_run_stats(sp_ctx* ctx, double* out) {
    if (!ctx || !out) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    for (int i = 0; i < 8; ++i) out[i] = ctx->last_stats[i];
    return SP_OK;
}

int sp_run(sp_ctx* ctx, const sp_request* req, const sp_particles* parts, sp_result* res) {
    if (!ctx || !res) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    HostRequest hr;
    if (int e = prepare_request(req, parts, hr)) return e;
    sp_tuning t;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        t = get_tuning(ctx, hr);
    }
    Outputs o;
    if (int e = run_plain(ctx, hr, o, (req->flags & SP_FLAG_KEEP_FINAL_STATE) != 0 || res->final_state, t)) return e;
    copy_result(hr, o, res);
    if (o.newton_failures) {
        set_error(SP_ERR_NOT_CONVERGED, "Implicit_Midpoint: Newton reached Max_Iterations in " +
                                            std::to_string(o.newton_failures) + " particle-steps");
    }
    return SP_OK;
}

int sp_prepare(sp_ctx* ctx, const sp_request* req, const sp_particles* parts, sp_prepared** out) {
    if (!ctx || !out) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    std::unique_ptr<sp_prepared> h(new sp_prepared());
    h->ctx = ctx;
    if (int e = prepare_request(req, parts, h->pr.hr)) return e;
    h->pr.keep_final = (req->flags & SP_FLAG_KEEP_FINAL_STATE) != 0;
    {
        RunLock lk(ctx);
        if (int e = prepare_plain(ctx, h->pr)) return e;
    }
    *out = h.release();
    return SP_OK;
}

int sp_execute(sp_prepared* h, sp_result* res) {
    if (!h || !res) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    sp_tuning t;
    {
        std::lock_guard<std::mutex> lk(h->ctx->mu);
        t = get_tuning(h->ctx, h->pr.hr);
    }
    h->pr.keep_final = h->pr.keep_final || res->final_state != nullptr;
    Outputs o;
    if (int e = execute_plain(h->ctx, h->pr, o, t)) return e;
    copy_result(h->pr.hr, o, res);
    return SP_OK;
}

int sp_prepared_inputs(const sp_prepared* h, double* params, double* initial_state) {
    if (!h) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    const HostRequest& hr = h->pr.hr;
    RunLock lk(h->ctx);
    SP_HIP(hipSetDevice(h->ctx->device));
    if (params) SP_HIP(hipMemcpy(params, h->pr.d.params, (size_t)NPARAM * hr.N * 8, hipMemcpyDeviceToHost));
    if (initial_state) {
        std::vector<double> soa((size_t)hr.dy * hr.N);
        SP_HIP(hipMemcpy(soa.data(), h->pr.state0, soa.size() * 8, hipMemcpyDeviceToHost));
        for (uint64_t p = 0; p < hr.N; ++p)
            for (uint32_t i = 0; i < hr.dy; ++i) initial_state[p * hr.dy + i] = soa[i * hr.N + p];
    }
    return SP_OK;
}

void sp_release(sp_prepared* h) {
    if (!h) return;
    {
        RunLock lk(h->ctx);
        hipSetDevice(h->ctx->device);
        free_prepared(h->pr);
    }
    delete h;
}

int sp_dev_philox(sp_ctx* ctx, const uint32_t ctr[4], const uint32_t key[2], uint64_t count, uint32_t* out) {
    if (!ctx || !ctr || !key || (!out && count)) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    RunLock lk(ctx);
    SP_HIP(hipSetDevice(ctx->device));
    uint32_t* d = nullptr;
    SP_HIP(hipMalloc(&d, std::max<uint64_t>(1, count) * 16));
    launch_philox_test(ctx->stream, ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], count, d);
    hipError_t e = hipMemcpyAsync(out, d, count * 16, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (e != hipSuccess) return set_error(SP_ERR_DEVICE, hipGetErrorString(e));
    return SP_OK;
}

int sp_dev_normals(sp_ctx* ctx, uint64_t seed, uint64_t particle, uint64_t first, uint64_t count, int precision, double* out) {
    if (!ctx || (!out && count)) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    RunLock lk(ctx);
    SP_HIP(hipSetDevice(ctx->device));
    double* d = nullptr;
    SP_HIP(hipMalloc(&d, std::max<uint64_t>(1, count) * 8));
    launch_normals_test(ctx->stream, seed, particle, first, count, precision == SP_PREC_MIXED, d);
    hipError_t e = hipMemcpyAsync(out, d, count * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (e != hipSuccess) return set_error(SP_ERR_DEVICE, hipGetErrorString(e));
    return SP_OK;
}

int sp_dev_box_muller(sp_ctx* ctx, const uint32_t* words, uint64_t n_blocks, int precision, double* out) {
    if (!ctx || (n_blocks && (!words || !out))) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    if (precision != SP_PREC_FP64 && precision != SP_PREC_MIXED) return set_error(SP_ERR_INVALID_ARG, "bad precision");
    RunLock lk(ctx);
    SP_HIP(hipSetDevice(ctx->device));
    const uint64_t nb = std::max<uint64_t>(1, n_blocks);
    uint32_t* dw = nullptr;
    double* d = nullptr;
    SP_HIP(hipMalloc(&dw, nb * 16));
    hipError_t e = hipMalloc(&d, nb * 32);
    if (e == hipSuccess) e = hipMemcpyAsync(dw, words, n_blocks * 16, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_box_muller_test(ctx->stream, dw, n_blocks, precision == SP_PREC_MIXED, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, n_blocks * 32, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dw);
    if (d) hipFree(d);
    if (e != hipSuccess) return set_error(SP_ERR_DEVICE, hipGetErrorString(e));
    return SP_OK;
}

int sp_host_box_muller(const uint32_t* words, uint64_t n_blocks, double* out) {
    if (n_blocks && (!words || !out)) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    for (uint64_t i = 0; i < n_blocks; ++i)
        box_muller4_f64(U4{words[4 * i], words[4 * i + 1], words[4 * i + 2], words[4 * i + 3]}, out + 4 * i);
    return SP_OK;
}

}  // extern "C"

extern "C" void sp_ctx_destroy(sp_ctx* ctx) {
    if (!ctx) return;
    stop_worker(ctx);
    hipSetDevice(ctx->device);
    destroy_res(ctx->res);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}
