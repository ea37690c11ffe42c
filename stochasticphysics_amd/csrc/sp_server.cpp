// sp_server.cpp -- batched execution and the continuous-batching request server.
//
// Both run the same iteration: every admitted request ("slot") advances by one
// chunk of steps; all 64-particle tiles of all slots of one kernel class form a
// work table that one persistent launch drains (waves pull tiles from an atomic
// counter, longest classes first), then one reduction launch folds every
// slot's per-wave partial sums into its output rows.  Particle state stays in
// HBM between chunks; Philox noise depends only on (seed, particle, step), so a
// request's result does not depend on what it shares the device with.
//
//  * sp_run_batch: the whole schedule of a static batch is planned up front and
//    captured in one hipGraph (chunk launches + reductions), launched once.
//  * sp_submit/sp_wait/sp_result_get: a worker thread admits requests by device
//    memory footprint as slots free up and runs iterations back to back; new
//    requests join at the next iteration boundary.
//  * mixed precision: each mixed request gets an FP64 shadow over its first
//    1/8 (512..4096 particles) on the same Philox stream.  Per output and
//    component, the difference of the two means over those particles and its
//    standard error (from the spread over the 64-particle tiles) are formed;
//    a request whose time-averaged difference exceeds check_tol by more than
//    its 99.9% confidence margin (or that holds a non-finite value) is re-run
//    in FP64.  Trajectories that pass a barrier or a pole can decorrelate
//    between FP32 and FP64 under the same noise, so the difference is partly
//    Monte-Carlo noise: the test asks for a bias beyond the tolerance, not for
//    trajectory agreement.
#include "sp_engine.h"

namespace sp {

// relative cost of one particle-step per kernel class (for longest-first ordering)
static double class_cost(const HostRequest& hr) {
    double c = hr.rq.problem == SP_PROBLEM_BROWN_AND_NEEL ? 3.0 : (hr.rq.problem == SP_PROBLEM_NEEL_SPHERICAL ? 1.6 : 1.0);
    if (hr.rq.problem == SP_PROBLEM_BROWN_AND_NEEL_EULER_SPHERICAL) c = 5.0;
    if (hr.rq.solver == SP_SOLVER_HEUN_STRONG) c *= 1.5;
    if (hr.rq.solver == SP_SOLVER_IMPLICIT_MIDPOINT) c *= 6.0;
    if (hr.mixed) c *= 0.6;
    return c;
}

struct IterPlan {
    std::vector<uint32_t> kinds;                      // one persistent launch per kind
    std::vector<uint32_t> kind_begin, kind_count;     // ranges into work (slot entries)
    std::vector<uint32_t> kind_tiles;                 // tiles per kind launch
    std::vector<double> kind_ps;                      // particle-steps per kind launch
    std::vector<SlotWork> work;
    std::vector<ReduceWork> reduce;
    uint32_t max_rows = 0;
    size_t partial_doubles = 0;
    std::vector<Slot*> finishing;                     // slots whose last chunk is in this iteration
    double particle_steps = 0.0;
};

// Plan one iteration for the given slots (indices into the DevRequest table = position in `slots`).
// Partial offsets are relative (pointer = base + offset, patched by the caller).
static void plan_iteration(const std::vector<Slot*>& slots, IterPlan& P) {
    P = IterPlan{};
    struct Item { double cost; uint32_t kind; uint32_t slot; };
    std::vector<Item> items;
    for (uint32_t i = 0; i < slots.size(); ++i) {
        Slot* s = slots[i];
        if (s->done) continue;
        items.push_back({class_cost(*s->hr) * (double)std::min<uint64_t>(s->chunk, s->hr->total_steps - s->next_step + 1), s->hr->kind, i});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.kind < b.kind || (a.kind == b.kind && a.cost > b.cost); });
    size_t off = 0;
    for (size_t k = 0; k < items.size(); ++k) {
        Slot* s = slots[items[k].slot];
        const HostRequest& hr = *s->hr;
        if (P.kinds.empty() || P.kinds.back() != hr.kind) {
            P.kinds.push_back(hr.kind);
            P.kind_begin.push_back((uint32_t)P.work.size());
            P.kind_count.push_back(0);
            P.kind_tiles.push_back(0);
            P.kind_ps.push_back(0.0);
        }
        const uint64_t OS = hr.rq.oversampling;
        const uint64_t sb = s->next_step;
        const uint64_t se = std::min(sb + s->chunk, hr.total_steps);
        const uint64_t first_window = sb == 0 ? 0 : sb / OS + 1;
        const uint64_t rows = (se - sb) / OS + (sb == 0 ? 1 : 0);
        SlotWork sw{};
        sw.step_begin = sb;
        sw.step_end = se;
        sw.partial = (double*)(uintptr_t)(off * 8);   // relative, patched later
        sw.req = items[k].slot;
        sw.tile_base = P.kind_tiles.back();
        sw.n_tiles = hr.n_tiles;
        P.work.push_back(sw);
        P.kind_count.back() += 1;
        P.kind_tiles.back() += hr.n_tiles;
        ReduceWork rw{};
        rw.partial = (const double*)(uintptr_t)(off * 8);
        rw.pstride = hr.n_tiles;
        rw.n_check = hr.n_check_tiles;
        rw.n_rows = (uint32_t)(rows * hr.dout);
        rw.row_base = (uint32_t)(first_window * hr.dout);
        rw.out = s->dev.out;
        rw.check_out = s->dev.check;
        P.reduce.push_back(rw);
        P.max_rows = std::max(P.max_rows, rw.n_rows);
        off += (size_t)rows * hr.dout * hr.n_tiles;
        P.particle_steps += (double)hr.N * (double)(se - sb);
        P.kind_ps.back() += (double)hr.N * (double)(se - sb);
        s->work_share = (double)hr.N * (double)(se - sb);
        s->next_step = se;
        if (se >= hr.total_steps) P.finishing.push_back(s);
    }
    P.partial_doubles = off;
}

// persistent grid for one kind, in work-groups of `block` threads
static dim3 persistent_grid(const sp_ctx* ctx, uint32_t n_work, const sp_tuning& t, uint32_t block) {
    const uint32_t wpb = std::max<uint32_t>(1, block / 64);
    const uint32_t per_simd = std::max<uint32_t>(1, t.persistent_waves_per_simd);
    const uint32_t resident = (uint32_t)ctx->prop.multiProcessorCount * 4u * per_simd;   // waves
    const uint32_t waves = std::min(n_work, resident);
    return dim3(std::max<uint32_t>(1, (waves + wpb - 1) / wpb));
}

// Issue the launches of one iteration whose tables are already on the device, on R's streams (s is R.stream, or the
// stream being captured).
// timing: record a timing-event pair around each class's launch (R.class_ev; the server's iterations, not captured
// in a graph)
static int record_iteration(sp_ctx* ctx, IterRes& R, hipStream_t s, const IterPlan& P, const DevRequest* d_reqs,
                            const SlotWork* d_work, const ReduceWork* d_reduce, uint32_t* d_counters,
                            const std::vector<sp_tuning>& tun, bool timing = false) {
    if (!P.kinds.empty()) launch_zero_u32(s, d_counters, (uint32_t)P.kinds.size());
    // fork: the classes of this iteration run side by side (the persistent launches of classes with few
    // waves, such as FP64 shadows, would otherwise serialise behind the others)
    const int n_side = (int)std::min<size_t>(P.kinds.size() > 0 ? P.kinds.size() - 1 : 0, (size_t)R.n_side);
    if (n_side > 0) {
        SP_HIP(hipEventRecord(R.fork_ev, s));
        for (int i = 0; i < n_side; ++i) SP_HIP(hipStreamWaitEvent(R.side[i], R.fork_ev, 0));
    }
    for (size_t k = 0; k < P.kinds.size(); ++k) {
        const sp_tuning& t = tun[k];
        const uint32_t block = launch_block(t.block_size, P.kind_tiles[k]);
        const dim3 grid = persistent_grid(ctx, P.kind_tiles[k], t, block);
        hipStream_t sk = (k == 0 || R.n_side == 0) ? s : R.side[(k - 1) % R.n_side];
        const bool timed = timing && k < (size_t)IterRes::N_CLASS_EV;
        if (timed) SP_HIP(hipEventRecord(R.class_ev[2 * k], sk));
        if (!launch_batch(P.kinds[k], grid, dim3(block), sk, d_reqs, d_work + P.kind_begin[k], P.kind_count[k], P.kind_tiles[k],
                          d_counters + k))
            return set_error(SP_ERR_UNSUPPORTED, "no batch kernel for a requested class");
        if (timed) SP_HIP(hipEventRecord(R.class_ev[2 * k + 1], sk));
    }
    for (int i = 0; i < n_side; ++i) {           // join
        SP_HIP(hipEventRecord(R.join_ev[i], R.side[i]));
        SP_HIP(hipStreamWaitEvent(s, R.join_ev[i], 0));
    }
    if (!P.reduce.empty()) launch_reduce(s, d_reduce, (uint32_t)P.reduce.size(), P.max_rows);
    return SP_OK;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Device tables of one iteration: [DevRequest x nreq][SlotWork x nwork][ReduceWork x nred]
struct TableLayout {
    size_t req_off, work_off, red_off, bytes;
};
static TableLayout layout_for(size_t nreq, size_t nwork, size_t nred) {
    TableLayout L;
    L.req_off = 0;
    L.work_off = align256(nreq * sizeof(DevRequest));
    L.red_off = L.work_off + align256(nwork * sizeof(SlotWork));
    L.bytes = L.red_off + align256(nred * sizeof(ReduceWork));
    return L;
}
static void fill_tables(char* host, const TableLayout& L, const std::vector<DevRequest>& reqs, const IterPlan& P,
                        double* partial_base) {
    std::memcpy(host + L.req_off, reqs.data(), reqs.size() * sizeof(DevRequest));
    SlotWork* w = (SlotWork*)(host + L.work_off);
    for (size_t i = 0; i < P.work.size(); ++i) {
        w[i] = P.work[i];
        w[i].partial = (double*)((char*)partial_base + (uintptr_t)P.work[i].partial);
    }
    ReduceWork* r = (ReduceWork*)(host + L.red_off);
    for (size_t i = 0; i < P.reduce.size(); ++i) {
        r[i] = P.reduce[i];
        r[i].partial = (const double*)((const char*)partial_base + (uintptr_t)P.reduce[i].partial);
    }
}

static std::vector<sp_tuning> tunings_for(sp_ctx* ctx, const IterPlan& P, const std::vector<Slot*>& slots) {
    std::vector<sp_tuning> t(P.kinds.size());
    for (size_t k = 0; k < P.kinds.size(); ++k) {
        const SlotWork& w = P.work[P.kind_begin[k]];
        t[k] = get_tuning(ctx, *slots[w.req]->hr);
    }
    return t;
}

static void make_shadow(const HostRequest& parent, HostRequest& sh) {
    sh = HostRequest{};
    sh.rq = parent.rq;
    sh.rq.precision = SP_PREC_FP64;
    sh.rq.flags |= SP_FLAG_SKIP_CHECK;
    sh.rq.save_interval = 0;                    // the shadow keeps no single-simulation outputs
    sh.n_saved = 0;
    sh.dy = parent.dy; sh.dw = parent.dw; sh.dout = parent.dout;
    sh.N = std::min<uint64_t>(parent.N, 64ull * parent.n_check_tiles);
    sh.rq.n_particles = sh.N;
    sh.n_out = parent.n_out;
    sh.total_steps = parent.total_steps;
    sh.n_tiles = (uint32_t)((sh.N + 63) / 64);
    sh.mixed = false;
    sh.kind = kind_for(parent.rq, false);
    if (parent.prov) {      // particle p is a function of (seed, p): the provider draws the same first particles
        auto spec = std::make_shared<ProvSpec>(*parent.prov);
        spec->base.N = sh.N;
        sh.prov = std::move(spec);
    } else {
        sh.params.resize((size_t)NPARAM * sh.N);
        for (uint32_t k = 0; k < NPARAM; ++k)
            for (uint64_t p = 0; p < sh.N; ++p) sh.params[k * sh.N + p] = parent.params[k * parent.N + p];
        sh.y0.resize((size_t)sh.dy * sh.N);
        for (uint32_t i = 0; i < sh.dy; ++i)
            for (uint64_t p = 0; p < sh.N; ++p) sh.y0[i * sh.N + p] = parent.y0[i * parent.N + p];
    }
    sh.n_check_tiles = sh.n_tiles;              // the same tiles, kept one by one
    for (uint32_t c = 0; c < sh.dout; ++c) { sh.wsum[c] = parent.wsum_check[c]; sh.wsum_check[c] = parent.wsum_check[c]; }
}

// Two-sided 99.9% quantiles of Student's t with dof = 1, 2, ... (dof > 120: the normal quantile).
static double t_quantile_999(uint32_t dof) {
    static const double q[] = {636.62, 31.599, 12.924, 8.610, 6.869, 5.959, 5.408, 5.041, 4.781, 4.587,
                               4.437, 4.318, 4.221, 4.140, 4.073, 4.015, 3.965, 3.922, 3.883, 3.850};
    if (dof == 0) return HUGE_VAL;
    if (dof <= 20) return q[dof - 1];
    if (dof <= 30) return 3.646;
    if (dof <= 40) return 3.551;
    if (dof <= 60) return 3.460;
    if (dof <= 120) return 3.373;
    return 3.291;
}

// Check of a mixed result against its FP64 shadow over the T check tiles.  Per output row k and
// component c, d_t = (mixed - fp64 weighted sum of tile t) / W and D = sum_t d_t (W = weight of the check
// particles); its standard error SE = sqrt(T/(T-1) sum_t (d_t - D/T)^2).  Reported: error = max |D|,
// noise = max 3 SE.  The test is on the time average: per component, the bias B = mean_k D over the
// rows k >= 1 and its standard error from the per-tile time averages; the result fails where
// |B| - t_999(T-1) SE(B) > tol (|B| > tol when T < 4), or where the mixed result is not finite (FP32
// overflow of a run-away particle).  A negative tol fails every check.
struct CheckResult { double error = 0.0, noise = 0.0, bias = 0.0, bias_se = 0.0; bool fail = false; };
static CheckResult check_of(const HostRequest& hr, const Outputs& mixed, const Outputs& shadow) {
    CheckResult r;
    const double tol = hr.rq.check_tol != 0.0 ? hr.rq.check_tol : 1e-3;
    for (double v : mixed.mean)
        if (!std::isfinite(v)) { r.error = HUGE_VAL; r.fail = true; return r; }
    const uint32_t T = hr.n_check_tiles;
    if (T == 0 || mixed.check_tiles.size() < (size_t)hr.n_out * hr.dout * T ||
        shadow.check_tiles.size() < (size_t)hr.n_out * hr.dout * T) { r.error = HUGE_VAL; r.fail = true; return r; }
    std::vector<double> d(T), avg(T);
    for (uint32_t c = 0; c < hr.dout; ++c) {
        const double W = hr.wsum_check[c];
        if (W == 0.0) continue;
        std::fill(avg.begin(), avg.end(), 0.0);
        for (uint64_t k = 0; k < hr.n_out; ++k) {
            const double* a = &mixed.check_tiles[(k * hr.dout + c) * T];
            const double* b = &shadow.check_tiles[(k * hr.dout + c) * T];
            double D = 0.0;
            for (uint32_t t = 0; t < T; ++t) { d[t] = (a[t] - b[t]) / W; D += d[t]; }
            double ss = 0.0;
            for (uint32_t t = 0; t < T; ++t) { const double x = d[t] - D / T; ss += x * x; }
            const double se = T > 1 ? std::sqrt(ss * T / (T - 1)) : 0.0;
            if (!std::isfinite(D) || !std::isfinite(se)) { r.error = HUGE_VAL; r.fail = true; return r; }
            r.error = std::max(r.error, std::fabs(D));
            r.noise = std::max(r.noise, 3.0 * se);
            if (k >= 1) for (uint32_t t = 0; t < T; ++t) avg[t] += d[t];
        }
        if (hr.n_out < 2) continue;
        double B = 0.0;
        for (uint32_t t = 0; t < T; ++t) { avg[t] /= (double)(hr.n_out - 1); B += avg[t]; }
        double ss = 0.0;
        for (uint32_t t = 0; t < T; ++t) { const double x = avg[t] - B / T; ss += x * x; }
        const double seB = T > 1 ? std::sqrt(ss * T / (T - 1)) : 0.0;
        r.bias = std::max(r.bias, std::fabs(B));
        r.bias_se = std::max(r.bias_se, seB);
        const double margin = T >= 4 ? t_quantile_999(T - 1) * seB : 0.0;
        if (std::fabs(B) - margin > tol || tol < 0.0) r.fail = true;
    }
    return r;
}

static uint64_t chunk_for(const HostRequest& hr, const sp_tuning& t) {
    const uint64_t align = lcm64(hr.rq.oversampling, 4);
    const uint64_t c = std::max<uint64_t>(align, (uint64_t)std::max<uint32_t>(t.chunk_steps, 1) / align * align);
    return c;
}

static bool wants_shadow(const HostRequest& hr) { return hr.mixed && !(hr.rq.flags & SP_FLAG_SKIP_CHECK); }

// Chunks of the next server iteration.  Every slot starts from its class's tuned chunk; with a budget, a class whose
// launch would take longer than the budget at that chunk (by its measured cost per particle-step times the
// particles of its slots) gets the largest aligned chunk that fits, at least one aligned block of steps.  A class
// not measured yet keeps the tuned chunk.  Each launch then ends within about the budget, so requests of cheap
// classes are not held in long iterations by an expensive one (such as Implicit_Midpoint beside Heun_Strong).
static void assign_chunks(const sp_ctx* ctx, double budget_ms, const std::vector<Slot*>& slots) {
    std::map<uint32_t, double> n_of_kind;
    for (Slot* s : slots)
        if (!s->done) n_of_kind[s->hr->kind] += (double)s->hr->N;
    for (Slot* s : slots) {
        if (s->done) continue;
        s->chunk = s->tuned_chunk;
        if (budget_ms <= 0.0) continue;
        const auto it = ctx->class_ns_per_ps.find(s->hr->kind);
        if (it == ctx->class_ns_per_ps.end() || !(it->second > 0.0)) continue;
        const double fit = budget_ms * 1e6 / (it->second * n_of_kind[s->hr->kind]);
        if (fit >= (double)s->tuned_chunk) continue;
        const uint64_t align = lcm64(s->hr->rq.oversampling, 4);
        s->chunk = std::max<uint64_t>(align, (uint64_t)fit / align * align);
    }
}

// smoothed device cost per particle-step of each class timed in the iteration just finished (on R's events)
static void update_class_costs(sp_ctx* ctx, const IterRes& R, const IterPlan& P) {
    const size_t n = std::min<size_t>(P.kinds.size(), (size_t)IterRes::N_CLASS_EV);
    for (size_t k = 0; k < n; ++k) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, R.class_ev[2 * k], R.class_ev[2 * k + 1]) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        if (!(P.kind_ps[k] > 0.0) || !(ms > 0.f)) continue;
        const double ns = (double)ms * 1e6 / P.kind_ps[k];
        auto it = ctx->class_ns_per_ps.find(P.kinds[k]);
        if (it == ctx->class_ns_per_ps.end()) ctx->class_ns_per_ps[P.kinds[k]] = ns;
        else it->second = 0.5 * it->second + 0.5 * ns;
    }
}

// ------------------------------------------------------------------ static batch (graph)
static int run_batch_static(sp_ctx* ctx, std::vector<HostRequest>& reqs, std::vector<Outputs>& outs, bool keep_final, bool use_graph) {
    RunLock lk(ctx);
    hipStream_t s = ctx->stream;
    SP_HIP(hipSetDevice(ctx->device));
    const size_t n = reqs.size();
    std::vector<std::unique_ptr<Slot>> slots;
    struct Cleanup { std::vector<std::unique_ptr<Slot>>* v; ~Cleanup() { for (auto& x : *v) free_dev(x->dev); } } cleanup{&slots};
    std::vector<Slot*> order;
    for (size_t i = 0; i < n; ++i) {
        auto sl = std::make_unique<Slot>();
        sl->hr = &reqs[i];
        if (int e = alloc_dev(reqs[i], sl->dev, s)) return e;
        sl->chunk = chunk_for(reqs[i], get_tuning(ctx, reqs[i]));
        slots.push_back(std::move(sl));
    }
    for (size_t i = 0; i < n; ++i) {
        if (!wants_shadow(reqs[i])) continue;
        auto sh = std::make_unique<Slot>();
        make_shadow(reqs[i], sh->shadow_req);
        sh->hr = &sh->shadow_req;
        sh->is_shadow = true;
        sh->parent = slots[i].get();
        slots[i]->shadow = sh.get();
        if (int e = alloc_dev(sh->shadow_req, sh->dev, s)) return e;
        sh->chunk = slots[i]->chunk;
        slots.push_back(std::move(sh));
    }
    for (auto& x : slots) order.push_back(x.get());
    std::vector<DevRequest> dreqs;
    for (Slot* x : order) dreqs.push_back(make_dev_request(*x->hr, x->dev));
    // plan every iteration
    std::vector<IterPlan> plans;
    for (;;) {
        IterPlan P;
        plan_iteration(order, P);
        if (P.work.empty()) break;
        for (Slot* f : P.finishing) f->done = true;
        plans.push_back(std::move(P));
    }
    size_t max_partial = 0;
    std::vector<TableLayout> lays;
    size_t table_total = 0;
    for (auto& P : plans) {
        max_partial = std::max(max_partial, P.partial_doubles);
        lays.push_back(layout_for(dreqs.size(), P.work.size(), P.reduce.size()));
        table_total += lays.back().bytes;
    }
    if (int e = ensure_partial(ctx, std::max<size_t>(8, max_partial * 8))) return e;
    if (int e = ensure_table(ctx, std::max<size_t>(256, table_total))) return e;
    size_t max_kinds = 1;
    for (auto& P : plans) max_kinds = std::max(max_kinds, P.kinds.size());
    if (int e = ensure_counters(ctx, max_kinds)) return e;
    size_t off = 0;
    std::vector<size_t> offs;
    for (size_t i = 0; i < plans.size(); ++i) {
        fill_tables((char*)ctx->res.pinned + off, lays[i], dreqs, plans[i], ctx->res.partial);
        offs.push_back(off);
        off += lays[i].bytes;
    }
    if (off) SP_HIP(hipMemcpyAsync(ctx->res.table, ctx->res.pinned, off, hipMemcpyHostToDevice, s));
    std::vector<std::vector<sp_tuning>> tuns;
    for (auto& P : plans) tuns.push_back(tunings_for(ctx, P, order));
    auto record_all = [&](hipStream_t st) -> int {
        for (size_t i = 0; i < plans.size(); ++i) {
            char* base = (char*)ctx->res.table + offs[i];
            if (int e = record_iteration(ctx, ctx->res, st, plans[i], (const DevRequest*)(base + lays[i].req_off),
                                         (const SlotWork*)(base + lays[i].work_off), (const ReduceWork*)(base + lays[i].red_off),
                                         ctx->res.counters, tuns[i]))
                return e;
        }
        return SP_OK;
    };
    hipEvent_t e0, e1;
    SP_HIP(hipEventCreate(&e0));
    SP_HIP(hipEventCreate(&e1));
    const auto t0 = Clock::now();
    SP_HIP(hipStreamSynchronize(s));
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    if (use_graph && !plans.empty()) {
        SP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = record_all(s);
        hipError_t ce = hipStreamEndCapture(s, &graph);
        if (rc) { if (graph) hipGraphDestroy(graph); return rc; }
        SP_HIP(ce);
        SP_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        SP_HIP(hipEventRecord(e0, s));
        SP_HIP(hipGraphLaunch(exec, s));
        SP_HIP(hipEventRecord(e1, s));
    } else {
        SP_HIP(hipEventRecord(e0, s));
        if (int e = record_all(s)) return e;
        SP_HIP(hipEventRecord(e1, s));
    }
    SP_HIP(hipStreamSynchronize(s));
    float dev_ms = 0.f;
    SP_HIP(hipEventElapsedTime(&dev_ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    const double wall = ms_since(t0, Clock::now());
    ctx->last_stats[0] = (double)plans.size();
    ctx->last_stats[1] = dev_ms;
    ctx->last_stats[2] = 0.0;
    ctx->last_stats[3] = wall;
    double ps = 0.0;
    for (auto& P : plans) ps += P.particle_steps;
    ctx->last_stats[4] = ps;
    outs.assign(n, Outputs{});
    double total_ps = 0.0;
    for (size_t i = 0; i < n; ++i) total_ps += (double)reqs[i].N * (double)reqs[i].total_steps;
    for (size_t i = 0; i < n; ++i) {
        Slot* sl = slots[i].get();
        if (int e = download_outputs(reqs[i], sl->dev, s, keep_final, sl->shadow != nullptr, outs[i])) return e;
        outs[i].device_ms = dev_ms * ((double)reqs[i].N * (double)reqs[i].total_steps) / std::max(1.0, total_ps);
        outs[i].latency_ms = wall;
        if (sl->shadow) {
            Outputs so;
            if (int e = download_outputs(*sl->shadow->hr, sl->shadow->dev, s, false, true, so)) return e;
            const CheckResult cr = check_of(reqs[i], outs[i], so);
            outs[i].check_error = cr.error;
            outs[i].check_noise = cr.noise;
            outs[i].check_bias = cr.bias;
            outs[i].check_bias_se = cr.bias_se;
            outs[i].check_failed = cr.fail;
        }
    }
    return SP_OK;
}

// ------------------------------------------------------------------ server worker
static void finish_job(sp_ctx* ctx, Slot* sl, hipStream_t s) {
    // called by the worker without ctx->mu (s: the stream of the slot's lane); publishes under the lock.  A request
    // that has already ended (failed, cancelled) is never published again.
    std::shared_ptr<Job> job = sl->job;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (job->status != SP_ERR_NOT_READY) return;
    }
    Outputs o;
    int rc = download_outputs(*sl->hr, sl->dev, s, job->keep_final, sl->shadow != nullptr, o);
    bool requeue = false;
    if (rc == SP_OK && sl->shadow) {
        Outputs so;
        rc = download_outputs(*sl->shadow->hr, sl->shadow->dev, s, false, true, so);
        if (rc == SP_OK) {
            const CheckResult cr = check_of(*sl->hr, o, so);
            o.check_error = cr.error;
            o.check_noise = cr.noise;
            o.check_bias = cr.bias;
            o.check_bias_se = cr.bias_se;
            if (cr.fail && !(sl->hr->rq.flags & SP_FLAG_NO_FALLBACK)) requeue = true;
        }
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (requeue) {
        // per-request FP64 fallback: run the request again in FP64 ahead of the queue
        job->out.check_error = o.check_error;
        job->out.check_noise = o.check_noise;
        job->out.check_bias = o.check_bias;
        job->out.check_bias_se = o.check_bias_se;
        job->out.used_fallback = 1;
        job->hr.rq.precision = SP_PREC_FP64;
        job->hr.mixed = false;
        job->hr.kind = kind_for(job->hr.rq, false);
        ctx->pending.push_front(job);
        ctx->stat_fallbacks += 1;
        return;
    }
    if (rc == SP_OK) {
        const Outputs prev = job->out;
        job->out = std::move(o);
        if (prev.used_fallback) {
            job->out.used_fallback = 1;
            job->out.check_error = prev.check_error;
            job->out.check_noise = prev.check_noise;
            job->out.check_bias = prev.check_bias;
            job->out.check_bias_se = prev.check_bias_se;
        }
        job->out.device_ms = job->device_ms_acc;
        job->out.latency_ms = ms_since(job->t_submit, Clock::now());
        job->status = SP_OK;
        ctx->stat_completed += 1;
        ctx->stat_particle_steps += job->out.particle_steps;
    } else {
        job->status = rc;
    }
    ctx->done_cv.notify_all();
}

// The worker runs two lanes of iterations: Implicit_Midpoint requests (a Newton solve per step, several times the
// device time of an explicit step) iterate on lane 1 with their own streams and scratch, every other request
// (explicit solvers, mixed requests with their FP64 shadows) on lane 0.  The lanes' iterations overlap on the device
// and each lane starts its next iteration as soon as its previous one ends, so requests of the cheap classes are not
// held at the iteration boundaries of the expensive one.  A request's result does not depend on its lane or its
// chunking.
static int lane_of(const HostRequest& hr) { return hr.rq.solver == SP_SOLVER_IMPLICIT_MIDPOINT ? 1 : 0; }

struct Lane {
    IterRes* R = nullptr;
    std::vector<std::unique_ptr<Slot>> active;
    bool inflight = false;
    IterPlan P;
    std::vector<Slot*> slots;         // the in-flight iteration's slots (their DevRequest table positions)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BlockCache* cache = nullptr;      // where finished slots' blocks go
    bool has_work() const {
        for (auto& a : active) if (!a->done) return true;
        return false;
    }
};

// plan and issue one iteration of lane L (no wait); returns the error of the issue (no launch is then in flight)
static int issue_iteration(sp_ctx* ctx, Lane& L) {
    hipStream_t s = L.R->stream;
    L.slots.clear();
    for (auto& a : L.active) L.slots.push_back(a.get());
    double budget_ms = 0.0;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        budget_ms = ctx->iteration_budget_ms;
    }
    assign_chunks(ctx, L.R->class_timing ? budget_ms : 0.0, L.slots);
    plan_iteration(L.slots, L.P);
    if (L.P.work.empty()) return SP_OK;
    std::vector<DevRequest> dreqs;
    for (Slot* x : L.slots) dreqs.push_back(make_dev_request(*x->hr, x->dev));
    const TableLayout T = layout_for(dreqs.size(), L.P.work.size(), L.P.reduce.size());
    std::vector<sp_tuning> tun;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        tun = tunings_for(ctx, L.P, L.slots);
        if (ctx->inject_failures > 0) {
            --ctx->inject_failures;
            return set_error(SP_ERR_DEVICE, "iteration failed (injected)");
        }
    }
    if (int e = ensure_partial(*L.R, std::max<size_t>(8, L.P.partial_doubles * 8))) return e;
    if (int e = ensure_table(*L.R, T.bytes)) return e;
    if (int e = ensure_counters(*L.R, std::max<size_t>(1, L.P.kinds.size()))) return e;
    fill_tables((char*)L.R->pinned, T, dreqs, L.P, L.R->partial);
    SP_HIP(hipMemcpyAsync(L.R->table, L.R->pinned, T.bytes, hipMemcpyHostToDevice, s));
    SP_HIP(hipEventRecord(L.e0, s));
    char* base = (char*)L.R->table;
    if (int e = record_iteration(ctx, *L.R, s, L.P, (const DevRequest*)(base + T.req_off), (const SlotWork*)(base + T.work_off),
                                 (const ReduceWork*)(base + T.red_off), L.R->counters, tun, L.R->class_timing))
        return e;
    SP_HIP(hipEventRecord(L.e1, s));
    L.inflight = true;
    return SP_OK;
}

static void release_done(Lane& L);

// bookkeeping after lane L's iteration ended with status rc (the device work of a successful one has completed)
static void complete_iteration(sp_ctx* ctx, Lane& L, int rc) {
    float ms = 0.f;
    if (rc == SP_OK && L.inflight) {
        if (hipEventElapsedTime(&ms, L.e0, L.e1) != hipSuccess) { (void)hipGetLastError(); ms = 0.f; }
        if (L.R->class_timing) update_class_costs(ctx, *L.R, L.P);
    }
    L.inflight = false;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->stat_iterations += 1;
        ctx->stat_busy_ms += ms;
    }
    for (Slot* x : L.slots) {     // device time share of this iteration by particle-steps
        if (x->is_shadow || !x->job || L.P.particle_steps <= 0.0) continue;
        x->job->device_ms_acc += ms * x->work_share / L.P.particle_steps;
        x->work_share = 0.0;
    }
    if (rc == SP_OK) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (Slot* x : L.slots)
            if (!x->is_shadow && x->job) x->job->steps_done = x->next_step;
    }
    if (rc != SP_OK) {
        // every slot of the failed iteration ends, FP64 shadows included: a shadow left running would
        // later publish its parent's partial sums
        std::lock_guard<std::mutex> lk(ctx->mu);
        for (Slot* x : L.slots) {
            if (x->done) continue;
            if (x->job && x->job->status == SP_ERR_NOT_READY) x->job->status = rc;
            x->done = true;
        }
        ctx->done_cv.notify_all();
    } else {
        const auto& F = L.P.finishing;
        for (Slot* f : F) f->done = true;
        for (Slot* f : F)
            if (!f->is_shadow && f->job && (!f->shadow || f->shadow->done)) finish_job(ctx, f, L.R->stream);
        for (Slot* f : F)
            if (f->is_shadow && f->parent && f->parent->done && f->parent->job &&
                std::find(F.begin(), F.end(), f->parent) == F.end())
                finish_job(ctx, f->parent, L.R->stream);
    }
    L.slots.clear();
    release_done(L);
}

// drain every stream of R (after an issue that failed part-way, launches may be queued on any of them)
static void drain(IterRes& R) {
    if (hipStreamSynchronize(R.stream) != hipSuccess) (void)hipGetLastError();
    for (int i = 0; i < IterRes::N_SIDE; ++i)
        if (R.side[i] && hipStreamSynchronize(R.side[i]) != hipSuccess) (void)hipGetLastError();
}

static void release_done(Lane& L) {
    // release finished slots (parent and shadow together: a mixed request and its shadow share a lane)
    std::vector<std::unique_ptr<Slot>> keep;
    for (auto& a : L.active) {
        const bool done_pair = a->done && (a->is_shadow ? (a->parent == nullptr || a->parent->done) : (!a->shadow || a->shadow->done));
        if (done_pair) free_dev(a->dev, L.cache);      // no work in flight uses a finished slot
        else keep.push_back(std::move(a));
    }
    // parents point at shadows: clear dangling links of the survivors
    for (auto& a : keep) {
        bool has = false;
        for (auto& b : keep) if (b.get() == a->shadow || b.get() == a->parent) has = true;
        if (!has) { a->shadow = nullptr; a->parent = nullptr; }
    }
    L.active.swap(keep);
}

static void worker_main(sp_ctx* ctx) {
    hipSetDevice(ctx->device);
    IterRes own;                                    // lane 1's streams and scratch
    const bool two_lanes = init_res(own, nullptr, 0) == SP_OK;
    if (!two_lanes) { (void)hipGetLastError(); destroy_res(own); }
    Lane lanes[2];
    lanes[0].R = &ctx->res;
    lanes[1].R = two_lanes ? &own : &ctx->res;
    lanes[0].cache = lanes[1].cache = &ctx->blocks;
    const int n_lanes = two_lanes ? 2 : 1;
    for (int l = 0; l < n_lanes; ++l) {
        init_class_timing(*lanes[l].R);             // per-class launch timing for the iteration budget
        hipEventCreate(&lanes[l].e0);
        hipEventCreate(&lanes[l].e1);
    }
    auto any_inflight = [&] { return lanes[0].inflight || lanes[1].inflight; };
    size_t pending_left = 0;                        // requests left queued by the last admission pass
    std::unique_lock<std::mutex> device_lock(ctx->run_mu, std::defer_lock);
    for (;;) {
        std::vector<std::shared_ptr<Job>> admitted;
        std::vector<size_t> admitted_need;
        {
            std::unique_lock<std::mutex> lk(ctx->mu);
            if (!any_inflight())
                ctx->cv.wait(lk, [&] { return ctx->stop || !ctx->pending.empty() || lanes[0].has_work() || lanes[1].has_work(); });
            if (ctx->stop) {
                for (auto& j : ctx->pending) { j->status = SP_ERR_CANCELLED; }
                ctx->pending.clear();
                for (auto& L : lanes)
                    for (auto& a : L.active) if (!a->is_shadow && a->job) a->job->status = SP_ERR_CANCELLED;
                ctx->done_cv.notify_all();
                break;
            }
            // drop cancelled active requests (the active lists are the worker's own)
            for (auto& L : lanes)
                for (auto& a : L.active) {
                    if (!a->is_shadow && a->job && a->job->cancel && !a->done) {
                        a->done = true;
                        if (a->shadow) a->shadow->done = true;
                        a->job->status = SP_ERR_CANCELLED;
                        ctx->done_cv.notify_all();
                    }
                }
            // admission by device-memory footprint (state + parameters + outputs, shadow included)
            while (!ctx->pending.empty()) {
                auto& j = ctx->pending.front();
                if (j->cancel) { j->status = SP_ERR_CANCELLED; ctx->pending.pop_front(); ctx->done_cv.notify_all(); continue; }
                size_t need = request_bytes(j->hr) +
                              (wants_shadow(j->hr) ? request_bytes(j->hr) / std::max<uint64_t>(1, j->hr.n_tiles) * j->hr.n_check_tiles * 2 : 0);
                size_t n_active = admitted.size();
                for (auto& L : lanes)
                    for (auto& a : L.active) if (!a->is_shadow && !a->done) ++n_active;
                if (n_active > 0 && (ctx->mem_used + need > ctx->mem_budget || n_active >= ctx->max_active)) break;
                admitted.push_back(j);
                admitted_need.push_back(need);
                ctx->mem_used += need;
                ctx->pending.pop_front();
            }
            pending_left = ctx->pending.size();
        }
        if (!device_lock.owns_lock()) device_lock.lock();   // the worker owns the streams and scratch from here
        for (size_t ai = 0; ai < admitted.size(); ++ai) {
            auto& j = admitted[ai];
            Lane& L = lanes[two_lanes ? lane_of(j->hr) : 0];
            hipStream_t s = L.R->stream;
            auto sl = std::make_unique<Slot>();
            sl->job = j;
            sl->hr = &j->hr;
            int rc = alloc_dev(j->hr, sl->dev, s, &ctx->blocks);
            sp_tuning t;
            {
                std::lock_guard<std::mutex> lk(ctx->mu);
                t = get_tuning(ctx, j->hr);
            }
            sl->tuned_chunk = sl->chunk = chunk_for(j->hr, t);
            if (rc == SP_OK && wants_shadow(j->hr)) {
                auto sh = std::make_unique<Slot>();
                make_shadow(j->hr, sh->shadow_req);
                sh->hr = &sh->shadow_req;
                sh->is_shadow = true;
                sh->parent = sl.get();
                sh->tuned_chunk = sh->chunk = sl->chunk;
                rc = alloc_dev(sh->shadow_req, sh->dev, s, &ctx->blocks);
                sl->shadow = sh.get();
                if (rc == SP_OK) L.active.push_back(std::move(sh));
            }
            if (rc != SP_OK) {
                std::lock_guard<std::mutex> lk(ctx->mu);
                j->status = rc;
                ctx->mem_used -= std::min<uint64_t>(ctx->mem_used, admitted_need[ai]);   // never admitted
                ctx->done_cv.notify_all();
                free_dev(sl->dev);
                continue;
            }
            L.active.push_back(std::move(sl));
        }
        // a synchronous call waiting for the device gets it once the lanes in flight have ended
        const bool yield_device = ctx->run_waiters.load() > 0;
        if (!yield_device) {
            for (int l = 0; l < n_lanes; ++l) {
                Lane& L = lanes[l];
                if (L.inflight || !L.has_work()) continue;
                const int rc = issue_iteration(ctx, L);
                if (rc != SP_OK) {
                    drain(*L.R);                    // launches issued before the failure end first
                    L.inflight = false;
                    complete_iteration(ctx, L, rc);
                }
            }
        }
        // cancelled or finished slots of the lanes not in flight
        for (int l = 0; l < n_lanes; ++l)
            if (!lanes[l].inflight) release_done(lanes[l]);
        // wait until a lane's iteration ends, or a request that arrived since the admission above can start an idle lane
        while (any_inflight()) {
            bool ended = false;
            for (int l = 0; l < n_lanes; ++l) {
                Lane& L = lanes[l];
                if (!L.inflight) continue;
                const hipError_t q = hipEventQuery(L.e1);
                if (q == hipErrorNotReady) continue;
                int rc = SP_OK;
                if (q != hipSuccess || hipStreamSynchronize(L.R->stream) != hipSuccess) {
                    (void)hipGetLastError();
                    drain(*L.R);
                    rc = set_error(SP_ERR_DEVICE, "server iteration failed on the device");
                }
                complete_iteration(ctx, L, rc);
                ended = true;
            }
            if (ended) break;
            {
                std::lock_guard<std::mutex> lk(ctx->mu);
                bool idle_lane = false;
                for (int l = 0; l < n_lanes; ++l) idle_lane = idle_lane || !lanes[l].inflight;
                if (ctx->stop || (ctx->pending.size() > pending_left && idle_lane)) break;
                bool cancelled = false;
                for (auto& L : lanes)
                    for (auto& a : L.active) cancelled = cancelled || (!a->is_shadow && a->job && a->job->cancel && !a->done);
                if (cancelled) break;
            }
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        if (!any_inflight()) {
            device_lock.unlock();
            while (ctx->run_waiters.load() > 0) {       // let the waiting synchronous calls take the device first
                {
                    std::lock_guard<std::mutex> lk(ctx->mu);
                    if (ctx->stop) break;
                }
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
        }
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            size_t acc = 0, n_act = 0;
            for (auto& L : lanes)
                for (auto& a : L.active) { acc += a->dev.bytes; if (!a->is_shadow && !a->done) ++n_act; }
            ctx->mem_used = acc;
            ctx->stat_active = (double)n_act;
        }
    }
    if (!device_lock.owns_lock()) device_lock.lock();
    for (auto& L : lanes) {
        if (L.inflight) hipStreamSynchronize(L.R->stream);
        L.inflight = false;
        for (auto& a : L.active) free_dev(a->dev);
        L.active.clear();
    }
    ctx->blocks.clear();
    for (int l = 0; l < n_lanes; ++l) {
        if (lanes[l].e0) hipEventDestroy(lanes[l].e0);
        if (lanes[l].e1) hipEventDestroy(lanes[l].e1);
    }
    if (two_lanes) destroy_res(own);
}

static void ensure_worker(sp_ctx* ctx) {
    if (!ctx->worker_started) {
        ctx->worker_started = true;
        ctx->t_start = Clock::now();
        ctx->worker = std::thread(worker_main, ctx);
    }
}

void stop_worker(sp_ctx* ctx) {
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->stop = true;
    }
    ctx->cv.notify_all();
    if (ctx->worker.joinable()) ctx->worker.join();
}

}  // namespace sp

using namespace sp;

extern "C" {

int sp_run_batch(sp_ctx* ctx, uint32_t n_req, const sp_request* reqs, const sp_particles* parts, sp_result* res) {
    if (!ctx || (!reqs && n_req) || (!parts && n_req) || (!res && n_req)) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    std::vector<HostRequest> hrs(n_req);
    bool keep_final = false, use_graph = true;
    for (uint32_t i = 0; i < n_req; ++i) {
        if (int e = prepare_request(&reqs[i], &parts[i], hrs[i])) return e;
        keep_final = keep_final || res[i].final_state != nullptr || (reqs[i].flags & SP_FLAG_KEEP_FINAL_STATE);
        std::lock_guard<std::mutex> lk(ctx->mu);
        use_graph = use_graph && get_tuning(ctx, hrs[i]).use_graph;
    }
    std::vector<Outputs> outs;
    if (int e = run_batch_static(ctx, hrs, outs, keep_final, use_graph)) return e;
    // per-request FP64 fallback for mixed results that failed the shadow check
    for (uint32_t i = 0; i < n_req; ++i) {
        if (outs[i].check_failed && !(hrs[i].rq.flags & SP_FLAG_NO_FALLBACK)) {
            HostRequest h2;
            sp_request r2 = reqs[i];
            r2.precision = SP_PREC_FP64;
            if (int e = prepare_request(&r2, &parts[i], h2)) return e;
            Outputs o2;
            sp_tuning t;
            {
                std::lock_guard<std::mutex> lk(ctx->mu);
                t = get_tuning(ctx, h2);
            }
            if (int e = run_plain(ctx, h2, o2, keep_final, t)) return e;
            o2.check_error = outs[i].check_error;
            o2.check_noise = outs[i].check_noise;
            o2.check_bias = outs[i].check_bias;
            o2.check_bias_se = outs[i].check_bias_se;
            o2.used_fallback = 1;
            outs[i] = std::move(o2);
        }
        copy_result(hrs[i], outs[i], &res[i]);
    }
    return SP_OK;
}

int sp_submit(sp_ctx* ctx, const sp_request* req, const sp_particles* parts, uint64_t* id) {
    if (!ctx || !id) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    HostRequest hr;
    if (int e = prepare_request(req, parts, hr)) return e;
    return submit_request(ctx, std::move(hr), (req->flags & SP_FLAG_KEEP_FINAL_STATE) != 0, id);
}

}  // extern "C"

namespace sp {
int submit_request(sp_ctx* ctx, HostRequest&& hr, bool keep_final, uint64_t* id) {
    auto job = std::make_shared<Job>();
    job->hr = std::move(hr);
    job->keep_final = keep_final;
    job->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->stop) return set_error(SP_ERR_CANCELLED, "context is shutting down");
        job->id = ctx->next_id++;
        ctx->jobs[job->id] = job;
        ctx->pending.push_back(job);
        ensure_worker(ctx);
    }
    ctx->cv.notify_all();
    *id = job->id;
    return SP_OK;
}
}  // namespace sp

extern "C" {

int sp_wait(sp_ctx* ctx, uint64_t id, int timeout_ms) {
    if (!ctx) return set_error(SP_ERR_INVALID_ARG, "NULL context");
    std::unique_lock<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(id);
    if (it == ctx->jobs.end()) return set_error(SP_ERR_NOT_FOUND, "unknown request id");
    std::shared_ptr<Job> job = it->second;
    auto ready = [&] { return job->status != SP_ERR_NOT_READY; };
    if (timeout_ms < 0) ctx->done_cv.wait(lk, ready);
    else if (!ctx->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return set_error(SP_ERR_TIMEOUT, "request not finished");
    if (job->status != SP_OK) return set_error(job->status, "request failed or was cancelled");
    return SP_OK;
}

int sp_result_get(sp_ctx* ctx, uint64_t id, sp_result* res) {
    if (!ctx || !res) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(id);
    if (it == ctx->jobs.end()) return set_error(SP_ERR_NOT_FOUND, "unknown request id");
    std::shared_ptr<Job> job = it->second;
    if (job->status == SP_ERR_NOT_READY) return set_error(SP_ERR_NOT_READY, "request not finished");
    if (job->status != SP_OK) {
        const int st = job->status;
        ctx->jobs.erase(it);
        return set_error(st, "request failed or was cancelled");
    }
    copy_result(job->hr, job->out, res);
    ctx->jobs.erase(it);
    return SP_OK;
}

int sp_cancel(sp_ctx* ctx, uint64_t id) {
    if (!ctx) return set_error(SP_ERR_INVALID_ARG, "NULL context");
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->jobs.find(id);
        if (it == ctx->jobs.end()) return set_error(SP_ERR_NOT_FOUND, "unknown request id");
        it->second->cancel = true;
    }
    ctx->cv.notify_all();
    return SP_OK;
}

int sp_progress(sp_ctx* ctx, uint64_t id, double* fraction) {
    if (!ctx || !fraction) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(id);
    if (it == ctx->jobs.end()) return set_error(SP_ERR_NOT_FOUND, "unknown request id");
    const Job& j = *it->second;
    if (j.status == SP_OK) *fraction = 1.0;
    else if (j.hr.total_steps == 0) *fraction = 0.0;
    else *fraction = (double)std::min(j.steps_done, j.hr.total_steps) / (double)j.hr.total_steps;
    return SP_OK;
}

int sp_server_stats(sp_ctx* ctx, double* out) {
    if (!ctx || !out) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    out[0] = ctx->stat_completed;
    out[1] = ctx->stat_active;
    out[2] = (double)ctx->pending.size();
    out[3] = ctx->stat_iterations;
    out[4] = ctx->stat_busy_ms;
    out[5] = ctx->worker_started ? ms_since(ctx->t_start, Clock::now()) : 0.0;
    out[6] = ctx->stat_particle_steps;
    out[7] = ctx->stat_fallbacks;
    return SP_OK;
}

// Autotuner: times the batch path on a shortened copy of the request for each
// candidate (work-group size x persistent waves per SIMD x chunk length), times
// the four fastest again on a longer copy, and stores the fastest for the
// request's (problem, solver, precision) class.  Rates are device rates (the
// launches between two events).
int sp_autotune(sp_ctx* ctx, const sp_request* req, const sp_particles* parts, sp_tuning* best, double* best_rate) {
    if (!ctx || !req || !parts || !best) return set_error(SP_ERR_INVALID_ARG, "NULL argument");
    HostRequest base;
    if (int e = prepare_request(req, parts, base)) return e;
    // shorten to at most ~2^31 particle-steps per trial
    sp_request r = *req;
    const uint64_t OS = r.oversampling;
    const uint64_t max_steps = std::max<uint64_t>(8 * OS, (uint64_t)(2e9 / (double)base.N) / OS * OS);
    if (r.n_steps > max_steps) r.n_steps = max_steps;
    r.flags |= SP_FLAG_SKIP_CHECK;
    HostRequest probe;
    if (int e = prepare_request(&r, parts, probe)) return e;
    const int prec = probe.mixed ? SP_PREC_MIXED : SP_PREC_FP64;
    const uint32_t blocks[] = {64, 128, 256};          // the kernels' __launch_bounds__ is MAX_BLOCK = 256
    const uint32_t waves[] = {1, 2, 3, 4};   // the kernel classes are built for 1 to 4 waves per SIMD
    const uint32_t chunks[] = {2000, 8000, 32000};
    double best_r = -1.0;
    sp_tuning bt{};
    int err = SP_OK;
    const uint64_t key = tuning_key(r.problem, r.solver, prec);
    bool had_prev = false;
    sp_tuning prev{};
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->tuning.find(key);
        if (it != ctx->tuning.end()) { had_prev = true; prev = it->second; }
    }
    // the device rate of one run of request q (the event-timed launches of run_batch_static: allocation, uploads,
    // graph instantiation and downloads excluded, which a trial of tens of milliseconds would otherwise be noisy with)
    auto trial = [&](const sp_tuning& t, const sp_request& q, double ps) -> double {
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            ctx->tuning[tuning_key(r.problem, r.solver, prec)] = t;
        }
        std::vector<HostRequest> one(1);
        if ((err = prepare_request(&q, parts, one[0]))) return -1.0;
        std::vector<Outputs> outs;
        if ((err = run_batch_static(ctx, one, outs, false, t.use_graph != 0))) return -1.0;
        const double ms = ctx->last_stats[1];
        return ms > 0.0 ? ps / (ms * 1e-3) : -1.0;
    };
    auto fail = [&]() {                   // a failed trial ends the search with its error
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (had_prev) ctx->tuning[key] = prev;     // the caller's tuning stays as it was
        else ctx->tuning.erase(key);
        return err;
    };
    const double probe_ps = (double)probe.N * (double)probe.total_steps;
    std::vector<std::pair<double, sp_tuning>> cand;
    for (uint32_t c : chunks)
        for (uint32_t b : blocks)
            for (uint32_t w : waves) {
                sp_tuning t{b, c, 1, w};
                trial(t, r, probe_ps);           // warm (allocations, code load)
                // the better of two timed runs
                double rate = err == SP_OK ? trial(t, r, probe_ps) : -1.0;
                if (err == SP_OK) rate = std::max(rate, trial(t, r, probe_ps));
                if (err != SP_OK) return fail();
                cand.emplace_back(rate, t);
            }
    // the four fastest again, on a run up to four times longer: the short trials rank them only roughly
    std::sort(cand.begin(), cand.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    sp_request r2 = *req;
    r2.flags |= SP_FLAG_SKIP_CHECK;
    if (r2.n_steps > 4 * max_steps) r2.n_steps = 4 * max_steps;
    HostRequest probe2;
    if (int e = prepare_request(&r2, parts, probe2)) return e;
    const double probe2_ps = (double)probe2.N * (double)probe2.total_steps;
    for (size_t i = 0; i < std::min<size_t>(4, cand.size()); ++i) {
        const sp_tuning& t = cand[i].second;
        double rate = trial(t, r2, probe2_ps);
        if (err == SP_OK) rate = std::max(rate, trial(t, r2, probe2_ps));
        if (err != SP_OK) return fail();
        if (rate > best_r) { best_r = rate; bt = t; }
    }
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->tuning[tuning_key(r.problem, r.solver, prec)] = bt;
    }
    *best = bt;
    if (best_rate) *best_rate = best_r;
    return best_r > 0 ? SP_OK : set_error(SP_ERR_DEVICE, "autotuner: every trial failed");
}

}  // extern "C"
