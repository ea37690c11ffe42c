"""Benchmark of the gfx950 ensemble SDE engine.

Workload of the headline number (one "step" = one full pass of the hot path
over one batch): the coupled Brown+Neel ensemble, BrownAndNeel full model,
uniaxial, Heun_Strong, FP64 -- 65,536 particles, 2e5 steps (199,950 executed:
output[0] is the start state), dt 1e-11 s, 12 mT sinusoidal z-field at
100 kHz, output every 50 steps: 1.3104e10 particle-steps per pass.  Inputs
are uploaded once; the timed region runs K passes with the inputs resident
in HBM.

One JSON line on stdout: particle-steps/s of the plain FP64 path (value), the
vector-ALU roofline of its kernel, a CPU baseline on host cores, and
  * "mixed": the 512 uneven sets of the batching distribution, plain FP64 one
    request at a time vs the batched persistent path in FP64 and in mixed
    precision (with its FP64 check), autotuned, and the <m> error of mixed
    against FP64 beside the rounding floor of FP64 itself;
  * "precision": the NeelSpherical cubic precision study at full size
    (1,048,576 particles x 5e5 steps), Heun_Strong and Millstein, FP64 vs mixed;
  * "serving": Poisson arrivals of spherical-cubic requests (Implicit_Midpoint
    for damping >= 0.3) at 200 req/s for 60 s through the continuous-batching
    server; the server's capacity from a run whose arrivals outpace it; then
    50%, 80% and 95% of that capacity: p50/p99 end-to-end latency and
    throughput;
  * "neel": the 4,096-particle Neel relaxation, FP64 and mixed, beside the CPU
    oracle on the whole ensemble;
  * "provider": the particles of a parameter file drawn on the device vs
    materialised on the host and uploaded;
  * "problems" (--problems): FP64 and mixed throughput of every problem class.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--serve-seconds S]
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic FP64 flops per particle-step of the BrownAndNeel (full, uniaxial) Heun_Strong step
# (add/sub/mul = 1, fma = 2; divisions, square roots, transcendentals and the RNG's integer work
# excluded): two det+drift evaluations 2 x 107, one S.dW 76, the Heun combination 36, Kahan 24.
FLOPS_PER_STEP = 350.0
PEAK_FP64_VALU_TFLOPS = 78.6      # MI355X vector FP64
PEAK_HBM_GBS = 8000.0
TRAFFIC_PROFILE = ROOT / "profiles" / "k_plain_bn_heun_fp64_traffic.md"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(e, budget_s=12.0):
    """Oracle (reference-style C++: one particle per task, one thread per core) on a bounded
    sample of the same workload: all steps, fewer particles; rate in particle-steps/s."""
    import oracle
    subprocess.run(["make", "-s", "-B", "-C", str(ROOT / "oracle"), "liboracle_native.so"], check=True,
                   stdout=subprocess.DEVNULL)
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    # a one-GPU lease owns 16 host cores whatever nproc shows (the rest belong to other GPUs' leases)
    cores = max(1, min(visible, 16))
    solver = {"EulerMaruyama": 1, "Implicit_Midpoint": 3, "Millstein": 6, "Heun_Strong": 7, "Heun_NotConsistent": 8}[e.solver]

    def run(sub, steps):
        t0 = time.perf_counter()
        oracle.simulate(sub.pid, solver, 1, steps, sub.oversampling, sub.dt, sub.seed, sub.oracle_props(), sub.y0,
                        axes=sub.axes, field=sub.field_dict(), threads=cores, native=True)
        return time.perf_counter() - t0
    probe = e.subset(np.arange(cores * 4))
    dt_probe = run(probe, 20_000)
    per_ps = dt_probe / (cores * 4 * 20_000)             # seconds per particle-step over all cores
    n = int(max(cores, min(e.n, budget_s / (per_ps * e.total_steps))))
    n = max(cores, n // cores * cores)
    sub = e.subset(np.arange(n))
    secs = run(sub, e.n_steps)
    rate = n * sub.total_steps / secs
    model = ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count() or 0
    return dict(value=rate, unit="particle-steps/s", cores=cores, kind="port", nproc=nproc, affinity_cores=visible,
                core_note="threads = the 16 host cores of a one-GPU lease (nproc and the affinity mask show the host's)",
                per_core=rate / cores, ns_per_particle_step=1e9 * cores / rate, cpu_model=model,
                sample=f"{n} of {e.n} particles x {sub.total_steps} steps ({secs:.1f} s, oracle -O3 -march=native, "
                       f"{cores} threads)")


def traffic_per_launch():
    """HBM bytes per launch of the plain kernel from the committed rocprofv3 PMC summary (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy kernel), or None when it is absent."""
    try:
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", TRAFFIC_PROFILE.read_text())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def perturbed(e, ulps=1):
    """The same ensemble with every start-state component moved by `ulps` units in the last place: the distance of its
    FP64 result from the unperturbed one is the rounding floor of the comparison (what any other rounding of the
    same FP64 arithmetic would give on the same noise stream)."""
    import copy
    p = copy.copy(e)
    y = np.array(e.y0, dtype=np.float64)
    for _ in range(ulps):
        y = np.nextafter(y, np.inf)
    p.y0 = y
    return p


def m_columns(e):
    """Columns of the ensemble mean that hold <m> (BrownAndNeel: 3..5; the particle axis n is 0..2)."""
    return slice(3, 6) if e.problem == "BrownAndNeel" else slice(0, 3)


def diff_stats(a, b, e):
    d = (a["mean"] - b["mean"])[1:, m_columns(e)]           # window 0 is the shared start state
    return float(np.abs(d).max()), float(np.sqrt(np.mean(d ** 2))), float(np.abs(d.mean(axis=0)).max())


def mixed_leg(ctx, configs, n_sets, seed=44, floor_every=8, tune=True, literal=False):
    """The 512-set batching distribution (N log-uniform in [1e3, 1.28e5], steps U[5e4, 5e5], half Neel
    EulerMaruyama / Heun_Strong, half BrownAndNeel Heun_Strong).  As the reference integrates them (no
    renormalisation), the moments of many sets run away over up to 5e5 steps and their means overflow: one batched FP64
    run of that scheme counts them (literal).  The timed legs renormalise after every step (the reference's own
    normalize(), which its EulerMaruyama_Normalized applies): plain FP64 one request at a time (inputs resident) vs one
    batched persistent schedule in FP64 and in mixed precision (with its FP64 check), with tunings chosen by the
    autotuner per kernel class.  <m> error of mixed against FP64, beside the rounding floor (FP64 against FP64 from a
    start state moved by one ulp) on every floor_every-th set."""
    def sets(precision, renorm=True):
        out = configs.config4_sets(n_sets=n_sets, seed=seed, precision=precision)
        for e in out:
            e.renormalize = renorm
        return out
    s64, smx = sets("fp64"), sets("mixed")
    ps = sum(e.particle_steps() for e in s64)
    # the plain baseline runs first, with the engine's default tuning (the autotuner tunes the batched path)
    preps = [ctx.prepare(e) for e in s64]
    preps[0].execute()
    t0 = time.perf_counter()
    ref = [p.execute() for p in preps]
    t_plain = time.perf_counter() - t0
    for p in preps:
        p.release()
    log(f"bench: {n_sets} uneven sets ({ps:.3e} particle-steps): plain FP64 one by one {ps / t_plain:.3e} /s "
        f"({t_plain:.1f} s)")
    lit = None
    if literal:
        t0 = time.perf_counter()
        raw = ctx.run_batch(sets("fp64", renorm=False))
        t_lit = time.perf_counter() - t0
        fin_lit = [bool(np.isfinite(r["mean"]).all()) for r in raw]
        del raw
        lit = dict(sets_with_finite_mean=int(sum(fin_lit)), neel_sets_finite=int(sum(fin_lit[0::2])),
                   brown_and_neel_sets_finite=int(sum(fin_lit[1::2])), batched_fp64_value=ps / t_lit,
                   note="the reference's scheme (no renormalisation): sets whose FP64 ensemble mean stays finite")
        log(f"bench: reference scheme (no renormalisation): {lit['sets_with_finite_mean']} of {n_sets} sets finite "
            f"({lit['neel_sets_finite']} Neel, {lit['brown_and_neel_sets_finite']} BrownAndNeel); batched FP64 "
            f"{ps / t_lit:.3e} /s")
    tunings = {}
    t0 = time.perf_counter()
    if tune:
        for sets in (s64, smx):
            largest = {}
            for e in sets:
                key = (e.problem, e.solver, e.precision)
                if key not in largest or e.particle_steps() > largest[key].particle_steps():
                    largest[key] = e
            for key, e in largest.items():
                tun, rate = ctx.autotune(e)
                tunings["/".join(key)] = dict(tun, probe_rate=rate)
    log(f"bench: autotuned {len(tunings)} kernel classes in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    b64 = ctx.run_batch(s64)
    t_b64 = time.perf_counter() - t0
    same = all(np.array_equal(x["mean"], y["mean"], equal_nan=True) for x, y in zip(b64, ref))
    log(f"bench: batched FP64 {ps / t_b64:.3e} /s ({t_b64:.1f} s); equal to plain bit for bit: {same}")
    t0 = time.perf_counter()
    res = ctx.run_batch(smx)
    t_bmx = time.perf_counter() - t0
    log(f"bench: batched mixed {ps / t_bmx:.3e} /s ({t_bmx:.1f} s)")
    fin = [i for i, r in enumerate(ref) if np.isfinite(r["mean"]).all()]      # all of them once renormalised
    st = np.array([diff_stats(res[i], ref[i], s64[i]) for i in fin])
    sub = fin[::floor_every]
    fl_res = ctx.run_batch([perturbed(s64[i]) for i in sub])
    fl = np.array([diff_stats(f, ref[i], s64[i]) for f, i in zip(fl_res, sub)])
    mx_sub = np.array([diff_stats(res[i], ref[i], s64[i]) for i in sub])
    fb = sum(int(r["used_fallback"]) for r in res)
    inv_sqrt_n = np.array([1.0 / math.sqrt(s64[i].n) for i in fin])
    log(f"bench: mixed vs FP64 <m> over {len(fin)} finite sets: max {st[:, 0].max():.2e}, rms {np.sqrt(np.mean(st[:, 1] ** 2)):.2e}, "
        f"time-averaged bias max {st[:, 2].max():.2e}; rounding floor on {len(sub)} sets: max {fl[:, 0].max():.2e} rms "
        f"{np.sqrt(np.mean(fl[:, 1] ** 2)):.2e} (mixed on the same sets: max {mx_sub[:, 0].max():.2e} rms "
        f"{np.sqrt(np.mean(mx_sub[:, 1] ** 2)):.2e}); fallbacks {fb}/{n_sets}")
    return dict(value=ps / t_bmx, unit="particle-steps/s",
                workload=f"{n_sets} uneven sets of the batching distribution, {ps:.4g} particle-steps, renormalised "
                         "after every step", reference_scheme=lit,
                path="batched persistent launch, hipGraph, autotuned tiles", dtype="fp32 bulk, fp64 state and sums",
                fp64_plain_value=ps / t_plain, fp64_batched_value=ps / t_b64, batched_equals_plain=bool(same),
                speedup_vs_fp64_plain=t_plain / t_bmx, speedup_vs_fp64_batched=t_b64 / t_bmx,
                seconds=dict(plain_fp64=t_plain, batched_fp64=t_b64, batched_mixed=t_bmx),
                tunings=tunings,
                sets_with_finite_fp64_mean=len(fin),
                sets_with_finite_mixed_mean=int(sum(np.isfinite(r["mean"]).all() for r in res)),
                err_max_vs_fp64=float(st[:, 0].max()), err_median_of_set_max=float(np.median(st[:, 0])),
                err_rms_vs_fp64=float(np.sqrt(np.mean(st[:, 1] ** 2))), err_bias_max_vs_fp64=float(st[:, 2].max()),
                sets_within_1e3_per_window=int((st[:, 0] <= 1e-3).sum()),
                sets_within_1e3_bias=int((st[:, 2] <= 1e-3).sum()),
                rounding_floor=dict(sets=len(sub), err_max=float(fl[:, 0].max()), err_rms=float(np.sqrt(np.mean(fl[:, 1] ** 2))),
                                    bias_max=float(fl[:, 2].max()), mixed_err_max_same_sets=float(mx_sub[:, 0].max()),
                                    mixed_err_rms_same_sets=float(np.sqrt(np.mean(mx_sub[:, 1] ** 2))),
                                    how="FP64 from a start state moved by one ulp vs FP64, same noise stream"),
                err_max_over_inv_sqrt_n_max=float((st[:, 0] / inv_sqrt_n).max()),
                check_error_max=float(max(r["check_error"] for r in res)),
                check_bias_max=float(max(r["check_bias"] for r in res)),
                check_noise_max=float(max(r["check_noise"] for r in res)), fp64_fallbacks=fb,
                note="<m> = the magnetisation columns of the mean; trajectories that differ only by rounding decorrelate, "
                     "so the per-window error is Monte-Carlo noise of order sigma/sqrt(N) whose size the rounding floor "
                     "shows; the time-averaged bias is the precision figure (the one the FP64 check accepts on)")


def serving_leg(ctx, configs, seconds, rate, seed=55):
    """Open-loop Poisson arrivals into the continuous-batching server; latency = submit -> result ready."""
    rng = np.random.default_rng(seed)
    n_req = max(1, int(rng.poisson(rate * seconds)))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    reqs = [configs.serving_request(rng, i) for i in range(n_req)]
    log(f"bench: serving {n_req} requests over {arrivals[-1]:.1f} s ({rate:.0f} req/s offered)")
    ids = [None] * n_req
    st0 = ctx.server_stats()
    t0 = time.perf_counter()
    for i, e in enumerate(reqs):
        delay = arrivals[i] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids[i] = ctx.submit(e)
    t_sub = time.perf_counter() - t0
    lat, ps, fallbacks, im = [], 0.0, 0, 0
    for i, e in enumerate(reqs):
        ctx.wait(ids[i], -1)
        r = ctx.result(ids[i], e)
        lat.append(r["latency_ms"])
        ps += r["particle_steps"]
        fallbacks += int(r["used_fallback"])
        im += int(e.solver == "Implicit_Midpoint")
    wall = time.perf_counter() - t0
    st1 = ctx.server_stats()
    st = {k: st1[k] - st0[k] for k in ("busy_ms", "iterations", "completed")}
    lat = np.array(lat)
    offered = sum(e.particle_steps() for e in reqs) / arrivals[-1]
    by_solver = {}
    for solver in ("Heun_Strong", "Implicit_Midpoint"):
        x = np.array([l for l, e in zip(lat, reqs) if e.solver == solver])
        if len(x):
            by_solver[solver] = dict(requests=int(len(x)), p50_ms=float(np.percentile(x, 50)),
                                     p99_ms=float(np.percentile(x, 99)))
    log(f"bench: serving p50 {np.percentile(lat, 50):.1f} ms p99 {np.percentile(lat, 99):.1f} ms, "
        f"{ps / wall:.3e} particle-steps/s, busy {st['busy_ms'] / 1e3:.1f} s of {wall:.1f} s; "
        + ", ".join(f"{k} p50 {v['p50_ms']:.1f} p99 {v['p99_ms']:.1f} ms" for k, v in by_solver.items()))
    return dict(requests=n_req, offered_req_per_s=rate, arrival_window_s=float(arrivals[-1]), submit_s=t_sub,
                wall_s=wall, p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), throughput=ps / wall, unit="particle-steps/s",
                offered_particle_steps_per_s=offered, device_busy_frac=st["busy_ms"] / (wall * 1e3),
                implicit_midpoint_requests=im, fp64_fallbacks=fallbacks, iterations=int(st["iterations"]),
                latency_by_solver=by_solver,
                request_sizes="N log-uniform [256, 4096], steps U[2e3, 1e4], OS 100; NeelSpherical cubic Lissajous, "
                              "mixed (Heun_Strong) or FP64 Implicit_Midpoint for damping >= 0.3")


def serving_study(ctx, configs, seconds, rate, low_seconds, seed=55, sat_seconds=8.0):
    """The serving configuration (Poisson arrivals at `rate` for `seconds`); the server's capacity, measured as the
    served particle-steps/s of a run whose arrivals (8x `rate`, `sat_seconds`) outpace the device; then the same
    request distribution at 50%, 80% and 95% of that capacity for `low_seconds` each: p50/p99 latency against
    offered load."""
    main = serving_leg(ctx, configs, seconds, rate, seed)
    sat = serving_leg(ctx, configs, sat_seconds, 8.0 * rate, seed + 10)
    capacity = sat["throughput"]
    per_request = main["offered_particle_steps_per_s"] / rate
    main["offered_fraction_of_capacity"] = main["offered_particle_steps_per_s"] / capacity
    loads = []
    for k, frac in enumerate((0.5, 0.8, 0.95)):
        r = serving_leg(ctx, configs, low_seconds, frac * capacity / per_request, seed + 1 + k)
        loads.append({key: r[key] for key in ("requests", "offered_req_per_s", "p50_ms", "p99_ms", "mean_ms", "throughput",
                                              "offered_particle_steps_per_s", "device_busy_frac", "fp64_fallbacks",
                                              "latency_by_solver")})
        loads[-1]["offered_fraction_of_capacity"] = r["offered_particle_steps_per_s"] / capacity
    main.update(capacity_particle_steps_per_s=capacity,
                capacity_run=dict(offered_req_per_s=8.0 * rate, seconds=sat_seconds, requests=sat["requests"],
                                  p50_ms=sat["p50_ms"], p99_ms=sat["p99_ms"]),
                other_loads=loads,
                note="device_busy_frac sums the device time of the server's two lanes (explicit solvers, "
                     "Implicit_Midpoint), whose iterations overlap")
    return main


def precision_study(ctx, configs, n, n_steps):
    """NeelSpherical, cubic K1 = -1.1e4, log-normal radii, Lissajous field (the precision-study ensemble):
    FP64 and mixed on the same Philox stream, Heun_Strong and Millstein; throughput and <m> error."""
    out = {}
    for solver in ("Heun_Strong", "Millstein"):
        row = {}
        res = {}
        with_floor = solver == "Heun_Strong"
        for prec in ("fp64", "mixed"):
            e = configs.config3(n=n, n_steps=n_steps, solver=solver, precision=prec)
            e.skip_check = True
            p = ctx.prepare(e)
            t0 = time.perf_counter()
            res[prec] = p.execute()
            dt = time.perf_counter() - t0
            p.release()
            row[prec] = e.particle_steps() / dt
            log(f"bench: precision study {solver} {prec}: {row[prec]:.3e} particle-steps/s ({dt:.1f} s)")
        d = res["mixed"]["mean"] - res["fp64"]["mean"]
        row.update(err_max=float(np.abs(d).max()), err_rms=float(np.sqrt(np.mean(d ** 2))),
                   err_bias_max=float(np.abs(d[1:].mean(axis=0)).max()),     # time-averaged difference per component
                   speedup=row["mixed"] / row["fp64"], monte_carlo_sigma_over_sqrt_n=float(0.58 / math.sqrt(n)))
        if with_floor:      # rounding floor: FP64 from a start state moved by one ulp, on the same noise stream
            e = perturbed(configs.config3(n=n, n_steps=n_steps, solver=solver, precision="fp64"))
            p = ctx.prepare(e)
            f = p.execute()["mean"] - res["fp64"]["mean"]
            p.release()
            row.update(floor_max=float(np.abs(f).max()), floor_rms=float(np.sqrt(np.mean(f ** 2))),
                       floor_bias_max=float(np.abs(f[1:].mean(axis=0)).max()))
        log(f"bench: precision study {solver}: mixed - FP64 <m> max {row['err_max']:.2e} rms {row['err_rms']:.2e} bias "
            f"{row['err_bias_max']:.2e}" + (f"; floor (FP64 one ulp apart) max {row['floor_max']:.2e} rms "
                                           f"{row['floor_rms']:.2e}" if with_floor else ""))
        out[solver] = row
    return dict(workload=f"NeelSpherical cubic, {n} particles x {(n_steps // 100 - 1) * 100} steps, dt 2e-12 s, "
                         "Lissajous 12 mT at 25.25/24.5 kHz, output every 100 steps", unit="particle-steps/s",
                floor="FP64 from a start state moved by one ulp vs FP64, same noise stream", results=out)


def neel_relaxation(ctx, configs):
    """Plain Neel relaxation, 4,096 uniaxial particles, EulerMaruyama, 2e5 steps (the reference's CPU-sized
    case): FP64 and mixed throughput beside the CPU oracle on the full ensemble."""
    e = configs.config1()
    p = ctx.prepare(e)
    p.execute()
    t0 = time.perf_counter()
    r64 = p.execute()
    t64 = time.perf_counter() - t0
    p.release()
    em = configs.config1(precision="mixed")
    em.skip_check = True
    pm = ctx.prepare(em)
    pm.execute()
    t0 = time.perf_counter()
    rm = pm.execute()
    tm = time.perf_counter() - t0
    pm.release()
    cpu = cpu_baseline(e, budget_s=60.0)
    ps = e.particle_steps()
    d = rm["mean"] - r64["mean"]
    log(f"bench: Neel relaxation (4,096 particles): FP64 {ps / t64:.3e}, mixed {ps / tm:.3e} particle-steps/s; CPU "
        f"{cpu['value']:.3e} on {cpu['cores']} threads ({cpu['sample']})")
    return dict(workload="Neel uniaxial EulerMaruyama, 4096 particles x 199,900 steps, dt 1e-12 s, zero field",
                fp64=ps / t64, mixed=ps / tm, unit="particle-steps/s", cpu_baseline=cpu, fp64_vs_cpu=ps / t64 / cpu["value"],
                err_max=float(np.abs(d).max()), err_rms=float(np.sqrt(np.mean(d ** 2))),
                note="the GPU is latency bound here: 4,096 particles are 64 tiles on 1,024 SIMDs, each run as an integrator "
                     "wave and two noise waves (split noise)")


def problem_rates(ctx, configs, n=65_536, n_steps=20_000):
    """FP64 and mixed throughput of every problem class on one n-particle ensemble each (20,000 steps, inputs
    resident): Neel EulerMaruyama, NeelSpherical Heun_Strong (cubic), BrownAndNeel Heun_Strong, BrownAndNeelEulerSpherical
    Heun_Strong (both charts on), Neel and NeelSpherical with uniaxialcubic anisotropy; each rate is the best of three
    timed runs after an untimed one."""
    makers = {
        "Neel uniaxial EulerMaruyama": lambda prec: configs.config1(n=n, n_steps=n_steps, precision=prec),
        "NeelSpherical cubic Heun_Strong": lambda prec: configs.config3(n=n, n_steps=n_steps, precision=prec),
        "BrownAndNeel uniaxial Heun_Strong": lambda prec: configs.config2(n=n, n_steps=n_steps, precision=prec),
        "BrownAndNeelEulerSpherical uniaxial Heun_Strong": lambda prec: configs.bnes(n=n, n_steps=n_steps, precision=prec),
        "Neel uniaxialcubic Heun_Strong": lambda prec: configs.uniaxialcubic(n=n, n_steps=n_steps, precision=prec),
        "NeelSpherical uniaxialcubic Heun_Strong": lambda prec: configs.uniaxialcubic(
            n=n, problem="NeelSpherical", n_steps=n_steps, precision=prec),
    }
    out = {}
    for name, make in makers.items():
        row = {}
        for prec in ("fp64", "mixed"):
            e = make(prec)
            e.skip_check = True
            p = ctx.prepare(e)
            p.execute()
            dt = float("inf")
            for _ in range(3):      # a run takes 10-150 ms: the best of three, as host jitter only adds time
                t0 = time.perf_counter()
                p.execute()
                dt = min(dt, time.perf_counter() - t0)
            p.release()
            row[prec] = e.particle_steps() / dt
        log(f"bench: {name}: fp64 {row['fp64']:.3e}, mixed {row['mixed']:.3e} particle-steps/s")
        out[name] = row
    return dict(workload=f"{n} particles x {n_steps} steps per ensemble (inputs resident, mixed without its FP64 check, "
                         "best of three timed runs)",
                unit="particle-steps/s", results=out)


def implicit_rates(ctx, configs, n=1_048_576, n_steps=4_000):
    """The serving physics (NeelSpherical, cubic, Lissajous, chart rotation) with damping 0.4, where the serving
    distribution takes Implicit_Midpoint: Implicit_Midpoint (FP64 Newton with dual-number Jacobians, tolerance 1e-10,
    at most 20 iterations) beside Heun_Strong FP64 on the same particles, inputs resident; Newton failures are the
    particle-steps that end at the iteration limit."""
    rng = np.random.default_rng(1)
    e = configs.serving_request(rng, 0, n_range=(n, n), step_range=(n_steps, n_steps + 1), precision="fp64")
    e.alpha[:] = 0.4
    out = {}
    for solver in ("Implicit_Midpoint", "Heun_Strong"):
        e.solver = solver
        p = ctx.prepare(e)
        p.execute()
        t0 = time.perf_counter()
        r = p.execute()
        dt = time.perf_counter() - t0
        p.release()
        out[solver] = dict(value=e.particle_steps() / dt, newton_failures=int(r["newton_failures"]),
                           failure_fraction=r["newton_failures"] / e.particle_steps())
        log(f"bench: {solver} fp64: {out[solver]['value']:.3e} particle-steps/s, Newton failures "
            f"{out[solver]['newton_failures']}")
    return dict(workload=f"NeelSpherical cubic serving physics, damping 0.4, {n} particles x {e.n_steps} steps",
                unit="particle-steps/s", results=out)


PROVIDER_PARTICLE = """[Particle_Properties]
Temperature = 300
[Particle_Properties.Magnetic_Properties]
Magnetic_radius = {radius}
Saturation_magnetisation = 4.8e5
Damping_constant = 0.2
Gyromagnetic_ratio = 1.76e11
Type_of_anisotropy = 'uniaxial'
[Particle_Properties.Magnetic_Properties.Anisotropy_Properties]
K_uniaxial = 5e3
[Particle_Properties.Hydrodynamic_Properties]
Hydrodynamic_Radius = {r_h}
Viscosity = 1.0e-03
[Particle_Simulation_Initialization]
Use_random_initial_particle_orientation = TRUE
Use_random_initial_particle_magnetisation = TRUE
[Particle_Simulation_Settings]
Magnetic_radius_distribution_type = 'lognormal'
Magnetic_radius_distribution_width = 0.1
Use_relative_distribution_width_for_magnetic_radius = TRUE
Hydrodynamic_shell_distribution_type = 'normal'
Hydrodynamic_shell_distribution_width = 0.05
Use_relative_distribution_width_for_hydrodynamic_shell = TRUE
Distribution_type = 'normal'
Anisotropy_constants_distribution_width = 0.05
"""

PROVIDER_RUN = """[Solver_Settings]
Solver = 'Heun_Strong'
[Simulation_Settings]
Timestep = 1e-11
Number_of_steps = 200000
Oversampling = 50
Number_of_simulations = {n}
Seed = 7
[Field_Parameters]
Type_of_field = 'sinusoidal'
Offset_Field = {{0.0, 0.0, 0.0}}
Amplitude = {{0.0, 0.0, 0.012}}
Frequency = 1e5
Phase = 0.0
[Problem_Settings]
Type_of_problem = 'BrownAndNeel'
[Particle_Provider]
Use_discrete_distribution_to_select_particle = TRUE
[Particle_Provider.Particle_List]
Small = 'small.ini'
Large = 'large.ini'
[Particle_Provider.Particle_Numbers]
Small = 3
Large = 1
"""


def provider_leg(ctx, n_device=1 << 24, n_host=1 << 20):
    """Particles of a two-type BrownAndNeel parameter file (discrete type selection, lognormal radii, normal shells
    and K, random orientations and magnetisations): drawn by the provider kernel straight into HBM
    (sp_settings_prepare) vs materialised on the host, run through the parameter calculators and uploaded
    (sp_settings_request + sp_prepare), the path a host-side provider takes."""
    import ctypes as C
    import tempfile
    from stochasticphysics_amd import Settings, capi
    from stochasticphysics_amd.capi import Particles, Request, lib
    with tempfile.TemporaryDirectory() as d:
        d = Path(d)
        (d / "small.ini").write_text(PROVIDER_PARTICLE.format(radius=1.0e-8, r_h=2.2e-8))
        (d / "large.ini").write_text(PROVIDER_PARTICLE.format(radius=1.4e-8, r_h=2.5e-8))
        (d / "dev.ini").write_text(PROVIDER_RUN.format(n=n_device))
        (d / "host.ini").write_text(PROVIDER_RUN.format(n=n_host))
        s = Settings(d / "dev.ini")
        ctx.prepare_settings(s).release()                     # first use: allocation pools, code objects
        t0 = time.perf_counter()
        h = ctx.prepare_settings(s)
        dev_s = time.perf_counter() - t0
        h.release()
        t0 = time.perf_counter()
        hs = C.c_void_p()
        capi.check(lib().sp_settings_load(str(d / "host.ini").encode(), C.byref(hs)))
        r, p = Request(), Particles()
        capi.check(lib().sp_settings_request(hs, C.byref(r), C.byref(p)))
        t1 = time.perf_counter()
        hp = C.c_void_p()
        capi.check(lib().sp_prepare(ctx._ctx, C.byref(r), C.byref(p), C.byref(hp)))
        host_s = time.perf_counter() - t0
        lib().sp_release(hp)
        lib().sp_settings_free(hs)
    bytes_per_particle = (capi.NPARAM + 6) * 8
    out = dict(workload="BrownAndNeel parameter file, two particle types (discrete selection), lognormal radius, normal "
                        "shell and K, random orientation and magnetisation",
               device_particles=n_device, device_prepare_ms=1e3 * dev_s, device_particles_per_s=n_device / dev_s,
               host_particles=n_host, host_prepare_ms=1e3 * host_s, host_draw_ms=1e3 * (t1 - t0),
               host_particles_per_s=n_host / host_s, speedup=(n_device / dev_s) / (n_host / host_s),
               bytes_written_per_particle=bytes_per_particle,
               note="device: parameter block and start state drawn into HBM (alloc + kernel + per-group sums); "
                    "host: .ini draw, parameter calculators, upload")
    log(f"bench: provider: device {out['device_particles_per_s']:.3e} particles/s ({n_device} in "
        f"{out['device_prepare_ms']:.1f} ms), host {out['host_particles_per_s']:.3e} particles/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="reduced workload (tests of the script itself)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="headline run without the precision study")
    ap.add_argument("--no-neel", action="store_true", help="headline run without the Neel relaxation leg")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-low-seconds", type=float, default=8.0,
                    help="duration of each of the serving runs at 50%%, 80%% and 95%% of the measured capacity")
    ap.add_argument("--batch-sets", type=int, default=512, help="uneven sets in the mixed/batched leg")
    ap.add_argument("--no-tune", action="store_true", help="batched legs with the default tuning (no autotuner)")
    ap.add_argument("--reference-scheme", action="store_true",
                    help="also run the 512 sets as the reference integrates them (no renormalisation) and count the "
                         "sets whose FP64 mean stays finite")
    ap.add_argument("--problems", action="store_true", help="headline run with the per-problem leg")
    ap.add_argument("--workload", choices=("headline", "precision", "neel", "problems", "provider", "implicit"),
                    default="headline",
                    help="precision: the NeelSpherical precision study; neel: the 4,096-particle Neel relaxation; "
                         "implicit: Implicit_Midpoint beside Heun_Strong on the serving physics (1M particles); "
                         "problems: every problem class on a 65,536-particle ensemble; provider: particles of a "
                         "parameter file drawn on the device vs on the host (each prints its own JSON line instead "
                         "of the headline one)")
    ap.add_argument("--no-provider", action="store_true", help="headline run without the particle-provider leg")
    ap.add_argument("--particles", type=int, default=1_048_576, help="precision study: particles")
    ap.add_argument("--problem-particles", type=int, default=65_536, help="per-problem leg: particles per ensemble")
    ap.add_argument("--study-steps", type=int, default=500_000, help="precision study: steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (the reference kernel that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes=4 * n, fetch_expected=4 * n, write_expected=4 * n)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the ensemble engine shards by "
              f"particle range across GPUs only by design (one exchange of the window sums per output)",
              file=sys.stderr)
        sys.exit(2)

    from stochasticphysics_amd import Context, configs
    if args.workload != "headline":
        ctx = Context(0)
        if args.workload == "precision":
            res = precision_study(ctx, configs, args.particles, args.study_steps)
        elif args.workload == "problems":
            res = problem_rates(ctx, configs, n=args.problem_particles)
        elif args.workload == "provider":
            res = provider_leg(ctx)
        elif args.workload == "implicit":
            res = implicit_rates(ctx, configs, *((args.particles,) if args.particles != 1_048_576 else ()))
        else:
            res = neel_relaxation(ctx, configs)
        ctx.close()
        print(json.dumps(res), flush=True)
        return
    n_steps = 20_000 if args.quick else 200_000
    e = configs.config2(n_steps=n_steps)
    ctx = Context(0)
    log(f"bench: BrownAndNeel Heun_Strong FP64 N={e.n} S={e.total_steps}")
    # the timed passes return what the reference returns (the ensemble means); the per-particle final state is
    # read once afterwards, from a separate run, for the count of particles whose |m| ran away
    prep = ctx.prepare(e)
    for _ in range(args.warmup):
        prep.execute()
    kernel_ms, launches = 0.0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        ref = prep.execute()
        st = ctx.last_run_stats()
        kernel_ms += st["kernel_ms"]
        launches += int(st["launches"])
        log(f"bench: pass {i + 1}/{args.steps} done ({st['kernel_ms']:.1f} ms kernel)")
    elapsed = time.perf_counter() - t0
    prep.release()
    ref = ctx.prepare(e, keep_final=True).execute()
    ps = e.particle_steps()
    value = args.steps * ps / elapsed
    kern_avg_ms = kernel_ms / max(1, launches)
    ps_per_launch = ps / max(1, launches // args.steps)
    achieved = FLOPS_PER_STEP * ps_per_launch / (kern_avg_ms * 1e-3) / 1e12
    bad = int((~np.isfinite(ref["final_state"]).all(axis=1)).sum())
    out = dict(metric="particle-steps/s plain-FP64 vs batched-mixed at <=1e-3 <m> err; p50/p99 latency",
               value=value, unit="particle-steps/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=elapsed / args.steps * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64", data="synthetic",
               config=dict(workload="BrownAndNeel full model, uniaxial, Heun_Strong, FP64: 65536 particles x "
                                    f"{e.total_steps} steps, dt 1e-11 s, r 12 nm, r_h 25 nm, K 5e3, 12 mT 100 kHz "
                                    "sinusoidal z-field, output every 50 steps",
                           particles=e.n, steps=e.total_steps, particle_steps_per_pass=ps),
               roofline=dict(bound="valu", achieved=achieved, peak=PEAK_FP64_VALU_TFLOPS, unit="TFLOP/s",
                             frac=achieved / PEAK_FP64_VALU_TFLOPS, traffic=traffic_per_launch(),
                             kernel="k_plain<BNCfg<uniaxial,Heun_Strong,double>>", kernel_avg_ms=kern_avg_ms,
                             launches_per_pass=launches // args.steps, flops_per_particle_step=FLOPS_PER_STEP),
               reference_scheme=dict(nonfinite_particles=bad, finite_mean=bool(np.isfinite(ref["mean"]).all()),
                                     note="Heun_Strong as the reference writes it does not keep |m| = 1 for this "
                                          "problem; particles whose |m| runs away overflow, as in the reference"))
    if not args.no_mixed:
        out["mixed"] = mixed_leg(ctx, configs, args.batch_sets if not args.quick else 4, tune=not args.no_tune,
                                 literal=args.reference_scheme)
    if not args.no_precision:
        out["precision"] = precision_study(ctx, configs, *((1 << 14, 20_000) if args.quick else (args.particles,
                                                                                                 args.study_steps)))
    if not args.no_serving:
        out["serving"] = serving_study(ctx, configs, 5.0 if args.quick else args.serve_seconds, args.serve_rate,
                                       2.0 if args.quick else args.serve_low_seconds)
    if not args.no_neel and not args.quick:
        out["neel"] = neel_relaxation(ctx, configs)
    if args.problems:
        out["problems"] = problem_rates(ctx, configs)
    if not args.no_provider:
        out["provider"] = provider_leg(ctx, *((1 << 16, 1 << 12) if args.quick else ()))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(e)
        out["vs_cpu"] = value / out["cpu_baseline"]["value"]
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
