"""Benchmark of the gfx950 ensemble SDE engine.

Workload (one "step" = one full pass of the hot path over one batch): the
coupled Brown+Neel ensemble, Heun_Strong, FP64 -- 65,536 particles, 2e5 steps,
dt 1e-11 s, 12 mT sinusoidal z-field at 100 kHz, output every 50 steps
(1.3104e10 particle-steps per pass).  Inputs are uploaded once (sp_prepare);
the timed region runs K passes with the inputs resident in HBM.

Printed JSON line: particle-steps/s of the plain FP64 path (value), plus the
mixed-precision batched path and its <m> error, serving p50/p99 latency, the
vector-ALU roofline of the dominant kernel and a CPU baseline on host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# algorithmic FP64 flops per particle-step of the BrownAndNeel Heun_Strong step
# (add/sub/mul = 1, fma = 2; transcendentals and RNG integer work excluded),
# counted from the kernel formulation; breakdown in DESIGN.md.
FLOPS_PER_STEP = {"BrownAndNeel/Heun_Strong": 382.0}
PEAK_FP64_VALU_TFLOPS = 78.6      # MI355X vector FP64
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(e, budget_s=12.0):
    """Oracle (reference-style C++: one particle per task, one thread per core) on a bounded
    sample of the same workload: all 2e5 steps, fewer particles; rate in particle-steps/s."""
    import oracle
    subprocess.run(["make", "-s", "-B", "-C", str(ROOT / "oracle"), "liboracle_native.so"], check=True,
                   stdout=subprocess.DEVNULL)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    solver = {"EulerMaruyama": 1, "Implicit_Midpoint": 3, "Millstein": 6, "Heun_Strong": 7}[e.solver]

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
    return dict(value=rate, unit="particle-steps/s", cores=cores, kind="port",
                sample=f"{n} of {e.n} particles x {sub.total_steps} steps ({secs:.1f} s, oracle -O3 -march=native, "
                       f"{cores} threads)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="reduced workload (tests of the script itself)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the ensemble engine shards by particle range across GPUs "
              f"(design only, see DESIGN.md); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    from stochasticphysics_amd import Context, configs
    n_steps = 20_000 if args.quick else 200_000
    e = configs.config2(n_steps=n_steps)
    ctx = Context(0)
    log(f"bench: config2 BrownAndNeel Heun_Strong FP64 N={e.n} S={e.total_steps}")
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
    ps = e.particle_steps()
    value = args.steps * ps / elapsed
    kern_avg_ms = kernel_ms / max(1, launches)
    flops = FLOPS_PER_STEP["BrownAndNeel/Heun_Strong"]
    ps_per_launch = ps / max(1, launches // args.steps)
    achieved = flops * ps_per_launch / (kern_avg_ms * 1e-3) / 1e12
    out = dict(metric="particle-steps/s plain-FP64 vs batched-mixed at <=1e-3 <m> err; p50/p99 latency",
               value=value, unit="particle-steps/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=elapsed / args.steps * 1e3, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64", data="synthetic",
               config=dict(workload="config2: BrownAndNeel full model, uniaxial, Heun_Strong, FP64; 65536 particles x "
                                    f"{e.total_steps} steps, dt 1e-11 s, 12 mT 100 kHz sinusoidal z-field, output every 50 steps",
                           particles=e.n, steps=e.total_steps, particle_steps_per_pass=ps),
               roofline=dict(bound="valu", achieved=achieved, peak=PEAK_FP64_VALU_TFLOPS, unit="TFLOP/s",
                             frac=achieved / PEAK_FP64_VALU_TFLOPS, traffic=None,
                             kernel="k_plain<BNCfg<uniaxial,Heun_Strong,double>>", kernel_avg_ms=kern_avg_ms,
                             flops_per_particle_step=flops, vgprs=st["vgprs"]))
    prep.release()
    if not args.no_mixed:
        em = configs.config2(n_steps=n_steps, precision="mixed")
        em.skip_check = True
        pm = ctx.prepare(em)
        pm.execute()
        t0 = time.perf_counter()
        mres = pm.execute()
        t_mixed = time.perf_counter() - t0
        pm.release()
        err = float(np.max(np.abs(mres["mean"] - ref["mean"])))
        out["mixed"] = dict(value=ps / t_mixed, unit="particle-steps/s", path="plain", err_vs_fp64=err,
                            speedup_vs_fp64=(ps / t_mixed) / value)
        log(f"bench: mixed {ps / t_mixed:.3e} particle-steps/s, max |<m>_mixed - <m>_fp64| = {err:.2e}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(e)
        out["vs_cpu"] = value / out["cpu_baseline"]["value"]
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
