"""Batched persistent path and continuous-batching server on the GPU.

Results of a request must not depend on what it runs with: Philox noise is keyed
by (seed, particle, step) and partial sums are reduced per 64-particle tile in a
fixed order, so alone / batched / mixed-with-others results are bit-identical.
"""
import math
import time

import numpy as np
import pytest

from stochasticphysics_amd import Context, configs
from stochasticphysics_amd.capi import SP_OK, SpError

from test_gpu_parity import bn_ensemble, bnes_ensemble, neel_ensemble, run_oracle, sph_ensemble

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def mixed_bag():
    return [neel_ensemble(n=200, n_steps=3000, solver="EulerMaruyama", field="sinusoidal", seed=21),
            neel_ensemble(n=70, n_steps=5000, solver="Heun_Strong", aniso="cubic", seed=22),
            bn_ensemble(n=130, n_steps=1500, solver="Heun_Strong", seed=23),
            sph_ensemble(n=150, n_steps=2500, solver="Millstein", seed=24),
            sph_ensemble(n=64, n_steps=1000, solver="Implicit_Midpoint", seed=25),
            bn_ensemble(n=90, n_steps=1200, solver="Heun_NotConsistent", seed=26),
            bnes_ensemble(n=100, n_steps=1500, solver="Millstein", aniso="cubic", seed=27)]


def same(a, b):
    assert np.array_equal(a["mean"], b["mean"])
    assert np.array_equal(a["weight"], b["weight"])
    assert np.array_equal(a["time"], b["time"])


def test_batch_equals_plain(ctx):
    bag = mixed_bag()
    plain = [ctx.run(e) for e in bag]
    batch = ctx.run_batch(bag)
    for p, b in zip(plain, batch):
        same(p, b)


@pytest.mark.parametrize("n", [4096, 25_600, 70_000])
@pytest.mark.parametrize("case", ["neel_em_fp64", "neel_heun_mixed", "bn_heun_fp64", "bn_heun_mixed", "sph_heun_fp64",
                                  "sph_heun_mixed"])
def test_plain_variants_equal_batch(ctx, case, n):
    """A plain launch of 64 tiles runs split with two noise waves per tile, of 400 tiles split with one, of 1,094
    tiles in the usual one-wave-per-tile kernels; NeelSpherical launches of at most one tile per SIMD, plain or
    batched, take the classes' one-wave builds.  The persistent batch kernel never splits; batched beside a
    70,000-particle ensemble of the same class (more tiles than SIMDs) it takes the regular builds.  All give the
    same bits."""
    solver = "EulerMaruyama" if case.startswith("neel_em") else "Heun_Strong"
    precision = "mixed" if case.endswith("mixed") else "fp64"
    mk = {"neel": neel_ensemble, "bn": bn_ensemble, "sph": sph_ensemble}[case.split("_")[0]]
    e = mk(n=n, n_steps=400, solver=solver, precision=precision, seed=7)
    e.skip_check = True
    big = mk(n=70_000, n_steps=400, solver=solver, precision=precision, seed=8)
    big.skip_check = True
    plain = ctx.run(e)
    same(plain, ctx.run_batch([e])[0])
    same(plain, ctx.run_batch([e, big])[0])


def test_batch_graph_and_direct_launch_agree(ctx):
    bag = mixed_bag()[:3]
    g = ctx.run_batch(bag)
    for e in bag:
        ctx.set_tuning(e.problem, e.solver, e.precision, block_size=128, chunk_steps=1000,
                       use_graph=0, persistent_waves_per_simd=1)
    d = ctx.run_batch(bag)
    for x, y in zip(g, d):
        same(x, y)


def test_batch_mixed_precision_against_oracle(ctx):
    bag = [neel_ensemble(n=256, n_steps=4000, solver="Heun_Strong", field="sinusoidal", seed=31),
           bn_ensemble(n=128, n_steps=2000, solver="Heun_Strong", seed=32)]
    for e in bag:
        e.precision = "mixed"
    res = ctx.run_batch(bag)
    for e, r in zip(bag, res):
        o = run_oracle(e, keep_final=False)
        err = float(np.max(np.abs(r["mean"] - o["mean"])))
        assert r["precision_used"] == 1 and not r["used_fallback"]
        assert err <= 1e-3, err
        # the FP64 shadow check measured the same kind of error on its subset (here the whole request), and
        # its statistical resolution (3 standard errors over the 64-particle tiles) is stated beside it
        assert 0.0 <= r["check_error"] <= 1e-3 and r["check_noise"] >= 0.0


def test_check_accepts_decorrelated_but_unbiased_mixed(ctx):
    """Moments released on an unstable point (in the reference's convention K > 0 makes the +z axis hard, so the
    config-1 start is a maximum of the energy) fall off it differently in FP32 and FP64 under the same noise: the
    two means differ by Monte-Carlo noise well above 1e-3 per output, with no bias.  The check states that error
    and its noise level and keeps the mixed result instead of re-running it in FP64."""
    e = configs.config1(n=8192, n_steps=200_000, oversampling=100, precision="mixed")
    r = ctx.run_batch([e])[0]
    assert r["used_fallback"] == 0 and r["precision_used"] == 1
    assert r["check_error"] > 1e-3 and r["check_noise"] > 0.0
    f = ctx.run(configs.config1(n=8192, n_steps=200_000, oversampling=100))
    assert np.all(np.isfinite(r["mean"])) and np.abs(r["mean"] - f["mean"]).max() > 1e-3


def test_forced_fp64_fallback(ctx):
    e = neel_ensemble(n=256, n_steps=2000, solver="Heun_Strong", field="sinusoidal", seed=41)
    ref = ctx.run(e)                       # FP64 plain
    e.precision = "mixed"
    e.check_tol = -1.0                     # a negative tolerance fails every check
    r = ctx.run_batch([e])[0]
    assert r["used_fallback"] == 1 and r["precision_used"] == 0 and r["check_error"] > 0
    same(r, ref)


def test_server_alone_vs_mix(ctx):
    bag = mixed_bag()
    target = bag[2]
    rid = ctx.submit(target)
    assert ctx.wait(rid, 120_000) == SP_OK
    alone = ctx.result(rid, target)
    ids = [ctx.submit(e) for e in bag]        # the target again, now among six other requests
    outs = []
    for rid, e in zip(ids, bag):
        assert ctx.wait(rid, 120_000) == SP_OK
        outs.append(ctx.result(rid, e))
    same(alone, outs[2])
    plain = [ctx.run(e) for e in bag]
    for p, o in zip(plain, outs):
        same(p, o)
    st = ctx.server_stats()
    assert st["completed"] >= 8 and st["iterations"] >= 1


def test_synchronous_run_while_serving(ctx):
    """The server's worker keeps the device while iterations of its two lanes (explicit solvers; Implicit_Midpoint)
    are in flight; a synchronous run issued meanwhile waits for the lanes to drain, takes the device and returns
    the plain result, and the served requests then finish with theirs."""
    bag = [sph_ensemble(n=256, n_steps=4000, solver="Implicit_Midpoint", seed=90 + i) for i in range(3)] + \
          [sph_ensemble(n=512, n_steps=6000, solver="Heun_Strong", seed=95 + i) for i in range(3)]
    probe = neel_ensemble(n=128, n_steps=2000, solver="Heun_Strong", seed=99)
    plain_probe = ctx.run(probe)
    plain = [ctx.run(e) for e in bag]
    ids = [ctx.submit(e) for e in bag]
    during = ctx.run(probe)
    same(plain_probe, during)
    for rid, e, p in zip(ids, bag, plain):
        assert ctx.wait(rid, 120_000) == SP_OK
        same(p, ctx.result(rid, e))


@pytest.mark.parametrize("problem", ["NeelSpherical", "BrownAndNeelEulerSpherical"])
@pytest.mark.parametrize("solver", ["Heun_Strong", "Millstein"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chunk_boundaries_do_not_change_results(ctx, problem, solver, precision):
    """A launch restarts a tile from its stored state, and a state kept in a rotated chart within an output window
    is mapped back at every window end: NeelSpherical with the chart rotation at pi/4 (a third of the steps
    rotated), BrownAndNeelEulerSpherical with both charts at pi/4, run in launches of 100 steps and of 4,000 steps
    give the same bits."""
    if problem == "NeelSpherical":
        e = sph_ensemble(n=256, n_steps=4000, solver=solver, precision=precision, seed=81)
    else:
        e = bnes_ensemble(n=256, n_steps=4000, solver=solver, precision=precision, seed=81)
        e.brown_min_angle = math.pi / 4
    e.min_angle = math.pi / 4
    e.skip_check = True
    out = []
    for chunk in (100, 4000):
        ctx.set_tuning(e.problem, e.solver, e.precision, block_size=256, chunk_steps=chunk, use_graph=1,
                       persistent_waves_per_simd=2)
        out.append(ctx.run(e))
    ctx.set_tuning(e.problem, e.solver, e.precision, block_size=256, chunk_steps=4000, use_graph=1,
                   persistent_waves_per_simd=2)
    same(out[0], out[1])


def test_server_iteration_budget_keeps_results(ctx):
    """With an iteration budget the server cuts each class's chunk by its measured cost (down to one aligned block
    of steps for a budget this small), and the chunk changes from one iteration to the next: served results still
    equal the plain runs bit for bit.  The requests are submitted twice, so that the second round runs with
    measured costs.  Budget 0 keeps the tuned chunk."""
    bag = mixed_bag()
    plain = [ctx.run(e) for e in bag]
    try:
        for budget in (1e-6, 0.0):
            ctx.set_iteration_budget(budget)
            for rnd in range(2):
                ids = [ctx.submit(e) for e in bag]
                for i, (rid, e, p) in enumerate(zip(ids, bag, plain)):
                    assert ctx.wait(rid, 120_000) == SP_OK
                    g = ctx.result(rid, e)
                    assert np.array_equal(g["mean"], p["mean"]), (budget, rnd, i, e.problem, e.solver,
                                                                  float(np.abs(g["mean"] - p["mean"]).max()))
    finally:
        ctx.set_iteration_budget(5.0)
    with pytest.raises(SpError):
        ctx.set_iteration_budget(-1.0)


def test_server_admission_by_memory(ctx):
    # budget smaller than two requests: they must run one after the other and still complete
    bag = [neel_ensemble(n=512, n_steps=2000, seed=50 + i) for i in range(4)]
    ctx.set_memory_budget(1, max_active=1)
    try:
        ids = [ctx.submit(e) for e in bag]
        outs = []
        for rid, e in zip(ids, bag):
            assert ctx.wait(rid, 120_000) == SP_OK
            outs.append(ctx.result(rid, e))
    finally:
        ctx.set_memory_budget(8 << 30, max_active=64)
    for e, o in zip(bag, outs):
        same(ctx.run(e), o)


def test_server_cancel(ctx):
    big = neel_ensemble(n=4096, n_steps=400_000, os_=100, seed=60)
    rid = ctx.submit(big)
    time.sleep(0.05)
    assert ctx.cancel(rid) == SP_OK
    rc = ctx.wait(rid, 120_000)
    assert rc != SP_OK
    with pytest.raises(SpError):
        ctx.result(rid, big)
    # the context keeps serving after a cancellation
    small = neel_ensemble(n=64, n_steps=500, seed=61)
    rid2 = ctx.submit(small)
    assert ctx.wait(rid2, 60_000) == SP_OK
    same(ctx.result(rid2, small), ctx.run(small))


def test_whole_pass_and_chunked_reduction_agree(ctx):
    """The plain path reduces a whole pass at once when its partial sums fit the reduction limit, chunk by
    chunk otherwise: the two give the same means and mixed-precision check tiles bit for bit."""
    e = bn_ensemble(n=300, n_steps=6000, solver="Heun_Strong", seed=71)
    m = sph_ensemble(n=700, n_steps=5000, solver="Heun_Strong", seed=72)
    m.precision = "mixed"
    for x in (e, m):
        ctx.set_tuning(x.problem, x.solver, x.precision, block_size=256, chunk_steps=1000,
                       use_graph=1, persistent_waves_per_simd=2)
    try:
        whole = [ctx.run(x) for x in (e, m)]
        ctx.set_reduction_limit(0)
        chunked = [ctx.run(x) for x in (e, m)]
    finally:
        ctx.set_reduction_limit(4 << 30)
    for a, b in zip(whole, chunked):
        same(a, b)


def test_reduction_over_more_rows_than_one_launch_dimension_holds(ctx):
    """64 particles with an output every step: 5.7M windows x 3 components = 1.7e7 reduction rows, more than
    the 16,777,215 blocks of 256 work-items one launch dimension can hold.  Deterministic relaxation (T = 0),
    compared with the oracle; the chunked reduction gives the same sums."""
    e = neel_ensemble(n=64, n_steps=5_700_000, os_=1, T=0.0, seed=73)
    assert e.n_out * 3 > (1 << 24)
    g = ctx.run(e)
    o = run_oracle(e, keep_final=False, threads=16)
    assert np.max(np.abs(g["mean"] - o["mean"])) <= 1e-10
    ctx.set_reduction_limit(0)
    try:
        c = ctx.run(e)
    finally:
        ctx.set_reduction_limit(4 << 30)
    same(g, c)


def test_failed_server_iteration_ends_request_and_shadow(ctx):
    """An iteration that fails ends every request in it, the FP64 shadow of a mixed request included: the
    failed request stays failed (its shadow never publishes partial sums later) and the server goes on."""
    a = sph_ensemble(n=640, n_steps=20_000, solver="Heun_Strong", seed=74)
    a.precision = "mixed"
    ctx.inject_failures(1)
    rid = ctx.submit(a)
    assert ctx.wait(rid, 120_000) != SP_OK
    b = neel_ensemble(n=256, n_steps=3000, solver="Heun_Strong", seed=75)
    rb = ctx.submit(b)
    assert ctx.wait(rb, 120_000) == SP_OK
    same(ctx.result(rb, b), ctx.run(b))
    assert ctx.wait(rid, 0) != SP_OK
    with pytest.raises(SpError):
        ctx.result(rid, a)
    # a failure in the middle of a long mixed request
    c = sph_ensemble(n=640, n_steps=400_000, solver="Heun_Strong", seed=76)
    c.precision = "mixed"
    rc_ = ctx.submit(c)
    time.sleep(0.05)
    ctx.inject_failures(1)
    assert ctx.wait(rc_, 120_000) != SP_OK
    time.sleep(0.2)
    assert ctx.wait(rc_, 0) != SP_OK
    st = ctx.server_stats()
    assert st["active"] == 0
