"""The benchmark configurations themselves on the GPU, scaled to sizes the oracle runs in seconds.

* the 512-set batching distribution (uneven Neel EulerMaruyama / Heun_Strong and BrownAndNeel Heun_Strong sets):
  a draw of its first sets, with their own parameters and fewer particles and steps, through the batched
  persistent path in FP64 and in mixed precision;
* the serving distribution (NeelSpherical cubic Lissajous, damping U[0.01, 0.5], Implicit_Midpoint for damping
  >= 0.3): a Poisson burst through sp_submit;
* the precision-study ensemble (NeelSpherical cubic, log-normal radii, Lissajous, chart rotation), Heun_Strong and
  Millstein, FP64 and mixed.

FP64 results must equal the oracle on the same Philox stream within max(1e-10, 10 x the case's rounding floor),
the floor being the distance between the oracle and its FMA-contracted build (some sets amplify a last-bit
difference to 1e-10 within a few thousand steps).  Mixed results are compared with the FP64 oracle particle by
particle: the difference of the weighted means must stay within 1e-3 plus 4.5 standard errors of the per-particle
differences in every window (particles that decorrelate under FP32 rounding leave Monte-Carlo noise, not bias),
and its average over the run within 1e-3 plus 4.5 standard errors of the per-particle averages.
"""
import time

import numpy as np
import pytest

from stochasticphysics_amd import Context, compute_params, configs
from stochasticphysics_amd.capi import SP_OK

from test_gpu_parity import run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def weights_of(e):
    mu = compute_params(e)[4]
    w = np.repeat(mu[:, None], e.dim_out, axis=1)
    if e.problem == "BrownAndNeel":
        w[:, :3] = 1.0
    return w


def assert_fp64_parity(e, g, tag=""):
    o = run_oracle(e, keep_final=False)
    oc = run_oracle(e, keep_final=False, contracted=True)
    floor = float(np.max(np.abs(o["mean"] - oc["mean"])))
    err = float(np.max(np.abs(g["mean"] - o["mean"])))
    assert err <= max(1e-10, 10 * floor), (tag, err, floor)
    return err, floor


def assert_mixed_unbiased(e, g, tag=""):
    """g: a mixed result with per-particle window means (save_interval = 1)."""
    assert g["precision_used"] == 1, tag
    o = run_oracle(e, keep_final=False, per_particle=True)
    w = weights_of(e)
    gp = np.transpose(g["trajectories"], (2, 0, 1))                  # [N, n_out, dout]
    op = o["per_particle"] / w[:, None, :]                             # the oracle keeps weighted outputs
    d = (gp - op)[:, 1:, :]                                            # window 0 is the shared start state
    W = w.sum(0)
    D = (w[:, None, :] * d).sum(0) / W
    se = np.sqrt((w[:, None, :] ** 2 * (d - D) ** 2).sum(0)) / W
    assert np.all(np.abs(D) <= 1e-3 + 4.5 * se), (tag, np.abs(D).max(), se.max())
    dbar = d.mean(axis=1)                                              # per-particle time averages [N, dout]
    B = (w * dbar).sum(0) / W
    seb = np.sqrt((w ** 2 * (dbar - B) ** 2).sum(0)) / W
    assert np.all(np.abs(B) <= 1e-3 + 4.5 * seb), (tag, B, seb)
    assert np.allclose(g["mean"], o["mean"], atol=1e-3 + 4.5 * float(se.max()))
    return float(np.abs(D).max()), float(np.abs(B).max())


def config4_draw(precision, n_sets=16):
    """The first sets of the 512-set draw (same seeds and parameters), 64..512 particles, 2,000..6,000 steps."""
    return configs.config4_sets(n_sets=n_sets, seed=44, precision=precision, particle_range=(64, 512),
                                step_range=(2000, 6000))


def test_config4_draw_fp64_batched_matches_oracle(ctx):
    sets = config4_draw("fp64")
    res = ctx.run_batch(sets)
    for i, (e, g) in enumerate(zip(sets, res)):
        assert_fp64_parity(e, g, tag=(i, e.problem, e.solver, e.n, e.n_steps))


def test_config4_draw_mixed_batched_is_unbiased(ctx):
    sets = config4_draw("mixed")
    for e in sets:
        e.save_interval = 1
    res = ctx.run_batch(sets)
    for i, (e, g) in enumerate(zip(sets, res)):
        assert g["check_error"] >= 0.0 and g["check_noise"] >= 0.0, i
        if g["used_fallback"]:             # re-run in FP64: then it is the FP64 result
            assert_fp64_parity(e, g, tag=i)
        else:
            assert_mixed_unbiased(e, g, tag=(i, e.problem, e.solver, e.n, e.n_steps))


def test_serving_burst_matches_oracle(ctx):
    """24 requests of the serving distribution at 200 req/s (Poisson), 64..512 particles, 1,000..3,000 steps; every
    result against the oracle, and each equal to the same request run alone on the plain path."""
    rng = np.random.default_rng(55)
    n_req = 24
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, n_req))
    reqs = [configs.serving_request(rng, i, n_range=(64, 512), step_range=(1000, 3000)) for i in range(n_req)]
    for e in reqs:
        e.save_interval = 1
    assert any(e.solver == "Implicit_Midpoint" for e in reqs) and any(e.solver == "Heun_Strong" for e in reqs)
    ids = []
    t0 = time.perf_counter()
    for t, e in zip(arrivals, reqs):
        delay = t - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(ctx.submit(e))
    outs = []
    for rid, e in zip(ids, reqs):
        assert ctx.wait(rid, 120_000) == SP_OK
        outs.append(ctx.result(rid, e))
    for i, (e, g) in enumerate(zip(reqs, outs)):
        tag = (i, e.solver, e.n, e.n_steps)
        if e.solver == "Implicit_Midpoint" or g["used_fallback"]:
            assert g["precision_used"] == 0, tag
            assert_fp64_parity(e, g, tag=tag)
        else:
            assert_mixed_unbiased(e, g, tag=tag)
        if not g["used_fallback"]:
            assert np.array_equal(ctx.run(e)["mean"], g["mean"]), tag


@pytest.mark.parametrize("solver", ["Heun_Strong", "Millstein"])
def test_config3_ensemble_fp64_and_mixed(ctx, solver):
    """The precision-study ensemble itself (its radii, cube orientations, start states, field and chart rotation),
    4,096 particles, 3,000 steps."""
    e = configs.config3(n=4096, n_steps=3000, solver=solver)
    g = ctx.run(e)
    assert_fp64_parity(e, g, tag=solver)
    m = configs.config3(n=4096, n_steps=3000, solver=solver, precision="mixed")
    m.save_interval = 1
    gm = ctx.run(m)
    assert_mixed_unbiased(m, gm, tag=solver)
