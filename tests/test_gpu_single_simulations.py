"""Single-simulation outputs (Save_single_simulations / Save_intervall, Results/SimulationResultManager.h:115-149)
on the GPU: every save_interval-th particle's own window means, written by the simulation kernel at each window end.

Bars: the kept trajectories equal the oracle's per-particle window means within 1e-10 (the same FP64 bar as the
ensemble mean), row 0 is the start output exactly, and the plain, batched and served paths write bitwise the same
trajectories (the kernel writes each window once, from the same registers, whatever the schedule)."""
import numpy as np
import pytest

from stochasticphysics_amd import Context, configs
from test_gpu_parity import bn_ensemble, bnes_ensemble, run_oracle, sph_ensemble

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def kept(e):
    return np.arange(e.save_interval - 1, e.n, e.save_interval)


@pytest.mark.parametrize("make", [lambda: bn_ensemble(n=130, n_steps=9000, os_=50),
                                  lambda: bnes_ensemble(n=100, n_steps=3000, os_=100),
                                  lambda: sph_ensemble(n=96, n_steps=3000, os_=100),
                                  lambda: configs.uniaxialcubic(n=100, problem="Neel", n_steps=3000, oversampling=100)])
def test_trajectories_match_oracle(make):
    """The requests run as three or more chunks (1,000 steps each): window rows continue across chunk boundaries.
    An FMA-contracted oracle build differs from the plain one by < 2e-12 on the kept particles of these cases."""
    e = make()
    e.save_interval = 7
    with Context(0) as c:
        c.set_tuning(e.problem, e.solver, "fp64", block_size=256, chunk_steps=1000, use_graph=1,
                     persistent_waves_per_simd=2)
        g = c.run(e, keep_final=True)
    o = run_oracle(e, per_particle=True)
    idx = kept(e)
    assert g["n_saved"] == len(idx) == e.n // 7
    traj = g["trajectories"]                                    # [n_out, dim_out, n_saved]
    assert traj.shape == (e.n_out, e.dim_out, len(idx))
    w = np.stack([e.oracle_props()[:, 1] * 4.0 / 3.0 * np.pi * e.oracle_props()[:, 0] ** 3] * e.dim_out, 1)
    if e.dim_out == 6:
        w[:, :3] = 1.0
    ref = o["per_particle"][idx] / w[idx][:, None, :]           # the oracle keeps weighted outputs
    assert np.max(np.abs(np.transpose(traj, (2, 0, 1)) - ref)) <= 1e-10
    assert np.max(np.abs(g["mean"] - o["mean"])) <= 1e-10


def test_trajectory_row0_is_the_start_output(ctx):
    e = bn_ensemble(n=64, n_steps=200, os_=10)
    e.save_interval = 1
    g = ctx.run(e)
    assert np.array_equal(g["trajectories"][0].T, e.y0)


def test_trajectories_same_on_every_path(ctx):
    """Plain, batched (persistent launch, hipGraph) and served results carry bitwise the same trajectories."""
    sets = [bn_ensemble(n=96, n_steps=3000, os_=50, seed=21), sph_ensemble(n=70, n_steps=3000, os_=100, seed=22),
            configs.uniaxialcubic(n=90, problem="NeelSpherical", n_steps=3000, oversampling=100, seed=23),
            bnes_ensemble(n=80, n_steps=3000, os_=100, seed=24)]
    for k, e in enumerate(sets):
        e.save_interval = 3 + k
    plain = [ctx.run(e) for e in sets]
    batch = ctx.run_batch(sets)
    ids = [ctx.submit(e) for e in sets]
    served = []
    for i, e in zip(ids, sets):
        assert ctx.wait(i, 60000) == 0
        served.append(ctx.result(i, e))
    for p, b, s in zip(plain, batch, served):
        assert np.array_equal(p["trajectories"], b["trajectories"])
        assert np.array_equal(p["trajectories"], s["trajectories"])
        assert np.array_equal(p["mean"], b["mean"]) and np.array_equal(p["mean"], s["mean"])


def test_mixed_request_keeps_its_trajectories(ctx):
    """A mixed-precision request keeps mixed trajectories (its FP64 shadow keeps none); when the check fails (here
    forced by a negative tolerance) the FP64 re-run replaces them with exactly the FP64 plain ones."""
    def make(precision):
        e = sph_ensemble(n=600, n_steps=2000, os_=100, precision=precision)
        e.save_interval = 50
        return e
    ref = ctx.run(make("fp64"))
    mixed = ctx.run_batch([make("mixed")])[0]
    assert mixed["precision_used"] == 1 and mixed["trajectories"].shape == (20, 3, 12)
    assert np.median(np.abs(mixed["trajectories"] - ref["trajectories"])) < 1e-3
    forced = make("mixed")
    forced.check_tol = -1.0
    fb = ctx.run_batch([forced])[0]
    assert fb["used_fallback"] == 1 and fb["precision_used"] == 0
    assert np.array_equal(fb["trajectories"], ref["trajectories"])
