"""GPU parity of the gfx950 kernels against the CPU oracle (run with -m gpu on an MI355X).

Bars (stated in each test):
  * Philox4x32-10: bit-exact vs the published KAT vectors and the oracle's copy.
  * FP64 normals: bit-exact vs the oracle's Box-Muller (same stream definition).
  * FP64 trajectories, replayed or Philox noise: |gpu - oracle| <= 1e-10 (state and <m> are
    O(1) quantities, so this is a relative bar).
  * zero temperature (deterministic LLG): <= 1e-12.
  * mixed precision: max |<m>_mixed - <m>_oracle| <= 1e-3 on the same Philox stream.
"""
import math

import numpy as np
import pytest

import oracle
from stochasticphysics_amd import Context, Ensemble, compute_params, configs, host_box_muller
from stochasticphysics_amd.capi import DIM_W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def run_oracle(e: Ensemble, keep_final=True, threads=8, per_particle=False, contracted=False):
    return oracle.simulate(e.pid, {"EulerMaruyama": 1, "EulerMaruyama_Normalized": 2, "Implicit_Midpoint": 3, "Millstein": 6,
                                   "Heun_Strong": 7,
                                   "Heun_NotConsistent": 8}[e.solver],
                           {"uniaxial": 1, "cubic": 2, "uniaxialcubic": 4}[e.anisotropy], e.n_steps, e.oversampling,
                           e.dt, e.seed,
                           e.oracle_props(), e.y0, axes=e.axes, field=e.field_dict(), bn_simple=e.bn_simple,
                           coord_transform=e.coord_transform, min_angle=e.min_angle, euler_transform=e.euler_transform,
                           brown_min_angle=e.brown_min_angle, newton_tol=e.newton_tol,
                           newton_max_iter=e.newton_max_iter, renormalize=e.renormalize, replay=e.replay, threads=threads,
                           keep_final=keep_final, per_particle=per_particle, contracted=contracted)


# ---------------------------------------------------------------- noise
KAT = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
       ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
       ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]


@pytest.mark.parametrize("ctr,key,exp", KAT)
def test_philox_kat_device(ctx, ctr, key, exp):
    out = ctx.dev_philox(ctr, key, 1)[0]
    assert [int(x) for x in out] == exp


def test_philox_device_vs_oracle(ctx):
    rng = np.random.default_rng(5)
    for _ in range(4):
        ctr = [int(x) for x in rng.integers(0, 2 ** 32, 4)]
        key = [int(x) for x in rng.integers(0, 2 ** 32, 2)]
        dev = ctx.dev_philox(ctr, key, 257)
        for i in (0, 1, 100, 256):
            c = list(ctr)
            c[0] = (c[0] + i) % 2 ** 32
            assert np.array_equal(dev[i], oracle.philox(c, key))


def test_normals_fp64_bit_exact(ctx):
    for seed, particle, first, count in [(1, 0, 0, 4096), (123456789, 77777, 13, 1001), (2 ** 63 + 5, 2 ** 32 - 1, 2 ** 33, 64)]:
        dev = ctx.dev_normals(seed, particle, first, count, "fp64")
        ref = oracle.normals(seed, particle, first, count)
        assert np.array_equal(dev, ref), np.max(np.abs(dev - ref))


def test_normals_fp32_close(ctx):
    dev32 = ctx.dev_normals(9, 4, 0, 1 << 16, "mixed")
    ref = oracle.normals(9, 4, 0, 1 << 16)
    # FP32 hardware log2/sqrt/sin/cos on the same uniforms: ~1e-6 relative, worst in the far tail
    assert np.max(np.abs(dev32 - ref)) < 5e-5
    assert abs(dev32.mean()) < 0.02 and abs(dev32.std() - 1) < 0.02


def test_box_muller_extreme_words(ctx):
    """Every word that puts u0 within 2^-16 of 1 (where u rounds to 1 in FP32) and of 0, plus random words:
    the FP64 transform is bit-identical to the host copy, the FP32 one finite and close to it."""
    hi = np.arange(0xFFFF0000, 0x1_0000_0000, dtype=np.uint64).astype(np.uint32)
    lo = np.arange(0, 0x10000, dtype=np.uint32)
    rnd = np.random.default_rng(8).integers(0, 2 ** 32, (4, 1 << 16), dtype=np.uint64).astype(np.uint32)
    blocks = [np.stack([hi, rnd[0], hi[::-1], rnd[1]], 1), np.stack([lo, rnd[2], lo[::-1], rnd[3]], 1), rnd.T]
    w = np.concatenate(blocks)
    ref = host_box_muller(w)
    assert np.array_equal(ctx.dev_box_muller(w, "fp64"), ref)
    z32 = ctx.dev_box_muller(w, "mixed")
    assert np.all(np.isfinite(z32))
    err = np.abs(z32 - ref)
    # FP32 u = (w + 1/2) 2^-32 carries an absolute rounding error <= 2^-25 near 1, so r^2 = -2 ln u ~ 2 (1 - u)
    # is off by <= 2^-23 and r by <= 2^-11.5 when u -> 1 (the first 2^16 rows); elsewhere ~1e-6 relative
    assert err[:0x10000].max() < 2.0 ** -11.5 + 5e-5
    assert err[0x10000:].max() < 5e-5


# ---------------------------------------------------------------- helpers to build small ensembles
def neel_ensemble(n=96, n_steps=2000, os_=10, solver="EulerMaruyama", aniso="uniaxial", field="none",
                  T=295.0, seed=3, precision="fp64"):
    rng = np.random.default_rng(seed)
    axes = configs.random_axes(rng, n) if aniso == "cubic" else configs.easy_axis_axes(configs.unit_sphere(rng, n))
    e = Ensemble(problem="Neel", solver=solver, field=field, anisotropy=aniso, precision=precision, n_steps=n_steps,
                 oversampling=os_, dt=1e-12, seed=seed, radius=rng.uniform(8e-9, 12e-9, n), Ms=np.full(n, 4.8e5),
                 alpha=rng.uniform(0.05, 0.3, n), gamma=np.full(n, 1.76e11), T=np.full(n, T),
                 K=np.full(n, 1e4 if aniso == "uniaxial" else -1.1e4), axes=axes, y0=configs.unit_sphere(rng, n),
                 field_amplitude=(0.004, 0.007, 0.012), field_freq=(3e8, 1.5e8, 4e8) if field == "lissajous" else (2e8, 0, 0),
                 field_phase=(0.1, 0.2, 0.3), field_offset=(0.001, 0.0, -0.002))
    return e


def bn_ensemble(n=80, n_steps=1000, os_=10, solver="Heun_Strong", simple=False, aniso="uniaxial", field="sinusoidal",
                T=295.0, seed=4, precision="fp64"):
    rng = np.random.default_rng(seed)
    return Ensemble(problem="BrownAndNeel", solver=solver, field=field, anisotropy=aniso, precision=precision,
                    n_steps=n_steps, oversampling=os_, dt=1e-11, seed=seed, radius=rng.uniform(9e-9, 13e-9, n),
                    Ms=np.full(n, 4.8e5), alpha=rng.uniform(0.05, 0.3, n), gamma=np.full(n, 1.76e11), T=np.full(n, T),
                    K=np.full(n, 5e3), r_h=rng.uniform(20e-9, 30e-9, n), viscosity=np.full(n, 1e-3),
                    y0=np.concatenate([configs.unit_sphere(rng, n), configs.unit_sphere(rng, n)], axis=1),
                    bn_simple=simple, field_amplitude=(0.0, 0.003, 0.012), field_freq=(2e7, 0, 0))


def sph_ensemble(n=96, n_steps=2000, os_=10, solver="Heun_Strong", aniso="cubic", field="lissajous", T=300.0, seed=5,
                 rotate=True, precision="fp64"):
    rng = np.random.default_rng(seed)
    m0 = configs.unit_sphere(rng, n)
    m0[:8] = [[0, 0, 1]] * 4 + [[0, 0, -1]] * 4   # start some particles at the poles
    m0[:8] += 0.01 * rng.standard_normal((8, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    return Ensemble(problem="NeelSpherical", solver=solver, field=field, anisotropy=aniso, precision=precision,
                    n_steps=n_steps, oversampling=os_, dt=2e-12, seed=seed, radius=11e-9 * np.exp(0.1 * rng.standard_normal(n)),
                    Ms=np.full(n, 4.8e5), alpha=rng.uniform(0.05, 0.3, n), gamma=np.full(n, 1.76e11), T=np.full(n, T),
                    K=np.full(n, -1.1e4 if aniso == "cubic" else 1e4), axes=configs.random_axes(rng, n),
                    y0=configs.to_spherical(m0), coord_transform=rotate, min_angle=math.pi / 36,
                    field_amplitude=(0.012, 0.012, 0.0), field_freq=(2.5e8, 2.45e8, 0.0))


def bnes_ensemble(n=96, n_steps=2000, os_=10, solver="Heun_Strong", aniso="uniaxial", field="sinusoidal", T=295.0, seed=6,
                  precision="fp64", rotate=True, dt=2e-12):
    """BrownAndNeelEulerSpherical: random body orientations and moments, a few of each near the chart poles."""
    rng = np.random.default_rng(seed)
    eul = np.stack([rng.uniform(0, 2 * math.pi, n), np.arccos(rng.uniform(-1, 1, n)), rng.uniform(0, 2 * math.pi, n)], 1)
    eul[:6, 1] = [0.01, 0.03, 0.05, math.pi - 0.01, math.pi - 0.03, math.pi - 0.05]   # body theta near 0 / pi
    m0 = configs.unit_sphere(rng, n)
    m0[6:12] = [[0, 0, 1]] * 3 + [[0, 0, -1]] * 3                                      # moments near the poles
    m0[6:12] += 0.02 * rng.standard_normal((6, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    e = Ensemble(problem="BrownAndNeelEulerSpherical", solver=solver, field=field, anisotropy=aniso, precision=precision,
                 n_steps=n_steps, oversampling=os_, dt=dt, seed=seed, radius=rng.uniform(9e-9, 13e-9, n),
                 Ms=np.full(n, 4.8e5), alpha=rng.uniform(0.05, 0.3, n), gamma=np.full(n, 1.76e11), T=np.full(n, T),
                 K=np.full(n, {"uniaxial": 5e3, "cubic": -1.1e4, "uniaxialcubic": 6e3}[aniso]),
                 r_h=rng.uniform(20e-9, 30e-9, n), viscosity=np.full(n, 1e-3),
                 y0=np.concatenate([eul, configs.to_spherical(m0)], axis=1), coord_transform=rotate,
                 min_angle=math.pi / 36, euler_transform=rotate, brown_min_angle=math.pi / 36,
                 field_amplitude=(0.0, 0.004, 0.012), field_freq=(2e8, 0, 0))
    if aniso == "uniaxialcubic":
        e.K_cubic = np.full(n, -1.1e4)
        e.uniaxial_rotation = rng.uniform([0, 0, 0], [2 * math.pi, math.pi, 2 * math.pi], (n, 3))
    return e


def with_replay(e: Ensemble, seed=99):
    rng = np.random.default_rng(seed)
    e.replay = rng.standard_normal((e.total_steps * DIM_W[e.pid], e.n))
    return e


def compare(g, o, tol, state=True, spherical=False):
    assert np.max(np.abs(g["mean"] - o["mean"])) <= tol, np.max(np.abs(g["mean"] - o["mean"]))
    assert np.allclose(g["weight"], o["weight"], rtol=1e-13, atol=0)
    assert np.allclose(g["time"], o["times"], rtol=0, atol=0)
    assert np.allclose(g["field"], o["fields"], rtol=1e-12, atol=1e-15)
    if state:
        gs, os_ = g["final_state"], o["final_state"]
        if spherical:   # compare the points on the sphere (angles are chart-dependent near poles)
            def pt(y):
                return np.stack([np.sin(y[:, 0]) * np.cos(y[:, 1]), np.sin(y[:, 0]) * np.sin(y[:, 1]), np.cos(y[:, 0])], 1)
            gs, os_ = pt(gs), pt(os_)
        assert np.max(np.abs(gs - os_)) <= tol, np.max(np.abs(gs - os_))


# ---------------------------------------------------------------- FP64 trajectory parity
@pytest.mark.parametrize("solver", ["EulerMaruyama", "Heun_Strong", "Millstein"])
@pytest.mark.parametrize("aniso", ["uniaxial", "cubic"])
def test_neel_replay_parity(ctx, solver, aniso):
    e = with_replay(neel_ensemble(solver=solver, aniso=aniso, field="lissajous"))
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-10)


@pytest.mark.parametrize("solver", ["EulerMaruyama", "Heun_Strong", "Millstein"])
def test_neel_philox_parity(ctx, solver):
    e = neel_ensemble(solver=solver, field="sinusoidal", n=130)   # ragged: 3 tiles, last one partial
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-10)


@pytest.mark.parametrize("solver,simple", [("EulerMaruyama", False), ("Heun_Strong", False), ("Heun_Strong", True),
                                           ("Millstein", False)])
def test_bn_parity(ctx, solver, simple):
    e = bn_ensemble(solver=solver, simple=simple)
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-10)


def test_bn_cubic_zero_det(ctx):
    """BrownAndNeel + cubic: the reference's deterministic vector is zero (BrownAndNeelRelaxation.inl:225-229)."""
    e = bn_ensemble(aniso="cubic", solver="EulerMaruyama", n=64)
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-10)


@pytest.mark.parametrize("solver", ["EulerMaruyama", "Heun_Strong", "Millstein"])
@pytest.mark.parametrize("aniso", ["uniaxial", "cubic"])
def test_spherical_parity(ctx, solver, aniso):
    e = sph_ensemble(solver=solver, aniso=aniso)
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-10, spherical=True)


@pytest.mark.parametrize("solver", ["Heun_Strong", "Millstein"])
def test_spherical_parity_wide_window(ctx, solver):
    """Chart rotation whenever theta is within pi/4 of a pole (about a third of the steps): the kernels keep a
    rotated state in its chart for the rest of its output window, where the reference maps it back and in again
    at every step (NeelRelaxationSpherical.h:440-471, 568-602); the two agree up to rounding."""
    e = sph_ensemble(solver=solver)
    e.min_angle = math.pi / 4
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-10, spherical=True)


def test_spherical_no_rotation(ctx):
    """Without chart rotation theta is never folded back: particles pass through the poles and theta leaves
    [0, pi] (here one reaches theta < -1).  Near a pole the reference's 1/sin(theta) terms amplify rounding
    differences, so T = 100 K keeps that amplification below 1e-13 (an FMA-contracted oracle build differs
    from the plain one by 9e-14 on this case)."""
    e = sph_ensemble(solver="EulerMaruyama", rotate=False, aniso="uniaxial", field="none", T=100.0)
    rng = np.random.default_rng(3)
    e.y0 = np.stack([rng.uniform(math.pi / 6, 5 * math.pi / 6, e.n), rng.uniform(0, 2 * math.pi, e.n)], 1)
    g, o = ctx.run(e, keep_final=True), run_oracle(e)
    assert (o["final_state"][:, 0] < -1.0).any()
    compare(g, o, 1e-10, spherical=True)


@pytest.mark.parametrize("make", [lambda: neel_ensemble(solver="EulerMaruyama_Normalized", field="sinusoidal", n=130),
                                  lambda: bn_ensemble(solver="EulerMaruyama_Normalized"),
                                  lambda: sph_ensemble(solver="EulerMaruyama_Normalized", n=64)])
def test_euler_maruyama_normalized_parity(ctx, make):
    """EulerMaruyama_Normalized (EulerMaruyama_Normalized.inl:40-55): EulerMaruyama plus normalize() each step;
    the Cartesian states stay on the unit sphere."""
    e = make()
    g, o = ctx.run(e, keep_final=True), run_oracle(e)
    compare(g, o, 1e-10, spherical=e.problem == "NeelSpherical")
    if e.problem != "NeelSpherical":
        for b in range(e.dim_y // 3):
            assert np.allclose(np.linalg.norm(g["final_state"][:, 3 * b:3 * b + 3], axis=1), 1.0, atol=1e-14)


@pytest.mark.parametrize("solver", ["Heun_Strong", "Millstein", "Implicit_Midpoint"])
def test_renormalize_flag_parity(ctx, solver):
    """The renormalize option applies normalize() after every step of any solver, here on BrownAndNeel."""
    e = bn_ensemble(solver=solver, n=72, n_steps=600 if solver != "Implicit_Midpoint" else 200)
    e.renormalize = True
    g, o = ctx.run(e, keep_final=True), run_oracle(e)
    assert g["newton_failures"] == o["newton_failures"]
    compare(g, o, 1e-10)
    assert np.allclose(np.linalg.norm(g["final_state"][:, 3:], axis=1), 1.0, atol=1e-14)


FIELD_CASES = {
    "constant": dict(field_offset=(0.004, -0.003, 0.008)),
    "triangular": dict(field_offset=(0.001, 0.0, 0.0), field_amplitude=(0.0, 0.004, 0.006), field_period=7e-10,
                       field_time_offset=1e-10),
    "rectangular": dict(field_offset=(0.0, 0.0, 0.001), field_amplitude=(0.005, 0.0, 0.008), field_period=8e-10,
                        field_time_offset=5e-11, field_tau=6e-11),
    "rectangular_alternating": dict(field_offset=(0.0, 0.0, 0.001), field_amplitude=(0.005, 0.0, 0.008),
                                    field_period=8e-10, field_time_offset=5e-11, field_tau=6e-11, field_alternating=True),
    "sinc": dict(field_offset=(0.0, 0.001, 0.0), field_amplitude=(0.006, 0.0, 0.009), field_period=1.1e-9),
}


def field_ensemble(name, **kw):
    e = neel_ensemble(solver="Heun_Strong", n=64, n_steps=2000, **kw)
    e.field = name.split("_")[0]
    e.field_offset, e.field_amplitude, e.field_freq, e.field_phase = (0.0,) * 3, (0.0,) * 3, (0.0,) * 3, (0.0,) * 3
    for k, v in FIELD_CASES[name].items():
        setattr(e, k, v)
    return e


@pytest.mark.parametrize("name", list(FIELD_CASES))
def test_field_types_parity(ctx, name):
    """Constant, triangular, rectangular (plain and alternating) and sinc fields (Fields/*Field.h): results and the
    window-mean field axis equal the oracle's; an FMA-contracted oracle build differs from the plain one by < 1e-13."""
    e = field_ensemble(name)
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-10)


@pytest.mark.parametrize("problem", ["Neel", "NeelSpherical", "NeelSpherical_uniaxial", "BrownAndNeel"])
def test_implicit_midpoint_parity(ctx, problem):
    """Implicit_Midpoint (Implicit_Midpoint.inl:38-188).  The NeelSpherical kernels form the Newton Jacobian by the
    chain rule written out, the oracle by dual numbers: the same exact Jacobian, so the iterations stop on the same
    particle-steps.  An FMA-contracted oracle build moves the NeelSpherical means by < 1e-14 (final states < 7e-13)
    and stops on the same particle-steps (uniaxialcubic: test_uniaxialcubic_parity)."""
    if problem == "Neel":
        e = neel_ensemble(solver="Implicit_Midpoint", n=64, n_steps=500, field="sinusoidal")
    elif problem == "NeelSpherical":
        e = sph_ensemble(solver="Implicit_Midpoint", n=64, n_steps=500)
    elif problem == "NeelSpherical_uniaxial":
        e = sph_ensemble(solver="Implicit_Midpoint", n=64, n_steps=500, aniso="uniaxial")
    else:
        e = bn_ensemble(solver="Implicit_Midpoint", n=64, n_steps=300)
    g, o = ctx.run(e, keep_final=True), run_oracle(e)
    # near a chart switch the reference's residual is discontinuous and Newton can stall above 1e-10:
    # the GPU must stall on exactly the same particle-steps as the oracle
    assert g["newton_failures"] == o["newton_failures"]
    compare(g, o, 1e-10, spherical=problem.startswith("NeelSpherical"))



def hnc_replay(e: Ensemble, seed=99):
    """Replayed noise for Heun_NotConsistent: two dim_W blocks of normals per step (dW1, then dW2)."""
    rng = np.random.default_rng(seed)
    e.replay = rng.standard_normal((e.total_steps * 2 * DIM_W[e.pid], e.n))
    return e


HNC = "Heun_NotConsistent"
HNC_CASES = {"neel_uniaxial": lambda: neel_ensemble(solver=HNC, field="sinusoidal", n=130),
             "neel_cubic_replay": lambda: hnc_replay(neel_ensemble(solver=HNC, aniso="cubic", field="lissajous")),
             "spherical_cubic": lambda: sph_ensemble(solver=HNC),
             "spherical_uniaxial": lambda: sph_ensemble(solver=HNC, aniso="uniaxial"),
             "brown_and_neel": lambda: bn_ensemble(solver=HNC),
             "brown_and_neel_simple": lambda: bn_ensemble(solver=HNC, simple=True),
             "brown_and_neel_replay": lambda: hnc_replay(bn_ensemble(solver=HNC, n=70)),
             "euler_spherical_uniaxial": lambda: bnes_ensemble(solver=HNC),
             "euler_spherical_uniaxialcubic": lambda: bnes_ensemble(solver=HNC, aniso="uniaxialcubic"),
             "uniaxialcubic_neel": lambda: configs.uniaxialcubic(n=130, problem="Neel", solver=HNC, n_steps=2000,
                                                                 oversampling=10),
             "uniaxialcubic_spherical": lambda: configs.uniaxialcubic(n=96, problem="NeelSpherical", solver=HNC,
                                                                      n_steps=2000, oversampling=10)}


@pytest.mark.parametrize("case", list(HNC_CASES))
def test_heun_not_consistent_parity(ctx, case):
    """Heun_NotConsistent (Heun_NotConsistent.inl:34-45): predictor on dW1, corrector with the stochastic matrix
    at both points on an independent dW2; every problem class, Philox and replayed noise.  An FMA-contracted oracle
    build differs from the plain one by at most 1.1e-11 (final state) and 1.6e-13 (mean) on these cases."""
    e = HNC_CASES[case]()
    g, o = ctx.run(e, keep_final=True), run_oracle(e)
    if e.problem == "BrownAndNeelEulerSpherical":
        compare_bnes(g, o, 1e-10)
    else:
        compare(g, o, 1e-10, spherical=e.problem == "NeelSpherical")


def test_heun_not_consistent_mixed_request_runs_fp64(ctx):
    """Heun_NotConsistent has no mixed-precision kernels: a mixed request runs the FP64 ones and equals the FP64
    run bit for bit."""
    e = neel_ensemble(solver=HNC, field="sinusoidal", n=130)
    a = ctx.run(e, keep_final=True)
    e.precision = "mixed"
    b = ctx.run(e, keep_final=True)
    assert b["precision_used"] == 0 and not b["used_fallback"]
    assert np.array_equal(a["mean"], b["mean"]) and np.array_equal(a["final_state"], b["final_state"])

UC_CASES = {"neel_em": ("Neel", "EulerMaruyama", 130, 2000), "neel_heun": ("Neel", "Heun_Strong", 130, 2000),
            "neel_millstein": ("Neel", "Millstein", 130, 2000), "neel_im": ("Neel", "Implicit_Midpoint", 64, 500),
            "sph_em": ("NeelSpherical", "EulerMaruyama", 96, 2000), "sph_heun": ("NeelSpherical", "Heun_Strong", 96, 2000),
            "sph_im": ("NeelSpherical", "Implicit_Midpoint", 64, 500)}


@pytest.mark.parametrize("case", list(UC_CASES))
def test_uniaxialcubic_parity(ctx, case):
    """Uniaxialcubic anisotropy (UniaxialCubicAnisotropy.hpp:114): random cube axes, random rotations of the
    uniaxial part, Lissajous drive, Philox noise at 300 K; an FMA-contracted oracle build differs from the plain one
    by at most 6e-12 on these cases, and stalls Newton on the same particle-steps."""
    problem, solver, n, steps = UC_CASES[case]
    e = configs.uniaxialcubic(n=n, problem=problem, solver=solver, n_steps=steps, oversampling=10)
    g, o = ctx.run(e, keep_final=True), run_oracle(e)
    assert g["newton_failures"] == o["newton_failures"]
    compare(g, o, 1e-10, spherical=problem == "NeelSpherical")


def test_uniaxialcubic_mixed_precision(ctx):
    e = configs.uniaxialcubic(n=4096, problem="NeelSpherical", solver="Heun_Strong", n_steps=4000, oversampling=10,
                              precision="mixed")
    g = ctx.run(e)
    o = run_oracle(e, keep_final=False)
    assert g["precision_used"] == 1
    assert np.max(np.abs(g["mean"] - o["mean"])) <= 1e-3


def bnes_points(y):
    """Body x and y axes and the moment of BrownAndNeelEulerSpherical states (angles are chart dependent)."""
    c, s_ = np.cos(y), np.sin(y)
    cphi, cth, cpsi, sphi, sth, spsi = c[:, 0], c[:, 1], c[:, 2], s_[:, 0], s_[:, 1], s_[:, 2]
    return np.stack([cphi * cpsi - cth * sphi * spsi, cth * sphi * cpsi + cphi * spsi, sth * sphi,
                     -sphi * cpsi - cth * cphi * spsi, cth * cphi * cpsi - sphi * spsi, sth * cphi,
                     s_[:, 3] * c[:, 4], s_[:, 3] * s_[:, 4], c[:, 3]], 1)


def compare_bnes(g, o, tol):
    compare(g, o, tol, state=False)
    assert np.max(np.abs(bnes_points(g["final_state"]) - bnes_points(o["final_state"]))) <= tol


@pytest.mark.parametrize("solver", ["EulerMaruyama", "Heun_Strong", "Millstein"])
@pytest.mark.parametrize("aniso", ["uniaxial", "cubic", "uniaxialcubic"])
def test_bnes_parity(ctx, solver, aniso):
    """BrownAndNeelEulerSpherical (BrownAndNeelRelaxationEulerSpherical.h): both charts rotate near their poles
    (theta_min = pi/36, particles start there); an FMA-contracted oracle build differs from the plain one by at most
    7e-12 on these cases.  Ten sines and cosines per stage make them sensitive to the last bit of sin and cos as
    well: the kernels' table-driven sin/cos (which equals the C library's on 98% of arguments) put into the oracle
    in place of the C library's moves the cubic Heun_Strong case by 7e-11 on the CPU, hence 3e-10 here."""
    e = bnes_ensemble(solver=solver, aniso=aniso)
    compare_bnes(ctx.run(e, keep_final=True), run_oracle(e), 3e-10)


@pytest.mark.parametrize("solver", ["Heun_Strong", "Millstein"])
def test_bnes_parity_wide_windows(ctx, solver):
    """Both charts rotate whenever their theta is within pi/4 of a pole: the kernels keep a rotated block in its
    chart for the rest of its output window, where the reference maps it back and in again at every step
    (BrownAndNeelRelaxationEulerSpherical.h:142-169, 635-665); the two agree up to rounding."""
    e = bnes_ensemble(solver=solver)
    e.min_angle = e.brown_min_angle = math.pi / 4
    compare_bnes(ctx.run(e, keep_final=True), run_oracle(e), 3e-10)


def test_bnes_replay_parity(ctx):
    e = with_replay(bnes_ensemble(solver="EulerMaruyama", n=130))   # ragged: 3 tiles
    compare_bnes(ctx.run(e, keep_final=True), run_oracle(e), 1e-10)


def test_bnes_mixed_precision(ctx):
    e = bnes_ensemble(solver="Heun_Strong", n=4096, n_steps=4000, precision="mixed")
    g = ctx.run(e)
    o = run_oracle(e, keep_final=False)
    assert g["precision_used"] == 1
    assert np.max(np.abs(g["mean"] - o["mean"])) <= 1e-3


def stable_t0(solver, n=64, n_steps=5000):
    """T = 0 ensemble whose particles stay in their anisotropy well (m0 within 50 degrees of the easy axis,
    4 mT drive): away from saddle points the deterministic LLG does not amplify rounding differences."""
    e = neel_ensemble(n=n, n_steps=n_steps, os_=100, solver=solver, field="sinusoidal", T=0.0)
    rng = np.random.default_rng(9)
    z = e.axes[:, 6:9]
    t = rng.standard_normal((n, 3))
    t -= (t * z).sum(1, keepdims=True) * z
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    ang = rng.uniform(0.1, 0.87, n)[:, None]
    e.y0 = np.cos(ang) * z + np.sin(ang) * t
    e.field_amplitude, e.field_freq = (0.0, 0.0, 0.004), (1e8, 0.0, 0.0)
    return e


@pytest.mark.parametrize("solver", ["EulerMaruyama", "Heun_Strong", "Millstein"])
def test_zero_temperature_parity(ctx, solver):
    """T = 0: deterministic LLG; GPU and oracle agree to 1e-12."""
    e = stable_t0(solver)
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-12)


def test_zero_temperature_brown_and_neel(ctx):
    """2000 steps: an FMA-contracted oracle build differs from the plain one by 1e-14 here (3.5e-13 at 5000
    steps, where the driven orbit has started to amplify rounding)."""
    e = bn_ensemble(n=64, n_steps=2000, T=0.0)
    compare(ctx.run(e, keep_final=True), run_oracle(e), 1e-12)


def test_zero_temperature_fixed_point(ctx):
    """Config-1 setup at T = 0: m0 on the easy axis stays exactly there."""
    e = configs.config1(n=64, n_steps=2000, oversampling=100)
    e.T = np.zeros(64)
    g = ctx.run(e, keep_final=True)
    assert np.array_equal(g["final_state"], np.tile([0.0, 0.0, 1.0], (64, 1)))
    # the mean is sum(mu m_z) / sum(mu) with the two sums taken in different orders: each carries a relative
    # rounding error of at most (n - 1) u for n = 64 positive terms, so the ratio is 1 within 2 * 63 * 2^-53
    assert np.max(np.abs(g["mean"] - np.tile([0.0, 0.0, 1.0], (e.n_out, 1)))) <= 2 * 63 * 2.0**-53


# ---------------------------------------------------------------- full-noise statistics
def weighted_mean_and_se(per_particle, weights):
    """per_particle [N, n_out, dout], weights [N, dout] -> the weighted ensemble mean [n_out, dout] and its
    standard error from the particles' spread."""
    w = weights[:, None, :]
    mean = (w * per_particle).sum(0) / w.sum(0)
    var = (w * w * (per_particle - mean) ** 2).sum(0) / w.sum(0) ** 2
    return mean, np.sqrt(var)


@pytest.mark.parametrize("make", [lambda s: configs.config1(n=8192, n_steps=6000, oversampling=100, seed=s),
                                  lambda s: configs.config2(n=8192, n_steps=3000, oversampling=50, seed=s),
                                  lambda s: configs.config3(n=8192, n_steps=4000, oversampling=100, seed=s)])
def test_full_noise_ensemble_means_agree(ctx, make):
    """Ensemble means at full noise: the GPU on one Philox stream and the oracle on an independent one agree
    within 3 sigma / sqrt(N) (sigma from both ensembles' spread) in >= 90% of the output windows and within
    4.5 sigma in every window (the windows are correlated in time; 4.5 sigma keeps a false alarm below 1e-3 over
    them).  Config 1 (Neel EM, m0 on the easy axis), config 2 (BrownAndNeel Heun, sinusoidal drive) and config 3
    (NeelSpherical cubic Heun, Lissajous drive) with their particle distributions, 8,192 particles."""
    g_e, o_e = make(101), make(101)     # the same particles ...
    o_e.seed = 303                      # ... and an independent noise stream for the oracle
    g_e.save_interval = 1
    g = ctx.run(g_e)
    o = run_oracle(o_e, keep_final=False, per_particle=True)
    mu = compute_params(g_e)[4]
    w = np.repeat(mu[:, None], g_e.dim_out, axis=1)
    if g_e.problem == "BrownAndNeel":
        w[:, :3] = 1.0
    gm, gse = weighted_mean_and_se(np.transpose(g["trajectories"], (2, 0, 1)), w)
    om, ose = weighted_mean_and_se(o["per_particle"] / w[:, None, :], w)    # the oracle keeps weighted outputs
    assert np.allclose(gm, g["mean"], rtol=0, atol=1e-12)           # the kept trajectories are the mean's inputs
    z = np.abs(gm - om) / np.sqrt(gse ** 2 + ose ** 2 + 1e-300)
    z = z[1:]                           # window 0 is the shared start state
    assert np.mean(z <= 3.0) >= 0.9, np.mean(z <= 3.0)
    assert z.max() <= 4.5, z.max()


# ---------------------------------------------------------------- mixed precision
@pytest.mark.parametrize("make", [lambda: neel_ensemble(solver="Heun_Strong", n=256, n_steps=4000, field="sinusoidal"),
                                  lambda: bn_ensemble(solver="Heun_Strong", n=256, n_steps=2000),
                                  lambda: sph_ensemble(solver="Heun_Strong", n=4096, n_steps=4000),
                                  lambda: sph_ensemble(solver="Millstein", n=4096, n_steps=4000)])
def test_mixed_precision_error(ctx, make):
    """FP32 rounding seeds the same trajectory sensitivity as any perturbation (chart switches, sensitive
    Brown+Neel particles), so a few particles decorrelate; the ensemble mean stays within 1e-3."""
    e = make()
    e.precision = "mixed"
    g = ctx.run(e)
    o = run_oracle(e, keep_final=False)
    err = np.max(np.abs(g["mean"] - o["mean"]))
    assert g["precision_used"] == 1
    assert err <= 1e-3, err
