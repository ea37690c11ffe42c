"""The device-side particle provider (sp_settings_prepare / sp_settings_submit) on the GPU.

Bars: the particle type of every particle and every type constant bit-exact; drawn properties, parameter rows
and start states within a few ulp of the host materialisation (the device math library's exp, sin, cos, acos and
atan2 may round differently from the C library's by an ulp); the ensemble run from the device-drawn particles
within 1e-10 of the oracle run from the oracle's own restatement of the provider (an FMA-contracted oracle build,
or a one-ulp change of every radius, moves these runs by < 2e-13)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from provider_cases import CASES, host_particles, write_run
from stochasticphysics_amd import Context, Settings, capi
from stochasticphysics_amd.capi import Particles, Request, lib
from test_gpu_parity import compare, compare_bnes

pytestmark = pytest.mark.gpu

FIELD = dict(offset=(0.0, 0.0, 0.001), amp=(0.0, 0.0, 0.01), freq=2e8)
ORACLE_FIELD = dict(type=3, offset=[0.0, 0.0, 0.001], amp=[0.0, 0.0, 0.01], freq=[2e8, 0, 0], phase=[0, 0, 0])


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def host_params(path):
    """sp_compute_particle_params on the host materialisation of the file."""
    h = C.c_void_p()
    assert lib().sp_settings_load(str(path).encode(), C.byref(h)) == 0
    try:
        r, p = Request(), Particles()
        assert lib().sp_settings_request(h, C.byref(r), C.byref(p)) == 0
        out = np.zeros(capi.NPARAM * r.n_particles)
        assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(out), None) == 0
        return out.reshape(capi.NPARAM, r.n_particles)
    finally:
        lib().sp_settings_free(h)


UNIT_ROWS = list(range(capi.P_AX, capi.P_AX + 9)) + list(range(capi.P_U, capi.P_U + 3))   # axes, easy axis R^T z


def check_params(params, ref):
    """Scalar rows within 2e-15 relative, the rows holding unit-vector components within 2e-15 absolute."""
    scalar = [k for k in range(capi.NPARAM) if k not in UNIT_ROWS]
    rel = np.abs(params[scalar] - ref[scalar]) / np.maximum(np.abs(ref[scalar]), 1e-300)
    rel[(params[scalar] == 0) & (ref[scalar] == 0)] = 0.0
    assert np.max(rel) <= 2e-15, np.max(rel)
    assert np.max(np.abs(params[UNIT_ROWS] - ref[UNIT_ROWS])) <= 2e-15


def oracle_run(problem, aniso, types, n, seed, discrete, n_steps=1000, os_=50, dt=1e-12):
    o = oracle.provide(problem, aniso, types, seed, n, discrete=discrete)
    props = o["props"] if aniso == 4 else o["props"][:, :8]
    return o, oracle.simulate(problem, 7, aniso, n_steps, os_, dt, seed, props, o["y0"], axes=o["axes"],
                              field=ORACLE_FIELD, coord_transform=problem in (2, 4), min_angle=math.pi / 36,
                              euler_transform=problem == 4, brown_min_angle=math.pi / 36, threads=8, keep_final=True)


@pytest.mark.parametrize("name", list(CASES))
def test_device_particles_equal_host_materialisation(ctx, tmp_path, name):
    problem, aniso, discrete, types = CASES[name]
    n = 700                                            # ragged: 2.7 work-groups of 256, 10.9 tiles of 64
    path = write_run(tmp_path, name, problem, aniso, types, n, seed=12345, discrete=discrete, field=FIELD)
    s = Settings(path)
    h = ctx.prepare_settings(s)
    params, y0 = h.inputs()
    ref = host_params(path)
    r, parts = host_particles(path)
    # the rows built from type constants only (damping, gyromagnetic ratio) bit-exact
    assert np.array_equal(params[capi.P_A], ref[capi.P_A]) and np.array_equal(params[capi.P_B], ref[capi.P_B])
    check_params(params, ref)
    assert np.max(np.abs(y0 - parts["y0"])) <= 4e-15


@pytest.mark.parametrize("name", list(CASES))
def test_device_provided_run_matches_oracle(ctx, tmp_path, name):
    problem, aniso, discrete, types = CASES[name]
    n = 700
    path = write_run(tmp_path, name, problem, aniso, types, n, seed=12345, discrete=discrete, field=FIELD)
    s = Settings(path)
    h = ctx.prepare_settings(s)
    g = h.execute(keep_final=True)
    o, ref = oracle_run(problem, aniso, types, n, 12345, discrete)
    if problem == 4:
        compare_bnes(g, ref, 1e-10)
    else:
        compare(g, ref, 1e-10, spherical=problem == 2)
    # Mean_Properties: T, radius, Ms, alpha, gamma, K, r_h, viscosity over the particles
    means = o["props"][:, [4, 0, 1, 2, 3, 5, 6, 7]].mean(0)
    assert np.allclose(g["particle_means"], means, rtol=1e-13, atol=0)
    again = h.execute(keep_final=True)                 # the resident inputs restart every execution
    assert np.array_equal(again["mean"], g["mean"]) and np.array_equal(again["final_state"], g["final_state"])


def test_server_path_and_mixed_shadow(ctx, tmp_path):
    """sp_settings_submit: an FP64 request equals the prepared path bit for bit; a mixed request gets an FP64
    shadow over its first particles, drawn again on the device, and stays within 1e-3 of the oracle."""
    problem, aniso, discrete, types = CASES["neel_lognormal_sequential"]
    n = 4100
    path = write_run(tmp_path, "fp64", problem, aniso, types, n, seed=7, discrete=discrete, field=FIELD, n_steps=2000)
    s = Settings(path)
    a = ctx.prepare_settings(s).execute()
    rid = ctx.submit_settings(s)
    assert ctx.wait(rid, 120_000) == 0
    b = ctx.settings_result(rid, s)
    assert np.array_equal(a["mean"], b["mean"]) and np.array_equal(a["weight"], b["weight"])
    assert np.array_equal(a["particle_means"], b["particle_means"])
    path = write_run(tmp_path, "mixed", problem, aniso, types, n, seed=7, discrete=discrete, field=FIELD, n_steps=2000,
                     precision="mixed")
    sm = Settings(path)
    rid = ctx.submit_settings(sm)
    assert ctx.wait(rid, 120_000) == 0
    m = ctx.settings_result(rid, sm)
    assert m["precision_used"] == 1 and m["check_error"] >= 0.0 and not m["used_fallback"]
    o, ref = oracle_run(problem, aniso, types, n, 7, discrete, n_steps=2000)
    assert np.max(np.abs(m["mean"] - ref["mean"])) <= 1e-3


def test_large_ensemble(ctx, tmp_path):
    """2^20 + 37 particles of two types with lognormal radii: every parameter row of the device draw within the
    bars above of the host materialisation."""
    problem, aniso, discrete, types = CASES["brown_and_neel_shell"]
    n = (1 << 20) + 37
    path = write_run(tmp_path, "large", problem, aniso, types, n, seed=99, discrete=True, field=FIELD, n_steps=100,
                     oversampling=50)
    h = ctx.prepare_settings(Settings(path))
    params, y0 = h.inputs()
    ref = host_params(path)
    check_params(params, ref)
    r, parts = host_particles(path)
    assert np.max(np.abs(y0 - parts["y0"])) <= 4e-15
