"""The command-line front-end (reference .ini in, Mean_Results out) and the autotuner, on the GPU."""
import ctypes as C
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from stochasticphysics_amd import Context
from stochasticphysics_amd.capi import Particles, Request, SpError, lib

from test_gpu_parity import neel_ensemble, sph_ensemble

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
INI = ROOT / "tests" / "golden" / "ini"
APP = ROOT / "stochasticphysics_amd" / "stophys_gpu_app"
DIM_Y = {1: 3, 2: 2, 3: 6}


def arr(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


def oracle_of_ini(name):
    """Run the oracle on exactly the request and particles the library's .ini reader produces."""
    h = C.c_void_p()
    assert lib().sp_settings_load(str(INI / name).encode(), C.byref(h)) == 0, lib().sp_last_error()
    try:
        r, p = Request(), Particles()
        assert lib().sp_settings_request(h, C.byref(r), C.byref(p)) == 0
        n, dy = r.n_particles, DIM_Y[r.problem]
        props = np.stack([arr(getattr(p, k), n) for k in ("magnetic_radius", "saturation_magnetisation", "damping",
                                                           "gyromagnetic_ratio", "temperature", "anisotropy_constant",
                                                           "hydrodynamic_radius", "viscosity")], 1)
        if r.anisotropy == 4:   # uniaxialcubic: K_cubic and the uniaxial part's Euler angles
            props = np.concatenate([props, arr(p.anisotropy_constant_cubic, n)[:, None],
                                    arr(p.uniaxial_rotation, 3 * n).reshape(n, 3)], 1)
        field = dict(type=r.field, offset=list(r.field_offset), amp=list(r.field_amplitude), freq=list(r.field_freq),
                     phase=list(r.field_phase))
        return oracle.simulate(r.problem, r.solver, r.anisotropy, r.n_steps, r.oversampling, r.dt, r.seed, props,
                               arr(p.initial_state, dy * n).reshape(n, dy), axes=arr(p.axes, 9 * n).reshape(n, 9),
                               field=field, bn_simple=bool(r.flags & 1), coord_transform=bool(r.flags & 2),
                               min_angle=r.min_angle, newton_tol=r.newton_tol, newton_max_iter=r.newton_max_iter,
                               threads=8, per_particle=True)
    finally:
        lib().sp_settings_free(h)


@pytest.mark.parametrize("name", ["neel_config1.ini", "bn_lissajous.ini"])
def test_app_result_file_matches_oracle(tmp_path, name):
    """stophys_gpu_app with the reference app's options writes Mean_Results equal to the oracle's within 1e-10
    (an FMA-contracted oracle build differs from the plain one by < 1e-13 on both files)."""
    out = tmp_path / "result.json"
    run = subprocess.run([str(APP), f"--parameter_file={INI / name}", "--instruction_set=GFX950",
                          f"--result_file={out}"], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    js = json.loads(out.read_text())["Mean_Results"]
    o = oracle_of_ini(name)
    assert np.max(np.abs(np.array(js["Results"]) - o["mean"])) <= 1e-10
    assert np.array_equal(np.array(js["Time"]), o["times"])
    assert np.allclose(np.array(js["Weight"]), o["weight"], rtol=1e-13, atol=0)


def test_app_uniaxialcubic_with_single_simulations(tmp_path):
    """A NeelSpherical uniaxialcubic parameter file with Save_single_simulations: Mean_Results and every kept
    particle's own window means equal the oracle's within 1e-10 (an FMA-contracted oracle build differs from the plain
    one by < 1e-12 on the mean and < 1e-13 on the kept particles); the single results go to their own file, named
    after the particles."""
    out = tmp_path / "result.json"
    run = subprocess.run([str(APP), f"--parameter_file={INI / 'uc_spherical.ini'}", f"--result_file={out}"],
                         capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert run.returncode == 0, run.stderr
    o = oracle_of_ini("uc_spherical.ini")
    assert np.max(np.abs(np.array(json.loads(out.read_text())["Mean_Results"]["Results"]) - o["mean"])) <= 1e-10
    single = json.loads((tmp_path / "uc_spherical_single.json").read_text())   # Single_simulations_savefile
    assert len(single) == 5                                   # 200 particles, Save_intervall = 40
    for k in range(5):
        p = 40 * (k + 1) - 1
        e = single[f"Single_{p + 1}"]
        w = np.array(e["Weight"])
        assert np.max(np.abs(np.array(e["Results"]) - o["per_particle"][p] / w)) <= 1e-10


def test_tuning_outside_the_kernels_launch_bounds_is_rejected():
    ctx = Context(0)
    try:
        for bad in (dict(block_size=512), dict(block_size=96), dict(chunk_steps=0)):
            kw = dict(block_size=128, chunk_steps=1000, use_graph=1, persistent_waves_per_simd=1)
            kw.update(bad)
            with pytest.raises(SpError):
                ctx.set_tuning("Neel", "Heun_Strong", "fp64", **kw)
    finally:
        ctx.close()


def test_autotune_keeps_results():
    """The autotuner times block size, waves per SIMD and chunk length and stores the best; results computed with
    the stored tuning are bitwise those of the plain path (the reduction order does not depend on tuning)."""
    ctx = Context(0)
    try:
        for e in (neel_ensemble(solver="Heun_Strong", field="sinusoidal", n=700, n_steps=3000),
                  sph_ensemble(solver="Millstein", n=300, n_steps=2000)):
            plain = ctx.run(e)
            tun, rate = ctx.autotune(e)
            assert rate > 0 and tun["block_size"] in (64, 128, 256) and tun["chunk_steps"] > 0
            batch = ctx.run_batch([e])[0]
            assert np.array_equal(batch["mean"], plain["mean"])
    finally:
        ctx.close()
