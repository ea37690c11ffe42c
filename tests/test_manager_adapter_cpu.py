"""The SimulationManager<prec, InstructionSet::GFX950> specialisation (include/stophys_gpu_simulation_manager.hpp)
that the reference's runAppWithArch<GFX950> instantiates (app/StoPhysApp_MultiArch.cpp:27-38), compiled against a
test double of the reference's InstructionSet enumeration, primary template and file archive
(tests/manager_adapter_check.cpp): constructed from a settings object, it must hold the request that the same
parameter file gives when it is read directly.  CPU only (the device context opens at StartSimulationManager)."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from stochasticphysics_amd import Settings
from stochasticphysics_amd.build import LIB
from stochasticphysics_amd.capi import Particles, Request, check, lib

ROOT = Path(__file__).resolve().parent.parent
INI = ROOT / "tests" / "golden" / "ini"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx or not LIB.exists():
        pytest.skip("g++ or the library is missing")
    exe = tmp_path_factory.mktemp("adapter") / "manager_adapter_check"
    subprocess.run([cxx, "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "manager_adapter_check.cpp"),
                    "-o", str(exe), f"-L{LIB.parent}", "-lstophys_gpu", f"-Wl,-rpath,{LIB.parent}",
                    "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"], check=True, capture_output=True)
    return exe


def fields(r):
    return [r.problem, r.solver, r.field, r.anisotropy, r.precision, r.n_particles, r.n_steps, r.oversampling, r.seed,
            float(r.dt), r.flags]


def run(exe, path, cwd):
    out = subprocess.run([str(exe), str(path)], cwd=cwd, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    v = out.stdout.split()
    return [int(x) for x in v[:9]] + [float(v[9]), int(v[10])], v[11]


@pytest.mark.parametrize("name", ["bn_lissajous.ini", "neel_config1.ini", "uc_spherical.ini"])
def test_specialisation_reads_the_settings_like_the_parameter_file(checker, name):
    # the settings are written to a temporary file elsewhere: their particle files resolve relative to the working
    # directory, as the reference opens them (ParticleProvider.h:275-277)
    got, result_file = run(checker, INI / name, cwd=INI)
    assert got == fields(Settings(INI / name).request)


def test_particles_embedded_in_the_parameter_file(checker, tmp_path, monkeypatch):
    """An empty Particle_List entry takes the particle type from the sections under its name in the parameter file
    itself (ParticleProvider.h:276-280)."""
    run_text = (INI / "bn_lissajous.ini").read_text().replace("Small = 'bn_small.ini'", "Small = ''")
    small = (INI / "bn_small.ini").read_text()
    embedded = "\n".join(f"[Small.{line[1:]}" if line.startswith("[") else line for line in small.splitlines())
    path = tmp_path / "embedded.ini"
    path.write_text(run_text + "\n" + embedded + "\n")
    got, _ = run(checker, path, cwd=INI)
    assert got == fields(Settings(INI / "bn_lissajous.ini").request)
    monkeypatch.chdir(INI)                     # the other particle type's file is relative to the working directory
    a = Settings(path)
    b = Settings(INI / "bn_lissajous.ini")
    assert fields(a.request) == fields(b.request)
    pa, pb = particles(a), particles(b)                 # the same particles, drawn from the same types
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def particles(settings):
    req, parts = Request(), Particles()
    check(lib().sp_settings_request(settings._h, C.byref(req), C.byref(parts)))
    n = req.n_particles
    return {k: np.ctypeslib.as_array(getattr(parts, k), shape=(n,)).copy()
            for k in ("magnetic_radius", "damping", "anisotropy_constant", "hydrodynamic_radius", "temperature")}
