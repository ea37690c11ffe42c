"""CPU-side checks of the C ABI library (no GPU calls): it loads, exports every
function include/stophys_gpu.h declares, and its host-side pieces agree with the oracle."""
import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
from stochasticphysics_amd import capi, configs, compute_params, host_box_muller, host_normals
from stochasticphysics_amd.capi import lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "stophys_gpu.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    L = lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) <= set(names)


def test_abi_version():
    assert lib().sp_abi_version() == 8


@pytest.mark.parametrize("kind,name,value", [("solver", "EulerMaruyama", 1), ("solver", "Implicit_Midpoint", 3),
                                             ("solver", "Millstein", 6), ("solver", "Heun_Strong", 7),
                                             ("problem", "Neel", 1), ("problem", "NeelSpherical", 2),
                                             ("problem", "BrownAndNeel", 3), ("field", "none", 1),
                                             ("field", "sinusoidal", 3), ("field", "lissajous", 4),
                                             ("anisotropy", "uniaxial", 1), ("anisotropy", "cubic", 2)])
def test_enum_strings(kind, name, value):
    v = C.c_int32()
    assert lib().sp_parse_enum(kind.encode(), name.encode(), C.byref(v)) == 0 and v.value == value
    assert lib().sp_enum_name(kind.encode(), value).decode() == name


def test_unknown_enum_is_an_error():
    v = C.c_int32()
    assert lib().sp_parse_enum(b"solver", b"NoSuchSolver", C.byref(v)) == 1
    assert b"NoSuchSolver" in lib().sp_last_error()


def test_host_normals_equal_oracle():
    for seed, p, first, count in [(0, 0, 0, 1000), (2 ** 40 + 3, 12345, 7, 333)]:
        assert np.array_equal(host_normals(seed, p, first, count), oracle.normals(seed, p, first, count))


EXTREME_WORDS = np.array([[0, 0, 1, 1], [0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFF80, 0x7FFFFFFF],
                          [0x80000000, 0x40000000, 0xC0000000, 0x3FFFFFFF], [0xFFFFFFFE, 0x1, 0xFFFFFF7F, 0xBFFFFFFF]],
                         dtype=np.uint32)


def test_host_box_muller_closed_form():
    """Box-Muller of given words against the closed form sqrt(-2 ln u) (cos, sin)(2 pi u), u = (w + 1/2) 2^-32,
    including the words that give the smallest and the largest u."""
    w = np.concatenate([EXTREME_WORDS, np.random.default_rng(4).integers(0, 2 ** 32, (256, 4), dtype=np.uint64)
                        .astype(np.uint32)])
    z = host_box_muller(w)
    u = (w.astype(np.float64) + 0.5) * 2.0 ** -32
    r0, r1 = np.sqrt(-2 * np.log(u[:, 0])), np.sqrt(-2 * np.log(u[:, 2]))
    ref = np.stack([r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
                    r1 * np.cos(2 * np.pi * u[:, 3]), r1 * np.sin(2 * np.pi * u[:, 3])], 1)
    assert np.all(np.isfinite(z))
    assert np.max(np.abs(z - ref)) <= 1e-14 * np.maximum(1.0, np.abs(ref)).max()


def test_host_box_muller_is_the_stream_transform():
    """normals of a particle's stream = Box-Muller of its Philox blocks (counter (c, 0, particle, 0))."""
    seed, part = 0x1234_5678_9ABC, 77
    words = np.stack([oracle.philox([c, 0, part, 0], [seed & 0xFFFFFFFF, seed >> 32]) for c in range(64)])
    assert np.array_equal(host_box_muller(words).ravel(), host_normals(seed, part, 0, 256))


def table_words():
    """Words that visit every sin/cos table row (top 7 bits) and every ln table row (c_j = 1 + j/128,
    j = -32..64, at several binades), plus both ends of each reduction interval."""
    rng = np.random.default_rng(6)
    top = (np.arange(128, dtype=np.uint64) << 25)
    sc = np.concatenate([top, top | 0x1FFFFFF, top | 0x1000000, top | rng.integers(0, 1 << 25, 128, dtype=np.uint64)])
    lg = []
    for k in (1, 8, 20, 30, 31):
        for j in range(-32, 65):
            for off in (-0.49 / 128, 0.0, 0.49 / 128):
                w = int(round(2.0 ** k * (1.0 + j / 128.0 + off)))
                if 0 <= w < 2 ** 32:
                    lg.append(w)
    lg = np.array(lg + [0, 1, 2, 3, 2 ** 32 - 1, 2 ** 32 - 2, 2 ** 31 - 1, 2 ** 31], dtype=np.uint64)
    n = max(len(sc), len(lg))
    sc = np.resize(sc, n).astype(np.uint32)
    lg = np.resize(lg, n).astype(np.uint32)
    return np.stack([lg, sc, lg[::-1], sc[::-1]], 1)


def test_host_box_muller_equals_oracle_on_every_table_row():
    """The library's literal tables (sp_rng_tables.h) and the oracle's tables computed with the C library give
    bit-identical normals on words that use every row of both tables."""
    w = table_words()
    assert np.array_equal(host_box_muller(w), oracle.box_muller(w))


def test_box_muller_pieces_against_high_precision():
    """ln u against a 40-digit decimal logarithm (relative 4e-16, including u -> 1 where ln u ~ -1e-10), and
    sin/cos(2 pi u) against libm on the exactly reduced angle (absolute 1e-15)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    w = table_words().ravel()
    for x in np.unique(w):
        u = (Decimal(int(x)) + Decimal("0.5")) / Decimal(2 ** 32)
        ref = float(u.ln())
        assert abs(oracle.bm_log(int(x)) - ref) <= 4e-16 * abs(ref), (x, oracle.bm_log(int(x)), ref)
    for x in np.unique(w)[::7]:
        k = int(x)
        s, c = oracle.bm_sincos2pi(k)
        frac = ((k + 0.5) / 2 ** 32) % 1.0              # exact
        ang = 2 * math.pi * (frac if frac < 0.5 else frac - 1.0)
        assert abs(s - math.sin(ang)) <= 1e-15 and abs(c - math.cos(ang)) <= 1e-15, (k, s, c)


def test_particle_params_match_oracle():
    e = configs.config2(n=16, n_steps=200, oversampling=50)
    P = compute_params(e)
    for p in range(4):
        q = oracle.neel_params(e.radius[p], e.Ms[p], e.alpha[p], e.gamma[p], e.T[p])
        b = oracle.brown_params(e.r_h[p], e.viscosity[p], e.T[p])
        assert P[0, p] == q["NeelFactor1"] and P[1, p] == q["NeelFactor2"]
        assert P[2, p] == q["NoisePrefactor"] and P[3, p] == q["DriftPrefactor"]
        mu = e.Ms[p] * 4.0 / 3.0 * math.pi * e.radius[p] ** 3
        assert P[4, p] == pytest.approx(mu, rel=1e-15)
        assert P[7, p] == b["BrownPrefactor"] and P[8, p] == b["Brown_F_Noise"]
        assert P[9, p] == pytest.approx(b["BrownPrefactor"] * mu, rel=1e-15)


def test_uniaxialcubic_params():
    """Uniaxialcubic rows: the cubic part's prefactor -2K_c/Ms in P_K1, -2K_u/Ms in P_KU and the easy axis
    R^T z in P_U (UniaxialCubicAnisotropy.hpp:61-114), R checked against scipy's extrinsic z-x-z rotation."""
    from scipy.spatial.transform import Rotation
    e = configs.uniaxialcubic(n=32)
    P = compute_params(e)
    assert P.shape == (capi.NPARAM, e.n)
    for p in range(e.n):
        assert P[5, p] == -2.0 * e.K_cubic[p] / e.Ms[p] and P[19, p] == -2.0 * e.K[p] / e.Ms[p]
        u = Rotation.from_euler("zxz", e.uniaxial_rotation[p]).as_matrix().T @ e.axes[p].reshape(3, 3)[2]
        assert np.allclose(P[20:23, p], u, rtol=0, atol=1e-15)


def test_brown_and_neel_uniaxialcubic_is_rejected():
    """The reference's BrownAndNeel effective torque is 0/0 for uniaxialcubic (zero Euler sines passed by
    BrownAndNeelRelaxation.inl:236-242 into CubicAnisotropy.h:183-192): the engine refuses the combination."""
    e = configs.uniaxialcubic(n=8)
    b = configs.config2(n=8, n_steps=100, oversampling=50)
    b.anisotropy, b.K_cubic, b.uniaxial_rotation = "uniaxialcubic", e.K_cubic, e.uniaxial_rotation
    r, p, keep = b.to_c()
    out = np.zeros(capi.NPARAM * 8)
    assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(out), None) == 2
    assert "uniaxialcubic" in lib().sp_last_error().decode()
    e.K_cubic = None
    r, p, keep = e.to_c()
    assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(out), None) == 1


def test_bnes_implicit_midpoint_is_rejected():
    """The reference has no Jacobians for BrownAndNeelEulerSpherical (assert(false),
    BrownAndNeelRelaxationEulerSpherical.h:606-671): Implicit_Midpoint is refused for it."""
    import sys
    sys.path.insert(0, str(Path(__file__).parent))
    from test_gpu_parity import bnes_ensemble
    e = bnes_ensemble(n=16, n_steps=100, solver="Implicit_Midpoint")
    r, p, keep = e.to_c()
    assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(np.zeros(capi.NPARAM * 16)), None) == 2
    e.solver = "Heun_Strong"
    r, p, keep = e.to_c()
    assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(np.zeros(capi.NPARAM * 16)), None) == 0


def test_request_validation_errors():
    e = configs.config1(n=8, n_steps=100, oversampling=10)
    r, p, keep = e.to_c()
    r.oversampling = 0
    assert lib().sp_request_shape(C.byref(r), None, None, None, None) == 1
    r.oversampling = 10
    n_out, dy, dw, do = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    assert lib().sp_request_shape(C.byref(r), C.byref(n_out), C.byref(dy), C.byref(dw), C.byref(do)) == 0
    assert (n_out.value, dy.value, dw.value, do.value) == (10, 3, 3, 3)
    out = np.zeros(capi.NPARAM * 8)
    n_par = C.c_uint32()
    r.n_particles = 0  # an empty ensemble is rejected with INVALID_ARG
    assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(out), C.byref(n_par)) == 1
    assert "n_particles" in lib().sp_last_error().decode()
    r.n_particles = 8
    r.solver = 9      # WeakTest: not built for the GPU backend
    assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(out), C.byref(n_par)) == 2
    r.solver = 8      # Heun_NotConsistent
    assert lib().sp_compute_particle_params(C.byref(r), C.byref(p), capi.dptr(out), C.byref(n_par)) == 0
