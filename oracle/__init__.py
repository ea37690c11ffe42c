"""CPU oracle of the StochasticPhysics hot path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/liboracle.so`` (built from sp_oracle.hpp,
sp_oracle_sim.hpp and oracle_capi.cpp by ``oracle/Makefile``).  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
as the checker.  The product package ``stochasticphysics_amd`` never imports it.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
_LIBS = {}

DIM_Y = {1: 3, 2: 2, 3: 6, 4: 5}
DIM_W = {1: 3, 2: 3, 3: 6, 4: 6}
DIM_OUT = {1: 3, 2: 3, 3: 6, 4: 9}


def build(force=False):
    """Compile the oracle with its Makefile (g++)."""
    if force or not all((_DIR / n).exists() for n in ("liboracle.so", "liboracle_native.so", "liboracle_fma.so")):
        subprocess.run(["make", "-C", str(_DIR), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)


def lib(native=False, contracted=False):
    """native: the -O3 -march=native build (CPU baseline); contracted: the FMA-contracted build (rounding floor)."""
    name = "liboracle_native.so" if native else ("liboracle_fma.so" if contracted else "liboracle.so")
    if name not in _LIBS:
        path = _DIR / name
        if not path.exists():
            build()
        _LIBS[name] = _bind(C.CDLL(str(path)))
    return _LIBS[name]


def _bind(L):
    """Argument and result types of the oracle's C entry points on a loaded library."""
    d, u64, i64 = C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_int64)
    L.or_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
    L.or_normals.argtypes = [u64, u64, u64, u64, C.c_int, d]
    L.or_bm_log.argtypes = [C.c_uint32]
    L.or_bm_log.restype = C.c_double
    L.or_bm_sincos2pi.argtypes = [C.c_uint32, d, d]
    L.or_box_muller.argtypes = [C.POINTER(C.c_uint32), u64, d]
    L.or_neel_params.argtypes = [C.c_double] * 5 + [d]
    L.or_brown_params.argtypes = [C.c_double] * 3 + [d]
    L.or_aniso_field.argtypes = [C.c_int, d, d, d, d, d]
    L.or_euler_axes.argtypes = [C.c_double] * 3 + [d]
    L.or_field.argtypes = [d, C.c_double, d]
    L.or_eval.argtypes = [C.c_int, C.c_int, d, d, C.c_int, C.c_double, d, d, C.c_double, C.c_int, d]
    L.or_eval.restype = C.c_int
    L.or_newton_testfn.argtypes = [C.c_int, d, C.c_double, C.c_int, d]
    L.or_newton_testfn.restype = C.c_int
    L.or_simulate.argtypes = [i64, C.POINTER(C.c_uint64), d, d, d, d, d, d, d, d, d, d, d, d]
    L.or_provide.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, u64, u64, d, d, d, d, i64]
    L.or_provide.restype = None
    L.or_simulate.restype = C.c_int
    return L


class use_library:
    """Context manager: the oracle functions of this module call the shared library at `path` (a copy of the
    oracle built elsewhere, e.g. with a mistake planted on purpose by a test) instead of liboracle.so."""

    def __init__(self, path):
        self.path = str(path)

    def __enter__(self):
        self.saved = dict(_LIBS)
        _LIBS.clear()
        _LIBS["liboracle.so"] = _bind(C.CDLL(self.path))
        return self

    def __exit__(self, *a):
        _LIBS.clear()
        _LIBS.update(self.saved)


def _p(a, ct=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def normals(seed, particle, first, count, libm=False):
    out = np.zeros(count)
    lib().or_normals(seed, particle, first, count, int(libm), _p(out))
    return out


def bm_log(w):
    """ln u of the engine's Box-Muller, u = (w + 1/2) 2^-32, for a 32-bit Philox word w."""
    return lib().or_bm_log(int(w))


def bm_sincos2pi(w):
    s, c = C.c_double(), C.c_double()
    lib().or_bm_sincos2pi(int(w), C.byref(s), C.byref(c))
    return s.value, c.value


def box_muller(words):
    """[n, 4] Philox words -> [n, 4] standard normals (the engine's FP64 transform)."""
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1, 4)
    out = np.zeros(w.shape)
    lib().or_box_muller(_p(w, C.c_uint32), w.shape[0], _p(out))
    return out


def field_at(field, t):
    """The oracle's applied field H(t) [T] for a field dict as simulate() takes it."""
    d = np.zeros(20)
    d[3] = field.get("type", 1)
    for j, key in enumerate(("offset", "amp", "freq", "phase")):
        d[4 + 3 * j:7 + 3 * j] = np.asarray(field.get(key, [0.0, 0.0, 0.0]), dtype=np.float64)
    d[16], d[17], d[18] = field.get("period", 0.0), field.get("time_offset", 0.0), field.get("tau", 0.0)
    d[19] = 1.0 if field.get("alternating", False) else 0.0
    out = np.zeros(3)
    lib().or_field(_p(d), float(t), _p(out))
    return out


def neel_params(radius, Ms, alpha, gamma, T):
    out = np.zeros(5)
    lib().or_neel_params(radius, Ms, alpha, gamma, T, _p(out))
    return dict(NeelFactor1=out[0], NeelFactor2=out[1], DriftPrefactor=out[2], NoisePrefactor=out[3], Damping=out[4])


def brown_params(r_h, viscosity, T):
    out = np.zeros(3)
    lib().or_brown_params(r_h, viscosity, T, _p(out))
    return dict(BrownPrefactor=out[0], BrownDiffusion=out[1], Brown_F_Noise=out[2])


def props_array(radius, Ms, alpha, gamma, T, K, r_h=0.0, viscosity=0.0):
    return np.array([radius, Ms, alpha, gamma, T, K, r_h, viscosity], dtype=np.float64)


def aniso_field(aniso, props, m, axes=None):
    out, jac = np.zeros(3), np.zeros(9)
    ax = None if axes is None else np.ascontiguousarray(axes, dtype=np.float64).ravel()
    lib().or_aniso_field(aniso, _p(np.asarray(props, float)), _p(np.asarray(m, float)), _p(ax), _p(out), _p(jac))
    return out, jac.reshape(3, 3)


def euler_axes(phi, theta, psi):
    out = np.zeros(9)
    lib().or_euler_axes(phi, theta, psi, _p(out))
    return out.reshape(3, 3)


WHAT = dict(det=0, drift=1, sto=2, jacdet_ref=3, jacsto_ref=4, finish=5, rotated=6, output=7)


def evaluate(problem, aniso, props, what, y, h=None, axes=None, simple=False, rotate=False, min_angle=0.0, dt=0.0,
             euler_rotate=False, brown_min_angle=0.0):
    """One problem term at state y.  BrownAndNeelEulerSpherical (problem 4): rotate / min_angle are the Neel chart
    settings, euler_rotate / brown_min_angle the Brown ones; 'finish' returns the prepared state then the state after
    finishCalculations (10 values), 'rotated' the (Brown, Neel) flags."""
    dy = DIM_Y[problem]
    bnes = problem == 4
    size = {0: dy, 1: dy, 2: dy * DIM_W[problem], 3: 9, 4: 9, 5: 10 if bnes else 2, 6: 2 if bnes else 1, 7: 9}[WHAT[what]]
    out = np.zeros(size)
    y = np.asarray(y, dtype=np.float64)
    hh = None if h is None else np.asarray(h, dtype=np.float64)
    ax = None if axes is None else np.ascontiguousarray(axes, dtype=np.float64).ravel()
    flags = (1 if simple else 0) | (2 if rotate else 0) | (8 if euler_rotate else 0)
    if bnes:
        dt = brown_min_angle
    rc = lib().or_eval(problem, aniso, _p(np.asarray(props, float)), _p(ax), flags, min_angle, _p(y), _p(hh), dt,
                       WHAT[what], _p(out))
    if rc:
        raise ValueError(f"or_eval({problem},{aniso},{what}) -> {rc}")
    return out


def newton_testfn(fn, x0, tol, max_iter):
    x0 = np.asarray(x0, dtype=np.float64)
    out = np.zeros_like(x0)
    it = lib().or_newton_testfn(fn, _p(x0), tol, max_iter, _p(out))
    return out, it


def simulate(problem, solver, aniso, n_steps, oversampling, dt, seed, props, y0, axes=None, field=None,
             bn_simple=False, coord_transform=False, min_angle=0.0, newton_tol=1e-10, newton_max_iter=20, renormalize=False,
             replay=None, threads=1, keep_final=False, per_particle=False, native=False, euler_transform=False,
             brown_min_angle=0.0, contracted=False):
    """Run the oracle ensemble.  props: [N,8] (radius, Ms, alpha, gamma, T, K, r_h, viscosity; [N,12] with K_cubic and
    the uniaxial rotation for uniaxialcubic); BrownAndNeelEulerSpherical: coord_transform / min_angle for its Neel part,
    euler_transform / brown_min_angle for its Brown part;
    y0: [N, dim_y]; axes: [N, 9] or None; field: dict(type, offset, amp, freq, phase) or None;
    replay: [K, N] standard normals or None; contracted: run the FMA-contracted build (the same arithmetic rounded
    differently)."""
    props = np.ascontiguousarray(props, dtype=np.float64)
    N = props.shape[0]
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    dy, dout = DIM_Y[problem], DIM_OUT[problem]
    n_out = n_steps // oversampling
    f = field or dict(type=1)
    dspec = np.zeros(21)
    dspec[0], dspec[1], dspec[2], dspec[3] = dt, min_angle, newton_tol, f.get("type", 1)
    dspec[20] = brown_min_angle
    for j, key in enumerate(("offset", "amp", "freq", "phase")):
        dspec[4 + 3 * j:7 + 3 * j] = np.asarray(f.get(key, [0.0, 0.0, 0.0]), dtype=np.float64)
    dspec[16], dspec[17], dspec[18] = f.get("period", 0.0), f.get("time_offset", 0.0), f.get("tau", 0.0)
    dspec[19] = 1.0 if f.get("alternating", False) else 0.0
    flags = (1 if bn_simple else 0) | (2 if coord_transform else 0) | (4 if renormalize else 0) | (8 if euler_transform else 0)
    ispec = np.array([problem, solver, aniso, flags, newton_max_iter, threads], dtype=np.int64)
    rep = None if replay is None else np.ascontiguousarray(replay, dtype=np.float64)
    uspec = np.array([N, n_steps, oversampling, seed, 0 if rep is None else rep.shape[0]], dtype=np.uint64)
    mean, weight = np.zeros(n_out * dout), np.zeros(dout)
    times, fields = np.zeros(n_out), np.zeros(n_out * 3)
    final = np.zeros(N * dy) if keep_final else None
    pp = np.zeros(N * n_out * dout) if per_particle else None
    stats = np.zeros(2)
    ax = None if axes is None else np.ascontiguousarray(axes, dtype=np.float64)
    rc = lib(native, contracted).or_simulate(_p(ispec, C.c_int64), _p(uspec, C.c_uint64), _p(dspec), _p(props), _p(ax), _p(y0),
                                 _p(rep), _p(mean), _p(weight), _p(times), _p(fields), _p(final), _p(pp), _p(stats))
    if rc:
        raise ValueError(f"or_simulate -> {rc}")
    res = dict(mean=mean.reshape(n_out, dout), weight=weight, times=times, fields=fields.reshape(n_out, 3),
               seconds=stats[0], newton_failures=int(stats[1]))
    if keep_final:
        res["final_state"] = final.reshape(N, dy)
    if per_particle:
        res["per_particle"] = pp.reshape(N, n_out, dout)
    return res


DIST_CODE = {"delta": 0, "normal": 1, "lognormal": 2}


def provide(problem, aniso, types, seed, n, discrete=False):
    """The particles of an ensemble from its particle types (or_provide).  types: list of dicts with count, T,
    radius, Ms, alpha, gamma, K, r_h, viscosity and optionally K_cubic, euler, mag, urot, rand_orient, rand_mag,
    rand_urot and distributions rad / shell / K / K_cubic given as (type, width, relative).  Returns props [n, 12]
    (the layout of simulate), axes [n, 9], y0 [n, dim_y] and the type index of every particle."""
    rows = []
    for t in types:
        row = [t["count"], t["T"], t["radius"], t["Ms"], t["alpha"], t["gamma"], t["K"], t.get("r_h", 0.0),
               t.get("viscosity", 0.0), t.get("K_cubic", 0.0), *t.get("euler", (0.0, 0.0, 0.0)),
               *t.get("mag", (0.0, 0.0, 1.0)), *t.get("urot", (0.0, 0.0, 0.0)), float(t.get("rand_orient", False)),
               float(t.get("rand_mag", False)), float(t.get("rand_urot", False))]
        for key in ("rad", "shell", "K_dist", "K_cubic_dist"):
            kind, width, rel = t.get(key, ("delta", 0.0, True))
            row += [DIST_CODE[kind], width, float(rel)]
        rows.append(row)
    tab = np.ascontiguousarray(rows, dtype=np.float64)
    assert tab.shape[1] == 34
    props, axes = np.zeros((n, 12)), np.zeros((n, 9))
    y0, typ = np.zeros((n, DIM_Y[problem])), np.zeros(n, dtype=np.int64)
    lib().or_provide(problem, aniso, len(types), int(discrete), seed, n, _p(tab), _p(props), _p(axes), _p(y0),
                     _p(typ, C.c_int64))
    return dict(props=props, axes=axes, y0=y0, type=typ)
