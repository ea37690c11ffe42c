"""Python front of the gfx950 engine (tests, benchmarks, notebooks).

``Ensemble`` holds one request in the reference's vocabulary (problem, solver,
field, anisotropy, per-particle properties); ``Context`` drives one GPU through
the C ABI.  Everything numeric runs in libstophys_gpu.so.
"""
import ctypes as C
import math
from dataclasses import dataclass, field as dfield
from typing import Optional

import numpy as np

from . import capi
from .capi import (ANISO, DIM_OUT, DIM_W, DIM_Y, FIELD, NOISE, PREC, PROBLEM, SOLVER, Particles, Request, Result,
                   Tuning, check, dptr, lib)


@dataclass
class Ensemble:
    """One simulation request; arrays are per particle (length n)."""
    problem: str = "Neel"
    solver: str = "EulerMaruyama"
    field: str = "none"
    anisotropy: str = "uniaxial"
    precision: str = "fp64"
    n_steps: int = 1000
    oversampling: int = 10
    dt: float = 1e-12
    seed: int = 1
    radius: np.ndarray = None
    Ms: np.ndarray = None
    alpha: np.ndarray = None
    gamma: np.ndarray = None
    T: np.ndarray = None
    K: np.ndarray = None
    r_h: Optional[np.ndarray] = None
    viscosity: Optional[np.ndarray] = None
    axes: Optional[np.ndarray] = None          # [n, 9]
    y0: np.ndarray = None                      # [n, dim_y]
    bn_simple: bool = False
    coord_transform: bool = False
    min_angle: float = 0.0
    newton_tol: float = 1e-10
    newton_max_iter: int = 20
    check_tol: float = 1e-3
    no_fallback: bool = False
    skip_check: bool = False
    renormalize: bool = False                  # problem.normalize() after every step (any solver)
    field_offset: tuple = (0.0, 0.0, 0.0)
    field_amplitude: tuple = (0.0, 0.0, 0.0)
    field_freq: tuple = (0.0, 0.0, 0.0)
    field_phase: tuple = (0.0, 0.0, 0.0)
    field_period: float = 0.0                  # triangular, rectangular, sinc: Periode [s]
    field_time_offset: float = 0.0             # triangular, rectangular: TimeOffset [s]
    field_tau: float = 0.0                     # rectangular: Tau [s]
    field_alternating: bool = False            # rectangular: Alternating
    replay: Optional[np.ndarray] = None        # [K, n] standard normals
    K_cubic: Optional[np.ndarray] = None       # uniaxialcubic: K of the cubic part (K is K_uniaxial)
    uniaxial_rotation: Optional[np.ndarray] = None   # uniaxialcubic: [n, 3] Euler angles of the uniaxial part
    save_interval: int = 0                     # Save_single_simulations: keep particles p with (p+1) % k == 0
    euler_transform: bool = False              # BrownAndNeelEulerSpherical: Use_euler_coordinate_transformation
    brown_min_angle: float = 0.0               # BrownAndNeelEulerSpherical: Brown_Min_angle_before_transformation

    @property
    def n(self):
        return int(len(self.radius))

    @property
    def pid(self):
        return PROBLEM[self.problem]

    @property
    def dim_y(self):
        return DIM_Y[self.pid]

    @property
    def dim_out(self):
        return DIM_OUT[self.pid]

    @property
    def n_out(self):
        return self.n_steps // self.oversampling

    @property
    def total_steps(self):
        return (self.n_out - 1) * self.oversampling

    def particle_steps(self):
        return self.n * self.total_steps

    def subset(self, idx):
        """Ensemble restricted to particles idx (same seeds; particle indices are renumbered)."""
        import copy
        e = copy.copy(self)
        for k in ("radius", "Ms", "alpha", "gamma", "T", "K", "r_h", "viscosity", "axes", "y0", "K_cubic",
                  "uniaxial_rotation"):
            v = getattr(self, k)
            if v is not None:
                setattr(e, k, np.ascontiguousarray(v[idx]))
        if self.replay is not None:
            e.replay = np.ascontiguousarray(self.replay[:, idx])
        return e

    def oracle_props(self):
        """[n, 8] property rows in the oracle's layout ([n, 12] with K_cubic and the uniaxial rotation for
        uniaxialcubic anisotropy)."""
        n = self.n
        z = np.zeros(n)
        cols = [self.radius, self.Ms, self.alpha, self.gamma, self.T, self.K,
                self.r_h if self.r_h is not None else z, self.viscosity if self.viscosity is not None else z]
        if self.anisotropy == "uniaxialcubic":
            rot = np.broadcast_to(np.asarray(self.uniaxial_rotation, dtype=np.float64), (n, 3))
            cols += [np.broadcast_to(np.asarray(self.K_cubic, dtype=np.float64), (n,)), rot[:, 0], rot[:, 1], rot[:, 2]]
        return np.stack([np.broadcast_to(np.asarray(c, dtype=np.float64), (n,)) for c in cols], axis=1)

    @property
    def n_saved(self):
        return self.n // self.save_interval if self.save_interval else 0

    def field_dict(self):
        return dict(type=FIELD[self.field], offset=self.field_offset, amp=self.field_amplitude,
                    freq=self.field_freq, phase=self.field_phase, period=self.field_period,
                    time_offset=self.field_time_offset, tau=self.field_tau, alternating=self.field_alternating)

    # ---- C structs
    def to_c(self):
        """(Request, Particles, keepalive) for the C ABI."""
        n = self.n
        keep = []

        def arr(a, shape=None):
            if a is None:
                return None
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape or (n,)), dtype=np.float64)
            keep.append(a)
            return dptr(a)
        r = Request()
        r.problem, r.solver, r.field = PROBLEM[self.problem], SOLVER[self.solver], FIELD[self.field]
        r.anisotropy, r.precision = ANISO[self.anisotropy], PREC[self.precision]
        r.noise = NOISE["replay"] if self.replay is not None else NOISE["philox"]
        r.flags = ((capi.FLAG_BN_SIMPLE if self.bn_simple else 0) | (capi.FLAG_COORD_TRANSFORM if self.coord_transform else 0)
                   | (capi.FLAG_NO_FALLBACK if self.no_fallback else 0) | (capi.FLAG_SKIP_CHECK if self.skip_check else 0)
                   | (capi.FLAG_RENORMALIZE if self.renormalize else 0)
                   | (capi.FLAG_EULER_TRANSFORM if self.euler_transform else 0))
        r.newton_max_iter = self.newton_max_iter
        r.n_particles, r.n_steps, r.oversampling, r.seed = n, self.n_steps, self.oversampling, self.seed
        r.dt, r.newton_tol, r.min_angle, r.check_tol = self.dt, self.newton_tol, self.min_angle, self.check_tol
        for name in ("field_offset", "field_amplitude", "field_freq", "field_phase"):
            getattr(r, name)[:] = [float(x) for x in getattr(self, name)]
        r.field_period, r.field_time_offset, r.field_tau = self.field_period, self.field_time_offset, self.field_tau
        r.field_alternating = 1 if self.field_alternating else 0
        r.save_interval = int(self.save_interval)
        r.brown_min_angle = self.brown_min_angle
        if self.replay is not None:
            rep = np.ascontiguousarray(self.replay, dtype=np.float64)
            keep.append(rep)
            r.replay_normals = dptr(rep)
            r.replay_count = rep.shape[0]
        p = Particles()
        p.magnetic_radius, p.saturation_magnetisation = arr(self.radius), arr(self.Ms)
        p.damping, p.gyromagnetic_ratio, p.temperature = arr(self.alpha), arr(self.gamma), arr(self.T)
        p.anisotropy_constant = arr(self.K)
        p.hydrodynamic_radius, p.viscosity = arr(self.r_h), arr(self.viscosity)
        p.axes = arr(self.axes, (n, 9)) if self.axes is not None else None
        p.initial_state = arr(self.y0, (n, self.dim_y))
        if self.anisotropy == "uniaxialcubic":
            p.anisotropy_constant_cubic = arr(self.K_cubic)
            p.uniaxial_rotation = arr(self.uniaxial_rotation, (n, 3))
        return r, p, keep


def _alloc_result(e: Ensemble, keep_final):
    o = dict(time=np.zeros(e.n_out), field=np.zeros((e.n_out, 3)), mean=np.zeros((e.n_out, e.dim_out)),
             weight=np.zeros(e.dim_out))
    if keep_final:
        o["final_state"] = np.zeros((e.n, e.dim_y))
    res = Result()
    res.time, res.field, res.mean, res.weight = dptr(o["time"]), dptr(o["field"]), dptr(o["mean"]), dptr(o["weight"])
    res.final_state = dptr(o["final_state"]) if keep_final else None
    if e.n_saved:
        o["trajectories"] = np.zeros((e.n_out, e.dim_out, e.n_saved))
        res.trajectories = dptr(o["trajectories"])
    return res, o


def _finish_result(res: Result, o):
    o.update(device_ms=res.device_ms, latency_ms=res.latency_ms, particle_steps=res.particle_steps,
             check_error=res.check_error, check_noise=res.check_noise, check_bias=res.check_bias,
             check_bias_se=res.check_bias_se, used_fallback=res.used_fallback,
             precision_used=res.precision_used,
             newton_failures=res.newton_failures, n_saved=res.n_saved)
    return o


def host_normals(seed, particle, first, count):
    out = np.zeros(count)
    check(lib().sp_host_normals(seed, particle, first, count, dptr(out)))
    return out


def host_box_muller(words):
    """FP64 Box-Muller of Philox words [n, 4] (uint32) -> [n, 4] standard normals."""
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1, 4)
    out = np.zeros(w.shape)
    check(lib().sp_host_box_muller(w.ctypes.data_as(C.POINTER(C.c_uint32)), w.shape[0], dptr(out)))
    return out


def compute_params(e: Ensemble):
    r, p, keep = e.to_c()
    out = np.zeros(capi.NPARAM * e.n)
    npar = C.c_uint32()
    check(lib().sp_compute_particle_params(C.byref(r), C.byref(p), dptr(out), C.byref(npar)))
    return out.reshape(npar.value, e.n)


class Prepared:
    def __init__(self, ctx, e: Ensemble, keep_final=False):
        self.ctx, self.e, self.keep_final = ctx, e, keep_final
        r, p, keep = e.to_c()
        if keep_final:
            r.flags |= capi.FLAG_KEEP_FINAL
        self._h = C.c_void_p()
        check(lib().sp_prepare(ctx._ctx, C.byref(r), C.byref(p), C.byref(self._h)))

    def execute(self):
        res, o = _alloc_result(self.e, self.keep_final)
        check(lib().sp_execute(self._h, C.byref(res)))
        return _finish_result(res, o)

    def release(self):
        if self._h:
            lib().sp_release(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Settings:
    """A reference parameter file (sp_settings): the request, and its particles on the host on demand."""

    def __init__(self, path):
        self._h = C.c_void_p()
        check(lib().sp_settings_load(str(path).encode(), C.byref(self._h)))
        self.request = Request()
        check(lib().sp_settings_request(self._h, C.byref(self.request), None))

    @property
    def n(self):
        return self.request.n_particles

    @property
    def dims(self):
        n_out, dy, dw, dout = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().sp_request_shape(C.byref(self.request), C.byref(n_out), C.byref(dy), C.byref(dw), C.byref(dout)))
        return n_out.value, dy.value, dout.value

    def alloc_result(self, keep_final=False):
        n_out, dy, dout = self.dims
        o = dict(time=np.zeros(n_out), field=np.zeros((n_out, 3)), mean=np.zeros((n_out, dout)), weight=np.zeros(dout))
        res = Result()
        res.time, res.field, res.mean, res.weight = dptr(o["time"]), dptr(o["field"]), dptr(o["mean"]), dptr(o["weight"])
        if keep_final:
            o["final_state"] = np.zeros((self.n, dy))
            res.final_state = dptr(o["final_state"])
        k = self.request.save_interval
        if k:
            o["trajectories"] = np.zeros((n_out, dout, self.n // k))
            res.trajectories = dptr(o["trajectories"])
        return res, o

    def finish(self, res, o):
        o = _finish_result(res, o)
        o["particle_means"] = np.array(res.particle_means[:])
        return o

    def write_result(self, res, path):
        check(lib().sp_settings_write_result(self._h, C.byref(res), str(path).encode()))

    def close(self):
        if self._h:
            lib().sp_settings_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ProvidedPrepared:
    """A parameter file's ensemble prepared on the device by the provider kernel (sp_settings_prepare)."""

    def __init__(self, ctx, settings: Settings):
        self.ctx, self.settings = ctx, settings
        self._h = C.c_void_p()
        check(lib().sp_settings_prepare(ctx._ctx, settings._h, C.byref(self._h)))

    def execute(self, keep_final=False):
        res, o = self.settings.alloc_result(keep_final)
        check(lib().sp_execute(self._h, C.byref(res)))
        self.last_result = res
        return self.settings.finish(res, o)

    def inputs(self):
        """The resident parameter block [NPARAM, N] and initial state [N, dim_y] the provider drew."""
        n, (_, dy, _) = self.settings.n, self.settings.dims
        params, y0 = np.zeros(capi.NPARAM * n), np.zeros(n * dy)
        check(lib().sp_prepared_inputs(self._h, dptr(params), dptr(y0)))
        return params.reshape(capi.NPARAM, n), y0.reshape(n, dy)

    def release(self):
        if self._h:
            lib().sp_release(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Context:
    """One GPU (sp_ctx)."""

    def __init__(self, device=0):
        self._ctx = C.c_void_p()
        check(lib().sp_ctx_create(device, C.byref(self._ctx)))

    def close(self):
        if self._ctx:
            lib().sp_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- plain path
    def run(self, e: Ensemble, keep_final=False):
        r, p, keep = e.to_c()
        res, o = _alloc_result(e, keep_final)
        rc = lib().sp_run(self._ctx, C.byref(r), C.byref(p), C.byref(res))
        check(rc)
        return _finish_result(res, o)

    def prepare(self, e: Ensemble, keep_final=False):
        """Upload once; the returned handle executes with the inputs resident in HBM."""
        return Prepared(self, e, keep_final)

    def last_run_stats(self):
        s = np.zeros(8)
        check(lib().sp_ctx_last_run_stats(self._ctx, dptr(s)))
        keys = ("launches", "kernel_ms", "reduce_ms", "wall_ms", "particle_steps", "vgprs", "grid", "block")
        return dict(zip(keys, s.tolist()))

    def prepare_settings(self, settings: Settings):
        """Draw a parameter file's particles on the device; the handle executes with them resident in HBM."""
        return ProvidedPrepared(self, settings)

    def submit_settings(self, settings: Settings):
        rid = C.c_uint64()
        check(lib().sp_settings_submit(self._ctx, settings._h, C.byref(rid)))
        return rid.value

    def settings_result(self, rid, settings: Settings):
        res, o = settings.alloc_result(False)
        check(lib().sp_result_get(self._ctx, rid, C.byref(res)))
        return settings.finish(res, o)

    # ---- batched path
    def run_batch(self, ensembles, keep_final=False):
        n = len(ensembles)
        reqs, parts, keeps, results, outs = (Request * n)(), (Particles * n)(), [], (Result * n)(), []
        for i, e in enumerate(ensembles):
            r, p, k = e.to_c()
            reqs[i], parts[i] = r, p
            keeps.append(k)
            res, o = _alloc_result(e, keep_final)
            results[i] = res
            outs.append(o)
        check(lib().sp_run_batch(self._ctx, n, reqs, parts, results))
        return [_finish_result(results[i], outs[i]) for i in range(n)]

    # ---- serving
    def submit(self, e: Ensemble):
        r, p, keep = e.to_c()
        rid = C.c_uint64()
        check(lib().sp_submit(self._ctx, C.byref(r), C.byref(p), C.byref(rid)))
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        return lib().sp_wait(self._ctx, rid, timeout_ms)

    def result(self, rid, e: Ensemble, keep_final=False):
        res, o = _alloc_result(e, keep_final)
        check(lib().sp_result_get(self._ctx, rid, C.byref(res)))
        return _finish_result(res, o)

    def cancel(self, rid):
        return lib().sp_cancel(self._ctx, rid)

    def progress(self, rid):
        f = C.c_double()
        check(lib().sp_progress(self._ctx, rid, C.byref(f)))
        return f.value

    def server_stats(self):
        s = np.zeros(8)
        check(lib().sp_server_stats(self._ctx, dptr(s)))
        keys = ("completed", "active", "pending", "iterations", "busy_ms", "wall_ms", "particle_steps", "fallbacks")
        return dict(zip(keys, s.tolist()))

    def set_memory_budget(self, nbytes, max_active=512):
        check(lib().sp_ctx_set_memory_budget(self._ctx, int(nbytes), int(max_active)))

    def set_reduction_limit(self, nbytes):
        check(lib().sp_ctx_set_reduction_limit(self._ctx, int(nbytes)))

    def set_iteration_budget(self, ms):
        """Server: cap each kernel class's launch per scheduler iteration at about `ms` (0: the tuned chunk)."""
        check(lib().sp_ctx_set_iteration_budget(self._ctx, float(ms)))

    def inject_failures(self, n_iterations):
        check(lib().sp_ctx_inject_failures(self._ctx, int(n_iterations)))

    def set_tuning(self, problem, solver, precision, **kw):
        t = Tuning(**{k: int(v) for k, v in kw.items()})
        check(lib().sp_ctx_set_tuning(self._ctx, PROBLEM[problem], SOLVER[solver], PREC[precision], C.byref(t)))

    def autotune(self, e: Ensemble):
        r, p, keep = e.to_c()
        t, rate = Tuning(), C.c_double()
        check(lib().sp_autotune(self._ctx, C.byref(r), C.byref(p), C.byref(t), C.byref(rate)))
        return dict(block_size=t.block_size, chunk_steps=t.chunk_steps,
                    use_graph=t.use_graph, persistent_waves_per_simd=t.persistent_waves_per_simd), rate.value

    # ---- noise test hooks
    def dev_philox(self, ctr, key, count):
        c = (C.c_uint32 * 4)(*[int(x) for x in ctr])
        k = (C.c_uint32 * 2)(*[int(x) for x in key])
        out = np.zeros(4 * count, dtype=np.uint32)
        check(lib().sp_dev_philox(self._ctx, c, k, count, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out.reshape(count, 4)

    def dev_normals(self, seed, particle, first, count, precision="fp64"):
        out = np.zeros(count)
        check(lib().sp_dev_normals(self._ctx, seed, particle, first, count, PREC[precision], dptr(out)))
        return out

    def dev_box_muller(self, words, precision="fp64"):
        w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1, 4)
        out = np.zeros(w.shape)
        check(lib().sp_dev_box_muller(self._ctx, w.ctypes.data_as(C.POINTER(C.c_uint32)), w.shape[0], PREC[precision],
                                      dptr(out)))
        return out
