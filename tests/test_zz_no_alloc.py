"""The allocation-free entry point sp_execute_noalloc against sp_run.

The caller owns the device buffers (allocated here through the HIP runtime, as a caller under graph capture
would have them already); the library's scratch lives in its file-scope device array.  Two calls in a row, the
second with a smaller request, each equal to sp_run bit for bit (means and final state)."""
import ctypes as C

import numpy as np
import pytest

from stochasticphysics_amd import Context, compute_params
from stochasticphysics_amd.capi import SP_OK, lib

from test_gpu_parity import bn_ensemble, sph_ensemble

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def dev_from(h, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), max(8, a.nbytes)) == 0
    assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def run_noalloc(h, e):
    r, parts, keep = e.to_c()
    d_params = dev_from(h, compute_params(e))
    d_state = dev_from(h, np.asarray(e.y0, dtype=np.float64).reshape(e.n, e.dim_y).T)
    d_mean = dev_from(h, np.zeros((e.n_out, e.dim_out)))
    try:
        assert lib().sp_execute_noalloc(0, None, C.byref(r), d_params, d_state, d_mean) == SP_OK, \
            lib().sp_last_error()
        assert h.hipDeviceSynchronize() == 0
        mean = np.zeros((e.n_out, e.dim_out))
        state = np.zeros((e.dim_y, e.n))
        assert h.hipMemcpy(mean.ctypes.data, d_mean, mean.nbytes, D2H) == 0
        assert h.hipMemcpy(state.ctypes.data, d_state, state.nbytes, D2H) == 0
    finally:
        for p in (d_params, d_state, d_mean):
            h.hipFree(p)
    return mean, state.T


def test_noalloc_equals_run_twice_in_a_row():
    h = hip()
    big = bn_ensemble(n=3000, n_steps=10_000, os_=50, solver="Heun_Strong", seed=81)
    small = sph_ensemble(n=700, n_steps=2000, os_=10, solver="Heun_Strong", seed=82)
    with Context(0) as ctx:
        for e in (big, small):
            print(f"noalloc: {e.problem} {e.solver} n={e.n} steps={e.n_steps}", flush=True)
            mean, state = run_noalloc(h, e)
            ref = ctx.run(e, keep_final=True)
            assert np.array_equal(mean, ref["mean"])
            assert np.array_equal(state, ref["final_state"])


def test_noalloc_rejects_requests_beyond_its_scratch():
    e = bn_ensemble(n=16, n_steps=100, os_=1)
    r, parts, keep = e.to_c()
    r.n_particles = 50_000_000                       # 5 windows x 6 x 781,250 tiles x 8 B > 32 MiB
    rc = lib().sp_execute_noalloc(0, None, C.byref(r), C.c_void_p(8), C.c_void_p(8), C.c_void_p(8))
    assert rc != SP_OK and b"scratch" in lib().sp_last_error()
