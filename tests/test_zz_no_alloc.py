"""dmx_compute_dt_noalloc (the CFL timestep with no device allocation,
block minima in a static device array) against dmx_compute_dt of a context
holding the same state: bit for bit, on two calls in a row, the second on a
smaller domain, so the second call reads only minima it wrote itself."""
import ctypes

import numpy as np
import pytest

from dmx import abi
from dmx import problems as P

pytestmark = pytest.mark.gpu


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    vp = ctypes.c_void_p
    h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2
DMX_E_PARAM = -1   # include/dmx.h


def _noalloc_dt(h, p, uin):
    """Copies uin to device memory, runs the no-alloc entry point on the
    null stream and reads the timestep back."""
    L = abi.lib()
    u = np.ascontiguousarray(uin, dtype=np.float64)
    du, ddt = ctypes.c_void_p(), ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(du), u.nbytes) == 0
    assert h.hipMalloc(ctypes.byref(ddt), 8) == 0
    try:
        assert h.hipMemcpy(du, u.ctypes.data, u.nbytes, H2D) == 0
        cp = abi.params_to_c(p)
        rc = L.dmx_compute_dt_noalloc(ctypes.byref(cp), du, ddt, None)
        assert rc == 0, L.dmx_last_error()
        assert h.hipDeviceSynchronize() == 0
        out = np.zeros(1)
        assert h.hipMemcpy(out.ctypes.data, ddt, 8, D2H) == 0
        return float(out[0])
    finally:
        h.hipFree(du)
        h.hipFree(ddt)


def test_noalloc_dt_equals_context_dt():
    h = _hip()
    cases = [("Orszag-Tang 48^3", P.orszag_tang(48)),
             ("shearing box", P.mri_test(16, 32, 16)),
             ("random 13x7x5", P.random_smooth(13, 7, 5, seed=7))]
    for name, pb in cases:
        p = pb.params.replace(precision=0)
        print(f"[gpu] no-alloc dt, {name} {p.nx}x{p.ny}x{p.nz}", flush=True)
        d = abi.Domain(p, pb.uin)
        d.advance(2, 0.0)                 # a state the step has touched
        state = d.download()              # right faces = next left faces
        want = d.compute_dt()
        d.close()
        got = _noalloc_dt(h, p, state)
        assert got == want, (name, got, want)


def test_noalloc_dt_rejects_bad_input():
    L = abi.lib()
    pb = P.random_smooth(13, 7, 5, seed=7)
    cp = abi.params_to_c(pb.params.replace(precision=1))
    assert L.dmx_compute_dt_noalloc(ctypes.byref(cp), ctypes.c_void_p(8), ctypes.c_void_p(8), None) \
        == DMX_E_PARAM
    cp = abi.params_to_c(pb.params.replace(precision=0))
    assert L.dmx_compute_dt_noalloc(ctypes.byref(cp), None, None, None) == DMX_E_PARAM
