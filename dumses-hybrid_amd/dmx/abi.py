"""ctypes binding of libdmx.so (the C ABI declared in include/dmx.h).

This is the Python host side of the drop-in boundary: the same entry points a
Fortran iso_c_binding shim binds (see fortran/dmx_shim.f90), with the
reference's argument meaning.  Loading fails loudly if the HIP library is not
built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .problems import Params

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "build", "libdmx.so")


class DmxParams(ctypes.Structure):
    """Mirror of ``struct dmx_params`` (include/dmx.h)."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("iso", ctypes.c_int32), ("slope_type", ctypes.c_int32),
        ("riemann", ctypes.c_int32), ("riemann2d", ctypes.c_int32),
        ("bdtype", ctypes.c_int32 * 3), ("fargo", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("gamma", ctypes.c_double), ("ciso", ctypes.c_double),
        ("omega0", ctypes.c_double), ("courant", ctypes.c_double),
        ("nu", ctypes.c_double), ("eta", ctypes.c_double),
    ]


def params_to_c(p: Params) -> DmxParams:
    c = DmxParams()
    for k in ("nx", "ny", "nz", "iso", "slope_type", "riemann", "riemann2d", "fargo",
              "precision"):
        setattr(c, k, int(getattr(p, k)))
    for i in range(3):
        c.bdtype[i] = int(p.bdtype[i])
    for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "gamma", "ciso", "omega0",
              "courant", "nu", "eta"):
        setattr(c, k, float(getattr(p, k)))
    return c


# every entry point include/dmx.h declares
EXPORTS = ("dmx_create", "dmx_destroy", "dmx_upload", "dmx_download", "dmx_compute_dt",
           "dmx_compute_dt_noalloc",
           "dmx_godunov", "dmx_fargo", "dmx_boundary", "dmx_dissipation", "dmx_advance", "dmx_history",
           "dmx_history_shear", "dmx_write_dump", "dmx_read_dump",
           "dmx_last_advance_ms", "dmx_profile", "dmx_kernel_name", "dmx_tune", "dmx_set_shape",
           "dmx_set_limiter", "dmx_learned_slopes", "dmx_mlp_create", "dmx_mlp_destroy",
           "dmx_mlp_set_params", "dmx_mlp_get_params", "dmx_mlp_forward", "dmx_mlp_loss_grad",
           "dmx_mlp_load_data", "dmx_mlp_train", "dmx_mfma_selftest", "dmx_batch_create",
           "dmx_batch_destroy", "dmx_batch_advance", "dmx_batch_tune", "dmx_batch_last_ms",
           "dmx_last_error", "dmx_version")

_lib = None


class DmxError(RuntimeError):
    pass


def lib(path: str | None = None):
    """Loads libdmx.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("DMX_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise DmxError(f"libdmx.so not built at {path}: run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    L.dmx_create.argtypes = [ctypes.POINTER(DmxParams), ctypes.POINTER(vp)]
    L.dmx_destroy.argtypes = [vp]
    L.dmx_destroy.restype = None
    L.dmx_upload.argtypes = [vp, dp]
    L.dmx_download.argtypes = [vp, dp]
    L.dmx_compute_dt.argtypes = [vp, dp]
    L.dmx_compute_dt_noalloc.argtypes = [ctypes.POINTER(DmxParams), vp, vp, vp]
    L.dmx_godunov.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.dmx_fargo.argtypes = [vp, ctypes.c_double]
    L.dmx_boundary.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.dmx_dissipation.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.dmx_advance.argtypes = [vp, ctypes.c_int32, dp, dp]
    L.dmx_history.argtypes = [vp, dp]
    L.dmx_history_shear.argtypes = [vp, dp]
    L.dmx_last_advance_ms.argtypes = [vp, dp]
    L.dmx_write_dump.argtypes = [vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_double]
    L.dmx_read_dump.argtypes = [vp, ctypes.c_char_p, dp, dp]
    L.dmx_profile.argtypes = [vp, ctypes.c_int32, dp, dp, ctypes.c_int32]
    L.dmx_kernel_name.argtypes = [ctypes.c_int32]
    L.dmx_tune.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int32)]
    L.dmx_set_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    fp = ctypes.POINTER(ctypes.c_float)
    L.dmx_set_limiter.argtypes = [vp, fp, ctypes.c_int32]
    L.dmx_learned_slopes.argtypes = [vp, fp, dp]
    L.dmx_mlp_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp)]
    L.dmx_mlp_destroy.argtypes = [vp]
    L.dmx_mlp_destroy.restype = None
    L.dmx_mlp_set_params.argtypes = [vp, fp, ctypes.c_int32]
    L.dmx_mlp_get_params.argtypes = [vp, fp, ctypes.c_int32]
    L.dmx_mlp_forward.argtypes = [vp, fp, ctypes.c_int32, fp]
    L.dmx_mlp_loss_grad.argtypes = [vp, fp, fp, ctypes.c_int32, fp, fp]
    L.dmx_mlp_load_data.argtypes = [vp, fp, fp, ctypes.c_int64]
    L.dmx_mlp_train.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64, fp]
    L.dmx_mfma_selftest.argtypes = [fp]
    L.dmx_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(vp)]
    L.dmx_batch_destroy.argtypes = [vp]
    L.dmx_batch_destroy.restype = None
    L.dmx_batch_advance.argtypes = [vp, ctypes.c_int32, dp]
    L.dmx_batch_tune.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int32)]
    L.dmx_batch_last_ms.argtypes = [vp, dp]
    L.dmx_kernel_name.restype = ctypes.c_char_p
    L.dmx_last_error.restype = ctypes.c_char_p
    L.dmx_version.restype = ctypes.c_char_p
    _lib = L
    return L


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().dmx_last_error()
        raise DmxError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


class Domain:
    """One device-resident ghosted domain (a dmx_ctx)."""

    def __init__(self, p: Params, uin: np.ndarray | None = None):
        self.p = p
        self._cp = params_to_c(p)
        h = ctypes.c_void_p()
        _check(lib().dmx_create(ctypes.byref(self._cp), ctypes.byref(h)), "dmx_create")
        self.h = h
        self.time = 0.0
        if uin is not None:
            self.upload(uin)

    def upload(self, uin: np.ndarray):
        u = np.ascontiguousarray(uin, dtype=np.float64)
        p = self.p
        if u.shape != (11, p.nz + 6, p.ny + 6, p.nx + 6):
            raise DmxError(f"uin shape {u.shape} does not match the domain")
        _check(lib().dmx_upload(self.h, _ptr(u)), "dmx_upload")

    def download(self) -> np.ndarray:
        p = self.p
        u = np.zeros((11, p.nz + 6, p.ny + 6, p.nx + 6))
        _check(lib().dmx_download(self.h, _ptr(u)), "dmx_download")
        return u

    def compute_dt(self) -> float:
        dt = ctypes.c_double()
        _check(lib().dmx_compute_dt(self.h, ctypes.byref(dt)), "dmx_compute_dt")
        return dt.value

    def godunov(self, time: float, dt: float):
        _check(lib().dmx_godunov(self.h, time, dt), "dmx_godunov")

    def fargo(self, dt: float):
        _check(lib().dmx_fargo(self.h, dt), "dmx_fargo")

    def boundary(self, time: float, dt: float):
        _check(lib().dmx_boundary(self.h, time, dt), "dmx_boundary")

    def dissipation(self, time: float, dt: float) -> None:
        _check(lib().dmx_dissipation(self.h, time, dt), "dmx_dissipation")

    def advance(self, nsteps: int, time: float | None = None, want_dts: bool = True):
        t = ctypes.c_double(self.time if time is None else time)
        dts = np.zeros(max(nsteps, 1))
        _check(lib().dmx_advance(self.h, nsteps, ctypes.byref(t),
                                 _ptr(dts) if want_dts else None), "dmx_advance")
        self.time = t.value
        return t.value, dts[:nsteps]

    def write_dump(self, path: str, time: float, dt: float) -> None:
        """The reference's binary dump of the current state (dmx_write_dump)."""
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        _check(lib().dmx_write_dump(self.h, path.encode(), float(time), float(dt)), "dmx_write_dump")

    def read_dump(self, path: str) -> tuple[float, float]:
        """Restart from a binary dump (dmx_read_dump); returns (time, dt)."""
        t, dt = ctypes.c_double(), ctypes.c_double()
        _check(lib().dmx_read_dump(self.h, path.encode(), ctypes.byref(t), ctypes.byref(dt)), "dmx_read_dump")
        self.time = t.value
        return t.value, dt.value

    def history(self, shear: bool = False) -> np.ndarray:
        out = np.zeros(9)
        fn = lib().dmx_history_shear if shear else lib().dmx_history
        _check(fn(self.h, _ptr(out)), "dmx_history")
        return out

    def last_advance_ms(self) -> float:
        ms = ctypes.c_double()
        _check(lib().dmx_last_advance_ms(self.h, ctypes.byref(ms)), "dmx_last_advance_ms")
        return ms.value

    def profile(self, nsteps: int, time: float | None = None, ngroups: int = 12) -> dict:
        """Mean device ms per step of each kernel group over nsteps."""
        t = ctypes.c_double(self.time if time is None else time)
        ms = np.zeros(ngroups)
        _check(lib().dmx_profile(self.h, nsteps, ctypes.byref(t), _ptr(ms), ngroups), "dmx_profile")
        self.time = t.value
        return {lib().dmx_kernel_name(k).decode(): float(ms[k]) for k in range(ngroups)}

    def tune(self) -> float:
        """Autotunes the trace and face/edge launch shapes on the current
        state (dmx_tune); returns the best face/edge stage time in ms and
        keeps the chosen (mode, planes) in self.shape (face/edge stage) and
        self.trace_shape (trace)."""
        ms = ctypes.c_double()
        sh = (ctypes.c_int32 * 4)()
        _check(lib().dmx_tune(self.h, ctypes.byref(ms), sh), "dmx_tune")
        self.shape = (int(sh[0]), int(sh[1]))
        self.trace_shape = (int(sh[2]), int(sh[3]))
        return ms.value

    def set_shape(self, flux, trace) -> None:
        """Launch shapes without timing (dmx_set_shape): flux = (mode,
        planes) of the face/edge stage, trace = (mode, planes)."""
        sh = (ctypes.c_int32 * 4)(int(flux[0]), int(flux[1]), int(trace[0]), int(trace[1]))
        _check(lib().dmx_set_shape(self.h, sh), "dmx_set_shape")
        self.shape = (int(flux[0]), int(flux[1]))
        self.trace_shape = (int(trace[0]), int(trace[1]))

    def set_limiter(self, theta) -> None:
        """Weights of the learned limiter (slope_type 4), 7304 floats."""
        th = np.ascontiguousarray(theta, dtype=np.float32)
        _check(lib().dmx_set_limiter(self.h, _fptr(th), th.size), "dmx_set_limiter")

    def learned_slopes(self):
        """(slopes (3, 8, nz, ny, nx) float32, fraction of cell-directions
        with the positivity fallback, fraction of variable slopes replaced by
        the TVD rule) of the current state (dmx_learned_slopes)."""
        p = self.p
        out = np.zeros((3, 8, p.nz, p.ny, p.nx), dtype=np.float32)
        fb = np.zeros(2)
        _check(lib().dmx_learned_slopes(self.h, _fptr(out), _ptr(fb)), "dmx_learned_slopes")
        return out, float(fb[0]), float(fb[1])

    def close(self):
        if getattr(self, "h", None):
            lib().dmx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The limiter network's trainer on the GPU (dmx_mlp_*)."""

    def __init__(self, max_batch: int):
        h = ctypes.c_void_p()
        _check(lib().dmx_mlp_create(max_batch, ctypes.byref(h)), "dmx_mlp_create")
        self.h = h
        self.cap = max_batch

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        _check(lib().dmx_mlp_set_params(self.h, _fptr(th), th.size), "dmx_mlp_set_params")

    def get_params(self) -> np.ndarray:
        from .learned import NPARAM
        th = np.zeros(NPARAM, dtype=np.float32)
        _check(lib().dmx_mlp_get_params(self.h, _fptr(th), th.size), "dmx_mlp_get_params")
        return th

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 8), dtype=np.float32)
        _check(lib().dmx_mlp_forward(self.h, _fptr(X), X.shape[0], _fptr(Y)), "dmx_mlp_forward")
        return Y

    def loss_grad(self, X, T):
        from .learned import NPARAM
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = ctypes.c_float()
        _check(lib().dmx_mlp_loss_grad(self.h, _fptr(X), _fptr(T), X.shape[0], ctypes.byref(loss), _fptr(g)),
               "dmx_mlp_loss_grad")
        return loss.value, g

    def load_data(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        _check(lib().dmx_mlp_load_data(self.h, _fptr(X), _fptr(T), X.shape[0]), "dmx_mlp_load_data")

    def train(self, steps: int, batch: int, lr: float, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        _check(lib().dmx_mlp_train(self.h, steps, batch, lr, seed, _fptr(losses)), "dmx_mlp_train")
        return losses[:steps]

    def close(self):
        if getattr(self, "h", None):
            lib().dmx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Domains of one precision advanced together (dmx_batch_*): every kernel
    of the step is one launch over the tiles of all boxes.  The domains keep
    their state (download each one); they must stay open while the batch is
    used."""

    def __init__(self, domains, nstreams: int = 0):
        self.domains = list(domains)
        arr = (ctypes.c_void_p * len(self.domains))(*[d.h.value for d in self.domains])
        h = ctypes.c_void_p()
        _check(lib().dmx_batch_create(arr, len(self.domains), nstreams, ctypes.byref(h)), "dmx_batch_create")
        self.h = h
        self.shape = None
        self.trace_shape = None

    def advance(self, nsteps: int, times=None) -> np.ndarray:
        """nsteps steps of every box from times (default: each domain's
        time); returns the end times and stores them in the domains."""
        t = np.array([d.time for d in self.domains] if times is None else times, dtype=np.float64)
        _check(lib().dmx_batch_advance(self.h, nsteps, _ptr(t)), "dmx_batch_advance")
        for d, tb in zip(self.domains, t):
            d.time = float(tb)
        return t

    def tune(self) -> float:
        """Autotunes the batch's launch shapes on the current states
        (dmx_batch_tune: the trace, then the face/edge stage after it
        against the fused trace + face/edge kernel, mode 3); keeps them in
        self.shape and self.trace_shape and returns the chosen stage
        kernel's time in ms.  The states are unchanged."""
        ms = ctypes.c_double()
        sh = (ctypes.c_int32 * 4)()
        _check(lib().dmx_batch_tune(self.h, ctypes.byref(ms), sh), "dmx_batch_tune")
        self.shape = (int(sh[0]), int(sh[1]))
        self.trace_shape = (int(sh[2]), int(sh[3]))
        return ms.value

    def last_ms(self) -> float:
        ms = ctypes.c_double()
        _check(lib().dmx_batch_last_ms(self.h, ctypes.byref(ms)), "dmx_batch_last_ms")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().dmx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mfma_selftest() -> np.ndarray:
    out = np.zeros((16, 16), dtype=np.float32)
    _check(lib().dmx_mfma_selftest(_fptr(out)), "dmx_mfma_selftest")
    return out
