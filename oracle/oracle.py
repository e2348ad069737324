"""Python access to the oracle (test infrastructure only).

Two checkers live under oracle/:
  * ``Oracle``: the C restatement (oracle/dmx_oracle.c, built by
    ``make -C oracle`` into oracle/_build/libdmx_oracle.so), loaded by ctypes;
  * ``run_reference``: the compiled reference itself (oracle/ref.mk builds
    oracle/_ref/iso{0,1}/ref_harness from the reference sources), driven
    through its binary restart interface.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
them, as the checker; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "dumses-hybrid_amd"))

from dmx.dumpfile import Dump, dump_path, write_dump  # noqa: E402
from dmx.problems import Params  # noqa: E402
from dmx.abi import DmxParams, params_to_c  # noqa: E402

LIB_PATH = os.path.join(HERE, "_build", "libdmx_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.dmxo_create.restype = ctypes.c_void_p
        L.dmxo_create.argtypes = [ctypes.POINTER(DmxParams)]
        L.dmxo_destroy.argtypes = [ctypes.c_void_p]
        L.dmxo_load.argtypes = [ctypes.c_void_p, dp]
        L.dmxo_store.argtypes = [ctypes.c_void_p, dp]
        L.dmxo_compute_dt.restype = ctypes.c_double
        L.dmxo_compute_dt.argtypes = [ctypes.c_void_p]
        L.dmxo_godunov.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.dmxo_fargo.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.dmxo_boundary.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.dmxo_dissipation.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.dmxo_advance.restype = ctypes.c_double
        L.dmxo_advance.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, dp]
        L.dmxo_history.argtypes = [ctypes.c_void_p, dp]
        L.dmxo_history_variant.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.dmxo_riemann1d.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, dp, dp]
        L.dmxo_riemann2d.restype = ctypes.c_double
        L.dmxo_riemann2d.argtypes = [ctypes.c_int, dp, dp, dp, dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """FP64 CPU restatement of the reference step on one ghosted domain."""

    def __init__(self, p: Params, uin: np.ndarray):
        self.p = p
        self._cp = params_to_c(p)
        self.h = lib().dmxo_create(ctypes.byref(self._cp))
        self.load(uin)

    def load(self, uin):
        u = np.ascontiguousarray(uin, dtype=np.float64)
        lib().dmxo_load(self.h, _ptr(u))

    def store(self) -> np.ndarray:
        p = self.p
        u = np.zeros((11, p.nz + 6, p.ny + 6, p.nx + 6))
        lib().dmxo_store(self.h, _ptr(u))
        return u

    def compute_dt(self) -> float:
        return lib().dmxo_compute_dt(self.h)

    def godunov(self, time, dt):
        lib().dmxo_godunov(self.h, time, dt)

    def fargo(self, dt):
        lib().dmxo_fargo(self.h, dt)

    def boundary(self, time, dt):
        lib().dmxo_boundary(self.h, time, dt)

    def dissipation(self, time, dt):
        lib().dmxo_dissipation(self.h, time, dt)

    def advance(self, nsteps: int, time: float = 0.0):
        dts = np.zeros(max(nsteps, 1))
        t = lib().dmxo_advance(self.h, nsteps, time, _ptr(dts))
        return t, dts[:nsteps]

    def history(self, shear: bool = False) -> np.ndarray:
        out = np.zeros(9)
        lib().dmxo_history_variant(self.h, 1 if shear else 0, _ptr(out))
        return out

    def close(self):
        if self.h:
            lib().dmxo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def riemann1d(solver: int, ql: np.ndarray, qr: np.ndarray, gamma: float,
              ciso: float = 0.0, iso: int = 0):
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.zeros(8)
    st = np.zeros(7)
    lib().dmxo_riemann1d(solver, _ptr(ql), _ptr(qr), gamma, ciso, iso, _ptr(f), _ptr(st))
    return f, st


def riemann2d(solver: int, qs: np.ndarray, dt: float, gamma: float, ciso: float = 0.0,
              iso: int = 0) -> float:
    """qs: (4, 8) corner states LL, RL, LR, RR."""
    q = [np.ascontiguousarray(qs[c], dtype=np.float64) for c in range(4)]
    return lib().dmxo_riemann2d(solver, *(_ptr(a) for a in q), dt, gamma, ciso, iso)


# ---------------------------------------------------------------------------
# the compiled reference (available where it was built from /root/reference)
# ---------------------------------------------------------------------------
_BC = {0: "'periodic'", 1: "'zero_gradient'", 2: "'shearingbox'"}
_RS = {2: "'llf'", 3: "'hll'", 4: "'hlld'", 5: "'upwind'", 8: "'hlla'"}


def reference_available(iso: int = 0) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"iso{iso}", "ref_harness"))


def _fmt(x: float) -> str:
    return repr(float(x)).replace("e", "d") if "e" in repr(float(x)) else repr(float(x)) + "d0"


def write_input(path: str, p: Params, restart: int = 1) -> None:
    """The reference's namelist input file (src/subroutines/init.f90:153-173)."""
    with open(path, "w") as f:
        f.write(f"&start_params\n restart={restart}\n tlim=1.d30\n verbose=.false.\n debug=.false.\n/\n")
        f.write("&scheme_params\n")
        f.write(f" bdtypex={_BC[p.bdtype[0]]}\n bdtypey={_BC[p.bdtype[1]]}\n bdtypez={_BC[p.bdtype[2]]}\n")
        f.write(f" riemann={_RS[p.riemann]}\n riemann2d={_RS[p.riemann2d]}\n")
        f.write(f" slope_type={p.slope_type}\n courant={_fmt(p.courant)}\n")
        f.write(f" fargo={'.true.' if p.fargo else '.false.'}\n/\n")
        f.write("&model_params\n")
        f.write(f" nx={p.nx}\n ny={p.ny}\n nz={p.nz}\n")
        for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
            f.write(f" {k}={_fmt(getattr(p, k))}\n")
        f.write(f" Omega0={_fmt(p.omega0)}\n ciso={_fmt(p.ciso)}\n gamma={_fmt(p.gamma)}\n")
        f.write(f" nu={_fmt(p.nu)}\n eta={_fmt(p.eta)}\n/\n")
        f.write("&output_params\n dtdump=1.d30\n dthist=-1.d0\n dtspec=-1.d0\n io_type='binary'\n/\n")
        f.write("&mpi_params\n nxslice=1\n nyslice=1\n nzslice=1\n/\n")


def run_reference(p: Params, uin: np.ndarray, nsteps: int, dumpevery: int = 0,
                  threads: int = 1, time: float = 0.0, keep: str | None = None):
    """Runs the compiled reference for nsteps from uin (injected through its
    restart reader).  Returns (final uin, dts, {step: uin} for dumps)."""
    exe = os.path.join(REF_DIR, f"iso{int(p.iso)}", "ref_harness")
    if not os.path.exists(exe):
        raise RuntimeError(f"compiled reference missing: {exe} (make -f oracle/ref.mk)")
    work = tempfile.mkdtemp(prefix="dmxref_")
    try:
        write_input(os.path.join(work, "input"), p)
        with open(os.path.join(work, "harness"), "w") as f:
            f.write(f"&harness_params\n nsteps={nsteps}\n dumpevery={dumpevery}\n mode='step'\n/\n")
        d = Dump(p.nx, p.ny, p.nz, time, 0.0, p.dx, p.dy, p.dz,
                 (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax), p.gamma, p.ciso,
                 np.ascontiguousarray(uin, dtype=np.float64))
        write_dump(dump_path(1, 0, work), d)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([exe], cwd=work, env=env, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        shape = (11, p.nz + 6, p.ny + 6, p.nx + 6)
        dts = np.fromfile(os.path.join(work, "dt.raw"), dtype="<f8")
        dumps = {}
        for fn in os.listdir(work):
            if fn.startswith("uin_") and fn.endswith(".raw"):
                dumps[int(fn[4:10])] = np.fromfile(os.path.join(work, fn), dtype="<f8").reshape(shape)
        return dumps[nsteps], dts, dumps
    finally:
        if keep:
            shutil.copytree(work, keep, dirs_exist_ok=True)
        shutil.rmtree(work, ignore_errors=True)


def parse_history(text: str):
    """Rows of a history.txt written by the reference's history routine
    (src/subroutines/history.f90:146-163: time, dt, then mass, maxwell,
    reynolds, max+rey, magp, meanBx, meanBy, meanBz, divB in E14.5).  Returns
    (values, half_ulps): float arrays of shape (rows, 11), the second the
    rounding bound of each printed value, half a unit in its last printed
    digit."""
    vals, hulps = [], []
    for line in text.splitlines():
        if not line.strip() or line.lstrip().startswith("#"):
            continue
        v, h = [], []
        for tok in line.split():
            mant, _, exp = tok.upper().partition("E")
            digits = len(mant.split(".")[1]) if "." in mant else 0
            v.append(float(tok))
            h.append(0.5 * 10.0 ** (int(exp or 0) - digits))
        vals.append(v)
        hulps.append(h)
    return np.array(vals), np.array(hulps)


def run_reference_history(p: Params, uin: np.ndarray, time: float = 0.0):
    """The reference's own history routine on the state uin (injected through
    its restart reader): returns the printed row (values, half_ulps) as
    parse_history, columns time, dt, mass .. divB."""
    exe = os.path.join(REF_DIR, f"iso{int(p.iso)}", "ref_harness")
    if not os.path.exists(exe):
        raise RuntimeError(f"compiled reference missing: {exe} (make -f oracle/ref.mk)")
    work = tempfile.mkdtemp(prefix="dmxrefh_")
    try:
        write_input(os.path.join(work, "input"), p)
        with open(os.path.join(work, "harness"), "w") as f:
            f.write("&harness_params\n nsteps=0\n dumpevery=0\n mode='history'\n/\n")
        d = Dump(p.nx, p.ny, p.nz, time, 0.0, p.dx, p.dy, p.dz,
                 (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax), p.gamma, p.ciso,
                 np.ascontiguousarray(uin, dtype=np.float64))
        write_dump(dump_path(1, 0, work), d)
        r = subprocess.run([exe], cwd=work, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference history failed: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        vals, hulps = parse_history(open(os.path.join(work, "history.txt")).read())
        return vals[-1], hulps[-1]
    finally:
        shutil.rmtree(work, ignore_errors=True)


def run_reference_kernels(iso: int, gamma: float, ciso: float, dt: float,
                          s1: np.ndarray, s2: np.ndarray) -> np.ndarray:
    """Feeds n 1D state pairs s1 (n,16) and n 2D corner sets s2 (n,32) through
    the reference's Riemann solvers; returns the raw outputs (n*(15+24) + n*4)."""
    exe = os.path.join(REF_DIR, f"iso{int(iso)}", "ref_harness")
    work = tempfile.mkdtemp(prefix="dmxrefk_")
    try:
        n = s1.shape[0]
        with open(os.path.join(work, "harness"), "w") as f:
            f.write("&harness_params\n nsteps=0\n dumpevery=0\n mode='kernels'\n/\n")
        with open(os.path.join(work, "kin.raw"), "wb") as f:
            f.write(np.int32(n).tobytes())
            f.write(np.array([gamma, ciso, dt], dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(s1, dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(s2, dtype="<f8").tobytes())
        r = subprocess.run([exe], cwd=work, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference kernels failed: {r.stderr[-2000:]}")
        return np.fromfile(os.path.join(work, "kout.raw"), dtype="<f8")
    finally:
        shutil.rmtree(work, ignore_errors=True)
