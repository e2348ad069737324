"""The C ABI without a GPU: the library loads, exports every entry point
include/dmx.h declares, and rejects invalid parameters before touching the
device (the reference would silently continue or stop)."""
import ctypes
import os
import re

import pytest

from dmx import abi
from dmx import problems as P

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "dmx.h")

needs_lib = pytest.mark.skipif(not os.path.exists(abi.LIB_PATH), reason="libdmx.so not built")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dmx_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert sorted(abi.EXPORTS) == declared()


@needs_lib
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_struct_layout_matches_header(tmp_path):
    """Offsets of the ctypes mirror equal what a C compiler lays out."""
    fields = [f[0] for f in abi.DmxParams._fields_]
    src = tmp_path / "off.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "dmx.h"\nint main(){'
                   + "".join(f'printf("%zu\\n", offsetof(dmx_params, {f}));' for f in fields)
                   + 'printf("%zu\\n", sizeof(dmx_params));return 0;}')
    exe = tmp_path / "off"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [getattr(abi.DmxParams, f).offset for f in fields] + [ctypes.sizeof(abi.DmxParams)]
    assert out == want


@needs_lib
@pytest.mark.parametrize("bad", [
    dict(nx=2), dict(riemann=7), dict(riemann2d=3), dict(slope_type=5),
    dict(bdtype=(P.PERIODIC, P.SHEARINGBOX, P.PERIODIC)), dict(fargo=1, omega0=0.0),
    dict(gamma=1.0), dict(courant=0.0), dict(precision=2), dict(xmax=0.0),
    dict(bdtype=(P.SHEARINGBOX, P.ZERO_GRADIENT, P.PERIODIC), omega0=1.0),
    dict(nx=900, ny=900, nz=900),        # a float64 plane above 4 GiB
    dict(nu=-1e-3), dict(eta=float("nan")),
])
def test_invalid_parameters_rejected(bad):
    p = P.Params(nx=8, ny=8, nz=8).replace(**bad)
    lib = abi.lib()
    h = ctypes.c_void_p()
    rc = lib.dmx_create(ctypes.byref(abi.params_to_c(p)), ctypes.byref(h))
    assert rc == -1
    assert h.value is None
    assert lib.dmx_last_error()


@needs_lib
def test_null_context_is_an_error():
    lib = abi.lib()
    dt = ctypes.c_double()
    assert lib.dmx_compute_dt(None, ctypes.byref(dt)) == -5
    assert b"context" in lib.dmx_last_error()
    assert lib.dmx_write_dump(None, b"/nonexistent/x.bin", 0.0, 0.0) == -5
    t = ctypes.c_double()
    assert lib.dmx_read_dump(None, b"/nonexistent/x.bin", ctypes.byref(t), ctypes.byref(dt)) == -5
    sh = (ctypes.c_int32 * 4)(0, 8, 0, 4)
    assert lib.dmx_set_shape(None, sh) == -5


@needs_lib
def test_batch_rejects_missing_contexts():
    lib = abi.lib()
    h = ctypes.c_void_p()
    assert lib.dmx_batch_create(None, 2, 0, ctypes.byref(h)) == -1
    assert h.value is None
    arr = (ctypes.c_void_p * 2)(None, None)
    assert lib.dmx_batch_create(arr, 2, 0, ctypes.byref(h)) == -5
    assert b"null context" in lib.dmx_last_error()
    t = (ctypes.c_double * 1)()
    assert lib.dmx_batch_advance(None, 1, t) == -5
    ms = ctypes.c_double()
    assert lib.dmx_batch_last_ms(None, ctypes.byref(ms)) == -5
    lib.dmx_batch_destroy(None)
