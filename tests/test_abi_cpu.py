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


@needs_lib
def test_c_usage_compiles_and_links(tmp_path):
    """The C usage INTEGRATION.md documents (context, upload, advance,
    download, autotuner, batch) compiles against include/dmx.h with a plain
    C compiler and links against libdmx.so (not run: no GPU here)."""
    src = tmp_path / "use.c"
    src.write_text(r'''
#include <stdio.h>
#include <stdlib.h>
#include "dmx.h"
int main(void)
{
    dmx_params p = { .nx = 64, .ny = 64, .nz = 64, .iso = 0, .slope_type = 2,
                     .riemann = DMX_HLLD, .riemann2d = DMX_HLLD,
                     .bdtype = {DMX_BC_PERIODIC, DMX_BC_PERIODIC, DMX_BC_PERIODIC},
                     .fargo = 0, .precision = DMX_FP64,
                     .xmin = 0, .xmax = 1, .ymin = 0, .ymax = 1, .zmin = 0, .zmax = 1,
                     .gamma = 5.0 / 3.0, .ciso = 0, .omega0 = 0, .courant = 0.7,
                     .nu = 0, .eta = 0 };
    dmx_ctx* ctx[2];
    size_t n = (size_t)70 * 70 * 70 * 11;
    double* uin = calloc(n, sizeof(double));
    for (int b = 0; b < 2; b++) {
        if (dmx_create(&p, &ctx[b])) { fprintf(stderr, "%s\n", dmx_last_error()); return 1; }
        dmx_upload(ctx[b], uin);
    }
    double t = 0.0, dts[10], ms;
    int32_t shape[4];
    dmx_tune(ctx[0], &ms, shape);
    dmx_set_shape(ctx[0], shape);
    dmx_advance(ctx[0], 10, &t, dts);
    dmx_batch* bt;
    double times[2] = {0.0, 0.0};
    if (dmx_batch_create(ctx, 2, 0, &bt) == 0) {
        dmx_batch_advance(bt, 5, times);
        dmx_batch_destroy(bt);
    }
    dmx_download(ctx[0], uin);
    for (int b = 0; b < 2; b++) dmx_destroy(ctx[b]);
    free(uin);
    return 0;
}
''')
    import subprocess
    libdir = os.path.dirname(abi.LIB_PATH)
    exe = tmp_path / "use"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), str(src),
                        "-L", libdir, "-ldmx", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
