"""Binary dump writer and restart reader on the C ABI (dmx_write_dump,
dmx_read_dump): the reference's format (src/subroutines/output.f90:224-235,
src/subroutines/restart.f90:212-223), and checkpoint/resume that continues
bit for bit."""
import os

import numpy as np
import pytest

from dmx import abi
from dmx import problems as P
from dmx.dumpfile import Dump, read_dump, write_dump


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_dump_bytes_equal_reference_layout(tmp_path, precision):
    pb = P.orszag_tang(16)
    p = pb.params.replace(precision=precision)
    print(f"[gpu] dump layout precision={precision}", flush=True)
    d = abi.Domain(p, pb.uin)
    t, dts = d.advance(3, 0.0)
    path = str(tmp_path / "output_000002" / "data.000000.bin")
    d.write_dump(path, t, dts[-1])
    u = d.download()
    ref = str(tmp_path / "py.bin")
    write_dump(ref, Dump(p.nx, p.ny, p.nz, t, dts[-1], p.dx, p.dy, p.dz,
                         (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax), p.gamma, p.ciso, u))
    assert open(path, "rb").read() == open(ref, "rb").read()
    back = read_dump(path)
    assert back.time == t and np.array_equal(back.uin, u)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ot", "shear_fargo"])
def test_restart_continues_bitwise(tmp_path, name):
    if name == "ot":
        pb = P.orszag_tang(16)
    else:
        pb = P.shearing_box(16, 32, 16, fargo=True)
    p = pb.params
    print(f"[gpu] restart round trip {name}", flush=True)
    a = abi.Domain(p, pb.uin)
    t, dts = a.advance(5, 0.0)
    path = str(tmp_path / "output_000001" / "data.000000.bin")
    a.write_dump(path, t, dts[-1])
    ta, _ = a.advance(5, t)
    ua = a.download()
    b = abi.Domain(p, np.zeros_like(pb.uin))
    tb0, dtb = b.read_dump(path)
    assert tb0 == t and dtb == dts[-1]
    tb, _ = b.advance(5, tb0)
    assert tb == ta
    assert np.array_equal(b.download(), ua)


@pytest.mark.gpu
def test_read_dump_rejects_mismatched_domain(tmp_path):
    pb = P.orszag_tang(16)
    small = P.orszag_tang(8)
    print("[gpu] restart size mismatch", flush=True)
    path = str(tmp_path / "d.bin")
    p8 = small.params
    write_dump(path, Dump(p8.nx, p8.ny, p8.nz, 0.0, 0.0, p8.dx, p8.dy, p8.dz,
                          (p8.xmin, p8.xmax, p8.ymin, p8.ymax, p8.zmin, p8.zmax), p8.gamma, p8.ciso,
                          small.uin))
    d = abi.Domain(pb.params, pb.uin)
    with pytest.raises(abi.DmxError, match="does not match"):
        d.read_dump(path)
    with pytest.raises(abi.DmxError, match="cannot open"):
        d.read_dump(str(tmp_path / "missing.bin"))
