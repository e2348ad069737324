"""The drop-in boundary end to end: the reference's own Fortran driver
(src/dumses.f90 with its namelist input, restart reader and binary output),
linked against dmx_shim.f90 + libdmx.so instead of its physics files
(dumses-hybrid_amd/fortran/dropin.mk), must write the same dump as the
reference for the same run; the C ABI's restart reader (dmx_read_dump) must
read that Fortran-written file back."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from dmx import problems as P
from dmx.dumpfile import Dump, dump_path, read_dump, write_dump
from goldens import assert_history_matches_printed, history_scales, load_step

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(REPO, "dumses-hybrid_amd", "fortran", "_build")


def exe(iso):
    return os.path.join(BUILD, f"iso{iso}", "dumses_dmx")


def run_dropin(p, uin, dts, nsteps, env_extra=None):
    """Runs the hybrid driver for exactly nsteps (tlim and dtdump placed
    between the known step times) and returns (dump, history text)."""
    t = np.cumsum(dts)
    work = tempfile.mkdtemp(prefix="dmxdropin_")
    try:
        O.write_input(os.path.join(work, "input"), p)
        tlim = 0.5 * ((t[nsteps - 3] if nsteps >= 3 else 0.0) + t[nsteps - 2])
        tdump = 0.5 * (t[nsteps - 2] + t[nsteps - 1])
        txt = open(os.path.join(work, "input")).read()
        txt = txt.replace("tlim=1.d30", f"tlim={O._fmt(tlim)}").replace("dtdump=1.d30", f"dtdump={O._fmt(tdump)}")
        txt = txt.replace("dthist=-1.d0", f"dthist={O._fmt(tdump)}")
        open(os.path.join(work, "input"), "w").write(txt)
        write_dump(dump_path(1, 0, work), Dump(p.nx, p.ny, p.nz, 0.0, 0.0, p.dx, p.dy, p.dz,
                                               (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax),
                                               p.gamma, p.ciso, uin))
        env = dict(os.environ, **(env_extra or {}))
        r = subprocess.run([exe(p.iso)], cwd=work, env=env, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
        d = read_dump(dump_path(2, 0, work))
        hist = open(os.path.join(work, "history.txt")).read()
        # the same file through the C ABI's restart reader
        from dmx import abi
        dom = abi.Domain(p, uin)
        t_abi, dt_abi = dom.read_dump(dump_path(2, 0, work))
        assert (t_abi, dt_abi) == (d.time, d.dt)
        assert np.array_equal(P.interior(dom.download()), P.interior(d.uin))
        dom.close()
        return d, hist, r.stdout
    finally:
        shutil.rmtree(work, ignore_errors=True)


@pytest.mark.skipif(not os.path.exists(exe(0)), reason="drop-in driver not built")
def test_dropin_reports_device_errors_on_the_host(tmp_path):
    """Without a usable device the shim stops with the library's message
    (the reference has no error convention; the shim turns a nonzero status
    into error stop)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present here: the driver would run, not stop")
    pb = P.random_smooth(8, 8, 8, seed=1)
    p = pb.params
    O.write_input(str(tmp_path / "input"), p)
    write_dump(dump_path(1, 0, str(tmp_path)), Dump(p.nx, p.ny, p.nz, 0.0, 0.0, p.dx, p.dy, p.dz,
                                                    (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax),
                                                    p.gamma, p.ciso, pb.uin))
    r = subprocess.run([exe(0)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "dmx:" in r.stdout + r.stderr


@pytest.mark.skipif(not os.path.exists(exe(0)), reason="drop-in driver not built")
def test_dropin_refuses_source_terms(tmp_path):
    """rhs=.true. (source_term between godunov and the ghost fill,
    src/dumses.f90:84-93) has no device counterpart: the shim stops with a
    message before it creates a context, instead of stepping on a stale
    host copy."""
    pb = P.random_smooth(8, 8, 8, seed=1)
    p = pb.params
    O.write_input(str(tmp_path / "input"), p)
    txt = (tmp_path / "input").read_text().replace(" eta=", " rhs=.true.\n eta=")
    (tmp_path / "input").write_text(txt)
    write_dump(dump_path(1, 0, str(tmp_path)), Dump(p.nx, p.ny, p.nz, 0.0, 0.0, p.dx, p.dy, p.dz,
                                                    (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax),
                                                    p.gamma, p.ciso, pb.uin))
    r = subprocess.run([exe(0)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "rhs=.true. (source_term) is not supported" in r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["orszag_tang_16", "step_hlld_zerogradient", "shear_fargo",
                                  "shear_nofargo", "shear_fargo_iso", "diss_both",
                                  "diss_eta_iso_zg", "shear_fargo_diss"])
def test_dropin_driver_matches_reference_bitwise(name):
    p, u, dts, outs, full = load_step(name)
    if not os.path.exists(exe(p.iso)):
        pytest.skip("drop-in driver not built")
    n = int(max(outs))
    print(f"[gpu] drop-in driver {name} {n} steps", flush=True)
    d, hist, out = run_dropin(p, u, dts, n)
    assert abs(d.time - np.cumsum(dts)[n - 1]) == 0.0
    got = P.interior(d.uin)
    assert np.array_equal(got, outs[n]), float(np.abs(got - outs[n]).max())
    # the row the reference's own history routine (history.f90, linked into
    # the drop-in driver) wrote for that same state, against dmx_history on
    # the device and dmxo_history, at the printed precision
    rows, hulps = O.parse_history(hist)
    assert rows.shape[0] == 1 and rows[0, 0] == pytest.approx(d.time, rel=1e-5)
    from dmx import abi
    dom = abi.Domain(p, d.uin)
    scale = history_scales(p, d.uin)
    assert_history_matches_printed(dom.history(), rows[0], hulps[0], scale, name + " (device)")
    assert_history_matches_printed(O.Oracle(p, d.uin).history(), rows[0], hulps[0], scale, name + " (oracle)")
    dom.close()
