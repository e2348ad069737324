"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

from dmx import problems as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

STEP_FIXTURES = ["step_hlld_mc", "step_hll_llf_minmod", "step_llf_upwind_multid",
                 "step_upwind_hlla_flat", "step_hlld_zerogradient", "orszag_tang_16",
                 "shear_nofargo", "shear_fargo", "shear_iso", "shear_fargo_iso",
                 "diss_both", "diss_eta_iso_zg", "diss_nu_hll_minmod", "shear_fargo_diss"]


def params_from_json(s) -> P.Params:
    d = json.loads(str(s))
    d["bdtype"] = tuple(d["bdtype"])
    return P.Params(**d)


def load_step(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    p = params_from_json(z["params"])
    ic = z["ic"]
    u = np.zeros((11, p.nz + 6, p.ny + 6, p.nx + 6))
    u[:8, 3:-3, 3:-3, 3:-3] = ic
    P.fill_periodic(u)
    outs = {int(s): z[f"out_{int(s):04d}"] for s in z["steps"]}
    full = z["full_last"] if z["full_last"].size else None
    return p, u, z["dts"], outs, full


HIST_COLUMNS = ["mass", "maxwell", "reynolds", "max+rey", "magp", "meanBx", "meanBy", "meanBz", "divB"]


def load_history_cases():
    """The states of history_reference.npz with the row the reference's own
    history routine printed for each: [(name, params, uin, time, row, hulp)],
    row/hulp with the columns time, dt, mass .. divB."""
    z = np.load(os.path.join(GOLDEN, "history_reference.npz"), allow_pickle=False)
    out = []
    for name in z["cases"]:
        name = str(name)
        p = params_from_json(z[f"{name}_params"])
        u = np.zeros((11, p.nz + 6, p.ny + 6, p.nx + 6))
        u[:8, 3:-3, 3:-3, 3:-3] = z[f"{name}_ic"]
        P.fill_periodic(u)
        out.append((name, p, u, float(z[f"{name}_row"][0]), z[f"{name}_row"], z[f"{name}_hulp"]))
    return out


def history_scales(p, u):
    """Per history column (mass .. divB) the sum of the magnitudes of its
    terms: the scale of the round-off of a sum taken in another order."""
    nx, ny, nz = p.nx, p.ny, p.nz
    sl = (slice(3, 3 + nz), slice(3, 3 + ny), slice(3, 3 + nx))
    dtau = p.dx * p.dy * p.dz / ((p.xmax - p.xmin) * (p.ymax - p.ymin) * (p.zmax - p.zmin))
    r = u[0][sl]
    bxl, byl, bzl = u[5][sl], u[6][sl], u[7][sl]
    bx = bxl + u[5][3:3 + nz, 3:3 + ny, 4:4 + nx]
    by = byl + u[6][3:3 + nz, 4:4 + ny, 3:3 + nx]
    bz = bzl + u[7][4:4 + nz, 3:3 + ny, 3:3 + nx]
    vx, vz = u[1][sl] / r, u[3][sl] / r
    dvx = vx - vx.mean(axis=(0, 1), keepdims=True)
    dvz = vz - vz.mean(axis=(0, 1), keepdims=True)
    mx = np.abs(0.25 * bx * bz).sum() * dtau
    re = np.abs(r * dvx * dvz).sum() * dtau
    divb = (np.abs(bx - 2 * bxl) / p.dx + np.abs(by - 2 * byl) / p.dy + np.abs(bz - 2 * bzl) / p.dz).sum()
    return np.array([np.abs(r).sum() * dtau, mx, re, mx + re,
                     0.125 * (bx ** 2 + by ** 2 + bz ** 2).sum() * dtau,
                     np.abs(bxl).sum() * dtau, np.abs(byl).sum() * dtau, np.abs(bzl).sum() * dtau,
                     divb + (np.abs(bxl) / p.dx + np.abs(byl) / p.dy + np.abs(bzl) / p.dz).sum()])


def assert_history_matches_printed(got, row, hulp, scale, what=""):
    """got (mass .. divB) equals the reference's printed row to the printed
    precision: within half a unit of the last printed digit, plus the
    round-off of a sum in another order (1e-13 of the sum of magnitudes)."""
    for q in range(9):
        tol = hulp[2 + q] + 1e-13 * scale[q]
        assert abs(got[q] - row[2 + q]) <= tol, (what, HIST_COLUMNS[q], got[q], row[2 + q], tol)


def assert_f32_close(u, us, rtol=5e-6):
    """float32 states that may differ by the contraction of inlined copies
    of the same arithmetic: within rtol of the largest magnitude of the
    variable (rho, E) or of its vector (momentum, B)."""
    iu, ir = P.interior(u), P.interior(us)
    mom = np.sqrt((ir[1:4] ** 2).sum(axis=0)).max()
    mag = np.sqrt((ir[5:8] ** 2).sum(axis=0)).max()
    scales = [np.abs(ir[0]).max(), mom, mom, mom, np.abs(ir[4]).max(), mag, mag, mag]
    for v in range(8):
        assert np.abs(iu[v] - ir[v]).max() <= rtol * scales[v], v
