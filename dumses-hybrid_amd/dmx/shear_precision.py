"""float32 bulk against float64 on the shearing-box MRI with FARGO.

    python -m dmx.shear_precision [--orbits 5] [--json out.json]

The workload: 128 x 256 x 128 on a 1 x 4 x 1 box, Omega = 1, q = 1.5,
rho = 1, P = 1e-3, net vertical field at beta = 1000, velocity
perturbations of 1e-3 c_s (dmx.problems.shearing_box).  Both precisions
advance side by side through the step-by-step ABI in the reference's order
(compute_dt, godunov, fargo_update, boundary; src/dumses.f90:80-120), each
landing exactly on every orbit t = 2 pi m by shortening its last step.  At
each orbit the script reports the per-variable L1 difference (relative to
the L1 norm of rho, of E, of the B vector, and of rho c_s for momentum) and
the box-averaged MRI diagnostics of both runs (the sums of the MRI
problem's history, dmx_history_shear): magnetic energy, Maxwell and
Reynolds stress.  Once the instability grows out of its linear phase the
two runs diverge point by point, as any two runs whose round-off differs
do; the diagnostics are then the comparison that means something.

Two float64 controls run beside them and are compared with the float64 run
the same way: the initial state with every hydro value moved by one unit in
the last place (np.nextafter), and the initial state rounded to float32 (the
rounding the float32 run starts from, with float64 arithmetic after it).
They measure how much the flow itself amplifies a difference of that size.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import abi
from . import problems as P


def step_to(d, t: float, tend: float) -> tuple[float, int]:
    """Advances a Domain from t to exactly tend; returns (t, steps)."""
    n = 0
    while t < tend:
        dt = min(d.compute_dt(), tend - t)
        d.godunov(t, dt)
        d.fargo(dt)
        d.boundary(t, dt)
        t = t + dt
        n += 1
    return t, n


def _l1(u, ref, cs):
    """Per-variable L1 difference relative to the L1 norm of rho, of E, of
    the B vector, and of rho c_s for momentum."""
    mom = (np.abs(ref[0]) * cs).sum()
    mag = np.sqrt((ref[5:8] ** 2).sum(axis=0)).sum()
    norms = [np.abs(ref[0]).sum(), mom, mom, mom, np.abs(ref[4]).sum(), mag, mag, mag]
    return [float(np.abs(u[v] - ref[v]).sum() / norms[v]) for v in range(8)]


def run(orbits: int = 5, nx: int = 128, ny: int = 256, nz: int = 128, log=print, controls: bool = True) -> dict:
    pb = P.shearing_box(nx, ny, nz, fargo=1)
    cs = np.sqrt(pb.params.gamma * 1e-3)
    doms = [abi.Domain(pb.params.replace(precision=prec), pb.uin) for prec in (0, 1)]
    if controls:
        u_ulp = pb.uin.copy()
        u_ulp[:5] = np.nextafter(u_ulp[:5], np.inf)
        u_f32 = pb.uin.astype(np.float32).astype(np.float64)
        doms += [abi.Domain(pb.params, u_ulp), abi.Domain(pb.params, u_f32)]
    times = [0.0] * len(doms)
    rows = []
    t0 = time.perf_counter()
    for m in range(1, orbits + 1):
        tend = 2.0 * np.pi * m
        steps = []
        for q, d in enumerate(doms):
            times[q], n = step_to(d, times[q], tend)
            steps.append(n)
        u64, u32 = (P.interior(d.download()) for d in doms[:2])
        l1 = _l1(u32, u64, cs)
        h64, h32 = (d.history(shear=True) for d in doms[:2])
        row = dict(orbit=m, steps_f64=steps[0], steps_f32=steps[1], l1=l1, l1_max=max(l1),
                   magp=[float(h64[4]), float(h32[4])], maxwell=[float(h64[1]), float(h32[1])],
                   reynolds=[float(h64[2]), float(h32[2])], mass=[float(h64[0]), float(h32[0])])
        if controls:
            row["l1_max_ulp_control"] = max(_l1(P.interior(doms[2].download()), u64, cs))
            row["l1_max_f32ic_control"] = max(_l1(P.interior(doms[3].download()), u64, cs))
        rows.append(row)
        ctl = (f", float64 controls: 1-ulp {row['l1_max_ulp_control']:.3e}, float32-rounded start "
               f"{row['l1_max_f32ic_control']:.3e}") if controls else ""
        log(f"[shear] orbit {m}: steps {steps[0]}/{steps[1]}, max L1 {max(l1):.3e}{ctl}, "
            f"magnetic energy {h64[4]:.4e} / {h32[4]:.4e}, Maxwell {h64[1]:.4e} / {h32[1]:.4e}, "
            f"Reynolds {h64[2]:.4e} / {h32[2]:.4e} ({time.perf_counter() - t0:.1f} s)")
    for d in doms:
        d.close()
    return dict(workload=f"shearing-box MRI {nx}x{ny}x{nz}, FARGO, HLLD/HLLD, MC, CFL 0.7",
                orbits=rows)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--orbits", type=int, default=5)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    res = run(a.orbits)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
