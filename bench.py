"""Throughput of the MI355X MHD step on the 256^3 Orszag-Tang workload.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

A step is one pass of the reference time loop (compute_dt, godunov with
HLLD/HLLD fluxes and MC slopes, boundary) over the whole 256^3 periodic box,
state resident in HBM.  value = zone-cycles/s of the float64 path (the
reference's arithmetic); the float32-bulk path is reported beside it with its
per-variable L1 difference from float64.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "dumses-hybrid_amd"))

METRIC = "zone-cycles/s at 256^3 HLLD MHD; FP32 and learned vs FP64 at fixed L1 error"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# algorithmic bytes per cell of each kernel group (inputs read once, outputs
# written once; 8-byte words in float64, T-size words for the bulk in float32)
def algo_bytes(group: str, wt: int, wb: int = 8) -> float:
    hydro, bfield = 5 * wt, 3 * wb
    table = {
        "compute_dt": hydro + bfield,
        "trace": hydro + bfield + 47 * wt,
        "faces": 2 * 47 * wt // 2 + 3 * 6 * wt,        # summary read, 3 faces x 6 words
        "edges": 47 * wt + 3 * wt,                      # summary read, 3 EMFs written
        "update_ct": 2 * (hydro + bfield) + 3 * 6 * wt + 3 * wt,
        "boundary": 0.0,
    }
    return float(table.get(group, 0.0))


def run_gpu(pb, steps, warmup, precision, profile_steps):
    from dmx import abi
    p = pb.params.replace(precision=precision)
    d = abi.Domain(p, pb.uin)
    t, _ = d.advance(warmup, 0.0)
    t0 = time.perf_counter()
    t, dts = d.advance(steps, t)
    wall = time.perf_counter() - t0
    dev_ms = d.last_advance_ms()
    prof = d.profile(profile_steps, t) if profile_steps else {}
    state = d.download()
    d.close()
    return dict(wall=wall, dev_ms=dev_ms, dts=dts, prof=prof, state=state)


def cpu_baseline(n=64, steps=6):
    """The compiled reference (serial OpenMP build timed with 1 thread, its
    fastest setting) on a bounded sample of the same workload: Orszag-Tang
    at n^3 for a few steps.  Falls back to the C restatement."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    from dmx import problems as P
    pb = P.orszag_tang(n)
    exe = os.path.join(O.REF_DIR, "iso0_omp", "ref_harness")
    if os.path.exists(exe):
        import shutil
        import subprocess
        import tempfile
        from dmx.dumpfile import Dump, dump_path, write_dump
        p = pb.params
        work = tempfile.mkdtemp(prefix="dmxcpu_")
        try:
            O.write_input(os.path.join(work, "input"), p)
            with open(os.path.join(work, "harness"), "w") as f:
                f.write(f"&harness_params\n nsteps={steps}\n dumpevery=0\n mode='step'\n/\n")
            write_dump(dump_path(1, 0, work), Dump(p.nx, p.ny, p.nz, 0.0, 0.0, p.dx, p.dy, p.dz,
                                                   (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax),
                                                   p.gamma, p.ciso, pb.uin))
            env = dict(os.environ, OMP_NUM_THREADS="1")
            t0 = time.perf_counter()
            subprocess.run([exe], cwd=work, env=env, check=True, capture_output=True)
            wall = time.perf_counter() - t0
        finally:
            shutil.rmtree(work, ignore_errors=True)
        kind = "reference"
    else:
        o = O.Oracle(pb.params, pb.uin)
        t0 = time.perf_counter()
        o.advance(steps)
        wall = time.perf_counter() - t0
        kind = "port"
    return dict(value=n ** 3 * steps / wall, unit="zone-cycles/s", cores=1, kind=kind,
                sample=f"Orszag-Tang {n}^3, {steps} steps incl. init I/O, 1 thread")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain 1 GiB device copy (torch) to calibrate HBM byte counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs one domain on one GPU; --gpus N>1 is not "
                                   "supported (the multi-GPU decomposition is design-only)"}))
        sys.exit(2)
    from dmx import problems as P
    n = a.n
    if a.calibrate:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 3 x 1 GiB read + 1 GiB write")
    log(f"[bench] building Orszag-Tang {n}^3 initial state")
    pb = P.orszag_tang(n)
    cells = n ** 3
    log("[bench] float64 run")
    r64 = run_gpu(pb, a.steps, a.warmup, 0, a.profile_steps)
    v64 = cells * a.steps / r64["wall"]
    log(f"[bench] float64: {v64:.4g} zone-cycles/s, {1e3 * r64['wall'] / a.steps:.3f} ms/step")
    out = {}
    if not a.no_fp32:
        log("[bench] float32-bulk run")
        r32 = run_gpu(pb, a.steps, a.warmup, 1, a.profile_steps)
        v32 = cells * a.steps / r32["wall"]
        u64, u32 = P.interior(r64["state"]), P.interior(r32["state"])
        names = ["rho", "rhou", "rhov", "rhow", "E", "Bx", "By", "Bz"]
        # per-variable L1 difference over the L1 norm of the variable (rho, E)
        # or of its vector (momentum, B; Bz of Orszag-Tang starts at zero)
        mom = np.sqrt((u64[1:4] ** 2).sum(axis=0)).sum()
        mag = np.sqrt((u64[5:8] ** 2).sum(axis=0)).sum()
        norms = [np.abs(u64[0]).sum(), mom, mom, mom, np.abs(u64[4]).sum(), mag, mag, mag]
        l1 = {names[v]: float(np.abs(u32[v] - u64[v]).sum() / norms[v]) for v in range(8)}
        out["fp32"] = dict(value=v32, ms_per_step=1e3 * r32["wall"] / a.steps, speedup_vs_fp64=v32 / v64,
                           l1_rel_vs_fp64=l1, kernel_ms=r32["prof"],
                           dtype="f32 bulk, f64 face-B/CT/pressure recovery/sums")
        log(f"[bench] float32: {v32:.4g} zone-cycles/s ({v32 / v64:.2f}x), L1 {l1}")
    prof = r64["prof"]
    dom = max(prof, key=prof.get) if prof else None
    roof = None
    if dom:
        ms = prof[dom]
        bpc = algo_bytes(dom, 8)
        ach = bpc * cells / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        roof = dict(bound="hbm", kernel=dom, achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=ach / HBM_PEAK_GBS, traffic=None, algo_bytes_per_cell=bpc,
                    kernel_ms=ms, step_kernel_ms=prof)
    res = {
        "metric": METRIC, "value": v64, "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * r64["wall"] / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D Orszag-Tang {n}^3 periodic, HLLD/HLLD, MC slopes, CFL 0.7",
                   "cells": cells, "parallelism": "single domain, 1 GPU"},
        "device_ms_per_step": r64["dev_ms"] / a.steps,
        "roofline": roof,
    }
    res.update(out)
    if not a.no_cpu:
        log("[bench] CPU baseline")
        res["cpu_baseline"] = cpu_baseline()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
