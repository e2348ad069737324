"""Throughput of the MI355X MHD step on the 256^3 Orszag-Tang workload.

    python bench.py [--gpus 1] [--steps 100] [--warmup 5]

A step is one pass of the reference time loop (compute_dt, godunov with
HLLD/HLLD fluxes and MC slopes, boundary) over the whole 256^3 periodic box,
state resident in HBM.  value = zone-cycles/s of the float64 path (the
reference's arithmetic); the float32-bulk path is reported beside it with its
per-variable L1 difference from float64, and the float32 path with the
learned limiter beside that.  Prints one JSON line.
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


# Algorithmic bytes per zone-cycle: one read and one write of the stored
# words of every interior cell (ghosts, dt and re-reads excluded): 8 float64
# words in float64 mode; in float32-bulk mode rho, rho*u, rho*v, rho*w in
# float32 and E, Bx, By, Bz in float64 (4 x 4 B + 4 x 8 B).
ALG_BYTES_PER_ZONE_CYCLE = {"f64": 128, "f32": 96}


# bytes per interior cell of each kernel group's own staging, its inputs
# read once and its outputs written once (8-byte words in float64; T-size
# words for the bulk in float32, face B and E always float64).  The trace
# summary has 47 words per cell, the face fluxes 18, the edge EMFs 3.  This
# is this design's intermediate traffic, not algorithmic work: reported as
# stage_bytes_per_cell beside the roofline.
def stage_bytes(group: str, wt: int, wb: int = 8) -> float:
    state = 4 * wt + 4 * wb
    table = {
        "compute_dt": state,
        "trace": state + 47 * wt,
        "faces": 47 * wt + 18 * wt,
        "edges": 47 * wt + 3 * wt,
        "flux": 47 * wt + 18 * wt + 3 * wt,           # faces and edges in one pass
        "fused": state + 18 * wt + 3 * wt,            # trace + faces + edges in one pass
        "update_ct": 2 * state + 18 * wt + 3 * wt,
        # learned limiter: state -> 8 float primitives -> 24 float slopes
        "learned_limiter": state + 8 * 4 + 8 * 4 + 24 * 4,
        "boundary": 0.0,
    }
    return float(table.get(group, 0.0))


# bytes leaving L2 per interior cell and launch, (2 x FETCH_SIZE + WRITE_SIZE)
# x 1 KiB from the two rocprofv3 counter runs at 256^3 summarised in
# profiles/mhd_step_ot256_hbm_bytes.md (FETCH_SIZE scaled by 2.00, WRITE_SIZE
# by 1.00 against the known-size calibration kernels of the same runs), for
# the launch shapes the autotuner keeps at 256^3.  Constants copied from
# that file; COUNTERS_SOURCE names the file and the commit whose kernels it
# measured.
COUNTERS_SOURCE = ("profiles/mhd_step_ot256_hbm_bytes.md, profiles/mhd_step_ot256_sq_counters.md, "
                   "profiles/fused_trace_flux_ot256.md (kernel code of commit 0f60dc7; the files name the launch shapes)")
MEASURED_TRAFFIC = {
    "f64": {"trace": 570.6, "flux": 1876.8, "update_ct": 414.4, "compute_dt": 86.3,
            "boundary": 17.7},
    "f32": {"fused": 234.5, "update_ct": 295.0, "compute_dt": 69.5, "boundary": 14.8},
}


# VALU-busy fraction of each kernel group at 256^3 (SQ_ACTIVE_INST_VALU x 4
# / 1024 SIMDs / kernel cycles) from the rocprofv3 SQ counter runs summarised
# in profiles/mhd_step_ot256_sq_counters.md and profiles/fused_trace_flux_ot256.md:
# the bound that applies where the HBM fraction is low (bit-exact float64
# division and square root)
MEASURED_VALU_BUSY = {
    "f64": {"trace": 0.504, "flux": 0.702, "update_ct": 0.371, "compute_dt": 0.730, "boundary": 0.049,
            "fused": 0.502},
    "f32": {"trace": 0.413, "flux": 0.640, "update_ct": 0.232, "compute_dt": 0.401, "boundary": 0.057,
            "fused": 0.608},
}


def roofline(prof: dict, cells: int, prec: str) -> dict | None:
    """HBM roofline of the dominant kernel group of a profiled run: the
    algorithmic bytes of the zone-cycles one launch advances (cells x
    ALG_BYTES_PER_ZONE_CYCLE) over that launch's mean duration (HIP events on
    the library's stream), against the 8 TB/s HBM peak."""
    if not prof:
        return None
    prof = dict(prof)
    if prof.get("faces", 0.0) > 0.0:
        # separate face and edge kernels (launch mode 1): one face/edge stage
        prof["flux"] = prof.pop("faces") + prof.pop("edges")
    dom = max(prof, key=prof.get)
    ms = prof[dom]
    b = ALG_BYTES_PER_ZONE_CYCLE[prec]
    ach = b * cells / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    meas = MEASURED_TRAFFIC.get(prec, {}).get(dom)
    traffic = meas * cells if meas is not None else None
    return dict(bound="hbm", kernel=dom, achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=ach / HBM_PEAK_GBS, traffic=traffic, algo_bytes_per_zone_cycle=b, cells=cells,
                stage_bytes_per_cell=stage_bytes(dom, 8 if prec == "f64" else 4),
                valu_busy=MEASURED_VALU_BUSY.get(prec, {}).get(dom),
                counters_source=COUNTERS_SOURCE if meas is not None else None,
                traffic_gbs=(traffic / (ms * 1e-3) / 1e9 if traffic and ms > 0 else None),
                kernel_ms=ms, step_kernel_ms=prof)


# the launch shapes the autotuner keeps at 256^3 on MI355X ((face/edge mode,
# planes), (trace mode, planes)); --no-tune applies them without timing, so
# that profiler runs measure the kernels of the timed run (under a profiler
# the autotuner's timings, and so its choice, can change)
TUNED_256 = {0: ((0, 2), (1, 16)), 1: ((3, 16), (1, 32))}
# --fused: the trace and the face/edge stage in one kernel (mode 3, the trace
# summary kept in LDS), 16 planes per block, in place of the tuned pair
FUSED_SHAPE = ((3, 16), (1, 16))


REPEATS = 5


def run_gpu(pb, steps, warmup, precision, limiter=None, tune=True, repeats=REPEATS, shape=None):
    """Warm-up, then `repeats` timed regions of `steps` steps each of the
    instrumented advance (HIP events around every kernel group on the
    library's stream, read once after the run), each bracketed by device
    synchronisation; the median repeat is reported, per-kernel times are
    the medians over the repeats."""
    from dmx import abi
    p = pb.params.replace(precision=precision)
    if limiter is not None:
        p = p.replace(slope_type=4)
    d = abi.Domain(p, pb.uin)
    if limiter is not None:
        d.set_limiter(limiter)
    if shape is not None:
        d.set_shape(*shape)
        tuned_ms = None
        log(f"[bench] fixed shapes: face/edge {d.shape}, trace {d.trace_shape}")
    elif tune:
        tuned_ms = d.tune()                # launch-shape autotuner (state unchanged)
        log(f"[bench] tuned face/edge stage: mode {d.shape[0]}, {d.shape[1]} planes, {tuned_ms:.3f} ms; "
            f"trace: mode {d.trace_shape[0]}, {d.trace_shape[1]} planes")
    else:
        d.set_shape(*TUNED_256[precision])
        tuned_ms = None
        log(f"[bench] fixed shapes: face/edge {d.shape}, trace {d.trace_shape}")
    t, _ = d.advance(warmup, 0.0)          # ends with a stream synchronisation
    walls, devs, profs = [], [], []
    for r in range(repeats):
        t0 = time.perf_counter()
        prof = d.profile(steps, t)         # ends with a stream synchronisation
        walls.append(time.perf_counter() - t0)
        devs.append(d.last_advance_ms())
        profs.append(prof)
        t = d.time
        log(f"[bench]   repeat {r + 1}/{repeats}: {1e3 * walls[-1] / steps:.3f} ms/step")
    med = sorted(range(repeats), key=lambda q: walls[q])[repeats // 2]
    prof = {k: float(np.median([pr[k] for pr in profs])) for k in profs[0]}
    state = d.download()
    d.close()
    return dict(wall=walls[med], walls=walls, dev_ms=devs[med], prof=prof, state=state, tuned_ms=tuned_ms,
                shape=d.shape, trace_shape=d.trace_shape, steps_total=warmup + repeats * steps)


def run_ensemble(steps=50, warmup=2, nboxes=48):
    """The batched ensemble: 48 periodic turbulence boxes with sizes drawn
    from {32,48,64,96,128}^3 (seed 4242, Mach 0.5, beta 10), float32 bulk,
    advanced `steps` steps by one batch (dmx_batch_*, 2 streams, autotuned
    shape) and, for comparison, by one context after another."""
    from dmx import abi
    from dmx import problems as P
    boxes = P.ensemble(nboxes, precision=1)
    cells = sum(pb.params.nx * pb.params.ny * pb.params.nz for pb in boxes)
    doms = [abi.Domain(pb.params, pb.uin) for pb in boxes]
    for d in doms:
        d.advance(warmup, 0.0)
    t0 = time.perf_counter()
    for d in doms:
        d.advance(steps)
    seq = cells * steps / (time.perf_counter() - t0)
    for d in doms:
        d.close()
    doms = [abi.Domain(pb.params, pb.uin) for pb in boxes]
    b = abi.Batch(doms)
    b.tune()
    b.advance(warmup, [0.0] * nboxes)
    t0 = time.perf_counter()
    b.advance(steps)
    bat = cells * steps / (time.perf_counter() - t0)
    shape, tshape = b.shape, b.trace_shape
    b.close()
    for d in doms:
        d.close()
    return dict(workload=f"{nboxes} periodic turbulence boxes, sizes from {{32,48,64,96,128}}^3 "
                         f"(seed 4242), float32 bulk, HLLD/HLLD, MC, {steps} steps",
                cells=cells, batched=bat, per_context=seq, speedup=bat / seq, unit="zone-cycles/s",
                shape=list(shape), trace_shape=list(tshape))


def cpu_baseline(n=128, steps=3, threads=(1, 2, 4, 8, 16)):
    """The compiled reference (its OpenMP -O3 build) on a bounded sample of
    the same workload: Orszag-Tang at n^3 for a few steps, timed by the
    harness around its step loop only (no initialisation or dump I/O), at
    each OpenMP thread count of `threads` up to the host's CPUs; the best
    is reported, with every value.  Falls back to the C restatement
    (1 thread) when the reference build is absent."""
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
        sweep = {}
        try:
            O.write_input(os.path.join(work, "input"), p)
            with open(os.path.join(work, "harness"), "w") as f:
                f.write(f"&harness_params\n nsteps={steps}\n dumpevery=0\n mode='step'\n/\n")
            write_dump(dump_path(1, 0, work), Dump(p.nx, p.ny, p.nz, 0.0, 0.0, p.dx, p.dy, p.dz,
                                                   (p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax),
                                                   p.gamma, p.ciso, pb.uin))
            ncpu = os.cpu_count() or 1
            for t in threads:
                if t > ncpu:
                    break
                env = dict(os.environ, OMP_NUM_THREADS=str(t))
                subprocess.run([exe], cwd=work, env=env, check=True, capture_output=True)
                wall = float(open(os.path.join(work, "loop_seconds.txt")).read())
                sweep[t] = n ** 3 * steps / wall
                log(f"[bench] CPU reference, {t} thread(s): {sweep[t]:.4g} zone-cycles/s")
        finally:
            shutil.rmtree(work, ignore_errors=True)
        best = max(sweep, key=sweep.get)
        return dict(value=sweep[best], unit="zone-cycles/s", cores=best, kind="reference",
                    sample=f"Orszag-Tang {n}^3 (the same initial field as the 256^3 workload at 1/8 of "
                           f"its cells), {steps} steps (compute_dt+godunov+boundary), step loop only, "
                           f"reference OpenMP -O3 build; best of OMP_NUM_THREADS in {sorted(sweep)}",
                    sample_cells=n ** 3, sample_steps=steps, host_cpus=ncpu,
                    sweep_note=f"the sweep stops at {max(sweep)} threads: this host shows {ncpu} CPUs, "
                               f"a GPU job's share is 16, and the reference's throughput falls with "
                               f"threads beyond 1-2 (OpenMP regions entered inside every routine)",
                    threads_sweep={str(k): v for k, v in sweep.items()})
    o = O.Oracle(pb.params, pb.uin)
    t0 = time.perf_counter()
    o.advance(steps)
    wall = time.perf_counter() - t0
    return dict(value=n ** 3 * steps / wall, unit="zone-cycles/s", cores=1, kind="port",
                sample=f"Orszag-Tang {n}^3, {steps} steps, C restatement, 1 thread, {wall:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--repeats", type=int, default=REPEATS,
                    help="timed regions of --steps steps each; the median is reported")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--workload", choices=("ot", "shear"), default="ot",
                    help="ot: 3D Orszag-Tang n^3 (the metric's workload); shear: the shearing-box "
                         "MRI with FARGO at 128x256x128 (a parity workload, reported for reference)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="apply the launch shapes the autotuner keeps at 256^3 instead of tuning")
    ap.add_argument("--fused", action="store_true",
                    help="run the trace and the face/edge stage as one kernel (launch mode 3) "
                         "instead of the tuned pair")
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
        # 1 GiB written by a fill kernel, then 1 GiB read + 1 GiB written
        # by an elementwise kernel: the known sizes against which the HBM
        # byte counters (FETCH_SIZE, WRITE_SIZE) are scaled
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration: 1 GiB fill, 1 GiB read + 1 GiB write elementwise")
    if a.workload == "shear":
        log("[bench] building the shearing-box MRI 128x256x128 initial state")
        pb = P.shearing_box(128, 256, 128, fargo=1)
        workload = "shearing-box MRI 128x256x128, FARGO, HLLD/HLLD, MC, CFL 0.7"
    else:
        log(f"[bench] building Orszag-Tang {n}^3 initial state")
        pb = P.orszag_tang(n)
        workload = f"3D Orszag-Tang {n}^3 periodic, HLLD/HLLD, MC slopes, CFL 0.7"
    cells = pb.params.nx * pb.params.ny * pb.params.nz
    log("[bench] float64 run")
    fixed = FUSED_SHAPE if a.fused else None
    r64 = run_gpu(pb, a.steps, a.warmup, 0, tune=not a.no_tune, repeats=a.repeats, shape=fixed)
    v64 = cells * a.steps / r64["wall"]
    log(f"[bench] float64: {v64:.4g} zone-cycles/s, {1e3 * r64['wall'] / a.steps:.3f} ms/step "
        f"(median of {a.repeats})")
    out = {}
    if not a.no_fp32:
        log("[bench] float32-bulk run")
        r32 = run_gpu(pb, a.steps, a.warmup, 1, tune=not a.no_tune, repeats=a.repeats, shape=fixed)
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
                           ms_per_step_repeats=[1e3 * w / a.steps for w in r32["walls"]],
                           l1_rel_vs_fp64=l1,
                           dtype="f32 bulk; f64 E, face B/CT, pressure recovery, dt and history sums")
        log(f"[bench] float32: {v32:.4g} zone-cycles/s ({v32 / v64:.2f}x), L1 {l1}")
    if not a.no_fp32 and not a.no_learned:
        from dmx import train_limiter as TL
        if os.path.exists(TL.WEIGHTS):
            log("[bench] float32-bulk run with the learned limiter")
            theta = np.load(TL.WEIGHTS)["theta"]
            rl = run_gpu(pb, a.steps, a.warmup, 1, limiter=theta, tune=not a.no_tune, repeats=1, shape=fixed)
            vl = cells * a.steps / rl["wall"]
            ul = P.interior(rl["state"])
            l1l = {names[v]: float(np.abs(ul[v] - u64[v]).sum() / norms[v]) for v in range(8)}
            out["learned"] = dict(value=vl, ms_per_step=1e3 * rl["wall"] / a.steps, speedup_vs_fp64=vl / v64,
                                  l1_rel_vs_fp64_mc=l1l, kernel_ms=rl["prof"],
                                  note="float32 bulk with the learned limiter (slope_type 4); its L1 is "
                                       "against the float64 MC run, so it measures the limiter change")
            log(f"[bench] learned: {vl:.4g} zone-cycles/s ({vl / v64:.2f}x)")
    roof = roofline(r64["prof"], cells, "f64")
    if "fp32" in out:
        out["fp32"]["roofline"] = roofline(r32["prof"], cells, "f32")
    res = {
        "metric": METRIC, "value": v64, "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * r64["wall"] / a.steps,
        "repeats": a.repeats, "ms_per_step_repeats": [1e3 * w / a.steps for w in r64["walls"]],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "cells": cells, "parallelism": "single domain, 1 GPU"},
        "device_ms_per_step": r64["dev_ms"] / a.steps,
        "tuned_flux_shape": {"f64": list(r64["shape"]), **({"f32": list(r32["shape"])} if "fp32" in out else {})},
        "tuned_trace_shape": {"f64": list(r64["trace_shape"]),
                              **({"f32": list(r32["trace_shape"])} if "fp32" in out else {})},
        "roofline": roof,
        # the whole step against the compulsory state traffic: one read and one
        # write of the 8 stored words of every cell, 128 B per zone-cycle
        "step_roofline": {"bytes_per_zone_cycle": 128, "achieved": 128 * v64 / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": 128 * v64 / 1e9 / HBM_PEAK_GBS},
    }
    res.update(out)
    if not a.no_ensemble:
        log("[bench] batched ensemble")
        res["ensemble"] = run_ensemble()
        log(f"[bench] ensemble: batched {res['ensemble']['batched']:.4g}, per context "
            f"{res['ensemble']['per_context']:.4g} zone-cycles/s")
    if not a.no_cpu:
        log("[bench] CPU baseline")
        res["cpu_baseline"] = cpu_baseline()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
