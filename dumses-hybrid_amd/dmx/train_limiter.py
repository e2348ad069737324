"""Training and evaluation of the learned slope limiter.

    python -m dmx.train_limiter train [--out PATH] [--steps N] [--boxes B] [--sizes 32,64,96,128]
    python -m dmx.train_limiter eval  [--weights PATH] [--n 32] [--tend T] [--kmax K]

The defaults of train are the recipe of the shipped weights
(data/limiter_weights.npz).

train: an ensemble of decaying-turbulence boxes of uneven sizes is run at
fine resolution by the plain solver (float32 bulk, HLLD/HLLD, MC slopes) as
one batch (dmx_batch_*, every kernel one launch over all boxes); snapshots
of each fine state give coarse/fine training pairs
(dmx.learned.training_pairs), and the network is trained on the GPU by the
library's own forward/backward/Adam kernels (dmx_mlp_*).

eval: from one fine initial state, runs (a) the plain solver at twice the
resolution, restricted to the coarse grid (the reference), (b) the plain
solver (MC) on the coarse grid and (c) the learned limiter on the coarse
grid, to the same time, and reports the L1 difference of (b) and (c) from
(a) per variable, with the wall time of each run.  The network corrects
the van Leer slopes (dmx.learned.limited_slopes): its targets are the
sub-cell slope minus van Leer, over the stencil scale.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import abi
from . import learned as L
from . import problems as P

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
WEIGHTS = os.path.join(DATA, "limiter_weights.npz")


def ensemble_boxes(nboxes: int = 12, seed: int = 4242, sizes=(32, 48, 64)):
    """Fine sizes drawn from `sizes` (cubes; coarse half that), random-phase
    Gaussian v and A (k^-2, k <= n/8 in fine-grid units), Mach and beta
    drawn from {0.5, 1} x {1, 10}."""
    g = np.random.Generator(np.random.PCG64(seed))
    boxes = []
    for b in range(nboxes):
        n = int(g.choice(list(sizes)))
        mach = float(g.choice([0.5, 1.0]))
        beta = float(g.choice([1.0, 10.0]))
        boxes.append(P.turbulence(n, seed=seed + b, mach=mach, beta=beta, kmax=max(4, n // 8),
                                  precision=1))
    return boxes


def run_ensemble(boxes, nsnap: int = 5, every: int = 6):
    """Advances every box in one batch (dmx_batch_*: each kernel of the step
    one launch over the tiles of all boxes) and returns the fine snapshots
    [(uin, gamma)], box by box."""
    doms = [abi.Domain(pb.params, pb.uin) for pb in boxes]
    batch = abi.Batch(doms)
    snaps = [[] for _ in boxes]
    for _ in range(nsnap):
        batch.advance(every)
        for b, d in enumerate(doms):
            snaps[b].append(d.download())
    batch.close()
    for d in doms:
        d.close()
    return [(u, boxes[b].params.gamma) for b in range(len(boxes)) for u in snaps[b]]


def train(out: str = WEIGHTS, steps: int = 6000, batch: int = 4096, seed: int = 0, tvd: bool = True,
          margin: float = 1.0, nboxes: int = 24, sizes=(32, 64, 96, 128), per_snapshot: int = 60000,
          extrema_only: bool = False, log=print):
    """per_snapshot > 0 keeps that many samples of each snapshot, drawn
    at random (bounds the memory of large ensembles)."""
    t0 = time.perf_counter()
    boxes = ensemble_boxes(nboxes, sizes=sizes)
    snaps = run_ensemble(boxes)
    t1 = time.perf_counter()
    gs = np.random.Generator(np.random.PCG64(seed + 99))

    def pairs(u, gam):
        X, T = L.training_pairs(u, gam, tvd=tvd, margin=margin, extrema_only=extrema_only)
        if per_snapshot and X.shape[0] > per_snapshot:
            keep = gs.choice(X.shape[0], per_snapshot, replace=False)
            X, T = X[keep], T[keep]
        return X, T

    nsnap = len(snaps)
    Xs, Ts = [], []
    while snaps:
        X, T = pairs(*snaps.pop(0))
        Xs.append(X)
        Ts.append(T)
        if len(Xs) % 10 == 0:
            log(f"[train] pairs from {len(Xs)} of {nsnap} snapshots")
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    del Xs, Ts
    log(f"[train] {len(boxes)} boxes, {nsnap} snapshots, {X.shape[0]} samples "
        f"(ensemble {t1 - t0:.1f} s, pairs {time.perf_counter() - t1:.1f} s)")
    g = np.random.Generator(np.random.PCG64(seed))
    perm = g.permutation(X.shape[0])
    nval = min(20000, X.shape[0] // 10)
    Xv, Tv = X[perm[:nval]], T[perm[:nval]]
    Xt, Tt = X[perm[nval:]], T[perm[nval:]]
    m = abi.Mlp(max(batch, nval))
    m.set_params(L.init_params(seed))
    m.load_data(Xt, Tt)
    losses = []
    for phase, (n, lr) in enumerate(((steps * 2 // 3, 2e-3), (steps - steps * 2 // 3, 5e-4))):
        tp = time.perf_counter()
        ls = m.train(n, batch, lr, seed=seed + 17 * phase)
        losses.append(ls)
        vl, _ = m.loss_grad(Xv, Tv)
        log(f"[train] phase {phase}: {n} steps lr {lr:g} in {time.perf_counter() - tp:.1f} s, last minibatch "
            f"loss {ls[-50:].mean():.4e}, validation loss {vl:.4e}")
    theta = m.get_params()
    # the loss of a zero network (the van Leer slopes), for scale
    mc = np.zeros_like(Tv)
    base = 0.5 * ((mc - Tv) ** 2).sum() / Tv.shape[0]
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    np.savez(out, theta=theta, losses=np.concatenate(losses), val_loss=np.float64(vl),
             zero_loss=np.float64(base), samples=np.int64(X.shape[0]), tvd_targets=np.bool_(tvd),
             extrema_only=np.bool_(extrema_only),
             tvd_margin=np.float64(margin))
    log(f"[train] wrote {out}: validation loss {vl:.4e} (zero-network loss {base:.4e})")
    m.close()
    return theta


def run_to(d, tend: float, max_steps: int = 100000, log=None, every: int = 100):
    """Advances a Domain to exactly tend (last step shortened) through the
    step-by-step ABI; returns the number of steps (log: progress every
    `every` steps)."""
    t, n = 0.0, 0
    while t < tend and n < max_steps:
        dt = min(d.compute_dt(), tend - t)
        d.godunov(t, dt)
        d.boundary(t, dt)
        t = t + dt
        n += 1
        if log and n % every == 0:
            log(f"[eval] step {n}, t = {t:.4f}")
    return n


def evaluate(theta, n: int = 32, tend: float = 0.25, seed: int = 777, mach: float = 1.0,
             beta: float = 1.0, fine_precision: int = 1, kmax: int | None = None, log=print):
    """L1 error (per variable, relative to the reference's L1 norm) of the
    plain MC and the learned coarse runs (float32 bulk) against the fine run
    at twice the resolution (float32 bulk, or float64 with fine_precision
    0), restricted to the coarse grid; with the wall time of each run."""
    kmax = max(4, n // 4) if kmax is None else kmax
    log(f"[eval] fine initial state {2 * n}^3")
    fine = P.turbulence(2 * n, seed=seed, mach=mach, beta=beta, kmax=kmax, precision=fine_precision,
                        log=log if n >= 128 else None)
    uc0 = L.restrict_state(fine.uin)
    pc = P.Params(**{**fine.params.__dict__}).replace(nx=n, ny=n, nz=n, precision=1)
    res = {}
    t0 = time.perf_counter()
    log(f"[eval] fine run {2 * n}^3 (precision {fine_precision}) to t = {tend}")
    df = abi.Domain(fine.params, fine.uin)
    del fine
    nf = run_to(df, tend, log=log if n >= 128 else None)
    res["fine_wall_s"] = time.perf_counter() - t0
    res["fine_steps"] = nf
    ref = P.interior(L.restrict_state(df.download()))
    df.close()
    for name, st in (("mc", 2), ("learned", 4)):
        log(f"[eval] coarse {name} run {n}^3")
        t0 = time.perf_counter()
        d = abi.Domain(pc.replace(slope_type=st), uc0)
        if st == 4:
            d.set_limiter(theta)
        ns = run_to(d, tend, log=log if n >= 128 else None)
        u = P.interior(d.download())
        res[f"{name}_wall_s"] = time.perf_counter() - t0
        res[f"{name}_steps"] = ns
        l1 = [float(np.abs(u[v] - ref[v]).sum() / max(np.abs(ref[v]).sum(), 1e-300)) for v in range(8)]
        res[f"{name}_l1"] = l1
        res[f"{name}_l1_mean"] = float(np.mean(l1))
        if st == 4:
            _, fb, frep = d.learned_slopes()
            res["learned_fallback"] = fb
            res["learned_tvd_replaced"] = frep
        d.close()
    log(f"[eval] n={n} fine {2 * n}: mean L1 mc {res['mc_l1_mean']:.4e}, learned "
        f"{res['learned_l1_mean']:.4e} (positivity fallback {res['learned_fallback']:.3f}, "
        f"network slopes outside the TVD region {res['learned_tvd_replaced']:.3f})")
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=("train", "eval"))
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--tend", type=float, default=0.25)
    ap.add_argument("--fine-precision", type=int, default=1, help="0: float64 fine run, 1: float32 bulk")
    ap.add_argument("--kmax", type=int, default=None)
    ap.add_argument("--json", default=None, help="also write the result to this file")
    ap.add_argument("--no-tvd-targets", dest="tvd_targets", action="store_false",
                    help="train on unprojected targets (default: projected into the limiter's TVD region)")
    ap.add_argument("--tvd-margin", type=float, default=1.0,
                    help="fraction of the TVD bound the projected targets keep to")
    ap.add_argument("--boxes", type=int, default=24)
    ap.add_argument("--sizes", default="32,64,96,128", help="fine box sizes to draw from")
    ap.add_argument("--per-snapshot", type=int, default=60000, help="samples kept per snapshot (0: all)")
    ap.add_argument("--extrema-only", action="store_true",
                    help="train only on the cells the network acts on (not monotone in all 8 variables)")
    a = ap.parse_args(argv)
    if a.cmd == "train":
        train(a.out, a.steps, tvd=a.tvd_targets, margin=a.tvd_margin, nboxes=a.boxes,
              sizes=tuple(int(x) for x in a.sizes.split(",")), per_snapshot=a.per_snapshot,
              extrema_only=a.extrema_only)
    else:
        theta = np.load(a.weights)["theta"]
        res = evaluate(theta, a.n, tend=a.tend, fine_precision=a.fine_precision, kmax=a.kmax)
        print(res)
        if a.json:
            import json
            with open(a.json, "w") as f:
                json.dump(dict(n=a.n, tend=a.tend, fine_precision=a.fine_precision, **res), f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
