"""The batched-ensemble workload at its real sizes: 48 independent periodic
turbulence boxes with sizes drawn from {32, 48, 64, 96, 128}^3 (seed 4242,
random-phase k^-2 velocity and vector potential, k <= 8, rms Mach 0.5,
beta 10), HLLD/HLLD, MC.  In float64 the batch gives every box the bits of
its own single context, and the smallest box the bits of the oracle; in
float32 bulk every box stays within L1 1e-5 of float64."""
import numpy as np
import pytest

import oracle as O
from dmx import abi
from dmx import problems as P

pytestmark = pytest.mark.gpu

NSTEPS = 5


@pytest.fixture(scope="module")
def ensemble():
    print("[gpu] building the 48-box ensemble (sizes from {32,48,64,96,128}^3, seed 4242)", flush=True)
    return P.ensemble(48)


def _batch(pbs, precision, nsteps):
    doms = [abi.Domain(pb.params.replace(precision=precision), pb.uin) for pb in pbs]
    b = abi.Batch(doms)
    t = b.advance(nsteps, [0.0] * len(doms))
    out = [(tb, d.download()) for tb, d in zip(t, doms)]
    b.close()
    for d in doms:
        d.close()
    return out


def _l1(a, b):
    """Per-variable L1 difference over the L1 norm of the variable (rho, E)
    or of its vector (momentum, B)."""
    a, b = P.interior(a), P.interior(b)
    mom = np.sqrt((b[1:4] ** 2).sum(axis=0)).sum()
    mag = np.sqrt((b[5:8] ** 2).sum(axis=0)).sum()
    norms = [np.abs(b[0]).sum(), mom, mom, mom, np.abs(b[4]).sum(), mag, mag, mag]
    return np.array([np.abs(a[v] - b[v]).sum() / norms[v] for v in range(8)])


def test_ensemble48_fp64_batch_equals_single_contexts_and_oracle(ensemble):
    sizes = sorted({pb.params.nx for pb in ensemble})
    assert sizes == [32, 48, 64, 96, 128] and len(ensemble) == 48
    print(f"[gpu] ensemble float64 batch, {NSTEPS} steps", flush=True)
    got = _batch(ensemble, 0, NSTEPS)
    print("[gpu] ensemble float64 single contexts", flush=True)
    for b, (pb, (tb, ub)) in enumerate(zip(ensemble, got)):
        d = abi.Domain(pb.params, pb.uin)
        ts, _ = d.advance(NSTEPS, 0.0)
        us = d.download()
        d.close()
        assert tb == ts, (b, tb, ts)
        assert np.array_equal(P.interior(ub), P.interior(us)), (b, pb.params.nx)
    smallest = min(range(len(ensemble)), key=lambda b: ensemble[b].params.nx)
    pb = ensemble[smallest]
    print(f"[gpu] ensemble box {smallest} ({pb.params.nx}^3) against the oracle", flush=True)
    o = O.Oracle(pb.params, pb.uin)
    to, _ = o.advance(NSTEPS)
    assert got[smallest][0] == to
    assert np.array_equal(P.interior(got[smallest][1]), P.interior(o.store()))


def test_ensemble48_fp32_batch_within_l1_of_fp64(ensemble):
    print(f"[gpu] ensemble float64 and float32 batches, {NSTEPS} steps", flush=True)
    r64 = _batch(ensemble, 0, NSTEPS)
    r32 = _batch(ensemble, 1, NSTEPS)
    worst = 0.0
    for b, ((t64, u64), (t32, u32)) in enumerate(zip(r64, r32)):
        l1 = _l1(u32, u64)
        worst = max(worst, float(l1.max()))
        assert l1.max() <= 1e-5, (b, ensemble[b].params.nx, l1)
        assert abs(t32 - t64) <= 1e-6 * t64
    print(f"[gpu] ensemble float32 worst per-variable L1 {worst:.3g}", flush=True)
