"""Batched ensemble (dmx_batch_*): boxes of uneven sizes, limiters and
boundary types advanced by one launch per kernel give, box by box, the bits
of dmx_advance on each box's own context, in both precisions and for every
Riemann solver pair; what a batch cannot hold is rejected."""
import numpy as np
import pytest

from dmx import abi
from dmx import problems as P
from goldens import assert_f32_close

pytestmark = pytest.mark.gpu


def boxes(precision, riemann=P.HLLD, riemann2d=P.HLLD):
    kw = dict(precision=precision, riemann=riemann, riemann2d=riemann2d)
    return [
        P.turbulence(16, seed=1, mach=1.0, beta=1.0, kmax=4, **kw),
        P.random_smooth(24, 20, 12, seed=2, slope_type=1, **kw),
        P.random_smooth(37, 9, 18, seed=3, slope_type=3,
                        bdtype=(P.ZERO_GRADIENT, P.PERIODIC, P.ZERO_GRADIENT), **kw),
        P.orszag_tang(8, **kw),
    ]


def singles(pbs, nsteps):
    out = []
    for pb in pbs:
        d = abi.Domain(pb.params, pb.uin)
        t, _ = d.advance(nsteps, 0.0)
        out.append((t, d.download()))
        d.close()
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_contexts(precision):
    """Same launch shapes as the single contexts (the batch starts with
    context 0's): bit for bit in both precisions.  A single periodic
    context fuses the next dt into its update kernel and the batch does
    not; the update, CT and dt arithmetic is compiled without contraction
    or reassociation, so the two give the same bits in float32 too."""
    print(f"[gpu] batch of 4 uneven boxes, precision {precision}", flush=True)
    pbs = boxes(precision)
    ref = singles(pbs, 4)
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    b = abi.Batch(doms)
    b.advance(3, [0.0] * len(doms))
    t = b.advance(1)                    # continues from the stored times
    for d, tb, (ts, us) in zip(doms, t, ref):
        assert tb == ts
        assert np.array_equal(d.download(), us)
    assert b.last_ms() > 0.0
    b.close()
    for d in doms:
        d.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kt", [8, 16, 32])
def test_batch_z_march_trace(precision, kt):
    """The batch's z-marching trace (context 0 set to it before the batch
    is made) against single contexts run with the same trace and face/edge
    shapes: bit for bit in both precisions."""
    print(f"[gpu] batch with the z-march trace over {kt} planes, precision {precision}", flush=True)
    pbs = boxes(precision)
    shape = ((0, 8), (1, kt))
    ref = []
    for pb in pbs:
        d = abi.Domain(pb.params, pb.uin)
        d.set_shape(*shape)
        t, _ = d.advance(4, 0.0)
        ref.append((t, d.download()))
        d.close()
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    doms[0].set_shape(*shape)
    b = abi.Batch(doms)
    t = b.advance(4, [0.0] * len(doms))
    for d, tb, (ts, us) in zip(doms, t, ref):
        assert tb == ts
        assert np.array_equal(d.download(), us)
    b.close()
    for d in doms:
        d.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_update_planes_single_vs_batch(precision):
    """A single context's update marches fewer planes per block where nz is
    small (2 at nz 131, 4 at nz 261, 1 below 128; ragged last blocks), the
    batch always 8: the same bits in both precisions, with the next dt
    fused into the update (periodic) and without it (zero-gradient z)."""
    pbs = [P.random_smooth(12, 10, 131, seed=51, precision=precision),
           P.random_smooth(9, 8, 261, seed=52, precision=precision),
           P.random_smooth(11, 9, 131, seed=53, precision=precision,
                           bdtype=(P.PERIODIC, P.PERIODIC, P.ZERO_GRADIENT))]
    print(f"[gpu] update planes, single contexts against a batch, precision {precision}", flush=True)
    ref = singles(pbs, 3)
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    b = abi.Batch(doms)
    t = b.advance(3, [0.0] * len(doms))
    for d, tb, (ts, us) in zip(doms, t, ref):
        assert tb == ts
        assert np.array_equal(d.download(), us)
    b.close()
    for d in doms:
        d.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kz", [1, 2])
def test_batch_context_with_one_or_two_planes(precision, kz):
    """Context 0 with the face/edge kernel over 1 or 2 planes per block
    (the batch runs it over 4): single contexts with that shape and the
    batch agree bit for bit in both precisions, since which block computes
    a cell does not change its value."""
    print(f"[gpu] batch, context 0 with the face/edge kernel over {kz} planes, precision {precision}", flush=True)
    pbs = boxes(precision)
    shape = ((0, kz), (1, 8))
    ref = []
    for pb in pbs:
        d = abi.Domain(pb.params, pb.uin)
        d.set_shape(*shape)
        t, _ = d.advance(4, 0.0)
        ref.append((t, d.download()))
        d.close()
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    doms[0].set_shape(*shape)
    b = abi.Batch(doms)
    t = b.advance(4, [0.0] * len(doms))
    for d, tb, (ts, us) in zip(doms, t, ref):
        assert tb == ts
        assert np.array_equal(d.download(), us)
    b.close()
    for d in doms:
        d.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kt", [8, 16, 32])
def test_batch_fused_kernel(precision, kt):
    """The batch's fused trace + face/edge kernel (kb_fused, the body of
    k_fused over every box's tiles; context 0 set to it before the batch is
    made) against single contexts run with the same shape: bit for bit in
    both precisions."""
    print(f"[gpu] batch with the fused kernel over {kt} planes, precision {precision}", flush=True)
    pbs = boxes(precision)
    shape = ((3, kt), (1, 16))
    ref = []
    for pb in pbs:
        d = abi.Domain(pb.params, pb.uin)
        d.set_shape(*shape)
        t, _ = d.advance(4, 0.0)
        ref.append((t, d.download()))
        d.close()
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    doms[0].set_shape(*shape)
    b = abi.Batch(doms)
    t = b.advance(4, [0.0] * len(doms))
    for d, tb, (ts, us) in zip(doms, t, ref):
        assert tb == ts
        assert np.array_equal(d.download(), us)
    b.close()
    for d in doms:
        d.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nstreams", [1, 2])
def test_batch_fused_dt(precision, nstreams):
    """All boxes periodic: the batch fuses the next dt into its update
    (kb_update_dt, the faces alternating between each box's two buffers)
    over calls of 3 and 2 steps, and gives the bits and times of single
    contexts; the faces end in their home buffers (the contexts advance
    on alone afterwards and stay equal)."""
    print(f"[gpu] batch with the fused dt, precision {precision}, {nstreams} stream(s)", flush=True)
    pbs = [P.turbulence(16, seed=1, mach=1.0, beta=1.0, kmax=4, precision=precision),
           P.random_smooth(24, 20, 12, seed=2, slope_type=1, precision=precision),
           P.orszag_tang(8, precision=precision)]
    ref = singles(pbs, 7)
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    b = abi.Batch(doms, nstreams=nstreams)
    b.advance(3, [0.0] * len(doms))
    b.advance(2)
    b.close()
    for d, (ts, us) in zip(doms, ref):
        d.advance(2)
        assert d.time == ts
        assert np.array_equal(d.download(), us)
        d.close()


@pytest.mark.parametrize("nstreams", [1, 3, 8])
def test_batch_streams(nstreams):
    """The boxes dealt to 1, 3 or 8 (capped at the box count) streams: the
    same bits, and each box's time comes back in its own slot."""
    print(f"[gpu] batch on {nstreams} streams", flush=True)
    pbs = boxes(0)
    ref = singles(pbs, 3)
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    b = abi.Batch(doms, nstreams=nstreams)
    t = b.advance(3, [0.0] * len(doms))
    for d, tb, (ts, us) in zip(doms, t, ref):
        assert tb == ts
        assert np.array_equal(d.download(), us)
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_tuned_shape(precision):
    """After the batch autotuner (state unchanged): float64 bit for bit; in
    float32 the kept trace and face/edge kernels may differ from the single
    contexts' and contract their arithmetic differently, so the states agree to float32
    rounding: 5e-6 of the largest magnitude of the variable (rho, E) or of
    its vector (momentum, B) after 4 steps."""
    print(f"[gpu] tuned batch, precision {precision}", flush=True)
    pbs = boxes(precision)
    ref = singles(pbs, 4)
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    b = abi.Batch(doms)
    ms = b.tune()
    print(f"[gpu] tuned shape {b.shape}, {ms:.3f} ms")
    b.advance(4, [0.0] * len(doms))
    for d, (ts, us) in zip(doms, ref):
        u = d.download()
        if precision == 0:
            assert d.time == ts
            assert np.array_equal(u, us)
        else:
            assert abs(d.time - ts) <= 1e-6 * ts
            assert_f32_close(u, us)
    b.close()


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("pair", [(P.LLF, P.LLF), (P.HLL, P.HLLA), (P.UPWIND, P.UPWIND), (P.HLLD, P.LLF)])
def test_batch_solver_pairs(pair, fused):
    """Every other Riemann solver pair, with the default shapes and with the
    batched fused trace + face/edge kernel: the bits of single contexts in
    float64."""
    print(f"[gpu] batch solver pair {pair}, fused {fused}", flush=True)
    pbs = boxes(0, *pair)[:2]
    shape = ((3, 8), (1, 16)) if fused else None
    ref = []
    for pb in pbs:
        d = abi.Domain(pb.params, pb.uin)
        if shape:
            d.set_shape(*shape)
        t, _ = d.advance(2, 0.0)
        ref.append((t, d.download()))
        d.close()
    doms = [abi.Domain(pb.params, pb.uin) for pb in pbs]
    if shape:
        doms[0].set_shape(*shape)
    b = abi.Batch(doms)
    b.advance(2, [0.0, 0.0])
    for d, (ts, us) in zip(doms, ref):
        assert d.time == ts
        assert np.array_equal(d.download(), us)
    b.close()


def test_batch_rejections():
    print("[gpu] batch rejections", flush=True)
    a = abi.Domain(P.orszag_tang(8).params, P.orszag_tang(8).uin)
    f32 = P.orszag_tang(8, precision=1)
    c = abi.Domain(f32.params, f32.uin)
    llf = P.orszag_tang(8, riemann=P.LLF)
    e = abi.Domain(llf.params, llf.uin)
    mri = P.mri_test(8, 16, 8)
    s = abi.Domain(mri.params, mri.uin)
    empty = abi.Domain(P.orszag_tang(8).params)
    for group, what in (([a, c], b"precision"), ([a, e], b"Riemann"), ([a, s], b"shearing"),
                        ([a, empty], b"dmx_upload"), ([a, a], b"repeats")):
        with pytest.raises(abi.DmxError) as ei:
            abi.Batch(group)
        assert what.decode() in str(ei.value)
    with pytest.raises(abi.DmxError) as ei:
        abi.Batch([a], nstreams=9)
    assert "nstreams" in str(ei.value)
