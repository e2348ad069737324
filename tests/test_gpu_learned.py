"""The learned limiter on the GPU against its float64 reference
(dmx/learned.py, itself pinned by tests/test_learned_cpu.py): the bf16 MFMA
lane layout, the training kernels' forward pass and gradient, that Adam
reduces the loss, and the limiter kernel's slopes."""
import numpy as np
import pytest

from dmx import abi
from dmx import learned as L
from dmx import problems as P

pytestmark = pytest.mark.gpu


def test_mfma_lane_layout():
    print("[gpu] mfma selftest", flush=True)
    got = abi.mfma_selftest()
    i = np.arange(16)[:, None]
    k = np.arange(32)[None, :]
    A = ((i + 2 * k) % 7) / 8.0
    kk = np.arange(32)[:, None]
    j = np.arange(16)[None, :]
    B = ((3 * kk + j) % 5) / 4.0 - 0.5
    assert np.array_equal(got, (A @ B).astype(np.float32))


def _random_problem(n, seed):
    g = np.random.Generator(np.random.PCG64(seed))
    theta = (L.init_params(seed).astype(np.float64) + 0.05 * g.standard_normal(L.NPARAM)).astype(np.float32)
    X = np.clip(g.standard_normal((n, L.NIN)), -1, 1).astype(np.float32)
    T = g.standard_normal((n, L.NOUT)).astype(np.float32)
    return theta, X, T


def test_forward_and_gradient_match_reference():
    theta, X, T = _random_problem(1000, 1)
    print("[gpu] mlp forward/gradient", flush=True)
    m = abi.Mlp(1024)
    m.set_params(theta)
    y = m.forward(X)
    y_ref = L.forward(theta, X)
    assert np.abs(y - y_ref).max() <= 1e-5 * max(1.0, np.abs(y_ref).max())
    loss, g = m.loss_grad(X, T)
    loss_ref, g_ref = L.loss_and_grad(theta, X, T)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max(), float(np.abs(g - g_ref).max())


def test_adam_training_reduces_loss():
    """Regression onto a fixed random teacher network."""
    g = np.random.Generator(np.random.PCG64(2))
    teacher = L.init_params(7)
    X = np.clip(g.standard_normal((20000, L.NIN)), -1, 1).astype(np.float32)
    T = L.forward(teacher, X).astype(np.float32)
    print("[gpu] mlp training", flush=True)
    m = abi.Mlp(2048)
    m.set_params(L.init_params(3))
    m.load_data(X, T)
    losses = m.train(400, 1024, 3e-3, seed=5)
    print(f"[gpu] loss {losses[:5].mean():.4e} -> {losses[-5:].mean():.4e}")
    assert losses[-20:].mean() < 0.1 * losses[:5].mean()
    th = m.get_params()
    assert np.isfinite(th).all()


def test_limiter_kernel_matches_reference():
    """Device slopes (features in float, network on bf16 MFMA) against the
    float64 rule on the same float primitives: where both keep the network
    slope they agree to the bf16 rounding of weights and activations; the
    fallback decisions agree on almost every cell."""
    pb = P.turbulence(16, seed=11, mach=1.0, beta=1.0)
    p = pb.params.replace(slope_type=4, precision=1)
    g = np.random.Generator(np.random.PCG64(4))
    theta = (0.5 * L.init_params(9)).astype(np.float32)
    theta[L.NPARAM - 8:] += 0.1
    print("[gpu] limiter kernel vs reference", flush=True)
    d = abi.Domain(p, pb.uin)
    d.set_limiter(theta)
    got, fb, frep = d.learned_slopes()
    u = d.download()
    agree = 0
    total = 0
    rep_ref = 0
    for dd in range(3):
        q5 = L.stencils(u, p.gamma, dd).astype(np.float32).astype(np.float64)
        ref, fell, replaced = L.limited_slopes(theta, q5)
        rep_ref += replaced.sum()
        ref = ref.reshape(p.nz, p.ny, p.nx, 8).transpose(3, 0, 1, 2)
        gd = got[dd].astype(np.float64)
        X, s = L.features(q5)
        scale = s.reshape(p.nz, p.ny, p.nx, 8).transpose(3, 0, 1, 2) + 1e-30
        close = np.abs(gd - ref) <= 0.05 * scale
        agree += close.sum()
        total += close.size
    # the reference counts interior cells, the device [0, n+1]^3
    frep_ref = rep_ref / total
    print(f"[gpu] agreement {agree / total:.4f}, positivity fallback {fb:.4f}, outside the TVD region "
          f"{frep:.4f} (reference {frep_ref:.4f})")
    assert agree / total > 0.98
    assert 0.0 <= fb <= 1.0
    assert abs(frep - frep_ref) < 0.05


def test_zero_network_is_mc_or_van_leer():
    """A network that outputs zero leaves the MC slope of the float
    primitives in cells monotone in all 8 variables, the van Leer slope
    elsewhere."""
    pb = P.turbulence(16, seed=13, mach=1.0, beta=1.0)
    p = pb.params.replace(slope_type=4, precision=1)
    print("[gpu] zero network", flush=True)
    d = abi.Domain(p, pb.uin)
    d.set_limiter(np.zeros(L.NPARAM, np.float32))
    got, fb, frep = d.learned_slopes()
    u = d.download()
    d.close()
    for dd in range(3):
        q5 = L.stencils(u, p.gamma, dd).astype(np.float32).astype(np.float64)
        dl, dr = q5[:, 2] - q5[:, 1], q5[:, 3] - q5[:, 2]
        want = np.where(L.monotone(q5)[:, None], L.mc_slope(dl, dr), L.van_leer(dl, dr))
        want = want.reshape(p.nz, p.ny, p.nx, 8).transpose(3, 0, 1, 2)
        _, s = L.features(q5)
        scale = s.reshape(p.nz, p.ny, p.nx, 8).transpose(3, 0, 1, 2)
        assert np.all(np.abs(got[dd] - want) <= 1e-5 * scale + 1e-30)
    assert fb == 0.0


def test_learned_step_conserves_and_keeps_div_b():
    pb = P.turbulence(16, seed=12, mach=0.5, beta=10.0)
    p = pb.params.replace(slope_type=4, precision=1)
    print("[gpu] learned step", flush=True)
    d = abi.Domain(p, pb.uin)
    d.set_limiter(0.5 * L.init_params(9))
    m0 = P.interior(d.download())[0].sum()
    d.advance(5, 0.0)
    u = d.download()
    assert np.isfinite(u).all()
    assert abs(P.interior(u)[0].sum() - m0) <= 1e-5 * m0
    h = d.history()
    assert abs(h[8]) < 1e-9


def test_learned_limiter_needs_weights():
    pb = P.turbulence(8, seed=1)
    d = abi.Domain(pb.params.replace(slope_type=4), pb.uin)
    with pytest.raises(abi.DmxError, match="dmx_set_limiter"):
        d.advance(1, 0.0)


def test_trained_limiter_gains_over_mc():
    """The committed network (dmx/data/limiter_weights.npz, written by
    python -m dmx.train_limiter train) against the plain MC limiter: from a
    decaying-turbulence state (Mach 1, beta 1) not in the training ensemble,
    32^3 runs to t = 0.25 compared with the 64^3 plain run restricted to
    32^3.  The learned coarse run is closer to it than the MC coarse run."""
    from dmx import train_limiter as TL
    theta = np.load(TL.WEIGHTS)["theta"]
    print("[gpu] learned limiter vs MC, 32^3 against 64^3", flush=True)
    res = TL.evaluate(theta, n=32, tend=0.25, seed=777)
    print(f"[gpu] mean L1: mc {res['mc_l1_mean']:.5f}, learned {res['learned_l1_mean']:.5f}; "
          f"wall: fine {res['fine_wall_s']:.3f} s, mc {res['mc_wall_s']:.3f} s, "
          f"learned {res['learned_wall_s']:.3f} s")
    assert res["learned_l1_mean"] < res["mc_l1_mean"]
