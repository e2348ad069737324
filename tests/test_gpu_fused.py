"""The fused trace + face/edge kernel (launch shape mode 3, k_fused): the
trace summary stays in LDS.  In float64 it gives the bits of the compiled
reference on every step fixture and of the separate trace and face/edge
kernels on ragged domains for every solver pair, limiter and boundary type;
in float32 it stays within round-off of the separate kernels.  And the
float32-bulk path's pressure recovery at low beta (E and face B in float64)."""
import numpy as np
import pytest

from dmx import abi
from dmx import problems as P
from goldens import STEP_FIXTURES, assert_f32_close, load_step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", STEP_FIXTURES)
def test_fp64_fused_matches_reference_fixture_bitwise(name):
    p, u, dts, outs, full = load_step(name)
    print(f"[gpu] fused kernel, fixture {name} {p.nx}x{p.ny}x{p.nz}", flush=True)
    d = abi.Domain(p, u)
    d.set_shape((3, 4), (0, 4))
    t = 0.0
    for s in range(1, int(max(outs)) + 1):
        t, dd = d.advance(1, t)
        assert dd[0] == dts[s - 1], (name, s, dd[0], dts[s - 1])
        if s in outs:
            got = P.interior(d.download())
            bad = got != outs[s]
            assert not bad.any(), (name, s, int(bad.sum()), float(np.abs(got - outs[s]).max()))
    d.close()


CASES = [
    ("multidim, HLLD/HLLD", lambda: P.random_smooth(21, 13, 10, seed=31, slope_type=3)),
    ("zero-gradient x,z, MC", lambda: P.random_smooth(17, 9, 23, seed=32, slope_type=2,
                                                     bdtype=(P.ZERO_GRADIENT, P.PERIODIC, P.ZERO_GRADIENT))),
    ("shearing box", lambda: P.mri_test(16, 32, 16)),
    ("two x tiles, HLL/LLF, minmod", lambda: P.random_smooth(45, 11, 7, seed=33, slope_type=1,
                                                            riemann=P.HLL, riemann2d=P.LLF)),
    ("LLF/upwind", lambda: P.random_smooth(19, 18, 9, seed=34, riemann=P.LLF, riemann2d=P.UPWIND)),
    ("upwind/HLLA, flat", lambda: P.random_smooth(12, 20, 11, seed=35, slope_type=0,
                                                  riemann=P.UPWIND, riemann2d=P.HLLA)),
    ("isothermal", lambda: P.random_smooth(14, 10, 12, seed=36, iso=1, ciso=0.8)),
    # 65 faces a row: a 32- and a 33-column tile in float32 (7 rows), three
    # of 16 and one of 17 in float64; 24 rows cut into 8s or 6s
    ("wide and narrow tiles", lambda: P.random_smooth(64, 23, 9, seed=37)),
]


@pytest.mark.parametrize("kt", [4, 8])
@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_fused_equals_separate_kernels(case, kt):
    """Fused kernel (kt planes per block; 4 wraps its LDS rings mid-block on
    the ragged domains) against the separate trace and fused face/edge
    kernels: bit for bit in float64; in float32 within 1e-5 of the largest
    magnitude of each variable (rho, E) or of its vector (momentum, B) after
    3 steps (the shearing box's By is 1e-6 of |B|)."""
    pb = dict(CASES)[case]()
    for prec in (0, 1):
        p = pb.params.replace(precision=prec)
        print(f"[gpu] fused kt={kt} precision {prec} {case} {p.nx}x{p.ny}x{p.nz}", flush=True)
        out = []
        for flux in ((0, 8), (3, kt)):
            d = abi.Domain(p, pb.uin)
            d.set_shape(flux, (1, 8))
            d.advance(3, 0.0)
            out.append(d.download())
            d.close()
        if prec == 0:
            assert np.array_equal(out[0], out[1])
        else:
            assert_f32_close(out[1], out[0], rtol=1e-5)


def test_fused_learned_limiter():
    """The fused kernel with the learned limiter (slope_type 4: the trace
    part reads the network's slopes) against the separate kernels: bit for
    bit in float64 over 2 steps; in float32 within round-off after 1 step
    (after that a round-off difference in a nearly uniform variable can move
    a cell between the MC and the network branch)."""
    from dmx import learned as L
    pb = P.orszag_tang(24)
    theta = 0.5 * L.init_params(9)
    for prec in (0, 1):
        p = pb.params.replace(precision=prec, slope_type=4)
        print(f"[gpu] fused kernel, learned limiter, precision {prec}", flush=True)
        out = []
        for flux in ((0, 8), (3, 8)):
            d = abi.Domain(p, pb.uin)
            d.set_limiter(theta)
            d.set_shape(flux, (1, 8))
            d.advance(2 if prec == 0 else 1, 0.0)
            out.append(d.download())
            d.close()
        if prec == 0:
            assert np.array_equal(out[0], out[1])
        else:
            assert_f32_close(out[1], out[0], rtol=1e-5)


@pytest.mark.parametrize("flux", [(0, 8), (3, 8)])
@pytest.mark.parametrize("beta,bound", [(1e-2, 1e-5), (1e-3, 3e-5)])
def test_fp32_low_beta_pressure(flux, beta, bound):
    """Low beta: a sound wave of amplitude 1e-3 in a uniform oblique field
    at 64^3, 20 steps; the internal energy is beta of the magnetic energy.
    The float32 bulk keeps E and the face fields in float64, recovers the
    pressure from them in float64 and evaluates the Riemann solvers' energy
    chain in float64.  Measured pressure L1 against the float64 run
    (relative to the L1 norm of the pressure): 2.0e-6 at beta 1e-2 and
    2.5e-5 at beta 1e-3 (with E in float32: 6.5e-6 and 4.7e-5).  What
    remains is the float32 rounding of the reconstructed field, about
    6e-8 |B| per face and corner state, which the 1D energy flux and the 2D
    EMFs turn into upwind dissipation through different discretisations,
    so that E and B^2/2 drift apart by a few eps_32 |B|^2 per step: an error
    that scales as 1/beta (2.5e-5 at 1e-3 is still 1.3e-5 with the whole
    Riemann solvers in float64)."""
    pb = P.sound_wave(64, beta=beta, amp=1e-3)
    print(f"[gpu] low-beta sound wave, beta {beta}, flux shape {flux}", flush=True)
    states = []
    for prec in (0, 1):
        d = abi.Domain(pb.params.replace(precision=prec), pb.uin)
        d.set_shape(flux, (1, 8))
        d.advance(20, 0.0)
        states.append(d.download())
        d.close()
    p64 = P.pressure(pb.params, states[0])
    p32 = P.pressure(pb.params, states[1])
    l1 = np.abs(p32 - p64).sum() / np.abs(p64).sum()
    wave = np.abs(p64 - p64.mean()).sum() / np.abs(p64).sum()
    print(f"[gpu]   pressure L1 float32 vs float64: {l1:.3e} (wave amplitude in L1: {wave:.3e})", flush=True)
    assert wave > 1e-4
    assert l1 <= bound
