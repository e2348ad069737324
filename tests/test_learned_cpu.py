"""The learned limiter's float64 reference (dmx/learned.py), pinned by
mathematics rather than by restating it: finite differences for the
gradient, exact linear profiles for the targets, conservation and div B for
the restriction, invariances for the features."""
import numpy as np
import pytest

from dmx import learned as L
from dmx import problems as P


def test_gelu_derivative_matches_finite_differences():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (L.gelu(x + h) - L.gelu(x - h)) / (2 * h)
    assert np.allclose(L.gelu_grad(x), fd, rtol=1e-7, atol=1e-9)
    # tanh form of GELU: odd part and limits
    assert abs(L.gelu(0.0)) == 0.0 and abs(L.gelu(8.0) - 8.0) < 1e-9 and abs(L.gelu(-8.0)) < 1e-9


def test_gradient_matches_central_differences():
    g = np.random.Generator(np.random.PCG64(3))
    theta = L.init_params(1).astype(np.float64) + 0.01 * g.standard_normal(L.NPARAM)
    X = g.standard_normal((17, L.NIN))
    T = g.standard_normal((17, L.NOUT))
    loss, grad = L.loss_and_grad(theta, X, T)
    idx = g.choice(L.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(L.NPARAM)
        e[i] = h
        fd = (L.loss_and_grad(theta + e, X, T)[0] - L.loss_and_grad(theta - e, X, T)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_features_are_scale_free():
    g = np.random.Generator(np.random.PCG64(5))
    q5 = g.standard_normal((50, 5, 8)) + 3.0
    X1, s1 = L.features(q5)
    X2, s2 = L.features(7.5 * q5)
    assert np.allclose(X1, X2, rtol=1e-13, atol=1e-15)
    assert np.allclose(s2, 7.5 * s1)
    flat = np.ones((4, 5, 8))
    X3, s3 = L.features(flat)
    assert (X3 == 0).all() and (s3 == 0).all()


def test_mc_slope_is_the_reference_limiter():
    """mc_slope against the reference's slope formula with slope_type 2
    (scheme.f90:452-477: dlft = 2 dl, drgt = 2 dr, dcen = (dlft + drgt) / 4,
    dq = sign(dcen) min(min(|dlft|, |drgt|) [dlft drgt > 0], |dcen|)),
    brute force on random differences with ties and zeros."""
    g = np.random.Generator(np.random.PCG64(7))
    dl = np.round(g.normal(size=20000), 1)
    dr = np.round(g.normal(size=20000), 1)
    dlft, drgt = 2.0 * dl, 2.0 * dr
    dcen = 0.5 * (dlft + drgt) / 2.0
    dlim = np.where(dlft * drgt <= 0, 0.0, np.minimum(np.abs(dlft), np.abs(drgt)))
    ref = np.copysign(1.0, dcen) * np.minimum(dlim, np.abs(dcen))
    assert np.array_equal(L.mc_slope(dl, dr), ref)


def test_zero_network_is_mc_or_van_leer():
    """A network that outputs zero: MC in cells monotone in all 8 variables,
    van Leer elsewhere."""
    g = np.random.Generator(np.random.PCG64(6))
    q5 = np.abs(g.standard_normal((4000, 5, 8))) + 0.5
    q5[:100] = np.sort(q5[:100], axis=1)          # some monotone cells
    out, fb, rep = L.limited_slopes(np.zeros(L.NPARAM), q5)
    dl, dr = q5[:, 2] - q5[:, 1], q5[:, 3] - q5[:, 2]
    mono = L.monotone(q5)
    assert mono[:100].all() and (~mono).sum() > 3000
    assert np.array_equal(out, np.where(mono[:, None], L.mc_slope(dl, dr), L.van_leer(dl, dr)))


def test_rule_clamps_to_the_tvd_bound_and_falls_back():
    """A network whose output is a constant c (only the last bias set), in
    cells with an extremum in some variable: where van Leer + c s has the
    sign of the one-sided differences it is clamped to twice the smaller
    one; where it has the wrong sign, and at extrema, the slope is van
    Leer.  Monotone cells take MC whatever the network says."""
    g = np.random.Generator(np.random.PCG64(8))
    q5 = np.abs(g.standard_normal((4000, 5, 8))) + 0.5
    q5[:100] = np.sort(q5[:100], axis=1)          # some monotone cells
    dl, dr = q5[:, 2] - q5[:, 1], q5[:, 3] - q5[:, 2]
    vl = L.van_leer(dl, dr)
    bound = 2.0 * np.minimum(np.abs(dl), np.abs(dr))
    for c in (3.0, -3.0):
        theta = np.zeros(L.NPARAM)
        theta[L.NPARAM - 8:] = c
        out, fb, rep = L.limited_slopes(theta, q5)
        _, s = L.features(q5)
        y = vl + c * s
        cellmono = np.broadcast_to(L.monotone(q5)[:, None], y.shape)
        mono = dl * dr > 0
        right = mono & (y * dl > 0) & ~cellmono
        wrong = ~right & ~cellmono
        assert right.any() and (mono & wrong).any() and cellmono.any()
        assert np.allclose(out[right], np.copysign(np.minimum(np.abs(y), bound), y)[right], rtol=1e-15, atol=0)
        assert np.array_equal(out[wrong], vl[wrong])
        assert np.array_equal(out[cellmono], L.mc_slope(dl, dr)[cellmono])
        assert np.all(np.abs(out[mono]) <= bound[mono] * (1 + 1e-15))
        assert np.array_equal(rep, ~cellmono & ~(right & (np.abs(y) <= bound)))


def test_restriction_conserves_and_keeps_div_b_zero():
    pb = P.turbulence(16, seed=2, mach=1.0, beta=1.0)
    uc = L.restrict_state(pb.uin)
    fi, ci = P.interior(pb.uin), P.interior(uc)
    for v in range(5):
        assert abs(ci[v].sum() * 8 - fi[v].sum()) <= 1e-12 * np.abs(fi[v]).sum()
    bx, by, bz = uc[5], uc[6], uc[7]
    div = (bx[3:-3, 3:-3, 4:-2] - bx[3:-3, 3:-3, 3:-3]) + (by[3:-3, 4:-2, 3:-3] - by[3:-3, 3:-3, 3:-3]) \
        + (bz[4:-2, 3:-3, 3:-3] - bz[3:-3, 3:-3, 3:-3])
    assert np.abs(div).max() < 1e-13


def test_targets_of_a_linear_profile_are_exact():
    """rho = 1 + 0.1 x (linear, fine 16^3): the coarse x-slope of rho is 0.1
    times the coarse width; the other directions have zero slope."""
    pb = P.turbulence(16, seed=2, mach=0.0, beta=1e30)
    p = pb.params
    u = pb.uin.copy()
    x = p.xmin + p.dx * (np.arange(p.nx + 6) - 2.5)
    u[0] = 1.0 + 0.1 * x[None, None, :]
    u[1:4] = 0.0
    u[4] = 1.0 / (p.gamma - 1.0)
    t = L.target_slopes(u, p.gamma)
    assert np.allclose(t[0, 0], 0.1 * 2 * p.dx, rtol=1e-12)
    assert np.abs(t[1, 0]).max() < 1e-14 and np.abs(t[2, 0]).max() < 1e-14


def test_training_pairs_shapes():
    pb = P.turbulence(16, seed=3)
    X, T = L.training_pairs(pb.uin, pb.params.gamma)
    assert X.shape == (3 * 8 ** 3, 40) and T.shape == (3 * 8 ** 3, 8)
    assert np.isfinite(X).all() and np.isfinite(T).all()
    assert np.abs(X).max() <= 1.0 + 1e-6


def test_targets_are_corrections_to_van_leer():
    """van Leer + target x scale gives back the (projected) sub-cell slope
    of target_slopes wherever the target is not clipped: a network that
    fits the targets reproduces the fine slopes through the limiter rule."""
    pb = P.turbulence(16, seed=8, mach=1.0, beta=1.0)
    gam = pb.params.gamma
    uc = L.restrict_state(pb.uin)
    tgt = L.target_slopes(pb.uin, gam)
    for tvd in (False, True):
        X, T = L.training_pairs(pb.uin, gam, tvd=tvd)
        n = T.shape[0] // 3
        for d in range(3):
            q5 = L.stencils(uc, gam, d)
            _, s = L.features(q5)
            dl, dr = q5[:, 2] - q5[:, 1], q5[:, 3] - q5[:, 2]
            t = tgt[d].reshape(8, -1).T
            if tvd:
                t = L.tvd_project(t, dl, dr)
            rec = L.van_leer(dl, dr) + T[d * n:(d + 1) * n].astype(np.float64) * s
            keep = np.abs(T[d * n:(d + 1) * n]) < 4.0
            assert keep.mean() > 0.95
            assert np.allclose(rec[keep], t[keep], rtol=0, atol=1e-6 * np.broadcast_to(s, t.shape)[keep].max())


def test_tvd_projection_is_accepted_by_the_rule():
    """Targets projected by tvd_project pass the limiter's TVD test (away
    from extrema) and are zero at extrema, where van Leer is zero."""
    g = np.random.Generator(np.random.PCG64(5))
    dl, dr, t = g.normal(size=(3, 10000))
    p = L.tvd_project(t, dl, dr)
    mono = dl * dr > 0
    assert np.all((p[mono] * dl[mono] > 0) & (np.abs(p[mono]) <= 2.0 * np.minimum(np.abs(dl[mono]), np.abs(dr[mono]))))
    assert np.all(p[~mono] == 0.0)
    ok = mono & (t * dl > 0) & (np.abs(t) <= 2.0 * np.minimum(np.abs(dl), np.abs(dr)))
    assert np.array_equal(p[ok], t[ok])
