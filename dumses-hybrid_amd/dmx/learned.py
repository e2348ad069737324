"""The learned slope limiter: network definition, features, training data.

A small MLP maps the 5-cell stencil of the 8 primitive variables along one
direction (40 inputs) to the 8 limited slopes of the centre cell:

    x (40) -> GELU(W1 x + b1) (64) -> GELU(W2 h1 + b2) (64) -> W3 h2 + b3 (8)

Features are scale-free.  For variable v with centre value q0 and stencil
differences d_m = q_m - q0 (m = -2, -1, +1, +2), s_v = max_m |d_m|:

    x[5v + 0..4] = d_-2/s, d_-1/s, q0/(|q0| + s), d_+1/s, d_+2/s

(all zero when s = 0), and the slope is y_v * s_v.  GELU is the tanh
approximation of Hendrycks and Gimpel (2016),
0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).

Targets come from a run at twice the resolution: the fine state restricted
to the coarse grid gives the stencil, and the difference between the
primitive states of the two halves of each coarse cell gives the slope that
the linear reconstruction should have (for a linear profile q0 + s x the
half averages differ by s/2).

This module is the float64 reference of the device kernels (feature
extraction, forward, backward) and the host side of training-data
preparation.  Parameters are one float32 vector in the order
W1 (40x64, row-major [in][out]), b1 (64), W2 (64x64), b2 (64), W3 (64x8), b3 (8).
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 40, 64, 8
SHAPES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NOUT)), ("b3", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)   # 7304
GELU_C = np.sqrt(2.0 / np.pi)
SMALLR = 1e-10


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases (float32)."""
    g = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            parts.append(g.uniform(-lim, lim, shp).ravel())
        else:
            parts.append(np.zeros(shp).ravel())
    return np.concatenate(parts).astype(np.float32)


def gelu(x):
    t = np.tanh(GELU_C * (x + 0.044715 * x ** 3))
    return 0.5 * x * (1.0 + t)


def gelu_grad(x):
    u = GELU_C * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    du = GELU_C * (1.0 + 3.0 * 0.044715 * x * x)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * du


def forward(theta, X, keep: bool = False):
    """Network output (n, 8) for features X (n, 40), in float64."""
    P = unpack(np.asarray(theta, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ P["W1"] + P["b1"]
    h1 = gelu(a1)
    a2 = h1 @ P["W2"] + P["b2"]
    h2 = gelu(a2)
    y = h2 @ P["W3"] + P["b3"]
    if keep:
        return y, (X, a1, h1, a2, h2)
    return y


def loss_and_grad(theta, X, T):
    """Mean squared error over all outputs, 0.5 * mean((y - T)^2) summed over
    the 8 outputs, and its gradient with respect to theta (float64)."""
    P = unpack(np.asarray(theta, dtype=np.float64))
    y, (X, a1, h1, a2, h2) = forward(theta, X, keep=True)
    n = X.shape[0]
    r = y - np.asarray(T, dtype=np.float64)
    loss = 0.5 * (r * r).sum() / n
    gy = r / n
    gW3 = h2.T @ gy
    gb3 = gy.sum(axis=0)
    gh2 = gy @ P["W3"].T
    ga2 = gh2 * gelu_grad(a2)
    gW2 = h1.T @ ga2
    gb2 = ga2.sum(axis=0)
    gh1 = ga2 @ P["W2"].T
    ga1 = gh1 * gelu_grad(a1)
    gW1 = X.T @ ga1
    gb1 = ga1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad


# ---------------------------------------------------------------------------
# features, fallback and slopes
# ---------------------------------------------------------------------------
def features(q5: np.ndarray):
    """q5 (n, 5, 8): stencil q_-2..q_+2 of the 8 primitive variables.
    Returns X (n, 40) and the scales s (n, 8)."""
    q0 = q5[:, 2, :]
    d = q5 - q0[:, None, :]
    s = np.abs(d).max(axis=1)
    safe = np.where(s > 0, s, 1.0)
    X = np.empty((q5.shape[0], 5, 8))
    X[:, 0] = d[:, 0] / safe
    X[:, 1] = d[:, 1] / safe
    X[:, 2] = q0 / (np.abs(q0) + safe)
    X[:, 3] = d[:, 3] / safe
    X[:, 4] = d[:, 4] / safe
    X[np.broadcast_to((s == 0)[:, None, :], X.shape)] = 0.0
    return X.transpose(0, 2, 1).reshape(-1, NIN), s


def van_leer(dl, dr):
    p = dl * dr
    return np.where(p > 0, 2.0 * p / np.where(p > 0, dl + dr, 1.0), 0.0)


def mc_slope(dl, dr):
    """The MC slope (scheme.f90:452-477 with slope_type 2): sign(dl + dr)
    min(2|dl|, 2|dr|, |dl + dr| / 2) where dl dr > 0, else 0."""
    c = 0.5 * (dl + dr)
    m = np.minimum(np.minimum(2.0 * np.abs(dl), 2.0 * np.abs(dr)), np.abs(c))
    return np.where(dl * dr > 0, np.copysign(m, c), 0.0)


def monotone(q5):
    """Cells (n,) whose 8 variables have no extremum along the stencil
    (dl dr >= 0 for every variable): the limiter rule gives them MC."""
    return ((q5[:, 2] - q5[:, 1]) * (q5[:, 3] - q5[:, 2]) >= 0).all(axis=1)


def limited_slopes(theta, q5):
    """The device rule: a cell monotone in all 8 variables along the
    stencil takes the MC slopes.  Elsewhere the van Leer slope plus the
    network's correction (its output times the stencil scale), kept in the
    TVD region: a variable at an extremum or whose slope has the wrong sign
    takes the van Leer slope, one larger than twice the smaller one-sided
    difference is clamped to that bound.  A cell whose density or pressure
    would turn non-positive at a face takes van Leer for all 8 variables.
    Returns (slopes (n, 8), fell_back (n,), replaced (n, 8): the variables
    of non-monotone cells whose network slope left the TVD region)."""
    X, s = features(q5)
    dl = q5[:, 2] - q5[:, 1]
    dr = q5[:, 3] - q5[:, 2]
    vl = van_leer(dl, dr)
    y = vl + forward(theta, X) * s
    sgn = (dl * dr > 0) & (y * dl > 0)
    bound = 2.0 * np.minimum(np.abs(dl), np.abs(dr))
    mono = monotone(q5)[:, None]
    tvd = mono | (sgn & (np.abs(y) <= bound))
    out = np.where(mono, mc_slope(dl, dr), np.where(sgn, np.copysign(np.minimum(np.abs(y), bound), y), vl))
    q0 = q5[:, 2]
    pos = np.ones(q5.shape[0], dtype=bool)
    for v in (0, 4):
        pos &= (q0[:, v] - 0.5 * np.abs(out[:, v]) > 0)
    out = np.where(pos[:, None], out, vl)
    return out, ~pos, ~tvd


# ---------------------------------------------------------------------------
# training data from fine/coarse pairs
# ---------------------------------------------------------------------------
def conserved_cells(u: np.ndarray) -> np.ndarray:
    """Interior cell-centred conserved variables (8, nz, ny, nx) of a ghosted
    state: rho, momentum, E and the face-averaged cell B."""
    s = (slice(3, -3),) * 3
    c = np.empty((8,) + u[(0,) + s].shape)
    c[:5] = u[(slice(0, 5),) + s]
    c[5] = 0.5 * (u[(5,) + s] + u[(8,) + s])
    c[6] = 0.5 * (u[(6,) + s] + u[(9,) + s])
    c[7] = 0.5 * (u[(7,) + s] + u[(10,) + s])
    return c


def primitives(c: np.ndarray, gamma: float, iso: int = 0, ciso: float = 0.0) -> np.ndarray:
    """Primitive variables (rho, u, v, w, p, Bx, By, Bz) of cell-centred
    conserved variables, with the floors of the trace."""
    r = np.maximum(c[0], SMALLR)
    u, v, w = c[1] / c[0], c[2] / c[0], c[3] / c[0]
    if iso:
        p = r * ciso * ciso
    else:
        eken = 0.5 * (u * u + v * v + w * w)
        emag = 0.5 * (c[5] ** 2 + c[6] ** 2 + c[7] ** 2)
        eint = (c[4] - emag) / c[0] - eken
        p = np.maximum((gamma - 1.0) * r * eint, SMALLR * SMALLR * SMALLR / gamma)
    return np.stack([r, u, v, w, p, c[5], c[6], c[7]])


def restrict_state(u_fine: np.ndarray) -> np.ndarray:
    """Coarse ghosted uin (periodic ghosts) from a fine ghosted uin: cell
    averages of the 2x2x2 conserved children, face averages of the 2x2 child
    faces on each coarse face."""
    from .problems import fill_periodic
    f = u_fine[:, 3:-3, 3:-3, 3:-3]
    nz, ny, nx = (n // 2 for n in f.shape[1:])
    uc = np.zeros((11, nz + 6, ny + 6, nx + 6))
    cs = (slice(3, -3),) * 3
    for v in range(5):
        uc[(v,) + cs] = f[v].reshape(nz, 2, ny, 2, nx, 2).mean(axis=(1, 3, 5))
    uc[(5,) + cs] = f[5][:, :, 0::2].reshape(nz, 2, ny, 2, nx).mean(axis=(1, 3))
    uc[(6,) + cs] = f[6][:, 0::2, :].reshape(nz, 2, ny, nx, 2).mean(axis=(1, 4))
    uc[(7,) + cs] = f[7][0::2].reshape(nz, ny, 2, nx, 2).mean(axis=(2, 4))
    return fill_periodic(uc)


def target_slopes(u_fine: np.ndarray, gamma: float) -> np.ndarray:
    """True slopes (3, 8, nz, ny, nx) of the primitive variables on the
    coarse grid: 2 x (primitive of the right half - primitive of the left
    half) of each coarse cell, per direction."""
    c = conserved_cells(u_fine)
    nz, ny, nx = (n // 2 for n in c.shape[1:])
    cc = c.reshape(8, nz, 2, ny, 2, nx, 2)
    out = np.empty((3, 8, nz, ny, nx))
    for d, ax in enumerate((6, 4, 2)):       # x, y, z halves
        lo = np.take(cc, 0, axis=ax)
        hi = np.take(cc, 1, axis=ax)
        axes = tuple(a for a in (2, 4, 6) if a != ax)
        axes = tuple(a - (1 if a > ax else 0) for a in axes)
        ql = primitives(lo.mean(axis=axes), gamma)
        qr = primitives(hi.mean(axis=axes), gamma)
        out[d] = 2.0 * (qr - ql)
    return out


def stencils(u_coarse: np.ndarray, gamma: float, d: int) -> np.ndarray:
    """5-cell primitive stencils (n, 5, 8) along direction d (0 = x) of every
    interior cell of a ghosted coarse state (needs 2 ghost layers)."""
    c = np.empty((8,) + u_coarse.shape[1:])
    c[:5] = u_coarse[:5]
    # cell B from faces; the last ghost layer has no right face and is unused
    c[5] = 0.5 * (u_coarse[5] + u_coarse[8])
    c[6] = 0.5 * (u_coarse[6] + u_coarse[9])
    c[7] = 0.5 * (u_coarse[7] + u_coarse[10])
    q = primitives(c, gamma)
    ax = 3 - d
    sl = []
    for m in range(-2, 3):
        idx = [slice(None), slice(3, -3), slice(3, -3), slice(3, -3)]
        n = q.shape[ax]
        idx[ax] = slice(3 + m, n - 3 + m)
        sl.append(q[tuple(idx)])
    st = np.stack(sl, axis=1)                  # (8, 5, nz, ny, nx)
    return st.reshape(8, 5, -1).transpose(2, 1, 0)


def tvd_project(t, dl, dr, margin: float = 1.0):
    """The slope t moved into the region the limiter rule accepts: at an
    extremum (dl dr <= 0) the van Leer slope (zero); elsewhere the sign of
    dl and at most `margin` times twice the smaller of |dl|, |dr| (a slope
    of the wrong sign becomes 1e-3 of that bound)."""
    bound = margin * 2.0 * np.minimum(np.abs(dl), np.abs(dr))
    sg = np.sign(dl)
    mag = np.where(t * dl > 0, np.minimum(np.abs(t), bound), 1e-3 * bound)
    return np.where(dl * dr > 0, sg * mag, 0.0)


def training_pairs(u_fine: np.ndarray, gamma: float, tvd: bool = False, margin: float = 1.0,
                   extrema_only: bool = False):
    """Features X (n, 40) and normalised targets T (n, 8) of every coarse
    cell and direction of one fine state: the difference between the true
    slope and the van Leer slope over the stencil scale, clipped to [-4, 4];
    with tvd, the true slope first projected into the region the limiter
    rule accepts (tvd_project, with margin); with extrema_only, only the
    cells the network acts on (not monotone in all 8 variables)."""
    uc = restrict_state(u_fine)
    tgt = target_slopes(u_fine, gamma)
    Xs, Ts = [], []
    for d in range(3):
        q5 = stencils(uc, gamma, d)
        X, s = features(q5)
        t = tgt[d].reshape(8, -1).T
        dl, dr = q5[:, 2] - q5[:, 1], q5[:, 3] - q5[:, 2]
        if tvd:
            t = tvd_project(t, dl, dr, margin)
        t = t - van_leer(dl, dr)
        T = np.clip(np.where(s > 0, t / np.where(s > 0, s, 1.0), 0.0), -4.0, 4.0)
        if extrema_only:
            keep = ~monotone(q5)
            X, T = X[keep], T[keep]
        Ts.append(T)
        Xs.append(X)
    return np.concatenate(Xs).astype(np.float32), np.concatenate(Ts).astype(np.float32)
