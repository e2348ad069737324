"""Initial states for the MHD step, built on the reference grid.

Every generator returns a ``Problem``: the run parameters (the reference's
namelist values) and the full ghosted state ``uin`` in the reference layout
(11 variables, see dumpfile.py), with ghosts filled by the periodic copy of
src/subroutines/boundary.f90:96-325 and right faces by the copy of
boundary.f90:82-89.  Face-centred B is built as the discrete curl of an
edge-centred vector potential, so its discrete divergence vanishes to
rounding.

Grid: cell centres x(i) = xmin + dx/2 + (i-1)*dx for i = -2..nx+3
(src/subroutines/init.f90:464-468); the left face of cell i sits at
xmin + (i-1)*dx.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field

import numpy as np

NG = 3
PERIODIC, ZERO_GRADIENT, SHEARINGBOX = 0, 1, 2
LLF, HLL, HLLD, UPWIND, HLLA = 2, 3, 4, 5, 8
SOLVER_NAMES = {"llf": LLF, "hll": HLL, "hlld": HLLD, "upwind": UPWIND, "hlla": HLLA}
BC_NAMES = {"periodic": PERIODIC, "zero_gradient": ZERO_GRADIENT, "shearingbox": SHEARINGBOX}


@dataclass
class Params:
    """Values of the reference namelists scheme_params / model_params
    (src/subroutines/init.f90:153-173; defaults :264-299)."""
    nx: int = 256
    ny: int = 256
    nz: int = 256
    xmin: float = 0.0
    xmax: float = 1.0
    ymin: float = -0.5
    ymax: float = 0.5
    zmin: float = 0.0
    zmax: float = 1.0
    gamma: float = 5.0 / 3.0
    ciso: float = 0.0
    omega0: float = 0.0
    courant: float = 0.7
    slope_type: int = 2
    riemann: int = HLLD
    riemann2d: int = HLLD
    bdtype: tuple = (PERIODIC, PERIODIC, PERIODIC)
    fargo: int = 0
    iso: int = 0
    precision: int = 0  # 0 float64, 1 float32 bulk
    nu: float = 0.0     # viscosity (model_params nu)
    eta: float = 0.0    # resistivity (model_params eta)

    @property
    def dx(self):
        return (self.xmax - self.xmin) / self.nx

    @property
    def dy(self):
        return (self.ymax - self.ymin) / self.ny

    @property
    def dz(self):
        return (self.zmax - self.zmin) / self.nz

    def replace(self, **kw) -> "Params":
        q = copy.copy(self)
        for k, v in kw.items():
            setattr(q, k, v)
        return q

    def to_dict(self) -> dict:
        d = dict(self.__dict__)
        d["bdtype"] = list(d["bdtype"])
        return d


@dataclass
class Problem:
    params: Params
    uin: np.ndarray  # (11, nz+6, ny+6, nx+6) float64
    name: str = ""
    meta: dict = field(default_factory=dict)


def centers(lo: float, d: float, n: int) -> np.ndarray:
    """Cell centres for i = -2..n+3 as the reference computes them."""
    i = np.arange(-2, n + 4, dtype=np.float64)
    return (lo + d / 2) + (i - 1) * d


def faces(lo: float, d: float, n: int) -> np.ndarray:
    """Left-face positions for i = -2..n+3."""
    i = np.arange(-2, n + 4, dtype=np.float64)
    return lo + (i - 1) * d


def fill_periodic(u: np.ndarray) -> np.ndarray:
    """Periodic ghost copy of all variables in the reference order x, z, y
    (boundary.f90:27-80), then the right-face copy (boundary.f90:82-89)."""
    nzg, nyg, nxg = u.shape[1:]
    nx, ny, nz = nxg - 6, nyg - 6, nzg - 6
    # x: ghosts -2..0 <- nx-2..nx ; nx+1..nx+3 <- 1..3  (array index = i + 2)
    u[:, :, :, 0:3] = u[:, :, :, nx - 2 + 2: nx + 1 + 2]
    u[:, :, :, nx + 3: nx + 6] = u[:, :, :, 3:6]
    u[:, 0:3, :, :] = u[:, nz - 2 + 2: nz + 1 + 2, :, :]
    u[:, nz + 3: nz + 6, :, :] = u[:, 3:6, :, :]
    u[:, :, 0:3, :] = u[:, :, ny - 2 + 2: ny + 1 + 2, :]
    u[:, :, ny + 3: ny + 6, :] = u[:, :, 3:6, :]
    right_faces(u)
    return u


def right_faces(u: np.ndarray) -> None:
    """B_R(i) = B_L(i+1) for i < iu2 in each direction (boundary.f90:82-89)."""
    u[8, :, :, :-1] = u[5, :, :, 1:]
    u[9, :, :-1, :] = u[6, :, 1:, :]
    u[10, :-1, :, :] = u[7, 1:, :, :]


def assemble(p: Params, rho, vx, vy, vz, pres, bxl, byl, bzl) -> np.ndarray:
    """Builds uin from interior primitive fields (shape (nz, ny, nx)) and
    interior left-face B.  Total energy uses the face-averaged cell B, as the
    reference's primitive() inverts it (scheme.f90:55-65)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    u = np.zeros((11, nz + 6, ny + 6, nx + 6))
    s = (slice(3, nz + 3), slice(3, ny + 3), slice(3, nx + 3))
    u[(5,) + s] = bxl
    u[(6,) + s] = byl
    u[(7,) + s] = bzl
    fill_periodic(u)
    bxc = 0.5 * (u[(5,) + s] + u[(8,) + s])
    byc = 0.5 * (u[(6,) + s] + u[(9,) + s])
    bzc = 0.5 * (u[(7,) + s] + u[(10,) + s])
    u[(0,) + s] = rho
    u[(1,) + s] = rho * vx
    u[(2,) + s] = rho * vy
    u[(3,) + s] = rho * vz
    if p.iso:
        u[(4,) + s] = 0.0
    else:
        u[(4,) + s] = (pres / (p.gamma - 1.0) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
                       + 0.5 * (bxc * bxc + byc * byc + bzc * bzc))
    return fill_periodic(u)


def curl_faces(p: Params, ax, ay, az):
    """Face B = curl A on the staggered grid.  ax lives on x-edges
    (x_i, y_{j-1/2}, z_{k-1/2}), ay on y-edges, az on z-edges; inputs are
    periodic interior arrays of shape (nz, ny, nx)."""
    dx, dy, dz = p.dx, p.dy, p.dz
    r = np.roll
    bx = (r(az, -1, axis=1) - az) / dy - (r(ay, -1, axis=0) - ay) / dz
    by = (r(ax, -1, axis=0) - ax) / dz - (r(az, -1, axis=2) - az) / dx
    bz = (r(ay, -1, axis=2) - ay) / dx - (r(ax, -1, axis=1) - ax) / dy
    return bx, by, bz


def _mesh(p: Params):
    xc = centers(p.xmin, p.dx, p.nx)[3:-3]
    yc = centers(p.ymin, p.dy, p.ny)[3:-3]
    zc = centers(p.zmin, p.dz, p.nz)[3:-3]
    xf = faces(p.xmin, p.dx, p.nx)[3:-3]
    yf = faces(p.ymin, p.dy, p.ny)[3:-3]
    zf = faces(p.zmin, p.dz, p.nz)[3:-3]
    return xc, yc, zc, xf, yf, zf


def orszag_tang(n: int = 256, **kw) -> Problem:
    """3D Orszag-Tang-type turbulence on the periodic unit box:
    rho = 25/(36 pi), P = 5/(12 pi), v = (-sin 2piy, sin 2pix, 0.1 sin 2piz),
    A_z = B0 (cos 2piy / 2pi + cos 4pix / 4pi), B0 = 1/sqrt(4 pi)."""
    p = Params(nx=n, ny=n, nz=n, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0,
               zmin=0.0, zmax=1.0, gamma=5.0 / 3.0, courant=0.7).replace(**kw)
    xc, yc, zc, xf, yf, zf = _mesh(p)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    tp = 2.0 * np.pi
    b0 = 1.0 / np.sqrt(4.0 * np.pi)
    shape = (p.nz, p.ny, p.nx)
    rho = np.full(shape, 25.0 / (36.0 * np.pi))
    pres = np.full(shape, 5.0 / (12.0 * np.pi))
    vx = -np.sin(tp * Y)
    vy = np.sin(tp * X)
    vz = 0.1 * np.sin(tp * Z)
    _, YF, XF = np.meshgrid(zc, yf, xf, indexing="ij")
    az = b0 * (np.cos(tp * YF) / tp + np.cos(2 * tp * XF) / (2 * tp))
    zero = np.zeros(shape)
    bx, by, bz = curl_faces(p, zero, zero, az)
    return Problem(p, assemble(p, rho, vx, vy, vz, pres, bx, by, bz), "orszag_tang")


def magnetic_loop(nx: int = 128, ny: int = 128, nz: int = 4, **kw) -> Problem:
    """Field-loop advection: [-1,1]x[-0.5,0.5]x[0,0.03], rho = P = 1,
    v = (2, 1, 0), A_z = 1e-3 (0.3 - r) for r < 0.3."""
    p = Params(nx=nx, ny=ny, nz=nz, xmin=-1.0, xmax=1.0, ymin=-0.5, ymax=0.5,
               zmin=0.0, zmax=0.03, gamma=5.0 / 3.0, courant=0.7).replace(**kw)
    xc, yc, zc, xf, yf, zf = _mesh(p)
    shape = (p.nz, p.ny, p.nx)
    _, YF, XF = np.meshgrid(zc, yf, xf, indexing="ij")
    r = np.sqrt(XF * XF + YF * YF)
    az = np.where(r < 0.3, 1e-3 * (0.3 - r), 0.0)
    zero = np.zeros(shape)
    bx, by, bz = curl_faces(p, zero, zero, az)
    one = np.ones(shape)
    return Problem(p, assemble(p, one, 2.0 * one, one, zero, one, bx, by, bz), "magnetic_loop")


def sound_wave(n: int = 64, beta: float = 1e-3, amp: float = 1e-3, **kw) -> Problem:
    """A small sound wave in a strong uniform field (low beta) on the
    periodic unit box: rho = 1, P = 1, B uniform along (1, 0.5, 0.25) with
    |B|^2 = 2 P / beta, and along x the acoustic perturbation of amplitude
    amp: drho = amp rho sin(2 pi x), dvx = amp c_s sin(2 pi x),
    dP = amp gamma P sin(2 pi x), c_s = sqrt(gamma P / rho).  With P = 1 the
    total energy is 1/beta times the internal energy, so the float64
    pressure recovery E - B^2/2 is what keeps the pressure's digits."""
    p = Params(nx=n, ny=n, nz=n, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0,
               zmin=0.0, zmax=1.0, gamma=5.0 / 3.0, courant=0.7).replace(**kw)
    xc, yc, zc, xf, yf, zf = _mesh(p)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    shape = (p.nz, p.ny, p.nx)
    cs = np.sqrt(p.gamma)
    s = np.sin(2.0 * np.pi * X)
    rho = 1.0 + amp * s
    vx = amp * cs * s
    pres = 1.0 + amp * p.gamma * s
    b0 = np.sqrt(2.0 / beta)
    d = np.array([1.0, 0.5, 0.25])
    d = d / np.sqrt((d * d).sum())
    zero = np.zeros(shape)
    return Problem(p, assemble(p, rho, vx, zero, zero, pres, zero + b0 * d[0], zero + b0 * d[1],
                               zero + b0 * d[2]), "sound_wave")


def pressure(p: Params, u: np.ndarray) -> np.ndarray:
    """Interior gas pressure (gamma - 1)(E - rho v^2/2 - B^2/2) of a state,
    with the face-averaged cell field (scheme.f90:55-65), in float64."""
    r = u[0, 3:-3, 3:-3, 3:-3]
    ek = 0.5 * (u[1, 3:-3, 3:-3, 3:-3] ** 2 + u[2, 3:-3, 3:-3, 3:-3] ** 2 + u[3, 3:-3, 3:-3, 3:-3] ** 2) / r
    b = [0.5 * (u[5 + c, 3:-3, 3:-3, 3:-3] + u[8 + c, 3:-3, 3:-3, 3:-3]) for c in range(3)]
    em = 0.5 * (b[0] ** 2 + b[1] ** 2 + b[2] ** 2)
    return (p.gamma - 1.0) * (u[4, 3:-3, 3:-3, 3:-3] - ek - em)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0,1) from the SplitMix64 generator (Steele, Lea &
    Flood 2014): state_i = seed + (i+1)*0x9E3779B97F4A7C15 mod 2**64, output
    the top 53 bits of the mixed state.  Portable and seed-stable."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64)
                               * np.uint64(0x9E3779B97F4A7C15))
        z = (s ^ (s >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def shearing_box(nx: int = 128, ny: int = 256, nz: int = 128, seed: int = 20261202,
                 fargo: int = 1, **kw) -> Problem:
    """Unstratified shearing-box MRI: box 1 x 4 x 1, Omega = 1, q = 1.5,
    rho = 1, P = 1e-3, net vertical field at beta = 1000, uniform random
    velocity perturbations of amplitude 1e-3 c_s (c_s = sqrt(gamma P / rho)).
    Velocities are relative to the background shear, as in the reference's
    shearing-box formulation (src/problem/mri/condinit.f90:68-85)."""
    p = Params(nx=nx, ny=ny, nz=nz, xmin=-0.5, xmax=0.5, ymin=-2.0, ymax=2.0,
               zmin=-0.5, zmax=0.5, gamma=5.0 / 3.0, omega0=1.0, courant=0.7,
               bdtype=(SHEARINGBOX, PERIODIC, PERIODIC), fargo=fargo).replace(**kw)
    shape = (p.nz, p.ny, p.nx)
    n = p.nx * p.ny * p.nz
    p0 = 1e-3
    cs = np.sqrt(p.gamma * p0)
    amp = 1e-3 * cs
    rnd = splitmix64(seed, 3 * n).reshape((3,) + shape)
    vx, vy, vz = (amp * (rnd[c] - 0.5) for c in range(3))
    b0 = np.sqrt(2.0 * p0 / 1000.0)
    zero = np.zeros(shape)
    rho = np.ones(shape)
    return Problem(p, assemble(p, rho, vx, vy, vz, p0 * rho, zero, zero, b0 + zero),
                   "shearing_box", {"seed": seed})


def mri_test(nx: int = 32, ny: int = 64, nz: int = 32, **kw) -> Problem:
    """Isothermal shearing-box MRI of the reference's own test suite
    (utils/test/test.py:975-990: problem mri, type 'Test', amp = 0, ISO=1):
    box [-0.5,0.5]x[0,4]x[-0.5,0.5], Omega0 = ciso = 1e-3, beta = 1000,
    gamma = 1.001 (utils/test/test.py:112), courant 0.8, shearing-periodic x;
    rho = 1, rho*vx = 1e-5 cos(20 pi z) cos(10 pi x), Bz_L = Bz_R =
    B0 sin(2 pi x) with B0 = sqrt(2 p0 / beta), p0 = ciso^2
    (src/problem/mri/condinit.f90:38-108)."""
    p = Params(nx=nx, ny=ny, nz=nz, xmin=-0.5, xmax=0.5, ymin=0.0, ymax=4.0, zmin=-0.5,
               zmax=0.5, gamma=1.001, ciso=1e-3, omega0=1e-3, courant=0.8, iso=1,
               bdtype=(SHEARINGBOX, PERIODIC, PERIODIC)).replace(**kw)
    tp = 2.0 * np.pi
    p0 = 1.0 * p.ciso * p.ciso
    b0 = np.sqrt(2.0 * p0 / 1000.0)
    xc = centers(p.xmin, p.dx, p.nx)
    zc = centers(p.zmin, p.dz, p.nz)
    u = np.zeros((11, p.nz + 6, p.ny + 6, p.nx + 6))
    Z, X = zc[:, None, None], xc[None, None, :]
    u[0] = 1.0
    u[1] = 1.0e-2 * 1.0e-3 * np.cos(tp * 10.0 * Z) * np.cos(tp * 5.0 * X) + 0.0 * u[1]
    bz = b0 * np.sin(tp * X) + 0.0 * u[7]
    u[7] = bz
    u[10] = bz
    # condinit sets every cell, ghosts included; init() then applies boundary
    return Problem(p, fill_periodic(u), "mri_test")


def random_smooth(nx: int, ny: int, nz: int, seed: int = 1, kmax: int = 2,
                  mach: float = 1.0, beta: float = 1.0, b0=(0.3, 0.2, 0.1), **kw) -> Problem:
    """Random smooth periodic state for parity tests: a few Fourier modes of
    wavenumber <= kmax in rho, P, v and in the vector potential, plus a uniform
    field b0, with order-one Mach and plasma-beta numbers so every branch of
    the solvers is exercised."""
    p = Params(nx=nx, ny=ny, nz=nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0,
               zmin=0.0, zmax=1.0).replace(**kw)
    g = np.random.Generator(np.random.PCG64(seed))
    xc, yc, zc, xf, yf, zf = _mesh(p)
    tp = 2.0 * np.pi

    def field(xs, ys, zs):
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        f = np.zeros_like(X)
        for _ in range(4):
            k = g.integers(-kmax, kmax + 1, size=3)
            a, ph = g.normal(), g.uniform(0, tp)
            f += a * np.cos(tp * (k[0] * X + k[1] * Y + k[2] * Z) + ph)
        return f / 4.0

    rho = 1.0 + 0.4 * np.tanh(field(xc, yc, zc))
    pres = 1.0 + 0.4 * np.tanh(field(xc, yc, zc))
    cs = np.sqrt(p.gamma)
    vx, vy, vz = (mach * cs * field(xc, yc, zc) for _ in range(3))
    bamp = np.sqrt(2.0 / beta)
    ax = bamp * field(xc, yf, zf) / tp
    ay = bamp * field(xf, yc, zf) / tp
    az = bamp * field(xf, yf, zc) / tp
    bx, by, bz = curl_faces(p, ax, ay, az)
    bx = bx + b0[0]
    by = by + b0[1]
    bz = bz + b0[2]
    return Problem(p, assemble(p, rho, vx, vy, vz, pres, bx, by, bz), "random_smooth",
                   {"seed": seed})


def _spectral_field(shape, g, kmax: int, slope: float = -2.0) -> np.ndarray:
    """Periodic random-phase Gaussian field with power spectrum P(k) ~ k^slope
    for 1 <= |k| <= kmax (k in units of the box), zero mean, unit rms."""
    nz, ny, nx = shape
    kz = np.fft.fftfreq(nz) * nz
    ky = np.fft.fftfreq(ny) * ny
    kx = np.fft.rfftfreq(nx) * nx
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    amp = np.zeros_like(k)
    m = (k >= 1.0) & (k <= kmax)
    # P(k) per mode ~ k^slope / k^2 (shell area), amplitude = sqrt of that
    amp[m] = np.sqrt(k[m] ** (slope - 2.0))
    ph = g.uniform(0.0, 2.0 * np.pi, k.shape)
    spec = amp * g.standard_normal(k.shape) * np.exp(1j * ph)
    f = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2))
    f -= f.mean()
    return f / max(np.sqrt((f * f).mean()), 1e-300)


def turbulence(n, seed: int = 4242, mach: float = 0.5, beta: float = 10.0, kmax: int = 8,
               log=None, **kw) -> Problem:
    """Decaying MHD turbulence in the periodic unit box (n int or (nx,ny,nz)):
    rho = 1, P = 1/gamma (sound speed 1), random-phase Gaussian velocity and
    vector potential with power spectrum k^-2 for k <= kmax, velocity scaled
    to rms Mach `mach`, field scaled to plasma beta `beta` (B_rms^2 =
    2 P / beta); face B is the discrete curl of the potential."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    p = Params(nx=nx, ny=ny, nz=nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0,
               zmin=0.0, zmax=1.0).replace(**kw)
    g = np.random.Generator(np.random.PCG64(seed))
    shape = (p.nz, p.ny, p.nx)
    pres0 = 1.0 / p.gamma

    def field(name):
        f = _spectral_field(shape, g, kmax)
        if log:
            log(f"[turbulence] {nx}x{ny}x{nz}: field {name} done")
        return f

    v = [field(f"v{c}") for c in "xyz"]
    vrms = np.sqrt(sum((c * c).mean() for c in v))
    v = [mach * c / vrms for c in v]
    a = [field(f"A{c}") for c in "xyz"]
    bx, by, bz = curl_faces(p, *a)
    del a
    if log:
        log(f"[turbulence] {nx}x{ny}x{nz}: face field done")
    brms = np.sqrt((bx * bx + by * by + bz * bz).mean())
    scale = np.sqrt(2.0 * pres0 / beta) / max(brms, 1e-300)
    one = np.ones(shape)
    u = assemble(p, one, v[0], v[1], v[2], pres0 * one, scale * bx, scale * by, scale * bz)
    if log:
        log(f"[turbulence] {nx}x{ny}x{nz}: state assembled")
    return Problem(p, u, "turbulence", {"seed": seed, "mach": mach, "beta": beta})


def ensemble(nboxes: int = 48, seed: int = 4242, sizes=(32, 48, 64, 96, 128), mach: float = 0.5,
             beta: float = 10.0, kmax: int = 8, **kw) -> list:
    """Independent periodic turbulence boxes of uneven sizes: box b has size
    n_b^3 with n_b drawn from `sizes` by a PCG64 stream of `seed`, and the
    initial fields of turbulence(n_b, seed + b, mach, beta, kmax)."""
    g = np.random.Generator(np.random.PCG64(seed))
    ns = g.choice(np.asarray(sizes), size=nboxes)
    return [turbulence(int(n), seed=seed + b, mach=mach, beta=beta, kmax=kmax, **kw) for b, n in enumerate(ns)]


def interior(u: np.ndarray, nvar: int = 8) -> np.ndarray:
    """The interior cells of the first nvar variables."""
    return u[:nvar, 3:-3, 3:-3, 3:-3]
