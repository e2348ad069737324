"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in the build container, where /root/reference exists:
    make -f oracle/ref.mk && python tests/golden/make_golden.py

Every fixture is data only: the input state (interior cells; ghosts are
rebuilt with dmx.problems.fill_periodic, the same periodic copy the input was
built with), the run parameters, and the reference's outputs (per-step dt and
the interior state after the listed steps).  The reference runs its serial
build (oracle/_ref/iso{0,1}/ref_harness), injected through its own binary
restart reader.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "dumses-hybrid_amd"))

import oracle as O  # noqa: E402
from dmx import problems as P  # noqa: E402


def perturbed_shear(nx, ny, nz, fargo, iso, seed=7):
    pb = P.shearing_box(nx, ny, nz, fargo=fargo)
    p = pb.params
    if iso:
        p = p.replace(iso=1, ciso=0.05)
    u = pb.uin.copy()
    g = np.random.Generator(np.random.PCG64(seed))
    u[1:4, 3:-3, 3:-3, 3:-3] += 0.02 * g.standard_normal(u[1:4, 3:-3, 3:-3, 3:-3].shape)
    P.fill_periodic(u)
    return p, u


def step_cases():
    rs = P.random_smooth(16, 12, 8, seed=11)
    base = rs.params
    yield "step_hlld_mc", base, rs.uin, [1, 5]
    yield "step_hll_llf_minmod", base.replace(riemann=P.HLL, riemann2d=P.LLF, slope_type=1), rs.uin, [1, 5]
    yield "step_llf_upwind_multid", base.replace(riemann=P.LLF, riemann2d=P.UPWIND, slope_type=3), rs.uin, [1, 5]
    yield "step_upwind_hlla_flat", base.replace(riemann=P.UPWIND, riemann2d=P.HLLA, slope_type=0), rs.uin, [1, 5]
    yield "step_hlld_zerogradient", base.replace(bdtype=(P.ZERO_GRADIENT, P.PERIODIC, P.ZERO_GRADIENT)), rs.uin, [1, 5]
    ot = P.orszag_tang(16)
    yield "orszag_tang_16", ot.params, ot.uin, [1, 10]
    p, u = perturbed_shear(8, 16, 6, fargo=0, iso=0)
    yield "shear_nofargo", p, u, [1, 8]
    p, u = perturbed_shear(8, 16, 6, fargo=1, iso=0)
    yield "shear_fargo", p, u, [1, 8]
    p, u = perturbed_shear(8, 16, 6, fargo=0, iso=1)
    yield "shear_iso", p, u, [1, 8]
    p, u = perturbed_shear(8, 16, 6, fargo=1, iso=1)
    yield "shear_fargo_iso", p, u, [1, 8]
    # dissipation, physics.f90:272-664 (after boundary, dumses.f90:103-105)
    yield "diss_both", base.replace(nu=2e-3, eta=3e-3), rs.uin, [1, 5]
    yield "diss_eta_iso_zg", base.replace(eta=5e-3, iso=1, ciso=0.7,
                                          bdtype=(P.ZERO_GRADIENT, P.PERIODIC, P.PERIODIC)), rs.uin, [1, 5]
    yield "diss_nu_hll_minmod", base.replace(nu=4e-3, slope_type=1, riemann=P.HLL,
                                             riemann2d=P.LLF), rs.uin, [1, 5]
    p, u = perturbed_shear(8, 16, 6, fargo=1, iso=0)
    yield "shear_fargo_diss", p.replace(nu=1e-4, eta=2e-4), u, [1, 8]


def write_step(name, p, uin, steps):
    nmax = max(steps)
    ref, dts, dumps = O.run_reference(p, uin, nmax, dumpevery=1)
    out = {f"out_{s:04d}": P.interior(dumps[s]) for s in steps}
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        params=json.dumps(p.to_dict()), ic=P.interior(uin), dts=dts,
        steps=np.array(steps), full_last=ref if uin.size < 40000 else np.zeros(0), **out)
    print(f"{name}: {nmax} steps, dt {dts[0]:.6g} .. {dts[-1]:.6g}")


def write_loop():
    """Field-loop advection, 128x128x4, 200 steps: per-step dt, the history
    sums of the final state as the reference's own history routine prints
    them (mass .. divB, with the half-unit rounding bound of each printed
    value), a 1-in-4 subsample of the final interior and the sha256 of the
    full final interior (the full state is 4 MB)."""
    pb = P.magnetic_loop()
    p = pb.params
    ref, dts, _ = O.run_reference(p, pb.uin, 200)
    hv, hh = O.run_reference_history(p, ref, float(np.sum(dts)))
    inner = np.ascontiguousarray(P.interior(ref))
    np.savez_compressed(
        os.path.join(HERE, "magnetic_loop_200.npz"), params=json.dumps(p.to_dict()),
        dts=dts, history=hv[2:], history_hulp=hh[2:], sub=inner[:, :, ::4, ::4],
        sha256=hashlib.sha256(inner.tobytes()).hexdigest())
    print(f"magnetic_loop_200: dt {dts[0]:.6g} .. {dts[-1]:.6g}")


def history_cases():
    """States on which the reference's history routine is run: a random
    smooth field (distinct By/Bz and v/w everywhere, so a swapped component
    changes every stress), the Orszag-Tang 16^3 state after 10 steps and the
    shearing box with FARGO after 8 steps (both from their step fixtures)."""
    rs = P.random_smooth(16, 12, 8, seed=11)
    yield "random_smooth", rs.params, rs.uin, 0.25
    for name, step in (("orszag_tang_16", 10), ("shear_fargo", 8)):
        z = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
        d = json.loads(str(z["params"]))
        d["bdtype"] = tuple(d["bdtype"])
        p = P.Params(**d)
        u = np.zeros((11, p.nz + 6, p.ny + 6, p.nx + 6))
        u[:8, 3:-3, 3:-3, 3:-3] = z[f"out_{step:04d}"]
        P.fill_periodic(u)
        yield name, p, u, float(np.sum(z["dts"][:step]))


def write_history():
    """history_reference.npz: per case the interior state, its time and the
    row the reference's own history routine (src/subroutines/history.f90,
    called by the harness on the injected state) prints for it."""
    out = {}
    names = []
    for name, p, u, t in history_cases():
        hv, hh = O.run_reference_history(p, u, t)
        out[f"{name}_params"] = json.dumps(p.to_dict())
        out[f"{name}_ic"] = P.interior(u)
        out[f"{name}_row"] = hv
        out[f"{name}_hulp"] = hh
        names.append(name)
        print(f"history {name}: maxwell {hv[3]:.6g} reynolds {hv[4]:.6g}")
    np.savez_compressed(os.path.join(HERE, "history_reference.npz"), cases=np.array(names), **out)


def random_states(n, seed):
    """Physically valid random 1D state pairs and 2D corner sets that cover
    every sampling branch of the HLLD solvers (super/subsonic both ways,
    weak and strong fields, zero normal field)."""
    g = np.random.Generator(np.random.PCG64(seed))

    def state(m):
        r = np.exp(g.uniform(-1.5, 1.5, m))
        pr = np.exp(g.uniform(-2.0, 1.5, m))
        v = g.normal(0, 1, (m, 3)) * g.choice([0.1, 1.0, 4.0], (m, 1))
        b = g.normal(0, 1, (m, 3)) * g.choice([0.0, 0.3, 1.5], (m, 1))
        return r, pr, v, b

    r, pr, v, b = state(2 * n)
    s1 = np.empty((n, 16))
    for side in range(2):
        o = side * 8
        sl = slice(side * n, (side + 1) * n)
        s1[:, o + 0] = r[sl]
        s1[:, o + 1] = pr[sl]
        s1[:, o + 2] = v[sl, 0]
        s1[:, o + 3] = b[sl, 0]
        s1[:, o + 4] = v[sl, 1]
        s1[:, o + 5] = b[sl, 1]
        s1[:, o + 6] = v[sl, 2]
        s1[:, o + 7] = b[sl, 2]
    # the solvers are called with equal normal fields (bn_mean, solver.f90:139)
    s1[:, 11] = s1[:, 3]
    s1[: n // 10, 3] = 0.0
    s1[: n // 10, 11] = 0.0
    r, pr, v, b = state(4 * n)
    s2 = np.empty((n, 32))
    for c in range(4):
        sl = slice(c * n, (c + 1) * n)
        o = c * 8
        s2[:, o + 0] = r[sl]
        s2[:, o + 1] = pr[sl]
        s2[:, o + 2] = v[sl, 0]
        s2[:, o + 3] = v[sl, 1]
        s2[:, o + 4] = b[sl, 0]
        s2[:, o + 5] = b[sl, 1]
        s2[:, o + 6] = b[sl, 2]
    # continuity of the tangential field pairs, as packed by
    # riemann_solver_magnetic (solver_magnetic.f90:161-171):
    # A(LL)=A(RL), A(LR)=A(RR), B(LL)=B(LR), B(RL)=B(RR)
    s2[:, 8 + 4] = s2[:, 0 + 4]
    s2[:, 24 + 4] = s2[:, 16 + 4]
    s2[:, 16 + 5] = s2[:, 0 + 5]
    s2[:, 24 + 5] = s2[:, 8 + 5]
    for c in range(4):
        o = c * 8
        s2[:, o + 7] = s2[:, o + 2] * s2[:, o + 5] - s2[:, o + 3] * s2[:, o + 4]
    return s1, s2


def write_kernels():
    n = 1000
    s1, s2 = random_states(n, 2024)
    gamma, ciso, dt = 5.0 / 3.0, 0.7, 0.013
    res = {}
    for iso in (0, 1):
        raw = O.run_reference_kernels(iso, gamma, ciso, dt, s1, s2)
        one = raw[: n * 39].reshape(n, 39)
        two = raw[n * 39:].reshape(n, 4)
        res[f"iso{iso}_hlld"] = one[:, 0:15]
        res[f"iso{iso}_hll"] = one[:, 15:23]
        res[f"iso{iso}_llf"] = one[:, 23:31]
        res[f"iso{iso}_upwind"] = one[:, 31:39]
        res[f"iso{iso}_emf"] = two  # hlld, llf, upwind, hlla
    np.savez_compressed(os.path.join(HERE, "riemann_kernels.npz"), s1=s1, s2=s2,
                        gamma=gamma, ciso=ciso, dt=dt, **res)
    print("riemann_kernels: %d states" % n)


def copy_mri_history():
    """The reference's own numeric golden file: the history of its isothermal
    shearing-box MRI test (utils/test/fig/reference/shearing/mri/history.txt,
    written by src/subroutines/history.f90:148-165), kept as data."""
    src = "/root/reference/utils/test/fig/reference/shearing/mri/history.txt"
    rows = [ln for ln in open(src) if ln.strip() and not ln.lstrip().startswith("#")]
    vals = np.array([[float(x) for x in ln.split()] for ln in rows])
    np.savez_compressed(os.path.join(HERE, "mri_history_reference.npz"), history=vals,
                        columns=np.array(["time", "dt", "mass", "maxwell", "reynolds", "max+rey",
                                          "magp", "meanBx", "meanBy", "meanBz", "divB"]))
    print(f"mri_history_reference: {vals.shape[0]} rows")


def main(only=None):
    """Writes every fixture, or only those named in `only` (step fixture
    names, "loop", "history")."""
    if not (O.reference_available(0) and O.reference_available(1)):
        raise SystemExit("build the reference first: make -f oracle/ref.mk")
    if not only:
        copy_mri_history()
        write_kernels()
    for name, p, u, steps in step_cases():
        if not only or name in only:
            write_step(name, p, u, steps)
    if not only or "loop" in only:
        write_loop()
    if not only or "history" in only:
        write_history()


if __name__ == "__main__":
    main(sys.argv[1:])
