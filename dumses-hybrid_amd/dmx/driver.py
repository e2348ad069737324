"""The reference time loop around a domain (src/dumses.f90:75-160).

Works with any object exposing advance(nsteps, time) -> (time, dts) and
history() -> 9 sums: the HIP domain (dmx.abi.Domain) or the oracle.
"""
from __future__ import annotations

import numpy as np


def run_history(dom, tlim: float, dthist: float, time: float = 0.0, max_steps: int = 10 ** 9,
                shear: bool = False):
    """Advances until (time - dt) > tlim, recording a history row
    [time, dt, mass, maxwell, reynolds, max+rey, magp, meanBx, meanBy, meanBz,
    divB] at time 0 and at the first step past every multiple of dthist, with
    the cadence test of src/dumses.f90:120-131.  shear selects the
    shearing-box stresses of the MRI problem's history."""
    rows = []
    if time == 0.0:
        rows.append(np.concatenate([[time, 0.0], dom.history(shear)]))
    thist = time
    nstep = 0
    while nstep < max_steps:
        time, dts = dom.advance(1, time)
        dt = float(dts[0])
        nstep += 1
        if (time - dt) <= (thist + dthist) and time > (thist + dthist):
            rows.append(np.concatenate([[time, dt], dom.history(shear)]))
            thist = thist + dthist
        if (time - dt) > tlim:
            break
    return np.array(rows), nstep
