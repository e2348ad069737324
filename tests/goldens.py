"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

from dmx import problems as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

STEP_FIXTURES = ["step_hlld_mc", "step_hll_llf_minmod", "step_llf_upwind_multid",
                 "step_upwind_hlla_flat", "step_hlld_zerogradient", "orszag_tang_16",
                 "shear_nofargo", "shear_fargo", "shear_iso", "shear_fargo_iso",
                 "diss_both", "diss_eta_iso_zg", "diss_nu_hll_minmod", "shear_fargo_diss"]


def params_from_json(s) -> P.Params:
    d = json.loads(str(s))
    d["bdtype"] = tuple(d["bdtype"])
    return P.Params(**d)


def load_step(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    p = params_from_json(z["params"])
    ic = z["ic"]
    u = np.zeros((11, p.nz + 6, p.ny + 6, p.nx + 6))
    u[:8, 3:-3, 3:-3, 3:-3] = ic
    P.fill_periodic(u)
    outs = {int(s): z[f"out_{int(s):04d}"] for s in z["steps"]}
    full = z["full_last"] if z["full_last"].size else None
    return p, u, z["dts"], outs, full
