"""Writes tests/golden/*.npz: small FP64 oracle runs whose outputs later
runs of the oracle and of the GPU library are compared against.

The reference has no golden vectors (no tests, no recorded outputs), so
these fixtures are produced by the oracle itself; they pin its behaviour
against regressions and give the GPU tests inputs that need no oracle
build.  Run:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import write_deck  # noqa: E402
from oracle_binding import Oracle  # noqa: E402

CASES = {
    # plain FARGO physics on a 24 x 48 log grid, 8 substeps
    "plain_24x48": dict(Nrad=24, Nsec=48),
    # adiabatic with relaxation cooling, 20 x 40, 8 substeps
    "adiabatic_20x40": dict(Nrad=20, Nsec=40, Adiabatic="YES", Cooling="YES", CoolingTime0=62.8),
}
NSUB = 8


def run(case, tmp):
    path = write_deck(os.path.join(tmp, case + ".par"), OutputDir=tmp + "/", **CASES[case])
    o = Oracle(path)
    for _ in range(NSUB):
        o.substep()
    out = {n: o.field(n) for n in ("Dens", "Vrad", "Vtheta")}
    if "Adiabatic" in CASES[case]:
        out["Energy"] = o.field("Energy")
    out["planets"] = o.planets()
    out["scalars"] = np.array([o.get("PhysicalTime"), o.get("OmegaFrame"), o.get("dt")])
    out["nshift"] = o.ifield("Nshift")
    return out


if __name__ == "__main__":
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        for case in CASES:
            np.savez_compressed(os.path.join(HERE, case + ".npz"), **run(case, tmp))
            print("wrote", case)
