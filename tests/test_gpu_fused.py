"""Fused substep path (fg_step_fused) vs the FP64 oracle.

The fused kernels regroup the per-routine arithmetic into five passes;
every expression keeps the reference's evaluation order, so the results
must stay bit-identical to the oracle substep after substep.
"""
import os

import numpy as np
import pytest

from conftest import DECKS, write_planets
from oracle_binding import Oracle

pytestmark = pytest.mark.gpu


def run_fused(path, nsub, check_every=1):
    from cudaprotoplanetarydisks_amd import Simulation
    sim = Simulation(path, fused=True)
    assert sim.get("fused_active") == 1.0
    orc = Oracle(path)
    for k in range(nsub):
        sim.substep()
        orc.substep()
        if (k + 1) % check_every and k != nsub - 1:
            continue
        assert int(sim.get("last_cfl")) == orc.iget("last_cfl"), f"substep {k}: CFL count"
        for name in ("PhysicalTime", "OmegaFrame", "dt"):
            assert sim.get(name) == orc.get(name), f"substep {k}: {name}"
        d, vr, vt = sim.state()
        for nm, g in (("Dens", d), ("Vrad", vr), ("Vtheta", vt)):
            o = orc.field(nm)
            if not np.array_equal(g, o):
                diff = np.argwhere(g != o)
                err = np.abs(g - o).max() / np.abs(o).max()
                pytest.fail(f"substep {k}: {nm} differs at {diff[:6].tolist()} (rel Linf {err:.3e})")
        assert np.array_equal(sim.ctx.nshift(), orc.ifield("Nshift")[:-1]), f"substep {k}: Nshift"
        assert np.array_equal(sim.planets(), orc.planets())
    return sim, orc


@pytest.mark.parametrize("case", [
    dict(Nrad=37, Nsec=50),
    dict(Nrad=21, Nsec=130, Frame="FIXED", OmegaFrame=0.0),
    dict(Nrad=40, Nsec=96, Viscosity=None, AlphaViscosity=1e-3, SigmaSlope=0.5, FlaringIndex=0.1),
    dict(Nrad=33, Nsec=64, Transport="STANDARD"),
    dict(Nrad=48, Nsec=200, RadialSpacing="A", ExcludeHill="YES", IndirectTerm="NO"),
], ids=["odd", "fixed_frame", "alpha_flaring", "standard", "arith_nohill"])
def test_fused_cases_bitexact(deck, case):
    run_fused(deck(**case), 15)


@pytest.mark.parametrize("case", [
    dict(Nrad=12, Nsec=3000),  # k_fin's single sweep with its last staging chunk and cells partial
    dict(Nrad=10, Nsec=3072),  # the widest ring of k_fin's single sweep
    dict(Nrad=12, Nsec=3200),  # k_fin's two-pass ring kernel (NSEC > 3072)
    dict(Nrad=70, Nsec=4096),  # the widest ring the fused path takes
], ids=["nsec3000", "nsec3072", "nsec3200", "nsec4096"])
def test_fused_wide_rings_bitexact(deck, case):
    run_fused(deck(**case), 4)


def test_fused_planets_feel_disk(deck, tmp_path):
    cfg = write_planets(tmp_path / "p.cfg", [("A", 1.0, 1e-3, 0.0, "YES", "NO"), ("B", 1.7, 5e-4, 0.0, "YES", "YES")])
    run_fused(deck(Nrad=32, Nsec=96, PlanetConfig=cfg), 15)


def test_fused_plain_fargo_100_substeps():
    """c1 workload, 100 substeps on the fused path: bit-exact."""
    run_fused(os.path.join(DECKS, "c1_plain_fargo.par"), 100, check_every=10)


def test_fused_temperature_output(deck):
    from cudaprotoplanetarydisks_amd import Simulation
    path = deck(Nrad=24, Nsec=64, WriteTemperature="YES")
    sim = Simulation(path, fused=True)
    orc = Oracle(path)
    for _ in range(6):
        sim.substep()
        orc.substep()
    assert np.array_equal(sim.ctx.field("TEMPERATURE")[:-1], orc.field("Temperature")[:-1])
