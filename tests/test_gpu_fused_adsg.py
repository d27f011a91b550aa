"""Fused substep path on adiabatic and self-gravitating decks.

The fused kernels evaluate the energy equation (the heating of Substep2,
Substep3 with or without cooling, the energy transport) and the
self-gravity kick with the per-routine kernels' expressions, so a fused run
must give the same bits as the per-routine run of the same deck, substep
after substep, across the DT boundaries where an adiabatic deck refreshes
its sound speed (SourceEuler.cu:207-213) and its alpha viscosity.  Against
the FP64 oracle both paths are held to 1e-12 on adiabatic decks (the
device exp/log of Qplus, Kernels.cu:93) and are bit-exact on isothermal
self-gravitating ones.
"""
import os

import numpy as np
import pytest

from conftest import DECKS
from oracle_binding import Oracle

pytestmark = pytest.mark.gpu

SG = dict(SelfGravity="YES", ThicknessSmoothing=0.3)
ADI = dict(Adiabatic="YES", AdiabaticIndex=1.4)


def state_of(sim, adiabatic):
    d, vr, vt = sim.state()
    out = {"dens": d, "vrad": vr, "vtheta": vt}
    if adiabatic:
        out["energy"] = sim.ctx.field("ENERGY")
    return out


def run_fused_vs_routine(path, nsub, adiabatic, sg=False):
    from cudaprotoplanetarydisks_amd import Simulation
    closed = 0
    with Simulation(path, fused=True) as fu, Simulation(path) as pr:
        assert fu.get("fused_active") == 1.0 and pr.get("fused_active") == 0.0
        for k in range(nsub):
            fu.substep()
            pr.substep()
            closed += fu.get("in_dt") == 0.0
            assert fu.get("last_cfl") == pr.get("last_cfl"), f"substep {k}: CFL count"
            for name in ("PhysicalTime", "OmegaFrame", "dt"):
                assert fu.get(name) == pr.get(name), f"substep {k}: {name}"
            a, b = state_of(fu, adiabatic), state_of(pr, adiabatic)
            for name in a:
                x, y = a[name], b[name]
                if name == "energy":  # the ghost ring is no state
                    x, y = x[:-1], y[:-1]
                if not np.array_equal(x, y):
                    bad = np.argwhere(x != y)
                    err = np.abs(x - y).max() / np.abs(y).max()
                    pytest.fail(f"substep {k}: {name} differs at {bad[:6].tolist()} (rel Linf {err:.3e})")
            assert np.array_equal(fu.ctx.nshift(), pr.ctx.nshift()), f"substep {k}: Nshift"
            assert np.array_equal(fu.planets(), pr.planets()), f"substep {k}: planets"
            if sg:
                assert np.array_equal(fu.ctx.field("SG_ACCR")[:-1], pr.ctx.field("SG_ACCR")[:-1]), f"substep {k}"
                assert np.array_equal(fu.ctx.field("SG_ACCT")[:-1], pr.ctx.field("SG_ACCT")[:-1]), f"substep {k}"
    return closed


def run_fused_vs_oracle(path, nsub, tol, adiabatic):
    from cudaprotoplanetarydisks_amd import Simulation
    orc = Oracle(path)
    with Simulation(path, fused=True) as sim:
        for k in range(nsub):
            sim.substep()
            orc.substep()
            assert int(sim.get("last_cfl")) == orc.iget("last_cfl"), f"substep {k}: CFL count"
            d, vr, vt = sim.state()
            pairs = [("Dens", d), ("Vrad", vr), ("Vtheta", vt)]
            if adiabatic:
                pairs.append(("Energy", sim.ctx.field("ENERGY")))
            for name, g in pairs:
                o = orc.field(name)
                g, o = g[:-1], o[:-1]
                err = np.abs(g - o).max() / np.abs(o).max()
                assert err <= tol, f"substep {k}: {name} rel Linf {err:.3e} > {tol:g}"
            assert np.array_equal(sim.ctx.nshift(), orc.ifield("Nshift")[:-1]), f"substep {k}: Nshift"


CASES = {
    "adiabatic": dict(Nrad=32, Nsec=96, DT=0.02, **ADI),
    "adiabatic_odd": dict(Nrad=37, Nsec=50, DT=0.02, **ADI),
    # k_fin's single sweep with the energy, partial staging chunk and cells
    "adiabatic_wide": dict(Nrad=12, Nsec=3000, DT=0.01, **ADI),
    "relax_cooling": dict(Nrad=32, Nsec=96, DT=0.02, Cooling="YES", CoolingTime0=62.8, **ADI),
    "beta_cooling": dict(Nrad=32, Nsec=96, DT=0.02, Cooling="YES", CoolingBeta=10.0, **ADI),
    "alpha_visc": dict(Nrad=32, Nsec=96, DT=0.02, Viscosity=None, AlphaViscosity=1e-3, **ADI),
    "alpha_corotating": dict(Nrad=40, Nsec=130, DT=0.015, Viscosity=None, AlphaViscosity=2e-3,
                             Frame="COROTATING", SigmaSlope=1.0, **ADI),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fused_adiabatic_matches_per_routine(deck, case):
    closed = run_fused_vs_routine(deck(**CASES[case]), 14, adiabatic=True)
    assert closed >= 2  # the run crossed DT boundaries


def test_fused_selfgravity_matches_per_routine(deck):
    run_fused_vs_routine(deck(Nrad=24, Nsec=48, DT=0.05, **SG), 12, adiabatic=False, sg=True)


def test_fused_selfgravity_bitexact_vs_oracle(deck):
    run_fused_vs_oracle(deck(Nrad=24, Nsec=48, DT=0.05, **SG), 10, 0.0, adiabatic=False)


def test_fused_adiabatic_vs_oracle(deck):
    run_fused_vs_oracle(deck(Nrad=32, Nsec=96, DT=0.02, Viscosity=None, AlphaViscosity=1e-3, Cooling="YES",
                             CoolingBeta=10.0, **ADI), 12, 1e-12, adiabatic=True)


def adsg_small(tmp_path, nrad, nsec, dt=None):
    src = open(os.path.join(DECKS, "c2_adsg.par")).read()
    src = src.replace("Nrad                512", f"Nrad                {nrad}")
    src = src.replace("Nsec                1536", f"Nsec                {nsec}")
    if dt:
        src = src.replace("DT                  0.314159265359", f"DT                  {dt}")
    src = src.replace("jupiter_feeldisk.cfg", os.path.join(DECKS, "jupiter_feeldisk.cfg"))
    src = src.replace("out_c2/", str(tmp_path) + "/")
    path = tmp_path / "c2_small.par"
    path.write_text(src)
    return str(path)


def test_fused_adsg_small_matches_per_routine(tmp_path):
    """The c2 physics (adiabatic, beta cooling, alpha viscosity, SG, planet
    feeling the disk, corotating frame) on reduced grids."""
    run_fused_vs_routine(adsg_small(tmp_path, 32, 48, dt=0.05), 12, adiabatic=True, sg=True)


def test_fused_adsg_small_vs_oracle(tmp_path):
    run_fused_vs_oracle(adsg_small(tmp_path, 32, 48, dt=0.05), 12, 1e-12, adiabatic=True)


def test_fused_adsg_c2_size_matches_per_routine(tmp_path):
    """The ADSG workload at its own size (512 x 1536, SG on the 1024 x 1536
    padded grid): 3 substeps, fused and per-routine bit-identical."""
    run_fused_vs_routine(adsg_small(tmp_path, 512, 1536), 3, adiabatic=True, sg=True)
