"""Restart from an output set (fg_sim_restart, the -s N of Main.cu:152-160
that the reference does not implement): a run restarted from output N
reproduces the uninterrupted run bit for bit, because the output files
hold the full FP64 state at a DT boundary."""
import os

import numpy as np
import pytest

from cudaprotoplanetarydisks_amd import Simulation
from conftest import write_planets

pytestmark = pytest.mark.gpu


def read_out(outdir, name, n, shape):
    return np.fromfile(os.path.join(outdir, name, f"{name}{n}.dat"), dtype=np.float64).reshape(shape)


@pytest.mark.parametrize("kw", [dict(), dict(Adiabatic="YES", AdiabaticIndex=1.4, Cooling="YES", CoolingBeta=10.0,
                                               WriteEnergy="YES"),
                                dict(fused=True)], ids=["isothermal", "adiabatic", "fused"])
def test_restart_reproduces_the_uninterrupted_run(deck, tmp_path, kw):
    kw = dict(kw)
    fused = kw.pop("fused", False)
    cfg = write_planets(tmp_path / "p.cfg", [("A", 1.0, 1e-3, 0.0, "YES", "NO")])
    full_dir, rs_dir = str(tmp_path / "full") + "/", str(tmp_path / "rs") + "/"
    for dd in (full_dir, rs_dir):
        os.makedirs(dd, exist_ok=True)
    common = dict(Nrad=24, Nsec=48, Ninterm=2, Ntot=6, PlanetConfig=cfg, **kw)
    d_full = deck("full.par", OutputDir=full_dir, **common)
    with Simulation(d_full, fused=fused, write_files=True) as s:
        s.run(0, 6)
    # second run: outputs up to set 1, then a fresh process-state restarts from it
    d_rs = deck("rs.par", OutputDir=rs_dir, **common)
    with Simulation(d_rs, fused=fused, write_files=True) as s:
        s.run(0, 2)  # writes sets 0 and 1 (i = 0, 2) and advances to i = 3
    with Simulation(d_rs, fused=fused, write_files=True) as s:
        s.restart(1)
        s.run(2, 6)
    shape = (24, 48)
    for name in ("dens", "vrad", "vtheta") + (("energy",) if "Adiabatic" in kw else ()):
        for n in (2, 3):
            a, b = read_out(full_dir, name, n, shape), read_out(rs_dir, name, n, shape)
            assert np.array_equal(a, b), f"{name}{n}"
    pa = open(os.path.join(full_dir, "planet0.dat")).read().splitlines()
    pb = open(os.path.join(rs_dir, "planet0.dat")).read().splitlines()
    assert pa[-1] == pb[-1]


def test_restart_missing_set_is_an_io_error(deck, tmp_path):
    from cudaprotoplanetarydisks_amd.native import FargoError
    d = deck(OutputDir=str(tmp_path) + "/")
    with Simulation(d) as s:
        with pytest.raises(FargoError) as e:
            s.restart(7)
        assert e.value.status == 5
