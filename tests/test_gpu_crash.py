"""Crash detection and state changes in the middle of a batch.

A state with a negative or non-finite density must stop the run with
FG_E_CRASH instead of stepping on: the check of DetectCrash/CrashKernel
(SourceEuler.cu:453-465, Kernels.cu:953-964; the reference's call is
commented out at SourceEuler.cu:266-273), plus NaN, on every path.  A state
set into a batch disk between batch substeps must be stepped exactly like
the same state set into a single-disk run."""
import ctypes

import numpy as np
import pytest

from cudaprotoplanetarydisks_amd import Simulation
from cudaprotoplanetarydisks_amd.ensemble import Ensemble, write_disk_deck
from cudaprotoplanetarydisks_amd.native import FG_E_CRASH, FargoError, call

pytestmark = pytest.mark.gpu

_pd = ctypes.POINTER(ctypes.c_double)
PATHS = [dict(fused=False), dict(fused=True), dict(precision="fp32")]
PATH_IDS = ["per-routine-fp64", "fused-fp64", "fused-fp32"]


@pytest.mark.parametrize("bad", [np.nan, -1e-6, np.inf], ids=["nan", "negative", "inf"])
@pytest.mark.parametrize("opts", PATHS, ids=PATH_IDS)
def test_bad_density_returns_crash(deck, opts, bad):
    path = deck(Nrad=24, Nsec=48)
    with Simulation(path, **opts) as sim:
        sim.substep()
        dens, vr, vt = sim.state()
        good = dens.copy()
        dens[7, 11] = bad
        sim.set_state(dens, vr, vt)
        print(f"crash case: {PATH_IDS[PATHS.index(opts)]} cell (7, 11) = {bad}", flush=True)
        with pytest.raises(FargoError) as e:
            sim.substep()
        assert e.value.status == FG_E_CRASH
        # a good state runs again: the flag does not outlive the state
        sim.set_state(good, vr, vt)
        for _ in range(3):
            sim.substep()
        d2, _, _ = sim.state()
        assert np.isfinite(d2).all() and (d2[:-1] > 0).all()


@pytest.mark.parametrize("fused", [False, True], ids=["per-routine", "fused"])
def test_bad_energy_returns_crash(deck, fused):
    path = deck(Nrad=24, Nsec=48, Adiabatic="YES", AdiabaticIndex=1.4)
    with Simulation(path, fused=fused) as sim:
        sim.substep()
        d, vr, vt, e = sim.state(energy=True)
        good = e.copy()
        e[5, 7] = -1e-9
        sim.set_state(d, vr, vt, e)
        print(f"crash case: energy, fused={fused}, cell (5, 7) = -1e-9", flush=True)
        with pytest.raises(FargoError) as err:
            sim.substep()
        assert err.value.status == FG_E_CRASH
        sim.set_state(d, vr, vt, good)
        for _ in range(3):
            sim.substep()
        _, _, _, e2 = sim.state(energy=True)
        assert np.isfinite(e2).all() and (e2[:-1] > 0).all()


def test_crash_in_a_batch_disk(tmp_path):
    params = [dict(nrad=24, nsec=72, q=1e-3, alpha=1e-3, h=0.05, slope=0.0, flaring=0.0),
              dict(nrad=32, nsec=96, q=2e-3, alpha=2e-4, h=0.07, slope=-0.5, flaring=0.1)]
    decks = [write_disk_deck(str(tmp_path), k, p) for k, p in enumerate(params)]
    with Ensemble(decks, tiles={}) as ens:
        ens.substep(10.0)
        dens, vr, vt = ens.state(1)
        dens[5, 3] = -1.0
        call("fg_sim_set_state", ens.sim_handle(1), *[a.ctypes.data_as(_pd) for a in (dens, vr, vt)], None)
        with pytest.raises(FargoError) as e:
            ens.substep(10.0)
        assert e.value.status == FG_E_CRASH
        assert "disk 1" in str(e.value)
        # the crashed disk is left alone, the other one goes on
        assert ens.get(1, "crashed") == 1.0 and ens.get(0, "crashed") == 0.0
        t0 = ens.get(0, "PhysicalTime")
        assert ens.substep(10.0) == 1
        assert ens.get(0, "PhysicalTime") > t0
        # a good state brings it back
        dens[5, 3] = 1e-3
        call("fg_sim_set_state", ens.sim_handle(1), *[a.ctypes.data_as(_pd) for a in (dens, vr, vt)], None)
        assert ens.get(1, "crashed") == 0.0
        assert ens.substep(10.0) == 2


def test_state_set_between_batch_substeps_is_stepped_like_a_single_run(tmp_path):
    """fg_sim_set_state on a batch disk: the next batch substep takes its dt
    (CFL count) and indirect term from the new state, bit for bit as a
    single-disk FP32 run given the same state."""
    params = [dict(nrad=24, nsec=72, q=1e-3, alpha=1e-3, h=0.05, slope=0.0, flaring=0.0),
              dict(nrad=40, nsec=120, q=3e-4, alpha=3e-3, h=0.04, slope=-1.0, flaring=0.0)]
    decks = [write_disk_deck(str(tmp_path), k, p) for k, p in enumerate(params)]
    with Simulation(decks[0], precision="fp32") as fresh:
        init = fresh.state()
    # a state that differs from where the disk is: the initial one, perturbed
    rng = np.random.default_rng(4)
    new = [a.copy() for a in init]
    new[0][:-1] *= 1 + 0.05 * rng.random(new[0][:-1].shape)
    with Ensemble(decks, tiles={}) as ens, Simulation(decks[0], precision="fp32") as single:
        for _ in range(3):
            ens.substep(100.0)
            single.substep()
        call("fg_sim_set_state", ens.sim_handle(0), *[a.ctypes.data_as(_pd) for a in new], None)
        single.set_state(*new)
        for _ in range(4):
            ens.substep(100.0)
            single.substep()
        assert ens.get(0, "last_cfl") == single.get("last_cfl")
        assert ens.get(0, "PhysicalTime") == single.get("PhysicalTime")
        for a, b, name in zip(ens.state(0), single.state(), ("dens", "vrad", "vtheta")):
            assert np.array_equal(a, b), f"{name}: {np.abs(a - b).max():.3e}"
