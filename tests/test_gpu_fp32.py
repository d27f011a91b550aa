"""FP32 bulk mode (fused kernels with FP32 arithmetic, FP64 state and
FP64 FARGO residual / CFL / force) against the FP64 path, which is itself
bit-exact against the oracle.  Bounds: density within 1e-5 relative L-inf
of the FP64 run; velocities within 5e-5."""
import math

import numpy as np
import pytest

from cudaprotoplanetarydisks_amd import Simulation
from cudaprotoplanetarydisks_amd.native import FargoError
from cudaprotoplanetarydisks_amd.precision import compare

pytestmark = pytest.mark.gpu

DENS_LINF = 1e-5
VEL_LINF = 5e-5


def test_fp32_tracks_fp64_two_orbits(deck):
    d = deck()
    last = None
    for rep in compare(d, 40, every=10):
        e = rep["errors"]
        assert e["dens"]["linf"] <= DENS_LINF, rep
        assert e["vrad"]["linf"] <= VEL_LINF, rep
        assert e["vtheta"]["linf"] <= VEL_LINF, rep
        last = rep
    assert last["orbits"] == pytest.approx(2.0)
    # the FP32 kernels ran: the runs differ
    assert last["errors"]["dens"]["linf"] > 0.0


def test_fp32_odd_grid_and_corotating(deck):
    d = deck(Nrad=37, Nsec=100, Frame="COROTATING")
    for rep in compare(d, 10, every=5):
        e = rep["errors"]
        assert e["dens"]["linf"] <= DENS_LINF, rep
        assert e["vtheta"]["linf"] <= VEL_LINF, rep


@pytest.mark.parametrize("case", [dict(Nrad=16, Nsec=3200), dict(Nrad=24, Nsec=4096)], ids=["nsec3200", "nsec4096"])
def test_fp32_wide_rings(deck, case):
    # k_fin's ring pass past the overlapped CFL split, and the widest ring
    for rep in compare(deck(**case), 4, every=2):
        e = rep["errors"]
        assert e["dens"]["linf"] <= DENS_LINF, rep
        assert e["vtheta"]["linf"] <= VEL_LINF, rep


def test_fp32_same_fargo_shift_and_cfl(deck):
    d = deck()
    s64 = Simulation(d, fused=True, precision="fp64")
    s32 = Simulation(d, fused=True, precision="fp32")
    try:
        same_shift = 0
        total = 0
        for _ in range(60):
            s64.substep()
            s32.substep()
            assert s64.get("last_cfl") == s32.get("last_cfl")
            a, b = s64.ctx.nshift(), s32.ctx.nshift()
            same_shift += int((a == b).sum())
            total += a.size
        # the integer shift comes from the FP64 ring mean of nearly equal
        # velocity fields: a ring may flip only at a rounding boundary
        assert same_shift >= total - 2
    finally:
        s64.close()
        s32.close()


def test_fp32_sigma_stays_positive_and_finite(deck):
    d = deck(Nrad=64, Nsec=192)
    with Simulation(d, precision="fp32", noise_seed=7, noise_amp=1e-2) as s:
        for _ in range(20):
            s.algogas()
        dens, vr, vt = s.state()
    assert np.isfinite(dens).all() and np.isfinite(vr).all() and np.isfinite(vt).all()
    assert (dens[:-1] > 0).all()


def test_fp32_rejects_decks_without_fused_kernels(deck):
    d = deck(Adiabatic="YES", Cooling="NO")
    with pytest.raises(FargoError) as e:
        Simulation(d, precision="fp32")
    assert e.value.status == 3  # FG_E_UNSUPPORTED


def test_fp32_orbit_count_matches_time(deck):
    d = deck()
    with Simulation(d, precision="fp32") as s:
        for _ in range(20):
            s.algogas()
        assert s.get("PhysicalTime") == pytest.approx(20 * s.get("DT"), rel=1e-12)
        assert s.get("PhysicalTime") / (2 * math.pi) == pytest.approx(1.0, rel=1e-9)


def run_pair(d, pairs, n_dt):
    with Simulation(d, precision="fp32", noise_seed=11, noise_amp=1e-3, fp32_pairs=pairs) as s:
        for _ in range(n_dt):
            s.algogas()
        return s.state()


def test_fp32_pairs_active_near_planet(deck, tmp_path):
    """Sigma stored as float pairs near the planet changes the run (the pair
    path is taken); with a massless planet the band is empty and the
    setting changes nothing."""
    from cudaprotoplanetarydisks_amd.workloads import write_planets
    d = deck(Nrad=64, Nsec=192)
    a, b = run_pair(d, 1.0, 6), run_pair(d, 0.0, 6)
    assert not np.array_equal(a[0], b[0])
    assert np.abs(a[0] - b[0]).max() <= 1e-5 * np.abs(b[0]).max()
    cfg = write_planets(tmp_path / "massless.cfg", [("Z", 1.0, 0.0, 0.0, "NO", "NO")])
    nop = deck(name="nop.par", Nrad=64, Nsec=192, PlanetConfig=cfg)
    a, b = run_pair(nop, 1.0, 4), run_pair(nop, 0.0, 4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("where", ["inner", "outer", "whole"])
def test_fp32_pair_band_at_grid_edges(deck, tmp_path, where):
    """Float-pair bands that reach ring 0, the last ring, or every ring (the
    q = 1e-3 planet at r = 1 with grids that end within the band) against
    the FP64 run, the planet feeling the disk (the results then wait behind
    its force reductions)."""
    from cudaprotoplanetarydisks_amd.workloads import write_planets
    rmin, rmax, hill = {"inner": (0.85, 2.5, 2.5), "outer": (0.4, 1.15, 2.5), "whole": (0.8, 1.2, 3.0)}[where]
    cfg = write_planets(tmp_path / f"{where}.cfg", [("A", 1.0, 1e-3, 0.0, "YES", "NO")])
    d = deck(name=f"{where}.par", Nrad=48, Nsec=160, Rmin=rmin, Rmax=rmax, PlanetConfig=cfg)
    s64 = Simulation(d, fused=True, precision="fp64")
    s32 = Simulation(d, precision="fp32", fp32_pairs=hill)
    try:
        for _ in range(6):
            s64.algogas()
            s32.algogas()
        for x, y in zip(s32.state(), s64.state()):
            x, y = x[:-1], y[:-1]
            assert np.isfinite(x).all()
            assert np.abs(x - y).max() <= 5e-5 * np.abs(y).max()
    finally:
        s64.close()
        s32.close()


def test_fp32_pairs_rejects_bad_width(deck):
    from cudaprotoplanetarydisks_amd.native import lib
    with Simulation(deck(), precision="fp32") as s:
        with pytest.raises(FargoError):
            call_pairs(lib().fg_sim_ctx(s.h), -1.0)


def call_pairs(ctx, w):
    from cudaprotoplanetarydisks_amd.native import call
    call("fg_set_fp32_pairs", ctx, w)


def test_fp32_two_planets_track_fp64(deck, tmp_path):
    """Two planets (the float-pair band spans both planets' rings and the
    rings between) against the FP64 run."""
    from cudaprotoplanetarydisks_amd.workloads import write_planets
    cfg = write_planets(tmp_path / "two.cfg", [("A", 1.0, 1e-3, 0.0, "YES", "NO"), ("B", 1.7, 5e-4, 0.0, "YES", "YES")])
    d = deck(name="two.par", Nrad=64, Nsec=192, PlanetConfig=cfg)
    for rep in compare(d, 8, every=4):
        e = rep["errors"]
        assert e["dens"]["linf"] <= DENS_LINF, rep
        assert e["vtheta"]["linf"] <= VEL_LINF, rep
