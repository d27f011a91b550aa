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
