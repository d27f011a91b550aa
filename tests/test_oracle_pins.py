"""Pins of the FP64 oracle (oracle/fargo_oracle.c) that do not restate its
formulas: known answers from the reference's own decks, closed forms,
invariants, brute force and numpy.  A mistake such as a dropped term, a
wrong sign or a transposed index in the oracle fails one of these."""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from conftest import REFERENCE_IN, write_deck, write_planets
from oracle_binding import Oracle, device_reduce, fft2d


# ----------------------------------------------------------- parser ----
@pytest.mark.skipif(not os.path.isdir(REFERENCE_IN), reason="reference decks not mounted")
def test_parser_known_answers_template2(tmp_path):
    # in/template2.par: Nrad 128, Nsec 384, Rmin 0.4, Rmax 2.5, Sigma0
    # 6.3661977237e-4, Viscosity 1e-5, EVANESCENT, GUIDING-CENTER, adiabatic
    # with cooling, "RadialSpacing A" (arithmetic: only LOGARITHM selects a
    # log grid, Interpret.cu:136)
    cwd = os.getcwd()
    os.chdir(os.path.dirname(REFERENCE_IN))
    try:
        o = Oracle(os.path.join(REFERENCE_IN, "template2.par"))
    finally:
        os.chdir(cwd)
    assert (o.iget("NRAD"), o.iget("NSEC"), o.iget("NTOT"), o.iget("NINTERM")) == (128, 384, 10001, 20)
    assert o.get("RMIN") == 0.4 and o.get("RMAX") == 2.5
    assert o.get("SIGMA0") == 6.3661977237e-4 and o.get("VISCOSITY") == 1e-5
    assert o.get("ASPECTRATIO") == 0.05 and o.get("THICKNESSSMOOTHING") == 0.6
    assert o.iget("Evanescent") == 1 and o.iget("GuidingCenter") == 1 and o.iget("Corotating") == 1
    assert o.iget("Adiabatic") == 1 and o.iget("Cooling") == 1 and o.iget("LogGrid") == 0
    assert o.iget("Write_Qplus") == 1 and o.iget("Write_Temperature") == 1
    assert o.get("COOLINGTIME0") == 62.8 and o.get("DT") == 0.314159265359
    assert o.planets()[0].tolist()[4] == 0.001  # in/Jup.cfg


@pytest.mark.skipif(not os.path.isdir(REFERENCE_IN), reason="reference decks not mounted")
def test_parser_known_answers_adsg():
    cwd = os.getcwd()
    os.chdir(os.path.dirname(REFERENCE_IN))
    try:
        o = Oracle(os.path.join(REFERENCE_IN, "adsg.par"))
    finally:
        os.chdir(cwd)
    assert o.iget("ForcedCircular") == 1 and o.iget("Corotating") == 1 and o.iget("GuidingCenter") == 0
    assert o.get("SIGMASLOPE") == 0.5 and o.get("RMAX") == 1.8 and o.get("SIGMA0") == 2e-3
    assert o.iget("NTOT") == 200 and o.iget("SelfGravity") == 0
    assert o.planets()[0][4] == 2.15786376476132e-05


def test_parser_case_insensitive_keys_and_separators(tmp_path):
    """Interpret.cu:114-126: the key is the first whitespace-delimited
    token (upper-cased), then any of "\\t :=>_" is skipped before the value.
    A key glued to its separator ("NSEC:16") is therefore not recognised."""
    path = tmp_path / "d.par"
    path.write_text("nrad = 12\nNSEC :16\n# comment\nrMIN\t0.5\nRMAX =>2.0\nThicknessSmoothing 0.6\n"
                    "Sigma0:7\n"
                    f"PlanetConfig {write_planets(tmp_path / 'p.cfg', [('P', 1.0, 1e-4, 0.0, 'no', 'no')])}\n"
                    "RadialSpacing LOGARITHM\nInnerBoundary EVANESCENT\n")
    o = Oracle(str(path))
    assert (o.iget("NRAD"), o.iget("NSEC"), o.get("RMIN"), o.get("RMAX")) == (12, 16, 0.5, 2.0)
    assert o.iget("LogGrid") == 1
    assert o.get("SIGMA0") == 173.0  # Param.cuh default: "Sigma0:7" is one unknown token


# ------------------------------------------------------------- grid ----
@pytest.mark.parametrize("spacing", ["LOGARITHM", "A"])
def test_grid_closed_forms(deck, spacing):
    o = Oracle(deck(Nrad=50, Nsec=60, RadialSpacing=spacing))
    nr, ns = 50, 60
    r = o.field("Radii")
    if spacing == "LOGARITHM":
        expect = 0.4 * (2.5 / 0.4) ** (np.arange(nr + 1) / nr)
    else:
        expect = 0.4 + (2.5 - 0.4) / nr * np.arange(nr + 1)
    assert np.allclose(r, expect, rtol=1e-14, atol=0)
    surf = o.field("Surf")[:nr]
    area = float(np.sum(surf) * ns)
    assert abs(area - math.pi * (2.5 ** 2 - 0.4 ** 2)) < 1e-12
    rmed = o.field("Rmed")[:nr]
    assert np.all((rmed > r[:-1]) & (rmed < r[1:]))
    # Rmed is the area-weighted centre: 2/3 (r+^3 - r-^3)/(r+^2 - r-^2)
    assert np.allclose(rmed, 2 / 3 * (r[1:] ** 3 - r[:-1] ** 3) / (r[1:] ** 2 - r[:-1] ** 2), rtol=1e-14)
    assert np.allclose(o.field("invdiffRmed")[1:nr], 1 / np.diff(rmed), rtol=1e-13)


# --------------------------------------------------------- initial state --
def test_initial_state_closed_forms(deck):
    o = Oracle(deck(Nrad=40, Nsec=32, SigmaSlope=0.5, FlaringIndex=0.1, Frame="FIXED", OmegaFrame=0.25))
    rmed = o.field("Rmed")[:40]
    rinf = o.field("Rinf")[:40]
    h, s, f, nu, S0, OF = 0.05, 0.5, 0.1, 1e-5, 6.3662e-4, 0.25
    dens = o.field("Dens")[:40]
    assert np.allclose(dens, (S0 * rmed ** -s)[:, None], rtol=1e-14)
    vt = o.field("Vtheta")[:40]
    vt_expect = np.sqrt(1 / rmed) * np.sqrt(1 - h * h * rmed ** (2 * f) * (1 + s - 2 * f)) - OF * rmed
    assert np.allclose(vt, vt_expect[:, None], rtol=1e-13)
    vr = o.field("Vrad")
    vr_expect = -3 * nu / rmed * (0.5 - s)  # constant-nu drift, Kernels.cu:682
    assert np.all(vr[0] == 0) and np.all(vr[40] == 0)
    assert np.allclose(vr[1:40], vr_expect[1:, None], rtol=1e-13)
    cs = o.field("SoundSpeed")[:40]
    assert np.allclose(cs, (h * rmed ** f * np.sqrt(1 / rmed))[:, None], rtol=1e-14)


# ---------------------------------------------------- DeviceReduce ----
def exact_sum(a):
    return float(sum(Fraction(x) for x in a))


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 64, 100, 1000, 4096, 49152, 70001])
def test_device_reduce_close_to_exact_sum(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) * np.exp(rng.uniform(-5, 5, n))
    s = device_reduce(a)
    assert abs(s - exact_sum(a)) <= 4 * n * np.finfo(float).eps * np.abs(a).sum()


def test_device_reduce_order_hand_worked():
    """n = 100: pow2(100)/32 = 4 blocks of 32 threads, grid stride 256.
    Thread t of block b adds a[64b+t] then a[64b+t+32] (when < n); the 32
    partials meet in the shuffle tree (lane 0 first adds lane 16); the
    host adds the block sums in order.  Worked by hand:
      block 0: s0 = fl(2^53 + a[32]=1) = 2^53, s16 = a[16] + a[48] = 2,
               tree: s0 + s16 = 2^53 + 2 (exact) -> block sum 2^53 + 2
      block 1: s0 = a[64] = -2^53 -> block sum -2^53
      host:    0 + (2^53 + 2) - 2^53 = 2
    A left-to-right sum gives 0 and the exact sum is 3, so a wrong pairing
    or order changes the result."""
    a = np.zeros(100)
    a[0], a[16], a[32], a[48], a[64] = 2.0 ** 53, 1.0, 1.0, 1.0, -(2.0 ** 53)
    assert device_reduce(a) == 2.0
    seq = 0.0
    for x in a:
        seq += x
    assert seq == 0.0 and math.fsum(a) == 3.0


# ------------------------------------------------------------- CFL ----
def test_cfl_brute_force(deck):
    o = Oracle(deck(Nrad=30, Nsec=40))
    rng = np.random.default_rng(3)
    vr, vt = o.view("Vrad"), o.view("Vtheta")
    vr[1:30] += rng.normal(0, 3e-3, (29, 40))
    vt[:30] += rng.normal(0, 3e-3, (30, 40))
    DeltaT = 0.314159265359
    n = o.condition_cfl(DeltaT)
    # independent evaluation of ConditionCFL (SourceEuler.cu:520-546)
    rinf, rsup, rmed = o.field("Rinf"), o.field("Rsup"), o.field("Rmed")
    cs = o.field("SoundSpeed")
    vmoy = vt[:30].mean(axis=1)
    best = DeltaT
    for i in range(1, 30):
        dxr = rsup[i] - rinf[i]
        dxt = rmed[i] * 2 * math.pi / 40
        res = vt[i] - vmoy[i]
        dvr = vr[i + 1] - vr[i]
        dvt = np.roll(vt[i], -1) - vt[i]
        dvr = np.where(dvr >= 0, 1e-10, -dvr)
        dvt = np.where(dvt >= 0, 1e-10, -dvt)
        i4 = np.maximum(dvr / dxr, dvt / dxt) * 4 * 1.41 ** 2
        dt = 0.5 / np.sqrt((cs[i] / min(dxr, dxt)) ** 2 + (np.abs(vr[i]) / dxr) ** 2 + (np.abs(res) / dxt) ** 2 + i4 ** 2)
        best = min(best, dt.min())
    for i in range(29):
        best = min(best, 2 * math.pi * 0.5 / 40 / abs(vmoy[i] / rmed[i] - vmoy[i + 1] / rmed[i + 1]))
    assert n == math.ceil(DeltaT / best)


# ------------------------------------------------------------ Substep1 --
def test_substep1_uniform_pressure_no_potential(deck):
    """Uniform Sigma and P, zero potential: VradInt = Vrad + dt vt^2/Rinf
    and VthetaInt = Vtheta (no pressure or potential gradient)."""
    o = Oracle(deck(Nrad=20, Nsec=24, Viscosity=0.0, AlphaViscosity=None))
    o.view("Dens")[:] = 1e-3
    o.view("Pressure")[:] = 2.5e-6
    o.view("Potential")[:] = 0.0
    vt = o.field("Vtheta")
    vr = o.field("Vrad")
    dt = 0.01
    o.L.or_substep1(o.h, dt)
    rinf = o.field("Rinf")
    OF = o.get("OmegaFrame")
    avg = (vt[1:20] + np.roll(vt[1:20], -1, axis=1) + vt[:19] + np.roll(vt[:19], -1, axis=1)) / 4
    expect = vr[1:20] + dt * (avg + OF * rinf[1:20, None]) ** 2 / rinf[1:20, None]
    assert np.allclose(o.field("VradInt")[1:20], expect, rtol=1e-13, atol=1e-17)
    assert np.array_equal(o.field("VthetaInt")[:20], vt[:20])


# ----------------------------------------------------------- viscosity --
def test_solid_body_rotation_has_no_shear(deck):
    o = Oracle(deck(Nrad=24, Nsec=32))
    rmed = o.field("Rmed")
    o.view("VradInt")[:] = 0.0
    o.view("VthetaInt")[:] = 0.7 * rmed[:, None]
    o.view("Dens")[:] = 1e-3
    o.viscous_terms(1)
    for name in ("TAURP", "DivergenceVelocity", "TAURR", "TAUPP"):
        a = o.field(name)[1:24]
        assert np.abs(a).max() < 1e-16, name


# ------------------------------------------------ artificial viscosity --
def test_expanding_flow_has_no_artificial_pressure(deck):
    o = Oracle(deck(Nrad=16, Nsec=20))
    vri = o.view("VradInt")
    vri[:] = np.linspace(0, 1e-3, 17)[:, None]
    o.view("VthetaInt")[:] = 0.3
    before = o.field("VradInt"), o.field("VthetaInt")
    o.substep2(0.01)
    assert not o.field("DensInt")[:16].any() and not o.field("TemperInt")[:16].any()
    o.host_kernel(0.01)
    assert np.array_equal(o.field("VradNew")[1:16], before[0][1:16])
    assert np.array_equal(o.field("VthetaNew")[:16], before[1][:16])


# ----------------------------------------------------------- transport --
def test_transport_conserves_mass(deck):
    o = Oracle(deck(Nrad=40, Nsec=64))
    rng = np.random.default_rng(11)
    d = o.view("Dens")
    d[:40] *= 1 + 0.2 * rng.random((40, 64))
    o.view("Vrad")[1:40] += rng.normal(0, 2e-3, (39, 64))
    o.view("Vtheta")[:40] += rng.normal(0, 2e-3, (40, 64))
    surf = o.field("Surf")[:40, None]
    m0 = float((o.field("Dens")[:40] * surf).sum())
    o.transport(0.02)
    m1 = float((o.field("Dens")[:40] * surf).sum())
    assert abs(m1 - m0) / m0 < 1e-13


def test_integer_courant_azimuthal_advection_is_a_shift(deck):
    """vtheta = k cells per dt and vrad = 0: the transport is the exact
    orbital shift by k sectors (FARGO, Kernels.cu:1111-1114)."""
    ns, nr, k, dt = 48, 10, 3, 0.05
    o = Oracle(deck(Nrad=nr, Nsec=ns, Frame="FIXED", OmegaFrame=0.0))
    rmed = o.field("Rmed")
    o.view("Vrad")[:] = 0.0
    o.view("Vtheta")[:nr] = (k * 2 * math.pi * rmed[:nr] / ns / dt)[:, None]
    rng = np.random.default_rng(5)
    d = o.view("Dens")
    d[:nr] = 1e-3 * (1 + 0.5 * rng.random((nr, ns)))
    before = o.field("Dens")[:nr]
    o.transport(dt)
    assert np.array_equal(o.ifield("Nshift")[:nr], [k] * nr)
    assert np.allclose(o.field("Dens")[:nr], np.roll(before, k, axis=1), rtol=1e-12)


def test_van_leer_azimuthal_step_is_tvd(deck):
    ns, nr = 64, 8
    o = Oracle(deck(Nrad=nr, Nsec=ns, Frame="FIXED", OmegaFrame=0.0))
    rmed = o.field("Rmed")
    o.view("Vrad")[:] = 0.0
    # sub-cell azimuthal velocity (0.3 cells per step) on a square wave
    o.view("Vtheta")[:nr] = (0.3 * 2 * math.pi * rmed[:nr] / ns / 0.05)[:, None]
    d = o.view("Dens")
    d[:nr] = 1e-3
    d[:nr, 10:25] = 2e-3
    o.transport(0.05)
    after = o.field("Dens")[:nr]
    assert after.min() >= 1e-3 * (1 - 1e-12) and after.max() <= 2e-3 * (1 + 1e-12)


# ----------------------------------------------------------- boundaries --
def test_evanescent_damping_closed_form(deck):
    o = Oracle(deck(Nrad=30, Nsec=16))
    rng = np.random.default_rng(2)
    d = o.view("Dens")
    d[:30] *= 1 + 0.1 * rng.random((30, 16))
    before = o.field("Dens")
    dt = 0.03
    o.apply_boundary(dt)
    after = o.field("Dens")
    rmed = o.field("Rmed")[:30]
    smed = o.field("SigmaMed")[:30]
    drmin, drmax = rmed[0] * 1.25, rmed[29] * 0.84
    tin, tout = 2 * math.pi * rmed[0] ** 1.5, 2 * math.pi * rmed[29] ** 1.5
    for i in range(30):
        if rmed[i] < drmin:
            lam = ((rmed[i] - drmin) / (rmed[0] - drmin)) ** 2 * 10 * dt / tin
        elif rmed[i] > drmax:
            lam = ((rmed[i] - drmax) / (rmed[29] - drmax)) ** 2 * 10 * dt / tout
        else:
            assert np.array_equal(after[i], before[i])  # untouched rings bit-identical
            continue
        assert np.allclose(after[i], (before[i] + lam * smed[i]) / (1 + lam), rtol=1e-14)


# ---------------------------------------------------------------- RK5 --
def test_cash_karp_keeps_a_circular_orbit(tmp_path):
    """Gas off: the planet (q = 1e-3, a = 1) moves on a Kepler circle with
    GM = 1 + q (star + indirect term); the 5th-order step holds the radius
    to 1e-9 and returns to its start after one period."""
    cfg = write_planets(tmp_path / "p.cfg", [("P", 1.0, 1e-3, 0.0, "NO", "NO")])
    nsteps = 400
    period = 2 * math.pi / math.sqrt(1.001)
    path = write_deck(tmp_path / "d.par", Disk="NO", Frame="FIXED", OmegaFrame=0.0, PlanetConfig=cfg,
                      DT=repr(period / nsteps), Nrad=8, Nsec=8, OutputDir=str(tmp_path) + "/")
    o = Oracle(path)
    for k in range(nsteps):
        o.algogas()
        x, y = o.planets()[0][:2]
        assert abs(math.hypot(x, y) - 1.0) < 1e-9
    x, y, vx, vy, _ = o.planets()[0]
    assert abs(x - 1.0) < 1e-6 and abs(y) < 1e-6


# ---------------------------------------------------------------- FFT --
@pytest.mark.parametrize("shape", [(8, 12), (6, 10), (16, 48), (4, 9), (2, 30)])
def test_fft2d_matches_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.allclose(fft2d(z, -1), np.fft.fft2(z), rtol=1e-12, atol=1e-12)
    assert np.allclose(fft2d(z, +1), np.fft.ifft2(z) * z.size, rtol=1e-12, atol=1e-12)


def test_sg_convolution_matches_brute_force(tmp_path):
    """The zero-padded FFT product of SgMain.cu equals the direct circular
    convolution over the 2 NRAD x NSEC grid (small grid)."""
    nr, ns = 4, 6
    rng = np.random.default_rng(9)
    K = rng.standard_normal((2 * nr, ns))
    S = np.zeros((2 * nr, ns))
    S[:nr] = rng.random((nr, ns))
    via_fft = fft2d(fft2d(K.astype(complex), -1) * fft2d(S.astype(complex), -1), +1).real / (2 * nr * ns)
    direct = np.zeros_like(K)
    for a in range(2 * nr):
        for b in range(ns):
            direct[a, b] = sum(K[(a - p) % (2 * nr), (b - q) % ns] * S[p, q] for p in range(2 * nr) for q in range(ns))
    assert np.allclose(via_fft, direct, rtol=1e-11, atol=1e-12)
