"""Pins of the FP64 oracle's physics routines by what the mathematics fixes.

Each test sets up a state whose answer is known without the oracle's
formula: a closed form derived from the physics (the softened point-mass
potential, a pressure or potential gradient of a linear profile, the
viscous stress and its divergence for a power-law shear flow, the von
Neumann-Richtmyer pressure of a uniform compression, backward-Euler energy
relaxation, the equilibrium rotation of a pressure-supported disk), an
invariant (a divergence-free radial flow keeps a uniform density uniform,
a planet on a circular orbit stays put in the co-rotating frame), an
independent brute-force implementation (one column of radial van Leer
transport, the softened log-polar self-gravity as the finite-difference
gradient of a direct potential sum), or the reference's own rounding rule
stated in a different form.  A dropped term, a wrong sign, a wrong factor
or a shifted index in the oracle fails one of them.  Reference lines:
Kernels.cu:13-150, 409-479, 482-607, 893-932, 1193-1210, 1226-1253;
SideEuler.cu:45-213; Viscosity.cu:33-50; Psys.cu:174-265.
"""
import math

import numpy as np
import pytest

from conftest import write_deck, write_planets
from oracle_binding import Oracle, _pd

CVNR = 1.41  # Fondam.cuh:10


def planet_deck(tmp_path, name="d.par", feeldisk="NO", dist=1.0, q=1e-3, **kw):
    cfg = write_planets(tmp_path / (name + ".cfg"), [("P", dist, q, 0.0, feeldisk, "NO")])
    kw.setdefault("OutputDir", str(tmp_path) + "/")
    return write_deck(tmp_path / name, PlanetConfig=cfg, **kw)


def set_planet(o, k, x, y, vx, vy, m):
    p = np.array([x, y, vx, vy, m], dtype=np.float64)
    o.L.or_set_planet(o.h, k, p.ctypes.data_as(_pd))


# ------------------------------------------------------------ potential --
@pytest.mark.parametrize("roche", [False, True], ids=["thickness-smoothing", "roche-smoothing"])
def test_potential_softened_point_mass_and_indirect_terms(tmp_path, roche):
    """FillForcesArrays (Pframeforce.cu:56-83, Kernels.cu:1226-1253): the
    potential of the star (-GM/r), of the planet (softened point mass
    -G m / sqrt(d^2 + eps^2), eps = 0.6 H(r_p) or the Roche smoothing) and
    the two indirect terms (the planet's pull on the star, +G m r.r_p/r_p^3,
    and the disk's, -a_disk.r).  Planet placed exactly on a cell centre so
    that d = 0 there."""
    kw = dict(Nrad=20, Nsec=32, Frame="FIXED", OmegaFrame=0.0, AspectRatio=0.05, FlaringIndex=0.2)
    if roche:
        kw.update(ThicknessSmoothing=None, RocheSmoothing=0.5)
    o = Oracle(planet_deck(tmp_path, **kw))
    nr, ns = 20, 32
    rmed = o.field("Rmed")[:nr]
    i0, j0 = 11, 5
    phi0 = j0 / ns * 2.0 * math.pi  # the potential's own angle rule (Kernels.cu:1236)
    xp, yp = rmed[i0] * math.cos(phi0), rmed[i0] * math.sin(phi0)
    m, taper = 2e-3, 0.7
    set_planet(o, 0, xp, yp, 0.0, 1.0, m)
    o.set("MassTaper", taper)
    ax, ay = 3e-4, -2e-4  # disk acceleration of the star
    o.set("DiskOnPrimaryAcceleration.x", ax)
    o.set("DiskOnPrimaryAcceleration.y", ay)
    o.fill_forces_arrays()
    pot = o.field("Potential")[:nr]
    rp = rmed[i0]
    mt = m * taper
    eps = (rp * (mt / 3.0) ** (1.0 / 3.0) * 0.5) if roche else 0.6 * 0.05 * rp ** 0.2 * rp
    # the planet's own cell: d = 0
    x0, y0 = xp, yp
    expect0 = -mt / eps + mt * (x0 * xp + y0 * yp) / rp ** 3 - 1.0 / rp + (ax * x0 + ay * y0)
    assert pot[i0, j0] == pytest.approx(expect0, rel=1e-13)
    # every cell: the closed form with numpy positions
    phi = np.arange(ns) / ns * 2.0 * np.pi
    X = rmed[:, None] * np.cos(phi)[None, :]
    Y = rmed[:, None] * np.sin(phi)[None, :]
    d2 = (X - xp) ** 2 + (Y - yp) ** 2
    expect = -mt / np.sqrt(d2 + eps ** 2) + mt * (X * xp + Y * yp) / rp ** 3 - 1.0 / rmed[:, None] + (ax * X + ay * Y)
    assert np.allclose(pot, expect, rtol=1e-12, atol=0)
    # the planet term is the minimum and the far side the least negative
    planet_part = pot - (-1.0 / rmed[:, None] + ax * X + ay * Y + mt * (X * xp + Y * yp) / rp ** 3)
    assert np.unravel_index(np.argmin(planet_part), planet_part.shape) == (i0, j0)


def test_potential_star_only_without_planet_mass(tmp_path):
    """MassTaper = 0 (the first substep, Q1) and no disk acceleration: the
    potential is the star's -GM/r exactly, independent of the sector."""
    o = Oracle(planet_deck(tmp_path, Nrad=12, Nsec=16, Frame="FIXED", OmegaFrame=0.0))
    o.set("MassTaper", 0.0)
    o.set("DiskOnPrimaryAcceleration.x", 0.0)
    o.set("DiskOnPrimaryAcceleration.y", 0.0)
    o.fill_forces_arrays()
    rmed = o.field("Rmed")[:12]
    pot = o.field("Potential")[:12]
    assert np.array_equal(pot, np.broadcast_to((-1.0 / rmed)[:, None], pot.shape))


def test_indirect_term_off_drops_both_indirect_parts(tmp_path):
    o = Oracle(planet_deck(tmp_path, Nrad=12, Nsec=16, Frame="FIXED", OmegaFrame=0.0, IndirectTerm="NO"))
    rmed = o.field("Rmed")[:12]
    set_planet(o, 0, 1.3, 0.4, 0.0, 1.0, 1e-3)
    o.set("MassTaper", 1.0)
    o.set("DiskOnPrimaryAcceleration.x", 1e-3)
    o.set("DiskOnPrimaryAcceleration.y", 1e-3)
    o.fill_forces_arrays()
    phi = np.arange(16) / 16 * 2.0 * np.pi
    X, Y = rmed[:, None] * np.cos(phi), rmed[:, None] * np.sin(phi)
    eps = 0.6 * 0.05 * math.hypot(1.3, 0.4)
    expect = -1e-3 / np.sqrt((X - 1.3) ** 2 + (Y - 0.4) ** 2 + eps ** 2) - 1.0 / rmed[:, None]
    assert np.allclose(o.field("Potential")[:12], expect, rtol=1e-13, atol=0)


# ------------------------------------------------------------- Substep1 --
def substep1_deck(deck, **kw):
    return deck(Nrad=24, Nsec=32, Viscosity=0.0, Frame="FIXED", OmegaFrame=0.0, **kw)


def test_substep1_linear_pressure_gives_its_gradient(deck):
    """P = a + b r, uniform Sigma = rho, no potential, no rotation:
    dv_r/dt = -(1/rho) dP/dr = -b/rho exactly on any grid; azimuthal
    pressure sawtooth P_j = c j: dv_theta/dt = -c/(rho r dtheta) for j >= 1."""
    o = Oracle(substep1_deck(deck))
    nr, ns, rho, dt = 24, 32, 2e-3, 0.01
    rmed = o.field("Rmed")
    o.view("Dens")[:] = rho
    o.view("Vtheta")[:] = 0.0
    vr = np.random.default_rng(1).normal(0, 1e-4, (nr + 1, ns))
    o.view("Vrad")[:] = vr
    o.view("Potential")[:] = 0.0
    a, b = 3e-6, -1.5e-6
    o.view("Pressure")[:] = (a + b * rmed)[:, None]
    o.substep1(dt)
    assert np.allclose(o.field("VradInt")[1:nr], vr[1:nr] - dt * b / rho, rtol=1e-12, atol=1e-18)
    assert np.array_equal(o.field("VthetaInt")[:nr], np.zeros((nr, ns)))
    # azimuthal sawtooth
    o2 = Oracle(substep1_deck(deck))
    o2.view("Dens")[:] = rho
    o2.view("Vtheta")[:] = 0.0
    o2.view("Vrad")[:] = 0.0
    o2.view("Potential")[:] = 0.0
    c = 2e-7
    o2.view("Pressure")[:] = 1e-5 + c * np.arange(ns)[None, :]
    o2.substep1(dt)
    vti = o2.field("VthetaInt")[:nr]
    dxt = 2 * math.pi / ns * rmed[:nr]
    assert np.allclose(vti[:, 1:], (-dt * c / (rho * dxt))[:, None], rtol=1e-12, atol=0)
    assert np.allclose(vti[:, 0], dt * c * (ns - 1) / (rho * dxt), rtol=1e-12, atol=0)


def test_substep1_linear_potential_gives_its_gradient(deck):
    """Phi = k r and Phi_j = k' j with uniform P: the kick is -grad Phi."""
    o = Oracle(substep1_deck(deck))
    nr, ns, dt = 24, 32, 0.02
    rmed = o.field("Rmed")
    o.view("Dens")[:] = 1e-3
    o.view("Pressure")[:] = 2e-6
    o.view("Vtheta")[:] = 0.0
    o.view("Vrad")[:] = 0.0
    k = 0.37
    o.view("Potential")[:] = (k * rmed)[:, None]
    o.substep1(dt)
    assert np.allclose(o.field("VradInt")[1:nr], -dt * k, rtol=1e-12, atol=0)
    o.view("Potential")[:] = 0.01 * np.arange(ns)[None, :]
    o.substep1(dt)
    dxt = 2 * math.pi / ns * rmed[:nr]
    assert np.allclose(o.field("VthetaInt")[:nr, 1:], (-dt * 0.01 / dxt)[:, None], rtol=1e-12, atol=0)
    assert np.array_equal(o.field("VradInt")[1:nr], np.zeros((nr - 1, ns)))


def test_substep1_centrifugal_term_in_rotating_frame_and_imposed_drift(deck):
    """Uniform v_theta = w in a frame rotating at Omega_F: the radial kick is
    (w + Omega_F r)^2 / r at the inner face (the inertial azimuthal speed);
    ImposedDiskDrift D adds dt D/2 r^(-2.5+s) to v_theta (Kernels.cu:40)."""
    o = Oracle(deck(Nrad=24, Nsec=32, Viscosity=0.0, Frame="FIXED", SigmaSlope=0.5, ImposedDiskDrift=1e-4))
    nr, ns, dt, w = 24, 32, 0.01, 0.15
    OF = 0.8
    o.set("OmegaFrame", OF)
    rinf, rmed = o.field("Rinf"), o.field("Rmed")
    o.view("Dens")[:] = 1e-3
    o.view("Pressure")[:] = 1e-6
    o.view("Potential")[:] = 0.0
    o.view("Vrad")[:] = 0.0
    o.view("Vtheta")[:] = w
    o.substep1(dt)
    assert np.allclose(o.field("VradInt")[1:nr], (dt * (w + OF * rinf[1:nr]) ** 2 / rinf[1:nr])[:, None],
                       rtol=1e-13, atol=0)
    assert np.allclose(o.field("VthetaInt")[:nr], (w + dt * 0.5e-4 * rmed[:nr] ** (-2.5 + 0.5))[:, None],
                       rtol=1e-14, atol=0)


# ----------------------------------------------------- viscous stress --
def viscous_deck(deck, nr=256, ns=8, nu=1e-5):
    return deck(Nrad=nr, Nsec=ns, Viscosity=nu, Frame="FIXED", OmegaFrame=0.0, Rmin=0.5, Rmax=2.0)


@pytest.mark.parametrize("q", [-0.5, 1.0, 2.0], ids=["keplerian", "solid", "q2"])
def test_shear_flow_stress_and_its_divergence(deck, q):
    """v_theta = A r^q, v_r = 0, uniform rho and nu: the continuum stress is
    tau_rphi = rho nu r dOmega/dr = rho nu A (q - 1) r^(q-1), tau_rr =
    tau_phiphi = div v = 0; the azimuthal viscous acceleration is
    (1/(rho r^2)) d(r^2 tau_rphi)/dr = nu A (q^2 - 1) r^(q-2).  Second-order
    discretisation: within 1e-3 at 256 log rings over [0.5, 2]."""
    nr, ns, nu, rho, A, dt = 256, 8, 1e-5, 2e-3, 0.9, 1e-3
    o = Oracle(viscous_deck(deck, nr, ns, nu))
    rinf, rmed = o.field("Rinf"), o.field("Rmed")
    o.view("Dens")[:] = rho
    o.view("VradInt")[:] = 0.0
    o.view("VthetaInt")[:nr] = (A * rmed[:nr] ** q)[:, None]  # ghost ring stays 0 (Rmed[NRAD] = 0, Q8)
    o.viscous_terms(1)
    trp = o.field("TAURP")[1:nr, 0]
    expect = rho * nu * A * (q - 1) * rinf[1:nr] ** (q - 1)
    if q == 1.0:
        assert np.abs(trp).max() < 1e-20
    else:
        assert np.allclose(trp, expect, rtol=1e-3)
    for name in ("TAURR", "TAUPP", "DivergenceVelocity"):
        assert np.abs(o.field(name)[:nr]).max() < 1e-18, name
    # the update, through Substep1 with uniform P and no potential
    o = Oracle(viscous_deck(deck, nr, ns, nu))
    o.view("Dens")[:] = rho
    o.view("Pressure")[:] = 1e-6
    o.view("Potential")[:] = 0.0
    o.view("Vrad")[:] = 0.0
    vt0 = np.zeros((nr + 1, ns))
    vt0[:nr] = (A * rmed[:nr] ** q)[:, None]
    o.view("Vtheta")[:] = vt0
    o.substep1(dt)
    acc = (o.field("VthetaInt") - vt0)[2:nr - 2, 0] / dt
    expect = nu * A * (q * q - 1) * rmed[2:nr - 2] ** (q - 2)
    if q == 1.0:
        assert np.abs(acc).max() < 1e-15
    else:
        assert np.allclose(acc, expect, rtol=2e-3)


def test_radial_compression_stress_and_its_divergence(deck):
    """v_r = v0 r^2, v_theta = 0: d v_r/dr = 2 v0 r, v_r/r = v0 r, div v =
    3 v0 r, so tau_rr = 2 rho nu (2 v0 r - v0 r) = 2 rho nu v0 r and
    tau_phiphi = 0; the radial viscous acceleration
    (1/rho) [(1/r) d(r tau_rr)/dr - tau_phiphi/r] = 4 nu v0."""
    nr, ns, nu, rho, v0, dt = 256, 8, 1e-5, 2e-3, 1e-3, 1e-3
    o = Oracle(viscous_deck(deck, nr, ns, nu))
    rinf, rmed = o.field("Rinf"), o.field("Rmed")
    o.view("Dens")[:] = rho
    o.view("VradInt")[:] = (v0 * rinf ** 2)[:, None]
    o.view("VthetaInt")[:] = 0.0
    o.viscous_terms(1)
    assert np.allclose(o.field("TAURR")[1:nr - 1, 0], 2 * rho * nu * v0 * rmed[1:nr - 1], rtol=1e-3)
    assert np.abs(o.field("TAUPP")[1:nr - 1, 0]).max() < 2e-3 * 2 * rho * nu * v0 * rmed.max()
    assert np.allclose(o.field("DivergenceVelocity")[1:nr - 1, 0], 3 * v0 * rmed[1:nr - 1], rtol=1e-3)
    assert np.abs(o.field("TAURP")[1:nr]).max() < 1e-22
    # the update: Substep1 with uniform P, no potential, no rotation keeps
    # VradInt = Vrad before the viscous step
    o = Oracle(viscous_deck(deck, nr, ns, nu))
    o.view("Dens")[:] = rho
    o.view("Pressure")[:] = 1e-6
    o.view("Potential")[:] = 0.0
    o.view("Vtheta")[:] = 0.0
    vr0 = (v0 * rinf ** 2)[:, None] * np.ones((1, ns))
    o.view("Vrad")[:] = vr0
    o.substep1(dt)
    acc = (o.field("VradInt") - vr0)[2:nr - 1, 0] / dt
    assert np.allclose(acc, 4 * nu * v0, rtol=2e-3)


# ------------------------------------------------ artificial viscosity --
def test_compressive_flow_von_neumann_richtmyer_pressure(deck):
    """v_r = -a r (converging) and v_theta_j = -b j (compressive except at
    the wrap): the artificial pressures are q = CVNR^2 rho (dv)^2 with
    CVNR = 1.41 (Fondam.cuh:10), zero where the flow expands; the velocity
    update is -dt grad(q)/rho and the energy gains dt q |dv|/dx
    (Kernels.cu:409-479)."""
    nr, ns, rho, a, b, dt = 16, 24, 3e-3, 2e-2, 1e-3, 0.01
    o = Oracle(deck(Nrad=nr, Nsec=ns, Adiabatic="YES", AdiabaticIndex=1.4, Frame="FIXED", OmegaFrame=0.0))
    rinf, rsup, rmed = o.field("Rinf"), o.field("Rsup"), o.field("Rmed")
    o.view("Dens")[:] = rho
    vri = -a * rinf[:, None] * np.ones((1, ns))
    vti = -b * np.arange(ns)[None, :] * np.ones((nr + 1, 1))
    o.view("VradInt")[:] = vri
    o.view("VthetaInt")[:] = vti
    e0 = 5e-5
    o.view("Energy")[:] = e0
    o.substep2(dt)
    qr = o.field("DensInt")[:nr]
    qt = o.field("TemperInt")[:nr]
    dr = rsup[:nr] - rinf[:nr]
    assert np.allclose(qr, (CVNR ** 2 * rho * (a * dr) ** 2)[:, None], rtol=1e-13, atol=0)
    assert np.allclose(qt[:, :ns - 1], CVNR ** 2 * rho * b * b, rtol=1e-13, atol=0)
    assert not qt[:, ns - 1].any()  # expanding across the wrap
    o.host_kernel(dt)
    dq = (qr[1:] - qr[:-1])[:, 0]
    expect_vr = vri[1:nr] - dt * (dq / (rho * (rmed[1:nr] - rmed[:nr - 1])))[:, None]
    assert np.allclose(o.field("VradNew")[1:nr], expect_vr, rtol=1e-13, atol=1e-19)
    dxt = 2 * math.pi / ns * rmed[:nr]
    vtn = o.field("VthetaNew")[:nr]
    Q = CVNR ** 2 * rho * b * b
    assert np.allclose(vtn[:, 1:ns - 1], vti[:nr, 1:ns - 1], rtol=1e-15, atol=0)
    assert np.allclose(vtn[:, 0], vti[:nr, 0] - dt * Q / (rho * dxt), rtol=1e-13, atol=1e-19)
    assert np.allclose(vtn[:, ns - 1], vti[:nr, ns - 1] + dt * Q / (rho * dxt), rtol=1e-13, atol=0)
    heat = (dt * qr[:, 0] * a * dr / dr)[:, None] + dt * Q * b / dxt[:, None]
    expect_e = np.full((nr, ns), e0) + heat
    expect_e[:, ns - 1] = e0 + dt * qr[:, 0] * a  # no azimuthal heating where it expands
    assert np.allclose(o.field("EnergyInt")[:nr], expect_e, rtol=1e-13, atol=0)


# ------------------------------------------------------------- Substep3 --
def test_energy_compression_is_implicit_euler(deck):
    """nu = 0 (no heating), v_r = v0 r: the update solves
    e1 = e0 - dt (gamma - 1) e1 div v (backward Euler of de/dt =
    -(gamma-1) e div v), with div v = (r v_r)'/r over the cell."""
    nr, ns, v0, dt, g = 16, 12, 2e-3, 0.05, 1.4
    o = Oracle(deck(Nrad=nr, Nsec=ns, Adiabatic="YES", AdiabaticIndex=g, Viscosity=0.0, Frame="FIXED",
                    OmegaFrame=0.0))
    rinf, rsup, rmed = o.field("Rinf"), o.field("Rsup"), o.field("Rmed")
    o.view("Dens")[:] = 1e-3
    o.view("Vrad")[:] = (v0 * rinf)[:, None]
    o.view("Vtheta")[:] = 0.0
    eint = 1e-4 * (1 + 0.1 * np.random.default_rng(2).random((nr + 1, ns)))
    o.view("EnergyInt")[:] = eint
    o.substep3(dt)
    divv = v0 * (rsup[:nr] ** 2 - rinf[:nr] ** 2) / (rsup[:nr] - rinf[:nr]) / rmed[:nr]
    assert np.allclose(o.field("DivergenceVelocity")[:nr, 0], divv, rtol=1e-13)
    e1 = o.field("Energy")[:nr]
    assert np.allclose(e1, eint[:nr] / (1 + (g - 1) * dt * divv[:, None]), rtol=1e-14, atol=0)
    # the change itself (rounding of the difference: ~1e-16 / (dt (g-1) div v))
    assert np.allclose(e1 - eint[:nr], -dt * (g - 1) * e1 * divv[:, None], rtol=1e-10, atol=0)


def test_cooling_relaxation_is_implicit_euler(deck):
    """Static gas (no stress, no compression): the cooling update is backward
    Euler of de/dt = -(e - e_med Sigma/Sigma_med)/tau - Qplus_med
    Sigma/Sigma_med, tau = CoolingTime0 r^(2 + 2f) (Theo.cu:62-66),
    Qplus_med = (9/4) nu Sigma_0 r^(-s-3) (Theo.cu:79-85).  Ring 0 takes the
    power-law extrapolation of a zero heating, 0 exp(log(0/0)...) = NaN in
    the reference's arithmetic (Kernels.cu:89-95), and is left out."""
    nr, ns, dt, tau0, nu, s0 = 16, 12, 0.3, 6.28, 1e-5, 6.3662e-4
    o = Oracle(deck(Nrad=nr, Nsec=ns, Adiabatic="YES", AdiabaticIndex=1.4, Cooling="YES", CoolingTime0=tau0,
                    Viscosity=nu, Frame="FIXED", OmegaFrame=0.0, FlaringIndex=0.1, SigmaSlope=0.5))
    rmed = o.field("Rmed")[:nr]
    tau = o.field("CoolingTimeMed")[:nr]
    assert np.allclose(tau, tau0 * rmed ** (2 + 2 * 0.1), rtol=1e-14)
    smed, emed, qmed = o.field("SigmaMed")[:nr], o.field("EnergyMed")[:nr], o.field("QplusMed")[:nr]
    assert np.allclose(qmed, 2.25 * nu * s0 * rmed ** (-0.5 - 3.0), rtol=1e-14)
    rng = np.random.default_rng(8)
    sig = smed[:, None] * (1 + 0.3 * rng.random((nr, ns)))
    o.view("Dens")[:nr] = sig
    o.view("Vrad")[:] = 0.0
    o.view("Vtheta")[:] = 0.0
    e0 = emed[:, None] * (0.5 + rng.random((nr, ns)))
    o.view("EnergyInt")[:nr] = e0
    o.substep3(dt)
    x = sig / smed[:, None]
    # solve e1 = e0 + dt * (-(e1 - emed x)/tau - qmed x) for e1
    expect = (e0 + dt * emed[:, None] * x / tau[:, None] - dt * qmed[:, None] * x) / (1 + dt / tau[:, None])
    e1 = o.field("Energy")[:nr]
    assert np.allclose(e1[1:], expect[1:], rtol=1e-13, atol=0)
    assert np.isnan(e1[0]).all()


def test_cooling_equilibrium_is_a_fixed_point(deck):
    """nu = 0 (Qplus_med = 0): e = e_med Sigma/Sigma_med does not move."""
    nr, ns = 16, 12
    o = Oracle(deck(Nrad=nr, Nsec=ns, Adiabatic="YES", AdiabaticIndex=1.4, Cooling="YES", CoolingTime0=6.28,
                    Viscosity=0.0, Frame="FIXED", OmegaFrame=0.0))
    smed, emed = o.field("SigmaMed")[:nr], o.field("EnergyMed")[:nr]
    assert not o.field("QplusMed")[:nr].any()
    sig = smed[:, None] * (1 + 0.3 * np.random.default_rng(9).random((nr, ns)))
    o.view("Dens")[:nr] = sig
    o.view("Vrad")[:] = 0.0
    o.view("Vtheta")[:] = 0.0
    eq = emed[:, None] * sig / smed[:, None]
    o.view("EnergyInt")[:nr] = eq
    o.substep3(0.3)
    assert np.allclose(o.field("Energy")[:nr], eq, rtol=1e-15, atol=0)


def test_viscous_heating_of_a_shear_flow_and_its_inner_extrapolation(deck):
    """v_theta = A r^q, v_r = 0: the heating Kernels.cu:66-70 forms is
    0.5/(nu rho) (tau_rr^2 + tau_rphi^2 + tau_phiphi^2) + (2/9) nu rho (div v)^2
    = 0.5 rho nu (r dOmega/dr)^2 for this flow; ring 0 takes the power law
    through rings 1 and 2 (Kernels.cu:89-95)."""
    nr, ns, nu, rho, A, q = 128, 8, 1e-5, 2e-3, 1.0, -0.5
    o = Oracle(deck(Nrad=nr, Nsec=ns, Adiabatic="YES", AdiabaticIndex=1.4, Viscosity=nu, Frame="FIXED",
                    OmegaFrame=0.0, Rmin=0.5, Rmax=2.0))
    rinf, rmed = o.field("Rinf"), o.field("Rmed")
    o.view("Dens")[:] = rho
    o.view("Vrad")[:] = 0.0
    o.view("Vtheta")[:nr] = (A * rmed[:nr] ** q)[:, None]
    o.view("EnergyInt")[:] = 1e-4
    o.substep3(1e-3)
    qp = o.field("Qplus")[:, 0]
    # TAURP lives on the inner face: its square is averaged by nothing, so
    # compare with the face value
    expect = 0.5 * rho * nu * (A * (q - 1) * rinf[1:nr] ** (q - 1)) ** 2
    assert np.allclose(qp[1:nr], expect, rtol=2e-3)
    p = math.log(qp[1] / qp[2]) / math.log(rmed[1] / rmed[2])
    assert qp[0] == pytest.approx(qp[1] * (rmed[0] / rmed[1]) ** p, rel=1e-13)


# ------------------------------------------------------- alpha viscosity --
def test_alpha_viscosity_is_alpha_cs2_over_omega(deck):
    """nu = alpha c_s^2 / Omega_K with the ring mean of c_s (Viscosity.cu:33-50):
    isothermal c_s = h r^f Omega_K r gives nu = alpha h^2 r^(2f + 1/2); a
    non-uniform ring uses its arithmetic mean."""
    nr, ns, alpha, h, f = 20, 16, 3e-3, 0.06, 0.15
    o = Oracle(deck(Nrad=nr, Nsec=ns, Viscosity=None, AlphaViscosity=alpha, AspectRatio=h, FlaringIndex=f))
    rmed = o.field("Rmed")[:nr]
    o.view("VradInt")[:] = 0.0
    o.view("VthetaInt")[:] = 0.0
    o.viscous_terms(1)
    assert np.allclose(o.field("viscosity_array")[:nr], alpha * h * h * rmed ** (2 * f + 0.5), rtol=1e-13)
    cs = o.view("SoundSpeed")
    rng = np.random.default_rng(5)
    cs[:nr] *= 1 + 0.2 * rng.random((nr, ns))
    mean = np.array([math.fsum(row) / ns for row in o.field("SoundSpeed")[:nr]])
    o.viscous_terms(1)
    assert np.allclose(o.field("viscosity_array")[:nr], alpha * mean ** 2 * rmed ** 1.5, rtol=1e-14)


def test_viscosity_cavity_taper(deck):
    """CavityRadius/Ratio/Width: nu is multiplied by the ratio inside
    r_c - w h and by ratio^((r_max - r)/(r_max - r_min)) across the
    transition (Viscosity.cu:42-47)."""
    nr, ns, nu = 40, 8, 1e-5
    rc, ratio, w, h = 1.0, 0.1, 4.0, 0.05
    o = Oracle(deck(Nrad=nr, Nsec=ns, Viscosity=nu, CavityRadius=rc, CavityRatio=ratio, CavityWidth=w,
                    AspectRatio=h, SigmaSlope=0.0))
    o.viscous_terms(1)
    rmed = o.field("Rmed")[:nr]
    lo, hi = rc - w * h, rc + w * h
    expect = np.where(rmed < lo, nu * ratio, np.where(rmed <= hi, nu * ratio ** ((hi - rmed) / (hi - lo)), nu))
    assert (rmed < lo).any() and ((rmed >= lo) & (rmed <= hi)).any() and (rmed > hi).any()
    assert np.allclose(o.field("viscosity_array")[:nr], expect, rtol=1e-13)


# --------------------------------------------------------- radial transport --
def vanleer_column(q_spec, sig, v, rmed, rinf, rsup, dt):
    """Brute-force single-column radial van Leer step of FARGO's transport
    (TransportEuler.cu:106-156): the interface value of a quantity is its
    upwind cell's value plus the limited slope times the distance from that
    cell's centre to the middle of the swept region, measured with half the
    centre spacing; the quantity moves as interface density x interface
    specific value x velocity x face length x dt, zero through the edges."""
    n = len(sig)

    def slopes(f):
        s = np.zeros(n)
        for i in range(1, n - 1):
            left = (f[i] - f[i - 1]) / (rmed[i] - rmed[i - 1])
            right = (f[i + 1] - f[i]) / (rmed[i + 1] - rmed[i])
            s[i] = 2 * left * right / (left + right) if left * right > 0 else 0.0
        return s

    def face(f):
        s = slopes(f)
        out = np.zeros(n + 1)
        for i in range(1, n):
            if v[i] > 0:
                out[i] = f[i - 1] + 0.5 * (rmed[i] - rmed[i - 1] - v[i] * dt) * s[i - 1]
            else:
                out[i] = f[i] - 0.5 * (rmed[i + 1] - rmed[i] + v[i] * dt) * s[i]
        return out

    sig_face = face(sig)
    q_face = face(q_spec)
    return sig_face, q_face


def test_radial_van_leer_matches_a_brute_force_column(deck):
    """With v_theta = 0 (no azimuthal motion, no shift) and fields uniform in
    sector, transport reduces to the 1D radial sweep of every column: the
    oracle's density and radial momenta equal the brute-force column."""
    nr, ns, dt = 40, 6, 0.05
    o = Oracle(deck(Nrad=nr, Nsec=ns, Frame="FIXED", OmegaFrame=0.0))
    rinf, rsup, rmed = o.field("Rinf"), o.field("Rsup"), o.field("Rmed")
    rng = np.random.default_rng(21)
    sig = 1e-3 * (1 + 0.5 * np.sin(np.arange(nr) * 0.7) + 0.2 * rng.random(nr))
    v = np.zeros(nr + 1)
    v[1:nr] = 3e-2 * np.sin(np.arange(1, nr) * 0.45) * rng.uniform(0.5, 1.0, nr - 1)
    o.view("Dens")[:nr] = sig[:, None]
    o.view("Dens")[nr] = 0.0
    o.view("Vrad")[:] = v[:, None]
    o.view("Vtheta")[:] = 0.0
    o.transport(dt)
    dphi = 2 * math.pi / ns
    surf = math.pi * (rsup[:nr] ** 2 - rinf[:nr] ** 2) / ns
    # the conserved quantities of the sweep: Sigma and the radial momenta
    # Sigma v_r(outer face) and Sigma v_r(inner face) (ComputeLRMomenta)
    momp = sig * v[1:nr + 1]
    momm = sig * v[:nr]
    rmed_ext = np.append(rmed[:nr], 0.0)  # Rmed[NRAD] = 0 (Q8): multiplied by a zero slope
    sig_face, one_face = vanleer_column(np.ones(nr), sig, v, rmed_ext, rinf, rsup, dt)

    def step(Q):
        _, qf = vanleer_column(Q / sig, sig, v, rmed_ext, rinf, rsup, dt)
        flux = dt * dphi * rinf[:nr + 1] * qf * sig_face * v
        flux[0] = flux[nr] = 0.0
        return Q + (flux[:nr] - flux[1:]) / surf

    new_sig = step(sig)
    new_p, new_m = step(momp), step(momm)
    dens = o.field("Dens")[:nr]
    assert np.allclose(dens, new_sig[:, None], rtol=1e-13, atol=0)
    # velocities come back from the momenta (ComputeVelocities, Kernels.cu:1200-1205)
    vr_new = (new_p[:-1] + new_m[1:]) / (new_sig[1:] + new_sig[:-1] + 1e-20)
    assert np.allclose(o.field("Vrad")[1:nr], vr_new[:, None], rtol=1e-12, atol=1e-20)
    assert not o.field("Vrad")[0].any()


def test_divergence_free_radial_flow_keeps_uniform_density(deck):
    """v_r = C / r carries the same mass through every face (r v_r = C), so
    a uniform Sigma stays uniform in every ring whose two faces are open."""
    nr, ns, dt = 30, 8, 0.05
    o = Oracle(deck(Nrad=nr, Nsec=ns, Frame="FIXED", OmegaFrame=0.0))
    rinf = o.field("Rinf")
    o.view("Dens")[:nr] = 2e-3
    vr = o.view("Vrad")
    vr[:] = (4e-3 / rinf)[:, None]
    vr[0] = 0.0
    vr[nr] = 0.0
    o.view("Vtheta")[:] = 0.0
    o.transport(dt)
    d = o.field("Dens")[1:nr - 1]
    assert np.allclose(d, 2e-3, rtol=1e-14, atol=0)


def test_velocity_recovery_from_momenta(deck):
    """ComputeVelocities (Kernels.cu:1193-1210) undoes ComputeLRMomenta: for
    a vanishing dt the transport is the identity on any Sigma, v_r, v_theta,
    in a rotating frame too (v_theta + r Omega_F is what the momenta carry)."""
    nr, ns = 20, 24
    o = Oracle(deck(Nrad=nr, Nsec=ns, Frame="COROTATING", OmegaFrame=0.9))
    rng = np.random.default_rng(6)
    d = o.view("Dens")
    d[:nr] = 1e-3 * (1 + 0.5 * rng.random((nr, ns)))
    vr = o.view("Vrad")
    vr[1:nr] = rng.normal(0, 1e-3, (nr - 1, ns))
    vt = o.view("Vtheta")
    vt[:nr] = 0.2 + rng.normal(0, 1e-3, (nr, ns))
    before = [o.field(n) for n in ("Dens", "Vrad", "Vtheta")]
    o.transport(1e-13)
    assert o.ifield("Nshift")[:nr].tolist() == [0] * nr
    for name, b in zip(("Dens", "Vrad", "Vtheta"), before):
        a = o.field(name)[:nr]
        assert np.allclose(a, b[:nr], rtol=1e-9, atol=1e-15), name
    assert not o.field("Vrad")[0].any()


# ----------------------------------------------------------- boundaries --
def test_open_inner_boundary_ghost_values(deck):
    """OPEN (SideEuler.cu:127, Kernels.cu:482-500): ring 0 copies ring 1's
    density and energy; the inner face lets matter out (v_r copied from the
    next face when that face flows inward and ring 1 is denser than the
    initial edge value) and never in (v_r = 0 otherwise)."""
    nr, ns = 12, 16
    o = Oracle(deck(Nrad=nr, Nsec=ns, InnerBoundary="OPEN"))
    rng = np.random.default_rng(3)
    smed0 = o.field("SigmaMed")[0]
    d = o.view("Dens")
    d[:nr] = smed0 * rng.uniform(0.8, 1.2, (nr, ns))
    vr = o.view("Vrad")
    vr[1:nr] = rng.normal(0, 1e-3, (nr - 1, ns))
    before_d, before_v = o.field("Dens"), o.field("Vrad")
    o.apply_boundary(0.01)
    assert np.array_equal(o.field("Dens")[0], before_d[1])
    out = (before_v[2] <= 0.0) & (before_d[1] >= smed0)
    assert np.array_equal(o.field("Vrad")[1], np.where(out, before_v[2], 0.0))
    assert np.array_equal(o.field("Dens")[1:], before_d[1:]) and np.array_equal(o.field("Vrad")[2:], before_v[2:])


def test_non_reflecting_boundary_characteristics(deck):
    """NONREFLECTING (SideEuler.cu:45-100): the ghost ring takes the next
    ring's pattern advanced by the angle a sound wave drifts at the local
    Keplerian rate relative to the frame while crossing the gap,
    (Omega_K - 1) dr / <c_s>, rounded the reference's way ((int)(x + 0.5));
    its mean is then reset to the initial edge value; the edge face
    velocity averages to the outgoing-characteristic value
    -c_s (Sigma - Sigma_med) / Sigma_med."""
    nr, ns = 16, 64
    o = Oracle(deck(Nrad=nr, Nsec=ns, InnerBoundary="NONREFLECTING", Rmin=0.3))
    rinf, rmed = o.field("Rinf"), o.field("Rmed")
    smed = o.field("SigmaMed")
    cs = o.field("SoundSpeed")
    rng = np.random.default_rng(12)
    d = o.view("Dens")
    d[:nr] = smed[:nr, None] * (1 + 0.05 * rng.random((nr, ns)))
    d[1, 9] *= 1.5  # a feature to follow
    vr = o.view("Vrad")
    vr[1:nr] = rng.normal(0, 1e-3, (nr - 1, ns))
    before_d, before_v = o.field("Dens"), o.field("Vrad")
    o.apply_boundary(0.01)
    after_d, after_v = o.field("Dens"), o.field("Vrad")
    cbar = 0.5 * (cs[0].mean() + cs[1].mean())
    x = (rinf[1] ** -1.5 - 1.0) / cbar * (rmed[1] - rmed[0]) / (2 * math.pi) * ns
    shift = int(x + 0.5)
    assert shift > 0
    moved = np.roll(before_d[1], shift)
    assert np.argmax(after_d[0]) == (9 + shift) % ns
    assert np.allclose(after_d[0] - moved, smed[0] - moved.mean(), rtol=0, atol=1e-15 * smed[0])
    assert after_d[0].mean() == pytest.approx(smed[0], rel=1e-14)
    # outer edge: ring NRAD-1 takes ring NRAD-2 moved the same way (here
    # backwards: Omega_K < 1 outside r = 1), (int)(x + 0.5) truncating
    # towards zero for x < 0
    cbar2 = 0.5 * (cs[nr - 1].mean() + cs[nr - 2].mean())
    x2 = (rinf[nr - 2] ** -1.5 - 1.0) / cbar2 * (rmed[nr - 1] - rmed[nr - 2]) / (2 * math.pi) * ns
    assert x2 < 0
    moved2 = np.roll(before_d[nr - 2], int(x2 + 0.5))
    assert after_d[nr - 1].mean() == pytest.approx(smed[nr - 1], rel=1e-14)
    assert np.allclose(after_d[nr - 1] - moved2, smed[nr - 1] - moved2.mean(), rtol=0, atol=1e-15 * smed[0])
    # inner face: the mean of the ghost face and the next face is the
    # characteristic value of ring 1
    mid = 0.5 * (after_v[1] + before_v[2])
    assert np.allclose(mid, -cs[1] * (before_d[1] - smed[1]) / smed[1], rtol=1e-12, atol=1e-18)
    assert np.array_equal(after_d[2:nr - 1], before_d[2:nr - 1])


def test_sub_keplerian_edges_are_pressure_supported_rotation(deck):
    """Without evanescent edges, Substep1 sets v_theta of the two edge rings
    to the rotation in equilibrium with gravity and the pressure gradient of
    the initial profile (SideEuler.cu:189-213): v^2 = GM/r + c_s^2
    dlnP/dlnr with P ~ r^(-s + 2f - 1), minus the frame's r Omega_F."""
    nr, ns, h, s, f = 16, 12, 0.05, 0.7, 0.1
    o = Oracle(deck(Nrad=nr, Nsec=ns, InnerBoundary="OPEN", AspectRatio=h, SigmaSlope=s, FlaringIndex=f,
                    Frame="COROTATING", OmegaFrame=0.6))
    rmed = o.field("Rmed")
    OF = o.get("OmegaFrame")
    o.view("Pressure")[:] = 1e-6
    o.view("Potential")[:] = 0.0
    o.substep1(0.01)
    vti = o.field("VthetaInt")
    for i in (0, nr - 1):
        r = rmed[i]
        c2 = (h * r ** f) ** 2 / r
        v = math.sqrt(1.0 / r + c2 * (-s + 2 * f - 1)) - r * OF
        assert np.allclose(vti[i], v, rtol=1e-14), i


# --------------------------------------------------------------- frames --
def test_planet_is_fixed_in_the_corotating_frame(tmp_path):
    """COROTATING (Psys.cu:174-265, SourceEuler.cu:238-261): the frame turns
    with the planet, so a planet on a circular orbit keeps its angle and
    radius, and the frame rate is its Kepler rate sqrt(G(1 + q)/a^3)."""
    q = 1e-3
    dt = 2 * math.pi / math.sqrt(1 + q) / 400
    path = planet_deck(tmp_path, Disk="NO", Frame="COROTATING", OmegaFrame=1.0, Nrad=8, Nsec=8, q=q, DT=repr(dt))
    o = Oracle(path)
    for _ in range(100):
        o.algogas()
        x, y = o.planets()[0][:2]
        assert abs(x - 1.0) < 1e-9 and abs(y) < 1e-13  # radius: the Cash-Karp step's own error
    assert o.get("OmegaFrame") == pytest.approx(math.sqrt(1 + q), rel=1e-9)
    assert o.get("PhysicalTime") == pytest.approx(100 * dt, rel=1e-14)


def test_guiding_centre_frame_turns_at_the_mean_motion(tmp_path):
    """GUIDING-CENTER: the frame follows the planet's mean longitude, so for
    an eccentric Kepler orbit (e = 0.05, starting at apocentre) the frame
    rate is the mean motion n = sqrt(G(1 + q)/a^3), the planet's mean
    longitude stays at its initial 180 deg - 180 deg = 0 (it starts at
    apocentre, mean anomaly pi, on the +x axis) and the planet returns to
    its start after one period."""
    q, e = 1e-3, 0.05
    n = math.sqrt(1 + q)
    period = 2 * math.pi / n
    steps = 200
    path = planet_deck(tmp_path, Disk="NO", Frame="GUIDING-CENTER", OmegaFrame=0.0, Eccentricity=e, Nrad=8,
                       Nsec=8, q=q, DT=repr(period / steps))
    o = Oracle(path)
    x0, y0 = o.planets()[0][:2]
    assert (x0, y0) == (1.0 + e, 0.0)
    for k in range(steps):
        o.algogas()
        x, y, vx, vy, _ = o.planets()[0]
        # E = acos((1 - d/a)/e) is ill-conditioned at the apsides (argument
        # +-1), where the rate is good to a few 1e-6; elsewhere to 1e-8
        assert o.get("OmegaFrame") == pytest.approx(n, rel=1e-8 if abs(y) > 0.01 else 5e-6)
        # the planet's mean longitude from its state (Kepler's equation)
        r = math.hypot(x, y)
        a = 1.0 / (2.0 / r - (vx * vx + vy * vy) / (1 + q))
        cosE = (1 - r / a) / e
        sinE = (x * vx + y * vy) / (e * math.sqrt((1 + q) * a))
        E = math.atan2(sinE, cosE)
        M = E - e * math.sin(E)
        ex = ((vx * vx + vy * vy) / (1 + q) - 1 / r) * x - (x * vx + y * vy) * vx / (1 + q)
        ey = ((vx * vx + vy * vy) / (1 + q) - 1 / r) * y - (x * vx + y * vy) * vy / (1 + q)
        varpi = math.atan2(ey, ex)
        lam = math.remainder(M + varpi, 2 * math.pi)
        assert abs(lam) < 1e-5, (k, lam)
    x, y = o.planets()[0][:2]
    assert abs(x - (1 + e)) < 1e-6 and abs(y) < 1e-6


# --------------------------------------------------------- self-gravity --
def test_self_gravity_kick_and_ring_mean(deck):
    """Update_sgvelocityKernel (Kernels.cu:915-932): v_r at the inner face
    gains dt times g_r linearly interpolated in r to that face, v_theta at
    theta_(j-1/2) gains dt times the mean of its two neighbours' g_theta
    (periodic in j, Q4); the axisymmetric part used elsewhere is each ring's
    mean of g_r (Q6)."""
    nr, ns, dt = 10, 16, 0.02
    o = Oracle(deck(Nrad=nr, Nsec=ns, SelfGravity="YES", ThicknessSmoothing=0.3))
    rng = np.random.default_rng(23)
    o.view("Dens")[:nr] = 1e-3 * rng.uniform(0.5, 1.5, (nr, ns))
    vr0 = rng.normal(0, 1e-3, (nr + 1, ns))
    vt0 = rng.normal(0, 1e-3, (nr + 1, ns))
    o.view("Vradial")[:] = vr0
    o.view("Vazimutal")[:] = vt0
    o.selfgravity(dt, 1)
    gr, gt = o.field("SG_Accr")[:nr], o.field("SG_Acct")[:nr]
    rmed, rinf = o.field("Rmed")[:nr], o.field("Rinf")
    interp = np.array([[np.interp(rinf[i], rmed[i - 1:i + 1], gr[i - 1:i + 1, j]) for j in range(ns)]
                       for i in range(1, nr)])
    assert np.allclose(o.field("Vradial")[1:nr], vr0[1:nr] + dt * interp, rtol=1e-12, atol=1e-18)
    assert np.array_equal(o.field("Vradial")[0], vr0[0])
    assert np.allclose(o.field("Vazimutal")[:nr], vt0[:nr] + dt * 0.5 * (gt + np.roll(gt, 1, axis=1)),
                       rtol=1e-12, atol=1e-18)
    axi = o.field("GLOBAL_AxiSGAccr")[:nr]
    assert np.allclose(axi, gr.mean(axis=1), rtol=1e-14)



def test_self_gravity_is_the_gradient_of_the_softened_potential(deck):
    """SgMain.cu:23-81 + Kernels.cu:840-912 (Baruteau & Masset 2008): each
    cell carries the mass Sigma r^2 du dphi of the log-polar grid (u = ln r);
    the field point at radius r feels the potential -G m / sqrt(|x - x'|^2 +
    eps^2 r^2) of every other cell (the self cell is removed, which is what
    the + G Sigma du dphi / eps correction does).  The oracle's FFT
    convolution must equal -dPhi/dr and -(1/r) dPhi/dphi of the direct sum,
    taken by central differences.  Tolerance 1e-5 of the field's scale: the
    kernel table's FP32 coshf (Q3) perturbs 2(cosh u - cos theta) of the
    nearest radial neighbours (u = 0.17) by ~4e-6 relative; the differences
    add O(h^2) ~ 1e-10."""
    nr, ns = 8, 12
    o = Oracle(deck(Nrad=nr, Nsec=ns, SelfGravity="YES", ThicknessSmoothing=0.3, AspectRatio=0.05, Rmin=0.5,
                    Rmax=2.0))
    rng = np.random.default_rng(17)
    dens = o.view("Dens")
    dens[:nr] = 1e-3 * rng.uniform(0.5, 1.5, (nr, ns))
    o.selfgravity(0.0, 0)
    eps = o.get("SGP_eps")
    du, dphi = o.get("SGP_rstep"), o.get("SGP_tstep")
    assert eps == pytest.approx(0.3 * 0.05) and dphi == pytest.approx(2 * math.pi / ns)
    assert du == pytest.approx(math.log(2.0 / 0.5) / nr)
    rmed = o.field("Rmed")[:nr]
    S = o.field("Dens")[:nr]
    phis = 2 * math.pi * np.arange(ns) / ns
    mass = S * (rmed ** 2)[:, None] * du * dphi
    xs = (rmed[:, None] * np.cos(phis)[None, :]).ravel()
    ys = (rmed[:, None] * np.sin(phis)[None, :]).ravel()
    ms = mass.ravel()

    def pot(r, phi, skip):
        x, y = r * math.cos(phi), r * math.sin(phi)
        d2 = (x - xs) ** 2 + (y - ys) ** 2 + eps * eps * r * r
        t = -ms / np.sqrt(d2)
        t[skip] = 0.0
        return t.sum()

    gr = o.field("SG_Accr")[:nr]
    gt = o.field("SG_Acct")[:nr]
    scale = max(np.abs(gr).max(), np.abs(gt).max())
    for i in range(nr):
        for j in range(ns):
            r, phi, k = rmed[i], phis[j], i * ns + j
            hr, hp = 1e-5 * r, 1e-5
            dr = (pot(r + hr, phi, k) - pot(r - hr, phi, k)) / (2 * hr)
            dp = (pot(r, phi + hp, k) - pot(r, phi - hp, k)) / (2 * hp)
            assert gr[i, j] == pytest.approx(-dr, abs=1e-5 * scale), (i, j)
            assert gt[i, j] == pytest.approx(-dp / r, abs=1e-5 * scale), (i, j)


# ----------------------------------------------------- profiles and EOS --
def test_aspect_ratio_transition_in_the_smoothing_length(tmp_path):
    """TransitionRadius/Ratio/Width (AspectRatioHost, Theo.cu; Force.cu:137):
    h is multiplied by the ratio inside r_t - w h and by
    ratio^((r_max - r)/(r_max - r_min)) across the transition; the planet's
    potential smoothing is 0.6 h(r_p) r_p^(1+f)."""
    rt, ratio, w, h = 1.0, 0.5, 3.0, 0.05
    lo, hi = rt - w * h, rt + w * h
    for rp in (0.7, 0.95, 1.1, 1.6):
        o = Oracle(planet_deck(tmp_path, Nrad=16, Nsec=16, Frame="FIXED", OmegaFrame=0.0, TransitionRadius=rt,
                               TransitionRatio=ratio, TransitionWidth=w, AspectRatio=h))
        hr = h * ratio if rp < lo else (h * ratio ** ((hi - rp) / (hi - lo)) if rp <= hi else h)
        eps = 0.6 * hr * rp
        set_planet(o, 0, rp, 0.0, 0.0, 1.0, 1e-3)
        o.set("MassTaper", 1.0)
        o.set("DiskOnPrimaryAcceleration.x", 0.0)
        o.set("DiskOnPrimaryAcceleration.y", 0.0)
        o.fill_forces_arrays()
        rmed = o.field("Rmed")[:16]
        X = rmed[:, None] * np.cos(np.arange(16) / 16 * 2 * np.pi)
        Y = rmed[:, None] * np.sin(np.arange(16) / 16 * 2 * np.pi)
        expect = -1e-3 / np.sqrt((X - rp) ** 2 + Y ** 2 + eps ** 2) + 1e-3 * X * rp / rp ** 3 - 1 / rmed[:, None]
        assert np.allclose(o.field("Potential")[:16], expect, rtol=1e-13), rp


def test_equation_of_state(deck):
    """Isothermal: c_s = h r^f v_K, P = Sigma c_s^2, T = P/Sigma (MU/R = 1);
    adiabatic: c_s = sqrt(gamma (gamma-1) e / Sigma), P = (gamma-1) e,
    T = (gamma-1) e / Sigma (Kernels.cu:167-213); the initial energy is
    e = Sigma0 h^2 r^(-s-1+2f) / (gamma-1) (Theo.cu:37-50)."""
    nr, ns, h, f, s = 12, 10, 0.07, 0.2, 0.4
    o = Oracle(deck(Nrad=nr, Nsec=ns, AspectRatio=h, FlaringIndex=f, SigmaSlope=s))
    rmed = o.field("Rmed")[:nr]
    rng = np.random.default_rng(31)
    o.view("Dens")[:nr] = 1e-3 * (1 + rng.random((nr, ns)))
    o.sound_speed()
    o.pressure()
    o.temperature()
    cs = (h * rmed ** f * np.sqrt(1 / rmed))[:, None]
    d = o.field("Dens")[:nr]
    assert np.allclose(o.field("SoundSpeed")[:nr], cs, rtol=1e-14)
    assert np.allclose(o.field("Pressure")[:nr], d * cs * cs, rtol=1e-14)
    assert np.allclose(o.field("Temperature")[:nr], cs * cs, rtol=1e-14)
    g = 1.4
    oa = Oracle(deck(name="a.par", Nrad=nr, Nsec=ns, AspectRatio=h, FlaringIndex=f, SigmaSlope=s, Adiabatic="YES",
                     AdiabaticIndex=g))
    assert np.allclose(oa.field("EnergyMed")[:nr], 6.3662e-4 * h * h * rmed ** (-s - 1 + 2 * f) / (g - 1),
                       rtol=1e-14)
    oa.view("Dens")[:nr] = d
    e = 1e-5 * (1 + rng.random((nr, ns)))
    oa.view("Energy")[:nr] = e
    oa.sound_speed()
    oa.pressure()
    oa.temperature()
    assert np.allclose(oa.field("SoundSpeed")[:nr], np.sqrt(g * (g - 1) * e / d), rtol=1e-14)
    assert np.allclose(oa.field("Pressure")[:nr], (g - 1) * e, rtol=1e-15)
    assert np.allclose(oa.field("Temperature")[:nr], (g - 1) * e / d, rtol=1e-14)


def test_initial_density_cavity(deck):
    """CavityRadius/Ratio: Sigma = Sigma0 r^-s / ratio inside the cavity
    radius (Theo.cu:13-23)."""
    nr = 30
    o = Oracle(deck(Nrad=nr, Nsec=8, SigmaSlope=0.5, CavityRadius=1.0, CavityRatio=0.2, CavityWidth=1.0))
    rmed = o.field("Rmed")[:nr]
    expect = 6.3662e-4 * rmed ** -0.5 * np.where(rmed < 1.0, 1 / 0.2, 1.0)
    assert np.allclose(o.field("SigmaMed")[:nr], expect, rtol=1e-14)
    assert np.allclose(o.field("Dens")[:nr], expect[:, None], rtol=1e-14)


def test_correct_vtheta_removes_the_frame_change(deck):
    """CorrectVtheta (SideEuler.cu:181): a frame that speeds up by d Omega
    sees every ring slow down by r d Omega."""
    o = Oracle(deck(Nrad=10, Nsec=8))
    rmed = o.field("Rmed")
    before = o.field("Vtheta")
    o.correct_vtheta(0.125)
    assert np.allclose(o.field("Vtheta")[:10], before[:10] - 0.125 * rmed[:10, None], rtol=1e-15)
    assert np.array_equal(o.field("Vtheta")[10], before[10])


# ---------------------------------------------------------------- planets --
def test_forced_circular_orbit(tmp_path):
    """ForcedCircular (Pframeforce.cu:142-158): each step turns the planet on
    its circle by d theta = omega dt, omega = (x v_y - y v_x)/r^2, and turns
    its velocity by the planet's new absolute angle theta + d theta (the
    reference's rotation; it equals the increment only on the first step
    from theta = 0).  Both are rotations: r and |v| are kept."""
    dt = 0.1
    o = Oracle(planet_deck(tmp_path, Disk="NO", Frame="FIXED", OmegaFrame=0.0, ForcedCircular="YES", Nrad=8,
                           Nsec=8, DT=repr(dt), dist=1.3))
    x, y, vx, vy, m = o.planets()[0]
    assert (x, y) == (1.3, 0.0)
    speed = math.hypot(vx, vy)

    def rot(a, u, v):
        return u * math.cos(a) - v * math.sin(a), u * math.sin(a) + v * math.cos(a)

    for k in range(12):
        theta = math.atan2(y, x)
        w = (x * vy - y * vx) / (x * x + y * y)
        if k == 0:
            assert w == pytest.approx(math.sqrt((1 + m) / 1.3 ** 3), rel=1e-12)
        ex, ey = rot(theta + w * dt, 1.3, 0.0)
        evx, evy = rot(theta + w * dt, vx, vy)
        o.algogas()
        x, y, vx, vy, _ = o.planets()[0]
        assert (x, y) == (pytest.approx(ex, abs=1e-14), pytest.approx(ey, abs=1e-14)), k
        assert (vx, vy) == (pytest.approx(evx, abs=1e-14), pytest.approx(evy, abs=1e-14)), k
        assert math.hypot(x, y) == pytest.approx(1.3, rel=1e-14)
        assert math.hypot(vx, vy) == pytest.approx(speed, rel=1e-14)


def test_release_migration_moves_the_radius_linearly(tmp_path):
    """ReleaseDate/ReleaseRadius (Pframeforce.cu:160-186): before the release
    date the planet's radius goes linearly from its start to the release
    radius, arriving at the release date."""
    dt, T, R1 = 0.05, 2.0, 1.4
    o = Oracle(planet_deck(tmp_path, Disk="NO", Frame="FIXED", OmegaFrame=0.0, ReleaseDate=T, ReleaseRadius=R1,
                           Nrad=8, Nsec=8, DT=repr(dt)))
    for k in range(1, 41):
        o.algogas()
        x, y = o.planets()[0][:2]
        t = k * dt
        assert math.hypot(x, y) == pytest.approx(1.0 + (R1 - 1.0) * t / T, rel=1e-12), k


def test_planet_planet_term_uses_the_partner_stellar_distance(tmp_path):
    """DerivMotionRK5 (RungeKutta.cu:55-60) adds G m_j (x_j - x_i)/|x_j|^3 for
    FeelOthers planets: the partner's distance to the star, not the mutual
    distance it computes and leaves unused.  A vanishing step measures the
    acceleration: star + that term (indirect term off)."""
    dt = 1e-7
    cfg = write_planets(tmp_path / "two.cfg", [("A", 1.0, 1e-3, 0.0, "NO", "YES"), ("B", 1.6, 2e-3, 0.0, "NO", "YES")])
    o = Oracle(write_deck(tmp_path / "two.par", Disk="NO", Frame="FIXED", OmegaFrame=0.0, IndirectTerm="NO",
                          PlanetConfig=cfg, Nrad=8, Nsec=8, DT=repr(dt), OutputDir=str(tmp_path) + "/"))
    p0 = o.planets()
    o.algogas()
    p1 = o.planets()
    for i, j in ((0, 1), (1, 0)):
        xi, xj, mj = p0[i][0], p0[j][0], p0[j][4]
        ax = -1.0 / xi ** 2 + mj * (xj - xi) / abs(xj) ** 3
        mutual = -1.0 / xi ** 2 + mj * (xj - xi) / abs(xj - xi) ** 3
        got = (p1[i][2] - p0[i][2]) / dt
        assert got == pytest.approx(ax, rel=1e-6)
        assert abs(got - mutual) > 1e-3 * abs(mj / (xj - xi) ** 2)


def test_substep_bookkeeping(deck):
    """AlgoGas (SourceEuler.cu:199-307): the first substep sees MassTaper 0
    (Q1); the sub-steps of one DT add up to DT; main()'s loop makes NTOT+1
    AlgoGas calls (Q16)."""
    o = Oracle(deck(Nrad=16, Nsec=24, Ntot=3, Ninterm=1, MassTaper=5.0))
    DT = o.get("DT")
    t = [0.0]
    first = True
    while True:
        closed = o.substep()
        if first:
            assert o.get("MassTaper") == 0.0
            first = False
        t.append(o.get("PhysicalTime"))
        if closed:
            break
    assert all(b > a for a, b in zip(t, t[1:]))
    assert t[-1] == pytest.approx(DT, rel=1e-14)
    assert o.get("MassTaper") > 0.0
    o2 = Oracle(deck(name="b.par", Nrad=16, Nsec=24, Ntot=3, Ninterm=1))
    o2.run(0, o2.iget("NTOT"))
    assert o2.get("PhysicalTime") == pytest.approx(4 * DT, rel=1e-13)


# ----------------------------------------------------- azimuthal transport --
def azimuthal_sweep(Q, sig, V, dxt, area_per_surf, dt, q_is_one=False):
    """One periodic van Leer sweep of a ring (TransportEuler.cu:229-269):
    the face value between sectors j-1 and j is the upwind sector's value
    plus the limited slope times the distance to the middle of the swept
    arc; the quantity moves as face density x face specific value x
    velocity x dt through each face."""
    def face(f):
        dm = f - np.roll(f, 1)
        dp = np.roll(f, -1) - f
        slope = np.where(dp * dm > 0, dp * dm / np.where(dp * dm > 0, dp + dm, 1.0) / dxt, 0.0)
        ksi = V * dt
        up = np.roll(f, 1) + (dxt - ksi) * np.roll(slope, 1)
        down = f - (dxt + ksi) * slope
        return np.where(ksi > 0, up, down)

    sig_face = face(sig)
    q_face = np.ones_like(sig) if q_is_one else face(Q / sig)
    flux = area_per_surf * q_face * sig_face * V
    return Q + flux - np.roll(flux, -1)


def test_fargo_azimuthal_transport_matches_a_brute_force_ring(deck):
    """v_r = 0: transport is, ring by ring, the FARGO split of v_theta into
    the ring mean (an integer shift Nshift = floor(N~ + 1/2) plus a uniform
    residual of N~ - Nshift sectors per step) and the residual field, each
    swept by periodic van Leer (residual first, then uniform), then the
    shift, then the velocity from the azimuthal momenta
    (TransportEuler.cu:177-245, Kernels.cu:1094-1210); a brute-force ring
    gives the same Sigma and v_theta."""
    nr, ns, dt, OF = 6, 48, 0.05, 0.3
    o = Oracle(deck(Nrad=nr, Nsec=ns, Frame="FIXED"))
    o.set("OmegaFrame", OF)
    rinf, rsup, rmed, surf = o.field("Rinf"), o.field("Rsup"), o.field("Rmed"), o.field("Surf")
    rng = np.random.default_rng(44)
    sig = 1e-3 * (1 + 0.4 * np.sin(np.arange(ns) * 2 * np.pi / ns * 3)[None, :] + 0.3 * rng.random((nr, ns)))
    # mean motion of 2.3 sectors per step (different per ring), plus noise
    cells = 2.3 + 0.5 * np.arange(nr)
    vt = (cells * 2 * np.pi * rmed[:nr] / ns / dt)[:, None] + rng.normal(0, 2e-3, (nr, ns))
    o.view("Dens")[:nr] = sig
    o.view("Vrad")[:] = 0.0
    o.view("Vtheta")[:nr] = vt
    o.transport(dt)
    nshift = o.ifield("Nshift")[:nr]
    for i in range(nr):
        R, dxt = rmed[i], 2 * np.pi / ns * rmed[i]
        vmed = sum(vt[i]) / ns  # sequential j sum, as the reference's ring loop
        ntilde = vmed / R * dt * ns / 2 / np.pi
        nround = math.floor(ntilde + 0.5)
        assert nshift[i] == nround
        vres = vt[i] - vmed
        vuni = np.full(ns, (ntilde - nround) * R / dt * 2 * np.pi / ns)
        tmp = sig[i] * (np.roll(vt[i], -1) + R * OF) * R
        tmm = sig[i] * (vt[i] + R * OF) * R
        d = sig[i].copy()
        a = (rsup[i] - rinf[i]) * dt / surf[i]
        for V in (vres, vuni):
            tmp = azimuthal_sweep(tmp, d, V, dxt, a, dt)
            tmm = azimuthal_sweep(tmm, d, V, dxt, a, dt)
            d = azimuthal_sweep(d, d, V, dxt, a, dt, q_is_one=True)
        d, tmp, tmm = (np.roll(f, nround) for f in (d, tmp, tmm))
        v_new = (np.roll(tmp, 1) + tmm) / (d + np.roll(d, 1) + 1e-15) / R - R * OF
        assert np.allclose(o.field("Dens")[i], d, rtol=1e-12, atol=0), i
        assert np.allclose(o.field("Vtheta")[i], v_new, rtol=1e-11, atol=1e-15), i
    assert not o.field("Vrad")[:nr].any()
