"""Self-gravity on the GPU (hand-written LDS FFT, fg_sg.hip) against the
oracle's FFT convolution (SgMain.cu:23-45, Kernels.cu:840-932), and a
cross-check of the FFT against rocFFT (torch.fft on the GPU).

The GPU FFT runs the oracle's butterflies in the oracle's order with the
oracle's twiddles, so the SG accelerations and whole SG substeps are
expected to match it bit for bit (isothermal) or within 1e-12 (adiabatic,
for the Qplus extrapolation of Kernels.cu:93)."""
import os

import numpy as np
import pytest

from conftest import DECKS
from oracle_binding import Oracle
from test_gpu_parity import run_pair

pytestmark = pytest.mark.gpu

SG = dict(SelfGravity="YES", ThicknessSmoothing=0.3)


def sg_pair(path, seed=3):
    from cudaprotoplanetarydisks_amd import Simulation
    orc = Oracle(path)
    sim = Simulation(path)
    rng = np.random.default_rng(seed)
    dens = orc.view("Dens")
    dens[:-1] = orc.field("SigmaMed")[:-1, None] * rng.uniform(0.5, 1.5, dens[:-1].shape)
    sim.set_state(dens=dens)
    orc.selfgravity(0.0, 0)
    sim.ctx.selfgravity(0.0, 0)
    return sim, orc


@pytest.mark.parametrize("grid", [(24, 48), (20, 60), (16, 42), (9, 25)])
def test_sg_accelerations_bitexact(deck, grid):
    """2NRAD x NSEC = 48x48, 40x60 (radix 2,3,5), 32x42 (a radix-7 direct
    DFT), 18x25 (5x5)."""
    nr, ns = grid
    sim, orc = sg_pair(deck(Nrad=nr, Nsec=ns, **SG))
    for g, o in (("SG_ACCR", "SG_Accr"), ("SG_ACCT", "SG_Acct")):
        a, b = sim.ctx.field(g)[:nr], orc.field(o)[:nr]
        assert np.array_equal(a, b), f"{g}: max diff {np.abs(a - b).max():.3e}"
    axi = sim.ctx.rings()["sgacc_mean"]
    assert np.array_equal(axi, orc.field("GLOBAL_AxiSGAccr")[:nr])


def test_sg_fft_matches_rocfft(deck, tmp_path):
    """The same convolution with torch.fft (rocFFT) in FP64, computed in a
    child process: the GPU FFT agrees to 1e-12 of the field's scale."""
    nr, ns = 32, 96
    sim, orc = sg_pair(deck(Nrad=nr, Nsec=ns, **SG), seed=11)
    check_against_rocfft(sim, orc, tmp_path)


def check_against_rocfft(sim, orc, tmp_path, tol=1e-12):
    """SG accelerations of the GPU against the oracle's inputs convolved by
    rocFFT (torch.fft in a child process)."""
    import subprocess
    import sys
    nr, ns = orc.nrad, orc.nsec
    dens = orc.field("Dens")[:nr]
    Rmed = orc.field("Rmed")[:nr]
    ratio = Rmed / Rmed[0]
    var = dens * np.sqrt(ratio)[:, None]
    Sr = np.zeros((2 * nr, ns))
    St = np.zeros((2 * nr, ns))
    Sr[:nr] = var
    St[:nr] = var * ratio[:, None]
    Kr = orc.field("Kr_aux")[: 2 * nr * ns].reshape(2 * nr, ns)
    Kt = orc.field("Kt_aux")[: 2 * nr * ns].reshape(2 * nr, ns)
    src, dst = tmp_path / "in.npz", tmp_path / "out.npz"
    np.savez(src, Sr=Sr, St=St, Kr=Kr, Kt=Kt)
    child = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rocfft_reference.py")
    r = subprocess.run([sys.executable, child, str(src), str(dst)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    z = np.load(dst)
    rstep, tstep, eps = orc.get("SGP_rstep"), orc.get("SGP_tstep"), orc.get("SGP_eps")
    norm = rstep * tstep / (2 * nr * ns)
    accr = z["Gr"] * (norm / np.sqrt(ratio))[:, None] + dens * rstep * tstep / eps
    acct = z["Gt"] * (norm / (ratio * np.sqrt(ratio)))[:, None]
    for g, ref in (("SG_ACCR", accr), ("SG_ACCT", acct)):
        a = sim.ctx.field(g)[:nr]
        err = np.abs(a - ref).max() / np.abs(ref).max()
        assert err < tol, f"{g}: {err:.3e}"


def test_sg_isothermal_substeps_bitexact(deck):
    run_pair(deck(Nrad=24, Nsec=48, **SG), 10, 0.0)


def test_sg_initial_state_matches(deck):
    from cudaprotoplanetarydisks_amd import Simulation
    path = deck(Nrad=24, Nsec=48, Frame="COROTATING", **SG)
    orc, sim = Oracle(path), Simulation(path)
    _, _, vt = sim.state()
    assert np.array_equal(vt, orc.field("Vtheta"))
    assert np.array_equal(sim.planets(), orc.planets())
    assert sim.get("OmegaFrame") == orc.get("OmegaFrame")


def test_adsg_deck_substeps(tmp_path):
    """The c2 physics (adiabatic, beta cooling, alpha viscosity, SG, planet
    feeling the disk, corotating frame) on a reduced 32 x 48 grid."""
    src = open(os.path.join(DECKS, "c2_adsg.par")).read()
    src = src.replace("Nrad                512", "Nrad                32")
    src = src.replace("Nsec                1536", "Nsec                48")
    src = src.replace("jupiter_feeldisk.cfg", os.path.join(DECKS, "jupiter_feeldisk.cfg"))
    src = src.replace("out_c2/", str(tmp_path) + "/")
    path = tmp_path / "c2_small.par"
    path.write_text(src)
    run_pair(str(path), 12, 1e-12, adiabatic=True)
