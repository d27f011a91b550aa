"""The five workloads at their own sizes, one short run each.

configs[0] (128 x 384) runs 100 substeps in test_gpu_parity.py and
test_gpu_fused.py.  Here:
  configs[1]  ADSG + self-gravity at 512 x 1536 (decks/c2_adsg.par): two
              per-routine substeps against the oracle (rel L-inf <= 1e-12,
              CFL count and Nshift equal), and the SG field on the
              1024 x 1536 padded grid against rocFFT (<= 1e-12);
  configs[2]  1024 x 3072: fused FP64 against the oracle bit for bit over
              three substeps (Nshift equal), FP32 bulk against FP64 over
              one DT (Sigma rel L-inf <= 1e-5);
  configs[3]  the 64 disks of seed 20261202: a few batch substeps, finite
              positive fields, and every disk bit-identical to its
              single-disk run;
  configs[4]  256 x 768 emulator disks (flaring drawn): a warm state set
              through fg_sim_set_state runs one orbit finite.
"""
import os

import numpy as np
import pytest

from conftest import DECKS
from oracle_binding import Oracle

pytestmark = pytest.mark.gpu
DT = 0.314159265359


def c2_deck(tmp_path):
    src = open(os.path.join(DECKS, "c2_adsg.par")).read()
    src = src.replace("jupiter_feeldisk.cfg", os.path.join(DECKS, "jupiter_feeldisk.cfg"))
    src = src.replace("out_c2/", str(tmp_path) + "/")
    path = tmp_path / "c2_adsg.par"
    path.write_text(src)
    return str(path)


def rel_linf(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_config1_adsg_512x1536_against_the_oracle(tmp_path):
    from cudaprotoplanetarydisks_amd import Simulation
    path = c2_deck(tmp_path)
    orc = Oracle(path, omp=True)
    with Simulation(path) as sim:
        assert (sim.nrad, sim.nsec) == (512, 1536)
        assert sim.get("SelfGravity") == 1.0 and sim.get("Adiabatic") == 1.0
        d, vr, vt, e = sim.state(energy=True)
        for g, name in ((d, "Dens"), (vr, "Vrad"), (vt, "Vtheta"), (e, "Energy")):
            assert np.array_equal(g, orc.field(name)), f"initial {name}"
        for k in range(2):
            print(f"configs[1] substep {k}", flush=True)
            sim.substep()
            orc.substep()
            assert int(sim.get("last_cfl")) == orc.iget("last_cfl"), k
            assert np.array_equal(sim.ctx.nshift(), orc.ifield("Nshift")[:-1]), k
            assert np.allclose(sim.planets(), orc.planets(), rtol=1e-12, atol=1e-15), k
            d, vr, vt, e = sim.state(energy=True)
            for g, name in ((d, "Dens"), (vr, "Vrad"), (vt, "Vtheta"), (e, "Energy")):
                err = rel_linf(g, orc.field(name))
                assert err <= 1e-12, f"substep {k}: {name} rel Linf {err:.3e}"
            for g, name in (("SG_ACCR", "SG_Accr"), ("SG_ACCT", "SG_Acct")):
                err = rel_linf(sim.ctx.field(g)[:512], orc.field(name)[:512])
                assert err <= 1e-12, f"substep {k}: {g} rel Linf {err:.3e}"


def test_config1_sg_field_1024x1536_against_rocfft(tmp_path):
    from cudaprotoplanetarydisks_amd import Simulation
    from test_gpu_selfgravity import check_against_rocfft
    path = c2_deck(tmp_path)
    orc = Oracle(path, omp=True)
    with Simulation(path) as sim:
        rng = np.random.default_rng(12)
        dens = orc.view("Dens")
        dens[:-1] = orc.field("SigmaMed")[:-1, None] * rng.uniform(0.5, 1.5, dens[:-1].shape)
        sim.set_state(dens=dens)
        sim.ctx.selfgravity(0.0, 0)
        check_against_rocfft(sim, orc, tmp_path)


def c3_deck(tmp_path, name="c3.par"):
    from cudaprotoplanetarydisks_amd.workloads import write_deck
    return write_deck(tmp_path / name, Nrad=1024, Nsec=3072, OutputDir=str(tmp_path) + "/")


def test_config2_fused_fp64_1024x3072_bitexact(tmp_path):
    from cudaprotoplanetarydisks_amd import Simulation
    from cudaprotoplanetarydisks_amd.workloads import NOISE_AMP, NOISE_SEED
    path = c3_deck(tmp_path)
    with Simulation(path, fused=True, noise_seed=NOISE_SEED, noise_amp=NOISE_AMP) as sim:
        assert sim.get("fused_active") == 1.0
        orc = Oracle(path, omp=True)
        d0, _, _ = sim.state()
        orc.view("Dens")[:] = d0  # the same seeded noise (generated on the host)
        assert not np.array_equal(d0[:-1], orc.field("SigmaMed")[:-1, None] * np.ones((1, 3072)))
        for k in range(3):
            print(f"configs[2] fp64 substep {k}", flush=True)
            sim.substep()
            orc.substep()
            assert int(sim.get("last_cfl")) == orc.iget("last_cfl"), k
            # the disk's force on the star (DeviceReduce order) of this substep
            for ax in ("IndirectTerm.x", "IndirectTerm.y"):
                assert sim.get(ax) == orc.get(ax), (k, ax, sim.get(ax), orc.get(ax))
            assert np.array_equal(sim.ctx.nshift(), orc.ifield("Nshift")[:-1]), k
            d, vr, vt = sim.state()
            for g, name in ((d, "Dens"), (vr, "Vrad"), (vt, "Vtheta")):
                assert np.array_equal(g, orc.field(name)), f"substep {k}: {name} {rel_linf(g, orc.field(name)):.3e}"


def test_config2_fp32_bulk_within_1e5_of_fp64_over_one_dt(tmp_path):
    from cudaprotoplanetarydisks_amd import Simulation
    from cudaprotoplanetarydisks_amd.workloads import NOISE_AMP, NOISE_SEED
    path = c3_deck(tmp_path)
    out = {}
    for prec in ("fp64", "fp32"):
        with Simulation(path, fused=True, precision=prec, noise_seed=NOISE_SEED, noise_amp=NOISE_AMP) as sim:
            sim.algogas()
            out[prec] = (sim.state(), sim.get("PhysicalTime"), sim.get("substeps"))
    (d64, vr64, vt64), t64, n64 = out["fp64"]
    (d32, vr32, vt32), t32, n32 = out["fp32"]
    assert t64 == pytest.approx(DT, rel=1e-12) and t32 == pytest.approx(DT, rel=1e-12)
    err = rel_linf(d32[:-1], d64[:-1])
    print(f"configs[2] fp32 vs fp64 after one DT ({int(n32)} / {int(n64)} substeps): Sigma rel Linf {err:.3e}")
    assert err <= 1e-5
    assert rel_linf(vt32[:-1], vt64[:-1]) <= 1e-5


def test_config3_64_disk_ensemble(tmp_path):
    from cudaprotoplanetarydisks_amd import Simulation
    from cudaprotoplanetarydisks_amd.ensemble import make_ensemble
    ens, params, directory = make_ensemble(64, 20261202, directory=str(tmp_path), tiles={})
    try:
        assert ens.n == 64 and sorted({p["nrad"] for p in params}) == [128, 192, 256, 384]
        steps = 3
        for s in range(steps):
            print(f"configs[3] batch substep {s}", flush=True)
            assert ens.substep(1e30) == 64
        for k in range(64):
            d, vr, vt = ens.state(k)
            assert np.isfinite(d).all() and np.isfinite(vr).all() and np.isfinite(vt).all(), k
            assert (d[:-1] > 0).all(), k
        # every disk against the same deck run alone
        for k, p in enumerate(params):
            nrad = p["nrad"]
            deck = os.path.join(directory, f"disk{k}.par")
            with Simulation(deck, precision="fp32") as alone:
                for _ in range(steps):
                    alone.substep()
                for a, b, name in zip(ens.state(k), alone.state(), ("dens", "vrad", "vtheta")):
                    assert np.array_equal(a, b), f"disk {k} ({nrad} rings) {name}"
    finally:
        ens.close()


def test_config4_warm_start_256x768_runs_finite(tmp_path):
    import ctypes
    from cudaprotoplanetarydisks_amd import emulator as em
    from cudaprotoplanetarydisks_amd.native import call
    pd = ctypes.POINTER(ctypes.c_double)
    ps = em.draw(2, 7, 256)
    assert all(p["nsec"] == 768 and 0.0 <= p["flaring"] <= 0.25 for p in ps)
    with em.open_ensemble(ps, str(tmp_path)) as ens:
        y = np.zeros(2 * em.NOUT, np.float32)
        y[:em.NOUT] = -0.4 * np.exp(-((em.R_OUT - 1.0) / 0.08) ** 2)  # a predicted gap
        y[em.NOUT:] = -1.0
        for k, p in enumerate(ps):
            d, vr, vt = em.warm_state(ens, k, p, y)
            call("fg_sim_set_state", ens.sim_handle(k), d.ctypes.data_as(pd), vr.ctypes.data_as(pd),
                 vt.ctypes.data_as(pd), None)
        ens.run(em.ORBIT)
        for k in range(len(ps)):
            d, vr, vt = ens.state(k)
            assert np.isfinite(d).all() and np.isfinite(vr).all() and np.isfinite(vt).all()
            assert (d[:-1] > 0).all()
            assert ens.get(k, "PhysicalTime") >= em.ORBIT - 1e-9
