"""Emulator on small grids: the network trained on batched runs predicts
held-out profiles better than the plain initial state does, and a warm
start through fg_sim_set_state gives a finite run that starts closer to
steady than the cold start."""
import ctypes
import math

import numpy as np
import pytest

from cudaprotoplanetarydisks_amd import emulator as em
from cudaprotoplanetarydisks_amd.native import call

pytestmark = pytest.mark.gpu
_pd = ctypes.POINTER(ctypes.c_double)


def test_emulator_beats_the_plain_initial_state(tmp_path):
    train = em.draw(64, 1, 32)
    held = em.draw(8, 2, 32)
    y = em.generate(train, 10.0)
    yh = em.generate(held, 10.0)
    emu = em.Emulator(max_batch=64)
    try:
        losses = emu.fit(em.features(train), y, steps=1500, lr=1e-3)
        assert losses[-1] < 0.2 * losses[0]
        pred = emu.predict(em.features(held))
    finally:
        emu.close()
    err_emu = np.abs(pred[:, :em.NOUT] - yh[:, :em.NOUT]).max(axis=1)
    err_plain = np.abs(yh[:, :em.NOUT]).max(axis=1)
    assert np.median(err_emu) < 0.5 * np.median(err_plain), (err_emu, err_plain)


def test_warm_state_is_a_valid_restart(tmp_path):
    ps = em.draw(2, 3, 32)
    with em.open_ensemble(ps, str(tmp_path)) as ens:
        rmed = ens.ring(0, "Rmed")[:32]
        y = np.zeros(2 * em.NOUT, np.float32)
        y[:em.NOUT] = -0.3 * np.exp(-((em.R_OUT - 1.0) / 0.1) ** 2)  # a gap
        for k, p in enumerate(ps):
            d, vr, vt = em.warm_state(ens, k, p, y)
            assert np.isfinite(vt).all() and (d[:-1] > 0).all()
            call("fg_sim_set_state", ens.sim_handle(k), d.ctypes.data_as(_pd), vr.ctypes.data_as(_pd),
                 vt.ctypes.data_as(_pd), None)
        sig0 = em.ring_profiles(ens, 0)[0]
        gap = sig0[np.argmin(np.abs(rmed - 1.0))]
        assert gap < 0.8
        ens.run(2 * em.ORBIT)
        for k in range(len(ps)):
            d, vr, vt = ens.state(k)
            assert np.isfinite(d).all() and np.isfinite(vt).all()


def test_steadiness_history_and_thresholds(tmp_path):
    ps = em.draw(3, 4, 24)
    with em.open_ensemble(ps, str(tmp_path)) as ens:
        R, hist, wall = em.orbit_history(ens, 4)
    assert len(R) == 3 and all(len(r) == 3 for r in R)
    assert all(len(h) == 4 for h in hist)
    assert all(np.isfinite(x) and x >= 0 for r in R for x in r)
    assert em.first_below([0.5, 0.2, 0.05], 0.1) == 4
    assert em.first_below([0.5, 0.2], 0.1) is None
