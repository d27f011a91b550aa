"""Ensemble parameter draws (CPU): ranges and shapes of the distributions."""
import math

import numpy as np

from cudaprotoplanetarydisks_amd.ensemble import draw


def test_draws_follow_the_stated_distributions():
    ps = draw(4000, seed=1)
    q = np.array([p["q"] for p in ps])
    a = np.array([p["alpha"] for p in ps])
    h = np.array([p["h"] for p in ps])
    s = np.array([p["slope"] for p in ps])
    nr = np.array([p["nrad"] for p in ps])
    assert q.min() >= 1e-4 and q.max() <= 3e-3
    assert abs(np.median(np.log(q)) - 0.5 * (math.log(1e-4) + math.log(3e-3))) < 0.15
    assert a.min() >= 1e-4 and a.max() <= 1e-2
    assert h.min() >= 0.03 and h.max() <= 0.08
    assert s.min() >= -1.5 and s.max() <= 0.0
    assert set(nr) == {128, 192, 256, 384}
    assert all(p["nsec"] == 3 * p["nrad"] for p in ps)


def test_draws_are_reproducible():
    assert draw(8, seed=3) == draw(8, seed=3)
    assert draw(8, seed=3) != draw(8, seed=4)
