"""The fused kernels divide several numerators by one denominator with one
reciprocal refinement (recip64 in fg_cell.h: v_rcp_f64 and two Newton
steps on d, then the corrected quotient and v_div_fixup per numerator).
Parity with the oracle needs every such quotient to have the bits of the
plain division n / d; fg_check_shared_division compares the two on the GPU.

Inputs: random mantissas over the whole range the claim covers (numerator
exponents -969..700, denominator exponents -70..70, both signs), the
physical range of the kernels' operands, signed zeros, infinities and NaN."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mismatches(num, den):
    from cudaprotoplanetarydisks_amd.native import call
    num = np.ascontiguousarray(num, dtype=np.float64)
    den = np.ascontiguousarray(den, dtype=np.float64)
    bad = ctypes.c_long(-1)
    pd = ctypes.POINTER(ctypes.c_double)
    call("fg_check_shared_division", num.ctypes.data_as(pd), den.ctypes.data_as(pd), len(num), ctypes.byref(bad))
    return bad.value


def test_wide_range_is_bit_identical():
    rng = np.random.default_rng(7)
    n = 1 << 22
    num = np.ldexp(1.0 + rng.random(n), rng.integers(-969, 701, n)) * rng.choice([-1.0, 1.0], n)
    den = np.ldexp(1.0 + rng.random(n), rng.integers(-70, 71, n)) * rng.choice([-1.0, 1.0], n)
    assert mismatches(num, den) == 0


def test_kernel_operands_are_bit_identical():
    # momenta / (Sigma + 1e-20), slopes / ring spacings, vtheta sums / Rmed
    rng = np.random.default_rng(8)
    n = 1 << 20
    sigma = np.ldexp(1.0 + rng.random(n), rng.integers(-40, 0, n))
    num = np.ldexp(1.0 + rng.random(n), rng.integers(-60, 10, n)) * rng.choice([-1.0, 1.0], n)
    assert mismatches(num, sigma + 1e-20) == 0
    assert mismatches(num, np.full(n, 1e-20)) == 0


def test_special_numerators():
    den = np.array([1e-3, 2.5, 1e-20, 7.0, 3e-5, 1.0])
    for special in (0.0, -0.0, np.inf, -np.inf, np.nan):
        assert mismatches(np.full(len(den), special), den) == 0
