"""fg_disk_force_hill_noalloc: the allocation-free variant of
fg_disk_force_hill (scratch in the library's per-GPU device array, results
to a device pointer) gives the same bits as fg_disk_force_hill, on two calls
in a row, the second on a smaller disk with fewer Hill-cut variants (the
scratch keeps the first call's sums; the second must not read them)."""
import ctypes
import os

import numpy as np
import pytest

from cudaprotoplanetarydisks_amd import Simulation
from cudaprotoplanetarydisks_amd.native import FG_E_INVALID, FargoError, call, lib

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime libfargo_hip.so is linked against (the caller's side
    of the test: it owns the output buffer)."""
    lib()  # loads libamdhip64 as a dependency first
    for name in ("libamdhip64.so.7", "libamdhip64.so.6", "libamdhip64.so",
                 "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.fail("libamdhip64 not found")


def noalloc_forces(hip, sim, args, dimfxy):
    dev = ctypes.c_void_p()
    nbytes = 8 * 4 * dimfxy
    assert hip.hipMalloc(ctypes.byref(dev), ctypes.c_size_t(nbytes)) == 0
    try:
        call("fg_disk_force_hill_noalloc", sim.ctx.h, *args, dimfxy, 0.9, 0.07, dev)
        call("fg_sync", sim.ctx.h)
        out = np.empty(4 * dimfxy)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, ctypes.c_size_t(nbytes), 2) == 0
    finally:
        hip.hipFree(dev)
    return out


@pytest.mark.parametrize("fused", [False, True], ids=["per-routine", "fused"])
def test_noalloc_matches_fg_disk_force_hill_twice(deck, fused):
    hip = hip_runtime()
    cases = [(deck(name="big.par", Nrad=96, Nsec=256), 11, (0.93, 0.21, 0.03)),
             (deck(name="small.par", Nrad=24, Nsec=48), 5, (-0.4, 1.1, 0.05))]
    for k, (path, dimfxy, args) in enumerate(cases):
        with Simulation(path, fused=fused) as sim:
            for _ in range(2):
                sim.substep()
            print(f"noalloc call {k}: grid {sim.nrad}x{sim.nsec} dimfxy {dimfxy}", flush=True)
            ref = sim.ctx.disk_force_hill(*args, dimfxy, 0.9, 0.07)
            got = noalloc_forces(hip, sim, args, dimfxy)
            assert np.array_equal(got, ref), np.abs(got - ref).max()
            assert np.isfinite(got).all() and np.abs(got).max() > 0


def test_noalloc_rejects_too_many_variants(deck):
    with Simulation(deck(Nrad=16, Nsec=32)) as sim:
        with pytest.raises(FargoError) as e:
            call("fg_disk_force_hill_noalloc", sim.ctx.h, 1.0, 0.0, 0.03, 65, 0.9, 0.07, ctypes.c_void_p(8))
        assert e.value.status == FG_E_INVALID and "dimfxy" in str(e.value)
