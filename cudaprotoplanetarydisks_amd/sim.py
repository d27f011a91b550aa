"""Python handle on the C++ host driver (fg_sim_* of fargo_hip.h).

Simulation mirrors the reference's main()/AlgoGas (Main.cu:89-363,
SourceEuler.cu:199-307): it reads a .par deck, keeps the planets and the
frame on the host in FP64 and runs the substep path on the GPU through
libfargo_hip.so.
"""
import ctypes
import os

import numpy as np

from . import native
from .native import FargoError, call, lib

_pd = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_pd) if a is not None else None


class Simulation:
    def __init__(self, parfile, device=0, precision="fp64", fused=False, write_files=False, outdir=None,
                 cli_flags="", noise_seed=0, noise_amp=0.0, diagnostics=False, fp32_pairs=None):
        opt = native.FgSimOptions()
        opt.device = device
        opt.precision = native.FG_PREC_FP32 if precision == "fp32" else native.FG_PREC_FP64
        opt.fused = 1 if fused else 0
        opt.write_files = 1 if write_files else 0
        self._outdir = outdir.encode() if outdir else None
        opt.outdir = self._outdir
        self._flags = cli_flags.encode()
        opt.cli_flags = self._flags
        opt.noise_seed = noise_seed
        opt.noise_amp = noise_amp
        opt.diagnostics = 1 if diagnostics else 0
        h = ctypes.c_void_p()
        call("fg_sim_create", os.fsencode(parfile), ctypes.byref(opt), ctypes.byref(h))
        self.h = h
        self.nrad = int(self.get("NRAD"))
        self.nsec = int(self.get("NSEC"))
        if precision == "fp32":  # autotuned tile shapes of this grid, if any
            from .ensemble import load_tiles
            t = load_tiles().get(f"{self.nrad}x{self.nsec}")
            if t:
                call("fg_set_tiles", lib().fg_sim_ctx(self.h), int(t["src_rows"]), int(t["trad_band"]))
            if fp32_pairs is not None:  # Hill radii of float-pair Sigma storage (default 1)
                call("fg_set_fp32_pairs", lib().fg_sim_ctx(self.h), float(fp32_pairs))

    def close(self):
        if getattr(self, "h", None):
            lib().fg_sim_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        return (self.nrad + 1, self.nsec)

    def get(self, name):
        v = ctypes.c_double()
        call("fg_sim_get", self.h, name.encode(), ctypes.byref(v))
        return v.value

    def substep(self):
        """One iteration of AlgoGas's sub-cycle; True when the DT closed."""
        closed = ctypes.c_int()
        call("fg_sim_substep", self.h, ctypes.byref(closed))
        return bool(closed.value)

    def algogas(self):
        call("fg_sim_algogas", self.h)

    def run(self, it0, it1):
        call("fg_sim_run", self.h, it0, it1)

    def write_outputs(self, index):
        call("fg_sim_write_outputs", self.h, index)

    def restart(self, index):
        """Continue from output set `index` of this deck's output directory."""
        call("fg_sim_restart", self.h, index)

    def state(self, energy=False):
        out = [np.empty(self.shape) for _ in range(4 if energy else 3)]
        call("fg_sim_get_state", self.h, *[_ptr(a) for a in out], *([] if energy else [None]))
        return out

    def set_state(self, dens=None, vrad=None, vtheta=None, energy=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (dens, vrad, vtheta, energy)]
        call("fg_sim_set_state", self.h, *[_ptr(a) for a in arrs])

    def planets(self):
        nb = ctypes.c_int()
        call("fg_sim_nplanets", self.h, ctypes.byref(nb))
        out = np.empty((nb.value, 5))
        for k in range(nb.value):
            call("fg_sim_planet", self.h, k, _ptr(out[k]))
        return out

    def ring(self, name):
        out = np.empty(self.nrad + 1)
        call("fg_sim_ring", self.h, name.encode(), _ptr(out))
        return out

    @property
    def ctx(self):
        return Context(lib().fg_sim_ctx(self.h), self.shape, owner=self)


class Context:
    """Device-side routines of one fg_ctx (one entry per reference routine)."""

    def __init__(self, handle, shape, owner=None):
        self.h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle
        self.shape = shape
        self._owner = owner

    def field(self, name):
        out = np.empty(self.shape)
        call("fg_get_field", self.h, native.FIELD[name], _ptr(out))
        return out

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        call("fg_set_field", self.h, native.FIELD[name], _ptr(a))

    def rings(self):
        n = self.shape[0] - 1
        out = {k: np.empty(n) for k in ("vmoy", "vmed", "viscosity", "cs_mean", "sgacc_mean")}
        call("fg_get_rings", self.h, *[_ptr(out[k]) for k in ("vmoy", "vmed", "viscosity", "cs_mean", "sgacc_mean")])
        return out

    def nshift(self):
        out = np.empty(self.shape[0] - 1, dtype=np.int32)
        call("fg_get_nshift", self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def cfl(self, DeltaT):
        n = ctypes.c_int()
        call("fg_cfl", self.h, DeltaT, ctypes.byref(n))
        return n.value

    def disk_force(self, x=0.0, y=0.0, rs=0.0):
        out = np.empty(2)
        call("fg_disk_force", self.h, x, y, rs, _ptr(out))
        return out

    def disk_force_hill(self, x, y, rs, dimfxy, a, rh):
        """ComputeForce (Force.cu:102-135): the 4*dimfxy GlobalForce sums."""
        out = np.empty(4 * dimfxy)
        call("fg_disk_force_hill", self.h, x, y, rs, int(dimfxy), a, rh, _ptr(out))
        return out

    def circumplanetary_mass(self, xpl, ypl, hill_radius):
        """CircumPlanetaryMass (SourceEuler.cu:549-563)."""
        out = ctypes.c_double()
        call("fg_circumplanetary_mass", self.h, xpl, ypl, hill_radius, ctypes.byref(out))
        return out.value

    def potential(self, planets, ix, iy):
        arr = (native.FgPlanetPot * max(1, len(planets)))()
        for k, p in enumerate(planets):
            arr[k] = native.FgPlanetPot(*p)
        call("fg_potential", self.h, arr, len(planets), ix, iy)

    def __getattr__(self, name):
        fn = "fg_" + name
        if fn in native.SIGNATURES:
            return lambda *a: call(fn, self.h, *a)
        raise AttributeError(name)
