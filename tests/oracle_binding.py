"""ctypes handle on the FP64 oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_pd = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_libs = {}


def oracle_lib(omp=False):
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name not in _libs:
        path = os.path.join(ORACLE_DIR, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run make -C oracle")
        L = ctypes.CDLL(path)
        L.or_create.restype = _vp
        L.or_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.or_destroy.argtypes = [_vp]
        L.or_last_error.restype = ctypes.c_char_p
        L.or_substep.argtypes = [_vp]
        L.or_algogas.argtypes = [_vp]
        L.or_run.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.or_write_outputs.argtypes = [_vp, ctypes.c_int]
        L.or_field.restype = _pd
        L.or_field.argtypes = [_vp, ctypes.c_char_p]
        L.or_field_len.restype = ctypes.c_long
        L.or_field_len.argtypes = [_vp, ctypes.c_char_p]
        L.or_ifield.restype = ctypes.POINTER(ctypes.c_int)
        L.or_ifield.argtypes = [_vp, ctypes.c_char_p]
        L.or_get.restype = ctypes.c_double
        L.or_get.argtypes = [_vp, ctypes.c_char_p]
        L.or_set.argtypes = [_vp, ctypes.c_char_p, ctypes.c_double]
        L.or_iget.argtypes = [_vp, ctypes.c_char_p]
        L.or_nrad.argtypes = [_vp]
        L.or_nsec.argtypes = [_vp]
        L.or_nplanets.argtypes = [_vp]
        L.or_planet.argtypes = [_vp, ctypes.c_int, _pd]
        L.or_set_planet.argtypes = [_vp, ctypes.c_int, _pd]
        L.or_device_reduce.restype = ctypes.c_double
        L.or_device_reduce.argtypes = [_pd, ctypes.c_int]
        L.or_condition_cfl.argtypes = [_vp, ctypes.c_double]
        L.or_compute_accel.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _pd]
        L.or_compute_force_hill.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_double, _pd]
        L.or_circumplanetary_mass_at.restype = ctypes.c_double
        L.or_circumplanetary_mass_at.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.or_circumplanetary_mass.restype = ctypes.c_double
        L.or_circumplanetary_mass.argtypes = [_vp]
        L.or_enable_diagnostics.argtypes = [_vp]
        L.or_update_log.argtypes = [_vp, ctypes.c_int]
        L.or_write_bigplanet.argtypes = [_vp, ctypes.c_int]
        for fn in ("or_fill_forces_arrays", "or_sound_speed", "or_pressure", "or_temperature",
                   "or_actualise_velocities", "or_make1d_cs"):
            getattr(L, fn).argtypes = [_vp]
        for fn in ("or_apply_boundary", "or_correct_vtheta", "or_substep1", "or_substep2", "or_host_kernel",
                   "or_substep3", "or_transport"):
            getattr(L, fn).argtypes = [_vp, ctypes.c_double]
        L.or_viscous_terms.argtypes = [_vp, ctypes.c_int]
        L.or_selfgravity.argtypes = [_vp, ctypes.c_double, ctypes.c_int]
        L.or_fft2d.argtypes = [_pd, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _libs[name] = L
    return _libs[name]


class Oracle:
    """One oracle simulation (the reference's main() state)."""

    def __init__(self, parfile, cli_flags="", outdir=None, write_files=False, omp=False):
        self.L = oracle_lib(omp)
        h = self.L.or_create(os.fsencode(parfile), cli_flags.encode(), outdir.encode() if outdir else None,
                             1 if write_files else 0)
        if not h:
            raise RuntimeError("oracle: " + self.L.or_last_error().decode())
        self.h = h
        self.nrad = self.L.or_nrad(h)
        self.nsec = self.L.or_nsec(h)

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    __del__ = close

    def view(self, name):
        """Writable numpy view of an internal array."""
        n = self.L.or_field_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        a = np.ctypeslib.as_array(self.L.or_field(self.h, name.encode()), shape=(n,))
        if n == (self.nrad + 1) * self.nsec:
            return a.reshape(self.nrad + 1, self.nsec)
        return a

    def field(self, name):
        return self.view(name).copy()

    def ifield(self, name):
        return np.ctypeslib.as_array(self.L.or_ifield(self.h, name.encode()), shape=(self.nrad + 1,)).copy()

    def get(self, name):
        return self.L.or_get(self.h, name.encode())

    def set(self, name, v):
        assert self.L.or_set(self.h, name.encode(), v) == 0, name

    def iget(self, name):
        return self.L.or_iget(self.h, name.encode())

    def planets(self):
        out = np.empty((self.L.or_nplanets(self.h), 5))
        for k in range(out.shape[0]):
            self.L.or_planet(self.h, k, out[k].ctypes.data_as(_pd))
        return out

    def substep(self):
        return bool(self.L.or_substep(self.h))

    def algogas(self):
        self.L.or_algogas(self.h)

    def run(self, it0, it1):
        self.L.or_run(self.h, it0, it1)

    def compute_accel(self, x, y, rs):
        out = np.empty(2)
        self.L.or_compute_accel(self.h, x, y, rs, 0.0, out.ctypes.data_as(_pd))
        return out

    def force_hill(self, x, y, rs, dimfxy, a, rh):
        """ComputeForce (Force.cu:102-135): GlobalForce, 4*dimfxy sums."""
        out = np.empty(4 * dimfxy)
        self.L.or_compute_force_hill(self.h, x, y, rs, dimfxy, a, rh, out.ctypes.data_as(_pd))
        return out

    def circumplanetary_mass(self, xpl, ypl, hill):
        return self.L.or_circumplanetary_mass_at(self.h, xpl, ypl, hill)

    def __getattr__(self, name):
        fn = getattr(self.__dict__.get("L") or oracle_lib(), "or_" + name)
        return lambda *a: fn(self.h, *a)


def device_reduce(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return oracle_lib().or_device_reduce(a.ctypes.data_as(_pd), a.size)


def fft2d(z, sign):
    """Oracle 2D DFT of complex array z (n1 x n2)."""
    a = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1), dtype=np.float64)
    oracle_lib().or_fft2d(a.ctypes.data_as(_pd), z.shape[0], z.shape[1], sign)
    return a[..., 0] + 1j * a[..., 1]
