"""ctypes binding of libfargo_hip.so (include/fargo_hip.h).

The library is the product; this module only loads it and mirrors its
structs.  Loading fails loudly when the library has not been built: there
is no Python or CPU fallback for any entry point.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libfargo_hip.so")
EXE_PATH = os.path.join(_PKG, "lib", "fargo_mi355x")
DECKS = os.path.join(_PKG, "decks")

FG_OK = 0
FG_E_INVALID, FG_E_HIP, FG_E_UNSUPPORTED, FG_E_PARAM, FG_E_IO, FG_E_CRASH, FG_E_NOMEM = range(1, 8)
FG_PREC_FP64, FG_PREC_FP32 = 0, 1
FG_MAX_PLANETS = 16

FIELDS = ["DENS", "VRAD", "VTHETA", "ENERGY", "PRESSURE", "SOUNDSPEED", "TEMPERATURE", "POTENTIAL",
          "VRADINT", "VTHETAINT", "VRADNEW", "VTHETANEW", "ENERGYINT", "ENERGYNEW", "DENSINT",
          "TEMPERINT", "QPLUS", "DIVV", "TAURR", "TAUPP", "TAURP", "RADMOMP", "RADMOMM",
          "THETAMOMP", "THETAMOMM", "VTHETARES", "SG_ACCR", "SG_ACCT"]
FIELD = {n: k for k, n in enumerate(FIELDS)}

_d = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_pf = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class FgParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "nrad", "nsec", "precision", "adiabatic", "cooling", "self_gravity", "viscosity_alpha",
        "fast_transport", "open_inner", "non_reflecting", "evanescent", "indirect_term",
        "exclude_hill")] + [(n, _d) for n in (
            "adiabatic_index", "alphaviscosity", "sigmaslope", "flaringindex", "sg_eps", "sg_rstep",
            "sg_tstep")]


class FgPlanetPot(ctypes.Structure):
    _fields_ = [("x", _d), ("y", _d), ("mass", _d), ("smooth", _d), ("inv_d3", _d)]


class FgSimOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("precision", ctypes.c_int), ("fused", ctypes.c_int),
                ("write_files", ctypes.c_int), ("outdir", ctypes.c_char_p),
                ("cli_flags", ctypes.c_char_p), ("noise_seed", ctypes.c_ulonglong),
                ("noise_amp", _d), ("diagnostics", ctypes.c_int)]


# exported symbols with (restype, argtypes); every one returns fg_status
# unless noted
SIGNATURES = {
    "fg_create": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "fg_destroy": (ctypes.c_int, [_vp]),
    "fg_last_error": (ctypes.c_char_p, []),
    "fg_sync": (ctypes.c_int, [_vp]),
    "fg_upload_state": (ctypes.c_int, [_vp, _pd, _pd, _pd, _pd]),
    "fg_download_state": (ctypes.c_int, [_vp, _pd, _pd, _pd, _pd, _pd, _pd]),
    "fg_get_field": (ctypes.c_int, [_vp, ctypes.c_int, _pd]),
    "fg_set_field": (ctypes.c_int, [_vp, ctypes.c_int, _pd]),
    "fg_get_rings": (ctypes.c_int, [_vp, _pd, _pd, _pd, _pd, _pd]),
    "fg_get_nshift": (ctypes.c_int, [_vp, _pi]),
    "fg_sound_speed": (ctypes.c_int, [_vp]),
    "fg_pressure": (ctypes.c_int, [_vp]),
    "fg_temperature": (ctypes.c_int, [_vp]),
    "fg_ring_mean_cs": (ctypes.c_int, [_vp]),
    "fg_cfl": (ctypes.c_int, [_vp, _d, _pi]),
    "fg_disk_force": (ctypes.c_int, [_vp, _d, _d, _d, _pd]),
    "fg_disk_force_hill": (ctypes.c_int, [_vp, _d, _d, _d, ctypes.c_int, _d, _d, _pd]),
    "fg_check_shared_division": (ctypes.c_int, [_pd, _pd, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
    "fg_circumplanetary_mass": (ctypes.c_int, [_vp, _d, _d, _d, _pd]),
    "fg_disk_force_hill_noalloc": (ctypes.c_int, [_vp, _d, _d, _d, ctypes.c_int, _d, _d, _vp]),
    "fg_potential": (ctypes.c_int, [_vp, ctypes.POINTER(FgPlanetPot), ctypes.c_int, _d, _d]),
    "fg_correct_vtheta": (ctypes.c_int, [_vp, _d]),
    "fg_boundary": (ctypes.c_int, [_vp, _d, _d]),
    "fg_substep1": (ctypes.c_int, [_vp, _d, _d]),
    "fg_substep2": (ctypes.c_int, [_vp, _d]),
    "fg_substep3": (ctypes.c_int, [_vp, _d]),
    "fg_transport": (ctypes.c_int, [_vp, _d, _d, _pd]),
    "fg_selfgravity": (ctypes.c_int, [_vp, _d, ctypes.c_int]),
    "fg_step_fused": (ctypes.c_int, [_vp, _d, _d, _d, ctypes.POINTER(FgPlanetPot), ctypes.c_int, _d, _d, _d,
                                     ctypes.c_int]),
    "fg_fused_prepare": (ctypes.c_int, [_vp, _d]),
    "fg_fused_results": (ctypes.c_int, [_vp, _pi, _pd, _pd]),
    "fg_set_tiles": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "fg_set_fp32_pairs": (ctypes.c_int, [_vp, ctypes.c_double]),
    "fg_profile": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fg_profile_read": (ctypes.c_int, [_vp, _pd, ctypes.POINTER(ctypes.c_long), ctypes.c_int]),
    "fg_sim_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(FgSimOptions), ctypes.POINTER(_vp)]),
    "fg_sim_destroy": (ctypes.c_int, [_vp]),
    "fg_sim_substep": (ctypes.c_int, [_vp, _pi]),
    "fg_sim_algogas": (ctypes.c_int, [_vp]),
    "fg_sim_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "fg_sim_write_outputs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fg_sim_restart": (ctypes.c_int, [_vp, ctypes.c_int]),
    "fg_sim_get_state": (ctypes.c_int, [_vp, _pd, _pd, _pd, _pd]),
    "fg_sim_set_state": (ctypes.c_int, [_vp, _pd, _pd, _pd, _pd]),
    "fg_sim_get": (ctypes.c_int, [_vp, ctypes.c_char_p, _pd]),
    "fg_sim_nplanets": (ctypes.c_int, [_vp, _pi]),
    "fg_sim_planet": (ctypes.c_int, [_vp, ctypes.c_int, _pd]),
    "fg_sim_ctx": (_vp, [_vp]),
    "fg_sim_ring": (ctypes.c_int, [_vp, ctypes.c_char_p, _pd]),
    "fg_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(FgSimOptions),
                                       ctypes.POINTER(_vp)]),
    "fg_batch_destroy": (ctypes.c_int, [_vp]),
    "fg_batch_set_tiles": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fg_batch_substep": (ctypes.c_int, [_vp, _d, _pi]),
    "fg_batch_run": (ctypes.c_int, [_vp, _d, ctypes.POINTER(ctypes.c_long)]),
    "fg_batch_sim": (_vp, [_vp, ctypes.c_int]),
    "fg_mlp_create": (ctypes.c_int, [_pi, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int,
                                     ctypes.POINTER(_vp)]),
    "fg_mlp_destroy": (ctypes.c_int, [_vp]),
    "fg_mlp_nparam": (ctypes.c_long, [_vp]),
    "fg_mlp_get_params": (ctypes.c_int, [_vp, _pf]),
    "fg_mlp_set_params": (ctypes.c_int, [_vp, _pf]),
    "fg_mlp_forward": (ctypes.c_int, [_vp, _pf, ctypes.c_int, _pf]),
    "fg_mlp_gradient": (ctypes.c_int, [_vp, _pf, _pf, ctypes.c_int, _pf, _pf]),
    "fg_mlp_train": (ctypes.c_int, [_vp, _pf, _pf, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _pf]),
    "fg_bf16_gemm": (ctypes.c_int, [_pf, _pf, ctypes.c_int, ctypes.c_int, ctypes.c_int, _pf]),
}


class FargoError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__(f"{where}: status {status}: {message}")
        self.status = status


_lib = None


def lib():
    """Load libfargo_hip.so (once).  Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} not built; run python -m cudaprotoplanetarydisks_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != FG_OK:
        msg = lib().fg_last_error()
        raise FargoError(status, where, msg.decode() if msg else "")
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
