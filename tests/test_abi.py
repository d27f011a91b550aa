"""C-ABI of libfargo_hip.so: the library loads, exports every entry point
include/fargo_hip.h declares, and the host-only error paths (deck parsing,
which runs before any HIP call) return the documented status codes."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, write_deck, write_planets

HEADER = os.path.join(ROOT, "include", "fargo_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:fg_status|const char \*|int|long|fg_ctx \*|fg_sim \*)\s*(fg_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_substep_path():
    names = declared_functions()
    for required in ("fg_create", "fg_destroy", "fg_cfl", "fg_disk_force", "fg_potential", "fg_boundary",
                     "fg_substep1", "fg_substep2", "fg_substep3", "fg_transport", "fg_selfgravity",
                     "fg_step_fused", "fg_sim_create", "fg_sim_algogas"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from cudaprotoplanetarydisks_amd import lib
    L = lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from cudaprotoplanetarydisks_amd.native import SIGNATURES
    assert set(declared_functions()) <= set(SIGNATURES)


def test_last_error_is_callable_without_gpu():
    from cudaprotoplanetarydisks_amd import lib
    assert isinstance(lib().fg_last_error(), (bytes, type(None)))


def _create_status(path):
    from cudaprotoplanetarydisks_amd import lib
    h = ctypes.c_void_p()
    return lib().fg_sim_create(os.fsencode(path), None, ctypes.byref(h))


def test_missing_deck_is_an_io_error(tmp_path):
    assert _create_status(tmp_path / "nope.par") == 5  # FG_E_IO


def test_viscosity_and_alpha_together_rejected(tmp_path):
    # Interpret.cu:177-183 exit(1)s; the library returns FG_E_PARAM
    path = write_deck(tmp_path / "d.par", AlphaViscosity=1e-3)
    assert _create_status(path) == 4


def test_both_smoothings_rejected(tmp_path):
    path = write_deck(tmp_path / "d.par", RocheSmoothing=0.5)  # Interpret.cu:188-194
    assert _create_status(path) == 4


def test_no_smoothing_rejected(tmp_path):
    path = write_deck(tmp_path / "d.par", ThicknessSmoothing=0.0)  # Interpret.cu:195-201
    assert _create_status(path) == 4


def test_missing_planet_file_is_an_io_error(tmp_path):
    path = write_deck(tmp_path / "d.par", PlanetConfig=str(tmp_path / "none.cfg"))  # Psys.cu:17-20
    assert _create_status(path) == 5


def test_fused_signature_matches_header():
    text = open(HEADER).read()
    m = re.search(r"fg_status fg_step_fused\(([^)]*)\)", text, flags=re.S)
    nargs = len([a for a in m.group(1).split(",") if a.strip()])
    from cudaprotoplanetarydisks_amd.native import SIGNATURES
    assert len(SIGNATURES["fg_step_fused"][1]) == nargs
