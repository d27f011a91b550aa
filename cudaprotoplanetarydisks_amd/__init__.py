"""MI355X (gfx950) implementation of the FARGO / FARGO-ADSG hydro substep.

The product is libfargo_hip.so (HIP kernels + C-ABI, include/fargo_hip.h)
and the fargo_mi355x executable, a drop-in for the reference's
bin/fargoGPU.  This package builds them and exposes thin ctypes handles.
"""
from .native import DECKS, EXE_PATH, LIB_PATH, FargoError, lib  # noqa: F401
from .sim import Context, Simulation  # noqa: F401

__all__ = ["Simulation", "Context", "FargoError", "lib", "LIB_PATH", "EXE_PATH", "DECKS"]
