"""piro_amd -- MI355X (gfx950) implementation of PIRO's operator-and-solve path.

The compute lives in libpiro_hip.so (HIP kernels behind the C ABI of
include/piro_hip.h).  ``piro_amd.core.Context`` is a thin torch-tensor
front-end over that ABI; importing it without the built library raises.
"""
from ._lib import LIB_PATH, PiroError, check, declared_symbols, lib  # noqa: F401

__all__ = ["lib", "LIB_PATH", "PiroError", "check", "declared_symbols"]
