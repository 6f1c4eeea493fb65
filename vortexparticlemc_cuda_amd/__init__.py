"""MI355X-native (gfx950) vortex-particle Monte Carlo hot path.

The product is libvpmc.so (hand-written HIP behind the C ABI in
include/vpmc.h).  This package is its Python binding: `_capi` loads the
library (raising if it is absent), `ops` wraps each entry point for torch
device tensors, `sim` mirrors the reference's SimParam / Simulation.
"""
from . import _capi  # noqa: F401
from ._capi import MODE_CONSISTENT, MODE_REFERENCE, VpmcError, lib  # noqa: F401

__all__ = ["lib", "MODE_CONSISTENT", "MODE_REFERENCE", "VpmcError"]
