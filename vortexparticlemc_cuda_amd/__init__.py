"""MI355X-native (gfx950) vortex-particle Monte Carlo hot path.

The product is libvpmc.so (hand-written HIP behind the C ABI in
include/vpmc.h).  This package is its Python binding: `_capi` loads the
library (raising if it is absent), `ops` wraps each entry point for torch
device tensors, `sim` mirrors the reference's SimParam / Simulation.
`phash()` loads the reference's bindings module (lib/phash.so: the C++ API
and the differentiable rollout as a C++ torch.autograd.Function).
"""
from . import _capi  # noqa: F401
from ._capi import MODE_CONSISTENT, MODE_REFERENCE, VpmcError, lib  # noqa: F401

__all__ = ["lib", "phash", "MODE_CONSISTENT", "MODE_REFERENCE", "VpmcError"]


def phash():
    """the bindings module phash (bindings/phash.cpp of the reference, built
    in-tree as lib/phash.so), registered as the top-level module `phash` so
    that `import phash` works afterwards"""
    import importlib.machinery
    import importlib.util
    import os
    import sys

    if "phash" in sys.modules:
        return sys.modules["phash"]
    import torch  # noqa: F401  (phash links libtorch; load it first)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "phash.so")
    if not os.path.exists(path):
        raise ImportError("phash.so is not built (%s); run "
                          "`python -c 'import __graft_entry__ as g; g.build()'`" % path)
    loader = importlib.machinery.ExtensionFileLoader("phash", path)
    spec = importlib.util.spec_from_file_location("phash", path, loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    sys.modules["phash"] = mod
    return mod
