"""MI355X-native coupled SPH + heightfield-wave step.

The compute path is the HIP library built from ``csrc/`` (libcwa.so, C ABI
in ``include/cwa.h``); this package is its Python binding and the scene
builders.  Importing it does not load the library; the first call that
needs it does, and fails loudly if it is missing.
"""
from ._lib import CwaError, Params, lib  # noqa: F401
from .scene import Scene, reference_params  # noqa: F401

__all__ = ["CwaError", "Params", "Scene", "lib", "reference_params"]
