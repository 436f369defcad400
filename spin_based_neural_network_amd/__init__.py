"""MI355X-native NQS variational Monte Carlo engine (libnqs_hip.so).

The product is the C-ABI shared library built from ``csrc/`` (hand-written
gfx950 HIP kernels plus the ``include/nqs`` entry points).  This package
holds the ctypes mirror of that ABI (``abi``) and the build recipe
(``build``); importing it loads nothing.
"""
from .abi import (NqsConfig, SrInfo, MarshallWrapper, HipOptions,  # noqa: F401
                  config_defaults, options_defaults, load_library, fnptr, ptr)

__all__ = ["NqsConfig", "SrInfo", "MarshallWrapper", "HipOptions",
           "config_defaults", "options_defaults", "load_library", "fnptr", "ptr"]
