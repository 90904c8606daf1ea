"""MI355X-native IIR filtering hot path of torchscience (sosfilt / lfilter).

Hand-written gfx950 HIP kernels behind the C ABI in include/tsiir.h, with a
thin Python layer that keeps the reference's operator signatures.

Importing the package loads nothing native (so the build step can import
`torchscience_amd.build` before any library exists); the filtering modules
load libtsiir.so and the operator library when they are imported, and raise
if they have not been built.
"""

import importlib

__version__ = "0.1.0"


def __getattr__(name):
    if name == "signal_processing":
        return importlib.import_module(".signal_processing", __name__)
    if name in ("TsiirError", "check_errors"):
        from . import _lib
        return getattr(_lib, name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
