"""MI355X-native IIR filtering hot path of torchscience (sosfilt / lfilter).

Hand-written gfx950 HIP kernels behind the C ABI in include/tsiir.h, with a
thin Python layer that keeps the reference's operator signatures.
"""

__version__ = "0.1.0"

from . import signal_processing  # noqa: F401,E402
from ._lib import TsiirError, check_errors  # noqa: F401,E402
