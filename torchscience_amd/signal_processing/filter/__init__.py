"""IIR filtering on MI355X: drop-ins for torchscience.signal_processing.filter
(src/torchscience/signal_processing/filter/__init__.py exports
batched_filter_apply, fftfilt, filtfilt, lfilter, sosfilt, sosfiltfilt and
the zi helpers)."""

from ._batched import batched_filter_apply
from ._fir import fftfilt
from ._filtfilt import filtfilt, sosfiltfilt
from ._lfilter import lfilter
from ._sosfilt import sosfilt
from ._zi import lfilter_zi, sosfilt_zi

__all__ = ["batched_filter_apply", "fftfilt", "filtfilt", "lfilter", "lfilter_zi", "sosfilt",
           "sosfilt_zi", "sosfiltfilt"]
