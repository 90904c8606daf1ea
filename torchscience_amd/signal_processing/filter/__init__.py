"""IIR filtering on MI355X: drop-ins for torchscience.signal_processing.filter
(src/torchscience/signal_processing/filter/__init__.py:9,77,93 exports
batched_filter_apply, lfilter, sosfilt)."""

from ._batched import batched_filter_apply
from ._lfilter import lfilter
from ._sosfilt import sosfilt
from ._zi import lfilter_zi, sosfilt_zi

__all__ = ["batched_filter_apply", "lfilter", "lfilter_zi", "sosfilt", "sosfilt_zi"]
