"""IIR filtering on MI355X: drop-ins for torchscience.signal_processing.filter
(src/torchscience/signal_processing/filter/__init__.py exports
batched_filter_apply, fftfilt, filtfilt, lfilter, sosfilt, sosfiltfilt, the
zi helpers, the impulse / step / frequency responses and the batched IIR
designs)."""

from ._batched import batched_filter_apply
from ._batched_design import (batched_butterworth_design, batched_chebyshev_type_1_design,
                              batched_chebyshev_type_2_design)
from ._fir import fftfilt
from ._filtfilt import filtfilt, sosfiltfilt
from ._lfilter import lfilter
from ._response import (FilterDesignError, SOSNormalizationError, frequency_response,
                        frequency_response_fir, frequency_response_sos, impulse_response,
                        impulse_response_sos, step_response)
from ._sosfilt import sosfilt
from ._zi import lfilter_zi, lfiltic, sosfilt_zi

__all__ = ["FilterDesignError", "SOSNormalizationError", "batched_butterworth_design",
           "batched_chebyshev_type_1_design", "batched_chebyshev_type_2_design",
           "batched_filter_apply", "fftfilt",
           "filtfilt", "frequency_response", "frequency_response_fir", "frequency_response_sos", "impulse_response", "impulse_response_sos",
           "lfilter", "lfilter_zi", "lfiltic", "sosfilt", "sosfilt_zi", "sosfiltfilt", "step_response"]
