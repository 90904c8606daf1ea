"""Impulse, step and frequency responses on MI355X.

Drop-ins for impulse_response / step_response / impulse_response_sos
(src/torchscience/signal_processing/filter/_impulse_response.py:11-223),
frequency_response_sos (_frequency_response_sos.py:12-140),
frequency_response (_frequency_response.py:10-118) and
frequency_response_fir (_frequency_response_fir.py:11-79).

The time responses filter an impulse or a step in FP64 through the same
kernels as lfilter / sosfilt.  The reference's per-section helper
(_impulse_response.py:226-271) ignores a section's a0 (it assumes a0 = 1),
so impulse_response_sos sets the a0 column to 1 before filtering.  The
frequency response is one HIP launch (tsiir_sos_frequency_response) over all
(filter, frequency) pairs; a batch of filters (F, S, 6) is accepted as well
as one cascade (S, 6).  frequency_response evaluates B(z^-1) / A(z^-1) by
Horner's rule in one launch (tsiir_ba_frequency_response).

Results come back on `device` (CPU by default for the time responses, as in
the reference), computed on the HIP device; there is no CPU path.
"""

from __future__ import annotations

import math
from typing import Optional, Union

import torch
from torch import Tensor

from ... import _lib


class FilterDesignError(Exception):
    """Base exception for filter design errors (_exceptions.py:4-7)."""


class SOSNormalizationError(FilterDesignError):
    """Second-order sections are not normalised (a0 != 1) (_exceptions.py:90-97)."""


def _hip_device(*tensors: Optional[Tensor]) -> torch.device:
    for t in tensors:
        if t is not None and t.device.type == "cuda":
            return t.device
    if not torch.cuda.is_available():
        raise RuntimeError("torchscience_amd filters run on MI355X (HIP) devices; none is "
                           "available")
    return torch.device("cuda", torch.cuda.current_device())


def _filter_signal(b: Tensor, a: Tensor, x: Tensor) -> Tensor:
    from ._lfilter import lfilter
    return lfilter(b, a, x)


def _time_response(numerator, denominator, signal, n_samples, dtype, device):
    if dtype is None:
        dtype = torch.float64
    if device is None:
        device = torch.device("cpu")
    dev = _hip_device(numerator, denominator)
    b = numerator.to(dtype=torch.float64, device=dev)
    a = torch.ones(1, dtype=torch.float64, device=dev) if denominator is None else \
        denominator.to(dtype=torch.float64, device=dev)
    if signal == "impulse":
        x = torch.zeros(n_samples, dtype=torch.float64, device=dev)
        if n_samples > 0:
            x[0] = 1.0
    else:
        x = torch.ones(n_samples, dtype=torch.float64, device=dev)
    response = _filter_signal(b, a, x)
    t = torch.arange(n_samples, dtype=dtype, device=device)
    return t, response.to(dtype).to(device)


def impulse_response(numerator: Tensor, denominator: Optional[Tensor] = None,
                     n_samples: int = 100, *, dtype: Optional[torch.dtype] = None,
                     device: Optional[torch.device] = None) -> tuple[Tensor, Tensor]:
    """(t, h): the first n_samples of the impulse response of b / a
    (_impulse_response.py:11-90); a defaults to [1] (FIR)."""
    return _time_response(numerator, denominator, "impulse", n_samples, dtype, device)


def step_response(numerator: Tensor, denominator: Optional[Tensor] = None,
                  n_samples: int = 100, *, dtype: Optional[torch.dtype] = None,
                  device: Optional[torch.device] = None) -> tuple[Tensor, Tensor]:
    """(t, s): the response to a unit step (_impulse_response.py:93-170)."""
    return _time_response(numerator, denominator, "step", n_samples, dtype, device)


def impulse_response_sos(sos: Tensor, n_samples: int = 100, *,
                         dtype: Optional[torch.dtype] = None,
                         device: Optional[torch.device] = None) -> tuple[Tensor, Tensor]:
    """(t, h) of the cascade sos (n_sections, 6) (_impulse_response.py:173-223),
    each section filtered with its a0 taken as 1, as the reference's helper
    does."""
    from ._sosfilt import sosfilt
    if dtype is None:
        dtype = torch.float64
    if device is None:
        device = torch.device("cpu")
    dev = _hip_device(sos)
    s = sos.to(dtype=torch.float64, device=dev).clone()
    x = torch.zeros(n_samples, dtype=torch.float64, device=dev)
    if n_samples > 0:
        x[0] = 1.0
    if s.shape[0] > 0 and n_samples > 0:
        s[:, 3] = 1.0
        response = sosfilt(s, x)
    else:
        response = x
    t = torch.arange(n_samples, dtype=dtype, device=device)
    return t, response.to(dtype).to(device)


def _grid(frequencies, whole, sampling_frequency, dev):
    """Frequency points (count: 0 up to, not including, Nyquist or the
    sampling frequency with whole) and their angular frequencies
    (_frequency_response_sos.py:91-116)."""
    if isinstance(frequencies, int):
        if whole:
            max_freq = 2.0 if sampling_frequency is None else sampling_frequency
        else:
            max_freq = 1.0 if sampling_frequency is None else sampling_frequency / 2.0
        freq_points = torch.linspace(0, max_freq, frequencies + 1, dtype=torch.float64,
                                     device=dev)[:-1]
    else:
        freq_points = frequencies.to(dtype=torch.float64, device=dev)
    if sampling_frequency is not None:
        w = 2 * math.pi * freq_points / sampling_frequency
    else:
        w = math.pi * freq_points
    return freq_points, w


@torch.compiler.disable
def frequency_response_sos(sos: Tensor, frequencies: Union[Tensor, int] = 512,
                           whole: bool = False, sampling_frequency: Optional[float] = None,
                           validate: bool = True, *, dtype: Optional[torch.dtype] = None,
                           device: Optional[torch.device] = None) -> tuple[Tensor, Tensor]:
    """(frequencies, H) of the cascade sos (n_sections, 6), or of each of a
    batch (n_filters, n_sections, 6) (then H is (n_filters, n_freq)).

    frequencies: a count of points from 0 up to (not including) Nyquist
    (the sampling frequency with whole=True), or the points themselves;
    normalised to Nyquist = 1 unless sampling_frequency is given (Hz).
    validate raises SOSNormalizationError when some a0 differs from 1 by
    more than 1e-10 (_frequency_response_sos.py:84-89)."""
    if device is None:
        device = sos.device
    if dtype is None:
        dtype = torch.complex128 if sos.dtype == torch.float64 else torch.complex64
    if validate and sos.numel() > 0:
        a0 = sos[..., 3]
        if not torch.allclose(a0, torch.ones_like(a0), atol=1e-10):
            raise SOSNormalizationError(
                f"SOS sections must have a0 = 1, got a0 values: {a0.tolist()}")
    dev = _hip_device(sos, frequencies if isinstance(frequencies, Tensor) else None)
    freq_points, w = _grid(frequencies, whole, sampling_frequency, dev)
    batched = sos.dim() == 3
    s = sos if batched else sos.reshape(1, -1, 6)
    coef = torch.float64 if s.dtype == torch.float64 else torch.float32
    s = s.to(device=dev, dtype=coef).contiguous()
    w = w.reshape(-1).contiguous()
    n_f, n_w = s.shape[0], w.numel()
    out = torch.empty((n_f, n_w, 2), dtype=torch.float64, device=dev)
    L = _lib.load()
    _lib.check(L.tsiir_sos_frequency_response(_lib.ptr(s), _lib.io_code(coef), n_f, s.shape[1],
                                              _lib.ptr(w), n_w, _lib.ptr(out),
                                              _lib.stream_handle(dev)),
               "tsiir_sos_frequency_response")
    h = torch.view_as_complex(out)
    h = h if batched else h.reshape(freq_points.shape)
    return freq_points.to(sos.dtype).to(device), h.to(dtype).to(device)


@torch.compiler.disable
def frequency_response(numerator: Tensor, denominator: Tensor,
                       frequencies: Union[Tensor, int] = 512, whole: bool = False,
                       sampling_frequency: Optional[float] = None, *,
                       dtype: Optional[torch.dtype] = None,
                       device: Optional[torch.device] = None) -> tuple[Tensor, Tensor]:
    """(frequencies, H) of the direct-form filter b / a: H = B(z^-1) /
    A(z^-1) at z = exp(jw), same frequency grid as frequency_response_sos
    (_frequency_response.py:10-118)."""
    if device is None:
        device = numerator.device
    if dtype is None:
        dtype = torch.complex128 if numerator.dtype == torch.float64 else torch.complex64
    if numerator.is_complex() or denominator.is_complex():
        raise NotImplementedError(
            "complex filter coefficients are not supported on the MI355X path")
    if numerator.dim() > 1 or denominator.dim() > 1:
        raise ValueError("numerator and denominator must be 1-D coefficient vectors, got "
                         f"shapes {tuple(numerator.shape)} and {tuple(denominator.shape)}")
    dev = _hip_device(numerator, denominator,
                      frequencies if isinstance(frequencies, Tensor) else None)
    freq_points, w = _grid(frequencies, whole, sampling_frequency, dev)
    b = numerator.to(device=dev, dtype=torch.float64).reshape(1, -1).contiguous()
    a = denominator.to(device=dev, dtype=torch.float64).reshape(1, -1).contiguous()
    w = w.reshape(-1).contiguous()
    out = torch.empty((1, w.numel(), 2), dtype=torch.float64, device=dev)
    L = _lib.load()
    _lib.check(L.tsiir_ba_frequency_response(_lib.ptr(b), b.shape[1], _lib.ptr(a), a.shape[1],
                                             1, _lib.ptr(w), w.numel(), _lib.ptr(out),
                                             _lib.stream_handle(dev)),
               "tsiir_ba_frequency_response")
    h = torch.view_as_complex(out).reshape(freq_points.shape)
    return freq_points.to(numerator.dtype).to(device), h.to(dtype).to(device)


def frequency_response_fir(coefficients: Tensor, frequencies: Union[Tensor, int] = 512,
                           whole: bool = False, sampling_frequency: Optional[float] = None,
                           *, dtype: Optional[torch.dtype] = None,
                           device: Optional[torch.device] = None) -> tuple[Tensor, Tensor]:
    """frequency_response with the denominator [1] (_frequency_response_fir.py:66-79)."""
    one = torch.ones(1, dtype=coefficients.dtype, device=coefficients.device)
    return frequency_response(coefficients, one, frequencies=frequencies, whole=whole,
                              sampling_frequency=sampling_frequency, dtype=dtype, device=device)
