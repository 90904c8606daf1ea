"""Batched IIR design on MI355X: drop-ins for batched_butterworth_design,
batched_chebyshev_type_1_design and batched_chebyshev_type_2_design
(src/torchscience/signal_processing/filter/_batched.py:22-344).

The reference designs the filters one at a time in Python (prototype, band
transform, bilinear map and pairing on small tensors for every cutoff);
here one HIP launch designs the whole batch in FP64, one lane per filter
(tsiir_iir_design_sos), following the same chain and pairing rules.  The
argument checks and their messages follow the reference
(_batched.py:84-117, _butterworth_design.py:97-113, the prototypes' order /
ripple / attenuation checks, the "Invalid filter_type" check).

Cutoffs that require a gradient get one: the same launch writes the
Jacobian of every coefficient with respect to the filter's cutoffs
(forward-mode differentiation in the kernel), which the backward contracts
with the incoming gradient.

Difference: every step is FP64 (the reference builds its prototypes at
torch's default dtype unless dtype is given, which leaves float32-accurate
poles in the float64 result).

Under torch.compile the native call runs outside the traced graph (a graph
break), as it does for the other ctypes-bound kernels of this package.
"""

from __future__ import annotations

from typing import Optional

import torch
from torch import Tensor

from ... import _lib

_FAMILIES = {"butterworth": 0, "chebyshev1": 1, "chebyshev2": 2}
_BANDS = {"lowpass": 0, "highpass": 1, "bandpass": 2, "bandstop": 3}


def _cutoff_rows(cutoffs: Tensor, filter_type: str) -> Tensor:
    """(batch,) or (batch, 2) cutoffs, the reference's shape checks."""
    if cutoffs.numel() == 0:
        raise ValueError("cutoffs tensor cannot be empty")
    cutoffs = torch.atleast_1d(cutoffs)
    if filter_type in ("bandpass", "bandstop"):
        if cutoffs.ndim == 1:
            raise ValueError(f"For {filter_type} filters, cutoffs must have shape (batch, 2)")
        if cutoffs.shape[-1] != 2:
            raise ValueError(
                f"For {filter_type} filters, cutoffs must have shape (batch, 2), "
                f"got {cutoffs.shape}")
        return cutoffs
    if cutoffs.ndim == 2:
        if cutoffs.shape[-1] == 1:
            return cutoffs.squeeze(-1)
        raise ValueError(
            f"For {filter_type} filters, cutoffs must be 1D or (batch, 1), got {cutoffs.shape}")
    return cutoffs


def _check_range(c: Tensor, band: bool) -> None:
    """Every filter's cutoff(s) inside (0, 1) (_butterworth_design.py:102-113);
    one host transfer for the whole batch."""
    if band:
        ok = (c[:, 0] > 0) & (c[:, 0] < c[:, 1]) & (c[:, 1] < 1)
    else:
        ok = (c > 0) & (c < 1)
    bad = torch.nonzero(~ok).reshape(-1)
    if bad.numel():
        i = int(bad[0])
        if band:
            raise ValueError(
                f"Cutoff frequencies must satisfy 0 < low < high < 1, got {c[i].tolist()}")
        raise ValueError(
            f"Cutoff frequency must be between 0 and 1 (Nyquist), got {c[i].item()}")


@torch.compiler.disable
def _design(family: str, order: int, param: float, cutoffs: Tensor, filter_type: str,
            output: str, sampling_frequency: Optional[float], dtype, device) -> Tensor:
    if output != "sos":
        raise ValueError(f"Batched design only supports 'sos' output, got '{output}'")
    c = _cutoff_rows(cutoffs, filter_type)
    if device is None:
        device = cutoffs.device
    band = filter_type in ("bandpass", "bandstop")
    c = c.to(torch.float64)
    if sampling_frequency is not None:
        c = c / (sampling_frequency / 2.0)
    _check_range(c.detach(), band)
    if order < 1:
        raise ValueError(f"Filter order must be positive, got {order}")
    if family == "chebyshev1" and param <= 0:
        raise ValueError(f"Passband ripple must be positive, got {param}")
    if family == "chebyshev2" and param <= 0:
        raise ValueError(f"Stopband attenuation must be positive, got {param}")
    if filter_type not in _BANDS:
        raise ValueError(f"Invalid filter_type: {filter_type}")
    out_dtype = torch.float64 if dtype is None else dtype
    kdt = torch.float32 if out_dtype == torch.float32 else torch.float64
    dev = c.device if c.device.type == "cuda" else (
        device if torch.device(device).type == "cuda" else torch.device(
            "cuda", torch.cuda.current_device()))
    c = c.to(dev).contiguous()
    d = _lib.IirDesignDesc()
    d.n_filters = c.shape[0]
    d.family = _FAMILIES[family]
    d.band = _BANDS[filter_type]
    d.order = order
    d.out_dtype = _lib.io_code(kdt)
    d.param = float(param)
    L = _lib.load()
    n_sec = L.tsiir_iir_design_sections(d)
    if n_sec == 0:
        raise NotImplementedError(
            f"filter order {order} is above the largest order of the batched design (32)")
    if c.requires_grad and torch.is_grad_enabled():
        sos = _DesignWithGrad.apply(c, d, n_sec, band)
    else:
        sos = torch.empty((c.shape[0], n_sec, 6), dtype=kdt, device=dev)
        _lib.check(L.tsiir_iir_design_sos(d, _lib.ptr(c), _lib.ptr(sos),
                                          _lib.stream_handle(dev)), "tsiir_iir_design_sos")
    return sos.to(dtype=out_dtype, device=device)


class _DesignWithGrad(torch.autograd.Function):
    """FP64 design whose backward contracts the kernel's Jacobian
    d sos / d cutoffs with the incoming gradient."""

    @staticmethod
    def forward(ctx, c, d, n_sec, band):
        dev = c.device
        d64 = _lib.IirDesignDesc()
        for f in ("n_filters", "family", "band", "order", "param"):
            setattr(d64, f, getattr(d, f))
        d64.out_dtype = _lib.F64
        sos = torch.empty((c.shape[0], n_sec, 6), dtype=torch.float64, device=dev)
        jac = torch.empty((c.shape[0], n_sec, 6, 2), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().tsiir_iir_design_sos_jacobian(
            d64, _lib.ptr(c.detach().contiguous()), _lib.ptr(sos), _lib.ptr(jac),
            _lib.stream_handle(dev)), "tsiir_iir_design_sos_jacobian")
        ctx.save_for_backward(jac)
        ctx.band = band
        return sos

    @staticmethod
    def backward(ctx, g):
        (jac,) = ctx.saved_tensors
        gc = (g.to(jac.dtype).unsqueeze(-1) * jac).sum(dim=(1, 2))     # (n, 2)
        return (gc if ctx.band else gc[:, 0]), None, None, None


def batched_butterworth_design(order: int, cutoffs: Tensor, filter_type: str = "lowpass",
                               output: str = "sos",
                               sampling_frequency: Optional[float] = None, *,
                               dtype: Optional[torch.dtype] = None,
                               device: Optional[torch.device] = None) -> Tensor:
    """(batch, n_sections, 6) Butterworth designs, one per cutoff
    (_batched.py:22-140)."""
    return _design("butterworth", order, 0.0, cutoffs, filter_type, output, sampling_frequency,
                   dtype, device)


def batched_chebyshev_type_1_design(order: int, passband_ripple_db: float, cutoffs: Tensor,
                                    filter_type: str = "lowpass", output: str = "sos",
                                    sampling_frequency: Optional[float] = None, *,
                                    dtype: Optional[torch.dtype] = None,
                                    device: Optional[torch.device] = None) -> Tensor:
    """(batch, n_sections, 6) Chebyshev type I designs (_batched.py:143-242)."""
    return _design("chebyshev1", order, passband_ripple_db, cutoffs, filter_type, output,
                   sampling_frequency, dtype, device)


def batched_chebyshev_type_2_design(order: int, stopband_attenuation_db: float,
                                    cutoffs: Tensor, filter_type: str = "lowpass",
                                    output: str = "sos",
                                    sampling_frequency: Optional[float] = None, *,
                                    dtype: Optional[torch.dtype] = None,
                                    device: Optional[torch.device] = None) -> Tensor:
    """(batch, n_sections, 6) Chebyshev type II designs (_batched.py:245-344)."""
    return _design("chebyshev2", order, stopband_attenuation_db, cutoffs, filter_type, output,
                   sampling_frequency, dtype, device)
