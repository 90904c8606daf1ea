"""batched_filter_apply on MI355X.

Reference: src/torchscience/signal_processing/filter/_batched.py:449-566,
which loops over the filters in Python and calls sosfilt / lfilter once per
filter.  Here every (filter, signal) pair is one channel of a single kernel
launch: the C ABI maps output channel c to input row c % x_rows and filter
c // chans_per_filter, so

* broadcast=True  (output (F, *x.shape)): channel c = f * B + b reads x row b
  with filter f (chans_per_filter = B, x_rows = B);
* broadcast=False (output (F, *x.shape[1:])): channel c = f * B' + b reads
  row c of x with filter f (chans_per_filter = B', x_rows = F * B').

The FIR format (filters (F, taps), a = [1], _batched.py:533-561) maps the
same way onto one launch of the direct-form FIR kernel.
"""

from __future__ import annotations

import math

from typing import Literal

import torch
from torch import Tensor

from ._fir import fir_rows
from ._ops import SosfiltRows, sosfilt_rows_forward


def batched_filter_apply(
    filters: Tensor,
    x: Tensor,
    filter_format: Literal["sos", "fir"] = "sos",
    broadcast: bool = True,
    axis: int = -1,
) -> Tensor:
    n_filters = filters.shape[0]
    if filter_format == "sos":
        if filters.ndim != 3 or filters.shape[2] != 6:
            raise ValueError(
                f"SOS filters must have shape (batch, n_sections, 6), got {filters.shape}")
        if broadcast:
            return _apply_sos(filters, x, axis, broadcast=True)
        if x.shape[0] != n_filters:
            raise ValueError(
                f"When broadcast=False, x batch dim must match filters: "
                f"{x.shape[0]} != {n_filters}")
        return _apply_sos(filters, x, axis, broadcast=False)
    if filter_format == "fir":
        if filters.ndim != 2:
            raise ValueError(
                f"FIR filters must have shape (batch, num_taps), got {filters.shape}")
        if not broadcast and x.shape[0] != n_filters:
            raise ValueError(
                f"When broadcast=False, x batch dim must match filters: "
                f"{x.shape[0]} != {n_filters}")
        return _apply_fir(filters, x, axis, broadcast)
    raise ValueError(f"filter_format must be 'sos' or 'fir', got '{filter_format}'")


def _apply_sos(filters: Tensor, x: Tensor, axis: int, broadcast: bool) -> Tensor:
    F, S = filters.shape[0], filters.shape[1]
    sig = x if broadcast else x[0]          # shape of one filter's input
    nd = sig.dim()
    ax = axis % nd
    out = torch.promote_types(filters.dtype, x.dtype)
    if out.is_complex or not out.is_floating_point:
        # rare forms: one filter at a time through sosfilt
        from ._sosfilt import sosfilt
        if broadcast:
            return torch.stack([sosfilt(filters[i], x, axis=axis) for i in range(F)])
        return torch.stack([sosfilt(filters[i], x[i], axis=axis) for i in range(F)])
    comp = torch.float64 if out == torch.float64 else torch.float32
    T = sig.shape[ax]
    if broadcast:
        xm = torch.movedim(x, ax, -1)
        rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
        Bsig = rows.shape[0]
        channels, cpf = F * Bsig, Bsig
        out_shape = (F,) + tuple(xm.shape)
    else:
        xm = torch.movedim(x, ax + 1, -1)
        rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
        Bsig = rows.shape[0] // F
        channels, cpf = F * Bsig, Bsig
        out_shape = tuple(xm.shape)
    sos = filters.to(device=x.device, dtype=comp).contiguous()
    if channels == 0 or S == 0:
        y = rows.new_empty((channels, T)) if S > 0 else rows.repeat(channels // max(rows.shape[0], 1), 1)
    elif torch.is_grad_enabled() and (sos.requires_grad or rows.requires_grad):
        y, _ = SosfiltRows.apply(rows, sos, None, channels, cpf, False, False)
    else:
        y, _ = sosfilt_rows_forward(sos, rows, None, channels=channels, chans_per_filter=cpf,
                                    want_zf=False)
    y = y.reshape(out_shape)
    y = torch.movedim(y, -1, ax + 1)
    return y.to(out)


def _rows_of(x: Tensor, axis: int, broadcast: bool, F: int, comp: torch.dtype):
    """(rows, T) contiguous signal rows, channels per filter and output shape."""
    sig = x if broadcast else x[0]
    ax = axis % sig.dim()
    T = sig.shape[ax]
    if broadcast:
        xm = torch.movedim(x, ax, -1)
        rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
        return rows, rows.shape[0], (F,) + tuple(xm.shape), ax
    xm = torch.movedim(x, ax + 1, -1)
    rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
    return rows, rows.shape[0] // F, tuple(xm.shape), ax


def _apply_fir(filters: Tensor, x: Tensor, axis: int, broadcast: bool) -> Tensor:
    F = filters.shape[0]
    out = torch.promote_types(filters.dtype, x.dtype)
    if out.is_complex or not out.is_floating_point:
        from ._lfilter import lfilter
        one = torch.ones(1, dtype=filters.dtype, device=filters.device)
        if broadcast:
            return torch.stack([lfilter(filters[i], one, x, axis=axis) for i in range(F)])
        return torch.stack([lfilter(filters[i], one, x[i], axis=axis) for i in range(F)])
    comp = torch.float64 if out == torch.float64 else torch.float32
    rows, cpf, out_shape, ax = _rows_of(x, axis, broadcast, F, comp)
    taps = filters.to(device=x.device, dtype=comp).contiguous()
    channels = F * cpf
    if channels == 0 or rows.shape[-1] == 0:
        y = rows.new_zeros((channels, rows.shape[-1]))
    else:
        y, _ = fir_rows(taps, rows, None, channels=channels, chans_per_filter=cpf)
    y = torch.movedim(y.reshape(out_shape), -1, ax + 1)
    return y.to(out)
