"""lfilter on MI355X: drop-in for torchscience.signal_processing.filter.lfilter.

Reference: src/torchscience/signal_processing/filter/_lfilter.py:12-178.

Filters of order <= 2 (max(len(b), len(a)) <= 3) are one second-order
section and run through the single-pass SOS kernel, with autograd.  Higher
orders with a single denominator coefficient (FIR, a = [a0]) run the
direct-form FIR kernel, with autograd.  Other higher orders run the
general-order DF2T kernel (one lane per channel, serial in time), which has
no backward.

zi is (batch..., N) with x's batch dimensions in order (the documented
meaning; the reference's moveaxis of zi, _lfilter.py:137, is not followed),
or (N,) for every channel; zf has the same layout, in the output dtype as
the reference returns it (_lfilter.py:139-175).
"""

from __future__ import annotations

import math

from typing import Optional, Union

import torch
import torch.nn.functional as F
from torch import Tensor

from ... import _lib
from ._fir import fir_rows
from ._ops import _desc, _require_device
from ._plan import plan_lfilter
from ._sosfilt import sosfilt


def lfilter(
    b: Tensor,
    a: Tensor,
    x: Tensor,
    axis: int = -1,
    zi: Optional[Tensor] = None,
) -> Union[Tensor, tuple[Tensor, Tensor]]:
    """Filter x along `axis` with the IIR/FIR filter (b, a) in DF2T form.

    Normalises by a[0] when a[0] != 1; pads b and a to a common length;
    returns y, or (y, zf) when zi (shape (..., order)) is given."""
    b = torch.atleast_1d(b)
    a = torch.atleast_1d(a)
    order = plan_lfilter(b.shape, a.shape)
    if b.is_complex() or a.is_complex():
        raise NotImplementedError(
            "complex filter coefficients are not supported on the MI355X path")
    if order <= 2:
        return _lfilter_as_section(b, a, x, axis, zi, order)
    if a.shape[0] == 1:
        return _lfilter_fir(b, a, x, axis, zi, order)
    return _lfilter_general(b, a, x, axis, zi, order)


def _complex_split(b, a, x, axis, zi, out):
    """Real filter, complex signal: filter the real and imaginary parts."""
    ax = axis % x.dim()
    xr = torch.view_as_real(torch.movedim(x.to(out), ax, -1))
    zr = None if zi is None else torch.view_as_real(zi.to(out))
    parts, zfs = [], []
    for k in range(2):
        r = lfilter(b, a, torch.movedim(xr[..., k], -1, ax), axis=axis,
                    zi=None if zi is None else zr[..., k])
        if zi is None:
            parts.append(r)
        else:
            parts.append(r[0])
            zfs.append(r[1])
    y = torch.complex(parts[0], parts[1])
    if zi is None:
        return y
    return y, torch.complex(zfs[0], zfs[1])


def _state_rows(zi, order, dev, comp):
    """zi (batch..., N) -> (rows, N) contiguous (_lfilter.py:139-141)."""
    return zi.to(device=dev, dtype=comp).reshape(-1, order).contiguous()


def _state_out(zf, batch, order, out):
    """(B, N) final state -> (batch..., N) in the output dtype."""
    return zf.reshape(batch + (order,) if batch else (order,)).to(out)


def _lfilter_fir(b, a, x, axis, zi, order):
    dev = _require_device(x)
    ax = axis % x.dim()
    out = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    if out.is_complex:
        return _complex_split(b, a, x, axis, zi, out)
    comp = torch.float64 if out == torch.float64 else torch.float32
    a0 = a.to(device=dev, dtype=comp)[0]
    bb = b.to(device=dev, dtype=comp)
    # b / a0 only when a0 != 1, so that a0 gets no gradient then (_lfilter.py:96-98)
    taps = torch.where(a0 != 1, bb / a0, bb).reshape(1, -1).contiguous()
    xm = torch.movedim(x, ax, -1)
    batch = xm.shape[:-1]
    T = xm.shape[-1]
    rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
    B = rows.shape[0]
    zi_r = None
    if zi is not None:
        zi_r = _state_rows(zi, order, dev, comp)
    if B == 0:
        y = rows.new_empty((0, T))
        zf = rows.new_empty((0, order))
    else:
        y, zf = fir_rows(taps, rows, zi_r, channels=B, chans_per_filter=B,
                         want_zf=zi is not None,
                         zi_broadcast=zi_r is not None and zi_r.shape[0] == 1 and B > 1)
    y = torch.movedim(y.reshape(batch + (T,)), -1, ax).to(out)
    if zi is None:
        return y
    return y, _state_out(zf, batch, order, out)


def _lfilter_as_section(b, a, x, axis, zi, order):
    dt = torch.promote_types(b.dtype, a.dtype)
    b3 = F.pad(b.to(dt), (0, 3 - b.shape[0]))
    a3 = F.pad(a.to(dt), (0, 3 - a.shape[0]))
    sos = torch.cat([b3, a3]).reshape(1, 6)
    if zi is None:
        return sosfilt(sos, x, axis=axis)
    zi_in = zi
    if order < 2:
        # order-0/1 state padded with the (always zero) second delay
        pad = torch.zeros(zi.shape[:-1] + (2 - order,), dtype=zi.dtype, device=zi.device)
        zi_in = torch.cat([zi, pad], dim=-1)
    zi_s = zi_in.unsqueeze(-2)  # (..., 1, 2)
    y, zf = sosfilt(sos, x, axis=axis, zi=zi_s)
    zf = zf.squeeze(-2)
    if order == 0:
        return y, torch.zeros_like(zi)
    out = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    return y, zf[..., :order].reshape(zi.shape).to(out)


def _lfilter_general(b, a, x, axis, zi, order):
    if torch.is_grad_enabled() and (b.requires_grad or a.requires_grad or x.requires_grad
                                    or (zi is not None and zi.requires_grad)):
        raise NotImplementedError(
            "gradients of lfilter are available for filters of order <= 2")
    dev = _require_device(x)
    nd = x.dim()
    ax = axis % nd
    out = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    if out.is_complex:
        return _complex_split(b, a, x, axis, zi, out)
    comp = torch.float64 if out == torch.float64 else torch.float32
    xm = torch.movedim(x, ax, -1)
    batch = xm.shape[:-1]
    T = xm.shape[-1]
    rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
    B = rows.shape[0]
    n = order + 1
    coef = torch.cat([F.pad(b.to(comp), (0, n - b.shape[0])),
                      F.pad(a.to(comp), (0, n - a.shape[0]))]).to(dev).contiguous()
    y = torch.empty_like(rows)
    zi_r = None
    flags = 0
    if zi is not None:
        zi_r = _state_rows(zi, order, dev, comp)
        if zi_r.shape[0] == 1 and B > 1:
            flags |= _lib.ZI_BROADCAST
        flags |= _lib.HAS_ZI | _lib.WRITE_ZF
    zf = torch.empty((B, order), dtype=comp, device=dev) if zi is not None else None
    d = _desc(channels=B, n_samples=T, n_sections=order, io_dtype=comp, coef_dtype=comp,
              flags=flags, x_rows=B, x_row_stride=T, y_row_stride=T, chans_per_filter=B)
    L = _lib.load()
    _lib.check(L.tsiir_lfilter_forward(d, _lib.ptr(coef), _lib.ptr(rows), _lib.ptr(y),
                                       _lib.ptr(zi_r), _lib.ptr(zf), _lib.stream_handle(dev)),
               "tsiir_lfilter_forward")
    y = torch.movedim(y.reshape(batch + (T,)), -1, ax).to(out)
    if zi is None:
        return y
    return y, _state_out(zf, batch, order, out)
