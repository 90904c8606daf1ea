"""Zero-phase filtering on MI355X: drop-ins for sosfiltfilt and filtfilt.

References: src/torchscience/signal_processing/filter/_sosfiltfilt.py:14-173
(edge padding _pad_signal :176-217) and _filtfilt.py:14-165 (:168-214).

sosfiltfilt makes two passes of the SOS kernel over the padded signal: the
forward pass from zi = sosfilt_zi(sos) * xp[0], then the backward pass in
the kernel's reverse-time mode (input read and output written from the end)
from zi = sosfilt_zi(sos) * y1[-1].  No flipped copies are made; the result
is the central T samples of the second pass.  When gradients are needed it
composes the differentiable sosfilt with flips, as the reference does.
filtfilt follows the reference's composition of lfilter passes.
"""

from __future__ import annotations

from typing import Literal, Optional

import torch
from torch import Tensor

from ._lfilter import lfilter
from ._ops import sosfilt_rows_forward
from ._plan import plan_sosfilt
from ._sosfilt import sosfilt
from ._zi import lfilter_zi, sosfilt_zi

PadType = Optional[Literal["odd", "even", "constant"]]


def _pad_signal(x: Tensor, padtype: str, padlen: int) -> Tensor:
    """Edge extension along the last axis (_sosfiltfilt.py:176-217)."""
    if padtype == "odd":
        left = 2 * x[..., 0:1] - x[..., 1:padlen + 1].flip(-1)
        right = 2 * x[..., -1:] - x[..., -padlen - 1:-1].flip(-1)
        return torch.cat([left, x, right], dim=-1)
    if padtype == "even":
        left = x[..., 1:padlen + 1].flip(-1)
        right = x[..., -padlen - 1:-1].flip(-1)
        return torch.cat([left, x, right], dim=-1)
    if padtype == "constant":
        left = x[..., 0:1].expand(*x.shape[:-1], padlen)
        right = x[..., -1:].expand(*x.shape[:-1], padlen)
        return torch.cat([left, x, right], dim=-1)
    raise ValueError(f"Unknown padtype: {padtype}")


def _needs_grad(*ts) -> bool:
    return torch.is_grad_enabled() and any(t is not None and t.requires_grad for t in ts)


def sosfiltfilt(sos: Tensor, x: Tensor, axis: int = -1, padtype: PadType = "odd",
                padlen: Optional[int] = None) -> Tensor:
    """Forward-backward (zero-phase) filtering with second-order sections;
    same signature, padding, defaults and errors as the reference."""
    if sos.ndim != 2 or sos.shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    S = sos.shape[0]
    if padlen is None:
        ntaps = 2 * S + 1
        ntaps -= min(int((sos[:, 2] == 0).sum().item()), int((sos[:, 5] == 0).sum().item()))
        padlen = 3 * ntaps
    xm = torch.moveaxis(x, axis, -1)
    T = xm.shape[-1]
    if padtype is not None and padlen >= T:
        raise ValueError(f"padlen ({padlen}) must be less than x.shape[axis] ({T})")
    xp = _pad_signal(xm, padtype, padlen) if padtype is not None and padlen > 0 else xm
    if S == 0:
        y = xp
    elif _needs_grad(sos, x) or xp.is_complex() or sos.is_complex():
        zi = sosfilt_zi(sos)
        zi_f = zi.unsqueeze(0) * xp[..., 0:1].unsqueeze(-1)
        if xp.ndim == 1:
            zi_f = zi_f.squeeze(0)
        y1, _ = sosfilt(sos, xp, axis=-1, zi=zi_f)
        yr = y1.flip(-1)
        zi_b = zi.unsqueeze(0) * yr[..., 0:1].unsqueeze(-1)
        if xp.ndim == 1:
            zi_b = zi_b.squeeze(0)
        y2, _ = sosfilt(sos, yr, axis=-1, zi=zi_b)
        y = y2.flip(-1)
    else:
        y = _sosfiltfilt_rows(sos, xp)
    if padtype is not None and padlen > 0:
        y = y[..., padlen:-padlen]
    return torch.moveaxis(y, -1, axis)


def _sosfiltfilt_rows(sos: Tensor, xp: Tensor) -> Tensor:
    """Both passes on the HIP kernel; xp is the padded signal, time last."""
    plan = plan_sosfilt(sos.shape, sos.dtype, xp.shape, xp.dtype, -1, None, None, None)
    S, Tp, B = plan.n_sections, plan.n_samples, plan.channels
    comp = plan.compute_dtype
    sos_c = sos.to(device=xp.device, dtype=comp).reshape(1, S, 6).contiguous()
    rows = xp.reshape(B, Tp).to(comp).contiguous()
    zi = sosfilt_zi(sos).to(device=xp.device, dtype=comp)
    zi_f = (zi.unsqueeze(0) * rows[:, 0:1].unsqueeze(-1)).contiguous()
    y1, _ = sosfilt_rows_forward(sos_c, rows, zi_f, channels=B, chans_per_filter=B,
                                 want_zf=False)
    zi_b = (zi.unsqueeze(0) * y1[:, -1:].unsqueeze(-1)).contiguous()
    y2, _ = sosfilt_rows_forward(sos_c, y1, zi_b, channels=B, chans_per_filter=B,
                                 want_zf=False, reverse=True)
    y = y2.reshape(plan.batch_shape + (Tp,))
    return y if y.dtype == plan.out_dtype else y.to(plan.out_dtype)


def filtfilt(b: Tensor, a: Tensor, x: Tensor, axis: int = -1, padtype: PadType = "odd",
             padlen: Optional[int] = None) -> Tensor:
    """Forward-backward (zero-phase) filtering with (b, a); same signature,
    padding, defaults and errors as the reference."""
    b = torch.atleast_1d(b)
    a = torch.atleast_1d(a)
    if b.ndim != 1:
        raise ValueError("b must be a 1-D tensor")
    if a.ndim != 1:
        raise ValueError("a must be a 1-D tensor")
    if bool(a[0] != 1.0):
        b = b / a[0]
        a = a / a[0]
    n_b, n_a = b.shape[0], a.shape[0]
    order = max(n_b, n_a) - 1
    if padlen is None:
        padlen = 3 * max(n_b, n_a)
    xm = torch.moveaxis(x, axis, -1)
    T = xm.shape[-1]
    if padtype is not None and padlen >= T:
        raise ValueError(f"padlen ({padlen}) must be less than x.shape[axis] ({T})")
    xp = _pad_signal(xm, padtype, padlen) if padtype is not None and padlen > 0 else xm
    if order > 0:
        zi = lfilter_zi(b, a)
        zi_f = zi.unsqueeze(0) * xp[..., 0:1]
        if xp.ndim == 1:
            zi_f = zi_f.squeeze(0)
        y1, _ = lfilter(b, a, xp, axis=-1, zi=zi_f)
        yr = y1.flip(-1)
        zi_b = zi.unsqueeze(0) * yr[..., 0:1]
        if xp.ndim == 1:
            zi_b = zi_b.squeeze(0)
        y2, _ = lfilter(b, a, yr, axis=-1, zi=zi_b)
        y = y2.flip(-1)
    else:
        y = b[0] * b[0] * xp
    if padtype is not None and padlen > 0:
        y = y[..., padlen:-padlen]
    return torch.moveaxis(y, -1, axis)
