"""Zero-phase filtering on MI355X: drop-ins for sosfiltfilt and filtfilt.

References: src/torchscience/signal_processing/filter/_sosfiltfilt.py:14-173
(edge padding _pad_signal :176-217) and _filtfilt.py:14-165 (:168-214).

sosfiltfilt makes two passes of the SOS kernel over the padded signal
xp = [left, x, right]: the forward pass from zi = sosfilt_zi(sos) * xp[0],
then the backward pass in the kernel's reverse-time mode (input read and
output written from the end) from zi = sosfilt_zi(sos) * y1[-1]; the result
is the central T samples of the second pass.  The padded signal is never
built: filtering a concatenation is filtering its pieces with the state
carried from one to the next, so the forward pass runs over the left
extension (for its final state only), x and the right extension, and the
backward pass over the right extension (for its state) and the central part,
the only output kept.  The operator tsiir::sosfiltfilt (C ABI
tsiir_sosfiltfilt) does this in one call: a thread-per-channel kernel per
extension and the two sosfilt passes over the rows; beyond its bounds (more
than 8 sections, padlen x sections above 16384) the same pieces run as
separate sosfilt calls.  No padded or flipped copies are made.  When
gradients are needed it composes the differentiable sosfilt with flips, as
the reference does.  filtfilt follows the reference's composition of lfilter
passes.
"""

from __future__ import annotations

from typing import Literal, Optional

import torch
from torch import Tensor

from ... import _lib
from ._lfilter import lfilter
from ._ops import _require_device, sosfilt_rows_forward
from ._plan import plan_sosfilt
from ._sosfilt import sosfilt
from ._zi import lfilter_zi, sosfilt_zi, sosfilt_zi_host

PadType = Optional[Literal["odd", "even", "constant"]]


def _extensions(x: Tensor, padtype: str, padlen: int) -> tuple[Tensor, Tensor]:
    """The left and right edge extensions along the last axis, padlen
    samples each (_sosfiltfilt.py:176-217)."""
    if padtype == "odd":
        left = 2 * x[..., 0:1] - x[..., 1:padlen + 1].flip(-1)
        right = 2 * x[..., -1:] - x[..., -padlen - 1:-1].flip(-1)
        return left, right
    if padtype == "even":
        return x[..., 1:padlen + 1].flip(-1), x[..., -padlen - 1:-1].flip(-1)
    if padtype == "constant":
        return (x[..., 0:1].expand(*x.shape[:-1], padlen),
                x[..., -1:].expand(*x.shape[:-1], padlen))
    raise ValueError(f"Unknown padtype: {padtype}")


def _pad_signal(x: Tensor, padtype: str, padlen: int) -> Tensor:
    """[left extension, x, right extension] along the last axis."""
    left, right = _extensions(x, padtype, padlen)
    return torch.cat([left, x, right], dim=-1)


def _needs_grad(*ts) -> bool:
    return torch.is_grad_enabled() and any(t is not None and t.requires_grad for t in ts)


def sosfiltfilt(sos: Tensor, x: Tensor, axis: int = -1, padtype: PadType = "odd",
                padlen: Optional[int] = None) -> Tensor:
    """Forward-backward (zero-phase) filtering with second-order sections;
    same signature, padding, defaults and errors as the reference."""
    if sos.ndim != 2 or sos.shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    S = sos.shape[0]
    fast = S > 0 and not (_needs_grad(sos, x) or x.is_complex() or sos.is_complex())
    xm = torch.moveaxis(x, axis, -1)
    T = xm.shape[-1]
    dev_pad = _device_padlen(sos, xm, padtype, padlen) if fast else None
    if dev_pad is not None:
        return torch.moveaxis(_sosfiltfilt_rows(sos, None, xm, padtype, dev_pad), -1, axis)
    # Otherwise the kernel form takes the coefficients to the host once: the default
    # padlen and sosfilt_zi are small host computations there, where on the
    # device each of their scalar tests and solves would wait for the device.
    sos_h = sos.detach().cpu() if fast else sos
    if padlen is None:
        ntaps = 2 * S + 1
        if fast:
            c = sos_h.to(torch.float64).tolist()
            ntaps -= min(sum(r[2] == 0 for r in c), sum(r[5] == 0 for r in c))
        else:
            ntaps -= min(int((sos[:, 2] == 0).sum().item()), int((sos[:, 5] == 0).sum().item()))
        padlen = 3 * ntaps
    if padtype is not None and padlen >= T:
        raise ValueError(f"padlen ({padlen}) must be less than x.shape[axis] ({T})")
    # the reference pads (and trims) only for padlen > 0
    pad = max(padlen, 0) if padtype is not None else 0
    if pad > 0 and padtype not in ("odd", "even", "constant"):
        raise ValueError(f"Unknown padtype: {padtype}")
    if fast:
        return torch.moveaxis(_sosfiltfilt_rows(sos, sos_h, xm, padtype, pad), -1, axis)
    xp = _pad_signal(xm, padtype, padlen) if pad > 0 else xm
    if S == 0:
        y = xp
    else:
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
    if pad > 0:
        y = y[..., padlen:-padlen]
    return torch.moveaxis(y, -1, axis)


def _device_padlen(sos: Tensor, x: Tensor, padtype: PadType,
                   padlen: Optional[int]) -> Optional[int]:
    """The padlen argument of the operator when it can form the default
    padlen and sosfilt_zi from coefficients on the device (-1: the default),
    so that the call never waits for the device; None where a host copy of
    the coefficients is needed: sos or x not on the device, beyond the
    operator's bounds, or a reference error that depends on the values
    (padlen >= T for some default padlen, 3 (2 S + 1) >= T).  A filter with
    no steady state then gives NaN rows and a TsiirError from the stream's
    next call (or check_errors) where the reference raises at once."""
    S, T = sos.shape[0], x.shape[-1]
    if not (sos.is_cuda and x.is_cuda) or S > _lib.SOSFILTFILT_MAX_SECTIONS:
        return None
    if padtype is None:
        return 0
    if padlen is None:
        pad, need = -1, 3 * (2 * S + 1)
    else:
        pad = need = max(padlen, 0)
    if need >= T or need * S > _lib.SOSFILTFILT_MAX_WORK:
        return None
    if pad != 0 and padtype not in ("odd", "even", "constant"):
        raise ValueError(f"Unknown padtype: {padtype}")
    return pad


def _sosfiltfilt_rows(sos: Tensor, sos_h: Optional[Tensor], x: Tensor, padtype: PadType,
                      padlen: int) -> Tensor:
    """Both passes on the HIP kernel over [left, x, right] (padlen samples
    of each extension, none when padlen is 0), time last; returns the
    central T samples of the second pass.  sos_h: the coefficients on the
    host, or None for the operator's device-side sosfilt_zi (and padlen -1
    for its default padlen, within the operator's bounds)."""
    plan = plan_sosfilt(sos.shape, sos.dtype, x.shape, x.dtype, -1, None, None, None)
    S, T, B = plan.n_sections, plan.n_samples, plan.channels
    comp = plan.compute_dtype
    sos_c = sos.to(device=x.device, dtype=comp).reshape(1, S, 6).contiguous()
    rows = x.reshape(B, T).to(comp)
    if rows.stride(-1) != 1:
        rows = rows.contiguous()
    if sos_h is None or (S <= _lib.SOSFILTFILT_MAX_SECTIONS
                         and padlen * S <= _lib.SOSFILTFILT_MAX_WORK):
        # one operator: edge kernels plus the two passes over the rows
        zu = None
        if sos_h is not None:
            zu = sosfilt_zi_host(sos_h).reshape(1, S, 2)
            zu = zu.pin_memory().to(x.device, non_blocking=True) if x.is_cuda else zu.to(x.device)
        pad = {None: _lib.PAD_NONE, "odd": _lib.PAD_ODD, "even": _lib.PAD_EVEN,
               "constant": _lib.PAD_CONSTANT}[padtype if padlen != 0 else None]
        _require_device(sos_c, rows)
        _lib.ops()
        y = torch.ops.tsiir.sosfiltfilt(rows, sos_c, zu, pad, padlen)
        y = y.reshape(plan.batch_shape + (T,))
        return y if y.dtype == plan.out_dtype else y.to(plan.out_dtype)
    zi = sosfilt_zi_host(sos_h).to(comp)
    zi = zi.pin_memory().to(x.device, non_blocking=True) if x.is_cuda else zi.to(x.device)

    def run(inp, z, want_zf, reverse=False):
        return sosfilt_rows_forward(sos_c, inp, z.contiguous(), channels=B, chans_per_filter=B,
                                    want_zf=want_zf, reverse=reverse)

    if padlen > 0:
        left, right = (e.contiguous() for e in _extensions(rows, padtype, padlen))
        _, z = run(left, zi.unsqueeze(0) * left[:, :1].unsqueeze(-1), True)
        y1, z = run(rows, z, True)
        y1r, _ = run(right, z, False)
        _, z = run(y1r, zi.unsqueeze(0) * y1r[:, -1:].unsqueeze(-1), True, reverse=True)
    else:
        y1, _ = run(rows, zi.unsqueeze(0) * rows[:, :1].unsqueeze(-1), False)
        z = zi.unsqueeze(0) * y1[:, -1:].unsqueeze(-1)
    y2, _ = run(y1, z, False, reverse=True)
    y = y2.reshape(plan.batch_shape + (T,))
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
    if (max(b.shape[0], a.shape[0]) >= 2 and b.is_floating_point() and
            a.is_floating_point() and not (_needs_grad(b, a, x) or x.is_complex())):
        return _filtfilt_kernel_form(b, a, x, axis, padtype, padlen)
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


def _filtfilt_kernel_form(b: Tensor, a: Tensor, x: Tensor, axis: int, padtype: PadType,
                          padlen: Optional[int]) -> Tensor:
    """filtfilt of order >= 1 without gradients, its pieces on the device.
    b and a come to the host once and are normalised by a[0] there
    (_filtfilt.py:93-95); padlen defaults to 3 * max(len(a), len(b)).
    Orders 1 and 2: lfilter runs the filter as one second-order section, so
    this is sosfiltfilt's kernel form with the section [b, a] (zi =
    lfilter_zi(b, a) is sosfilt_zi of the section).  Higher orders: the same
    pieces through the general-order kernel (tsiir_lfilter_forward_ws), the
    backward passes in its reverse-time mode."""
    b_h, a_h = b.detach().cpu(), a.detach().cpu()
    if bool(a_h[0] != 1.0):
        b_h = b_h / a_h[0]
        a_h = a_h / a_h[0]
    n_b, n_a = b_h.shape[0], a_h.shape[0]
    if padlen is None:
        padlen = 3 * max(n_b, n_a)
    xm = torch.moveaxis(x, axis, -1)
    T = xm.shape[-1]
    if padtype is not None and padlen >= T:
        raise ValueError(f"padlen ({padlen}) must be less than x.shape[axis] ({T})")
    pad = padlen if padtype is not None else 0
    if pad > 0 and padtype not in ("odd", "even", "constant"):
        raise ValueError(f"Unknown padtype: {padtype}")
    if max(n_b, n_a) <= 3:
        sos_h = torch.zeros(1, 6, dtype=torch.promote_types(b_h.dtype, a_h.dtype))
        sos_h[0, :n_b] = b_h
        sos_h[0, 3:3 + n_a] = a_h
        sos = _to_device(sos_h, x.device)
        return torch.moveaxis(_sosfiltfilt_rows(sos, sos_h, xm, padtype, pad), -1, axis)
    return torch.moveaxis(_filtfilt_general_rows(b_h, a_h, xm, padtype, pad), -1, axis)


def _to_device(t: Tensor, dev: torch.device) -> Tensor:
    """A small host tensor on dev, through pinned memory (no synchronisation)."""
    return t.pin_memory().to(dev, non_blocking=True) if dev.type == "cuda" else t.to(dev)


def _filtfilt_general_rows(b_h: Tensor, a_h: Tensor, x: Tensor, padtype: PadType,
                           padlen: int) -> Tensor:
    """Both passes of filtfilt (order >= 3) over [left, x, right] with the
    general-order kernel, time last, as _sosfiltfilt_rows does for sections:
    the left extension for its state, the rows, the right extension forward
    and then backward for the backward pass's state, the rows backward."""
    from ._lfilter import _general_rows
    _require_device(x)
    out = torch.promote_types(torch.promote_types(b_h.dtype, a_h.dtype), x.dtype)
    comp = torch.float64 if out == torch.float64 else torch.float32
    n = max(b_h.shape[0], a_h.shape[0])
    pad_to = lambda v: torch.nn.functional.pad(v.to(torch.float64), (0, n - v.shape[0]))
    coef = _to_device(torch.cat([pad_to(b_h), pad_to(a_h)]).to(comp), x.device)
    zi = _to_device(lfilter_zi(b_h.to(torch.float64), a_h.to(torch.float64)).to(comp),
                    x.device)
    batch, T = x.shape[:-1], x.shape[-1]
    rows = x.reshape(-1, T).to(comp).contiguous()

    def run(inp, z, want_zf, reverse=False):
        return _general_rows(coef, inp, z.contiguous(), want_zf, False, reverse=reverse)

    if padlen > 0:
        left, right = (e.contiguous() for e in _extensions(rows, padtype, padlen))
        _, z = run(left, zi.unsqueeze(0) * left[:, :1], True)
        y1, z = run(rows, z, True)
        y1r, _ = run(right, z, False)
        _, z = run(y1r, zi.unsqueeze(0) * y1r[:, -1:], True, reverse=True)
    else:
        y1, _ = run(rows, zi.unsqueeze(0) * rows[:, :1], False)
        z = zi.unsqueeze(0) * y1[:, -1:]
    y2, _ = run(y1, z, False, reverse=True)
    return y2.reshape(batch + (T,)).to(out)
