"""Row-level calls into libtsiir and the autograd Function around them.

Inputs here are already planned: signals are (rows, T) contiguous on a HIP
device, sos is (n_filters, S, 6) contiguous, zi is (rows or 1, S, 2).  The
public functions in _sosfilt.py / _lfilter.py / _batched.py reshape the
reference's argument forms into these and back.
"""

from __future__ import annotations

import torch

from ... import _lib

# the operators tsiir::sosfilt_forward / sosfilt_backward (HIP and Meta
# implementations, libtsiir_torch.so), registered when this module loads
_lib.ops()


def _require_device(*tensors: torch.Tensor) -> torch.device:
    """The common device of the tensors: a HIP device, or "meta" (shape
    inference, tracing)."""
    dev = None
    for t in tensors:
        if t is None:
            continue
        if t.device.type not in ("cuda", "meta"):
            raise RuntimeError(
                "torchscience_amd filters run on MI355X (HIP) devices; got a tensor on "
                f"{t.device}. Move the inputs to a 'cuda' device.")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise RuntimeError(f"tensors on different devices: {dev} and {t.device}")
    return dev


def _desc(*, channels, n_samples, n_sections, io_dtype, coef_dtype, flags, x_rows,
          x_row_stride, y_row_stride, chans_per_filter, mode=_lib.MODE_PLAIN):
    d = _lib.SosfiltDesc()
    d.channels = channels
    d.n_samples = n_samples
    d.n_sections = n_sections
    d.io_dtype = _lib.io_code(io_dtype)
    d.coef_dtype = _lib.io_code(coef_dtype)
    d.mode = mode
    d.flags = flags
    d.reserved = 0
    d.x_rows = x_rows
    d.x_row_stride = x_row_stride
    d.y_row_stride = y_row_stride
    d.chans_per_filter = chans_per_filter
    return d


def sosfilt_rows_forward(sos: torch.Tensor, x: torch.Tensor, zi: torch.Tensor | None,
                         *, channels: int, chans_per_filter: int, want_zf: bool,
                         zi_broadcast: bool = False, io_dtype: torch.dtype | None = None,
                         mode: int = _lib.MODE_PLAIN, path: str | None = None,
                         keep_sections: bool = False, reverse: bool = False,
                         _debug_give_up: bool = False):
    """y (channels, T) and zf (channels, S, 2) for planned row inputs,
    through the operator tsiir::sosfilt_forward.

    Channel c reads x row c % x.shape[0] and filter c // chans_per_filter.
    path: None (chosen from the shape), "window" or "channel" (forces the
    work decomposition; results agree to rounding), or "lookback" (the
    window decomposition without the segmented form for few long rows).  keep_sections: also
    return the outputs of sections 0..S-2, (S-1, channels, T), for the
    backward (else None).  reverse: filter along decreasing time (input read
    and output written from the end; zi is the state at the end).
    _debug_give_up (tests only): every look-back wait gives up at once."""
    _require_device(sos, x, zi)
    io = io_dtype or x.dtype
    if x.dtype != io:
        x = x.to(io)
    flags = 0
    if path is not None:
        flags |= {"window": _lib.PATH_WINDOW, "channel": _lib.PATH_CHANNEL,
                  "lookback": _lib.PATH_WINDOW | _lib.NO_SEGMENT}[path]
    if reverse:
        flags |= _lib.REVERSE
    if _debug_give_up:
        flags |= _lib.DEBUG_GIVE_UP
    y, zf, saved = torch.ops.tsiir.sosfilt_forward(x, sos, zi, channels, chans_per_filter, want_zf,
                                              zi_broadcast, mode, flags, keep_sections)
    zf = zf if want_zf else None
    if keep_sections:
        return y, zf, saved
    return y, zf


def sosfilt_rows_backward(sos, x, zi, dy, dzf, *, chans_per_filter: int, y, saved,
                          zi_broadcast: bool = False):
    """(dx, dsos, dzi) through the operator tsiir::sosfilt_backward, from the
    forward output y and the section outputs `saved` that
    sosfilt_rows_forward(keep_sections=True) returned (no forward recompute);
    x has one row per channel."""
    _require_device(sos, x, zi, dy, dzf)
    return torch.ops.tsiir.sosfilt_backward(x, y, saved, sos, zi, dy, dzf, chans_per_filter,
                                       zi_broadcast)


class SosfiltRows(torch.autograd.Function):
    """Autograd over the operators.  The forward (tsiir::sosfilt_forward with
    keep_sections) keeps every section's output; the backward
    (tsiir::sosfilt_backward) runs the hand-written adjoint kernels
    (time-reversed section filters with coefficient-gradient accumulation)
    on them."""

    @staticmethod
    def forward(ctx, x, sos, zi, channels, chans_per_filter, want_zf, zi_broadcast):
        y, zf, saved = sosfilt_rows_forward(sos, x, zi, channels=channels,
                                            chans_per_filter=chans_per_filter, want_zf=want_zf,
                                            zi_broadcast=zi_broadcast, keep_sections=True)
        ctx.save_for_backward(x, sos, zi, y, saved)
        ctx.meta = (channels, chans_per_filter, zi_broadcast, want_zf)
        # an unused zf passes no gradient (None) rather than a filled zero tensor
        ctx.set_materialize_grads(False)
        if want_zf:
            return y, zf
        return y, None

    @staticmethod
    def backward(ctx, dy, dzf):
        x, sos, zi, y, saved = ctx.saved_tensors
        channels, cpf, zi_b, want_zf = ctx.meta
        if dy is None:
            dy = torch.zeros((channels, x.shape[1]), dtype=x.dtype, device=x.device)
        dy = dy.contiguous()
        if dzf is not None:
            dzf = dzf.contiguous().to(torch.float64 if x.dtype == torch.float64 else torch.float32)
        xr = x
        if x.shape[0] != channels:
            xr = x.repeat(channels // x.shape[0], 1) if channels % x.shape[0] == 0 else \
                x[torch.arange(channels, device=x.device) % x.shape[0]]
        dx, dsos, dzi = sosfilt_rows_backward(sos, xr.contiguous(), zi, dy, dzf,
                                              chans_per_filter=cpf, zi_broadcast=zi_b,
                                              y=y, saved=saved)
        gx = gsos = gzi = None
        if ctx.needs_input_grad[0]:
            gx = dx if x.shape[0] == channels else \
                dx.reshape(channels // x.shape[0], x.shape[0], -1).sum(0)
        if ctx.needs_input_grad[1]:
            gsos = dsos
        if zi is not None and ctx.needs_input_grad[2]:
            gzi = dzi.sum(0, keepdim=True).to(zi.dtype) if zi_b else dzi.to(zi.dtype)
        return gx, gsos, gzi, None, None, None, None
