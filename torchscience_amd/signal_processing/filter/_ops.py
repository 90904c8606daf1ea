"""Row-level calls into libtsiir and the autograd Function around them.

Inputs here are already planned: signals are (rows, T) contiguous on a HIP
device, sos is (n_filters, S, 6) contiguous, zi is (rows or 1, S, 2).  The
public functions in _sosfilt.py / _lfilter.py / _batched.py reshape the
reference's argument forms into these and back.
"""

from __future__ import annotations

import torch

from ... import _lib
from ._plan import fp32_needs_df2t


def _require_device(*tensors: torch.Tensor) -> torch.device:
    dev = None
    for t in tensors:
        if t is None:
            continue
        if t.device.type != "cuda":
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
    """y (channels, T) and zf (channels, S, 2) for planned row inputs.

    Channel c reads x row c % x.shape[0] and filter c // chans_per_filter.
    path: None (chosen from the shape), "window" or "channel" (forces the
    work decomposition; results agree to rounding).  keep_sections: also
    return the outputs of sections 0..S-2, (S-1, channels, T), for the
    backward (else None).  reverse: filter along decreasing time (input read
    and output written from the end; zi is the state at the end).
    _debug_give_up (tests only): every look-back wait gives up at once."""
    dev = _require_device(sos, x, zi)
    L = _lib.load()
    io = io_dtype or x.dtype
    T = x.shape[1]
    S = sos.shape[1]
    acc = torch.float64 if io == torch.float64 else torch.float32
    y = torch.empty((channels, T), dtype=io, device=dev)
    zf = torch.empty((channels, S, 2), dtype=acc, device=dev) if want_zf else None
    flags = 0
    if zi is not None:
        flags |= _lib.HAS_ZI
        if zi_broadcast:
            flags |= _lib.ZI_BROADCAST
    if want_zf:
        flags |= _lib.WRITE_ZF
    if path is not None:
        flags |= {"window": _lib.PATH_WINDOW, "channel": _lib.PATH_CHANNEL}[path]
    if reverse:
        flags |= _lib.REVERSE
    if io != torch.float64 and mode == _lib.MODE_PLAIN and fp32_needs_df2t(sos):
        flags |= _lib.FP32_DF2T
    if _debug_give_up:
        flags |= _lib.DEBUG_GIVE_UP
    d = _desc(channels=channels, n_samples=T, n_sections=S, io_dtype=io, coef_dtype=sos.dtype,
              flags=flags, x_rows=max(x.shape[0], 1),
              x_row_stride=x.stride(0) if x.dim() == 2 else T, y_row_stride=T,
              chans_per_filter=max(chans_per_filter, 1), mode=mode)
    nbytes = L.tsiir_sosfilt_workspace_bytes(d)
    if nbytes == 0:
        raise _lib.TsiirError("invalid sosfilt descriptor")
    ws = _lib.workspaces.get(dev, nbytes, "fwd")
    st = _lib.stream_handle(dev)
    if keep_sections:
        saved = torch.empty((max(S - 1, 0), channels, T), dtype=io, device=dev)
        _lib.check(L.tsiir_sosfilt_forward_saved(d, _lib.ptr(x), _lib.ptr(y), _lib.ptr(sos),
                                                 _lib.ptr(zi), _lib.ptr(zf),
                                                 _lib.ptr(saved) if S > 1 else None,
                                                 ws.data_ptr(), ws.numel(), st),
                   "tsiir_sosfilt_forward_saved")
        return y, zf, saved
    _lib.check(L.tsiir_sosfilt_forward(d, _lib.ptr(x), _lib.ptr(y), _lib.ptr(sos), _lib.ptr(zi),
                                       _lib.ptr(zf), ws.data_ptr(), ws.numel(), st),
               "tsiir_sosfilt_forward")
    return y, zf


def sosfilt_rows_backward(sos, x, zi, dy, dzf, *, chans_per_filter: int,
                          zi_broadcast: bool = False, y=None, saved=None):
    """(dx, dsos, dzi) for sosfilt_rows_forward with x rows == channels.

    With y (the forward output) and saved (the section outputs that
    sosfilt_rows_forward(keep_sections=True) returned) the forward is not
    recomputed."""
    dev = _require_device(sos, x, zi, dy, dzf)
    L = _lib.load()
    B, T = dy.shape
    S = sos.shape[1]
    acc = torch.float64 if x.dtype == torch.float64 else torch.float32
    flags = 0
    if zi is not None:
        flags |= _lib.HAS_ZI
        if zi_broadcast:
            flags |= _lib.ZI_BROADCAST
    if dzf is not None:
        flags |= _lib.WRITE_ZF
    d = _desc(channels=B, n_samples=T, n_sections=S, io_dtype=x.dtype, coef_dtype=sos.dtype,
              flags=flags, x_rows=max(B, 1), x_row_stride=T, y_row_stride=T,
              chans_per_filter=max(chans_per_filter, 1))
    use_saved = y is not None
    nbytes = (L.tsiir_sosfilt_backward_saved_workspace_bytes(d) if use_saved
              else L.tsiir_sosfilt_backward_workspace_bytes(d))
    if nbytes == 0:
        raise _lib.TsiirError("invalid sosfilt backward descriptor")
    ws = _lib.workspaces.get(dev, nbytes, "bwd")
    dx = torch.empty_like(dy)
    dsos = torch.empty_like(sos)
    dzi = torch.empty((B, S, 2), dtype=acc, device=dev)
    st = _lib.stream_handle(dev)
    if use_saved:
        _lib.check(L.tsiir_sosfilt_backward_saved(
            d, _lib.ptr(x), _lib.ptr(y), _lib.ptr(saved) if S > 1 else None, _lib.ptr(sos),
            _lib.ptr(zi), _lib.ptr(dy), _lib.ptr(dzf), _lib.ptr(dx), _lib.ptr(dsos),
            _lib.ptr(dzi), ws.data_ptr(), ws.numel(), st), "tsiir_sosfilt_backward_saved")
        return dx, dsos, dzi
    _lib.check(L.tsiir_sosfilt_backward(d, _lib.ptr(x), _lib.ptr(sos), _lib.ptr(zi),
                                        _lib.ptr(dy), _lib.ptr(dzf), _lib.ptr(dx),
                                        _lib.ptr(dsos), _lib.ptr(dzi), ws.data_ptr(),
                                        ws.numel(), st),
               "tsiir_sosfilt_backward")
    return dx, dsos, dzi


class SosfiltRows(torch.autograd.Function):
    """Autograd over sosfilt_rows_forward.  The forward keeps every section's
    output (tsiir_sosfilt_forward_saved); backward runs the hand-written
    adjoint kernels (time-reversed section filters with coefficient-gradient
    accumulation) on them through tsiir_sosfilt_backward_saved."""

    @staticmethod
    def forward(ctx, x, sos, zi, channels, chans_per_filter, want_zf, zi_broadcast):
        y, zf, saved = sosfilt_rows_forward(sos, x, zi, channels=channels,
                                            chans_per_filter=chans_per_filter, want_zf=want_zf,
                                            zi_broadcast=zi_broadcast, keep_sections=True)
        ctx.save_for_backward(x, sos, zi, y, saved)
        ctx.meta = (channels, chans_per_filter, zi_broadcast, want_zf)
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
