"""FIR filtering on MI355X: the row-level call into libtsiir, its autograd
Function, and fftfilt.

One direct-form kernel (tsiir_fir_forward) serves every FIR use of the
filtering path:

* lfilter(b, a, x, zi) with len(a) == 1 (_lfilter.py:96-166 with a = [a0];
  taps b / a0, zi / zf the DF2T state of order len(b) - 1),
* batched_filter_apply(filter_format="fir") (_batched.py:533-561),
* fftfilt(b, x) (_fftfilt.py:9-99): the 'same' slice of the full convolution
  (offset (K - 1) // 2).

Filters of FFT_MIN_TAPS..FFT_MAX_TAPS taps without state (one shared filter,
or one per equal group of channels) go through the overlap-save FFT kernel
instead
(tsiir_fir_forward_fft), whose cost per output does not grow with K; this
covers the forward and the input gradient (the same filter, reversed).

Rows here are planned already: signals (rows, T) contiguous in the compute
dtype on a HIP device, taps (n_filters, K) contiguous.  Channel c reads row
c % rows and tap row c // chans_per_filter.
"""

from __future__ import annotations

import math

import torch
from torch import Tensor

from ... import _lib
from ._ops import _require_device

# Tap counts routed to the FFT kernel; below FFT_MIN_TAPS the direct kernel
# is faster (measured on MI355X over 4096 x 65536 rows: the FFT kernel
# faster from 48 taps in FP32 and FP64, the direct kernel at 16; at 32 they
# are within 5%).
FFT_MIN_TAPS = 48
FFT_MAX_TAPS = 2048


def _fir_desc(*, channels, n_samples, x_len, n_taps, dtype, shift, flags, x_rows,
              x_row_stride, y_row_stride, chans_per_filter):
    d = _lib.FirDesc()
    d.channels = channels
    d.n_samples = n_samples
    d.x_len = x_len
    d.n_taps = n_taps
    d.io_dtype = _lib.io_code(dtype)
    d.shift = shift
    d.flags = flags
    d.x_rows = x_rows
    d.x_row_stride = x_row_stride
    d.y_row_stride = y_row_stride
    d.chans_per_filter = chans_per_filter
    return d


@torch.compiler.disable
def fir_rows_forward(taps: Tensor, x: Tensor, zi: Tensor | None, *, channels: int,
                     chans_per_filter: int, shift: int = 0, want_zf: bool = False,
                     zi_broadcast: bool = False, reverse_taps: bool = False,
                     n_out: int | None = None):
    """y (channels, n_out) and zf (channels, K - 1) (or None).

    y[c][n] = sum_j h[j] x[c % rows][n + shift - j] (x zero outside its row),
    plus zi[c][n] for n < K - 1; zf is the DF2T final state.  reverse_taps
    uses h[K - 1 - j].  n_out defaults to the input length."""
    dev = _require_device(x, taps, zi)
    K = taps.shape[-1]
    T_in = x.shape[-1]
    T = T_in if n_out is None else n_out
    dt = x.dtype
    y = torch.empty((channels, T), dtype=dt, device=dev)
    N = K - 1
    zf = torch.empty((channels, N), dtype=dt, device=dev) if want_zf else None
    flags = 0
    if zi is not None:
        flags |= _lib.HAS_ZI | (_lib.ZI_BROADCAST if zi_broadcast else 0)
    if want_zf:
        flags |= _lib.WRITE_ZF
    if reverse_taps:
        flags |= _lib.FIR_REVERSE_TAPS
    d = _fir_desc(channels=channels, n_samples=T, x_len=T_in, n_taps=K, dtype=dt, shift=shift,
                  flags=flags, x_rows=x.shape[0], x_row_stride=T_in, y_row_stride=T,
                  chans_per_filter=chans_per_filter)
    L = _lib.load()
    if (zi is None and not want_zf and FFT_MIN_TAPS <= K <= FFT_MAX_TAPS
            and (chans_per_filter >= channels or channels % chans_per_filter == 0)):
        nbytes = L.tsiir_fir_fft_workspace_bytes(d)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        _lib.check(L.tsiir_fir_forward_fft(d, _lib.ptr(taps), _lib.ptr(x), _lib.ptr(y),
                                           _lib.ptr(ws), nbytes, _lib.stream_handle(dev)),
                   "tsiir_fir_forward_fft")
        return y, zf
    _lib.check(L.tsiir_fir_forward(d, _lib.ptr(taps), _lib.ptr(x), _lib.ptr(y), _lib.ptr(zi),
                                   _lib.ptr(zf), _lib.stream_handle(dev)), "tsiir_fir_forward")
    return y, zf


@torch.compiler.disable
def fir_rows_backward_taps(taps_shape, x: Tensor, g: Tensor, *, channels: int,
                           chans_per_filter: int, shift: int = 0) -> Tensor:
    """dtaps[f][j] = sum over channels c of filter f, n < G of
    g[c][n] x[c % rows][n + shift - j]; g is (channels, G) contiguous."""
    dev = _require_device(x, g)
    F, K = taps_shape
    dt = x.dtype
    d = _fir_desc(channels=channels, n_samples=g.shape[-1], x_len=x.shape[-1], n_taps=K,
                  dtype=dt, shift=shift, flags=0, x_rows=x.shape[0],
                  x_row_stride=x.shape[-1], y_row_stride=g.shape[-1],
                  chans_per_filter=chans_per_filter)
    dtaps = torch.empty((F, K), dtype=dt, device=dev)
    L = _lib.load()
    nbytes = L.tsiir_fir_backward_taps_workspace_bytes(d)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    _lib.check(L.tsiir_fir_backward_taps(d, _lib.ptr(x), _lib.ptr(g), _lib.ptr(dtaps),
                                         _lib.ptr(ws), nbytes, _lib.stream_handle(dev)),
               "tsiir_fir_backward_taps")
    return dtaps


class FirRows(torch.autograd.Function):
    """Autograd over fir_rows_forward.  The output is linear in x, the taps
    and zi: dx is the FIR of the (zf-extended) output gradient with the taps
    reversed (the same kernel), dtaps its correlation with x
    (tsiir_fir_backward_taps), dzi its first K - 1 samples."""

    @staticmethod
    def forward(ctx, x, taps, zi, channels, chans_per_filter, shift, want_zf, zi_broadcast):
        y, zf = fir_rows_forward(taps, x, zi, channels=channels,
                                 chans_per_filter=chans_per_filter, shift=shift,
                                 want_zf=want_zf, zi_broadcast=zi_broadcast)
        ctx.save_for_backward(x, taps)
        ctx.meta = (channels, chans_per_filter, shift, want_zf, zi_broadcast, zi is not None)
        return y, zf

    @staticmethod
    def backward(ctx, dy, dzf):
        x, taps = ctx.saved_tensors
        channels, cpf, shift, want_zf, zi_b, has_zi = ctx.meta
        T = x.shape[-1]
        K = taps.shape[-1]
        N = K - 1
        if dy is None:
            dy = torch.zeros((channels, T), dtype=x.dtype, device=x.device)
        g = dy.to(x.dtype).contiguous()
        if want_zf and N > 0:
            dz = dzf if dzf is not None else torch.zeros((channels, N), dtype=x.dtype,
                                                         device=x.device)
            g = torch.cat([g, dz.to(x.dtype)], dim=-1).contiguous()
        gx = gtaps = gzi = None
        if ctx.needs_input_grad[0]:
            dx, _ = fir_rows_forward(taps, g, None, channels=channels, chans_per_filter=cpf,
                                     shift=K - 1 - shift, reverse_taps=True, n_out=T)
            gx = dx if x.shape[0] == channels else \
                dx.reshape(channels // x.shape[0], x.shape[0], T).sum(0)
        if ctx.needs_input_grad[1]:
            gtaps = fir_rows_backward_taps(tuple(taps.shape), x, g, channels=channels,
                                           chans_per_filter=cpf, shift=shift)
        if has_zi and ctx.needs_input_grad[2]:
            dzi = torch.zeros((channels, N), dtype=x.dtype, device=x.device)
            m = min(N, g.shape[-1])
            dzi[:, :m] = g[:, :m]
            gzi = dzi.sum(0, keepdim=True) if zi_b else dzi
        return gx, gtaps, gzi, None, None, None, None, None


def fir_rows(taps: Tensor, x: Tensor, zi: Tensor | None, *, channels: int,
             chans_per_filter: int, shift: int = 0, want_zf: bool = False,
             zi_broadcast: bool = False):
    """fir_rows_forward, through FirRows when a gradient is needed."""
    if torch.is_grad_enabled() and (x.requires_grad or taps.requires_grad
                                    or (zi is not None and zi.requires_grad)):
        return FirRows.apply(x, taps, zi, channels, chans_per_filter, shift, want_zf,
                             zi_broadcast)
    return fir_rows_forward(taps, x, zi, channels=channels, chans_per_filter=chans_per_filter,
                            shift=shift, want_zf=want_zf, zi_broadcast=zi_broadcast)


def _compute_dtype(dt: torch.dtype) -> torch.dtype:
    return torch.float64 if dt in (torch.float64, torch.complex128) else torch.float32


def fftfilt(b: Tensor, x: Tensor, axis: int = -1) -> Tensor:
    """FIR filter x along `axis` with taps b; output the shape of x.

    Equals scipy.signal.fftconvolve(x, b, mode='same') along the axis: the
    full convolution sliced from (len(b) - 1) // 2 (_fftfilt.py:60-92).
    Computed in FP32 (FP64 for float64 data): by overlap-save FFTs for
    FFT_MIN_TAPS..FFT_MAX_TAPS taps, directly otherwise.
    Differentiable with respect to b and x."""
    if b.is_complex():
        raise NotImplementedError(
            "complex filter coefficients are not supported on the MI355X path")
    xm = torch.movedim(x, axis, -1)
    out = torch.promote_types(b.dtype, x.dtype)
    if out.is_complex:
        xr = torch.view_as_real(xm.to(out))
        parts = [fftfilt(b, xr[..., k], axis=-1) for k in range(2)]
        return torch.movedim(torch.complex(parts[0], parts[1]), -1, axis)
    comp = _compute_dtype(out)
    shape = xm.shape
    T = shape[-1]
    K = b.shape[0]
    rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
    taps = b.to(device=x.device, dtype=comp).reshape(1, K).contiguous()
    B = rows.shape[0]
    if B == 0 or T == 0:
        y = rows.new_zeros((B, T))
    else:
        y, _ = fir_rows(taps, rows, None, channels=B, chans_per_filter=B, shift=(K - 1) // 2)
    return torch.movedim(y.reshape(shape), -1, axis).to(out)
