"""PyTorch-op restatement of the reference's filtering loop (TEST
INFRASTRUCTURE ONLY: tests and bench.py's cpu_baseline leg use it).

Follows src/torchscience/signal_processing/filter/_lfilter.py:86-178 and
_sosfilt.py:84-133 op for op: a0 normalisation only when a[0] != 1, the
per-sample DF2T update as separate tensor ops vectorised over the batch,
sections in order.  Run on the CPU, its autograd tape is the reference's
backward (gradients reach x, b0..b2, a1, a2, zi; a0 gets 0 when a0 == 1).
``fftfilt_fft`` restates _fftfilt.py:60-92 (real FFTs of the next power of
two at or above the full length, product, inverse, 'same' slice) for the
gradients of fftfilt.
"""

from __future__ import annotations

import torch


def lfilter_loop(b, a, x, zi=None):
    """x (B, T); zi (B, N) or None.  Returns (y, zf)."""
    if bool(a[0] != 1.0):
        b = b / a[0]
        a = a / a[0]
    n = max(b.shape[0], a.shape[0])
    if b.shape[0] < n:
        b = torch.nn.functional.pad(b, (0, n - b.shape[0]))
    if a.shape[0] < n:
        a = torch.nn.functional.pad(a, (0, n - a.shape[0]))
    order = n - 1
    out_dtype = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    B, T = x.shape
    z = zi.to(out_dtype).clone() if zi is not None else torch.zeros(B, order, dtype=out_dtype)
    y = torch.zeros(B, T, dtype=out_dtype)
    for t in range(T):
        xn = x[:, t]
        yn = b[0] * xn + z[:, 0]
        y[:, t] = yn
        for k in range(order - 1):
            z[:, k] = b[k + 1] * xn - a[k + 1] * yn + z[:, k + 1]
        z[:, order - 1] = b[order] * xn - a[order] * yn
    return y, z


def sosfilt_loop(sos, x, zi=None):
    """sos (S, 6), x (B, T), zi (B, S, 2) or None.  Returns (y, zf)."""
    y = x.clone()
    zfs = []
    for i in range(sos.shape[0]):
        y, zf = lfilter_loop(sos[i, :3], sos[i, 3:], y, None if zi is None else zi[:, i, :])
        zfs.append(zf)
    return y, torch.stack(zfs, dim=1)


def sosfilt_loop_timed(sos, x, n_samples):
    """The reference loop over the first n_samples of x (CPU baseline)."""
    with torch.no_grad():
        return sosfilt_loop(sos, x[:, :n_samples])[0]


def fftfilt_fft(b, x):
    """x (B, T) real, b (K,) real: the reference's FFT convolution, sliced."""
    T = x.shape[-1]
    K = b.shape[0]
    n_fft = 1
    while n_fft < T + K - 1:
        n_fft *= 2
    y_full = torch.fft.irfft(torch.fft.rfft(b, n=n_fft).unsqueeze(0)
                             * torch.fft.rfft(x, n=n_fft), n=n_fft)
    start = (K - 1) // 2
    return y_full[:, start:start + T]
