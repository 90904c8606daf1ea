"""Steady-state initial conditions (host-side set-up math, not the hot path).

lfilter_zi: src/torchscience/signal_processing/filter/_lfilter_zi.py:70-128
  solve (I - A^T) zi = b[1:] - a[1:] b[0], A the companion matrix of a.
sosfilt_zi: src/torchscience/signal_processing/filter/_sosfilt_zi.py:54-86
  per section lfilter_zi scaled by the DC gain of the preceding sections.
lfiltic: src/torchscience/signal_processing/filter/_lfilter.py:181-305
  the DF2T state that continues a filter from its past outputs y and inputs
  x (most recent first): zi[k] = sum_i b[k+i+1] x[i] - a[k+i+1] y[i].
All are small dense computations done with torch ops on the tensors' device;
sosfilt_zi_host is sosfilt_zi for coefficients already on the host.
"""

from __future__ import annotations

from typing import Optional

import torch
import torch.nn.functional as F
from torch import Tensor


def lfilter_zi(b: Tensor, a: Tensor) -> Tensor:
    b = torch.atleast_1d(b)
    if b.ndim != 1:
        raise ValueError("Numerator b must be 1-D.")
    a = torch.atleast_1d(a)
    if a.ndim != 1:
        raise ValueError("Denominator a must be 1-D.")
    while a.numel() > 1 and bool(a[0] == 0.0):
        a = a[1:]
    if a.numel() < 1:
        raise ValueError("There must be at least one nonzero `a` coefficient.")
    if bool(a[0] != 1.0):
        b = b / a[0]
        a = a / a[0]
    n = max(a.numel(), b.numel())
    if a.numel() < n:
        a = torch.cat([a, torch.zeros(n - a.numel(), dtype=a.dtype, device=a.device)])
    elif b.numel() < n:
        b = torch.cat([b, torch.zeros(n - b.numel(), dtype=b.dtype, device=b.device)])
    dt = torch.result_type(a, b)
    comp = torch.zeros((n - 1, n - 1), dtype=dt, device=a.device)
    if n > 1:
        comp[0, :] = -a[1:].to(dt)
        if n > 2:
            comp[1:, :-1] += torch.eye(n - 2, dtype=dt, device=a.device)
    m = torch.eye(n - 1, dtype=dt, device=a.device) - comp.T
    rhs = (b[1:] - a[1:] * b[0]).to(dt)
    return torch.linalg.solve(m, rhs)


def sosfilt_zi(sos: Tensor) -> Tensor:
    if sos.ndim != 2 or sos.shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    if sos.dtype in (torch.int8, torch.int16, torch.int32, torch.int64, torch.bool):
        sos = sos.to(torch.float64)
    n = sos.shape[0]
    zi = torch.empty((n, 2), dtype=sos.dtype, device=sos.device)
    scale = 1.0
    for s in range(n):
        b = sos[s, :3]
        a = sos[s, 3:]
        zi[s, :] = scale * lfilter_zi(b, a)
        scale = scale * torch.sum(b) / torch.sum(a)
    return zi


def sosfilt_zi_host(sos: Tensor) -> Tensor:
    """sosfilt_zi of host coefficients (S, 6), in FP64 (Python floats: the
    arrays are a few dozen values), as a CPU tensor.  Each section's system
    (I - A^T) zi = b[1:] - a[1:] b[0] (_lfilter_zi.py:100-128; for a section
    [[1 + a1, -1], [a2, 1]]) is solved in closed form, then scaled by the DC
    gains sum(b)/sum(a) of the sections before it (_sosfilt_zi.py:55-86).
    Sections with a0 == 0, a singular system or a zero sum(a) go through
    sosfilt_zi, with its errors and infinities."""
    if sos.ndim != 2 or sos.shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    c = sos.detach().to(device="cpu", dtype=torch.float64)
    out = []
    scale = 1.0
    for b0, b1, b2, a0, a1, a2 in c.tolist():
        sa = a0 + a1 + a2
        if a0 == 0.0 or sa == 0.0:
            return sosfilt_zi(c)
        n0, n1, n2, m1, m2 = b0 / a0, b1 / a0, b2 / a0, a1 / a0, a2 / a0
        det = (1.0 + m1) + m2
        if det == 0.0:
            return sosfilt_zi(c)
        r0, r1 = n1 - m1 * n0, n2 - m2 * n0
        out.append((scale * ((r0 + r1) / det), scale * (((1.0 + m1) * r1 - m2 * r0) / det)))
        scale = scale * ((b0 + b1 + b2) / sa)
    return torch.tensor(out, dtype=torch.float64).reshape(-1, 2)


def lfiltic(b: Tensor, a: Tensor, y: Tensor, x: Optional[Tensor] = None) -> Tensor:
    """Initial DF2T state (max(len(b), len(a)) - 1,) under which lfilter
    continues a filtering whose most recent outputs were y[0], y[1], ... and
    inputs x[0], x[1], ... (x zeros when None); short histories are padded
    with zeros (_lfilter.py:181-305).  b and a are normalised by a[0] when
    a[0] != 1.  zi[k] = sum_{i <= N-1-k} b[k+i+1] x[i] - a[k+i+1] y[i],
    formed as two Hankel products."""
    b = torch.atleast_1d(b)
    a = torch.atleast_1d(a)
    y = torch.atleast_1d(y)
    if b.ndim != 1:
        raise ValueError("b must be a 1-D tensor")
    if a.ndim != 1:
        raise ValueError("a must be a 1-D tensor")
    if y.ndim != 1:
        raise ValueError("y must be a 1-D tensor")
    dtype = torch.promote_types(b.dtype, torch.promote_types(a.dtype, y.dtype))
    if not dtype.is_floating_point:
        dtype = torch.float64
    device = y.device
    b = b.to(dtype=dtype, device=device)
    a = a.to(dtype=dtype, device=device)
    y = y.to(dtype=dtype)
    a0 = a[0]
    b = torch.where(a0 != 1, b / a0, b)
    a = torch.where(a0 != 1, a / a0, a)
    n_coef = max(b.shape[0], a.shape[0])
    N = n_coef - 1
    if N == 0:
        return torch.tensor([], dtype=dtype, device=device)
    b = F.pad(b, (0, n_coef - b.shape[0]))
    a = F.pad(a, (0, n_coef - a.shape[0]))
    x = torch.zeros(N, dtype=dtype, device=device) if x is None else \
        torch.atleast_1d(x).to(dtype=dtype, device=device)
    y = F.pad(y, (0, max(N - y.shape[0], 0)))[:N]
    x = F.pad(x, (0, max(N - x.shape[0], 0)))[:N]
    k = torch.arange(N, device=device)[:, None]
    i = torch.arange(N, device=device)[None, :]
    idx = k + i + 1
    mask = idx <= N
    zero = torch.zeros((), dtype=dtype, device=device)
    hb = torch.where(mask, b[idx.clamp(max=N)], zero)
    ha = torch.where(mask, a[idx.clamp(max=N)], zero)
    return hb @ x - ha @ y
