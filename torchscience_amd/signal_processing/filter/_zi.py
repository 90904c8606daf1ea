"""Steady-state initial conditions (host-side set-up math, not the hot path).

lfilter_zi: src/torchscience/signal_processing/filter/_lfilter_zi.py:70-128
  solve (I - A^T) zi = b[1:] - a[1:] b[0], A the companion matrix of a.
sosfilt_zi: src/torchscience/signal_processing/filter/_sosfilt_zi.py:54-86
  per section lfilter_zi scaled by the DC gain of the preceding sections.
Both are small dense solves done with torch ops on the tensors' device.
"""

from __future__ import annotations

import torch
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
