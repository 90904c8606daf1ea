"""Host-side planning for the filtering ops: shapes, dtypes, broadcasting.

Pure shape/dtype logic (no device work), so it is tested on the CPU.  It
encodes the reference semantics the drop-in must keep:

* sos must be (n_sections, 6), else ValueError("sos must be shape
  (n_sections, 6)") (_sosfilt.py:84-86);
* the filtered axis is moved last and the rest flattened to channels
  (_sosfilt.py:91-93, _lfilter.py:120-127);
* the output dtype is torch.promote_types(coefficients, x)
  (_lfilter.py:131-133), applied once per section, so for sosfilt it is
  promote(sos.dtype, x.dtype);
* an (n_sections, 2) zi is broadcast over a batched x; otherwise zi carries
  the batch dims of x followed by (n_sections, 2) (_sosfilt.py:98-112);
* zf has zi's (broadcast) shape and zi's dtype (_sosfilt.py:114, 126).
"""

from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass(frozen=True)
class SosPlan:
    n_sections: int
    axis: int                      # normalised, non-negative
    batch_shape: tuple[int, ...]
    n_samples: int
    channels: int
    out_dtype: torch.dtype
    compute_dtype: torch.dtype     # float32 or float64
    io_dtype: torch.dtype          # dtype the kernel reads and writes
    complex_input: bool
    zi_broadcast: bool
    zf_shape: tuple[int, ...] | None
    zf_dtype: torch.dtype | None
    compensated: bool = False      # FP32 storage, FP64 recurrence updates


_REAL_FLOATS = (torch.float16, torch.bfloat16, torch.float32, torch.float64)


def _compute_dtype(out_dtype: torch.dtype) -> torch.dtype:
    if out_dtype in (torch.float64, torch.complex128):
        return torch.float64
    return torch.float32


def plan_sosfilt(sos_shape, sos_dtype: torch.dtype, x_shape, x_dtype: torch.dtype,
                 axis: int = -1, zi_shape=None, zi_dtype: torch.dtype | None = None,
                 precision: str | None = None) -> SosPlan:
    sos_shape = tuple(sos_shape)
    if len(sos_shape) != 2 or sos_shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    S = sos_shape[0]
    x_shape = tuple(x_shape)
    nd = len(x_shape)
    if nd == 0:
        raise ValueError("x must have at least one dimension")
    if not -nd <= axis < nd:
        raise IndexError(f"axis {axis} is out of bounds for x of dimension {nd}")
    ax = axis % nd
    batch = x_shape[:ax] + x_shape[ax + 1:]
    T = x_shape[ax]
    B = 1
    for d in batch:
        B *= d
    if sos_dtype.is_complex:
        raise NotImplementedError(
            "complex filter coefficients are not supported on the MI355X path")
    if not (sos_dtype.is_floating_point):
        raise TypeError(f"sos must be a floating point tensor, got {sos_dtype}")
    out = torch.promote_types(sos_dtype, x_dtype)
    complex_in = out.is_complex
    compute = _compute_dtype(out)
    io = compute
    if precision is not None:
        if precision not in ("fp32", "fp32c", "fp64", "bf16io", "bf16c"):
            raise ValueError(
                "precision must be 'fp32', 'fp32c', 'fp64', 'bf16io' or 'bf16c', "
                f"got {precision!r}")
        if complex_in:
            raise NotImplementedError("precision= is for real signals")
        if precision == "fp64":
            compute, io, out = torch.float64, torch.float64, torch.float64
        elif precision in ("fp32", "fp32c"):
            # a forced float32 form returns float32 also for float64 sos
            compute, io, out = torch.float32, torch.float32, torch.float32
        else:
            compute, io, out = torch.float32, torch.bfloat16, torch.bfloat16
    zi_b = False
    zf_shape = None
    if zi_shape is not None:
        zi_shape = tuple(zi_shape)
        if zi_shape == (S, 2) and len(batch) > 0:
            zi_b = True
            zf_shape = batch + (S, 2)
        else:
            if zi_shape != batch + (S, 2):
                raise RuntimeError(
                    f"zi must have shape {batch + (S, 2)} (or ({S}, 2)), got {zi_shape}")
            zf_shape = zi_shape
    return SosPlan(S, ax, batch, T, B, out, compute, io, complex_in, zi_b, zf_shape,
                   zi_dtype if zi_shape is not None else None,
                   precision in ("fp32c", "bf16c"))


def plan_lfilter(b_shape, a_shape):
    """1-D checks of lfilter (_lfilter.py:86-93) and its order."""
    b_shape, a_shape = tuple(b_shape), tuple(a_shape)
    if len(b_shape) == 0:
        b_shape = (1,)
    if len(a_shape) == 0:
        a_shape = (1,)
    if len(b_shape) != 1:
        raise ValueError("b must be a 1-D tensor")
    if len(a_shape) != 1:
        raise ValueError("a must be a 1-D tensor")
    return max(b_shape[0], a_shape[0]) - 1


def plan_lfilter_zi(zi_shape, batch_shape, order: int) -> bool:
    """Check an lfilter zi against the signal's batch shape; True when one
    state row serves every channel.

    The reference reads zi as ``zi.reshape(-1, order)`` rows, one per
    flattened channel (_lfilter.py:136-141), so a zi whose row count is
    neither the channel count nor 1 fails there; here it raises before any
    device work (a kernel would read past or short of the state rows)."""
    zi_shape = tuple(zi_shape)
    B = 1
    for d in batch_shape:
        B *= d
    numel = 1
    for d in zi_shape:
        numel *= d
    if order < 1 or len(zi_shape) == 0 or zi_shape[-1] != order or numel % order:
        raise RuntimeError(
            f"zi must have shape {tuple(batch_shape) + (order,)} (or ({order},)), "
            f"got {zi_shape}")
    rows = numel // order
    if rows != B and rows != 1:
        raise RuntimeError(
            f"zi must have shape {tuple(batch_shape) + (order,)} (or ({order},)), "
            f"got {zi_shape}: {rows} state rows for {B} channels")
    return rows == 1 and B > 1
