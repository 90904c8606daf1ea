"""sosfilt on MI355X: drop-in for torchscience.signal_processing.filter.sosfilt.

Reference: src/torchscience/signal_processing/filter/_sosfilt.py:13-133.
Same signature, return arity ((y) or (y, zf) iff zi is given), ValueError
text, zi broadcasting, dtype promotion and autograd reach (x, sos, zi).
The cascade runs in one kernel pass per call instead of S x T Python
iterations; gradients come from the adjoint kernels.
"""

from __future__ import annotations

from typing import Optional, Union

import torch
from torch import Tensor

from ... import _lib
from ._ops import SosfiltRows, sosfilt_rows_forward
from ._plan import plan_sosfilt


def _to_rows(x: Tensor, ax: int, B: int, T: int) -> Tensor:
    return torch.movedim(x, ax, -1).reshape(B, T)


def sosfilt(
    sos: Tensor,
    x: Tensor,
    axis: int = -1,
    zi: Optional[Tensor] = None,
    *,
    precision: Optional[str] = None,
) -> Union[Tensor, tuple[Tensor, Tensor]]:
    """Filter x along `axis` with cascaded second-order sections.

    sos is (n_sections, 6) with rows [b0, b1, b2, a0, a1, a2]; each row is
    normalised by a0 when a0 != 1.  zi, if given, is (..., n_sections, 2)
    with x's batch dims, or (n_sections, 2) for every channel; then
    (y, zf) is returned, zf with zi's dtype.

    ``precision`` (keyword-only, not in the reference) selects the kernel
    form: None follows the reference's dtype promotion; "fp32" / "fp64"
    force the compute and result type; "fp32c" keeps float32 I/O and inter-section
    storage but runs every recurrence update in float64 with float64
    coefficients (the compensated form for poles near the unit circle) and
    returns float32; "bf16io" reads and writes bfloat16
    with float32 arithmetic and returns bfloat16; "bf16c" reads and writes
    bfloat16 with float64 recurrence updates (the compensated form with
    bfloat16 I/O).  Float32 recurrences of filters with a pole beyond radius
    0.995 run in the DF2T update order with float64 carries.
    """
    if sos.is_complex():
        return _sosfilt_complex_coef(sos, x, axis, zi, precision)
    plan = plan_sosfilt(sos.shape, sos.dtype, x.shape, x.dtype, axis,
                        None if zi is None else zi.shape,
                        None if zi is None else zi.dtype, precision)
    S, T, B = plan.n_sections, plan.n_samples, plan.channels
    comp = plan.compute_dtype
    dev = x.device
    if S == 0:
        # no sections: y = x.clone() and an (empty) zf, as _sosfilt.py:96,114
        y = x.clone()
        if zi is None:
            return y
        return y, torch.zeros(plan.zf_shape, dtype=zi.dtype, device=zi.device)
    # the compensated form reads its coefficients in float64 (exact for
    # float32 sos; keeps float64-designed filters at full precision)
    cdt = torch.float64 if plan.compensated else comp
    sos_c = sos.to(device=dev, dtype=cdt).reshape(1, S, 6).contiguous()

    if plan.complex_input:
        # Real coefficients act on the real and imaginary parts separately.
        xc = _to_rows(x.to(plan.out_dtype), plan.axis, B, T)
        xr = torch.view_as_real(xc).permute(0, 2, 1).reshape(2 * B, T)
        zr = None
        if zi is not None:
            zc = zi.to(device=dev, dtype=plan.out_dtype)
            zc = zc.expand(plan.zf_shape) if plan.zi_broadcast else zc
            zr = torch.view_as_real(zc.reshape(B, S, 2)).permute(0, 3, 1, 2).reshape(2 * B, S, 2)
            zr = zr.to(comp).contiguous()
        y2, zf2 = _run(sos_c, xr.to(comp).contiguous(), zr, 2 * B, comp, comp, zi is not None,
                       False)
        y = torch.view_as_complex(y2.reshape(B, 2, T).permute(0, 2, 1).contiguous())
        y = torch.movedim(y.reshape(plan.batch_shape + (T,)), -1, plan.axis)
        if zi is None:
            return y
        zf = torch.view_as_complex(zf2.reshape(B, 2, S, 2).permute(0, 2, 3, 1).contiguous())
        return y, zf.reshape(plan.zf_shape).to(zi.dtype)

    xr = _to_rows(x, plan.axis, B, T).to(plan.io_dtype).contiguous()
    zr = None
    if zi is not None:
        # the compensated forms carry a float64 zi as it is (zf then keeps
        # float64, as zf has zi's dtype); otherwise the state is in the
        # compute type
        sdt = torch.float64 if (plan.compensated and zi.dtype == torch.float64) else comp
        zr = zi.to(device=dev, dtype=sdt)
        zr = zr.reshape(1, S, 2) if plan.zi_broadcast else zr.reshape(B, S, 2)
        zr = zr.contiguous()
    y, zf = _run(sos_c, xr, zr, B, comp, plan.io_dtype, zi is not None, plan.zi_broadcast,
                 plan.compensated)
    y = torch.movedim(y.reshape(plan.batch_shape + (T,)), -1, plan.axis)
    if y.dtype != plan.out_dtype:
        y = y.to(plan.out_dtype)
    if zi is None:
        return y
    return y, zf.reshape(plan.zf_shape).to(zi.dtype)


def _sosfilt_complex_coef(sos, x, axis, zi, precision):
    """Complex sections: the reference's cascade of lfilter calls, one per
    section (_sosfilt.py:116-128), each through the complex-coefficient
    lfilter (real kernels on the conjugate-extended sections).  zf has zi's
    dtype, as in the reference, whose zf is allocated like zi
    (_sosfilt.py:114) so a real zi keeps only the real part of the state."""
    from ._lfilter import lfilter
    if sos.dim() != 2 or sos.shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    if precision is not None:
        raise NotImplementedError("precision= is for real filter coefficients")
    S = sos.shape[0]
    nd = x.dim()
    if nd == 0:
        raise ValueError("x must have at least one dimension")
    ax = axis % nd
    batch = tuple(x.shape[:ax] + x.shape[ax + 1:])
    zs = None
    if zi is not None:
        if tuple(zi.shape) == (S, 2) and len(batch) > 0:
            zs = zi.expand(batch + (S, 2))
        elif tuple(zi.shape) == batch + (S, 2):
            zs = zi
        else:
            raise RuntimeError(
                f"zi must have shape {batch + (S, 2)} (or ({S}, 2)), got {tuple(zi.shape)}")
    y = x.clone()
    zfs = []
    for i in range(S):
        if zs is None:
            y = lfilter(sos[i, :3], sos[i, 3:], y, axis=axis)
        else:
            y, zf_i = lfilter(sos[i, :3], sos[i, 3:], y, axis=axis, zi=zs[..., i, :])
            zfs.append(zf_i)
    if zi is None:
        return y
    if S == 0:
        return y, torch.zeros(batch + (0, 2), dtype=zi.dtype, device=zi.device)
    zf = torch.stack(zfs, dim=-2)
    if not zi.dtype.is_complex:
        zf = zf.real
    return y, zf.to(zi.dtype)


def _run(sos_c, xr, zr, B, comp, io, want_zf, zi_b, compensated=False):
    needs_grad = torch.is_grad_enabled() and (
        sos_c.requires_grad or xr.requires_grad or (zr is not None and zr.requires_grad))
    if needs_grad and io == comp and not compensated:
        y, zf = SosfiltRows.apply(xr, sos_c, zr, B, B, want_zf, zi_b)
        return y, zf
    if needs_grad:
        raise NotImplementedError(
            "gradients are available for the default, 'fp32' and 'fp64' precisions")
    return sosfilt_rows_forward(sos_c, xr, zr, channels=B, chans_per_filter=B,
                                want_zf=want_zf, zi_broadcast=zi_b, io_dtype=io,
                                mode=_lib.MODE_COMPENSATED if compensated else _lib.MODE_PLAIN)
