"""lfilter on MI355X: drop-in for torchscience.signal_processing.filter.lfilter.

Reference: src/torchscience/signal_processing/filter/_lfilter.py:12-178.

Filters of order <= 2 (max(len(b), len(a)) <= 3) are one second-order
section and run through the single-pass SOS kernel, with autograd.  Higher
orders with a single denominator coefficient (FIR, a = [a0]) run the
direct-form FIR kernel, with autograd.  Other higher orders run the
general-order DF2T kernel (time-parallel chunks up to order 16, one lane
per channel above), with an autograd whose backward reuses that kernel
(reverse-time all-pole pass), the FIR kernel and the tap-gradient kernel.
Complex coefficients run as four of these real filters on the
conjugate-extended polynomials (see _lfilter_complex_coef).

zi is (batch..., N) with x's batch dimensions in order (the documented
meaning; the reference's moveaxis of zi, _lfilter.py:137, is not followed),
or (N,) for every channel; zf has the same layout, in the output dtype as
the reference returns it (_lfilter.py:139-175).
"""

from __future__ import annotations

import math

from typing import Optional, Union

import torch
import torch.nn.functional as F
from torch import Tensor

from ... import _lib
from ._fir import fir_rows
from ._ops import _desc, _require_device
from ._plan import plan_lfilter, plan_lfilter_zi
from ._sosfilt import sosfilt


def lfilter(
    b: Tensor,
    a: Tensor,
    x: Tensor,
    axis: int = -1,
    zi: Optional[Tensor] = None,
) -> Union[Tensor, tuple[Tensor, Tensor]]:
    """Filter x along `axis` with the IIR/FIR filter (b, a) in DF2T form.

    Normalises by a[0] when a[0] != 1; pads b and a to a common length;
    returns y, or (y, zf) when zi (shape (..., order)) is given."""
    b = torch.atleast_1d(b)
    a = torch.atleast_1d(a)
    order = plan_lfilter(b.shape, a.shape)
    if b.is_complex() or a.is_complex():
        return _lfilter_complex_coef(b, a, x, axis, zi, order)
    if order <= 2:
        return _lfilter_as_section(b, a, x, axis, zi, order)
    if a.shape[0] == 1:
        return _lfilter_fir(b, a, x, axis, zi, order)
    return _lfilter_general(b, a, x, axis, zi, order)


def _complex_split(b, a, x, axis, zi, out):
    """Real filter, complex signal: filter the real and imaginary parts."""
    ax = axis % x.dim()
    xr = torch.view_as_real(torch.movedim(x.to(out), ax, -1))
    zr = None if zi is None else torch.view_as_real(zi.to(out))
    parts, zfs = [], []
    for k in range(2):
        r = lfilter(b, a, torch.movedim(xr[..., k], -1, ax), axis=axis,
                    zi=None if zi is None else zr[..., k])
        if zi is None:
            parts.append(r)
        else:
            parts.append(r[0])
            zfs.append(r[1])
    y = torch.complex(parts[0], parts[1])
    if zi is None:
        return y
    return y, torch.complex(zfs[0], zfs[1])


def _polymul(p, q):
    """Coefficients of P(z) Q(z) for rows p (..., n) and one q (m,):
    (..., n + m - 1), as sums of shifted products (differentiable)."""
    n, m = p.shape[-1], q.shape[-1]
    return sum(F.pad(p * q[j], (j, m - 1 - j)) for j in range(m))


def _lfilter_complex_coef(b, a, x, axis, zi, order):
    """lfilter with complex coefficients, through the real kernels.

    With B(z), A(z) the normalised polynomials (a0 = 1) and Z(z) the
    polynomial of zi's N state values, the DF2T output is
    Y = (B X + Z) / A.  Multiplying above and below by A~(z), A's
    polynomial with conjugated coefficients, gives a real denominator
    A A~ (its coefficient of z^-m, sum_k a_k conj(a_(k-m')), equals its own
    conjugate) and the numerator (B A~) X + Z A~.  So, with c = B A~ and
    s = Z A~ split into real and imaginary parts and d = A A~,
      Re y = lfilter(Re c, d, Re x, zi = Re s) - lfilter(Im c, d, Im x),
      Im y = lfilter(Re c, d, Im x, zi = Im s) + lfilter(Im c, d, Re x),
    four real filters of order N + len(a) - 1 (N with an FIR a = [a0]) that
    run on the HIP kernels, two launches each over stacked channels.  The
    poles that A~ adds are the conjugates of A's, inside the unit circle
    when A's are, and are cancelled by A~'s zeros in the numerator to
    rounding.  zf follows from the DF2T state definition on the last N
    samples of x and y (_lfilter.py:153-166):
      zf[k] = sum_(j > k) (b_j x[T-1+k+1-j] - a_j y[T-1+k+1-j]) + zi[k + T].
    Complex coefficients of order above 32 (real order above 64) raise."""
    nd = x.dim()
    ax = axis % nd
    xm = torch.movedim(x, ax, -1)
    batch = xm.shape[:-1]
    T = xm.shape[-1]
    zb = False
    if zi is not None:
        zb = plan_lfilter_zi(zi.shape, batch, order) if order > 0 else False
    dev = _require_device(x)
    out = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    comp = torch.float64 if out == torch.complex128 else torch.float32
    na = a.shape[0]
    if order + na - 1 > 64:
        raise NotImplementedError(
            "complex filter coefficients are supported up to order 32 on the MI355X path")
    bc = b.to(device=dev, dtype=out)
    ac = a.to(device=dev, dtype=out)
    a0 = ac[0]
    if bool(a0 != 1):
        # the reference's normalisation (_lfilter.py:96-98); a[0] is then 1
        bc = bc / a0
        ac = torch.cat([torch.ones(1, dtype=out, device=dev), ac[1:] / a0])
    bc = F.pad(bc, (0, order + 1 - bc.shape[0]))
    ab = ac.conj().resolve_conj()
    d = _polymul(ac, ab).real.to(comp).contiguous()
    c = _polymul(bc, ab)
    cr = c.real.to(comp).contiguous()
    ci = c.imag.to(comp).contiguous()
    R = max(c.shape[0], d.shape[0]) - 1
    B = 1
    for n_ in batch:
        B *= n_
    if order == 0:
        y = torch.movedim(bc[0] * xm.to(out), -1, ax)
        if zi is None:
            return y
        return y, torch.zeros_like(zi)
    xc = xm.to(out)
    xr, xi = xc.real.to(comp), xc.imag.to(comp)
    sr = si = None
    if zi is not None:
        zrow = zi.to(device=dev, dtype=out).reshape(-1, order)
        s = _polymul(zrow, ab)                               # (rows, N + na - 1)
        s = F.pad(s, (0, R - s.shape[-1]))
        s = s.expand(B, R) if zb else s
        sr = s.real.to(comp).reshape(batch + (R,))
        si = s.imag.to(comp).reshape(batch + (R,))
    # P = F(Re c) over [Re x, Im x] with the state parts; Q = F(Im c) over [Im x, Re x]
    xs = torch.stack([xr, xi])
    zs = None if zi is None else torch.stack([sr, si])
    P = lfilter(cr, d, xs, axis=-1, zi=zs)
    P = P[0] if zi is not None else P
    Q = lfilter(ci, d, torch.stack([xi, xr]), axis=-1)
    y = torch.complex(P[0] - Q[0], P[1] + Q[1]).to(out)
    y_out = torch.movedim(y, -1, ax)
    if zi is None:
        return y_out
    # the final DF2T state from the last N samples of x and y
    N = order
    ac_p = F.pad(ac, (0, N + 1 - ac.shape[0]))
    xt = _tail(xc.reshape(B, T), N)
    yt = _tail(y.reshape(B, T), N)
    zf = xt @ _tail_toeplitz(bc, N).transpose(0, 1) - yt @ _tail_toeplitz(ac_p, N).transpose(0, 1)
    if T < N:
        zrow = zi.to(device=dev, dtype=out).reshape(-1, N).expand(B, N)
        zf = zf + F.pad(zrow[:, T:], (0, T))
    return y_out, zf.reshape(batch + (N,))


def _state_rows(zi, order, dev, comp):
    """zi (batch..., N) -> (rows, N) contiguous (_lfilter.py:139-141)."""
    return zi.to(device=dev, dtype=comp).reshape(-1, order).contiguous()


def _state_out(zf, batch, order, out):
    """(B, N) final state -> (batch..., N) in the output dtype."""
    return zf.reshape(batch + (order,) if batch else (order,)).to(out)


def _lfilter_fir(b, a, x, axis, zi, order):
    ax = axis % x.dim()
    zb = False
    if zi is not None:
        bshape = x.shape[:ax] + x.shape[ax + 1:]
        zb = plan_lfilter_zi(zi.shape, bshape, order)
    dev = _require_device(x)
    out = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    if out.is_complex:
        return _complex_split(b, a, x, axis, zi, out)
    comp = torch.float64 if out == torch.float64 else torch.float32
    a0 = a.to(device=dev, dtype=comp)[0]
    bb = b.to(device=dev, dtype=comp)
    # b / a0 only when a0 != 1, so that a0 gets no gradient then (_lfilter.py:96-98)
    taps = torch.where(a0 != 1, bb / a0, bb).reshape(1, -1).contiguous()
    xm = torch.movedim(x, ax, -1)
    batch = xm.shape[:-1]
    T = xm.shape[-1]
    rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
    B = rows.shape[0]
    zi_r = None
    if zi is not None:
        zi_r = _state_rows(zi, order, dev, comp)
    if B == 0:
        y = rows.new_empty((0, T))
        zf = rows.new_empty((0, order))
    else:
        y, zf = fir_rows(taps, rows, zi_r, channels=B, chans_per_filter=B,
                         want_zf=zi is not None, zi_broadcast=zb)
    y = torch.movedim(y.reshape(batch + (T,)), -1, ax).to(out)
    if zi is None:
        return y
    return y, _state_out(zf, batch, order, out)


def _lfilter_as_section(b, a, x, axis, zi, order):
    dt = torch.promote_types(b.dtype, a.dtype)
    b3 = F.pad(b.to(dt), (0, 3 - b.shape[0]))
    a3 = F.pad(a.to(dt), (0, 3 - a.shape[0]))
    sos = torch.cat([b3, a3]).reshape(1, 6)
    if zi is None:
        return sosfilt(sos, x, axis=axis)
    zi_in = zi
    if order < 2:
        # order-0/1 state padded with the (always zero) second delay
        pad = torch.zeros(zi.shape[:-1] + (2 - order,), dtype=zi.dtype, device=zi.device)
        zi_in = torch.cat([zi, pad], dim=-1)
    zi_s = zi_in.unsqueeze(-2)  # (..., 1, 2)
    y, zf = sosfilt(sos, x, axis=axis, zi=zi_s)
    zf = zf.squeeze(-2)
    if order == 0:
        return y, torch.zeros_like(zi)
    out = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    return y, zf[..., :order].reshape(zi.shape).to(out)


@torch.compiler.disable
def _general_rows(coef, rows, zi_r, want_zf, zi_broadcast, reverse=False):
    """DF2T of order N = len(coef) // 2 - 1 (coef = b then a, padded) over
    rows (B, T) through tsiir_lfilter_forward_ws (time-parallel for orders up
    to 16); reverse: along decreasing time (zi the state at the end, zf after
    the first sample).  Returns (y, zf or None)."""
    B, T = rows.shape
    N = coef.numel() // 2 - 1
    dev = rows.device
    y = torch.empty_like(rows)
    flags = 0
    if zi_r is not None:
        flags |= _lib.HAS_ZI | (_lib.ZI_BROADCAST if zi_broadcast else 0)
    if want_zf:
        flags |= _lib.WRITE_ZF
    if reverse:
        flags |= _lib.REVERSE
    zf = torch.empty((B, N), dtype=rows.dtype, device=dev) if want_zf else None
    if B == 0:
        return y, zf
    d = _desc(channels=B, n_samples=T, n_sections=N, io_dtype=rows.dtype,
              coef_dtype=rows.dtype, flags=flags, x_rows=B, x_row_stride=T, y_row_stride=T,
              chans_per_filter=B)
    L = _lib.load()
    nbytes = L.tsiir_lfilter_workspace_bytes(d)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
    _lib.check(L.tsiir_lfilter_forward_ws(d, _lib.ptr(coef), _lib.ptr(rows), _lib.ptr(y),
                                          _lib.ptr(zi_r), _lib.ptr(zf), _lib.ptr(ws), nbytes,
                                          _lib.stream_handle(dev)), "tsiir_lfilter_forward_ws")
    return y, zf


def _tail_toeplitz(c, N):
    """M[k, l] = c[N + k - l] for k <= l, else 0 (N x N): the weights with
    which zf[k] reads the sample T - N + l of a signal through c."""
    k = torch.arange(N, device=c.device)[:, None]
    l_ = torch.arange(N, device=c.device)[None, :]
    return torch.where(k <= l_, c[(N + k - l_).clamp(0, N)], torch.zeros((), dtype=c.dtype,
                                                                         device=c.device))


def _tail(v, N):
    """The last N samples of rows v (B, T), left-padded with zeros if T < N."""
    T = v.shape[-1]
    if T == 0:
        # (padding an empty tensor is not relied on to write the pad values)
        return v.new_zeros((v.shape[0], N))
    return v[:, T - N:] if T >= N else F.pad(v, (N - T, 0))


def _direct_taps(dz, v, N):
    """d[m] = sum_c sum_{k < m} dz[c, k] v[c, T + k - m] for m = 1..N (d[0] = 0):
    the part of a coefficient gradient that reaches it through zf directly."""
    G = dz.transpose(0, 1) @ _tail(v, N)              # G[k, l] = sum_c dz[c,k] vtail[c,l]
    m = torch.arange(1, N + 1, device=dz.device)[:, None]
    k = torch.arange(N, device=dz.device)[None, :]
    # entries with k >= m are masked below; clamp their indices into G
    vals = G[k.expand(N, N), (N + k - m).clamp(0, N - 1)]
    vals = torch.where(k < m, vals, torch.zeros((), dtype=G.dtype, device=G.device))
    return F.pad(vals.sum(1), (1, 0))


class LfilterRows(torch.autograd.Function):
    """Autograd of the general-order DF2T over rows, with normalised
    coefficients (a[0] = 1).  With g the output gradient run backwards in
    time through the all-pole part 1/A (plus what zf feeds back through
    y), dx is g through the taps b reversed (the FIR kernel), db and da are
    the correlations of g with x and with -y (the tap-gradient kernel), dzi
    is g's first N samples; zf's gradient also reaches x, y, b and a
    directly through the last N samples (_lfilter.py:153-166 recorded by
    the reference's autograd)."""

    @staticmethod
    def forward(ctx, rows, bn, an, zi_r, want_zf, zi_broadcast):
        y, zf = _general_rows(torch.cat([bn, an]).contiguous(), rows, zi_r, want_zf,
                              zi_broadcast)
        ctx.save_for_backward(rows, bn, an, y)
        ctx.meta = (want_zf, zi_broadcast, zi_r is not None)
        return y, zf

    @staticmethod
    def backward(ctx, dy, dzf):
        from ._fir import fir_rows_backward_taps, fir_rows_forward
        x, bn, an, y = ctx.saved_tensors
        want_zf, zi_b, has_zi = ctx.meta
        B, T = x.shape
        N = bn.numel() - 1
        dt = x.dtype
        dyh = torch.zeros_like(x) if dy is None else dy.to(dt).clone()
        dz = dzf.to(dt) if (want_zf and dzf is not None) else None
        lo = max(N - T, 0)                 # tail columns l >= lo exist in the signal
        if dz is not None and T > 0:
            dyh[:, T - N + lo:] -= (dz @ _tail_toeplitz(an, N))[:, lo:]
        # g[n] = dyh[n] - sum_k a_k g[n + k]: the all-pole filter in reverse time
        one = F.pad(torch.ones(1, dtype=dt, device=x.device), (0, N))
        g = dyh
        if T > 0 and B > 0:
            gr, _ = _general_rows(torch.cat([one, an]).contiguous(),
                                  dyh.flip(-1).contiguous(), None, False, False)
            g = gr.flip(-1).contiguous()
        gx = gb = ga = gzi = None
        if ctx.needs_input_grad[0]:
            gx = torch.zeros_like(x)
            if T > 0 and B > 0:
                gx, _ = fir_rows_forward(bn.reshape(1, -1).contiguous(), g, None, channels=B,
                                         chans_per_filter=B, shift=N, reverse_taps=True)
            if dz is not None and T > 0:
                gx[:, T - N + lo:] += (dz @ _tail_toeplitz(bn, N))[:, lo:]
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            db = torch.zeros(N + 1, dtype=dt, device=x.device)
            da = torch.zeros(N + 1, dtype=dt, device=x.device)
            if T > 0 and B > 0:
                db = fir_rows_backward_taps((1, N + 1), x, g, channels=B,
                                            chans_per_filter=B)[0]
                da = -fir_rows_backward_taps((1, N + 1), y, g, channels=B,
                                             chans_per_filter=B)[0]
            if dz is not None:
                db = db + _direct_taps(dz, x, N)
                da = da - _direct_taps(dz, y, N)
            da = torch.cat([torch.zeros(1, dtype=dt, device=x.device), da[1:]])
            gb, ga = db, da
        if has_zi and ctx.needs_input_grad[3]:
            dzi = torch.zeros((B, N), dtype=dt, device=x.device)
            m = min(N, T)
            dzi[:, :m] = g[:, :m]
            if dz is not None and T < N:
                dzi[:, T:] += dz[:, :N - T]
            gzi = dzi.sum(0, keepdim=True) if zi_b else dzi
        return gx, gb, ga, gzi, None, None


def _lfilter_general(b, a, x, axis, zi, order):
    nd = x.dim()
    ax = axis % nd
    zb = False
    if zi is not None:
        zb = plan_lfilter_zi(zi.shape, x.shape[:ax] + x.shape[ax + 1:], order)
    dev = _require_device(x)
    out = torch.promote_types(torch.promote_types(b.dtype, a.dtype), x.dtype)
    if out.is_complex:
        return _complex_split(b, a, x, axis, zi, out)
    comp = torch.float64 if out == torch.float64 else torch.float32
    xm = torch.movedim(x, ax, -1)
    batch = xm.shape[:-1]
    T = xm.shape[-1]
    rows = xm.reshape(math.prod(xm.shape[:-1]), T).to(comp).contiguous()
    B = rows.shape[0]
    n = order + 1
    bb = F.pad(b.to(device=dev, dtype=comp), (0, n - b.shape[0]))
    aa = F.pad(a.to(device=dev, dtype=comp), (0, n - a.shape[0]))
    a0 = aa[0]
    # normalise by a0 only when a0 != 1 (_lfilter.py:96-98), so that a0 gets
    # no gradient then
    bn = torch.where(a0 != 1, bb / a0, bb)
    an = torch.where(a0 != 1, aa / a0, aa)
    zi_r = None
    if zi is not None:
        zi_r = _state_rows(zi, order, dev, comp)
    y, zf = LfilterRows.apply(rows, bn, an, zi_r, zi is not None, zb)
    y = torch.movedim(y.reshape(batch + (T,)), -1, ax).to(out)
    if zi is None:
        return y
    return y, _state_out(zf, batch, order, out)
