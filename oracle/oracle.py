"""FP64 CPU oracle for torchscience's IIR filtering path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: tests/, the smoke
check of ``__graft_entry__`` and the ``cpu_baseline`` leg of bench.py import
it; the product package (``torchscience_amd``) never does.

It restates, in numpy over the C DF2T core in ``oracle.c``:

* ``sosfilt``  - src/torchscience/signal_processing/filter/_sosfilt.py:84-133
  (shape check and message, axis move, zi broadcast of an (S, 2) zi over a
  batched x, section cascade, zf per section);
* ``lfilter``  - src/torchscience/signal_processing/filter/_lfilter.py:86-178
  (1-D checks and messages, a0 normalisation, b/a padding, order-0 gain,
  zi reshape, DF2T loop);
* ``batched_filter_apply`` - src/torchscience/signal_processing/filter/_batched.py:449-566;
* ``sosfilt_zi`` / ``lfilter_zi`` - _sosfilt_zi.py:54-86 / _lfilter_zi.py:70-128;
* ``lfiltic`` - _lfilter.py:181-305 (a0 normalisation, zero padding of the
  histories, the state sums worked backwards from the last delay);
* ``sosfiltfilt`` - _sosfiltfilt.py:96-173 (default padlen from the count of
  zeros at the origin, padding _pad_signal :176-217, forward pass from
  sosfilt_zi * first sample, reversed pass from sosfilt_zi * last output,
  crop);
* ``filtfilt`` - _filtfilt.py:86-165 (a0 normalisation, padlen
  3 * max(len(a), len(b)), the same two passes with lfilter_zi, order-0
  gain b0^2);
* ``fir_rows`` - lfilter with a = [a0] (_lfilter.py:96-166: the DF2T loop
  with a zero denominator tail is a convolution plus the state terms) and
  the 'same' slice of fftfilt (_fftfilt.py:60-92), as full convolutions
  (numpy.convolve) sliced at an offset;
* ``fftfilt`` - _fftfilt.py:9-99 (axis move, 'same' slice from
  (len(b) - 1) // 2);
* ``impulse_response_sos`` - _impulse_response.py:173-271 (sections filtered
  in order with a0 taken as 1);
* ``sos_frequency_response`` - _frequency_response_sos.py:128-137 (product
  of the section ratios at z = exp(jw));
* ``ba_frequency_response`` - _frequency_response.py:104-116 (B(z^-1) /
  A(z^-1) at z = exp(jw));
* ``iir_design_sos`` - one filter of batched_butterworth_design /
  batched_chebyshev_type_1_design / batched_chebyshev_type_2_design
  (_batched.py:22-344): prototypes (_butterworth_prototype.py:84-91,
  _chebyshev_type_1_prototype.py:103-124, _chebyshev_type_2_prototype.py:
  108-142), pre-warp (_butterworth_design.py:118), band maps
  (_lowpass_to_*_zpk.py, bandstop keeping the gain), bilinear map at fs = 2
  (_bilinear_transform_zpk.py:57-87) and zpk_to_sos (_zpk_to_sos.py:44-197),
  all in float64.

Deliberate difference from the reference: zi with ``axis != -1`` follows the
documented meaning (zi batch dims are x's batch dims in order), not the
reference's ``moveaxis(zi, axis, -1)`` (_lfilter.py:137), which misorders
state for x of two or more dims and is not covered by any reference test.

All results are float64 (complex128 for complex input); the product's dtype
promotion is checked separately.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_lfilter_f64.argtypes = [_f64p, ctypes.c_int, _f64p, ctypes.c_int,
                                         _f64p, _f64p, ctypes.c_int64, ctypes.c_int64,
                                         _f64p, _f64p, ctypes.c_int]
        L.oracle_sosfilt_f64.argtypes = [_f64p, ctypes.c_int, ctypes.c_int64,
                                         _f64p, _f64p, ctypes.c_int64, ctypes.c_int64,
                                         _f64p, _f64p, ctypes.c_int]
        L.oracle_sosfilt_f32io.argtypes = [_f64p, ctypes.c_int, ctypes.c_int64,
                                           _f32p, _f32p, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_int]
        L.oracle_sosfilt_f32.argtypes = [_f32p, ctypes.c_int, ctypes.c_int64,
                                         _f32p, _f32p, ctypes.c_int64, ctypes.c_int64,
                                         _f32p, _f32p, ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray | None, t=_f64p):
    return None if a is None else a.ctypes.data_as(t)


def default_threads() -> int:
    return max(1, os.cpu_count() or 1)


# --------------------------------------------------------------------------
# flat (channels, samples) cores
# --------------------------------------------------------------------------

def sosfilt_rows(sos: np.ndarray, x: np.ndarray, zi: np.ndarray | None = None,
                 chans_per_filter: int | None = None, n_threads: int = 1):
    """Real FP64 sosfilt over rows of x (B, T).

    sos is (S, 6) or (n_filters, S, 6) with channel c using filter
    c // chans_per_filter.  zi is (B, S, 2) or None.  Returns (y, zf)."""
    sos = np.ascontiguousarray(sos, dtype=np.float64)
    if sos.ndim == 2:
        sos = sos[None]
    x = np.ascontiguousarray(x, dtype=np.float64)
    B, T = x.shape
    S = sos.shape[1]
    cpf = B if chans_per_filter is None else chans_per_filter
    cpf = max(cpf, 1)
    y = np.empty_like(x)
    zf = np.empty((B, S, 2), dtype=np.float64)
    zi_c = None if zi is None else np.ascontiguousarray(zi, dtype=np.float64)
    lib().oracle_sosfilt_f64(_p(sos), S, cpf, _p(x), _p(y), B, T, _p(zi_c), _p(zf),
                             n_threads)
    return y, zf


def sosfilt_rows_f32io(sos: np.ndarray, x: np.ndarray, chans_per_filter: int | None = None,
                       n_threads: int = 1) -> np.ndarray:
    """Float32 rows in and out, FP64 arithmetic (the CPU baseline form)."""
    sos = np.ascontiguousarray(sos, dtype=np.float64)
    if sos.ndim == 2:
        sos = sos[None]
    x = np.ascontiguousarray(x, dtype=np.float32)
    B, T = x.shape
    cpf = B if chans_per_filter is None else max(chans_per_filter, 1)
    y = np.empty_like(x)
    lib().oracle_sosfilt_f32io(_p(sos), sos.shape[1], cpf, _p(x, _f32p), _p(y, _f32p),
                               B, T, n_threads)
    return y


def sosfilt_rows_f32(sos: np.ndarray, x: np.ndarray, zi: np.ndarray | None = None,
                     chans_per_filter: int | None = None, n_threads: int = 1):
    """The reference's FP32 arithmetic: float32 sos and x, every operation of
    the a0 normalisation and the DF2T update rounded to float32
    (_lfilter.py:96-98, 153-166; _sosfilt.py:117-128).  sos (S, 6) or
    (n_filters, S, 6); zi (B, S, 2) or None.  Returns float32 (y, zf)."""
    sos = np.ascontiguousarray(sos, dtype=np.float32)
    if sos.ndim == 2:
        sos = sos[None]
    x = np.ascontiguousarray(x, dtype=np.float32)
    B, T = x.shape
    S = sos.shape[1]
    cpf = B if chans_per_filter is None else max(chans_per_filter, 1)
    y = np.empty_like(x)
    zf = np.empty((B, S, 2), dtype=np.float32)
    zi_c = None if zi is None else np.ascontiguousarray(zi, dtype=np.float32)
    lib().oracle_sosfilt_f32(_p(sos, _f32p), S, cpf, _p(x, _f32p), _p(y, _f32p), B, T,
                             _p(zi_c, _f32p), _p(zf, _f32p), n_threads)
    return y, zf


def lfilter_rows(b: np.ndarray, a: np.ndarray, x: np.ndarray,
                 zi: np.ndarray | None = None, n_threads: int = 1):
    """Real FP64 lfilter over rows of x (B, T); zi (B, N).  Returns (y, zf)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    B, T = x.shape
    order = max(b.size, a.size) - 1
    if order > 63:
        raise ValueError("oracle supports filter order <= 63")
    y = np.empty_like(x)
    zf = np.empty((B, max(order, 0)), dtype=np.float64)
    zi_c = None if zi is None else np.ascontiguousarray(zi, dtype=np.float64)
    lib().oracle_lfilter_f64(_p(b), b.size, _p(a), a.size, _p(x), _p(y), B, T,
                             _p(zi_c), _p(zf) if order > 0 else None, n_threads)
    return y, zf


def _lfilter_rows_complex(b, a, x, zi):
    """DF2T with complex coefficients, numpy-vectorised over rows (small
    inputs only).  Same op order as _lfilter.py:153-166."""
    b = np.asarray(b, dtype=np.complex128)
    a = np.asarray(a, dtype=np.complex128)
    if a[0] != 1.0:
        b = b / a[0]
        a = a / a[0]
    n = max(b.size, a.size)
    b = np.pad(b, (0, n - b.size))
    a = np.pad(a, (0, n - a.size))
    order = n - 1
    x = np.asarray(x, dtype=np.complex128)
    B, T = x.shape
    y = np.empty_like(x)
    if order == 0:
        return b[0] * x, np.zeros((B, 0), np.complex128)
    z = np.zeros((B, order), np.complex128) if zi is None else np.array(zi, np.complex128)
    for t in range(T):
        xn = x[:, t]
        yn = b[0] * xn + z[:, 0]
        y[:, t] = yn
        for k in range(order - 1):
            z[:, k] = b[k + 1] * xn - a[k + 1] * yn + z[:, k + 1]
        z[:, order - 1] = b[order] * xn - a[order] * yn
    return y, z


# --------------------------------------------------------------------------
# reference-semantics wrappers
# --------------------------------------------------------------------------

def _is_complex(*arrs) -> bool:
    return any(np.iscomplexobj(a) for a in arrs if a is not None)


def lfilter(b, a, x, axis: int = -1, zi=None):
    """lfilter with the reference's semantics (_lfilter.py:12-178)."""
    b = np.atleast_1d(np.asarray(b))
    a = np.atleast_1d(np.asarray(a))
    if b.ndim != 1:
        raise ValueError("b must be a 1-D tensor")
    if a.ndim != 1:
        raise ValueError("a must be a 1-D tensor")
    x = np.asarray(x)
    xm = np.moveaxis(x, axis, -1)
    batch = xm.shape[:-1]
    T = xm.shape[-1]
    rows = xm.reshape(-1, T)
    order = max(b.size, a.size) - 1
    zi_rows = None
    if zi is not None:
        zi_rows = np.asarray(zi).reshape(-1, order) if order > 0 else None
        if zi_rows is not None and zi_rows.shape[0] == 1 and rows.shape[0] > 1:
            zi_rows = np.broadcast_to(zi_rows, (rows.shape[0], order))
    if _is_complex(b, a):
        y, zf = _lfilter_rows_complex(b, a, rows, zi_rows)
    elif _is_complex(x, zi):
        zr = None if zi_rows is None else np.real(zi_rows)
        zim = None if zi_rows is None else np.imag(zi_rows)
        yr, zfr = lfilter_rows(np.real(b), np.real(a), np.real(rows), zr)
        yi, zfi = lfilter_rows(np.real(b), np.real(a), np.imag(rows), zim)
        y, zf = yr + 1j * yi, zfr + 1j * zfi
    else:
        y, zf = lfilter_rows(b, a, rows, zi_rows)
    y = np.moveaxis(y.reshape(xm.shape), -1, axis)
    if zi is None:
        return y
    if order == 0:
        return y, np.zeros_like(np.asarray(zi), dtype=y.dtype)
    return y, zf.reshape(batch + (order,)) if batch else zf.reshape(order)


def sosfilt(sos, x, axis: int = -1, zi=None):
    """sosfilt with the reference's semantics (_sosfilt.py:13-133)."""
    sos = np.asarray(sos)
    if sos.ndim != 2 or sos.shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    S = sos.shape[0]
    x = np.asarray(x)
    xm = np.moveaxis(x, axis, -1)
    batch = xm.shape[:-1]
    T = xm.shape[-1]
    rows = xm.reshape(-1, T)
    B = rows.shape[0]
    zi_rows = None
    if zi is not None:
        zi = np.asarray(zi)
        if zi.shape == (S, 2) and len(batch) > 0:
            zi_rows = np.broadcast_to(zi, (B, S, 2))
        else:
            zi_rows = zi.reshape(B, S, 2)
    if _is_complex(sos):
        y = rows.astype(np.complex128)
        zf = np.zeros((B, S, 2), np.complex128)
        for s in range(S):
            y, zfs = _lfilter_rows_complex(sos[s, :3], sos[s, 3:], y,
                                           None if zi_rows is None else zi_rows[:, s, :])
            zf[:, s, :] = zfs
    elif _is_complex(x, zi):
        zr = None if zi_rows is None else np.real(zi_rows)
        zim = None if zi_rows is None else np.imag(zi_rows)
        yr, zfr = sosfilt_rows(np.real(sos), np.real(rows), zr)
        yi, zfi = sosfilt_rows(np.real(sos), np.imag(rows), zim)
        y, zf = yr + 1j * yi, zfr + 1j * zfi
    else:
        y, zf = sosfilt_rows(sos, rows, zi_rows)
    y = np.moveaxis(y.reshape(xm.shape), -1, axis)
    if zi is None:
        return y
    zf_shape = batch + (S, 2) if (zi.shape != (S, 2) or len(batch) > 0) else (S, 2)
    return y, zf.reshape(zf_shape)


def batched_filter_apply(filters, x, filter_format: str = "sos", broadcast: bool = True,
                         axis: int = -1):
    """batched_filter_apply with the reference's semantics (_batched.py:449-566)."""
    filters = np.asarray(filters)
    x = np.asarray(x)
    F = filters.shape[0]
    if filter_format == "sos":
        if filters.ndim != 3 or filters.shape[2] != 6:
            raise ValueError(
                f"SOS filters must have shape (batch, n_sections, 6), got {filters.shape}")
        if broadcast:
            return np.stack([sosfilt(filters[i], x, axis=axis) for i in range(F)])
        if x.shape[0] != F:
            raise ValueError(
                f"When broadcast=False, x batch dim must match filters: {x.shape[0]} != {F}")
        return np.stack([sosfilt(filters[i], x[i], axis=axis) for i in range(F)])
    if filter_format == "fir":
        if filters.ndim != 2:
            raise ValueError(
                f"FIR filters must have shape (batch, num_taps), got {filters.shape}")
        one = np.ones(1)
        if broadcast:
            return np.stack([lfilter(filters[i], one, x, axis=axis) for i in range(F)])
        if x.shape[0] != F:
            raise ValueError(
                f"When broadcast=False, x batch dim must match filters: {x.shape[0]} != {F}")
        return np.stack([lfilter(filters[i], one, x[i], axis=axis) for i in range(F)])
    raise ValueError(f"filter_format must be 'sos' or 'fir', got '{filter_format}'")


def lfilter_zi(b, a) -> np.ndarray:
    """Steady-state initial state (_lfilter_zi.py:70-128): solve
    (I - A^T) zi = b[1:] - a[1:] b[0] with A the companion matrix."""
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    while a.size > 1 and a[0] == 0.0:
        a = a[1:]
    if a[0] != 1.0:
        b = b / a[0]
        a = a / a[0]
    n = max(a.size, b.size)
    a = np.pad(a, (0, n - a.size))
    b = np.pad(b, (0, n - b.size))
    comp = np.zeros((n - 1, n - 1))
    comp[0, :] = -a[1:]
    comp[1:, :-1] += np.eye(n - 2) if n > 2 else 0.0
    m = np.eye(n - 1) - comp.T
    rhs = b[1:] - a[1:] * b[0]
    return np.linalg.solve(m, rhs)


def sosfilt_zi(sos) -> np.ndarray:
    """Cascaded steady state (_sosfilt_zi.py:54-86)."""
    sos = np.asarray(sos, dtype=np.float64)
    zi = np.empty((sos.shape[0], 2))
    scale = 1.0
    for s in range(sos.shape[0]):
        b, a = sos[s, :3], sos[s, 3:]
        zi[s] = scale * lfilter_zi(b, a)
        scale = scale * b.sum() / a.sum()
    return zi


def _pad_signal(x, padtype, padlen):
    """Edge extension along the last axis (_sosfiltfilt.py:176-217,
    _filtfilt.py:168-214)."""
    if padtype == "odd":
        left = 2 * x[..., 0:1] - x[..., 1:padlen + 1][..., ::-1]
        right = 2 * x[..., -1:] - x[..., -padlen - 1:-1][..., ::-1]
        return np.concatenate([left, x, right], axis=-1)
    if padtype == "even":
        left = x[..., 1:padlen + 1][..., ::-1]
        right = x[..., -padlen - 1:-1][..., ::-1]
        return np.concatenate([left, x, right], axis=-1)
    if padtype == "constant":
        left = np.repeat(x[..., 0:1], padlen, axis=-1)
        right = np.repeat(x[..., -1:], padlen, axis=-1)
        return np.concatenate([left, x, right], axis=-1)
    raise ValueError(f"Unknown padtype: {padtype}")


def sosfiltfilt(sos, x, axis: int = -1, padtype="odd", padlen=None):
    """Zero-phase SOS filtering with the reference's semantics
    (_sosfiltfilt.py:96-173)."""
    sos = np.asarray(sos, dtype=np.float64)
    if sos.ndim != 2 or sos.shape[1] != 6:
        raise ValueError("sos must be shape (n_sections, 6)")
    S = sos.shape[0]
    if padlen is None:
        ntaps = 2 * S + 1
        ntaps -= min(int((sos[:, 2] == 0).sum()), int((sos[:, 5] == 0).sum()))
        padlen = 3 * ntaps
    xm = np.moveaxis(np.asarray(x, dtype=np.float64), axis, -1)
    T = xm.shape[-1]
    if padtype is not None and padlen >= T:
        raise ValueError(f"padlen ({padlen}) must be less than x.shape[axis] ({T})")
    xp = _pad_signal(xm, padtype, padlen) if padtype is not None and padlen > 0 else xm
    batch, Tp = xp.shape[:-1], xp.shape[-1]
    rows = np.ascontiguousarray(xp.reshape(-1, Tp))
    if S > 0:
        zi = sosfilt_zi(sos)
        y1, _ = sosfilt_rows(sos, rows, zi[None] * rows[:, 0, None, None])
        yr = np.ascontiguousarray(y1[:, ::-1])
        y2, _ = sosfilt_rows(sos, yr, zi[None] * yr[:, 0, None, None])
        y = y2[:, ::-1]
    else:
        y = rows
    y = y.reshape(batch + (Tp,))
    if padtype is not None and padlen > 0:
        y = y[..., padlen:-padlen]
    return np.moveaxis(y, -1, axis)


def filtfilt(b, a, x, axis: int = -1, padtype="odd", padlen=None):
    """Zero-phase (b, a) filtering with the reference's semantics
    (_filtfilt.py:86-165)."""
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    if b.ndim != 1:
        raise ValueError("b must be a 1-D tensor")
    if a.ndim != 1:
        raise ValueError("a must be a 1-D tensor")
    if a[0] != 1.0:
        b = b / a[0]
        a = a / a[0]
    order = max(b.size, a.size) - 1
    if padlen is None:
        padlen = 3 * max(b.size, a.size)
    xm = np.moveaxis(np.asarray(x, dtype=np.float64), axis, -1)
    T = xm.shape[-1]
    if padtype is not None and padlen >= T:
        raise ValueError(f"padlen ({padlen}) must be less than x.shape[axis] ({T})")
    xp = _pad_signal(xm, padtype, padlen) if padtype is not None and padlen > 0 else xm
    batch, Tp = xp.shape[:-1], xp.shape[-1]
    rows = np.ascontiguousarray(xp.reshape(-1, Tp))
    if order > 0:
        zi = lfilter_zi(b, a)
        y1, _ = lfilter_rows(b, a, rows, zi[None] * rows[:, 0:1])
        yr = np.ascontiguousarray(y1[:, ::-1])
        y2, _ = lfilter_rows(b, a, yr, zi[None] * yr[:, 0:1])
        y = y2[:, ::-1]
    else:
        y = b[0] * b[0] * rows
    y = y.reshape(batch + (Tp,))
    if padtype is not None and padlen > 0:
        y = y[..., padlen:-padlen]
    return np.moveaxis(y, -1, axis)


def impulse_state_response(sos_row: np.ndarray, n: int) -> np.ndarray:
    """Markov parameters of one DF2T biquad: h[k] for k < n (closed form via
    the state-space matrices, used for brute-force checks)."""
    b0, b1, b2, a0, a1, a2 = [float(v) for v in sos_row]
    if a0 != 1.0:
        b0, b1, b2, a1, a2 = b0 / a0, b1 / a0, b2 / a0, a1 / a0, a2 / a0
    A = np.array([[-a1, 1.0], [-a2, 0.0]])
    Bv = np.array([b1 - a1 * b0, b2 - a2 * b0])
    C = np.array([1.0, 0.0])
    h = np.empty(n)
    h[0] = b0
    P = np.eye(2)
    for k in range(1, n):
        h[k] = C @ P @ Bv
        P = A @ P
    return h


def fir_rows(taps: np.ndarray, x: np.ndarray, zi: np.ndarray | None = None, shift: int = 0,
             want_zf: bool = False):
    """FIR over rows of x (B, T) with taps (K,) or (B, K) per row:
    y[n] = sum_j h[j] x[n + shift - j] (x zero outside the row), plus zi[n]
    for n < K - 1; zf[k] = the full convolution at T + k plus zi[T + k]
    (the DF2T final state of _lfilter.py:153-166 with a = [1]).
    Returns (y, zf or None)."""
    x = np.asarray(x, dtype=np.float64)
    B, T = x.shape
    h = np.asarray(taps, dtype=np.float64)
    if h.ndim == 1:
        h = np.broadcast_to(h, (B, h.size))
    K = h.shape[1]
    N = K - 1
    if shift < 0:
        raise ValueError("shift >= 0 only")
    ext = np.zeros((B, T + N + shift))         # full convolution, zero-extended
    if T > 0:
        for i in range(B):
            ext[i, :T + N] = np.convolve(x[i], h[i])
    y = ext[:, shift:shift + T].copy()
    zf = None
    if zi is not None or want_zf:
        if shift != 0:
            raise ValueError("state terms need shift == 0")
        tail = ext[:, T:T + N].copy()
        if zi is not None:
            z = np.broadcast_to(np.asarray(zi, dtype=np.float64), (B, N))
            m = min(N, T)
            y[:, :m] += z[:, :m]
            if T < N:
                tail[:, :N - T] += z[:, T:]
        zf = tail
    return y, zf


def fftfilt(b, x, axis: int = -1):
    """fftfilt with the reference's semantics (_fftfilt.py:9-99): the full
    convolution of each row with b, sliced from (len(b) - 1) // 2."""
    b = np.asarray(b)
    x = np.asarray(x)
    xm = np.moveaxis(x, axis, -1)
    T = xm.shape[-1]
    rows = xm.reshape(-1, T)
    if np.iscomplexobj(rows):
        yr, _ = fir_rows(b, np.real(rows), shift=(b.size - 1) // 2)
        yi, _ = fir_rows(b, np.imag(rows), shift=(b.size - 1) // 2)
        y = yr + 1j * yi
    else:
        y, _ = fir_rows(b, rows, shift=(b.size - 1) // 2)
    return np.moveaxis(y.reshape(xm.shape), -1, axis)


def impulse_response_sos(sos, n: int) -> np.ndarray:
    """The cascade's impulse response with each section's a0 taken as 1, as
    the reference's per-section helper filters (_impulse_response.py:215-220,
    226-271)."""
    s = np.array(sos, dtype=np.float64).reshape(-1, 6)
    s[:, 3] = 1.0
    x = np.zeros((1, n))
    if n:
        x[0, 0] = 1.0
    if s.shape[0] == 0 or n == 0:
        return x[0]
    y, _ = sosfilt_rows(s, x)
    return y[0]


def sos_frequency_response(sos, w) -> np.ndarray:
    """H(exp(jw)) = prod over sections of (b0 + b1 z^-1 + b2 z^-2) /
    (a0 + a1 z^-1 + a2 z^-2), a0 as given."""
    s = np.asarray(sos, dtype=np.float64).reshape(-1, 6)
    zi1 = np.exp(-1j * np.asarray(w, dtype=np.float64))
    h = np.ones_like(zi1)
    for b0, b1, b2, a0, a1, a2 in s:
        h = h * (b0 + b1 * zi1 + b2 * zi1 ** 2) / (a0 + a1 * zi1 + a2 * zi1 ** 2)
    return h


def _design_separate(x):
    """Real values and positive-imaginary complex values, in order
    (_zpk_to_sos.py:176-197)."""
    x = np.asarray(x, dtype=np.complex128)
    if x.size == 0:
        return np.zeros(0), np.zeros(0, np.complex128)
    is_real = (np.abs(x.imag) < 1e-6 * np.abs(x.real) + 1e-10) | (np.abs(x) < 1e-6)
    return x[is_real].real, x[~is_real & (x.imag > 0)]


def _design_zpk_to_sos(z, p, k):
    n_sec = (p.size + 1) // 2
    if z.size < p.size:
        z = np.concatenate([z, -np.ones(p.size - z.size, np.complex128)])
    rp, cp = _design_separate(p)
    rz, cz = _design_separate(z)
    rp, cp, rz, cz = list(rp), list(cp), list(rz), list(cz)
    rows = []
    for p1 in cp:
        if cz:
            z1 = cz.pop(0)
            num = [1.0, -2 * z1.real, z1.real ** 2 + z1.imag ** 2]
        elif len(rz) >= 2:
            z1, z2 = rz.pop(0), rz.pop(0)
            num = [1.0, -(z1 + z2), z1 * z2]
        else:
            num = [1.0, 2.0, 1.0]
        rows.append(num + [1.0, -2 * p1.real, p1.real ** 2 + p1.imag ** 2])
    while len(rp) >= 2:
        p1, p2 = rp.pop(0), rp.pop(0)
        z1, z2 = (rz.pop(0), rz.pop(0)) if len(rz) >= 2 else (-1.0, -1.0)
        rows.append([1.0, -(z1 + z2), z1 * z2, 1.0, -(p1 + p2), p1 * p2])
    if len(rp) == 1:
        z1 = rz[0] if rz else -1.0
        rows.append([1.0, -z1, 0.0, 1.0, -rp[0], 0.0])
    sos = np.array(rows, dtype=np.float64).reshape(-1, 6)
    sos[:, :3] *= abs(k) ** (1.0 / n_sec)
    if k < 0:
        sos[0, 0] = -sos[0, 0]
    return sos


def iir_design_sos(family: str, order: int, cutoff, band: str = "lowpass",
                   param: float = 0.0) -> np.ndarray:
    """One digital IIR design, family "butterworth" | "chebyshev1" (param:
    passband ripple dB) | "chebyshev2" (param: stopband attenuation dB),
    cutoff(s) as fractions of Nyquist, into the reference's SOS layout."""
    N = int(order)
    kk = np.arange(N, dtype=np.float64)
    z = np.zeros(0, np.complex128)
    if family == "butterworth":
        ang = np.pi * (2 * kk + N + 1) / (2 * N)
        p = np.cos(ang) + 1j * np.sin(ang)
        k = 1.0
    else:
        e = np.sqrt(10 ** (param / 10) - 1)
        eps = e if family == "chebyshev1" else 1.0 / e
        a = np.arcsinh(1.0 / eps) / N
        th = np.pi * (2 * kk + 1) / (2 * N)
        p1 = -np.sinh(a) * np.sin(th) + 1j * np.cosh(a) * np.cos(th)
        if family == "chebyshev1":
            p = p1
            k = float(np.real(np.prod(-p)))
            if N % 2 == 0:
                k = k / np.sqrt(1 + eps ** 2)
        else:
            p = 1.0 / p1
            ct = np.cos(th)
            if N % 2 == 1:
                ct = np.concatenate([ct[:N // 2], ct[N // 2 + 1:]])
            z = 1j * (1.0 / ct)
            k = float(np.real(np.prod(-p) / np.prod(-z))) if z.size else float(np.real(np.prod(-p)))
    c = np.atleast_1d(np.asarray(cutoff, dtype=np.float64))
    w = 4.0 * np.tan(np.pi * c / 2.0)
    deg = p.size - z.size
    if band == "lowpass":
        z, p, k = z * w[0], p * w[0], k * w[0] ** deg
    elif band == "highpass":
        pz = np.prod(-z) if z.size else 1.0
        k = k * float(np.real(pz / np.prod(-p)))
        z = np.concatenate([w[0] / z, np.zeros(deg, np.complex128)])
        p = w[0] / p
    else:
        bw = w[1] - w[0]
        wc = np.sqrt(w[0] * w[1])

        def roots(v, stop):
            h = bw / (2 * v) if stop else bw * v / 2
            d = np.sqrt(h * h - wc * wc + 0j)
            return np.concatenate([h + d, h - d])
        stop = band == "bandstop"
        p = roots(p, stop)
        zt = roots(z, stop) if z.size else np.zeros(0, np.complex128)
        if stop:
            z = np.concatenate([zt, np.tile(np.array([1j * wc, -1j * wc]), deg)])
        else:
            z = np.concatenate([zt, np.zeros(deg, np.complex128)])
            k = k * bw ** deg
    deg = p.size - z.size
    num = np.prod(4.0 - z) if z.size else 1.0
    k = k * float(np.real(num / np.prod(4.0 - p)))
    zd = (1 + z / 4.0) / (1 - z / 4.0)
    zd = np.concatenate([zd, -np.ones(deg, np.complex128)])
    pd = (1 + p / 4.0) / (1 - p / 4.0)
    return _design_zpk_to_sos(zd, pd, k)


def lfiltic(b, a, y, x=None) -> np.ndarray:
    """DF2T initial state from past outputs / inputs, most recent first
    (_lfilter.py:181-305): zi[k] = sum_{j=k}^{N-1} b[j+1] x[j-k] -
    a[j+1] y[j-k], histories zero-padded to N."""
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    y = np.atleast_1d(np.asarray(y, dtype=np.float64))
    if a[0] != 1.0:
        b, a = b / a[0], a / a[0]
    n = max(b.size, a.size)
    N = n - 1
    if N == 0:
        return np.zeros(0)
    b = np.concatenate([b, np.zeros(n - b.size)])
    a = np.concatenate([a, np.zeros(n - a.size)])
    x = np.zeros(N) if x is None else np.atleast_1d(np.asarray(x, dtype=np.float64))
    y = np.concatenate([y, np.zeros(max(N - y.size, 0))])
    x = np.concatenate([x, np.zeros(max(N - x.size, 0))])
    zi = np.zeros(N)
    for k in range(N - 1, -1, -1):
        acc = 0.0
        for j in range(k, N):
            acc += b[j + 1] * x[j - k] - a[j + 1] * y[j - k]
        zi[k] = acc
    return zi


def ba_frequency_response(b, a, w) -> np.ndarray:
    """B(z^-1) / A(z^-1) at z = exp(jw) (numpy.polyval of the reversed
    coefficients); an empty polynomial is 1 (_frequency_response.py:139-140)."""
    zi1 = np.exp(-1j * np.asarray(w, dtype=np.float64))

    def ev(c):
        c = np.atleast_1d(np.asarray(c, dtype=np.float64))
        return np.ones_like(zi1) if c.size == 0 else np.polyval(c[::-1], zi1)
    return ev(b) / ev(a)
