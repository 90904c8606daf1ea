"""FP64 IIR design for building filter coefficients on the host.

Host-side set-up math, not the filtering path: it produces the SOS
coefficient tables the benchmark workloads and tests filter with.  It
follows the reference's design chain step by step, always in float64:

* analog prototypes: Butterworth poles exp(j*pi*(2k+n+1)/(2n))
  (_butterworth_prototype.py:84-91); Chebyshev type I poles
  -sinh(a) sin(theta_k) + j cosh(a) cos(theta_k), gain prod(-p) and
  1/sqrt(1+eps^2) for even order (_chebyshev_type_1_prototype.py:93-124);
* pre-warp 4*tan(pi*f/2) (_butterworth_design.py:130);
* lowpass->lowpass scaling (_lowpass_to_lowpass_zpk.py:55-60) and
  lowpass->bandpass (_lowpass_to_bandpass_zpk.py:69-96);
* bilinear transform at fs = 2 (_bilinear_transform_zpk.py:57-85);
* zpk -> sos with the reference's pairing: conjugate poles in the order they
  appear, zeros padded with -1, and the gain split evenly as |k|^(1/S) over
  the sections (_zpk_to_sos.py:44-168).

The reference builds its prototype poles at torch's default dtype (float32
unless changed, _butterworth_prototype.py:66-77), which leaves float32-accurate
poles inside a float64 SOS; this module does not reproduce that and designs
every step in float64.  Its magnitude responses are pinned against
scipy.signal's designs in tests/test_host.py.
"""

from __future__ import annotations

import math

import numpy as np


def _butterworth_prototype(order: int):
    k = np.arange(order, dtype=np.float64)
    ang = math.pi * (2 * k + order + 1) / (2 * order)
    return np.zeros(0, np.complex128), np.cos(ang) + 1j * np.sin(ang), 1.0


def _chebyshev1_prototype(order: int, ripple_db: float):
    eps = math.sqrt(10 ** (ripple_db / 10) - 1)
    a = math.asinh(1.0 / eps) / order
    k = np.arange(order, dtype=np.float64)
    theta = math.pi * (2 * k + 1) / (2 * order)
    poles = -math.sinh(a) * np.sin(theta) + 1j * (math.cosh(a) * np.cos(theta))
    gain = float(np.real(np.prod(-poles)))
    if order % 2 == 0:
        gain = gain / math.sqrt(1 + eps ** 2)
    return np.zeros(0, np.complex128), poles, gain


def _lp2lp(z, p, k, wc):
    return z * wc, p * wc, k * wc ** (p.size - z.size)


def _lp2bp(z, p, k, w0, bw):
    deg = p.size - z.size
    hp = bw * p / 2
    sd = np.sqrt(hp * hp - w0 * w0 + 0j)
    p_new = np.concatenate([hp + sd, hp - sd])
    if z.size:
        hz = bw * z / 2
        sz = np.sqrt(hz * hz - w0 * w0 + 0j)
        zt = np.concatenate([hz + sz, hz - sz])
    else:
        zt = np.zeros(0, np.complex128)
    return np.concatenate([zt, np.zeros(deg, np.complex128)]), p_new, k * bw ** deg


def _bilinear(z, p, k, fs: float = 2.0):
    fs2 = 2 * fs
    deg = p.size - z.size
    pd = (1 + p / fs2) / (1 - p / fs2)
    zd = (1 + z / fs2) / (1 - z / fs2) if z.size else np.zeros(0, np.complex128)
    zd = np.concatenate([zd, -np.ones(deg, np.complex128)])
    num = np.prod(fs2 - z) if z.size else 1.0
    kd = k * float(np.real(num / np.prod(fs2 - p)))
    return zd, pd, kd


def _separate(x):
    if x.size == 0:
        return np.zeros(0), np.zeros(0, np.complex128)
    tol = 1e-6 * np.abs(x.real) + 1e-10
    is_real = (np.abs(x.imag) < tol) | (np.abs(x) < 1e-6)
    return x[is_real].real, x[~is_real & (x.imag > 0)]


def zpk_to_sos(z, p, k) -> np.ndarray:
    """Reference pairing and even gain split (_zpk_to_sos.py:44-168)."""
    z = np.asarray(z, np.complex128)
    p = np.asarray(p, np.complex128)
    n_sec = (p.size + 1) // 2
    rp, cp = _separate(p)
    if z.size < p.size:
        z = np.concatenate([z, -np.ones(p.size - z.size, np.complex128)])
    rz, cz = _separate(z)
    rp, cp, rz, cz = list(rp), list(cp), list(rz), list(cz)
    rows = []
    while cp:
        p1 = cp.pop(0)
        if cz:
            z1 = cz.pop(0)
        elif len(rz) >= 2:
            z1, z2 = rz.pop(0), rz.pop(0)
            rows.append([1.0, -(z1 + z2), z1 * z2, 1.0, -2 * p1.real, p1.real ** 2 + p1.imag ** 2])
            continue
        else:
            z1 = -1.0 + 0j
        z1 = complex(z1)
        rows.append([1.0, -2 * z1.real, z1.real ** 2 + z1.imag ** 2,
                     1.0, -2 * p1.real, p1.real ** 2 + p1.imag ** 2])
    while len(rp) >= 2:
        p1, p2 = rp.pop(0), rp.pop(0)
        if len(rz) >= 2:
            z1, z2 = rz.pop(0), rz.pop(0)
        else:
            z1 = z2 = -1.0
        rows.append([1.0, -(z1 + z2), z1 * z2, 1.0, -(p1 + p2), p1 * p2])
    if len(rp) == 1:
        p1 = rp.pop(0)
        z1 = rz[0] if rz else -1.0
        rows.append([1.0, -z1, 0.0, 1.0, -p1, 0.0])
    if not rows:
        return np.zeros((0, 6))
    sos = np.array(rows, dtype=np.float64)
    sos[:, :3] *= abs(k) ** (1.0 / n_sec)
    if k < 0:
        sos[0, 0] = -sos[0, 0]
    return sos


def _design(z, p, k, cutoff, filter_type):
    c = np.atleast_1d(np.asarray(cutoff, dtype=np.float64))
    if c.size == 1:
        if not 0 < c[0] < 1:
            raise ValueError(f"Cutoff frequency must be between 0 and 1 (Nyquist), got {c[0]}")
    elif not 0 < c[0] < c[1] < 1:
        raise ValueError(f"Cutoff frequencies must satisfy 0 < low < high < 1, got {c.tolist()}")
    warped = 4.0 * np.tan(np.pi * c / 2.0)
    if filter_type == "lowpass":
        z, p, k = _lp2lp(z, p, k, warped[0])
    elif filter_type == "bandpass":
        z, p, k = _lp2bp(z, p, k, math.sqrt(warped[0] * warped[1]), warped[1] - warped[0])
    else:
        raise ValueError(f"filter_type {filter_type!r} is not supported by the FP64 designer")
    return zpk_to_sos(*_bilinear(z, p, k, 2.0))


def butterworth_sos(order: int, cutoff, filter_type: str = "lowpass") -> np.ndarray:
    """Butterworth SOS, as butterworth_design(order, cutoff, filter_type,
    output="sos") computes it (_butterworth_design.py:90-174), in float64."""
    if order < 1:
        raise ValueError(f"Filter order must be positive, got {order}")
    return _design(*_butterworth_prototype(order), cutoff, filter_type)


def chebyshev1_sos(order: int, cutoff, ripple_db: float,
                   filter_type: str = "lowpass") -> np.ndarray:
    """Chebyshev type I SOS, as chebyshev_type_1_design(order, cutoff,
    passband_ripple_db, filter_type) computes it (_chebyshev_type_1_design.py:95-178)."""
    if order < 1:
        raise ValueError(f"Filter order must be positive, got {order}")
    return _design(*_chebyshev1_prototype(order, ripple_db), cutoff, filter_type)


def batched_butterworth_sos(order: int, cutoffs) -> np.ndarray:
    """(n, S, 6) lowpass Butterworth designs, one per cutoff, as
    batched_butterworth_design stacks them (_batched.py:83-140)."""
    cutoffs = np.atleast_1d(np.asarray(cutoffs, dtype=np.float64))
    return np.stack([butterworth_sos(order, float(c)) for c in cutoffs])
