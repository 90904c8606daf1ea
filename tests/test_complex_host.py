"""The host algebra of complex filter coefficients (CPU): lfilter and
sosfilt with complex coefficients run as four real filters on the
conjugate-extended polynomials (numerator B conj(A), real denominator
A conj(A)), with zf from the DF2T state definition.  Here the real filters
are the FP64 oracle's lfilter (a test-only substitute for the HIP kernels,
which need a GPU), so the decomposition, the state polynomial Z conj(A), the
zf formula and the a0 normalisation are checked against scipy.signal's
complex DF2T on the CPU: orders 0-8, FIR and full denominators, a0 != 1,
real and complex x, zi per row and shared, T < order, another axis."""

import numpy as np
import pytest
import torch


@pytest.fixture
def host_complex(monkeypatch, oracle_mod):
    import torchscience_amd.signal_processing.filter._lfilter as L
    real = L.lfilter

    def oracle_real(b, a, x, axis=-1, zi=None):
        if b.is_complex() or a.is_complex():
            return real(b, a, x, axis, zi)
        out = oracle_mod.lfilter(b.detach().numpy(), a.detach().numpy(), x.detach().numpy(),
                                 axis=axis, zi=None if zi is None else zi.detach().numpy())
        if zi is None:
            return torch.from_numpy(np.asarray(out))
        return torch.from_numpy(np.asarray(out[0])), torch.from_numpy(np.asarray(out[1]))

    monkeypatch.setattr(L, "lfilter", oracle_real)
    monkeypatch.setattr(L, "_require_device", lambda *t: torch.device("cpu"))
    return L


def _stable(rng, na):
    a = 0.3 * (rng.standard_normal(na) + 1j * rng.standard_normal(na))
    a[0] = 1.0
    if na > 1:
        r = np.max(np.abs(np.roots(a)))
        if r > 0.9:
            a[1:] = a[1:] * (0.9 / r) ** np.arange(1, na)
    return a * (1.2 - 0.4j)


@pytest.mark.parametrize("N", range(0, 9))
def test_complex_lfilter_algebra(host_complex, N):
    import scipy.signal as ss
    L = host_complex
    rng = np.random.default_rng(40 + N)
    for na in sorted({1, max(1, N), N + 1}):
        for T in (1, 3, 7, 120):
            for cx in (False, True):
                b = rng.standard_normal(N + 1) + 1j * rng.standard_normal(N + 1)
                a = _stable(rng, na)
                x = rng.standard_normal((2, T)) + (1j * rng.standard_normal((2, T)) if cx else 0)
                order = max(N + 1, na) - 1
                bt, at, xt = torch.from_numpy(b), torch.from_numpy(a), torch.from_numpy(x)
                y = L._lfilter_complex_coef(bt, at, xt, -1, None, order)
                np.testing.assert_allclose(y.numpy(), ss.lfilter(b, a, x), rtol=1e-10,
                                           atol=1e-10)
                if order == 0:
                    continue
                zi = rng.standard_normal((2, order)) + 1j * rng.standard_normal((2, order))
                y, zf = L._lfilter_complex_coef(bt, at, xt, -1, torch.from_numpy(zi), order)
                ys, zfs = ss.lfilter(b, a, x, zi=zi)
                np.testing.assert_allclose(y.numpy(), ys, rtol=1e-10, atol=1e-10)
                np.testing.assert_allclose(zf.numpy(), zfs, rtol=1e-10, atol=1e-10)
                # one state row for both channels
                y, zf = L._lfilter_complex_coef(bt, at, xt, -1, torch.from_numpy(zi[0]), order)
                ys, zfs = ss.lfilter(b, a, x, zi=np.broadcast_to(zi[0], (2, order)))
                np.testing.assert_allclose(y.numpy(), ys, rtol=1e-10, atol=1e-10)
                np.testing.assert_allclose(zf.numpy(), zfs, rtol=1e-10, atol=1e-10)


def test_complex_lfilter_other_axis_and_dtype(host_complex):
    import scipy.signal as ss
    L = host_complex
    rng = np.random.default_rng(3)
    b = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    a = _stable(rng, 3)
    x = rng.standard_normal((50, 4))
    y = L._lfilter_complex_coef(torch.from_numpy(b), torch.from_numpy(a), torch.from_numpy(x),
                                0, None, 2)
    np.testing.assert_allclose(y.numpy(), ss.lfilter(b, a, x, axis=0), rtol=1e-10, atol=1e-10)
    # complex64 coefficients with float32 x give complex64
    y = L._lfilter_complex_coef(torch.from_numpy(b.astype(np.complex64)),
                                torch.from_numpy(a.astype(np.complex64)),
                                torch.from_numpy(x.T.astype(np.float32).copy()), -1, None, 2)
    assert y.dtype == torch.complex64


def test_complex_sosfilt_algebra(host_complex):
    import scipy.signal as ss
    import torchscience_amd.signal_processing.filter._sosfilt as SF
    rng = np.random.default_rng(8)
    sos = np.array([[1.0, 0.5 + 0.2j, 0.1, 1.0, -0.3 + 0.1j, 0.05],
                    [0.7, -0.2, 0.3j, 1.2, 0.1 - 0.2j, 0.1],
                    [0.2 + 0.1j, 0.4, 0.2, 1.0, -0.5, 0.2 + 0.1j]])
    sn = sos / sos[:, 3:4]
    x = rng.standard_normal((3, 40)) + 1j * rng.standard_normal((3, 40))
    zi = rng.standard_normal((3, 3, 2)) + 1j * rng.standard_normal((3, 3, 2))
    y, zf = SF._sosfilt_complex_coef(torch.from_numpy(sos), torch.from_numpy(x), -1,
                                     torch.from_numpy(zi), None)
    ys, zfs = ss.sosfilt(sn, x, zi=zi.transpose(1, 0, 2))
    np.testing.assert_allclose(y.numpy(), ys, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(zf.numpy(), zfs.transpose(1, 0, 2), rtol=1e-10, atol=1e-10)
    # a real zi keeps the real part of the state (zf is allocated like zi)
    y, zf = SF._sosfilt_complex_coef(torch.from_numpy(sos), torch.from_numpy(x), -1,
                                     torch.from_numpy(zi.real.copy()), None)
    ys, zfs = ss.sosfilt(sn, x, zi=zi.real.transpose(1, 0, 2))
    assert zf.dtype == torch.float64
    np.testing.assert_allclose(zf.numpy(), zfs.real.transpose(1, 0, 2), rtol=1e-10, atol=1e-10)
    with pytest.raises(ValueError, match="sos must be shape"):
        SF._sosfilt_complex_coef(torch.from_numpy(sos[:, :5]), torch.from_numpy(x), -1, None,
                                 None)
    with pytest.raises(NotImplementedError):
        SF._sosfilt_complex_coef(torch.from_numpy(sos), torch.from_numpy(x), -1, None, "fp32")
