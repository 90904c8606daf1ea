"""Impulse, step and frequency responses through the HIP path (GPU).

Mirrors the reference's tests (test__analysis_functions.py:180-380,
test__frequency_response.py:21-135, 176-290): known FIR coefficients and
cumulative sums, IIR responses against scipy, the Butterworth -3 dB point,
unit DC gain, stopband attenuation, the a0 validation, whole /
sampling-frequency grids; the direct-form and FIR responses against
scipy.signal.freqz and against the section form; plus the batched form
against the oracle per filter.  FP64 at 1e-10 unless
the reference's own test states a looser bound.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _F():
    from torchscience_amd.signal_processing import filter as F
    return F


def test_impulse_and_step_fir():
    F, _ = _F(), _dev()
    b = torch.tensor([0.25, 0.5, 0.25], dtype=torch.float64)
    t, h = F.impulse_response(b, n_samples=5)
    assert h.device.type == "cpu" and h.dtype == torch.float64
    torch.testing.assert_close(h, torch.tensor([0.25, 0.5, 0.25, 0.0, 0.0], dtype=torch.float64))
    torch.testing.assert_close(t, torch.arange(5, dtype=torch.float64))
    _, s = F.step_response(b, n_samples=5)
    torch.testing.assert_close(s, torch.tensor([0.25, 0.75, 1.0, 1.0, 1.0], dtype=torch.float64))
    _, h32 = F.impulse_response(b, dtype=torch.float32)
    assert h32.dtype == torch.float32 and h32.shape == (100,)
    bl = torch.tensor(np.hanning(40), dtype=torch.float64)
    _, hl = F.impulse_response(bl, n_samples=50)
    np.testing.assert_allclose(hl.numpy(), np.concatenate([np.hanning(40), np.zeros(10)]),
                               rtol=1e-12, atol=1e-14)


def test_impulse_and_step_iir_match_scipy(oracle_mod):
    F, dev = _F(), _dev()
    import scipy.signal as ss
    for order in (2, 4, 7):
        b, a = ss.butter(order, 0.25)
        imp = np.zeros(120)
        imp[0] = 1.0
        _, h = F.impulse_response(torch.tensor(b), torch.tensor(a), n_samples=120)
        np.testing.assert_allclose(h.numpy(), ss.lfilter(b, a, imp), rtol=1e-10, atol=1e-10)
        _, s = F.step_response(torch.tensor(b), torch.tensor(a), n_samples=120,
                               device=dev)
        assert s.device.type == "cuda"
        np.testing.assert_allclose(s.cpu().numpy(), ss.lfilter(b, a, np.ones(120)), rtol=1e-10,
                                   atol=1e-10)
    # a0 != 1 is normalised (_impulse_response.py:76-78)
    b2, a2 = np.array([0.5, 0.25]), np.array([2.0, -0.5])
    _, h = F.impulse_response(torch.tensor(b2), torch.tensor(a2), n_samples=30)
    np.testing.assert_allclose(h.numpy(), oracle_mod.lfilter(b2, a2, np.eye(1, 30)[0]),
                               rtol=1e-12, atol=1e-14)
    _, s = F.step_response(torch.tensor(b), torch.tensor(a), n_samples=400)
    assert abs(float(s[-1]) - 1.0) < 1e-6   # lowpass steady state


def test_impulse_response_sos(oracle_mod):
    F, _ = _F(), _dev()
    import scipy.signal as ss
    sos = ss.butter(6, 0.3, output="sos")
    _, h = F.impulse_response_sos(torch.tensor(sos), n_samples=50)
    np.testing.assert_allclose(h.numpy(), ss.sosfilt(sos, np.eye(1, 50)[0]), rtol=1e-10,
                               atol=1e-10)
    fir = torch.tensor([[0.5, 0.3, 0.2, 1.0, 0.0, 0.0]], dtype=torch.float64)
    _, h = F.impulse_response_sos(fir, n_samples=5)
    torch.testing.assert_close(h, torch.tensor([0.5, 0.3, 0.2, 0.0, 0.0], dtype=torch.float64))
    odd = np.array([[0.5, 0.2, 0.1, 2.0, 0.3, 0.1], [1.0, -0.4, 0.2, 1.0, -0.2, 0.05]])
    _, h = F.impulse_response_sos(torch.tensor(odd), n_samples=40)
    np.testing.assert_allclose(h.numpy(), oracle_mod.impulse_response_sos(odd, 40),
                               rtol=1e-12, atol=1e-14)
    _, h = F.impulse_response_sos(torch.tensor(sos), n_samples=100)
    assert h.shape == (100,)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("cutoff", [0.1, 0.3, 0.5, 0.7])
def test_frequency_response_matches_scipy(order, cutoff):
    F, dev = _F(), _dev()
    import scipy.signal as ss
    sos = torch.tensor(ss.butter(order, cutoff, output="sos"), device=dev)
    freqs, h = F.frequency_response_sos(sos, frequencies=256)
    w, hs = ss.sosfreqz(sos.cpu().numpy(), worN=256)
    assert h.dtype == torch.complex128 and h.device == sos.device
    np.testing.assert_allclose(h.cpu().numpy(), hs, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(freqs.cpu().numpy() * np.pi, w, rtol=1e-12, atol=1e-15)


def test_frequency_response_reference_checks():
    F, dev = _F(), _dev()
    import scipy.signal as ss
    sos = torch.tensor(ss.butter(4, 0.25, output="sos"), device=dev)
    _, h = F.frequency_response_sos(sos, frequencies=torch.tensor([0.25], device=dev))
    assert abs(20 * np.log10(abs(h.item())) + 3.0103) < 0.01
    sos = torch.tensor(ss.butter(4, 0.5, output="sos"), device=dev)
    _, h = F.frequency_response_sos(sos, frequencies=torch.tensor([0.0], device=dev))
    assert abs(abs(h.item()) - 1.0) < 1e-10
    sos = torch.tensor(ss.butter(4, 0.2, output="sos"), device=dev)
    _, h = F.frequency_response_sos(sos, frequencies=torch.tensor([0.8], device=dev))
    assert 20 * np.log10(abs(h.item())) < -40.0
    bad = torch.tensor([[1.0, 0.5, 0.25, 2.0, 0.1, 0.05]], device=dev)
    with pytest.raises(F.SOSNormalizationError):
        F.frequency_response_sos(bad, validate=True)
    f, h = F.frequency_response_sos(bad, validate=False)
    assert h.shape == f.shape and h.dtype == torch.complex64
    s2 = torch.tensor(ss.butter(2, 0.3, output="sos"), device=dev)
    fh, _ = F.frequency_response_sos(s2, frequencies=100, whole=False)
    fw, _ = F.frequency_response_sos(s2, frequencies=100, whole=True)
    assert fh[-1].item() == pytest.approx(0.99, rel=0.01)
    assert fw[-1].item() == pytest.approx(1.98, rel=0.01)
    ff, hz = F.frequency_response_sos(s2, frequencies=100, sampling_frequency=44100.0)
    assert ff[0].item() == 0.0 and ff[-1].item() == pytest.approx(22050 * 99 / 100, rel=0.01)
    _, hn = F.frequency_response_sos(s2, frequencies=100)
    np.testing.assert_allclose(hz.cpu().numpy(), hn.cpu().numpy(), rtol=1e-12, atol=1e-14)


def test_frequency_response_batched(oracle_mod):
    """(F, S, 6) in one launch equals the oracle per filter."""
    F, dev = _F(), _dev()
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    rng = np.random.default_rng(4)
    sos = batched_butterworth_sos(8, rng.uniform(0.02, 0.45, 300))
    w = np.linspace(0, np.pi, 1000, endpoint=False)
    _, h = F.frequency_response_sos(torch.tensor(sos, device=dev),
                                    frequencies=torch.tensor(w / np.pi, device=dev))
    assert h.shape == (300, 1000)
    ref = np.stack([oracle_mod.sos_frequency_response(s, w) for s in sos])
    np.testing.assert_allclose(h.cpu().numpy(), ref, rtol=1e-9, atol=1e-12)
    _, h32 = F.frequency_response_sos(torch.tensor(sos, dtype=torch.float32, device=dev),
                                      frequencies=64)
    assert h32.dtype == torch.complex64 and h32.shape == (300, 64)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 8])
def test_frequency_response_ba_matches_scipy_freqz(order, oracle_mod):
    F, dev = _F(), _dev()
    import scipy.signal as ss
    b, a = ss.butter(order, 0.3)
    a = a * 1.7
    freqs, h = F.frequency_response(torch.tensor(b, device=dev), torch.tensor(a, device=dev),
                                    frequencies=256)
    w, hs = ss.freqz(b, a, worN=256)
    assert h.dtype == torch.complex128 and h.device.type == "cuda" and h.shape == (256,)
    assert freqs.dtype == torch.float64
    np.testing.assert_allclose(h.cpu().numpy(), hs, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(h.cpu().numpy(), oracle_mod.ba_frequency_response(b, a, w),
                               rtol=1e-12, atol=1e-14)


def test_frequency_response_ba_matches_sections():
    """Direct form and sections of one Butterworth design give one response
    (test__frequency_response.py:195-209, 262-290)."""
    F, dev = _F(), _dev()
    import scipy.signal as ss
    for order in (2, 3, 4, 5):
        sos = torch.tensor(ss.butter(order, 0.4, output="sos"), device=dev)
        b, a = ss.butter(order, 0.4)
        _, hs = F.frequency_response_sos(sos, frequencies=128)
        _, hb = F.frequency_response(torch.tensor(b, device=dev), torch.tensor(a, device=dev),
                                     frequencies=128)
        np.testing.assert_allclose(hb.cpu().numpy(), hs.cpu().numpy(), rtol=1e-9, atol=1e-12)


def test_frequency_response_fir():
    """Moving average against freqz, the DC gain sum(h), a type I FIR's
    linear phase, float32 in -> complex64 out, and the grid options
    (test__frequency_response.py:211-257)."""
    F, dev = _F(), _dev()
    import scipy.signal as ss
    h = torch.full((5,), 0.2, dtype=torch.float64, device=dev)
    _, r = F.frequency_response_fir(h, frequencies=256)
    _, hs = ss.freqz(h.cpu().numpy(), [1.0], worN=256)
    np.testing.assert_allclose(r.cpu().numpy(), hs, rtol=1e-12, atol=1e-14)
    h32 = torch.tensor([0.1, 0.2, 0.4, 0.2, 0.1], device=dev)
    f, r = F.frequency_response_fir(h32, frequencies=torch.tensor([0.0], device=dev))
    assert r.dtype == torch.complex64 and f.dtype == torch.float32
    assert abs(abs(r.item()) - h32.sum().item()) < 1e-6
    h64 = h32.double()
    f, r = F.frequency_response_fir(h64, frequencies=50)
    w = f.cpu().numpy() * np.pi
    # symmetric about tap 2: H = real amplitude * exp(-2jw)
    amp = (r.cpu().numpy() * np.exp(2j * w))
    np.testing.assert_allclose(amp.imag, 0.0, atol=1e-14)
    fw, rw = F.frequency_response_fir(h64, frequencies=100, whole=True, sampling_frequency=8.0)
    assert fw[-1].item() == pytest.approx(8.0 * 99 / 100)
    _, rn = F.frequency_response_fir(h64, frequencies=100, whole=True)
    np.testing.assert_allclose(rw.cpu().numpy(), rn.cpu().numpy(), rtol=1e-12, atol=1e-14)
    _, rc = F.frequency_response_fir(h64.cpu(), frequencies=16)
    assert rc.device.type == "cpu"
    with pytest.raises(NotImplementedError):
        F.frequency_response(h64.to(torch.complex128), torch.ones(1, device=dev))
