"""Batched IIR design through the HIP kernel (tsiir_iir_design_sos) against
the FP64 oracle's design chain, which is pinned against scipy's designs.

Coefficients are compared one by one (the same pairing rules fix the row
order) for Butterworth, Chebyshev I and II, every band type and orders 1-16;
plus the reference's batched-design cases (test__batched.py: one cutoff,
shapes, band types, orders, dtype, large batch, boundary cutoffs, the empty
batch) and its argument checks.
"""

import time

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


def _call(F, family, order, param, cut, band, **kw):
    if family == "butterworth":
        return F.batched_butterworth_design(order, cut, band, **kw)
    if family == "chebyshev1":
        return F.batched_chebyshev_type_1_design(order, param, cut, band, **kw)
    return F.batched_chebyshev_type_2_design(order, param, cut, band, **kw)


@pytest.mark.parametrize("family,param", [("butterworth", 0.0), ("chebyshev1", 0.5),
                                          ("chebyshev2", 45.0)])
@pytest.mark.parametrize("band", ["lowpass", "highpass", "bandpass", "bandstop"])
def test_design_matches_oracle(oracle_mod, family, param, band):
    F, dev = _F(), _dev()
    rng = np.random.default_rng(hash((family, band)) % 1000)
    for order in (1, 2, 3, 4, 5, 6, 8, 11, 16):
        print("case design", family, band, order, flush=True)
        if band in ("bandpass", "bandstop"):
            lo = rng.uniform(0.02, 0.6, 24)
            cut = np.stack([lo, lo + rng.uniform(0.05, 0.97 - lo)], axis=1)
        else:
            cut = rng.uniform(0.02, 0.95, 24)
        sos = _call(F, family, order, param, torch.tensor(cut, device=dev), band)
        ref = np.stack([oracle_mod.iir_design_sos(family, order, c, band, param) for c in cut])
        assert sos.shape == ref.shape and sos.dtype == torch.float64
        np.testing.assert_allclose(sos.cpu().numpy(), ref, rtol=1e-9, atol=1e-11,
                                   err_msg=f"{family} {band} {order}")


def test_reference_batched_cases(oracle_mod):
    F, dev = _F(), _dev()
    s = F.batched_butterworth_design(4, torch.tensor([0.3]))
    assert s.shape == (1, 2, 6) and s.device.type == "cpu" and s.dtype == torch.float64
    # the float32 cutoff is designed at its float64 value, as the reference does
    np.testing.assert_allclose(s[0].numpy(), oracle_mod.iir_design_sos(
        "butterworth", 4, float(np.float32(0.3))), rtol=1e-12, atol=1e-14)
    assert F.batched_butterworth_design(4, torch.linspace(0.1, 0.4, 10)).shape == (10, 2, 6)
    c2 = torch.tensor([[0.1, 0.3], [0.2, 0.4], [0.15, 0.35]])
    assert F.batched_butterworth_design(2, c2, filter_type="bandpass").shape == (3, 2, 6)
    assert F.batched_butterworth_design(3, c2, filter_type="bandstop").shape == (3, 3, 6)
    assert F.batched_butterworth_design(5, torch.tensor([0.2, 0.3]),
                                        filter_type="highpass").shape == (2, 3, 6)
    assert F.batched_chebyshev_type_1_design(4, 1.0, torch.linspace(0.1, 0.4, 5)).shape == \
        (5, 2, 6)
    assert F.batched_chebyshev_type_2_design(4, 40.0, torch.linspace(0.1, 0.4, 5)).shape == \
        (5, 2, 6)
    s32 = F.batched_butterworth_design(4, torch.tensor([0.2, 0.3]), dtype=torch.float32)
    assert s32.dtype == torch.float32
    sg = F.batched_butterworth_design(4, torch.tensor([0.2, 0.3]), device=dev)
    assert sg.device.type == "cuda"
    s = F.batched_butterworth_design(4, torch.linspace(0.05, 0.45, 100))
    assert s.shape == (100, 2, 6)
    s = F.batched_butterworth_design(4, torch.tensor([0.01, 0.5, 0.99]))
    assert torch.isfinite(s).all()
    s = F.batched_butterworth_design(4, torch.tensor([1000.0, 4000.0]),
                                     sampling_frequency=16000.0)
    np.testing.assert_allclose(s.numpy(), np.stack(
        [oracle_mod.iir_design_sos("butterworth", 4, c) for c in (0.125, 0.5)]),
        rtol=1e-12, atol=1e-14)
    s = F.batched_butterworth_design(4, torch.tensor([[0.2], [0.3]]))
    assert s.shape == (2, 2, 6)


def test_compiled_batched_design():
    """test__batched.py:447-457: batched_butterworth_design under
    torch.compile (the native call runs outside the traced graph) gives the
    eager result, shape (5, 2, 6) for order 4 and five cutoffs."""
    from torchscience_amd.signal_processing.filter import batched_butterworth_design
    dev = _dev()

    @torch.compile
    def design_filters(cutoffs):
        return batched_butterworth_design(4, cutoffs)

    cutoffs = torch.linspace(0.1, 0.4, 5, device=dev)
    sos = design_filters(cutoffs)
    assert sos.shape == (5, 2, 6)
    assert torch.equal(sos, batched_butterworth_design(4, cutoffs))


def test_design_errors():
    F, _ = _F(), _dev()
    with pytest.raises(ValueError, match="cutoffs tensor cannot be empty"):
        F.batched_butterworth_design(4, torch.tensor([]))
    with pytest.raises(ValueError, match="only supports 'sos' output"):
        F.batched_butterworth_design(4, torch.tensor([0.2]), output="ba")
    with pytest.raises(ValueError, match="cutoffs must have shape"):
        F.batched_butterworth_design(4, torch.tensor([0.2, 0.3]), filter_type="bandpass")
    with pytest.raises(ValueError, match="between 0 and 1"):
        F.batched_butterworth_design(4, torch.tensor([0.2, 1.2]))
    with pytest.raises(ValueError, match="0 < low < high < 1"):
        F.batched_butterworth_design(4, torch.tensor([[0.4, 0.3]]), filter_type="bandstop")
    with pytest.raises(ValueError, match="Filter order must be positive"):
        F.batched_butterworth_design(0, torch.tensor([0.2]))
    with pytest.raises(ValueError, match="Passband ripple must be positive"):
        F.batched_chebyshev_type_1_design(4, 0.0, torch.tensor([0.2]))
    with pytest.raises(ValueError, match="Stopband attenuation must be positive"):
        F.batched_chebyshev_type_2_design(4, -1.0, torch.tensor([0.2]))
    with pytest.raises(ValueError, match="Invalid filter_type"):
        F.batched_butterworth_design(4, torch.tensor([0.2]), filter_type="notch")


def test_design_4096_filters_and_timing(oracle_mod):
    """The order-8 lowpass bank of 4096 filters used for per-channel
    filtering: equal to the oracle, and the time of one launch."""
    F, dev = _F(), _dev()
    rng = np.random.default_rng(2)
    cut = torch.tensor(rng.uniform(0.02, 0.45, 4096), device=dev)
    s = F.batched_butterworth_design(8, cut)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(20):
        s = F.batched_butterworth_design(8, cut)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 20
    idx = [0, 1, 2047, 4095]
    ref = np.stack([oracle_mod.iir_design_sos("butterworth", 8, float(cut[i])) for i in idx])
    np.testing.assert_allclose(s[idx].cpu().numpy(), ref, rtol=1e-9, atol=1e-12)
    print(f"design of 4096 order-8 lowpass filters: {dt * 1e3:.3f} ms per call", flush=True)


@pytest.mark.parametrize("family,param", [("butterworth", 0.0), ("chebyshev1", 1.0),
                                          ("chebyshev2", 40.0)])
@pytest.mark.parametrize("band", ["lowpass", "highpass", "bandpass", "bandstop"])
def test_design_gradients_match_central_differences(oracle_mod, family, param, band):
    """d loss / d cutoffs from the kernel's Jacobian against FP64 central
    differences of the oracle's design (step 1e-6, agreement 1e-6 of the
    gradient's scale), and the reference's checks that gradients flow
    (test__batched.py:353-405)."""
    F, dev = _F(), _dev()
    rng = np.random.default_rng(7)
    two = band in ("bandpass", "bandstop")
    cut = np.array([[0.15, 0.45], [0.3, 0.7], [0.05, 0.2]]) if two else \
        np.array([0.12, 0.3, 0.61])
    for order in (2, 3, 5):
        print("case design grad", family, band, order, flush=True)
        c = torch.tensor(cut, requires_grad=True)
        sos = _call(F, family, order, param, c, band)
        wts = rng.standard_normal(tuple(sos.shape))
        (sos * torch.tensor(wts)).sum().backward()
        assert c.grad is not None and torch.isfinite(c.grad).all()

        def loss(cc):
            return float(np.sum(np.stack([oracle_mod.iir_design_sos(family, order, x, band,
                                                                    param) for x in cc]) * wts))
        h = 1e-6
        fd = np.zeros_like(cut)
        for idx in np.ndindex(cut.shape):
            cp, cm = cut.copy(), cut.copy()
            cp[idx] += h
            cm[idx] -= h
            fd[idx] = (loss(cp) - loss(cm)) / (2 * h)
        scale = max(np.max(np.abs(fd)), 1.0)
        np.testing.assert_allclose(c.grad.numpy(), fd, rtol=0, atol=1e-6 * scale,
                                   err_msg=f"{family} {band} {order}")
