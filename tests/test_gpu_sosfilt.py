"""Parity of the HIP sosfilt / lfilter path against the FP64 oracle (GPU).

Tolerances: FP64 compute matches the scipy goldens at the reference tests'
rtol = atol = 1e-10; FP32 compute must stay within 1e-5 per-channel normwise
relative error max|y - y64| / max|y64| against the FP64 oracle run on the
same (float32-rounded) coefficients and inputs.
"""

import numpy as np
import pytest
import torch

from tests.conftest import golden_cases

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _filt():
    from torchscience_amd.signal_processing import filter as F
    return F


def relerr(y, y64):
    y = np.asarray(y, dtype=np.float64).reshape(-1, np.shape(y)[-1])
    y64 = np.asarray(y64, dtype=np.float64).reshape(y.shape)
    num = np.max(np.abs(y - y64), axis=-1)
    den = np.maximum(np.max(np.abs(y64), axis=-1), 1e-300)
    return float(np.max(num / den))


def test_sosfilt_goldens_fp64(golden):
    F, dev = _filt(), _dev()
    for name, c in golden_cases(golden, "sos").items():
        print("case", name, flush=True)
        sos = torch.tensor(c["sos"], device=dev)
        x = torch.tensor(c["x"], device=dev)
        axis = int(c["axis"])
        tol = 1e-9 if name == "butter12_many" else 1e-10
        if "zi" in c:
            y, zf = F.sosfilt(sos, x, axis=axis, zi=torch.tensor(c["zi"], device=dev))
        else:
            y = F.sosfilt(sos, x, axis=axis)
        assert y.dtype == torch.float64 and y.shape == x.shape
        if name == "cheby1_16_near_unit":
            # poles at r = 0.99972 and outputs up to ~2e3: FP64 rounding is
            # amplified ~1e6 by the noise gain, so compare normwise per channel
            # (max|y - y64| / max|y64|), the error measure of the FP32 claims
            assert relerr(y.cpu().numpy(), c["y"]) < 1e-9, name
            assert relerr(zf.cpu().numpy().reshape(2, -1), c["zf"].reshape(2, -1)) < 1e-9, name
            continue
        if "zi" in c:
            np.testing.assert_allclose(zf.cpu().numpy(), c["zf"], rtol=tol, atol=tol, err_msg=name)
        np.testing.assert_allclose(y.cpu().numpy(), c["y"], rtol=tol, atol=tol, err_msg=name)


def test_lfilter_goldens_fp64(golden):
    F, dev = _filt(), _dev()
    for name, c in golden_cases(golden, "lf").items():
        print("case", name, flush=True)
        b = torch.tensor(c["b"], device=dev)
        a = torch.tensor(c["a"], device=dev)
        x = torch.tensor(c["x"], device=dev)
        if "zi" in c:
            y, zf = F.lfilter(b, a, x, zi=torch.tensor(c["zi"], device=dev))
            np.testing.assert_allclose(zf.cpu().numpy(), c["zf"], rtol=1e-10, atol=1e-10,
                                       err_msg=name)
        else:
            y = F.lfilter(b, a, x)
        np.testing.assert_allclose(y.cpu().numpy(), c["y"], rtol=1e-10, atol=1e-10,
                                   err_msg=name)


def test_config1_plain_fp64(golden):
    F, dev = _filt(), _dev()
    g = torch.Generator().manual_seed(0)
    x = torch.randn(16, 4096, dtype=torch.float64, generator=g)
    sos = torch.tensor(golden["c1/sos"])
    y, zf = F.sosfilt(sos.to(dev), x.to(dev), zi=torch.zeros(16, 4, 2, dtype=torch.float64,
                                                             device=dev))
    np.testing.assert_allclose(y.cpu().numpy(), golden["c1/y"], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("T", [1, 3, 63, 64, 65, 4095, 4096, 4097, 8191, 20000, 70001])
def test_fp32_lengths_vs_oracle(oracle_mod, T):
    """Ragged lengths around lane (64) and window (4096) boundaries, with zi."""
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    rng = np.random.default_rng(T)
    B = 5
    x = rng.standard_normal((B, T)).astype(np.float32)
    zi = rng.standard_normal((B, 6, 2)).astype(np.float32)
    print("T", T, flush=True)
    y, zf = F.sosfilt(torch.tensor(sos, device=dev), torch.tensor(x, device=dev),
                      zi=torch.tensor(zi, device=dev))
    y64, zf64 = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.astype(np.float64),
                                        zi.astype(np.float64))
    assert relerr(y.cpu().numpy(), y64) < F32_TOL
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-4, atol=1e-4)


def test_fp64_long_sequence_lookback(oracle_mod):
    """2 channels x 1.5M samples: hundreds of windows chained by look-back."""
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass")
    rng = np.random.default_rng(7)
    x = rng.standard_normal((2, 1_500_001))
    zi = rng.standard_normal((2, 6, 2))
    y, zf = F.sosfilt(torch.tensor(sos, device=dev), torch.tensor(x, device=dev),
                      zi=torch.tensor(zi, device=dev))
    y64, zf64 = oracle_mod.sosfilt_rows(sos, x, zi, n_threads=2)
    np.testing.assert_allclose(y.cpu().numpy(), y64, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-9, atol=1e-9)


def test_fp32_near_unit_circle(oracle_mod):
    """Order-16 Chebyshev I at cutoff 0.01 (poles at r = 0.99972)."""
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import chebyshev1_sos
    sos = chebyshev1_sos(16, 0.01, 1.0)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((4, 300_000)).astype(np.float32)
    y = F.sosfilt(torch.tensor(sos, device=dev), torch.tensor(x, device=dev))  # fp64 compute
    y64, _ = oracle_mod.sosfilt_rows(sos, x.astype(np.float64), n_threads=4)
    # FP64 rounding amplified by the filter's noise gain ~1/(theta^2 (1-r)) ~ 4e6
    # per section: two correct FP64 evaluation orders differ by ~1e-9 here
    assert relerr(y.cpu().numpy(), y64) < 2e-8
    y32 = F.sosfilt(torch.tensor(sos, device=dev), torch.tensor(x, device=dev), precision="fp32")
    y64b, _ = oracle_mod.sosfilt_rows(sos.astype(np.float32).astype(np.float64),
                                      x.astype(np.float64), n_threads=4)
    err = relerr(y32.cpu().numpy(), y64b)
    # the reference's own FP32 arithmetic on the same input (its DF2T loop in
    # float32): rounding inside the loop is amplified by the noise gain
    ref32, _ = oracle_mod.sosfilt_rows_f32(sos.astype(np.float32), x, n_threads=4)
    err_ref = relerr(ref32, y64b)
    print("fp32 cheby16 relerr", err, "reference fp32", err_ref)
    # poles beyond radius 0.995 run the DF2T form with FP64 carries: no worse
    # than the reference's FP32 loop
    assert err <= err_ref
    # the compensated form consumes the float64 coefficients as given
    yc = F.sosfilt(torch.tensor(sos, device=dev), torch.tensor(x, device=dev), precision="fp32c")
    assert yc.dtype == torch.float32
    errc = relerr(yc.cpu().numpy(), y64)
    print("fp32c cheby16 relerr", errc)
    assert errc < 1e-5


def test_fp32_routing_follows_the_storage(oracle_mod):
    """The FP32 routing decision (DF2T with FP64 carries for poles beyond
    radius 0.995) is cached per coefficient storage: a coefficient tensor
    allocated where a freed one of the same size lived (the caching
    allocator usually hands out the same block) is judged again, so the
    order-16 Chebyshev I after a Butterworth stays no worse than the
    reference's FP32 loop."""
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import (butterworth_sos,
                                                                    chebyshev1_sos)
    x = np.random.default_rng(8).standard_normal((4, 300_000)).astype(np.float32)
    xt = torch.tensor(x, device=dev)
    cheb = chebyshev1_sos(16, 0.01, 1.0).astype(np.float32)
    y64, _ = oracle_mod.sosfilt_rows(cheb.astype(np.float64), x.astype(np.float64), n_threads=4)
    ref32, _ = oracle_mod.sosfilt_rows_f32(cheb, x, n_threads=4)
    err_ref = relerr(ref32, y64)
    for _ in range(2):
        bw = torch.tensor(butterworth_sos(16, 0.2).astype(np.float32), device=dev)
        assert bw.shape == cheb.shape
        F.sosfilt(bw, xt[:, :5000])
        addr = bw.data_ptr()
        del bw
        st = torch.tensor(cheb, device=dev)
        print("coefficient address reused:", st.data_ptr() == addr, flush=True)
        err = relerr(F.sosfilt(st, xt).cpu().numpy(), y64)
        print("fp32 cheby16 relerr", err, "reference fp32", err_ref, flush=True)
        assert err <= err_ref
        del st


def test_batched_per_channel_filters(oracle_mod):
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    rng = np.random.default_rng(9)
    cut = rng.uniform(0.02, 0.45, 37)
    filt = batched_butterworth_sos(8, cut).astype(np.float32)
    x = rng.standard_normal((37, 9000)).astype(np.float32)
    y = F.batched_filter_apply(torch.tensor(filt, device=dev), torch.tensor(x, device=dev),
                               broadcast=False)
    y64, _ = oracle_mod.sosfilt_rows(filt.astype(np.float64), x.astype(np.float64),
                                     chans_per_filter=1, n_threads=4)
    assert relerr(y.cpu().numpy(), y64) < F32_TOL
    # broadcast=True: every filter on the same (3, 5000) signal
    xs = rng.standard_normal((3, 5000))
    yb = F.batched_filter_apply(torch.tensor(filt[:4].astype(np.float64), device=dev),
                                torch.tensor(xs, device=dev))
    assert yb.shape == (4, 3, 5000)
    for f in range(4):
        ref = oracle_mod.sosfilt(filt[f].astype(np.float64), xs)
        np.testing.assert_allclose(yb[f].cpu().numpy(), ref, rtol=1e-10, atol=1e-10)


def test_zf_continuation_and_steady_state():
    F, dev = _filt(), _dev()
    import scipy.signal as ss
    sos = torch.tensor(ss.butter(4, 0.3, output="sos"), device=dev)
    x = torch.randn(3, 50_000, dtype=torch.float64, device=dev)
    y_full = F.sosfilt(sos, x)
    y1, zf1 = F.sosfilt(sos, x[:, :12_345], zi=torch.zeros(3, 2, 2, dtype=torch.float64,
                                                          device=dev))
    y2, _ = F.sosfilt(sos, x[:, 12_345:], zi=zf1)
    torch.testing.assert_close(torch.cat([y1, y2], -1), y_full, rtol=1e-10, atol=1e-10)
    sos6 = torch.tensor(ss.butter(6, 0.2, output="sos"), device=dev)
    y, _ = F.sosfilt(sos6, torch.ones(50, dtype=torch.float64, device=dev),
                     zi=F.sosfilt_zi(sos6))
    torch.testing.assert_close(y, torch.ones_like(y), rtol=1e-10, atol=1e-10)


def test_lfilter_empty_time_axis_keeps_state():
    """No samples: zf = zi on every lfilter path (one section, FIR, general
    order), as the reference's loop runs no step (_lfilter.py:153-175)."""
    F, dev = _filt(), _dev()
    x = torch.empty(3, 0, dtype=torch.float64, device=dev)
    for b, a in (([0.2, 0.3, 0.1], [1.0, -0.4, 0.1]),            # one section
                 ([0.1] * 9, [1.0]),                               # FIR
                 ([0.1, 0.2, 0.1, 0.05, 0.02], [1.0, -0.5, 0.2, -0.05, 0.01])):  # order 4
        N = max(len(b), len(a)) - 1
        zi = torch.randn(3, N, dtype=torch.float64, device=dev)
        y, zf = F.lfilter(torch.tensor(b, dtype=torch.float64, device=dev),
                          torch.tensor(a, dtype=torch.float64, device=dev), x, zi=zi)
        assert y.shape == (3, 0) and zf.shape == (3, N)
        torch.testing.assert_close(zf, zi, rtol=0, atol=0)


def test_edge_shapes_and_dtypes():
    F, dev = _filt(), _dev()
    import scipy.signal as ss
    sos64 = torch.tensor(ss.butter(3, 0.3, output="sos"), device=dev)
    # empty batch, empty time, zf = zi for T = 0
    assert F.sosfilt(sos64, torch.empty(0, 10, dtype=torch.float64, device=dev)).shape == (0, 10)
    zi = torch.randn(4, 2, 2, dtype=torch.float64, device=dev)
    y, zf = F.sosfilt(sos64[:2], torch.empty(4, 0, dtype=torch.float64, device=dev), zi=zi)
    assert y.shape == (4, 0)
    torch.testing.assert_close(zf, zi)
    # dtype promotion: sos f64, x f32 -> f64; zf keeps zi's dtype
    x32 = torch.randn(2, 300, device=dev)
    y = F.sosfilt(sos64, x32)
    assert y.dtype == torch.float64
    # sos64 has 2 sections; zi (2, 2, 2) = x's batch (2,) + (S, 2)
    y, zf = F.sosfilt(sos64.float(), x32, zi=torch.zeros(2, 2, 2, dtype=torch.float64,
                                                         device=dev))
    assert y.dtype == torch.float32 and zf.dtype == torch.float64
    # complex input, real coefficients
    xc = torch.randn(100, dtype=torch.complex128)
    yc = F.sosfilt(sos64, xc.to(dev))
    ref = ss.sosfilt(sos64.cpu().numpy(), xc.numpy())
    np.testing.assert_allclose(yc.cpu().numpy(), ref, rtol=1e-10, atol=1e-10)
    # sos with no sections
    xs = torch.randn(3, 10, device=dev)
    torch.testing.assert_close(F.sosfilt(torch.zeros(0, 6, device=dev), xs), xs)


def test_deterministic_repeat():
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    sos = torch.tensor(butterworth_sos(6, [0.05, 0.3], "bandpass"), dtype=torch.float32,
                       device=dev)
    x = torch.randn(64, 100_000, device=dev)
    y0 = F.sosfilt(sos, x)
    for _ in range(5):
        assert torch.equal(F.sosfilt(sos, x), y0)
    from torchscience_amd import check_errors
    check_errors(dev)


def test_bf16_io(oracle_mod):
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    sos = butterworth_sos(4, 0.2).astype(np.float32)
    x = torch.randn(8, 10_000, device=dev).bfloat16()
    y = F.sosfilt(torch.tensor(sos, device=dev), x, precision="bf16io")
    assert y.dtype == torch.bfloat16
    y64, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.float().cpu().numpy())
    assert relerr(y.float().cpu().numpy(), y64) < 1e-2  # bf16 output rounding (2^-8)


def test_lfilter_general_order_and_complex(golden):
    F, dev = _filt(), _dev()
    import scipy.signal as ss
    b, a = ss.butter(7, 0.3)
    x = torch.randn(4, 3000, dtype=torch.float64)
    zi = torch.randn(4, 7, dtype=torch.float64)
    y, zf = F.lfilter(torch.tensor(b, device=dev), torch.tensor(a, device=dev), x.to(dev),
                      zi=zi.to(dev))
    yr, zfr = ss.lfilter(b, a, x.numpy(), zi=zi.numpy())
    np.testing.assert_allclose(y.cpu().numpy(), yr, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(zf.cpu().numpy(), zfr, rtol=1e-9, atol=1e-9)
    h = torch.tensor(ss.firwin(51, 0.3), device=dev)
    xf = torch.randn(2000, dtype=torch.float64, device=dev)
    yf = F.lfilter(h, torch.ones(1, dtype=torch.float64, device=dev), xf)
    np.testing.assert_allclose(yf.cpu().numpy(), np.convolve(xf.cpu().numpy(), h.cpu().numpy())[:2000],
                               rtol=1e-10, atol=1e-10)
    xc = torch.randn(100, dtype=torch.complex128)
    bb = torch.tensor([0.5, 0.5], dtype=torch.float64)
    aa = torch.tensor([1.0, -0.3], dtype=torch.float64)
    yc = F.lfilter(bb.to(dev), aa.to(dev), xc.to(dev))
    np.testing.assert_allclose(yc.cpu().numpy(), ss.lfilter(bb.numpy(), aa.numpy(), xc.numpy()),
                               rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("order", [3, 8, 16])
def test_lfilter_general_chunked_matches_serial_oracle(oracle_mod, order):
    """Orders up to 16 over at least 512 samples run in time-parallel form
    (chunks of 256 from zero state, FP64 start-state scan, second pass):
    equal to the oracle's serial DF2T, zi per row and shared, zf, including
    a partial last chunk."""
    F, dev = _filt(), _dev()
    import scipy.signal as ss
    b, a = ss.butter(order, 0.23)
    a = a * 1.5      # a0 != 1: normalised
    b = b * 1.5
    rng = np.random.default_rng(order)
    for T in (512, 256 * 9 + 37, 65536 + 5):
        print("case lfilter chunked", order, T, flush=True)
        x = rng.standard_normal((3, T))
        for zi in (rng.standard_normal((3, order)), rng.standard_normal(order)):
            y, zf = F.lfilter(torch.tensor(b, device=dev), torch.tensor(a, device=dev),
                              torch.tensor(x, device=dev), zi=torch.tensor(zi, device=dev))
            yr, zfr = oracle_mod.lfilter(b, a, x, zi=np.broadcast_to(zi, (3, order)))
            scale = np.max(np.abs(yr))
            assert np.max(np.abs(y.cpu().numpy() - yr)) < 1e-9 * scale, (order, T)
            assert np.max(np.abs(zf.cpu().numpy() - zfr)) < 1e-9 * max(np.max(np.abs(zfr)), 1)
        if order <= 8:
            # (an order-16 direct form is beyond FP32 arithmetic in any order
            # of evaluation)
            y32 = F.lfilter(torch.tensor(b, dtype=torch.float32, device=dev),
                            torch.tensor(a, dtype=torch.float32, device=dev),
                            torch.tensor(x, dtype=torch.float32, device=dev))
            y64 = oracle_mod.lfilter(b.astype(np.float32), a.astype(np.float32),
                                     x.astype(np.float32))
            assert relerr(y32.cpu().numpy(), y64) < 1e-4, (order, T)


def test_lfiltic_continuation_on_device():
    """lfilter over a signal in two parts, the second from lfiltic of the
    first part's last outputs and inputs (the use shown in
    _lfilter.py:224-230), equals one pass: order 2 (one section), order 5
    (general form), a 5-tap FIR, with a0 != 1."""
    F, dev = _filt(), _dev()
    import scipy.signal as ss
    rng = np.random.default_rng(21)
    b5, a5 = ss.butter(5, 0.2)
    cases = [(np.array([0.1, 0.2, 0.1]), np.array([1.0, -0.5, 0.1])),
             (b5 * 2.0, a5 * 2.0), (np.array([0.2, 0.3, 0.1, -0.1, 0.05]), np.array([1.0]))]
    for b, a in cases:
        N = max(b.size, a.size) - 1
        x = torch.tensor(rng.standard_normal(700), device=dev)
        bt, at = torch.tensor(b, device=dev), torch.tensor(a, device=dev)
        full = F.lfilter(bt, at, x)
        y1 = F.lfilter(bt, at, x[:300])
        zi = F.lfiltic(bt, at, y1.flip(0)[:N], x[:300].flip(0)[:N])
        assert zi.shape == (N,) and zi.device == x.device
        y2, _ = F.lfilter(bt, at, x[300:], zi=zi)
        np.testing.assert_allclose(torch.cat([y1, y2]).cpu().numpy(), full.cpu().numpy(),
                                   rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("B", [5, 1100])
def test_compensated_form_carries_float64_state(oracle_mod, B):
    """fp32c with a float64 zi keeps the state in float64 between chunks
    (TSIIR_STATE_F64; zf has zi's dtype): streaming three uneven chunks equals
    one-shot FP64 filtering of the whole signal (few channels: windows with
    the look-back; many: chained windows)."""
    F, dev = _filt(), _dev()
    from torchscience_amd.signal_processing.filter._design import chebyshev1_sos
    sos = chebyshev1_sos(8, 0.02, 1.0)
    rng = np.random.default_rng(B)
    lens = [5000, 70001, 12345]
    x = rng.standard_normal((B, sum(lens))).astype(np.float32)
    zi = torch.zeros(B, 4, 2, dtype=torch.float64, device=dev)
    st = torch.tensor(sos, device=dev)
    ys, t0 = [], 0
    for n in lens:
        y, zi = F.sosfilt(st, torch.tensor(x[:, t0:t0 + n], device=dev), zi=zi,
                          precision="fp32c")
        assert y.dtype == torch.float32 and zi.dtype == torch.float64
        ys.append(y.cpu().numpy())
        t0 += n
    ref, zf_ref = oracle_mod.sosfilt_rows(sos, x.astype(np.float64), n_threads=8)
    err = relerr(np.concatenate(ys, axis=1), ref)
    print("fp32c streamed with float64 state", B, err)
    assert err < 1e-5
    np.testing.assert_allclose(zi.cpu().numpy(), zf_ref, rtol=1e-6,
                               atol=1e-6 * np.max(np.abs(zf_ref)))
