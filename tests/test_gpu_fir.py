"""Parity of the HIP FIR path (tsiir_fir_forward / tsiir_fir_backward_taps)
against the FP64 oracle (GPU).

Covers lfilter with a = [a0] above order 2 (zi per row, shared and absent;
a0 != 1; T < K, where zi reaches zf), filters longer than one LDS chunk,
fftfilt (the reference's own cases, test__fftfilt.py), the "fir" format of
batched_filter_apply (broadcast and per-signal), gradients (gradcheck with
the reference's settings; the reference loop's autograd on the CPU;
adjoint identities against the oracle forward at larger sizes) and the full
4096 x 65536 shape through sampled rows; the FFT (overlap-save) kernel
through its own entry point against the oracle and the direct kernel, and
fftfilt at long K against scipy.signal.fftconvolve.  FP64 at 1e-10, FP32 at 1e-5 of
the output's largest magnitude.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _F():
    from torchscience_amd.signal_processing import filter as F
    return F


def relerr(a, ref):
    scale = max(np.max(np.abs(ref)), 1e-30) if np.size(ref) else 1.0
    return float(np.max(np.abs(np.asarray(a, dtype=np.float64) - ref)) / scale) \
        if np.size(ref) else 0.0


@pytest.mark.parametrize("T", [1, 5, 2047, 2048, 2049, 5000])
@pytest.mark.parametrize("K", [4, 9, 63])
def test_lfilter_fir_fp64(oracle_mod, K, T):
    F, dev = _F(), _dev()
    print("case lfilter fir fp64", K, T, flush=True)
    rng = np.random.default_rng(K * 7919 + T)
    b = rng.standard_normal(K)
    x = rng.standard_normal((3, T))
    for a0 in (1.0, 2.5):
        a = np.array([a0])
        bt, at = torch.tensor(b, device=dev), torch.tensor(a, device=dev)
        xt = torch.tensor(x, device=dev)
        y = F.lfilter(bt, at, xt)
        np.testing.assert_allclose(y.cpu().numpy(), oracle_mod.lfilter(b, a, x), rtol=1e-10,
                                   atol=1e-10)
        for zi in (rng.standard_normal((3, K - 1)), rng.standard_normal(K - 1)):
            y, zf = F.lfilter(bt, at, xt, zi=torch.tensor(zi, device=dev))
            yr, zfr = oracle_mod.lfilter(b, a, x, zi=np.broadcast_to(zi, (3, K - 1)))
            np.testing.assert_allclose(y.cpu().numpy(), yr, rtol=1e-10, atol=1e-10)
            np.testing.assert_allclose(zf.cpu().numpy(), zfr, rtol=1e-10, atol=1e-10)
            assert zf.shape == (3, K - 1) and zf.dtype == torch.float64


@pytest.mark.parametrize("K", [101, 1024, 1025, 2500])
def test_fir_long_taps(oracle_mod, K):
    """Filters longer than one LDS chunk (1024 taps) and the zf tail."""
    F, dev = _F(), _dev()
    rng = np.random.default_rng(K)
    b = rng.standard_normal(K) / np.sqrt(K)
    for T in (3, 3000, 10000):
        print("case long taps", K, T, flush=True)
        x = rng.standard_normal((2, T))
        zi = rng.standard_normal((2, K - 1))
        yr, zfr = oracle_mod.fir_rows(b, x, zi, want_zf=True)
        one = torch.ones(1, dtype=torch.float64, device=dev)
        y, zf = F.lfilter(torch.tensor(b, device=dev), one, torch.tensor(x, device=dev),
                          zi=torch.tensor(zi, device=dev))
        np.testing.assert_allclose(y.cpu().numpy(), yr, rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(zf.cpu().numpy(), zfr, rtol=1e-10, atol=1e-10)
        y, zf = F.lfilter(torch.tensor(b, dtype=torch.float32, device=dev), one.float(),
                          torch.tensor(x, dtype=torch.float32, device=dev),
                          zi=torch.tensor(zi, dtype=torch.float32, device=dev))
        yr32, zfr32 = oracle_mod.fir_rows(b.astype(np.float32), x.astype(np.float32),
                                          zi.astype(np.float32), want_zf=True)
        assert y.dtype == torch.float32
        assert relerr(y.cpu().numpy(), yr32) < F32_TOL
        assert relerr(zf.cpu().numpy(), zfr32) < F32_TOL


def test_fftfilt_reference_cases(oracle_mod):
    """The reference's fftfilt tests (test__fftfilt.py:12-160) at 1e-10."""
    F, dev = _F(), _dev()
    import scipy.signal as ss
    rng = np.random.default_rng(5)
    for K, T in [(51, 1000), (11, 100), (201, 500), (20, 100), (2, 7), (1, 5), (300, 40)]:
        print("case fftfilt", K, T, flush=True)
        b = ss.firwin(K, 0.3) if K > 2 else rng.standard_normal(K)
        x = rng.standard_normal(T)
        y = F.fftfilt(torch.tensor(b, device=dev), torch.tensor(x, device=dev))
        np.testing.assert_allclose(y.cpu().numpy(), oracle_mod.fftfilt(b, x), rtol=1e-10,
                                   atol=1e-10)
    b = torch.tensor(ss.firwin(21, 0.3), device=dev)
    for shape in [(10, 100), (5, 8, 100), (2, 3, 4, 100)]:
        x = torch.randn(*shape, dtype=torch.float64, device=dev)
        y = F.fftfilt(b, x)
        assert y.shape == x.shape
        np.testing.assert_allclose(y.cpu().numpy(),
                                   oracle_mod.fftfilt(b.cpu().numpy(), x.cpu().numpy()),
                                   rtol=1e-10, atol=1e-10)
    x = torch.randn(100, 50, dtype=torch.float64, device=dev)
    y0 = F.fftfilt(b, x, axis=0)
    np.testing.assert_allclose(y0.cpu().numpy(),
                               oracle_mod.fftfilt(b.cpu().numpy(), x.cpu().numpy(), axis=0),
                               rtol=1e-10, atol=1e-10)
    x = torch.randn(10, 100, dtype=torch.float64, device=dev)
    torch.testing.assert_close(F.fftfilt(b, x, axis=-1), F.fftfilt(b, x, axis=1))
    b32 = torch.randn(21, device=dev)
    x32 = torch.randn(100, device=dev)
    y32 = F.fftfilt(b32, x32)
    assert y32.dtype == torch.float32 and y32.shape == x32.shape
    assert relerr(y32.cpu().numpy(),
                  oracle_mod.fftfilt(b32.cpu().numpy(), x32.cpu().numpy())) < F32_TOL
    xc = torch.randn(3, 64, dtype=torch.complex128, device=dev)
    yc = F.fftfilt(b, xc)
    np.testing.assert_allclose(yc.cpu().numpy(),
                               oracle_mod.fftfilt(b.cpu().numpy(), xc.cpu().numpy()),
                               rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("K", [1, 3, 17, 63])
def test_batched_fir(oracle_mod, K):
    F, dev = _F(), _dev()
    print("case batched fir", K, flush=True)
    rng = np.random.default_rng(100 + K)
    filt = rng.standard_normal((3, K))
    x = rng.standard_normal((4, 1000))
    y = F.batched_filter_apply(torch.tensor(filt, device=dev), torch.tensor(x, device=dev),
                               filter_format="fir", broadcast=True)
    assert y.shape == (3, 4, 1000)
    np.testing.assert_allclose(y.cpu().numpy(),
                               oracle_mod.batched_filter_apply(filt, x, "fir", True),
                               rtol=1e-10, atol=1e-10)
    x3 = rng.standard_normal((3, 2, 1000))
    y = F.batched_filter_apply(torch.tensor(filt, device=dev), torch.tensor(x3, device=dev),
                               filter_format="fir", broadcast=False)
    assert y.shape == (3, 2, 1000)
    np.testing.assert_allclose(y.cpu().numpy(),
                               oracle_mod.batched_filter_apply(filt, x3, "fir", False),
                               rtol=1e-10, atol=1e-10)
    xa = rng.standard_normal((3, 700, 2))
    y = F.batched_filter_apply(torch.tensor(filt, device=dev), torch.tensor(xa, device=dev),
                               filter_format="fir", broadcast=False, axis=0)
    np.testing.assert_allclose(y.cpu().numpy(),
                               oracle_mod.batched_filter_apply(filt, xa, "fir", False, axis=0),
                               rtol=1e-10, atol=1e-10)
    y32 = F.batched_filter_apply(torch.tensor(filt, dtype=torch.float32, device=dev),
                                 torch.tensor(x, dtype=torch.float32, device=dev),
                                 filter_format="fir", broadcast=True)
    assert y32.dtype == torch.float32
    ref32 = oracle_mod.batched_filter_apply(filt.astype(np.float32), x.astype(np.float32),
                                            "fir", True)
    assert relerr(y32.cpu().numpy(), ref32) < F32_TOL


def test_fftfilt_gradcheck():
    """The reference's gradchecks (test__fftfilt.py:62-100)."""
    F, dev = _F(), _dev()
    b = torch.randn(21, dtype=torch.float64, device=dev, requires_grad=True)
    x = torch.randn(100, dtype=torch.float64, device=dev, requires_grad=True)
    torch.autograd.gradcheck(lambda b_: F.fftfilt(b_, x.detach()), (b,), eps=1e-6, atol=1e-4,
                             rtol=1e-4)
    torch.autograd.gradcheck(lambda x_: F.fftfilt(b.detach(), x_), (x,), eps=1e-6, atol=1e-4,
                             rtol=1e-4)
    torch.autograd.gradcheck(F.fftfilt, (b, x), eps=1e-6, atol=1e-4, rtol=1e-4)


def test_fftfilt_gradients_match_fft_restatement():
    F, dev = _F(), _dev()
    from oracle import torch_ref
    # 2500 taps: the tap-gradient kernel's second 2048-tap pass
    for K, T in [(21, 100), (64, 3000), (1100, 2500), (2500, 3000)]:
        print("case fftfilt grad", K, T, flush=True)
        b = torch.randn(K, dtype=torch.float64)
        x = torch.randn(3, T, dtype=torch.float64)
        w = torch.randn(3, T, dtype=torch.float64)
        bc, xc = b.clone().requires_grad_(), x.clone().requires_grad_()
        (torch_ref.fftfilt_fft(bc, xc) * w).sum().backward()
        bg = b.to(dev).requires_grad_()
        xg = x.to(dev).requires_grad_()
        (F.fftfilt(bg, xg) * w.to(dev)).sum().backward()
        np.testing.assert_allclose(xg.grad.cpu().numpy(), xc.grad.numpy(), rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(bg.grad.cpu().numpy(), bc.grad.numpy(), rtol=1e-9, atol=1e-9)


def test_lfilter_fir_gradients_match_reference_loop():
    """Gradients through x, b, a0 and zi (and zf) equal the autograd of the
    reference's DF2T loop restated op for op (CPU); a0 gets 0 when a0 == 1."""
    F, dev = _F(), _dev()
    from oracle import torch_ref
    K, T = 9, 150
    for a0v in (1.0, 2.0):
        for shared_zi in (False, True):
            print("case lfilter fir grad", a0v, shared_zi, flush=True)
            b = torch.randn(K, dtype=torch.float64)
            a = torch.tensor([a0v], dtype=torch.float64)
            x = torch.randn(3, T, dtype=torch.float64)
            zi = torch.randn(1 if shared_zi else 3, K - 1, dtype=torch.float64)
            w = torch.randn(3, T, dtype=torch.float64)
            v = torch.randn(3, K - 1, dtype=torch.float64)
            leaves = [t.clone().requires_grad_() for t in (b, a, x, zi)]
            y, zf = torch_ref.lfilter_loop(leaves[0], leaves[1], leaves[2],
                                           leaves[3].expand(3, K - 1))
            ((y * w).sum() + (zf * v).sum()).backward()
            gl = [t.to(dev).requires_grad_() for t in (b, a, x, zi)]
            y2, zf2 = F.lfilter(gl[0], gl[1], gl[2], zi=gl[3])
            ((y2 * w.to(dev)).sum() + (zf2 * v.to(dev)).sum()).backward()
            for name, r, g in zip("b a x zi".split(), leaves, gl):
                np.testing.assert_allclose(g.grad.cpu().numpy(), r.grad.numpy(), rtol=1e-9,
                                           atol=1e-9, err_msg=name)


def test_batched_fir_gradients():
    """Broadcast and per-signal forms: the input gradient sums over filters."""
    F, dev = _F(), _dev()
    from oracle import torch_ref
    K, T = 12, 300
    filt = torch.randn(3, K, dtype=torch.float64)
    for bc in (True, False):
        print("case batched fir grad", bc, flush=True)
        x = torch.randn(*((2, T) if bc else (3, 2, T)), dtype=torch.float64)
        w = torch.randn(3, 2, T, dtype=torch.float64)
        fc, xc = filt.clone().requires_grad_(), x.clone().requires_grad_()
        one = torch.ones(1, dtype=torch.float64)
        ys = [torch_ref.lfilter_loop(fc[i], one, xc if bc else xc[i])[0] for i in range(3)]
        (torch.stack(ys) * w).sum().backward()
        fg, xg = filt.to(dev).requires_grad_(), x.to(dev).requires_grad_()
        y = F.batched_filter_apply(fg, xg, filter_format="fir", broadcast=bc)
        (y * w.to(dev)).sum().backward()
        np.testing.assert_allclose(xg.grad.cpu().numpy(), xc.grad.numpy(), rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(fg.grad.cpu().numpy(), fc.grad.numpy(), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fir_adjoint_identities_large(oracle_mod, dtype):
    """At 64 x 20000 with 101 and 1500 taps: <dL/dh, v> = <g, FIR_v(x)> and
    <dL/dx, u> = <g, FIR_h(u)> with the forward maps from the oracle."""
    F, dev = _F(), _dev()
    rng = np.random.default_rng(9)
    B, T = 64, 20000
    for K in (101, 1500):
        print("case adjoint", dtype, K, flush=True)
        h = rng.standard_normal(K) / np.sqrt(K)
        x = rng.standard_normal((B, T))
        g = rng.standard_normal((B, T))
        hg = torch.tensor(h, dtype=dtype, device=dev, requires_grad=True)
        xg = torch.tensor(x, dtype=dtype, device=dev, requires_grad=True)
        one = torch.ones(1, dtype=dtype, device=dev)
        y = F.lfilter(hg, one, xg)
        (y * torch.tensor(g, dtype=dtype, device=dev)).sum().backward()
        v = rng.standard_normal(K)
        u = rng.standard_normal((B, T))
        lhs_h = float(np.dot(hg.grad.double().cpu().numpy(), v))
        rhs_h = float(np.sum(g * oracle_mod.fir_rows(v, x)[0]))
        lhs_x = float(np.sum(xg.grad.double().cpu().numpy() * u))
        rhs_x = float(np.sum(g * oracle_mod.fir_rows(h, u)[0]))
        tol = 1e-10 if dtype == torch.float64 else 1e-5
        scale_h = np.sqrt(B * T) * np.sqrt(np.sum(v * v)) * np.sqrt(np.mean(x * x))
        scale_x = np.sqrt(B * T) * np.linalg.norm(h) * np.sqrt(np.mean(u * u))
        assert abs(lhs_h - rhs_h) <= tol * scale_h, (lhs_h, rhs_h)
        assert abs(lhs_x - rhs_x) <= tol * scale_x, (lhs_x, rhs_x)


def test_fir_full_shape_sampled_rows(oracle_mod):
    """4096 x 65536 FP32 with 101 taps (the batched-FIR measurement shape):
    sampled rows against the oracle, and time-shift invariance of row 0."""
    F, dev = _F(), _dev()
    import scipy.signal as ss
    B, T, K = 4096, 65536, 101
    h = ss.firwin(K, 0.2).astype(np.float32)
    g = torch.Generator(device=dev).manual_seed(1)
    x = torch.randn(B, T, device=dev, generator=g)
    y = F.lfilter(torch.tensor(h, device=dev), torch.ones(1, device=dev), x)
    for c in (0, 1, 2047, 4095):
        ref, _ = oracle_mod.fir_rows(h, x[c:c + 1].cpu().numpy())
        assert relerr(y[c:c + 1].cpu().numpy(), ref) < F32_TOL, c
    xs = torch.zeros(1, T, device=dev)
    xs[0, 1000:] = x[0, :T - 1000]
    ys = F.lfilter(torch.tensor(h, device=dev), torch.ones(1, device=dev), xs)
    # time-shift invariance; 101 taps run through the FFT kernel, whose
    # rounding depends on a sample's place in its block: to the FP32 bar
    torch.testing.assert_close(ys[0, 1000:], y[0, :T - 1000], rtol=0,
                               atol=F32_TOL * y[0].abs().max().item())


def test_fir_empty_and_errors():
    F, dev = _F(), _dev()
    b = torch.randn(8, dtype=torch.float64, device=dev)
    one = torch.ones(1, dtype=torch.float64, device=dev)
    y = F.lfilter(b, one, torch.zeros(3, 0, dtype=torch.float64, device=dev))
    assert y.shape == (3, 0)
    y, zf = F.lfilter(b, one, torch.zeros(0, 10, dtype=torch.float64, device=dev),
                      zi=torch.zeros(0, 7, dtype=torch.float64, device=dev))
    assert y.shape == (0, 10) and zf.shape == (0, 7)
    with pytest.raises(ValueError, match="FIR filters must have shape"):
        F.batched_filter_apply(torch.ones(2, 3, 4, device=dev), torch.ones(5, device=dev),
                               filter_format="fir")
    with pytest.raises(ValueError, match="x batch dim must match"):
        F.batched_filter_apply(torch.ones(2, 3, device=dev), torch.ones(3, 5, device=dev),
                               filter_format="fir", broadcast=False)


def _fft_call(h, x, *, channels, shift, n_out, reverse=False, ws_bytes=None):
    """tsiir_fir_forward_fft through the C-ABI: one shared filter h over
    `channels` channels reading rows c % rows of x."""
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._fir import _fir_desc
    dev = x.device
    y = torch.full((channels, n_out), float("nan"), dtype=x.dtype, device=dev)
    d = _fir_desc(channels=channels, n_samples=n_out, x_len=x.shape[1], n_taps=h.shape[-1],
                  dtype=x.dtype, shift=shift, flags=_lib.FIR_REVERSE_TAPS if reverse else 0,
                  x_rows=x.shape[0], x_row_stride=x.shape[1], y_row_stride=n_out,
                  chans_per_filter=channels)
    L = _lib.load()
    nbytes = L.tsiir_fir_fft_workspace_bytes(d)
    assert nbytes > 0
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    st = L.tsiir_fir_forward_fft(d, _lib.ptr(h), _lib.ptr(x), _lib.ptr(y), _lib.ptr(ws),
                                 nbytes if ws_bytes is None else ws_bytes,
                                 _lib.stream_handle(dev))
    return st, y


@pytest.mark.parametrize("K", [1, 2, 63, 101, 1000, 2048])
def test_fir_fft_entry_matches_oracle(oracle_mod, K):
    """Every output block boundary, odd block counts (the last block alone
    in its complex window), fewer input rows than channels, output
    longer or shorter than the input, 'same' and reversed-tap shifts; FP64
    at 1e-10 and FP32 at 1e-5 of the largest output."""
    from torchscience_amd import _lib
    dev = _dev()
    rng = np.random.default_rng(K)
    h = rng.standard_normal(K)
    L_blk = 4096 - K + 1
    for (rows, channels, T_in, T_out, shift, rev) in (
            (5, 5, 3 * L_blk + 7, 3 * L_blk + 7, 0, False),
            (3, 7, 2 * L_blk, 2 * L_blk - 1, (K - 1) // 2, False),
            (2, 2, 1, 1, 0, False),
            (1, 3, 900, 1200, K - 1, True),
            (4, 4, L_blk + 1, L_blk + 1, K - 1 - (K - 1) // 2, True)):
        print("case fir fft", K, rows, channels, T_in, T_out, shift, rev, flush=True)
        x = rng.standard_normal((rows, T_in))
        hh = h[::-1].copy() if rev else h
        xr = x[np.arange(channels) % rows]
        xe = np.concatenate([xr, np.zeros((channels, max(T_out - T_in, 0)))], axis=1)
        ref, _ = oracle_mod.fir_rows(hh, xe, shift=shift)
        ref = ref[:, :T_out]
        for dt, tol in ((torch.float64, 1e-10), (torch.float32, F32_TOL)):
            st, y = _fft_call(torch.tensor(h, dtype=dt, device=dev),
                              torch.tensor(x, dtype=dt, device=dev), channels=channels,
                              shift=shift, n_out=T_out, reverse=rev)
            assert st == _lib.OK
            got = y.cpu().numpy()
            assert np.isfinite(got).all()
            assert relerr(got, ref) < tol, (dt, relerr(got, ref))


def test_fir_fft_matches_direct_kernel_full_shape():
    """The FFT and direct kernels agree on 512 x 65536 FP32 rows with a
    401-tap lowpass (the same sums, different rounding)."""
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._fir import fir_rows_forward
    import scipy.signal as ss
    dev = _dev()
    h = torch.tensor(ss.firwin(401, 0.2), dtype=torch.float32, device=dev)
    x = torch.randn(512, 65536, dtype=torch.float32, device=dev)
    st, y_fft = _fft_call(h, x, channels=512, shift=0, n_out=65536)
    assert st == _lib.OK
    from torchscience_amd.signal_processing.filter import _fir
    old = _fir.FFT_MIN_TAPS
    try:
        _fir.FFT_MIN_TAPS = 1 << 30
        y_dir, _ = fir_rows_forward(h.reshape(1, -1), x, None, channels=512,
                                    chans_per_filter=512)
    finally:
        _fir.FFT_MIN_TAPS = old
    scale = y_dir.abs().max().item()
    assert (y_fft - y_dir).abs().max().item() < F32_TOL * scale


def test_fir_fft_workspace_too_small():
    from torchscience_amd import _lib
    dev = _dev()
    h = torch.randn(100, dtype=torch.float64, device=dev)
    x = torch.randn(2, 500, dtype=torch.float64, device=dev)
    st, _ = _fft_call(h, x, channels=2, shift=0, n_out=500, ws_bytes=100)
    assert st == _lib.EWORKSPACE


@pytest.mark.parametrize("K", [64, 1500, 2048])
def test_fftfilt_long_filters_match_scipy(K):
    """fftfilt routes these through the FFT kernel: against
    scipy.signal.fftconvolve(x, b, mode='same') (the reference's own oracle,
    test__fftfilt.py), FP64 at 1e-10 and FP32 at 1e-5; gradient of x at
    FP64 against the adjoint (correlation with the same taps)."""
    F, dev = _F(), _dev()
    import scipy.signal as ss
    rng = np.random.default_rng(K)
    b = rng.standard_normal(K)
    x = rng.standard_normal((3, 5000))
    ref = np.stack([ss.fftconvolve(r, b, mode="same") for r in x])
    for dt, tol in ((torch.float64, 1e-10), (torch.float32, F32_TOL)):
        y = F.fftfilt(torch.tensor(b, dtype=dt, device=dev), torch.tensor(x, dtype=dt, device=dev))
        assert y.dtype == dt
        assert relerr(y.cpu().numpy(), ref) < tol
    xt = torch.tensor(x, device=dev, requires_grad=True)
    g = rng.standard_normal(x.shape)
    F.fftfilt(torch.tensor(b, device=dev), xt).backward(torch.tensor(g, device=dev))
    # y = C x with C the 'same' convolution; dx = C^T g, C^T is correlation
    full = np.stack([np.correlate(r, b, mode="full") for r in g])
    s = K - 1 - (K - 1) // 2
    gx_ref = full[:, s:s + x.shape[1]]
    assert relerr(xt.grad.cpu().numpy(), gx_ref) < 1e-10


@pytest.mark.parametrize("cpf", [1, 2, 3])
def test_fir_fft_per_filter_matches_oracle(oracle_mod, cpf):
    """One filter per group of cpf channels (odd groups leave a channel
    alone in its complex pair), through the C-ABI, FP64 at 1e-10 and FP32 at
    1e-5; and batched_filter_apply's per-signal "fir" format at 700 taps
    (routed to the FFT kernel) against the oracle."""
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._fir import _fir_desc
    dev = _dev()
    rng = np.random.default_rng(cpf)
    nf, K, T = 3, 300, 9000
    channels = nf * cpf
    h = rng.standard_normal((nf, K))
    x = rng.standard_normal((channels, T))
    ref, _ = oracle_mod.fir_rows(np.repeat(h, cpf, axis=0), x, shift=5)
    for dt, tol in ((torch.float64, 1e-10), (torch.float32, F32_TOL)):
        print("case fir fft per filter", cpf, dt, flush=True)
        ht = torch.tensor(h, dtype=dt, device=dev)
        xt = torch.tensor(x, dtype=dt, device=dev)
        y = torch.full((channels, T), float("nan"), dtype=dt, device=dev)
        d = _fir_desc(channels=channels, n_samples=T, x_len=T, n_taps=K, dtype=dt, shift=5,
                      flags=0, x_rows=channels, x_row_stride=T, y_row_stride=T,
                      chans_per_filter=cpf)
        L = _lib.load()
        nbytes = L.tsiir_fir_fft_workspace_bytes(d)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        assert L.tsiir_fir_forward_fft(d, _lib.ptr(ht), _lib.ptr(xt), _lib.ptr(y),
                                       _lib.ptr(ws), nbytes, _lib.stream_handle(dev)) == _lib.OK
        got = y.cpu().numpy()
        assert np.isfinite(got).all() and relerr(got, ref) < tol
    F = _F()
    b = rng.standard_normal((4, 700))
    xs = rng.standard_normal((4, 5000))
    yb = F.batched_filter_apply(torch.tensor(b, device=dev), torch.tensor(xs, device=dev),
                                filter_format="fir", broadcast=False)
    refb, _ = oracle_mod.fir_rows(b, xs)     # lfilter(b[i], [1], x[i]) per row
    assert yb.shape == refb.shape and relerr(yb.cpu().numpy(), refb) < 1e-10


def test_fir_fft_channels_stay_apart():
    """The FFT kernel transforms two blocks of ONE channel per complex
    window, so channels never share rounding or non-finite values: rows at
    scales 1e6, 1, 1e-3 each match the direct kernel per channel at the
    FP32 tolerance of their own scale, and a NaN in one channel leaves every
    other channel finite and reaches only the outputs of its own channel's
    window (two blocks)."""
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter import _fir
    from torchscience_amd.signal_processing.filter._fir import fir_rows_forward
    import scipy.signal as ss
    dev = _dev()
    K = 101
    Lb = 4096 - K + 1
    h = torch.tensor(ss.firwin(K, 0.2), dtype=torch.float32, device=dev)
    x = torch.randn(4, 8 * Lb, dtype=torch.float32, device=dev)
    x *= torch.tensor([1e6, 1.0, 1e-3, 1.0], device=dev)[:, None]
    st, y = _fft_call(h, x, channels=4, shift=0, n_out=x.shape[1])
    assert st == _lib.OK
    old = _fir.FFT_MIN_TAPS
    try:
        _fir.FFT_MIN_TAPS = 1 << 30
        y_dir, _ = fir_rows_forward(h.reshape(1, -1), x, None, channels=4, chans_per_filter=4)
    finally:
        _fir.FFT_MIN_TAPS = old
    for c in range(4):
        scale = y_dir[c].abs().max().item()
        assert (y[c] - y_dir[c]).abs().max().item() < F32_TOL * scale, c
    bad = 3 * Lb + 100                       # in block 3 (window of blocks 2 and 3)
    x[0, bad] = float("nan")
    st, y = _fft_call(h, x, channels=4, shift=0, n_out=x.shape[1])
    assert st == _lib.OK
    assert torch.isfinite(y[1:]).all()
    nan_at = torch.isnan(y[0]).nonzero().flatten()
    assert nan_at.numel() > 0
    assert nan_at.min().item() >= 2 * Lb and nan_at.max().item() < 4 * Lb
    assert torch.isfinite(y[0, :2 * Lb]).all() and torch.isfinite(y[0, 4 * Lb:]).all()
