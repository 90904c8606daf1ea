"""The segmented form of few long rows (GPU): fewer than 1024 channels of
rows long enough for 1024 segments of 2^k * 16384 samples.  When every
filter forgets its state within 4096 samples (checked on the device: the
infinity norm of the cascade's Phi^4096 below 2^-64), each segment runs as an
independent row from the state a 4096-sample warm-up over the end of the
segment before it reaches; otherwise the look-back form runs.

Checks, against the FP64 oracle on every channel: FP32 (1e-5 normwise), FP64
(1e-10), the compensated form (1e-5), BF16 I/O (1e-2); zi per channel and
broadcast, zf; shared and per-channel filters; a channel count that is not a
power of two; one channel; the same within tolerance as the look-back form
(path="lookback": the segmented form off).  The two forms usually agree
bitwise: a window's incoming state differs between them by less than 2^-64
of it, which rounds away in FP32.  So which form ran is read from the
workspace header of a C-ABI call with a caller-owned workspace: its first
word is the device check's verdict (1: the filter forgets its state, the
segmented form ran) and its third the launch epoch, which every table
preparation that runs advances by one.  FP32 recurrences prepare the tables
once for the segmented main pass (which runs each segment's warm-up itself)
and the look-back form alike: 1, and 2 with a remainder pass long enough to
prepare its own (a short remainder runs the short-row kernel, which
prepares none).  A lowpass at cutoff 0.002 (||Phi^4096|| ~ 1e-7) must
fail the check and keep the look-back.
"""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def relerr(y, y64):
    y = np.asarray(y, dtype=np.float64)
    num = np.max(np.abs(y - y64), axis=-1)
    den = np.maximum(np.max(np.abs(y64), axis=-1), 1e-300)
    return float(np.max(num / den))


def _fwd(sos, x, zi=None, **kw):
    from torchscience_amd.signal_processing.filter._ops import sosfilt_rows_forward
    return sosfilt_rows_forward(sos, x, zi, **kw)


def _abi_header(sos, x, flags=0, ws=None):
    """One C-ABI forward over contiguous FP32 rows x (B, T) with a shared
    filter and a fresh zeroed workspace (or `ws`); returns (y, header
    words)."""
    from torchscience_amd import _lib
    L = _lib.load()
    B, T = x.shape
    d = _lib.SosfiltDesc(channels=B, n_samples=T, n_sections=sos.shape[-2], io_dtype=_lib.F32,
                         coef_dtype=_lib.F32, mode=0, flags=flags, reserved=0, x_rows=B,
                         x_row_stride=T, y_row_stride=T, chans_per_filter=B)
    n = L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d))
    if ws is None:
        ws = torch.zeros(n, dtype=torch.uint8, device=x.device)
    assert ws.numel() >= n
    y = torch.empty_like(x)
    _lib.check(L.tsiir_sosfilt_forward(ctypes.byref(d), _lib.ptr(x), _lib.ptr(y), _lib.ptr(sos),
                                       None, None, _lib.ptr(ws), ws.numel(),
                                       _lib.stream_handle(x.device)), "tsiir_sosfilt_forward")
    torch.cuda.synchronize(x.device)
    return y, ws[:16].cpu().numpy().view(np.uint32).tolist()


def test_device_check_picks_the_form(oracle_mod):
    from torchscience_amd import _lib
    dev = _dev()
    B, T = 8, 1 << 21
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    x = torch.randn(B, T, generator=g, device=dev)
    for name, sos, forgets in (("bandpass", _bandpass(), True),
                               ("lowpass 0.002", _lowpass(0.002), False)):
        print("device check:", name, flush=True)
        st = torch.tensor(sos.astype(np.float32)[None], device=dev)
        y, hdr = _abi_header(st, x)
        assert hdr[0] == (1 if forgets else 0) and hdr[2] == 1, (name, hdr)
        y_lb, hdr_lb = _abi_header(st, x, _lib.NO_SEGMENT)
        assert hdr_lb[2] == 1, hdr_lb
        # with a remainder after the segments (128 segments of 16384 and 78
        # samples): its pass runs the short-row kernel, which prepares no
        # tables; one of 12345 samples prepares them a second time
        _, hdr_r = _abi_header(st, torch.cat([x, x[:, :78]], dim=1))
        assert hdr_r[0] == (1 if forgets else 0) and hdr_r[2] == 1, hdr_r
        _, hdr_r = _abi_header(st, torch.cat([x, x[:, :12345]], dim=1))
        assert hdr_r[0] == (1 if forgets else 0) and hdr_r[2] == (2 if forgets else 1), hdr_r
        if not forgets:
            # the look-back form's output (its accuracy on this filter, whose
            # FP32 recurrence the drop-in routes to DF2T, is tested elsewhere)
            assert torch.equal(y, y_lb)
            continue
        ref, _ = oracle_mod.sosfilt_rows(sos.astype(np.float32).astype(np.float64),
                                         x[:2].double().cpu().numpy(), n_threads=16)
        assert relerr(y[:2].double().cpu().numpy(), ref) < 1e-5


def test_reused_workspace_judges_each_filter():
    """The header caches the verdict of the last shared filter judged (bit
    for bit); a workspace reused with a filter that does not forget its
    state, then with the first filter again, gets each verdict right: the
    look-back form's output for the former, the segmented form's for the
    latter."""
    from torchscience_amd import _lib
    dev = _dev()
    B, T = 8, 1 << 21
    g = torch.Generator(device=dev)
    g.manual_seed(12)
    x = torch.randn(B, T, generator=g, device=dev)
    band = torch.tensor(_bandpass().astype(np.float32)[None], device=dev)
    low = torch.tensor(_lowpass(0.002).astype(np.float32)[None], device=dev)
    y_fresh, _ = _abi_header(band, x)
    low_lb, _ = _abi_header(low, x, _lib.NO_SEGMENT)
    ws = torch.zeros(64 << 20, dtype=torch.uint8, device=dev)
    for sos, want_gate, want in ((band, 1, y_fresh), (band, 1, y_fresh), (low, 0, low_lb),
                                 (band, 1, y_fresh), (low, 0, low_lb)):
        y, hdr = _abi_header(sos, x, 0, ws)
        assert hdr[0] == want_gate, hdr
        assert torch.equal(y, want)


def _bandpass():
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    return butterworth_sos(6, [0.05, 0.3], "bandpass")


def _lowpass(cut, order=8):
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    return butterworth_sos(order, cut)


# (channels, samples, io dtype, mode, zi: None | "row" | "bcast", per-channel filters)
CASES = [
    (8, 1 << 21, torch.float32, 0, None, False),
    (8, 1 << 21, torch.float32, 0, "row", False),
    (3, 1 << 23, torch.float32, 0, "bcast", False),
    (1, 1 << 24, torch.float32, 0, "row", False),
    (8, 1 << 21, torch.float32, 0, "row", True),
    (8, 1 << 21, torch.float64, 0, "row", False),
    (8, 1 << 21, torch.float32, 1, "row", False),
    (8, 1 << 21, torch.bfloat16, 0, None, False),
    # rows with a remainder after the segments (78 and 12305 samples)
    (8, (1 << 21) + 78, torch.float32, 0, "row", False),
    (5, (1 << 22) + 12345, torch.float32, 0, "bcast", True),
]


@pytest.mark.parametrize("B,T,io,mode,zi_kind,per_channel", CASES)
def test_segmented_against_oracle(oracle_mod, B, T, io, mode, zi_kind, per_channel):
    dev = _dev()
    print(f"segmented: B={B} T={T} io={io} mode={mode} zi={zi_kind} "
          f"per_channel={per_channel}", flush=True)
    rng = np.random.default_rng(B * 7 + mode)
    if per_channel:
        sos = np.stack([_lowpass(float(c), 6) for c in rng.uniform(0.05, 0.4, B)])
        cpf = 1
    else:
        sos = _bandpass()[None]
        cpf = B
    cdt = torch.float64 if (io == torch.float64 or mode == 1) else torch.float32
    sos_k = sos if cdt == torch.float64 else sos.astype(np.float32).astype(np.float64)
    st = torch.tensor(sos_k, dtype=cdt, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(B + T)
    x = torch.randn(B, T, generator=g, device=dev, dtype=torch.float32).to(io)
    sdt = torch.float64 if io == torch.float64 else torch.float32
    zi = zi_np = None
    if zi_kind == "row":
        zi_np = rng.standard_normal((B, sos.shape[1], 2))
        zi = torch.tensor(zi_np, dtype=sdt, device=dev)
        zi_np = zi.double().cpu().numpy()
    elif zi_kind == "bcast":
        z1 = rng.standard_normal((1, sos.shape[1], 2))
        zi = torch.tensor(z1, dtype=sdt, device=dev)
        zi_np = np.broadcast_to(zi.double().cpu().numpy(), (B, sos.shape[1], 2))
    kw = dict(channels=B, chans_per_filter=cpf, want_zf=zi is not None,
              zi_broadcast=zi_kind == "bcast", mode=mode)
    y, zf = _fwd(st, x, zi, **kw)
    y_lb, zf_lb = _fwd(st, x, zi, path="lookback", **kw)
    from torchscience_amd import check_errors
    check_errors(dev)
    ref, zf_ref = oracle_mod.sosfilt_rows(sos_k, x.double().cpu().numpy(), zi_np,
                                          chans_per_filter=cpf, n_threads=16)
    tol = {torch.float64: 1e-10, torch.bfloat16: 1e-2}.get(io, 1e-5)
    err = relerr(y.double().cpu().numpy(), ref)
    err_lb = relerr(y_lb.double().cpu().numpy(), ref)
    print(f"  max rel err segmented {err:.3e}, look-back {err_lb:.3e}", flush=True)
    assert err < tol and err_lb < tol
    if zi is not None:
        scale = np.max(np.abs(ref))
        assert zf.shape == (B, sos.shape[1], 2) and zf.dtype == sdt
        assert np.max(np.abs(zf.double().cpu().numpy() - zf_ref)) < tol * scale


@pytest.mark.parametrize("B,T,io,zi_kind", [(8, 1 << 21, torch.float32, "row"),
                                             (3, 1 << 23, torch.float32, "bcast"),
                                             (8, 1 << 21, torch.float64, "row"),
                                             (8, (1 << 21) + 78, torch.float32, "row"),
                                             (5, (1 << 22) + 12345, torch.float64, "bcast")])
def test_segmented_reverse_time(oracle_mod, B, T, io, zi_kind):
    """Along decreasing time (the second pass of the zero-phase filters): each
    segment's warm-up runs over the start of the segment after it, the last
    segment in time starts from zi (the state at the end) and zf is the
    state after the first segment; equal to the oracle on the flipped rows,
    and to the look-back form within tolerance."""
    dev = _dev()
    print(f"segmented reverse: B={B} T={T} io={io} zi={zi_kind}", flush=True)
    rng = np.random.default_rng(B + 99)
    sos = _bandpass()
    cdt = torch.float64 if io == torch.float64 else torch.float32
    sos_k = sos if cdt == torch.float64 else sos.astype(np.float32).astype(np.float64)
    st = torch.tensor(sos_k[None], dtype=cdt, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(T + 1)
    x = torch.randn(B, T, generator=g, device=dev, dtype=torch.float32).to(io)
    S = sos.shape[0]
    if zi_kind == "row":
        zi = torch.tensor(rng.standard_normal((B, S, 2)), dtype=cdt, device=dev)
        zi_np = zi.double().cpu().numpy()
    else:
        zi = torch.tensor(rng.standard_normal((1, S, 2)), dtype=cdt, device=dev)
        zi_np = np.broadcast_to(zi.double().cpu().numpy(), (B, S, 2))
    kw = dict(channels=B, chans_per_filter=B, want_zf=True, zi_broadcast=zi_kind == "bcast",
              reverse=True)
    y, zf = _fwd(st, x, zi, **kw)
    y_lb, zf_lb = _fwd(st, x, zi, path="lookback", **kw)
    ref, zf_ref = oracle_mod.sosfilt_rows(sos_k, x.double().cpu().numpy()[:, ::-1], zi_np,
                                          n_threads=16)
    tol = 1e-10 if io == torch.float64 else 1e-5
    err = relerr(y.double().cpu().numpy()[:, ::-1], ref)
    err_lb = relerr(y_lb.double().cpu().numpy()[:, ::-1], ref)
    print(f"  max rel err segmented {err:.3e}, look-back {err_lb:.3e}", flush=True)
    assert err < tol and err_lb < tol
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(zf.double().cpu().numpy() - zf_ref)) < tol * scale


def test_sosfiltfilt_few_long_rows(oracle_mod):
    """sosfiltfilt on 8 x 2^21: both passes run segmented over the padded
    rows (2^21 + 2 x 39 samples: 128 segments and a 78-sample remainder),
    against the oracle's zero-phase filter."""
    from torchscience_amd.signal_processing.filter import sosfiltfilt
    dev = _dev()
    print("sosfiltfilt 8 x 2^21", flush=True)
    sos = _bandpass().astype(np.float32)
    g = torch.Generator(device=dev)
    g.manual_seed(21)
    x = torch.randn(8, 1 << 21, generator=g, device=dev)
    y = sosfiltfilt(torch.tensor(sos, device=dev), x)
    ref = oracle_mod.sosfiltfilt(sos.astype(np.float64), x.double().cpu().numpy())
    err = relerr(y.double().cpu().numpy(), ref)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < 1e-5


def test_public_sosfilt_takes_the_segmented_form(oracle_mod):
    """The drop-in call on the long-sequence shape (8 x 2^24, the benchmark's
    bandpass): all 8 channels against the oracle."""
    from torchscience_amd.signal_processing.filter import sosfilt
    dev = _dev()
    print("sosfilt 8 x 2^24 bandpass", flush=True)
    sos = _bandpass().astype(np.float32)
    st = torch.tensor(sos, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    x = torch.randn(8, 1 << 24, generator=g, device=dev)
    y = sosfilt(st, x)
    ref, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.double().cpu().numpy(),
                                     n_threads=16)
    err = relerr(y.cpu().numpy(), ref)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < 1e-5
