"""FP32-recurrence forward kernel (many channels, chained 4096-sample windows)
against the FP64 oracle (GPU).

FP32 and BF16 I/O with up to 64 sections select this kernel: with at least
1024 channels the 16384-sample groups of a channel are chained, with fewer
the 32768-sample groups (and the leftover windows) look back over the
groups before them.  Covers:
filters whose b0 can be normalised (first input tap 1, running b0 product
restored on output) and filters that keep b0 as a tap (a zero b0, or an
overall gain far from 1), mixed in one per-channel batch; lengths around the
window (4096 samples) boundary; zi per channel and broadcast; zf; input rows
shared by several channels; rows that are not 16-byte aligned; time reversal;
the saved section outputs of the training forward; few channels with many
predecessor groups (several look-back rounds of 4 groups each).

Tolerances are per-channel normwise relative errors max|y - y64| / max|y64|
against the FP64 oracle run on the same (rounded) coefficients and inputs:
FP32 1e-5, BF16 I/O 1e-2 (bfloat16 output rounding).
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


def relerr(y, y64):
    y = np.asarray(y, dtype=np.float64)
    num = np.max(np.abs(y - y64), axis=-1)
    den = np.maximum(np.max(np.abs(y64), axis=-1), 1e-300)
    return float(np.max(num / den))


def _fwd(sos, x, zi=None, **kw):
    from torchscience_amd.signal_processing.filter._ops import sosfilt_rows_forward
    return sosfilt_rows_forward(sos, x, zi, **kw)


def _butter(n, cut):
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    return batched_butterworth_sos(n, np.asarray(cut, dtype=np.float64))


def _nonunit_filters(n, S, rng):
    """Filters the kernel must run with b0 kept as a tap: section 0 has b0 = 0
    (a one-sample delay in front of a stable biquad), or the gain product is
    far below 2^-50."""
    out = []
    for i in range(n):
        sos = _butter(2 * S, [rng.uniform(0.05, 0.4)])[0].copy()
        if i % 2 == 0:
            sos[0, :3] = [0.0, 1.0, 0.5]
        else:
            sos[:, :3] *= 1e-5  # product of b0 ~ 1e-5^S * gain
        out.append(sos)
    return np.stack(out)


@pytest.mark.parametrize("T", [4096, 4097, 3 * 4096 + 5, 40000])
def test_shared_filter_with_zi_and_zf(oracle_mod, T):
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B = 1030  # not a multiple of 8 (windows of one channel on different dies)
    rng = np.random.default_rng(T)
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    x = rng.standard_normal((B, T)).astype(np.float32)
    zi = rng.standard_normal((B, 6, 2)).astype(np.float32)
    print("shared filter T", T, flush=True)
    y, zf = _fwd(torch.tensor(sos[None], device=dev), torch.tensor(x, device=dev),
                 torch.tensor(zi, device=dev), channels=B, chans_per_filter=B, want_zf=True)
    y64, zf64 = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.astype(np.float64),
                                        zi.astype(np.float64), n_threads=8)
    err = relerr(y.cpu().numpy(), y64)
    print("relerr", err)
    assert err < F32_TOL
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-4, atol=1e-4)


def test_zi_broadcast_and_shared_input_rows(oracle_mod):
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B, T, rows = 1024, 2 * 4096, 256
    rng = np.random.default_rng(1)
    sos = butterworth_sos(8, 0.2).astype(np.float32)
    x = rng.standard_normal((rows, T)).astype(np.float32)
    zi = rng.standard_normal((1, 4, 2)).astype(np.float32)
    y, zf = _fwd(torch.tensor(sos[None], device=dev), torch.tensor(x, device=dev),
                 torch.tensor(zi, device=dev), channels=B, chans_per_filter=B, want_zf=True,
                 zi_broadcast=True)
    xb = x[np.arange(B) % rows].astype(np.float64)
    y64, zf64 = oracle_mod.sosfilt_rows(sos.astype(np.float64), xb,
                                        np.repeat(zi.astype(np.float64), B, 0), n_threads=8)
    assert relerr(y.cpu().numpy(), y64) < F32_TOL
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-4, atol=1e-4)


def test_mixed_unit_and_b0_tap_filters(oracle_mod):
    """Per-channel filters: Butterworth lowpass designs (b0 normalised) and
    filters that keep b0 as a tap, interleaved across channels."""
    dev = _dev()
    B, T, S = 1024, 3 * 4096 + 100, 4
    rng = np.random.default_rng(2)
    sos = _butter(2 * S, rng.uniform(0.02, 0.45, B))
    bad = _nonunit_filters(B // 4, S, rng)
    sos[::4] = bad
    sos = sos.astype(np.float32)
    x = rng.standard_normal((B, T)).astype(np.float32)
    y, zf = _fwd(torch.tensor(sos, device=dev), torch.tensor(x, device=dev), None,
                 channels=B, chans_per_filter=1, want_zf=True)
    y64, zf64 = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.astype(np.float64),
                                        chans_per_filter=1, n_threads=8)
    err = relerr(y.cpu().numpy(), y64)
    print("mixed relerr", err)
    assert err < F32_TOL
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-4, atol=1e-4 * np.abs(zf64).max())


def test_unaligned_rows(oracle_mod):
    """Rows of a (B, T + 1) buffer read from column 1: not 16-byte aligned."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B, T = 1024, 2 * 4096 + 7
    rng = np.random.default_rng(3)
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    buf = rng.standard_normal((B, T + 1)).astype(np.float32)
    xt = torch.tensor(buf, device=dev)[:, 1:]
    assert xt.stride(0) == T + 1
    y, _ = _fwd(torch.tensor(sos[None], device=dev), xt, None, channels=B, chans_per_filter=B,
                want_zf=False)
    y64, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), buf[:, 1:].astype(np.float64),
                                     n_threads=8)
    assert relerr(y.cpu().numpy(), y64) < F32_TOL


@pytest.mark.parametrize("T", [4096, 2 * 4096 + 3])
def test_reverse(oracle_mod, T):
    """Filtering along decreasing time equals filtering the flipped rows."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B = 1024
    rng = np.random.default_rng(4)
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    x = rng.standard_normal((B, T)).astype(np.float32)
    zi = rng.standard_normal((B, 6, 2)).astype(np.float32)
    y, zf = _fwd(torch.tensor(sos[None], device=dev), torch.tensor(x, device=dev),
                 torch.tensor(zi, device=dev), channels=B, chans_per_filter=B, want_zf=True,
                 reverse=True)
    y64, zf64 = oracle_mod.sosfilt_rows(sos.astype(np.float64), x[:, ::-1].astype(np.float64),
                                        zi.astype(np.float64), n_threads=8)
    assert relerr(y.cpu().numpy()[:, ::-1], y64) < F32_TOL
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("T", [2 * 4096 + 11, 2 * 4096 + 12])
def test_saved_section_outputs(oracle_mod, T):
    """The training forward keeps the output of every section but the last,
    in true units (the b0 normalisation undone).  T = 8203: the saved rows
    are not 16-byte aligned (element stores); T = 8204: they are."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B = 1024
    print(f"saved sections T={T}", flush=True)
    rng = np.random.default_rng(5)
    sos = butterworth_sos(8, 0.2).astype(np.float32)
    x = rng.standard_normal((B, T)).astype(np.float32)
    y, _, saved = _fwd(torch.tensor(sos[None], device=dev), torch.tensor(x, device=dev), None,
                       channels=B, chans_per_filter=B, want_zf=False, keep_sections=True)
    assert saved.shape == (3, B, T)
    x64 = x.astype(np.float64)
    for s in range(3):
        ys, _ = oracle_mod.sosfilt_rows(sos[: s + 1].astype(np.float64), x64, n_threads=8)
        assert relerr(saved[s].cpu().numpy(), ys) < F32_TOL, s
    y64, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), x64, n_threads=8)
    assert relerr(y.cpu().numpy(), y64) < F32_TOL


def test_bf16_io(oracle_mod):
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B, T = 1024, 4096 + 64
    rng = np.random.default_rng(6)
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    x = torch.tensor(rng.standard_normal((B, T)), dtype=torch.bfloat16)
    y, _ = _fwd(torch.tensor(sos[None], device=dev), x.to(dev), None, channels=B,
                chans_per_filter=B, want_zf=False)
    assert y.dtype == torch.bfloat16
    y64, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.double().numpy(), n_threads=8)
    assert relerr(y.float().cpu().numpy(), y64) < 1e-2


def test_error_next_to_channel_persistent_kernel(oracle_mod):
    """Error against the FP64 oracle of this kernel (direct form I, FP32 lane
    scan) next to that of the channel-persistent kernel (DF2T recurrence,
    FP64 lane scan) on the same call: both within the FP32 tolerance, and
    the first no worse than 2x the second."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    g = torch.Generator().manual_seed(7)
    x = torch.randn(1024, 3 * 4096 + 9, generator=g)
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    st = torch.tensor(sos[None]).to(dev)
    y, _ = _fwd(st, x.to(dev), None, channels=1024, chans_per_filter=1024, want_zf=False)
    yc, _ = _fwd(st, x.to(dev), None, channels=1024, chans_per_filter=1024, want_zf=False,
                 path="channel")
    y64, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.double().numpy(), n_threads=8)
    e_new, e_chan = relerr(y.cpu().numpy(), y64), relerr(yc.cpu().numpy(), y64)
    print(f"relerr fp32 window kernel {e_new:.3e}, channel kernel {e_chan:.3e}")
    assert e_new < F32_TOL and e_chan < F32_TOL
    assert e_new <= 2 * e_chan


@pytest.mark.parametrize("reverse", [False, True])
def test_few_channels_lookback(oracle_mod, reverse):
    """5 channels of 282 full windows (35 groups of 8 windows, 2 leftover
    windows with their own look-back) and a partial one; a look-back round
    inspects 4 predecessors, so late groups can need several rounds.
    Per-channel filters, two of them keeping b0 as a tap; zi and zf."""
    dev = _dev()
    B, S = 5, 4
    T = 70 * 16384 + 2 * 4096 + 37
    rng = np.random.default_rng(8 + reverse)
    sos = _butter(2 * S, rng.uniform(0.02, 0.45, B))
    sos[1:3] = _nonunit_filters(2, S, rng)
    sos = sos.astype(np.float32)
    x = rng.standard_normal((B, T)).astype(np.float32)
    zi = rng.standard_normal((B, S, 2)).astype(np.float32)
    print("few channels lookback, reverse", reverse, flush=True)
    y, zf = _fwd(torch.tensor(sos, device=dev), torch.tensor(x, device=dev),
                 torch.tensor(zi, device=dev), channels=B, chans_per_filter=1, want_zf=True,
                 reverse=reverse)
    xo = x[:, ::-1] if reverse else x
    y64, zf64 = oracle_mod.sosfilt_rows(sos.astype(np.float64), xo.astype(np.float64),
                                        zi.astype(np.float64), chans_per_filter=1, n_threads=8)
    yh = y.cpu().numpy()
    err = relerr(yh[:, ::-1] if reverse else yh, y64)
    print("lookback relerr", err)
    assert err < F32_TOL
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-4, atol=1e-4 * np.abs(zf64).max())


@pytest.mark.parametrize("B", [1030, 6])
@pytest.mark.parametrize("S", [1, 17, 64])
def test_many_sections(oracle_mod, S, B):
    """Shared filter of S biquads (cutoffs U[0.3, 0.45]) on both forms
    (chained groups at 1030 channels, look-back groups at 6), over full
    groups, leftover windows and a partial window; zi and zf."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    rng = np.random.default_rng(100 * S + B)
    sos = np.concatenate([butterworth_sos(2, float(c)) for c in rng.uniform(0.3, 0.45, S)])
    sos = sos.astype(np.float32)
    T = 2 * 32768 + 3 * 4096 + 100 if B < 1024 else 16384 + 3 * 4096 + 100
    x = rng.standard_normal((B, T)).astype(np.float32)
    zi = rng.standard_normal((B, S, 2)).astype(np.float32)
    print("many sections S", S, "B", B, flush=True)
    y, zf = _fwd(torch.tensor(sos[None], device=dev), torch.tensor(x, device=dev),
                 torch.tensor(zi, device=dev), channels=B, chans_per_filter=B, want_zf=True)
    y64, zf64 = oracle_mod.sosfilt_rows(sos.astype(np.float64), x.astype(np.float64),
                                        zi.astype(np.float64), n_threads=8)
    err = relerr(y.cpu().numpy(), y64)
    print("relerr", err)
    assert err < F32_TOL
    np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-3,
                               atol=1e-4 * max(1.0, np.abs(zf64).max()))
