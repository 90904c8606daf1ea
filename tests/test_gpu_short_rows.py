"""Short rows (GPU): rows of at most 512 samples and 4096 samples x sections,
with at most 8 sections, run in the thread-per-channel kernel (every
section of a sample before the next, DF2T, the state in registers).

Against the FP64 oracle on every channel: FP64 I/O 1e-12, FP32 1e-5, the
compensated form (FP64 updates) 1e-6, BF16 I/O 1e-2 (per-channel normwise
relative errors); zf to the same bounds relative to the output scale.  Also:
zi per channel and broadcast, reverse time against the oracle on flipped
rows, input rows shared by several channels, per-channel filters with a0 !=
1, one channel and one sample, and agreement with the window form
(path="window") within the same tolerance.
"""

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


def _filters(n, S, rng):
    """n stable filters of S sections with a0 = 2 (the kernel normalises)."""
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    sos = batched_butterworth_sos(2 * S, rng.uniform(0.05, 0.4, n))
    sos = sos * 2.0
    return sos


TOL = {torch.float64: 1e-12, torch.float32: 1e-5, torch.bfloat16: 1e-2}


@pytest.mark.parametrize("io", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("B,T,S", [(4096, 39, 6), (3, 1, 1), (1, 512, 8), (257, 7, 8),
                                   (64, 200, 2)])
@pytest.mark.parametrize("reverse", [False, True])
def test_short_rows_against_oracle(oracle_mod, io, B, T, S, reverse):
    dev = _dev()
    print(f"short rows io={io} B={B} T={T} S={S} reverse={reverse}", flush=True)
    rng = np.random.default_rng(B * 31 + T + S)
    sos = _filters(B, S, rng)
    cdt = torch.float64 if io == torch.float64 else torch.float32
    sos_k = sos.astype(np.float32).astype(np.float64) if cdt == torch.float32 else sos
    x = torch.tensor(rng.standard_normal((B, T)), dtype=torch.float32 if io != torch.float64
                     else torch.float64, device=dev).to(io)
    zi = torch.tensor(rng.standard_normal((B, S, 2)), dtype=cdt, device=dev)
    y, zf = _fwd(torch.tensor(sos_k, dtype=cdt, device=dev), x, zi, channels=B,
                 chans_per_filter=1, want_zf=True, reverse=reverse)
    x64 = x.double().cpu().numpy()
    if reverse:
        x64 = x64[:, ::-1]
    ref, zf_ref = oracle_mod.sosfilt_rows(sos_k, np.ascontiguousarray(x64),
                                          zi.double().cpu().numpy(), chans_per_filter=1,
                                          n_threads=8)
    got = y.double().cpu().numpy()
    if reverse:
        got = got[:, ::-1]
    err = relerr(got, ref)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < TOL[io]
    scale = max(np.max(np.abs(ref)), 1.0)
    assert np.max(np.abs(zf.double().cpu().numpy() - zf_ref)) < TOL[io] * scale


@pytest.mark.parametrize("io", [torch.float32, torch.float64])
def test_short_rows_match_window_form(io):
    """The same call through the window form (path="window"): within the
    tolerance, zi broadcast, input rows shared by 4 channels each."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B, T = 1024, 300
    rng = np.random.default_rng(7)
    sos = torch.tensor(butterworth_sos(6, [0.05, 0.3], "bandpass")[None], dtype=io, device=dev)
    x = torch.tensor(rng.standard_normal((B // 4, T)), dtype=io, device=dev)
    zi = torch.tensor(rng.standard_normal((1, 6, 2)), dtype=io, device=dev)
    kw = dict(channels=B, chans_per_filter=B, want_zf=True, zi_broadcast=True)
    y, zf = _fwd(sos, x, zi, **kw)
    yw, zfw = _fwd(sos, x, zi, path="window", **kw)
    assert relerr(y.double().cpu().numpy(), yw.double().cpu().numpy()) < TOL[io]
    scale = float(yw.abs().max())
    assert float((zf - zfw).abs().max()) < TOL[io] * scale


def test_short_rows_compensated(oracle_mod):
    """FP32 I/O with FP64 recurrence updates (mode 1) on a filter with poles
    near the unit circle."""
    from torchscience_amd.signal_processing.filter._design import chebyshev1_sos
    dev = _dev()
    B, T = 512, 500
    rng = np.random.default_rng(9)
    sos = chebyshev1_sos(16, 0.01, 1.0)
    x = torch.tensor(rng.standard_normal((B, T)), dtype=torch.float32, device=dev)
    y, _ = _fwd(torch.tensor(sos[None], dtype=torch.float64, device=dev), x, None, channels=B,
                chans_per_filter=B, want_zf=False, mode=1)
    ref, _ = oracle_mod.sosfilt_rows(sos, x.double().cpu().numpy(), n_threads=8)
    assert relerr(y.double().cpu().numpy(), ref) < 1e-6
