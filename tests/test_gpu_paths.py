"""Both forward work decompositions (window-parallel with look-back, and
channel-persistent) on the same inputs, against the FP64 oracle (GPU).

Covers every I/O dtype and arithmetic mode, zi per channel and broadcast, zf,
per-channel filters, input rows shared by several channels, 1..64 sections,
and lengths around the lane (32 / 64 samples) and window (2048 / 4096
samples) boundaries.  Tolerances (per-channel normwise relative error
max|y - y64| / max|y64|): FP64 1e-10; FP32 and the compensated form 1e-5;
BF16 I/O 1e-2 (bfloat16 output rounding).
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


FORMS = {  # name: (io dtype, coefficient dtype, mode, tolerance)
    "fp64": (torch.float64, torch.float64, 0, 1e-10),
    "fp32": (torch.float32, torch.float32, 0, 1e-5),
    "fp32c": (torch.float32, torch.float64, 1, 1e-5),
    "bf16io": (torch.bfloat16, torch.float32, 0, 1e-2),
}


def _run(sos, x, zi, path, form, cpf, zi_b):
    from torchscience_amd.signal_processing.filter._ops import sosfilt_rows_forward
    io, cdt, mode, _ = FORMS[form]
    return sosfilt_rows_forward(sos.to(cdt), x.to(io), zi, channels=cpf * sos.shape[0],
                                chans_per_filter=cpf, want_zf=True, zi_broadcast=zi_b,
                                io_dtype=io, mode=mode, path=path)


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("T", [1, 31, 33, 2047, 2049, 4095, 4097, 3 * 4096 + 5, 40000])
def test_paths_agree_with_oracle(oracle_mod, form, T):
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    dev = _dev()
    rng = np.random.default_rng(T + 7 * len(form))
    n_filters, cpf, S = 3, 2, 4
    B = n_filters * cpf
    sos = batched_butterworth_sos(2 * S, rng.uniform(0.05, 0.45, n_filters))
    io, cdt, _, tol = FORMS[form]
    sos_used = torch.tensor(sos).to(cdt).double().numpy()
    x = rng.standard_normal((B, T))
    x_used = torch.tensor(x).to(io).double().numpy()
    acc = torch.float64 if io == torch.float64 else torch.float32
    zi = rng.standard_normal((B, S, 2))
    zi_used = torch.tensor(zi).to(acc).double().numpy()
    y64, zf64 = oracle_mod.sosfilt_rows(sos_used, x_used, zi_used, chans_per_filter=cpf)
    outs = {}
    for path in ("window", "channel"):
        print("form", form, "T", T, "path", path, flush=True)
        y, zf = _run(torch.tensor(sos, device=dev), torch.tensor(x, device=dev),
                     torch.tensor(zi, device=dev).to(acc), path, form, cpf, False)
        outs[path] = y.double().cpu().numpy()
        assert relerr(outs[path], y64) < tol, path
        zf_tol = 1e-9 if form == "fp64" else 1e-3
        np.testing.assert_allclose(zf.double().cpu().numpy(), zf64, rtol=zf_tol, atol=zf_tol)
    assert relerr(outs["window"], outs["channel"]) < 2 * tol


@pytest.mark.parametrize("S", [1, 7, 33, 64])
def test_channel_path_sections_and_broadcast(oracle_mod, S):
    """Shared filter of S sections, one input row read by every channel
    (x_rows = 1), broadcast zi."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    from torchscience_amd.signal_processing.filter._ops import sosfilt_rows_forward
    dev = _dev()
    rng = np.random.default_rng(S)
    sos = np.concatenate([butterworth_sos(2, float(c)) for c in rng.uniform(0.1, 0.4, S)])
    B, T = 5, 9000
    x = rng.standard_normal((1, T))
    zi = rng.standard_normal((1, S, 2))
    y64, zf64 = oracle_mod.sosfilt_rows(sos, np.repeat(x, B, 0), np.repeat(zi, B, 0))
    print("S", S, flush=True)
    for path in ("window", "channel"):
        y, zf = sosfilt_rows_forward(torch.tensor(sos, device=dev)[None],
                                     torch.tensor(x, device=dev), torch.tensor(zi, device=dev),
                                     channels=B, chans_per_filter=B, want_zf=True,
                                     zi_broadcast=True, io_dtype=torch.float64, path=path)
        assert relerr(y.cpu().numpy(), y64) < 1e-10, path
        np.testing.assert_allclose(zf.cpu().numpy(), zf64, rtol=1e-9, atol=1e-9)


def test_channel_path_unaligned_rows(oracle_mod):
    """Row starts not 16-byte aligned take the per-lane I/O route."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    from torchscience_amd.signal_processing.filter._ops import sosfilt_rows_forward
    dev = _dev()
    rng = np.random.default_rng(3)
    sos = butterworth_sos(8, 0.2).astype(np.float32)
    B, T = 4, 3 * 4096 + 77
    x = rng.standard_normal((B, T + 1)).astype(np.float32)
    xt = torch.tensor(x, device=dev)[:, 1:]  # row stride T + 1, offset 4 bytes
    y64, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), x[:, 1:].astype(np.float64))
    y, _ = sosfilt_rows_forward(torch.tensor(sos, device=dev)[None], xt, None, channels=B,
                                chans_per_filter=B, want_zf=False, path="channel")
    assert relerr(y.cpu().numpy(), y64) < 1e-5


def test_conflicting_path_flags_rejected():
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._ops import _desc
    dev = _dev()
    L = _lib.load()
    x = torch.zeros(2, 100, device=dev)
    sos = torch.tensor([[1.0, 0, 0, 1, 0, 0]], device=dev)
    d = _desc(channels=2, n_samples=100, n_sections=1, io_dtype=torch.float32,
              coef_dtype=torch.float32, flags=_lib.PATH_WINDOW | _lib.PATH_CHANNEL, x_rows=2,
              x_row_stride=100, y_row_stride=100, chans_per_filter=2, mode=0)
    n = L.tsiir_sosfilt_workspace_bytes(d)
    ws = torch.zeros(n, dtype=torch.uint8, device=dev)
    st = _lib.stream_handle(dev)
    rc = L.tsiir_sosfilt_forward(d, x.data_ptr(), x.data_ptr(), sos.data_ptr(), None, None,
                                 ws.data_ptr(), n, st)
    assert rc == 1  # TSIIR_EINVAL


@pytest.mark.parametrize("io", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("T", [4096, 4097, 3 * 4096 + 5, 20000])
def test_compensated_many_channels(oracle_mod, io, T):
    """The compensated form (FP32 or BF16 data, FP64 updates) with enough
    channels for chained single-wave windows of 64 samples per lane; zi and
    zf; lengths around the 4096-sample window."""
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    from torchscience_amd.signal_processing.filter._ops import sosfilt_rows_forward
    dev = _dev()
    rng = np.random.default_rng(T + (io == torch.bfloat16))
    B, S = 1030, 4
    sos = batched_butterworth_sos(2 * S, np.array([rng.uniform(0.05, 0.45)]))
    x = rng.standard_normal((B, T))
    x_used = torch.tensor(x).to(io).double().numpy()
    zi = rng.standard_normal((B, S, 2))
    zi_used = torch.tensor(zi).float().double().numpy()
    print("compensated many channels", io, "T", T, flush=True)
    y, zf = sosfilt_rows_forward(torch.tensor(sos, device=dev), torch.tensor(x, device=dev).to(io),
                                 torch.tensor(zi, device=dev).float(), channels=B,
                                 chans_per_filter=B, want_zf=True, io_dtype=io, mode=1)
    y64, zf64 = oracle_mod.sosfilt_rows(sos[0], x_used, zi_used, n_threads=8)
    tol = 1e-5 if io == torch.float32 else 1e-2
    assert relerr(y.double().cpu().numpy(), y64) < tol
    np.testing.assert_allclose(zf.double().cpu().numpy(), zf64, rtol=1e-4, atol=1e-4)
