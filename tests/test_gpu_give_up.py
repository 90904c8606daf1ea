"""Look-back waits that give up are reported, never silently wrong (GPU).

The host forces the give-up with the testing flag TSIIR_DEBUG_GIVE_UP
(every wait gives up at once, so no kernel waits or hangs): the first
window (or group) of every channel needs no wait and stays exact; every
window after it, and the final state, come out NaN; the workspace header's
error word and the registered mailbox are set; and the next product call on
that stream raises (without synchronising), after which a fresh
workspace serves calls normally again."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


# (channels, samples, dtype, extra flags, samples of the first span):
# FP32 chained 4-window groups, FP32 look-back 8-window groups, FP32 DF2T
# chained windows, FP64 4-wave look-back windows
CASES = [
    (1024, 2 * 16384, torch.float32, 0, 16384),
    (4, 3 * 32768, torch.float32, 0, 32768),
    (1024, 3 * 4096, torch.float32, 0x80, 4096),
    (4, 3 * 8192, torch.float64, 0, 8192),
]


@pytest.mark.parametrize("B,T,dtype,extra,first", CASES)
def test_give_up_makes_nan_and_sets_mailbox(oracle_mod, B, T, dtype, extra, first):
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    L = _lib.load()
    print(f"give-up: B={B} T={T} {dtype} flags+{extra:#x}", flush=True)
    sos_np = butterworth_sos(8, 0.2)
    sos = torch.tensor(sos_np, dtype=dtype, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(B + T)
    x = torch.randn(B, T, device=dev, generator=g, dtype=dtype)
    y = torch.empty_like(x)
    zf = torch.empty(B, 4, 2, device=dev, dtype=dtype)
    io = _lib.io_code(dtype)
    d = _lib.SosfiltDesc(channels=B, n_samples=T, n_sections=4, io_dtype=io, coef_dtype=io,
                         mode=0, flags=_lib.WRITE_ZF | _lib.DEBUG_GIVE_UP | extra, reserved=0,
                         x_rows=B, x_row_stride=T, y_row_stride=T, chans_per_filter=B)
    nbytes = L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    mailbox = torch.zeros(1, dtype=torch.int32, pin_memory=True)
    assert L.tsiir_workspace_init(ws.data_ptr(), nbytes, st) == _lib.OK
    assert L.tsiir_workspace_set_error_mailbox(ws.data_ptr(), nbytes, mailbox.data_ptr(),
                                               st) == _lib.OK
    assert L.tsiir_sosfilt_forward(ctypes.byref(d), x.data_ptr(), y.data_ptr(),
                                   sos.data_ptr(), None, zf.data_ptr(), ws.data_ptr(), nbytes,
                                   st) == _lib.OK
    torch.cuda.synchronize(dev)
    assert int(ws[12:16].view(torch.int32).item()) != 0
    assert int(mailbox[0]) != 0
    yc = y.cpu().numpy()
    # the first span of every channel waits for nothing: exact
    idx = [0, B - 1]
    ref, _ = oracle_mod.sosfilt_rows(sos_np, x[idx, :first].double().cpu().numpy())
    tol = 1e-10 if dtype == torch.float64 else 1e-5
    head = yc[idx, :first].astype(np.float64)
    assert np.isfinite(head).all()
    assert np.max(np.abs(head - ref)) <= tol * np.max(np.abs(ref))
    # everything after it depends on a state that was given up on
    assert np.isnan(yc[:, first:]).all()
    assert torch.isnan(zf).all()


def test_product_call_reports_an_earlier_give_up():
    from torchscience_amd import check_errors
    from torchscience_amd.signal_processing.filter import sosfilt
    from torchscience_amd.signal_processing.filter._ops import sosfilt_rows_forward
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    sos = torch.tensor(butterworth_sos(8, 0.2), dtype=torch.float32, device=dev)[None]
    x = torch.randn(8, 3 * 32768, device=dev)
    y, _ = sosfilt_rows_forward(sos, x, None, channels=8, chans_per_filter=8, want_zf=False,
                                _debug_give_up=True)
    torch.cuda.synchronize(dev)
    assert torch.isnan(y[:, 32768:]).all()
    with pytest.raises(RuntimeError, match="gave up"):
        sosfilt(sos[0], x)
    # the failed workspace was dropped: calls work again, and nothing is pending
    y2 = sosfilt(sos[0], x)
    torch.cuda.synchronize(dev)
    assert torch.isfinite(y2).all()
    check_errors(dev)
