"""tsiir_sosfilt_forward_noalloc (the entry point that allocates nothing:
its workspace is a fixed device array of the library) against
tsiir_sosfilt_forward with a caller-owned workspace, through the C ABI
(GPU).  Two calls in a row, the second smaller, must equal the workspace
entry point's output bitwise (same kernels, same work decomposition); a
request whose workspace exceeds the fixed array is rejected with
TSIIR_ESCRATCH before any launch."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _desc(B, T, S, io, flags=0):
    from torchscience_amd import _lib
    return _lib.SosfiltDesc(channels=B, n_samples=T, n_sections=S, io_dtype=io,
                            coef_dtype=io, mode=_lib.MODE_PLAIN, flags=flags, reserved=0,
                            x_rows=B, x_row_stride=T, y_row_stride=T, chans_per_filter=B)


def _both(B, T, S, dtype, with_state, seed):
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    L = _lib.load()
    io = _lib.io_code(dtype)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    x = torch.randn(B, T, device=dev, generator=g, dtype=dtype)
    sos = torch.tensor(butterworth_sos(2 * S, 0.17), dtype=dtype, device=dev)
    flags = 0
    zi = zf_a = zf_b = None
    if with_state:
        flags = _lib.HAS_ZI | _lib.WRITE_ZF
        zi = torch.randn(B, S, 2, device=dev, generator=g, dtype=dtype)
        zf_a = torch.empty(B, S, 2, device=dev, dtype=dtype)
        zf_b = torch.empty(B, S, 2, device=dev, dtype=dtype)
    d = _desc(B, T, S, io, flags)
    nbytes = L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d))
    assert 0 < nbytes <= _lib.NOALLOC_SCRATCH_BYTES
    ws = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    st = torch.cuda.current_stream(dev).cuda_stream
    ya = torch.empty_like(x)
    yb = torch.full_like(x, float("nan"))
    print(f"noalloc vs workspace: B={B} T={T} S={S} {dtype} state={with_state}", flush=True)
    assert L.tsiir_sosfilt_forward(ctypes.byref(d), _lib.ptr(x), _lib.ptr(ya), _lib.ptr(sos),
                                   _lib.ptr(zi), _lib.ptr(zf_a), ws.data_ptr(), nbytes,
                                   st) == _lib.OK
    assert L.tsiir_sosfilt_forward_noalloc(ctypes.byref(d), _lib.ptr(x), _lib.ptr(yb),
                                           _lib.ptr(sos), _lib.ptr(zi), _lib.ptr(zf_b),
                                           st) == _lib.OK
    torch.cuda.synchronize(dev)
    assert torch.equal(ya, yb)
    if with_state:
        assert torch.equal(zf_a, zf_b)
    return ya


def test_noalloc_equals_workspace_entry_point_two_calls():
    # first call: many channels (chained windows), large workspace
    _both(2048, 40000, 6, torch.float32, False, 1)
    # second call, smaller: few channels (look-back windows), zi / zf
    _both(37, 9001, 4, torch.float32, True, 2)
    # and a float64 one (the FP64 kernel)
    _both(16, 4096, 4, torch.float64, True, 3)


def test_noalloc_rejects_requests_beyond_its_scratch():
    from torchscience_amd import _lib
    _dev()
    L = _lib.load()
    d = _desc(2048, 1 << 20, 8, _lib.F32)
    assert L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d)) > _lib.NOALLOC_SCRATCH_BYTES
    # rejected on its descriptor alone: nothing is launched, no pointer read
    rc = L.tsiir_sosfilt_forward_noalloc(ctypes.byref(d), None, None, None, None, None, None)
    assert rc == _lib.ESCRATCH
    assert b"scratch" in L.tsiir_status_string(rc)
