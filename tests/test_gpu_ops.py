"""The operator boundary on the GPU: tsiir::sosfilt_forward / backward
called by the drop-in, torch.compile(fullgraph=True) of sosfilt equal to
eager bitwise (forward and gradients), and torch.library.opcheck's schema /
FakeTensor / dispatch checks of both operators."""

import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def test_compiled_sosfilt_equals_eager():
    from torchscience_amd.signal_processing.filter import sosfilt
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    torch._dynamo.reset()
    sos = torch.tensor(butterworth_sos(8, 0.2), dtype=torch.float32, device=dev)
    x = torch.randn(64, 20000, device=dev)
    zi = torch.randn(64, 4, 2, device=dev)
    y, zf = sosfilt(sos, x, zi=zi)
    f = torch.compile(sosfilt, fullgraph=True)
    yc, zfc = f(sos, x, zi=zi)
    assert torch.equal(y, yc) and torch.equal(zf, zfc)

    def loss(s, v, z):
        yy, zz = sosfilt(s, v, zi=z)
        return yy.square().mean() + zz.sum()
    grads = []
    for fn in (loss, torch.compile(loss, fullgraph=True)):
        s_ = sos.clone().requires_grad_(True)
        v_ = x.clone().requires_grad_(True)
        z_ = zi.clone().requires_grad_(True)
        fn(s_, v_, z_).backward()
        grads.append((s_.grad, v_.grad, z_.grad))
    for a, b in zip(*grads):
        torch.testing.assert_close(a, b, rtol=1e-6, atol=1e-7)


def test_opcheck_operators():
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    ops = _lib.ops()
    sos = torch.tensor(butterworth_sos(8, 0.2), dtype=torch.float32, device=dev)[None]
    x = torch.randn(16, 9000, device=dev)
    zi = torch.randn(16, 4, 2, device=dev)
    args = (x, sos, zi, 16, 16, True, False, 0, 0, True)
    torch.library.opcheck(ops.sosfilt_forward.default, args,
                          test_utils=("test_schema", "test_faketensor"))
    y, zf, saved = ops.sosfilt_forward(*args)
    dy = torch.randn_like(y)
    dzf = torch.randn_like(zf)
    bargs = (x, y, saved, sos, zi, dy, dzf, 16, False)
    torch.library.opcheck(ops.sosfilt_backward.default, bargs,
                          test_utils=("test_schema", "test_faketensor"))


def test_compiled_ctypes_bound_paths_equal_eager():
    """The paths bound through ctypes (general-order lfilter, FIR lfilter,
    frequency_response_sos) run under torch.compile with a graph break at
    the native call and give the eager results."""
    from torchscience_amd.signal_processing.filter import frequency_response_sos, lfilter
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    x = torch.randn(4, 3000, generator=g, device=dev)
    b = torch.tensor([0.02, 0.08, 0.12, 0.08, 0.02], device=dev)
    a = torch.tensor([1.0, -1.6, 1.3, -0.5, 0.08], device=dev)
    taps = torch.hamming_window(64, device=dev) / 32

    @torch.compile
    def run(x):
        return lfilter(b, a, x), lfilter(taps, torch.ones(1, device=dev), x)

    y1, y2 = run(x)
    assert torch.equal(y1, lfilter(b, a, x)) and torch.equal(y2, lfilter(taps, torch.ones(1, device=dev), x))
    sos = torch.tensor(butterworth_sos(6, 0.2), device=dev)
    w, h = torch.compile(lambda s: frequency_response_sos(s, 256))(sos)
    w0, h0 = frequency_response_sos(sos, 256)
    assert torch.equal(h, h0) and torch.equal(w, w0)
