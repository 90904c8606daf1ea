"""Host-side logic on the CPU: planning, error behaviour, and the C ABI
library's exported symbols (loaded, not called: no GPU here)."""

import ctypes
import os
import re
import subprocess

import pytest
import torch

from torchscience_amd import _lib
from torchscience_amd.signal_processing.filter import sosfilt, lfilter, batched_filter_apply
from torchscience_amd.signal_processing.filter._plan import plan_lfilter, plan_sosfilt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "tsiir.h")).read()
    return set(re.findall(r"\b(tsiir_[a-z_0-9]+)\s*\(", text))


def test_header_and_binding_agree():
    assert _header_symbols() == set(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        from torchscience_amd import build
        build.build()
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in _header_symbols():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for name in _header_symbols():
        assert re.search(rf"\bT {name}$", out, re.M), name
    assert _lib.load().tsiir_version() == 1


def test_binding_arity_matches_header():
    """Every declared function's parameter count equals its ctypes argtypes."""
    text = open(os.path.join(ROOT, "include", "tsiir.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    L = _lib.load()
    for name, params in re.findall(r"\b(tsiir_[a-z_0-9]+)\s*\(([^)]*)\)\s*;", text):
        n = 0 if params.strip() in ("", "void") else params.count(",") + 1
        assert len(getattr(L, name).argtypes or []) == n, name


def test_desc_layout_matches_header():
    assert ctypes.sizeof(_lib.SosfiltDesc) == 72
    assert _lib.SosfiltDesc.chans_per_filter.offset == 64


def test_workspace_size_query_without_gpu():
    L = _lib.load()
    d = _lib.SosfiltDesc(channels=4096, n_samples=65536, n_sections=6, io_dtype=0, coef_dtype=0,
                         mode=0, flags=0, reserved=0, x_rows=4096, x_row_stride=65536,
                         y_row_stride=65536, chans_per_filter=4096)
    n = L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d))
    # 1024-byte header + one 896-byte table entry per section (rounded to 256 B),
    # then one 1088-byte FP32-kernel entry per section (rounded to 256 B), then
    # 64 B of status per (window, channel, section); windows of 4096 samples
    tables = 1024 + 5376 + 6 * 1088
    assert n == (tables + 255) // 256 * 256 + 16 * 4096 * 6 * 64
    d.n_sections = 0
    assert L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d)) == 0
    assert L.tsiir_status_string(2) == b"unsupported request"


def test_segmented_workspace_plan_without_gpu():
    """Few long rows (8 x 2^24, the long-sequence shape) size the workspace
    for the segmented form as well: 1024 segments of 131072 samples (the
    longest power of two that still gives 1024 rows), the largest of the
    warm-up, main and remainder passes' own workspaces, the 4096-sample
    warm-up outputs, three FP64 state arrays over the segments and two over
    the channels; TSIIR_NO_SEGMENT, 1024 channels or rows too short for
    1024 segments keep the look-back form's size."""
    L = _lib.load()

    def ws(B, T, flags=0):
        d = _lib.SosfiltDesc(channels=B, n_samples=T, n_sections=6, io_dtype=_lib.F32,
                             coef_dtype=_lib.F32, mode=0, flags=flags, reserved=0, x_rows=B,
                             x_row_stride=T, y_row_stride=T, chans_per_filter=B)
        return L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d))

    def align(v):
        return (v + 255) // 256 * 256

    def plain(B, T):
        # header, one 896-byte table entry per section, one 1088-byte FP32
        # entry per section, 64 B of status per (4096-sample window, row, section)
        return align(1024 + 5376 + 6 * 1088) + (-(-T // 4096)) * B * 6 * 64

    def seg(B, T, rows, lseg, wrows, rem):
        base = max(plain(wrows, 4096), plain(rows, lseg), plain(B, rem) if rem else 0)
        states, cstates = rows * 6 * 2 * 8, B * 6 * 2 * 8
        off = align(align(base) + rows * 4096 * 4)       # warm-up outputs
        off = align(align(align(off + states) + states) + states)  # warm, zin, zall
        return align(off + cstates) + cstates             # remainder's zi, zf

    B, T = 8, 1 << 24
    assert ws(B, T, _lib.NO_SEGMENT) == plain(B, T)
    assert ws(B, T) == max(seg(B, T, 1024, 131072, 1024, 0), plain(B, T))
    assert ws(B, T, _lib.WRITE_ZF) == ws(B, T)
    # REVERSE: no warm-up after the very last segment
    assert ws(B, T, _lib.REVERSE) == max(seg(B, T, 1024, 131072, 1023, 0), plain(B, T))
    # a remainder: 8 x (2^21 + 78) splits into 128 segments of 16384 and 78 samples
    T2 = (1 << 21) + 78
    assert ws(B, T2) == max(seg(B, T2, 1024, 16384, 1024, 78), plain(B, T2))
    assert ws(1024, 65536) == plain(1024, 65536)
    # too short for 1024 segments of at least 16384 samples
    assert ws(8, 105 * 16384 + 4096) == plain(8, 105 * 16384 + 4096)


def test_fir_fft_host_checks_without_gpu():
    """The FFT FIR entry point's size query and the argument checks that
    return before any device work."""
    L = _lib.load()

    def desc(**kw):
        d = _lib.FirDesc(channels=7, n_samples=1000, x_len=1000, n_taps=101, io_dtype=_lib.F32,
                         shift=0, flags=0, x_rows=7, x_row_stride=1000, y_row_stride=1000,
                         chans_per_filter=7)
        for k, v in kw.items():
            setattr(d, k, v)
        return d
    # twiddles (1024) + one taps' transform (4096) per filter, complex in the
    # I/O type
    assert L.tsiir_fir_fft_workspace_bytes(desc()) == 5120 * 8
    assert L.tsiir_fir_fft_workspace_bytes(desc(io_dtype=_lib.F64)) == 5120 * 16
    assert L.tsiir_fir_fft_workspace_bytes(desc(chans_per_filter=1)) == (1024 + 7 * 4096) * 8
    assert L.tsiir_fir_fft_workspace_bytes(desc(chans_per_filter=9)) == 5120 * 8
    assert L.tsiir_fir_fft_workspace_bytes(desc(n_taps=2049)) == 0
    assert L.tsiir_fir_fft_workspace_bytes(desc(chans_per_filter=2)) == 0
    for d, code in ((desc(flags=_lib.HAS_ZI), _lib.EINVAL),
                    (desc(flags=_lib.WRITE_ZF), _lib.EINVAL),
                    (desc(chans_per_filter=3), _lib.EINVAL),
                    (desc(n_taps=2049), _lib.EUNSUPPORTED),
                    (desc(n_taps=0), _lib.EUNSUPPORTED),
                    (desc(io_dtype=_lib.BF16), _lib.EUNSUPPORTED)):
        assert L.tsiir_fir_forward_fft(d, None, None, None, None, 0, None) == code
    assert L.tsiir_fir_forward_fft(desc(channels=0, x_rows=1), None, None, None, None, 0,
                                   None) == _lib.OK


def test_sos_shape_error_matches_reference():
    with pytest.raises(ValueError, match="shape"):
        sosfilt(torch.ones(1, 3), torch.randn(10))
    with pytest.raises(ValueError, match=r"sos must be shape \(n_sections, 6\)"):
        sosfilt(torch.ones(2, 6, 1), torch.randn(10))


def test_lfilter_1d_errors():
    with pytest.raises(ValueError, match="b must be a 1-D tensor"):
        lfilter(torch.ones(2, 2), torch.ones(2), torch.randn(10))
    with pytest.raises(ValueError, match="a must be a 1-D tensor"):
        lfilter(torch.ones(2), torch.ones(2, 2), torch.randn(10))
    assert plan_lfilter((5,), (2,)) == 4
    assert plan_lfilter((), ()) == 0


def test_batched_errors():
    with pytest.raises(ValueError, match=r"SOS filters must have shape \(batch, n_sections, 6\)"):
        batched_filter_apply(torch.ones(2, 6), torch.randn(10))
    with pytest.raises(ValueError, match="x batch dim must match filters"):
        batched_filter_apply(torch.ones(3, 2, 6), torch.randn(2, 10), broadcast=False)
    with pytest.raises(ValueError, match="filter_format must be 'sos' or 'fir'"):
        batched_filter_apply(torch.ones(3, 2, 6), torch.randn(10), filter_format="ba")


def test_cpu_tensors_fail_loudly():
    with pytest.raises(RuntimeError, match="HIP"):
        sosfilt(torch.ones(2, 6, dtype=torch.float64), torch.randn(10, dtype=torch.float64))


def test_plan_promotion_and_zi_broadcast():
    f32, f64 = torch.float32, torch.float64
    p = plan_sosfilt((4, 6), f64, (3, 100), f32)
    assert p.out_dtype == f64 and p.compute_dtype == f64 and p.channels == 3
    p = plan_sosfilt((4, 6), f32, (100,), f32)
    assert p.out_dtype == f32 and p.batch_shape == () and p.channels == 1
    p = plan_sosfilt((4, 6), f32, (5, 100, 8), f32, axis=1, zi_shape=(4, 2), zi_dtype=f64)
    assert p.zi_broadcast and p.zf_shape == (5, 8, 4, 2) and p.zf_dtype == f64
    assert p.channels == 40 and p.n_samples == 100
    p = plan_sosfilt((4, 6), f32, (100,), f32, zi_shape=(4, 2), zi_dtype=f32)
    assert not p.zi_broadcast and p.zf_shape == (4, 2)
    p = plan_sosfilt((4, 6), f32, (3, 100), torch.complex64)
    assert p.complex_input and p.compute_dtype == f32
    p = plan_sosfilt((4, 6), f32, (3, 100), torch.int64)
    assert p.out_dtype == f32
    p = plan_sosfilt((4, 6), f32, (3, 100), f32, precision="bf16io")
    assert p.io_dtype == torch.bfloat16 and p.out_dtype == torch.bfloat16
    f64 = torch.float64
    for prec in ("fp32", "fp32c"):
        p = plan_sosfilt((4, 6), f64, (3, 100), f32, precision=prec)
        assert p.compute_dtype == p.io_dtype == p.out_dtype == f32
        assert p.compensated == (prec == "fp32c")
    p = plan_sosfilt((4, 6), f32, (3, 100), f32, precision="fp64")
    assert p.compute_dtype == p.io_dtype == p.out_dtype == f64
    with pytest.raises(RuntimeError, match="zi must have shape"):
        plan_sosfilt((4, 6), f32, (3, 100), f32, zi_shape=(2, 4, 2), zi_dtype=f32)
    with pytest.raises(NotImplementedError):
        plan_sosfilt((4, 6), torch.complex128, (3, 100), f32)


def test_design_matches_scipy_magnitude():
    import numpy as np
    import scipy.signal as ss
    from torchscience_amd.signal_processing.filter._design import (butterworth_sos,
                                                                    chebyshev1_sos)
    pairs = [(butterworth_sos(6, [0.05, 0.3], "bandpass"),
              ss.butter(6, [0.05, 0.3], "bandpass", output="sos")),
             (butterworth_sos(4, 0.125), ss.butter(4, 0.125, output="sos")),
             (butterworth_sos(8, 0.33), ss.butter(8, 0.33, output="sos")),
             (chebyshev1_sos(16, 0.01, 1.0), ss.cheby1(16, 1.0, 0.01, output="sos"))]
    for sos, ref in pairs:
        _, h = ss.sosfreqz(sos, 4096)
        _, h2 = ss.sosfreqz(ref, 4096)
        assert np.max(np.abs(np.abs(h) - np.abs(h2))) / np.max(np.abs(h2)) < 1e-10
    # the reference's even gain split: every section carries |k|^(1/S)
    sos = butterworth_sos(4, 0.125)
    assert np.allclose(sos[:, 0], sos[0, 0]) and sos.shape == (2, 6)


def test_zero_phase_validation_matches_reference():
    """sosfiltfilt / filtfilt raise the reference's ValueErrors before any
    device work (_sosfiltfilt.py:96-121, _filtfilt.py:86-115)."""
    from torchscience_amd.signal_processing.filter import filtfilt, sosfiltfilt
    sos = torch.tensor([[0.2, 0.4, 0.2, 1.0, -0.5, 0.2]], dtype=torch.float64)
    with pytest.raises(ValueError, match=r"sos must be shape \(n_sections, 6\)"):
        sosfiltfilt(torch.ones(2, 5), torch.randn(100))
    with pytest.raises(ValueError, match=r"padlen \(9\) must be less than x.shape\[axis\] \(9\)"):
        sosfiltfilt(sos, torch.randn(9, dtype=torch.float64))
    with pytest.raises(ValueError, match="Unknown padtype: wrap"):
        sosfiltfilt(sos, torch.randn(100, dtype=torch.float64), padtype="wrap")
    with pytest.raises(ValueError, match="a must be a 1-D tensor"):
        filtfilt(torch.ones(3), torch.ones(2, 2), torch.randn(100))
    with pytest.raises(ValueError, match=r"padlen \(12\) must be less"):
        filtfilt(torch.ones(3), torch.ones(4), torch.randn(12))
    with pytest.raises(RuntimeError, match="HIP"):
        sosfiltfilt(sos, torch.randn(100, dtype=torch.float64))


def test_batched_design_argument_checks():
    """The reference's ValueErrors of the batched designs (_batched.py:84-117,
    _butterworth_design.py:102-113, the prototypes' checks), raised before
    any device work."""
    from torchscience_amd.signal_processing.filter import (
        batched_butterworth_design, batched_chebyshev_type_1_design,
        batched_chebyshev_type_2_design)
    with pytest.raises(ValueError, match="cutoffs tensor cannot be empty"):
        batched_butterworth_design(4, torch.tensor([]))
    with pytest.raises(ValueError, match="Batched design only supports 'sos' output"):
        batched_butterworth_design(4, torch.tensor([0.2]), output="zpk")
    with pytest.raises(ValueError, match=r"cutoffs must have shape \(batch, 2\)"):
        batched_butterworth_design(4, torch.tensor([0.2, 0.3]), filter_type="bandstop")
    with pytest.raises(ValueError, match="must be 1D or"):
        batched_butterworth_design(4, torch.ones(3, 2) * 0.2)
    with pytest.raises(ValueError, match="between 0 and 1"):
        batched_butterworth_design(4, torch.tensor([0.0]))
    with pytest.raises(ValueError, match="0 < low < high < 1"):
        batched_butterworth_design(4, torch.tensor([[0.3, 0.2]]), filter_type="bandpass")
    with pytest.raises(ValueError, match="Filter order must be positive"):
        batched_butterworth_design(-1, torch.tensor([0.2]))
    with pytest.raises(ValueError, match="Passband ripple must be positive"):
        batched_chebyshev_type_1_design(4, -0.5, torch.tensor([0.2]))
    with pytest.raises(ValueError, match="Stopband attenuation must be positive"):
        batched_chebyshev_type_2_design(4, 0.0, torch.tensor([0.2]))
    with pytest.raises(ValueError, match="Invalid filter_type"):
        batched_butterworth_design(4, torch.tensor([0.2]), filter_type="allpass")
    with pytest.raises(ValueError, match="between 0 and 1"):
        batched_butterworth_design(4, torch.tensor([9000.0]), sampling_frequency=16000.0)


def test_frequency_response_validation_before_device_work():
    """a0 != 1 raises the reference's SOSNormalizationError
    (_frequency_response_sos.py:84-89), a FilterDesignError."""
    from torchscience_amd.signal_processing.filter import (
        FilterDesignError, SOSNormalizationError, frequency_response_sos)
    assert issubclass(SOSNormalizationError, FilterDesignError)
    with pytest.raises(SOSNormalizationError, match="a0 = 1"):
        frequency_response_sos(torch.tensor([[1.0, 0.5, 0.25, 2.0, 0.1, 0.05]]))


def test_lfiltic_host_math_and_errors():
    """lfiltic is set-up math on the tensors' device (here the CPU): equal
    to the oracle's restatement of _lfilter.py:181-305, the reference's
    ValueErrors, the empty state of a pure gain."""
    import numpy as np
    from oracle import oracle
    from torchscience_amd.signal_processing.filter import lfiltic
    rng = np.random.default_rng(3)
    b = np.array([0.1, 0.2, 0.1, 0.05])
    a = np.array([2.0, -0.5, 0.1])
    y = rng.standard_normal(5)
    x = rng.standard_normal(2)
    zi = lfiltic(torch.tensor(b), torch.tensor(a), torch.tensor(y), torch.tensor(x))
    np.testing.assert_allclose(zi.numpy(), oracle.lfiltic(b, a, y, x), rtol=1e-12, atol=1e-13)
    zi = lfiltic(torch.tensor(b), torch.tensor(a), torch.tensor(y))
    np.testing.assert_allclose(zi.numpy(), oracle.lfiltic(b, a, y), rtol=1e-12, atol=1e-13)
    assert zi.dtype == torch.float64
    assert lfiltic(torch.tensor([2.0]), torch.tensor([1.0]), torch.tensor([1.0])).numel() == 0
    with pytest.raises(ValueError, match="y must be a 1-D tensor"):
        lfiltic(torch.ones(2), torch.ones(2), torch.ones(2, 2))
    with pytest.raises(ValueError, match="b must be a 1-D tensor"):
        lfiltic(torch.ones(2, 2), torch.ones(2), torch.ones(2))


@pytest.mark.parametrize("zi_shape", [(2, 4), (5, 4), (4, 3), (3, 2, 4)])
def test_lfilter_zi_rows_checked_before_launch(zi_shape):
    """lfilter of order > 2 (FIR and general form) reads zi as one state row
    per channel, or one row for all: any other row count raises before any
    device work (the reference fails on it too, _lfilter.py:136-166)."""
    from torchscience_amd.signal_processing.filter._plan import plan_lfilter_zi
    x = torch.randn(4, 50)
    zi = torch.zeros(zi_shape)
    for a in (torch.ones(1), torch.tensor([1.0, -0.5, 0.1, 0.0, 0.01])):
        with pytest.raises(RuntimeError, match="zi must have shape"):
            lfilter(torch.ones(5), a, x, zi=zi)
    assert plan_lfilter_zi((4, 4), (4,), 4) is False
    assert plan_lfilter_zi((4,), (4,), 4) is True
    assert plan_lfilter_zi((1, 4), (4,), 4) is True
    assert plan_lfilter_zi((2, 2, 4), (2, 2), 4) is False


# ---- the operator boundary: TORCH_LIBRARY schemas, Meta implementations ----

def test_operator_schemas_registered():
    """libtsiir_torch.so registers tsiir::sosfilt_forward / sosfilt_backward
    (the counterpart of the reference's TORCH_LIBRARY registration,
    src/torchscience/csrc/torchscience.cpp:811) with HIP and Meta kernels."""
    ops = _lib.ops()
    fwd = str(ops.sosfilt_forward.default._schema)
    bwd = str(ops.sosfilt_backward.default._schema)
    assert fwd.startswith("tsiir::sosfilt_forward(Tensor x, Tensor sos, Tensor? zi")
    assert "-> (Tensor y, Tensor zf, Tensor saved)" in fwd
    assert "-> (Tensor dx, Tensor dsos, Tensor dzi)" in bwd
    for name in ("sosfilt_forward", "sosfilt_backward"):
        op = getattr(ops, name).default
        assert torch._C._dispatch_has_kernel_for_dispatch_key(op.name(), "CUDA")
        assert torch._C._dispatch_has_kernel_for_dispatch_key(op.name(), "Meta")


@pytest.mark.parametrize("sos_dt,x_dt,out_dt", [(torch.float32, torch.float32, torch.float32),
                                                (torch.float64, torch.float32, torch.float64),
                                                (torch.float32, torch.bfloat16, torch.float32)])
def test_meta_shape_inference(sos_dt, x_dt, out_dt):
    """sosfilt on meta tensors: output shape, dtype promotion and zf layout
    (zi with the batch dims of x and zi's dtype), with no device work."""
    sos = torch.empty(4, 6, device="meta", dtype=sos_dt)
    x = torch.empty(3, 7, 1000, device="meta", dtype=x_dt)
    y = sosfilt(sos, x, axis=1)
    assert y.shape == x.shape and y.device.type == "meta" and y.dtype == out_dt
    zi = torch.empty(3, 1000, 4, 2, device="meta", dtype=torch.float64)
    y, zf = sosfilt(sos, x, axis=1, zi=zi)
    assert y.shape == x.shape and zf.shape == zi.shape and zf.dtype == torch.float64
    yb, zfb = sosfilt(sos, x, zi=torch.empty(4, 2, device="meta"))
    assert zfb.shape == (3, 7, 4, 2)
    y2 = batched_filter_apply(torch.empty(5, 2, 6, device="meta"),
                              torch.empty(5, 300, device="meta"), broadcast=False)
    assert y2.shape == (5, 300)


def test_torch_compile_traces_sosfilt_fullgraph():
    """torch.compile(fullgraph=True) traces sosfilt (forward) and its
    gradient (the autograd.Function around the operators) without a graph
    break, on meta tensors (FakeTensor tracing through the Meta kernels)."""
    torch._dynamo.reset()
    sos = torch.randn(4, 6, device="meta")
    x = torch.randn(3, 5, 1000, device="meta")
    f = torch.compile(sosfilt, fullgraph=True, backend="eager")
    assert f(sos, x, axis=1).shape == x.shape
    sos_g = torch.randn(4, 6, device="meta", requires_grad=True)
    x_g = torch.randn(8, 1000, device="meta", requires_grad=True)
    zi_g = torch.zeros(8, 4, 2, device="meta", requires_grad=True)

    def loss(s, v, z):
        y, zf = sosfilt(s, v, zi=z)
        return y.square().sum() + zf.sum()
    g = torch.compile(loss, fullgraph=True, backend="aot_eager")
    g(sos_g, x_g, zi_g).backward()
    assert sos_g.grad.shape == (4, 6) and x_g.grad.shape == (8, 1000)
    assert zi_g.grad.shape == (8, 4, 2)


def test_pole_radius_routing_criterion():
    """tsiir::max_pole_radius (the FP32 DF2T routing criterion) equals the
    largest |root| of a0 z^2 + a1 z + a2 over the sections."""
    import numpy as np
    from torchscience_amd.signal_processing.filter._design import (butterworth_sos,
                                                                    chebyshev1_sos)
    for sos in (chebyshev1_sos(16, 0.01, 1.0), butterworth_sos(6, [0.05, 0.3], "bandpass"),
                butterworth_sos(8, 0.02), np.array([[1.0, 0, 0, 2.0, -3.0, 1.12]]),
                np.array([[1.0, 0, 0, 1.0, -1.9, 0.9]])):
        r = _lib.ops().max_pole_radius(torch.tensor(sos))
        ref = max(np.max(np.abs(np.roots(row[3:]))) for row in sos)
        assert abs(r - ref) < 1e-12
    assert _lib.ops().max_pole_radius(torch.tensor([[1.0, 0, 0, 0.0, 1.0, 0.5]])) == float("inf")


def test_sosfilt_zi_host_equals_sosfilt_zi():
    """The host steady state used by the kernel form of sosfiltfilt equals
    sosfilt_zi (torch ops, the reference's statement order) to rounding, and
    scipy's sosfilt_zi; a0 != 1 is normalised, a section with a0 == 0 and a
    singular system take sosfilt_zi's path and errors."""
    import numpy as np
    import scipy.signal as ss
    from torchscience_amd.signal_processing.filter._design import (butterworth_sos,
                                                                    chebyshev1_sos)
    from torchscience_amd.signal_processing.filter._zi import sosfilt_zi, sosfilt_zi_host
    cases = [butterworth_sos(6, [0.05, 0.3], "bandpass"), chebyshev1_sos(16, 0.01, 1.0),
             ss.butter(4, 0.2, output="sos") * 2.0, ss.cheby2(6, 40, 0.3, output="sos")]
    for sos in cases:
        s32 = sos.astype(np.float32).astype(np.float64)  # the values the host copy holds
        got = sosfilt_zi_host(torch.tensor(sos, dtype=torch.float32)).numpy()
        ref = sosfilt_zi(torch.tensor(s32, dtype=torch.float64)).numpy()
        sp = ss.sosfilt_zi(s32)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
        assert np.max(np.abs(got - sp)) <= 1e-9 * np.max(np.abs(sp))
    lead0 = torch.tensor([[1.0, 0.5, 0.2, 0.0, 1.0, -0.3]], dtype=torch.float64)
    assert torch.allclose(sosfilt_zi_host(lead0), sosfilt_zi(lead0))
    singular = torch.tensor([[1.0, 0.0, 0.0, 1.0, -2.0, 1.0]], dtype=torch.float64)
    with pytest.raises(Exception) as e_ref:
        sosfilt_zi(singular)
    with pytest.raises(type(e_ref.value)):
        sosfilt_zi_host(singular)
