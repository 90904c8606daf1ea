"""sosfiltfilt through the operator tsiir::sosfiltfilt (C ABI
tsiir_sosfiltfilt: edge kernels for the extensions, two sosfilt passes over
the rows) and, beyond its bounds, through separate sosfilt calls (GPU).

Against the oracle's sosfiltfilt (FP64, the reference's padding and
passes) on the same rounded inputs, per-row normwise relative error: FP64
1e-10, FP32 1e-5, BF16 1e-2 (bfloat16 coefficients and signal, FP32
arithmetic, the output rounded to bfloat16).
Cases: the four padtypes, padlen None / 0 / 5 / 2000 (a long extension in
the edge kernels), odd lengths, one row, 1-D input, axis 0, a 10-section
filter and padlen x sections above 16384 (both the separate-call form);
filtfilt of orders 1 and 2 (one section, the same kernel form) against the
oracle's filtfilt.
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
    y = np.asarray(y, dtype=np.float64).reshape(-1, np.shape(y)[-1])
    y64 = np.asarray(y64, dtype=np.float64).reshape(y.shape)
    return float(np.max(np.max(np.abs(y - y64), axis=-1) /
                        np.maximum(np.max(np.abs(y64), axis=-1), 1e-300)))


TOL = {torch.float64: 1e-10, torch.float32: 1e-5, torch.bfloat16: 1e-2}


def _bandpass():
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    return butterworth_sos(6, [0.05, 0.3], "bandpass")


def _run(oracle_mod, sos, x, io, **kw):
    from torchscience_amd.signal_processing.filter import sosfiltfilt
    dev = _dev()
    # coefficients in the I/O type, as the oracle receives them (BF16: both
    # bfloat16, the output bfloat16, the arithmetic FP32)
    st = torch.tensor(sos, device=dev).to(io)
    sos_k = st.double().cpu().numpy()
    xt = torch.tensor(x, device=dev).to(io)
    y = sosfiltfilt(st, xt, **kw)
    assert y.shape == xt.shape and y.dtype == io
    ref = oracle_mod.sosfiltfilt(sos_k, xt.double().cpu().numpy(), **kw)
    if kw.get("axis", -1) == 0:
        y, ref = y.T, ref.T
    return relerr(y.double().cpu().numpy(), ref)


@pytest.mark.parametrize("io", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("padtype", ["odd", "even", "constant", None])
@pytest.mark.parametrize("padlen", [None, 0, 5, 2000])
def test_padding_forms(oracle_mod, io, padtype, padlen):
    print(f"sosfiltfilt io={io} padtype={padtype} padlen={padlen}", flush=True)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((64, 4096 * 2 + 77))
    err = _run(oracle_mod, _bandpass(), x, io, padtype=padtype, padlen=padlen)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < TOL[io]


@pytest.mark.parametrize("shape,axis", [((1, 5000), -1), ((5001,), -1), ((3000, 7), 0),
                                        ((1024, 333), -1)])
def test_shapes(oracle_mod, shape, axis):
    print(f"sosfiltfilt shape={shape} axis={axis}", flush=True)
    x = np.random.default_rng(len(shape)).standard_normal(shape)
    assert _run(oracle_mod, _bandpass(), x, torch.float32, axis=axis) < 1e-5


def test_beyond_the_operator_bounds(oracle_mod):
    """Ten sections, and 6 sections with padlen 2999 (17994 > 16384): the
    pieces run as separate sosfilt calls."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    x = np.random.default_rng(5).standard_normal((32, 3000))
    assert _run(oracle_mod, butterworth_sos(20, 0.2), x, torch.float32) < 1e-5
    assert _run(oracle_mod, _bandpass(), x, torch.float32, padlen=2999) < 1e-5


@pytest.mark.parametrize("io", [torch.float64, torch.float32])
@pytest.mark.parametrize("padtype", ["odd", "even", "constant", None])
@pytest.mark.parametrize("ba", ["order2", "order1", "short_b_a0"])
def test_filtfilt_one_section(oracle_mod, io, padtype, ba):
    """filtfilt of orders 1 and 2 runs as the one-section kernel form:
    against the oracle's filtfilt (lfilter passes), with a[0] != 1 and b
    shorter than a."""
    import scipy.signal as ss
    from torchscience_amd.signal_processing.filter import filtfilt
    dev = _dev()
    b, a = {"order2": ss.butter(2, 0.15), "order1": ss.butter(1, 0.3),
            "short_b_a0": (np.array([0.5]), np.array([2.0, -1.2, 0.4]))}[ba]
    x = np.random.default_rng(4).standard_normal((48, 3000))
    bt, at = torch.tensor(b, device=dev).to(io), torch.tensor(a, device=dev).to(io)
    xt = torch.tensor(x, device=dev).to(io)
    y = filtfilt(bt, at, xt, padtype=padtype)
    assert y.shape == xt.shape and y.dtype == io
    ref = oracle_mod.filtfilt(bt.double().cpu().numpy(), at.double().cpu().numpy(),
                              xt.double().cpu().numpy(), padtype=padtype)
    err = relerr(y.double().cpu().numpy(), ref)
    print(f"filtfilt {ba} io={io} padtype={padtype}: {err:.3e}", flush=True)
    assert err < TOL[io]


@pytest.mark.parametrize("io", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [3, 8, 16, 20])
@pytest.mark.parametrize("T", [300, 5000, 5001])
def test_general_order_reverse_equals_flip(io, order, T):
    """The general-order kernel's reverse-time mode (serial form for short
    rows and order 20, chunked time-parallel form for 5000 samples up to
    order 16) equals the forward mode on flipped rows bit for bit, zi and zf
    included.  5000 samples: whole 16-byte vectors both ways; 5001: the
    reverse mode and the odd rows of the forward one take element
    accesses."""
    import scipy.signal as ss
    from torchscience_amd.signal_processing.filter._lfilter import _general_rows
    dev = _dev()
    b, a = ss.butter(order, 0.3) if order <= 16 else (np.r_[1.0, np.zeros(order - 1), 0.5],
                                                      np.r_[1.0, np.zeros(order - 1), -0.3])
    coef = torch.tensor(np.concatenate([b, a]), dtype=io, device=dev)
    rng = np.random.default_rng(order + T)
    x = torch.tensor(rng.standard_normal((37, T)), dtype=io, device=dev)
    zi = torch.tensor(rng.standard_normal((37, order)), dtype=io, device=dev)
    y, zf = _general_rows(coef, x, zi, True, False, reverse=True)
    yf, zff = _general_rows(coef, x.flip(-1).contiguous(), zi, True, False)
    assert torch.equal(y, yf.flip(-1)) and torch.equal(zf, zff)


@pytest.mark.parametrize("io", [torch.float64, torch.float32])
@pytest.mark.parametrize("padtype", ["odd", None])
@pytest.mark.parametrize("order", [4, 8])
def test_filtfilt_general_order(oracle_mod, io, padtype, order):
    """filtfilt of orders 4 and 8 (the general-order kernel, backward passes
    in its reverse-time mode) against the oracle's filtfilt: FP64 1e-10,
    FP32 (DF2T of the full polynomial, as the reference computes it) 1e-4
    for order 4 and 2e-3 for order 8 (measured 3.8e-6 and 2.9e-4)."""
    import scipy.signal as ss
    from torchscience_amd.signal_processing.filter import filtfilt
    dev = _dev()
    b, a = ss.butter(order, 0.2)
    x = np.random.default_rng(order).standard_normal((64, 6000))
    bt, at = torch.tensor(b, device=dev).to(io), torch.tensor(a, device=dev).to(io)
    xt = torch.tensor(x, device=dev).to(io)
    y = filtfilt(bt, at, xt, padtype=padtype)
    assert y.shape == xt.shape and y.dtype == io
    ref = oracle_mod.filtfilt(bt.double().cpu().numpy(), at.double().cpu().numpy(),
                              xt.double().cpu().numpy(), padtype=padtype)
    err = relerr(y.double().cpu().numpy(), ref)
    print(f"filtfilt order {order} io={io} padtype={padtype}: {err:.3e}", flush=True)
    tol = 1e-10 if io == torch.float64 else (1e-4 if order == 4 else 2e-3)
    assert err < tol


def _lowpass5():
    # the third section is first order (b2 == a2 == 0): default padlen 18, not 21
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    return butterworth_sos(5, 0.2)


@pytest.mark.parametrize("io", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("padtype,padlen", [("odd", None), ("even", None), ("odd", 7),
                                            (None, None), ("constant", 0)])
def test_device_steady_state_equals_host_route(io, padtype, padlen):
    """Coefficients on the device: the operator forms sosfilt_zi and the
    default padlen itself (no host copy); with them on the host, the host
    forms both.  The outputs agree bit for bit."""
    from torchscience_amd.signal_processing.filter import sosfiltfilt
    print(f"device steady state io={io} padtype={padtype} padlen={padlen}", flush=True)
    dev = _dev()
    sos = torch.tensor(_lowpass5()).to(io)
    x = torch.tensor(np.random.default_rng(11).standard_normal((37, 3001))).to(dev, io)
    y_dev = sosfiltfilt(sos.to(dev), x, padtype=padtype, padlen=padlen)
    y_host = sosfiltfilt(sos, x, padtype=padtype, padlen=padlen)
    assert torch.equal(y_dev, y_host)


def test_operator_default_padlen_and_steady_state():
    """tsiir::sosfiltfilt with zi_unit None and padlen -1 against the same
    operator given sosfilt_zi_host and the reference's default padlen (18
    for a filter with one first-order section; 3 * (2 S + 1) = 21 without
    the b2 / a2 zeros): equal bit for bit, and different from padlen 21."""
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter._zi import sosfilt_zi_host
    dev = _dev()
    _lib.ops()
    sos = torch.tensor(_lowpass5(), dtype=torch.float32)
    zu = sosfilt_zi_host(sos).reshape(1, 3, 2).to(dev)
    sos_d = sos.reshape(1, 3, 6).to(dev)
    x = torch.randn(64, 5000, device=dev)
    y_dev = torch.ops.tsiir.sosfiltfilt(x, sos_d, None, _lib.PAD_ODD, _lib.PADLEN_DEFAULT)
    y18 = torch.ops.tsiir.sosfiltfilt(x, sos_d, zu, _lib.PAD_ODD, 18)
    y21 = torch.ops.tsiir.sosfiltfilt(x, sos_d, zu, _lib.PAD_ODD, 21)
    assert torch.equal(y_dev, y18)
    assert not torch.equal(y_dev, y21)


@pytest.mark.parametrize("row", [[1.0, 0.0, 0.0, 1.0, -1.0, 0.0],    # pole at z = 1
                                 [1.0, 0.5, 0.0, 0.0, 1.0, -1.0]])   # a0 == 0, sum(a) == 0
def test_no_steady_state_reports(row):
    """A filter without a steady state (the reference's sosfilt_zi raises:
    its solve meets a singular system) on the device route: NaN outputs,
    then TsiirError from check_errors; the stream's next call runs
    normally.  With host coefficients the reference's immediate error
    stands."""
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter import sosfiltfilt
    print(f"no steady state: {row}", flush=True)
    dev = _dev()
    sos = torch.tensor([row], dtype=torch.float64)
    x = torch.randn(8, 1000, device=dev, dtype=torch.float64)
    y = sosfiltfilt(sos.to(dev), x)
    with pytest.raises(_lib.TsiirError, match="steady state"):
        _lib.check_errors(dev)
    assert torch.isnan(y).all()
    good = torch.tensor(_lowpass5(), dtype=torch.float64)
    assert torch.equal(sosfiltfilt(good.to(dev), x), sosfiltfilt(good, x))
    with pytest.raises(RuntimeError):
        sosfiltfilt(sos, x)


def test_leading_zero_a0_is_not_an_error():
    """a0 == 0 with sum(a) != 0: the reference's sosfilt_zi strips the zero
    and solves, and its passes divide by a0; no error on either route and
    no finite output."""
    from torchscience_amd import _lib
    from torchscience_amd.signal_processing.filter import sosfiltfilt
    dev = _dev()
    sos = torch.tensor([[1.0, 0.5, 0.0, 0.0, 1.0, 0.2]], dtype=torch.float64)
    x = torch.randn(8, 1000, device=dev, dtype=torch.float64)
    y_dev = sosfiltfilt(sos.to(dev), x)
    y_host = sosfiltfilt(sos, x)
    _lib.check_errors(dev)
    assert not torch.isfinite(y_dev).any() and not torch.isfinite(y_host).any()


def test_no_steady_state_next_call_raises():
    """Without check_errors, the stream's next call reports it."""
    from torchscience_amd.signal_processing.filter import sosfiltfilt
    dev = _dev()
    x = torch.randn(4, 500, device=dev)
    bad = torch.tensor([[1.0, 0.0, 0.0, 1.0, -1.0, 0.0]], device=dev)
    sosfiltfilt(bad, x)
    torch.cuda.synchronize(dev)
    with pytest.raises(RuntimeError, match="steady state"):
        sosfiltfilt(torch.tensor(_lowpass5(), dtype=torch.float32, device=dev), x)
    y = sosfiltfilt(torch.tensor(_lowpass5(), dtype=torch.float32, device=dev), x)
    assert torch.isfinite(y).all()


@pytest.mark.parametrize("io", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [4, 8, 16])
def test_general_order_unaligned_rows_equal_aligned(io, order):
    """The chunked general-order form reads and writes rows in 16-byte
    vectors through an LDS tile when they are aligned, element by element
    otherwise: rows of 6001 samples, every other one starting off a 16-byte
    boundary, give the same bits as each row alone in an aligned buffer, zf
    included."""
    import scipy.signal as ss
    from torchscience_amd.signal_processing.filter._lfilter import _general_rows
    dev = _dev()
    b, a = ss.butter(order, 0.2)
    coef = torch.tensor(np.concatenate([b, a]), dtype=io, device=dev)
    rng = np.random.default_rng(order)
    x = torch.tensor(rng.standard_normal((19, 6001)), dtype=io, device=dev)
    zi = torch.tensor(rng.standard_normal((19, order)), dtype=io, device=dev)
    y, zf = _general_rows(coef, x, zi, True, False)
    for c in (0, 1, 2, 3, 18):
        yc, zfc = _general_rows(coef, x[c:c + 1].contiguous(), zi[c:c + 1].contiguous(), True,
                                False)
        assert torch.equal(y[c], yc[0]) and torch.equal(zf[c], zfc[0]), c
