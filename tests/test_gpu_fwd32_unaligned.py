"""FP32-recurrence forward kernel on windows that do not start on a 16-byte
boundary (GPU).

A window r elements past a 16-byte boundary is staged as the aligned run r
elements earlier and moved by r in registers; its own elements are the only
ones written.  Rows of a contiguous (B, T) tensor with T not a multiple of the
16-byte vector (4 FP32, 8 BF16 elements) start at every such offset r in
turn, as do their ends (the windows of the reverse-time pass are counted from
the end).

Checks: the full 4096-sample windows equal, bit for bit, those of the same
rows placed at aligned addresses (zero padding after the row for forward
time, before it for reverse time: the windows hold the same samples); the
partial window, which runs in the FP64-carry kernel, and zf against the FP64
oracle (FP32 1e-5, BF16 1e-2 normwise); an unaligned input view with an
aligned output equal in full, zf included, to a contiguous copy; many
channels (chained windows) and few (look-back over the groups before); the
saved section outputs of the training forward at unaligned rows; and that
nothing outside a row is written (sentinels around output views).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

WIN = 4096


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


def _sos():
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    return butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)


def _bits(t):
    t = t.contiguous()
    return t.view({torch.bfloat16: torch.int16, torch.float64: torch.int64}.get(t.dtype,
                                                                            torch.int32))


VEC = {torch.float32: 4, torch.bfloat16: 8, torch.float64: 2}
TOL = {torch.float32: 1e-5, torch.bfloat16: 1e-2, torch.float64: 1e-12}


@pytest.mark.parametrize("io", [torch.float32, torch.bfloat16, torch.float64])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("B", [1024, 8])
@pytest.mark.parametrize("tail", [5, 1001])
def test_odd_length_rows_match_aligned(oracle_mod, io, reverse, B, tail):
    """T = 3 * 4096 + tail: row c starts (forward) or ends (reverse) c * tail
    mod (vector) elements past a 16-byte boundary.  The partial window runs
    in the FP64-carry kernel, whose whole 64-sample runs (a 1001-sample
    tail has 15) move as aligned vectors; FP64 rows take the FP64 kernels
    throughout."""
    dev = _dev()
    T = 3 * WIN + tail
    pad = (-T) % VEC[io]
    print(f"odd rows io={io} reverse={reverse} B={B} T={T}", flush=True)
    fdt = torch.float64 if io == torch.float64 else torch.float32
    rng = np.random.default_rng(B + 2 * reverse)
    sos = _sos()
    st = torch.tensor(sos[None], device=dev)
    st = st.to(fdt)
    x = torch.tensor(rng.standard_normal((B, T)), dtype=fdt, device=dev).to(io)
    zi = torch.tensor(rng.standard_normal((B, sos.shape[0], 2)), dtype=fdt, device=dev)
    kw = dict(channels=B, chans_per_filter=B, want_zf=True, reverse=reverse)
    y, zf = _fwd(st, x, zi, **kw)
    # the same rows at aligned addresses: the padding follows a forward row
    # (and does not reach its full windows) or precedes a reverse row
    xa = torch.zeros(B, T + pad, dtype=io, device=dev)
    lo = pad if reverse else 0
    xa[:, lo:lo + T] = x
    ya, _ = _fwd(st, xa, zi, **kw)
    ya = ya[:, lo:lo + T]
    # every output before T (after it, in reverse time) is computed from the
    # same samples and states either way: equal bit for bit, the partial
    # window included
    assert torch.equal(_bits(y), _bits(ya))
    x64 = x.double().cpu().numpy()
    zi64 = zi.double().cpu().numpy()
    if reverse:
        ref, zf_ref = oracle_mod.sosfilt_rows(sos.astype(np.float64), x64[:, ::-1], zi64,
                                              n_threads=8)
        got = y.double().cpu().numpy()[:, ::-1]
    else:
        ref, zf_ref = oracle_mod.sosfilt_rows(sos.astype(np.float64), x64, zi64, n_threads=8)
        got = y.double().cpu().numpy()
    err = relerr(got, ref)
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < TOL[io]
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(zf.double().cpu().numpy() - zf_ref)) < TOL[io] * scale


@pytest.mark.parametrize("io", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("reverse", [False, True])
def test_unaligned_input_view_matches_contiguous(io, reverse):
    """x a view at column 1 of a (B, T + 3) buffer (the row stride is odd:
    every offset r occurs), y contiguous and aligned (T a multiple of 8): the
    whole result, zf included, equals that of a contiguous copy of x."""
    dev = _dev()
    B, T = 1024, 2 * WIN + 8
    print(f"unaligned input view io={io} reverse={reverse}", flush=True)
    rng = np.random.default_rng(11 + reverse)
    st = torch.tensor(_sos()[None], device=dev)
    buf = torch.tensor(rng.standard_normal((B, T + 3)), dtype=torch.float32,
                       device=dev).to(io)
    xv = buf[:, 1:T + 1]
    zi = torch.tensor(rng.standard_normal((B, 6, 2)), dtype=torch.float32, device=dev)
    kw = dict(channels=B, chans_per_filter=B, want_zf=True, reverse=reverse)
    y, zf = _fwd(st, xv, zi, **kw)
    yc, zfc = _fwd(st, xv.contiguous(), zi, **kw)
    assert torch.equal(_bits(y), _bits(yc))
    assert torch.equal(zf, zfc)


@pytest.mark.parametrize("T", [2 * WIN + 1, 2 * WIN + 2, 2 * WIN + 3])
def test_saved_sections_odd_rows(oracle_mod, T):
    """The training forward (saved outputs of sections 0..S-2, forward time
    only) at every row offset: the kept rows and y against the oracle."""
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    B = 1024
    print(f"saved sections odd rows T={T}", flush=True)
    rng = np.random.default_rng(T)
    sos = butterworth_sos(8, 0.2).astype(np.float32)
    x = rng.standard_normal((B, T)).astype(np.float32)
    y, _, saved = _fwd(torch.tensor(sos[None], device=dev), torch.tensor(x, device=dev), None,
                       channels=B, chans_per_filter=B, want_zf=False, keep_sections=True)
    x64 = x.astype(np.float64)
    for s in range(sos.shape[0] - 1):
        ys, _ = oracle_mod.sosfilt_rows(sos[: s + 1].astype(np.float64), x64, n_threads=8)
        assert relerr(saved[s].cpu().numpy(), ys) < 1e-5, s
    y64, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64), x64, n_threads=8)
    assert relerr(y.cpu().numpy(), y64) < 1e-5


def _abi_forward(sos, x, y, y_row_stride, flags):
    """One C-ABI forward over FP32/BF16 rows x (B, T), contiguous, shared
    FP32 filter, into y (a pointer, any element offset) with the given output
    row stride."""
    import ctypes
    from torchscience_amd import _lib
    L = _lib.load()
    B, T = x.shape
    io = _lib.BF16 if x.dtype == torch.bfloat16 else _lib.F32
    d = _lib.SosfiltDesc(channels=B, n_samples=T, n_sections=sos.shape[-2], io_dtype=io,
                         coef_dtype=_lib.F32, mode=0, flags=flags, reserved=0, x_rows=B,
                         x_row_stride=T, y_row_stride=y_row_stride, chans_per_filter=B)
    n = L.tsiir_sosfilt_workspace_bytes(ctypes.byref(d))
    ws = torch.zeros(max(n, 1), dtype=torch.uint8, device=x.device)
    _lib.check(L.tsiir_sosfilt_forward(ctypes.byref(d), _lib.ptr(x), y, _lib.ptr(sos), None,
                                       None, _lib.ptr(ws), ws.numel(),
                                       _lib.stream_handle(x.device)), "tsiir_sosfilt_forward")
    torch.cuda.synchronize(x.device)


@pytest.mark.parametrize("io", [torch.float32, torch.bfloat16])
@pytest.mark.parametrize("reverse", [False, True])
@pytest.mark.parametrize("B", [1024, 8])
def test_writes_stay_inside_rows(io, reverse, B):
    """Output rows of T = 2 * 4096 + 5 samples written at column 1 of a
    (B, T + 3) buffer of sentinels (odd row stride: every offset r occurs):
    every element outside the rows keeps its sentinel, and the rows equal
    bit for bit the contiguous output of the same call."""
    from torchscience_amd import _lib
    dev = _dev()
    T = 2 * WIN + 5
    print(f"rows bounded io={io} reverse={reverse} B={B}", flush=True)
    rng = np.random.default_rng(17 + B)
    sos = torch.tensor(_sos()[None], device=dev)
    x = torch.tensor(rng.standard_normal((B, T)), dtype=torch.float32, device=dev).to(io)
    flags = _lib.REVERSE if reverse else 0
    y = torch.empty_like(x)
    _abi_forward(sos, x, _lib.ptr(y), T, flags)
    bits = torch.int16 if io == torch.bfloat16 else torch.int32
    sentinel = -12345
    buf = torch.full((B, T + 3), sentinel, dtype=bits, device=dev)
    _abi_forward(sos, x, buf.data_ptr() + buf.element_size(), T + 3, flags)
    assert torch.equal(buf[:, 1:T + 1], _bits(y))
    assert bool((buf[:, 0] == sentinel).all())
    assert bool((buf[:, T + 1:] == sentinel).all())
