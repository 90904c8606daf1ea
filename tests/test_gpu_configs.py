"""The benchmark configurations at their full sizes, through the public
drop-in calls, against the FP64 oracle (GPU).

Per configuration (sizes as in the benchmark):
  C2a  sosfilt, x [4096, 65536] FP32, shared 6-section order-12 Butterworth
       bandpass [0.05, 0.3]: 64 sampled channels <= 1e-5 normwise; scaling
       the input by 2 scales the output by 2 bitwise and reversing the
       channel order reverses the output rows bitwise (whole tensor);
  C2b  batched_filter_apply(broadcast=False), per-channel order-8
       Butterworth lowpass, cutoff U[0.02, 0.45]: 64 sampled channels
       <= 1e-5, the scaling property bitwise;
  C2c  sosfilt, x [8, 2^24] (the long-sequence look-back path): all 8
       channels <= 1e-5;
  C5   order-16 Chebyshev I (1 dB, cutoff 0.01, poles at r = 0.99972),
       x [2048, 2^20], 16 sampled channels, each precision form against its
       stated bound: fp32c (FP64 updates) <= 1e-5; fp32 (the DF2T form with
       FP64 carries that such poles are routed to) and bf16io no worse than
       the reference's own FP32 arithmetic (its DF2T loop restated in
       float32, oracle.sosfilt_rows_f32) on the same channels; bf16c (FP64
       updates, bfloat16 I/O) <= 3e-3, the bfloat16 output rounding (2^-9
       of the row maximum) plus margin;
  C4   10,000 Poisson requests over 2,000 streams through the serving
       engine: 16 sampled streams' concatenated outputs against one-shot
       FP64 filtering <= 1e-5; 4 streams' requests served alone equal their
       results in the mix bitwise.

Errors are per-channel normwise max|y - y64| / max|y64| (worst channel)
against the FP64 oracle on the coefficient values the kernel consumed.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _relerr(y, y64):
    y = np.asarray(y, dtype=np.float64)
    num = np.max(np.abs(y - y64), axis=-1)
    den = np.maximum(np.max(np.abs(y64), axis=-1), 1e-300)
    return float(np.max(num / den))


def _randn(B, T, seed, dev):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    return torch.randn(B, T, device=dev, generator=g, dtype=torch.float32)


def _sample(B, n):
    return np.unique(np.linspace(0, B - 1, n).astype(np.int64))


def test_config_c2a_full_size(oracle_mod):
    from torchscience_amd.signal_processing.filter import sosfilt
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    print("C2a: sosfilt fp32 x[4096,65536], 6-section bandpass", flush=True)
    x = _randn(4096, 65536, 1, dev)
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    st = torch.tensor(sos, device=dev)
    y = sosfilt(st, x)
    idx = _sample(4096, 64)
    ref, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64),
                                     x[idx].cpu().numpy().astype(np.float64), n_threads=16)
    err = _relerr(y[idx].cpu().numpy(), ref)
    print("C2a max rel err", err, flush=True)
    assert err < TOL
    # linearity under power-of-two scaling: every operation scales exactly
    assert torch.equal(sosfilt(st, x * 2.0), y * 2.0)
    # channels are independent: reversing the rows reverses the output
    assert torch.equal(sosfilt(st, x.flip(0)), y.flip(0))


def test_config_c2b_full_size(oracle_mod):
    from torchscience_amd.signal_processing.filter import batched_filter_apply
    from torchscience_amd.signal_processing.filter._design import batched_butterworth_sos
    dev = _dev()
    print("C2b: batched_filter_apply per-channel order-8 lowpass x[4096,65536]", flush=True)
    x = _randn(4096, 65536, 1, dev)
    rng = np.random.default_rng(2)
    sos = batched_butterworth_sos(8, rng.uniform(0.02, 0.45, 4096)).astype(np.float32)
    st = torch.tensor(sos, device=dev)
    y = batched_filter_apply(st, x, broadcast=False)
    idx = _sample(4096, 64)
    ref, _ = oracle_mod.sosfilt_rows(sos[idx].astype(np.float64),
                                     x[idx].cpu().numpy().astype(np.float64),
                                     chans_per_filter=1, n_threads=16)
    err = _relerr(y[idx].cpu().numpy(), ref)
    print("C2b max rel err", err, flush=True)
    assert err < TOL
    assert torch.equal(batched_filter_apply(st, x * 2.0, broadcast=False), y * 2.0)


def test_config_c2c_full_size(oracle_mod):
    from torchscience_amd.signal_processing.filter import sosfilt
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    print("C2c: sosfilt fp32 x[8,2^24], 6-section bandpass", flush=True)
    x = _randn(8, 1 << 24, 3, dev)
    sos = butterworth_sos(6, [0.05, 0.3], "bandpass").astype(np.float32)
    y = sosfilt(torch.tensor(sos, device=dev), x)
    ref, _ = oracle_mod.sosfilt_rows(sos.astype(np.float64),
                                     x.cpu().numpy().astype(np.float64), n_threads=16)
    err = _relerr(y.cpu().numpy(), ref)
    print("C2c max rel err (all 8 channels)", err, flush=True)
    assert err < TOL


@pytest.fixture(scope="module")
def c5_input():
    dev = _dev()
    return _randn(2048, 1 << 20, 5, dev)


@pytest.mark.parametrize("form", ["fp32c", "fp32", "bf16io", "bf16c"])
def test_config_c5_full_size(oracle_mod, c5_input, form):
    from torchscience_amd.signal_processing.filter import sosfilt
    from torchscience_amd.signal_processing.filter._design import chebyshev1_sos
    dev = c5_input.device
    print(f"C5: order-16 Chebyshev I x[2048,2^20], form {form}", flush=True)
    sos64 = chebyshev1_sos(16, 0.01, 1.0)
    comp = form in ("fp32c", "bf16c")
    sos = sos64 if comp else sos64.astype(np.float32).astype(np.float64)
    st = torch.tensor(sos64 if comp else sos64.astype(np.float32), device=dev)
    x = c5_input.bfloat16() if form.startswith("bf16") else c5_input
    y = sosfilt(st, x, precision=form)
    assert y.shape == x.shape and y.dtype == x.dtype
    idx = _sample(2048, 16)
    xs = x[idx].float().cpu().numpy()
    ref, _ = oracle_mod.sosfilt_rows(sos, xs.astype(np.float64), n_threads=16)
    err = _relerr(y[idx].float().cpu().numpy(), ref)
    ref32, _ = oracle_mod.sosfilt_rows_f32(sos64.astype(np.float32), xs, n_threads=16)
    ref_fp32_err = _relerr(ref32, ref)
    print(f"C5 {form}: max rel err {err:.3e}; the reference's FP32 DF2T "
          f"{ref_fp32_err:.3e}", flush=True)
    if form == "fp32c":
        assert err < TOL
    elif form == "bf16c":
        assert err < 3e-3
    else:
        assert err <= ref_fp32_err


def test_config_c4_full_size(oracle_mod):
    """The serving workload at size: 10,000 requests, 2,000 streams."""
    from torchscience_amd.serving import HipExecutor, Request, ServingEngine, StreamTable
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    print("C4: 10000 requests over 2000 streams", flush=True)
    rng = np.random.default_rng(4)
    n_req, n_str = 10000, 2000
    sos, chans = {}, {}
    for sid in range(n_str):
        sec = int(rng.integers(1, 9))
        sos[sid] = butterworth_sos(2 * sec, float(rng.uniform(0.02, 0.45)))
        chans[sid] = int(rng.integers(1, 65))
    sids = rng.integers(0, n_str, n_req)
    lens = np.exp(rng.uniform(np.log(256), np.log(65536), n_req)).astype(np.int64)
    arrivals = np.cumsum(rng.exponential(1.0 / 20000.0, n_req))
    sizes = np.array([chans[int(s)] * int(T) for s, T in zip(sids, lens)], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    g = torch.Generator(device=dev)
    g.manual_seed(4)
    xbuf = torch.randn(int(offs[-1]), device=dev, generator=g)

    def requests(only=None):
        out = []
        for i in range(n_req):
            s = int(sids[i])
            if only is not None and s not in only:
                continue
            x = xbuf[offs[i]:offs[i + 1]].view(chans[s], int(lens[i]))
            out.append(Request(i, s, x, float(arrivals[i])))
        return out

    def engine(streams):
        table = StreamTable(dev)
        for s in streams:
            table.register(s, sos[s], chans[s])
        return ServingEngine(table, HipExecutor(dev, 4, precision="fp32c"),
                             hbm_budget_bytes=32 << 30, max_batch_channels=8192,
                             max_inflight=4)

    reqs = requests()
    stats = engine(range(n_str)).run(reqs)
    torch.cuda.synchronize()
    print(f"C4 served {stats.n_requests} requests in {stats.batches} batches, p50 "
          f"{stats.p50 * 1e3:.3f} ms p99 {stats.p99 * 1e3:.3f} ms", flush=True)
    assert all(r.t_done >= 0 for r in reqs)
    by_stream = {}
    for r in reqs:
        by_stream.setdefault(r.stream_id, []).append(r)
    check = [int(s) for s in rng.choice(sorted(by_stream), 16, replace=False)]
    worst = 0.0
    for s in check:
        rs = sorted(by_stream[s], key=lambda r: (r.arrival, r.rid))
        xc = np.concatenate([r.x.cpu().numpy() for r in rs], axis=1).astype(np.float64)
        yc = np.concatenate([r.y.cpu().numpy() for r in rs], axis=1).astype(np.float64)
        ref, _ = oracle_mod.sosfilt_rows(sos[s], xc, n_threads=16)
        worst = max(worst, _relerr(yc, ref))
    print("C4 16 streams vs one-shot FP64: max rel err", worst, flush=True)
    assert worst < TOL
    alone = set(check[:4])
    solo = requests(alone)
    engine(sorted(alone)).run(solo)
    torch.cuda.synchronize()
    mixed = {r.rid: r for r in reqs}
    for r in solo:
        assert torch.equal(r.y, mixed[r.rid].y), f"request {r.rid} differs alone vs in mix"
    print(f"C4 {len(solo)} requests of 4 streams: alone == in mix (bitwise)", flush=True)
