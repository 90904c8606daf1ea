"""The persistent serving launch (tsiir_serve_create / submit / poll /
destroy) on the GPU: items served through the queue equal the ragged-batch
launch bitwise (outputs and states), also across several submits and after
the launch has drained for lack of work and is launched again; the engine
over the queue matches one-shot FP64 filtering and gives the same result
alone and in a mix.  Every wait here is bounded."""

import ctypes
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _items(reqs, table):
    from torchscience_amd.serving.engine import ITEM_DTYPE
    rows = []
    for r in reqs:
        S = table.sections(r.stream_id)
        for c in range(r.channels):
            rows.append((r.x.data_ptr() + 4 * c * r.samples, r.y.data_ptr() + 4 * c * r.samples,
                         table.table(r.stream_id).data_ptr(),
                         table.state(r.stream_id).data_ptr() + 16 * S * c, r.samples, S, 0))
    return np.array(rows, dtype=ITEM_DTYPE)


def _wait(L, q, first, n, limit_s=20.0):
    flag = ctypes.c_int32()
    t0 = time.perf_counter()
    while True:
        assert L.tsiir_serve_poll(q, first, n, ctypes.byref(flag)) == 0
        if flag.value:
            return
        assert time.perf_counter() - t0 < limit_s, "queue items not done in time"
        time.sleep(1e-4)


@pytest.mark.parametrize("mode", [0, 1])
def test_queue_equals_batch_launch(mode):
    from torchscience_amd import _lib
    from torchscience_amd.serving import Request, StreamTable
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    L = _lib.load()
    rng = np.random.default_rng(5 + mode)
    table_a, table_b = StreamTable(dev), StreamTable(dev)
    sos = {s: butterworth_sos(2 * int(rng.integers(1, 9)), float(rng.uniform(0.02, 0.45)))
           for s in range(12)}
    chans = {s: int(rng.integers(1, 9)) for s in sos}
    for s in sos:
        zi = rng.standard_normal((chans[s], sos[s].shape[0], 2))
        table_a.register(s, sos[s], chans[s], zi=zi)
        table_b.register(s, sos[s], chans[s], zi=zi)
    table_a.ready()
    # three rounds of chunks (one per stream per round), lengths 1 .. 40000
    rounds = [[Request(100 * k + s, s, torch.from_numpy(
        rng.standard_normal((chans[s], int(rng.integers(1, 40000)))).astype(np.float32)).to(dev))
        for s in sos] for k in range(3)]
    torch.cuda.synchronize()
    q = ctypes.c_void_p()
    assert L.tsiir_serve_create(4096, 256, mode, 2000, ctypes.byref(q)) == 0
    try:
        for k, rnd in enumerate(rounds):
            print("queue round", k, "mode", mode, flush=True)
            ya = [torch.empty_like(r.x) for r in rnd]
            yb = [torch.empty_like(r.x) for r in rnd]
            for r, y in zip(rnd, ya):
                r.y = y
            items = _items(rnd, table_a)
            dev_items = torch.from_numpy(items.view(np.uint8)).to(dev)
            assert L.tsiir_serve_batch(dev_items.data_ptr(), len(items), mode,
                                       torch.cuda.current_stream(dev).cuda_stream) == 0
            torch.cuda.synchronize()
            for r, y in zip(rnd, yb):
                r.y = y
            items = _items(rnd, table_b)
            half = len(items) // 2
            firsts = []
            for part in (np.ascontiguousarray(items[:half]), np.ascontiguousarray(items[half:])):
                f = ctypes.c_uint64()
                assert L.tsiir_serve_submit(q, part.ctypes.data, len(part), ctypes.byref(f)) == 0
                firsts.append(f.value)
            assert firsts[1] == firsts[0] + half
            _wait(L, q, firsts[0], len(items))
            torch.cuda.synchronize()
            for a, b in zip(ya, yb):
                assert torch.equal(a, b)
            for s in sos:
                assert torch.equal(table_a.state(s), table_b.state(s))
            if k == 1:
                time.sleep(0.05)  # past the idle time: the launch drains, the next submit relaunches
    finally:
        assert L.tsiir_serve_destroy(q) == 0


def test_queue_rejects_more_items_than_slots():
    from torchscience_amd import _lib
    _dev()
    L = _lib.load()
    q = ctypes.c_void_p()
    assert L.tsiir_serve_create(4, 16, 1, 1000, ctypes.byref(q)) == 0
    try:
        from torchscience_amd.serving.engine import ITEM_DTYPE
        items = np.zeros(5, dtype=ITEM_DTYPE)
        f = ctypes.c_uint64()
        # rejected before anything is written or launched
        assert L.tsiir_serve_submit(q, items.ctypes.data, 5, ctypes.byref(f)) == _lib.EAGAIN
    finally:
        assert L.tsiir_serve_destroy(q) == 0


def test_persistent_engine_matches_oneshot_and_alone(oracle_mod):
    from torchscience_amd.serving import PersistentExecutor, Request, ServingEngine, StreamTable
    from torchscience_amd.signal_processing.filter._design import butterworth_sos
    dev = _dev()
    rng = np.random.default_rng(31)
    sos = {s: butterworth_sos(2 * int(rng.integers(1, 9)), float(rng.uniform(0.02, 0.45)))
           for s in range(20)}
    chans = {s: int(rng.integers(1, 17)) for s in sos}
    table = StreamTable(dev)
    for s in sos:
        table.register(s, sos[s], chans[s])
    reqs = []
    t = np.cumsum(rng.exponential(1 / 5000.0, 150))
    for i in range(150):
        s = int(rng.integers(0, 20))
        T = int(np.exp(rng.uniform(np.log(40), np.log(30000))))
        x = torch.from_numpy(rng.standard_normal((chans[s], T)).astype(np.float32)).to(dev)
        reqs.append(Request(i, s, x, float(t[i])))
    ex = PersistentExecutor(dev, 4, workgroups=512, idle_us=5000)
    try:
        stats = ServingEngine(table, ex, long_samples=8192).run(reqs)
    finally:
        ex.close()
    torch.cuda.synchronize()
    assert all(r.status == "done" for r in reqs)
    print("persistent engine: batches", stats.batches, "p99 ms", stats.p99 * 1e3)
    worst = 0.0
    for s in sos:
        rs = sorted([r for r in reqs if r.stream_id == s], key=lambda r: (r.arrival, r.rid))
        if not rs:
            continue
        xc = np.concatenate([r.x.cpu().numpy() for r in rs], axis=1).astype(np.float64)
        yc = np.concatenate([r.y.cpu().numpy() for r in rs], axis=1).astype(np.float64)
        ref, _ = oracle_mod.sosfilt_rows(sos[s], xc)
        worst = max(worst, float(np.max(np.max(np.abs(yc - ref), axis=1)
                                        / np.max(np.abs(ref), axis=1))))
    assert worst < 1e-5
    # one stream's requests alone, from zero state, equal their results in the mix
    s0 = reqs[0].stream_id
    table2 = StreamTable(dev)
    table2.register(s0, sos[s0], chans[s0])
    solo = [Request(r.rid, r.stream_id, r.x, r.arrival) for r in reqs if r.stream_id == s0]
    ex2 = PersistentExecutor(dev, 4, workgroups=64, idle_us=5000)
    try:
        ServingEngine(table2, ex2, long_samples=8192).run(solo)
    finally:
        ex2.close()
    torch.cuda.synchronize()
    mixed = {r.rid: r for r in reqs}
    for r in solo:
        assert torch.equal(r.y, mixed[r.rid].y)
