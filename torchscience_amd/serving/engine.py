"""Continuous-batching server for streaming sosfilt.

Streams are persistent: each has a filter (1..32 sections) and a channel
count; its filter state (the zi/zf of the reference's sosfilt,
src/torchscience/signal_processing/filter/_sosfilt.py:98-133) stays resident on
the device between chunks, in FP64.  Requests are chunks (channels x samples)
of a stream.  The engine admits arrived requests while the HBM budget allows,
groups ready requests of distinct, idle streams into one batch (one kernel
launch over all their channels, `tsiir_serve_batch`), and keeps up to
`max_inflight` batches running on separate HIP streams, so requests that
arrive while others are on the device are launched as soon as slots, budget
and their stream free up, and their work overlaps on the device.

Ordering: a stream's chunks run one at a time, in arrival order (its state
threads through them).  Each batch item's result depends only on its own
chunk and state, never on what else is in the batch.

The executor interface (`launch`, `done`) separates the scheduler from the
device; `HipExecutor` runs the HIP kernel, and tests drive the scheduler with
a CPU executor.
"""

from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Protocol

import numpy as np
import torch

from .. import _lib


@dataclass
class Request:
    """One chunk of a stream: x is (channels, samples) float32 on the device."""

    rid: int
    stream_id: int
    x: torch.Tensor
    arrival: float = 0.0          # seconds after the start of the run
    y: torch.Tensor | None = None
    t_launch: float = -1.0
    t_done: float = -1.0

    @property
    def channels(self) -> int:
        return int(self.x.shape[0])

    @property
    def samples(self) -> int:
        return int(self.x.shape[1])

    @property
    def nbytes(self) -> int:
        return 2 * self.x.numel() * 4  # input + output

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


# numpy view of tsiir_serve_item (include/tsiir.h)
ITEM_DTYPE = np.dtype([("x", np.uint64), ("y", np.uint64), ("table", np.uint64),
                       ("state", np.uint64), ("n_samples", np.int64), ("n_sections", np.int32),
                       ("reserved", np.int32)])
assert ITEM_DTYPE.itemsize == ctypes.sizeof(_lib.ServeItem) == 48


class StreamTable:
    """Device-resident per-stream filter tables and FP64 states."""

    def __init__(self, device: torch.device):
        self.device = device
        self._table: dict[int, torch.Tensor] = {}
        self._state: dict[int, torch.Tensor] = {}
        self._sections: dict[int, int] = {}
        self._channels: dict[int, int] = {}

    def register(self, stream_id: int, sos, channels: int, zi=None) -> None:
        L = _lib.load()
        sos_t = torch.as_tensor(sos, dtype=torch.float64).to(self.device).contiguous()
        S = int(sos_t.shape[0])
        nbytes = L.tsiir_serve_table_bytes(S)
        if sos_t.ndim != 2 or sos_t.shape[1] != 6 or nbytes == 0:
            raise ValueError("stream filters must be (n_sections, 6) with 1..32 sections")
        table = torch.empty(nbytes // 8, dtype=torch.float64, device=self.device)
        st = _lib.stream_handle(self.device)
        _lib.check(L.tsiir_serve_prepare_filter(sos_t.data_ptr(), _lib.F64, S, table.data_ptr(),
                                                st), "tsiir_serve_prepare_filter")
        state = torch.zeros((channels, S, 2), dtype=torch.float64, device=self.device)
        if zi is not None:
            state.copy_(torch.as_tensor(zi, dtype=torch.float64).reshape(channels, S, 2))
        self._table[stream_id] = table
        self._state[stream_id] = state
        self._sections[stream_id] = S
        self._channels[stream_id] = channels
        self._sos_keep = getattr(self, "_sos_keep", [])
        self._sos_keep.append(sos_t)

    def sections(self, stream_id: int) -> int:
        return self._sections[stream_id]

    def channels(self, stream_id: int) -> int:
        return self._channels[stream_id]

    def table(self, stream_id: int) -> torch.Tensor:
        return self._table[stream_id]

    def state(self, stream_id: int) -> torch.Tensor:
        return self._state[stream_id]

    def ready(self) -> None:
        """Wait for the table builds (they run on the registering stream)."""
        torch.cuda.synchronize(self.device)

    def snapshot(self, stream_id: int) -> np.ndarray:
        """Host copy of a stream's state (checkpoint / eviction)."""
        return self._state[stream_id].cpu().numpy()

    def restore(self, stream_id: int, state) -> None:
        self._state[stream_id].copy_(torch.as_tensor(state, dtype=torch.float64))


class Executor(Protocol):
    def launch(self, batch: list[Request], table: StreamTable) -> object: ...

    def done(self, handle: object) -> bool: ...


class HipExecutor:
    """Launches one `tsiir_serve_batch` per batch.  Batches go round-robin
    over the HIP streams (default priority) of their lane: lane 0 for batches
    of short chunks, lane 1 for batches holding a long chunk, so short
    requests never queue behind a long one in a hardware queue.
    precision "fp32c" (default) runs FP64 recurrence updates on FP32 data;
    "fp32" runs the FP32 recurrence."""

    def __init__(self, device: torch.device, n_streams: int = 4, precision: str = "fp32c"):
        self.device = device
        n_long = max(1, n_streams // 2)
        n_short = max(1, n_streams - n_long)
        self.lanes = [[torch.cuda.Stream(device) for _ in range(n_short)],
                      [torch.cuda.Stream(device) for _ in range(n_long)]]
        self._next = [0, 0]
        self._lib = _lib.load()
        if precision not in ("fp32", "fp32c"):
            raise ValueError("precision must be 'fp32' or 'fp32c'")
        self.mode = _lib.MODE_COMPENSATED if precision == "fp32c" else _lib.MODE_PLAIN

    def launch(self, batch: list[Request], table: StreamTable, lane: int = 0):
        C = np.fromiter((r.channels for r in batch), dtype=np.int64, count=len(batch))
        T = np.fromiter((r.samples for r in batch), dtype=np.int64, count=len(batch))
        S = np.fromiter((table.sections(r.stream_id) for r in batch), dtype=np.int64,
                        count=len(batch))
        for r in batch:
            if r.y is None:
                r.y = torch.empty_like(r.x)
        xb = np.fromiter((r.x.data_ptr() for r in batch), dtype=np.uint64, count=len(batch))
        yb = np.fromiter((r.y.data_ptr() for r in batch), dtype=np.uint64, count=len(batch))
        tb = np.fromiter((table.table(r.stream_id).data_ptr() for r in batch), dtype=np.uint64,
                         count=len(batch))
        sb = np.fromiter((table.state(r.stream_id).data_ptr() for r in batch), dtype=np.uint64,
                         count=len(batch))
        n = int(C.sum())
        req = np.repeat(np.arange(len(batch)), C)            # request of each item
        ch = np.arange(n) - np.repeat(np.cumsum(C) - C, C)   # channel within its request
        host = torch.empty(n * ITEM_DTYPE.itemsize, dtype=torch.uint8, pin_memory=True)
        items = host.numpy().view(ITEM_DTYPE)
        items["x"] = xb[req] + (4 * ch * T[req]).astype(np.uint64)
        items["y"] = yb[req] + (4 * ch * T[req]).astype(np.uint64)
        items["table"] = tb[req]
        items["state"] = sb[req] + (16 * S[req] * ch).astype(np.uint64)
        items["n_samples"] = T[req]
        items["n_sections"] = S[req]
        items["reserved"] = 0
        streams = self.lanes[lane]
        stream = streams[self._next[lane]]
        self._next[lane] = (self._next[lane] + 1) % len(streams)
        # the inputs (and fresh outputs) were produced on the caller's stream
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            dev_items = host.to(self.device, non_blocking=True)
            _lib.check(self._lib.tsiir_serve_batch(dev_items.data_ptr(), n, self.mode,
                                                   stream.cuda_stream), "tsiir_serve_batch")
            ev = torch.cuda.Event()
            ev.record(stream)
        # keep the caching allocator from reusing these blocks before the
        # side stream is done with them
        for r in batch:
            r.x.record_stream(stream)
            r.y.record_stream(stream)
            table.state(r.stream_id).record_stream(stream)
            table.table(r.stream_id).record_stream(stream)
        dev_items.record_stream(stream)
        return (ev, host, dev_items)

    def done(self, handle) -> bool:
        return handle[0].query()


@dataclass
class ServeStats:
    n_requests: int
    wall_s: float
    sample_sections: float
    samples: float
    latencies: np.ndarray = field(repr=False)
    batches: int = 0
    mean_batch_requests: float = 0.0

    @property
    def p50(self) -> float:
        return float(np.percentile(self.latencies, 50))

    @property
    def p99(self) -> float:
        return float(np.percentile(self.latencies, 99))

    @property
    def gss(self) -> float:
        return self.sample_sections / self.wall_s / 1e9


class ServingEngine:
    """Continuous batching over an Executor.

    hbm_budget_bytes bounds the input+output bytes of requests on the device
    at once (admission control); max_batch_channels bounds one launch;
    max_inflight bounds concurrently running batches per lane.  Chunks longer
    than long_samples are batched separately (lane 1) from shorter ones
    (lane 0), so that latency-sensitive short chunks are not held behind long
    ones."""

    def __init__(self, table: StreamTable, executor: Executor, hbm_budget_bytes: int = 8 << 30,
                 max_batch_channels: int = 4096, max_inflight: int = 4,
                 long_samples: int = 16384,
                 clock: Callable[[], float] = time.perf_counter):
        self.table = table
        self.executor = executor
        self.budget = int(hbm_budget_bytes)
        self.max_channels = int(max_batch_channels)
        self.max_inflight = int(max_inflight)
        self.long_samples = int(long_samples)
        self.clock = clock
        self.batches: list[list[int]] = []  # request ids per launched batch (for tests)

    def validate(self, r: Request) -> None:
        """Reject a request the kernel cannot serve safely: the state slot
        and the x / y offsets it is given assume a registered stream, a
        (channels, samples) float32 contiguous chunk on the stream table's
        device whose channel count is the stream's."""
        dev = getattr(self.table, "device", None)
        if not hasattr(self.table, "channels"):
            return
        try:
            ch = self.table.channels(r.stream_id)
        except KeyError:
            raise ValueError(f"request {r.rid}: stream {r.stream_id} is not registered") from None
        for name, t in (("x", r.x), ("y", r.y)):
            if t is None:
                continue
            if t.dim() != 2 or t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError(f"request {r.rid}: {name} must be a contiguous float32 "
                                 f"(channels, samples) tensor, got {tuple(t.shape)} {t.dtype}")
            if dev is not None and t.device != torch.device(dev):
                raise ValueError(f"request {r.rid}: {name} is on {t.device}, the stream "
                                 f"table on {dev}")
        if r.x.shape[0] != ch:
            raise ValueError(f"request {r.rid}: {r.x.shape[0]} channels for stream "
                             f"{r.stream_id} of {ch} channels")
        if r.y is not None and r.y.shape != r.x.shape:
            raise ValueError(f"request {r.rid}: y {tuple(r.y.shape)} does not match x "
                             f"{tuple(r.x.shape)}")

    def run(self, requests: list[Request]) -> ServeStats:
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in reqs:
            self.validate(r)
            if r.nbytes > self.budget:
                raise ValueError(f"request {r.rid} needs {r.nbytes} B, above the HBM budget")
            if r.channels > self.max_channels:
                raise ValueError(f"request {r.rid} has more channels than a batch allows")
        if hasattr(self.table, "ready"):
            self.table.ready()
        n = len(reqs)
        i = 0
        ready: deque[Request] = deque()
        inflight: list[tuple[object, list[Request], int, int]] = []
        busy: set[int] = set()
        used = 0
        t0 = self.clock()
        while i < n or ready or inflight:
            now = self.clock() - t0
            while i < n and reqs[i].arrival <= now:
                ready.append(reqs[i])
                i += 1
            still = []
            for rec in inflight:
                h, batch, nb, _ = rec
                if self.executor.done(h):
                    tdone = self.clock() - t0
                    for r in batch:
                        r.t_done = tdone
                        busy.discard(r.stream_id)
                    used -= nb
                else:
                    still.append(rec)
            inflight = still
            launched = False
            # two lanes: chunks of at most long_samples samples, and longer ones
            for lane in (0, 1):
                n_lane = sum(1 for x in inflight if x[3] == lane)
                if not ready or n_lane >= self.max_inflight:
                    continue
                batch, nb, ch = [], 0, 0
                taken: set[int] = set()
                blocked: set[int] = set()  # keep each stream's arrival order
                for r in ready:
                    sid = r.stream_id
                    if sid in busy or sid in taken or sid in blocked:
                        blocked.add(sid)
                        continue
                    if (r.samples > self.long_samples) != (lane == 1):
                        blocked.add(sid)
                        continue
                    if used + nb + r.nbytes > self.budget or ch + r.channels > self.max_channels:
                        blocked.add(sid)
                        continue
                    batch.append(r)
                    taken.add(sid)
                    nb += r.nbytes
                    ch += r.channels
                if batch:
                    ids = {id(r) for r in batch}
                    ready = deque(r for r in ready if id(r) not in ids)
                    tl = self.clock() - t0
                    for r in batch:
                        r.t_launch = tl
                        busy.add(r.stream_id)
                    h = self.executor.launch(batch, self.table, lane)
                    inflight.append((h, batch, nb, lane))
                    used += nb
                    self.batches.append([r.rid for r in batch])
                    launched = True
            if not launched and not inflight and not ready and i < n:
                # idle until the next arrival
                wait = reqs[i].arrival - (self.clock() - t0)
                if wait > 2e-4:
                    time.sleep(wait - 1e-4)
        wall = self.clock() - t0
        lat = np.array([r.latency for r in reqs]) if reqs else np.zeros(1)
        ss = float(sum(r.channels * r.samples * self.table.sections(r.stream_id) for r in reqs))
        smp = float(sum(r.channels * r.samples for r in reqs))
        nbat = len(self.batches)
        return ServeStats(n, wall, ss, smp, lat, nbat, n / max(nbat, 1))


def run_plain(requests: list[Request], sos_by_stream: dict, table_channels: dict,
              device: torch.device, clock: Callable[[], float] = time.perf_counter,
              precision: str = "fp32c") -> ServeStats:
    """Per-request plain calls of the drop-in sosfilt (one call per request
    in arrival order, each from the stream's current zi, waiting for its
    result), the baseline the engine is compared with."""
    from ..signal_processing.filter import sosfilt
    reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
    zi = {sid: torch.zeros((c, sos_by_stream[sid].shape[0], 2), dtype=torch.float32,
                           device=device) for sid, c in table_channels.items()}
    cdt = torch.float64 if precision == "fp32c" else torch.float32
    sos_dev = {sid: torch.as_tensor(s, dtype=cdt).to(device) for sid, s in sos_by_stream.items()}
    t0 = clock()
    for r in reqs:
        wait = r.arrival - (clock() - t0)
        if wait > 0:
            time.sleep(wait)
        r.t_launch = clock() - t0
        y, zf = sosfilt(sos_dev[r.stream_id], r.x, zi=zi[r.stream_id], precision=precision)
        zi[r.stream_id] = zf
        torch.cuda.current_stream(device).synchronize()
        r.y = y
        r.t_done = clock() - t0
    wall = clock() - t0
    lat = np.array([r.latency for r in reqs])
    ss = float(sum(r.channels * r.samples * sos_by_stream[r.stream_id].shape[0] for r in reqs))
    smp = float(sum(r.channels * r.samples for r in reqs))
    return ServeStats(len(reqs), wall, ss, smp, lat, len(reqs), 1.0)
