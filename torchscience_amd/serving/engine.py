"""Continuous-batching server for streaming sosfilt.

Streams are persistent: each has a filter (1..32 sections) and a channel
count; its filter state (the zi/zf of the reference's sosfilt,
src/torchscience/signal_processing/filter/_sosfilt.py:98-133) stays resident on
the device between chunks, in FP64.  Requests are chunks (channels x samples)
of a stream.  The engine admits arrived requests while the HBM budget allows,
groups ready requests of distinct, idle streams into batches and keeps up to
`max_inflight` batches per lane running on separate HIP streams, so requests
that arrive while others are on the device are launched as soon as slots,
budget and their stream free up, and their work overlaps on the device.

Two lanes: chunks up to `long_samples` go to one ragged-batch launch
(`tsiir_serve_batch`: one wave per channel walks its chunk, no waits between
items); longer chunks run the window-parallel forward (`tsiir_sosfilt_forward`
with the stream's FP64 state, TSIIR_STATE_F64), which splits a chunk's
samples over many waves chained by the look-back, so a long chunk's latency
does not grow with one wave walking it.

Ordering: a stream's chunks run one at a time, in arrival order (its state
threads through them).  Each item's result depends only on its own chunk
and state, never on what else is in the batch.

Failure paths: a request can be cancelled (`cancel`) or given a deadline
until it is launched (then it is dropped, status "cancelled" / "expired",
and its stream's state does not advance); long chunks commit their new
state only when their launch completed without a look-back give-up
(status "failed" otherwise, state unchanged); a stream's state can be
evicted to the host and restored.

The executor interface (`launch`, `done`) separates the scheduler from the
device; `HipExecutor` runs the HIP kernels, and tests drive the scheduler
with a CPU executor.
"""

from __future__ import annotations

import ctypes
import gc
import threading
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
    deadline: float | None = None  # latest launch time (seconds after the start)
    status: str = "pending"        # done, failed, cancelled or expired at the end

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
        self._sos: dict[int, torch.Tensor] = {}

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
        self._sos[stream_id] = sos_t.reshape(1, S, 6)

    def sections(self, stream_id: int) -> int:
        return self._sections[stream_id]

    def channels(self, stream_id: int) -> int:
        return self._channels[stream_id]

    def table(self, stream_id: int) -> torch.Tensor:
        return self._table[stream_id]

    def state(self, stream_id: int) -> torch.Tensor:
        return self._state[stream_id]

    def sos(self, stream_id: int) -> torch.Tensor:
        """The stream's (1, S, 6) FP64 sections on the device."""
        return self._sos[stream_id]

    def commit(self, stream_id: int, state: torch.Tensor) -> None:
        """Make `state` the stream's state (after the chunk that wrote it
        completed)."""
        self._state[stream_id] = state

    def ready(self) -> None:
        """Wait for the table builds (they run on the registering stream)."""
        torch.cuda.synchronize(self.device)

    def snapshot(self, stream_id: int) -> np.ndarray:
        """Host copy of a stream's state (checkpoint / eviction)."""
        return self._state[stream_id].cpu().numpy()

    def restore(self, stream_id: int, state) -> None:
        """Load a stream's state from a host copy (after evict, or a checkpoint)."""
        if stream_id not in self._state:
            raise KeyError(f"stream {stream_id} is not registered")
        shape = (self._channels[stream_id], self._sections[stream_id], 2)
        st = torch.as_tensor(state, dtype=torch.float64).reshape(shape)
        self._state[stream_id] = st.to(self.device).contiguous()

    def evict(self, stream_id: int) -> np.ndarray:
        """Move a stream's state to the host and release its device copy; the
        stream stays registered and `restore` brings the state back."""
        host = self.snapshot(stream_id)
        self._state[stream_id] = torch.empty((0,), dtype=torch.float64, device=self.device)
        return host

    def resident(self, stream_id: int) -> bool:
        return self._state[stream_id].numel() > 0


class Executor(Protocol):
    def launch(self, batch: list[Request], table: StreamTable) -> object: ...

    def done(self, handle: object) -> bool: ...


class _LongWorkspace:
    """Workspace of the window-parallel forward on one HIP stream, with a
    pinned error mailbox read when a batch completes."""

    def __init__(self, device: torch.device, stream: torch.cuda.Stream, nbytes: int):
        L = _lib.load()
        size = max(int(nbytes), 1 << 20)
        size = (size + 255) // 256 * 256
        self.buf = torch.empty(size, dtype=torch.uint8, device=device)
        self.mailbox = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        _lib.check(L.tsiir_workspace_init(self.buf.data_ptr(), size, stream.cuda_stream),
                   "tsiir_workspace_init")
        _lib.check(L.tsiir_workspace_set_error_mailbox(self.buf.data_ptr(), size,
                                                       self.mailbox.data_ptr(),
                                                       stream.cuda_stream),
                   "tsiir_workspace_set_error_mailbox")

    def error(self) -> int:
        return int(self.mailbox[0])


class HipExecutor:
    """Runs batches on the device.  Lane 0 (short chunks): one
    `tsiir_serve_batch` launch per batch, states updated in place.  Lane 1
    (long chunks): per request, the window-parallel `tsiir_sosfilt_forward`
    from the stream's FP64 state into a fresh state buffer, committed when the
    batch completes without a look-back give-up.  Batches go round-robin over
    the HIP streams (default priority) of their lane, so short requests never
    queue behind a long one in a hardware queue.  precision "fp32c" (default)
    runs FP64 recurrence updates on FP32 data; "fp32" the FP32 recurrence."""

    def __init__(self, device: torch.device, n_streams: int = 4, precision: str = "fp32c",
                 parallel_long: bool = True, parallel_long_max_channels: int = 1024,
                 timing: bool = False, _debug_give_up: bool = False):
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
        self.parallel_long = parallel_long
        # a long-chunk batch with more channels than this fills the GPU with
        # one wave per channel, so it runs as one ragged-batch launch; smaller
        # ones split every chunk over many waves (latency)
        self.parallel_long_max_channels = int(parallel_long_max_channels)
        self._ws: dict[int, _LongWorkspace] = {}
        # per-launch device time (HIP events on the launch stream) and the
        # algorithmic bytes it moved (x read + y written), when timing
        self.timing = timing
        self.kernel_ms: list[float] = []
        self.kernel_bytes: list[int] = []
        self._timed: list[tuple] = []
        # staging of item arrays: pinned host / device slot pairs reused across
        # launches (a slot returns to the pool when its batch completes), so
        # the launch path allocates nothing in steady state
        self._slots: list[tuple[torch.Tensor, torch.Tensor]] = []
        self._free: list[int] = []
        # tests only: every look-back wait of the long lane gives up at once
        self._debug_flags = _lib.DEBUG_GIVE_UP if _debug_give_up else 0

    def _stream(self, lane: int) -> torch.cuda.Stream:
        streams = self.lanes[lane]
        stream = streams[self._next[lane]]
        self._next[lane] = (self._next[lane] + 1) % len(streams)
        # the inputs (and fresh outputs) were produced on the caller's stream
        stream.wait_stream(torch.cuda.current_stream(self.device))
        return stream

    def launch(self, batch: list[Request], table: StreamTable, lane: int = 0):
        for r in batch:
            if r.y is None:
                r.y = torch.empty_like(r.x)
        if (lane == 1 and self.parallel_long and
                sum(r.channels for r in batch) <= self.parallel_long_max_channels):
            return self._launch_long(batch, table)
        C = np.fromiter((r.channels for r in batch), dtype=np.int64, count=len(batch))
        T = np.fromiter((r.samples for r in batch), dtype=np.int64, count=len(batch))
        S = np.fromiter((table.sections(r.stream_id) for r in batch), dtype=np.int64,
                        count=len(batch))
        xb = np.fromiter((r.x.data_ptr() for r in batch), dtype=np.uint64, count=len(batch))
        yb = np.fromiter((r.y.data_ptr() for r in batch), dtype=np.uint64, count=len(batch))
        tb = np.fromiter((table.table(r.stream_id).data_ptr() for r in batch), dtype=np.uint64,
                         count=len(batch))
        sb = np.fromiter((table.state(r.stream_id).data_ptr() for r in batch), dtype=np.uint64,
                         count=len(batch))
        n = int(C.sum())
        req = np.repeat(np.arange(len(batch)), C)            # request of each item
        ch = np.arange(n) - np.repeat(np.cumsum(C) - C, C)   # channel within its request
        slot = self._acquire(n * ITEM_DTYPE.itemsize)
        host_full, dev_full = self._slots[slot]
        host = host_full[:n * ITEM_DTYPE.itemsize]
        items = host.numpy().view(ITEM_DTYPE)
        items["x"] = xb[req] + (4 * ch * T[req]).astype(np.uint64)
        items["y"] = yb[req] + (4 * ch * T[req]).astype(np.uint64)
        items["table"] = tb[req]
        items["state"] = sb[req] + (16 * S[req] * ch).astype(np.uint64)
        items["n_samples"] = T[req]
        items["n_sections"] = S[req]
        items["reserved"] = 0
        stream = self._stream(lane)
        with torch.cuda.stream(stream):
            dev_items = dev_full[:n * ITEM_DTYPE.itemsize]
            dev_items.copy_(host, non_blocking=True)
            ev0 = self._start(stream)
            _lib.check(self._lib.tsiir_serve_batch(dev_items.data_ptr(), n, self.mode,
                                                   stream.cuda_stream), "tsiir_serve_batch")
            ev = torch.cuda.Event(enable_timing=self.timing)
            ev.record(stream)
            self._stop(ev0, ev, batch)
        # keep the caching allocator from reusing these blocks before the
        # side stream is done with them
        for r in batch:
            r.x.record_stream(stream)
            r.y.record_stream(stream)
            table.state(r.stream_id).record_stream(stream)
            table.table(r.stream_id).record_stream(stream)
        return {"event": ev, "slot": slot, "batch": batch, "commit": None, "table": table}

    def _launch_long(self, batch: list[Request], table: StreamTable):
        stream = self._stream(1)
        L = self._lib
        news = []
        with torch.cuda.stream(stream):
            ev0 = self._start(stream)
            for r in batch:
                C, T, S = r.channels, r.samples, table.sections(r.stream_id)
                zi = table.state(r.stream_id)
                zf = torch.empty_like(zi)
                flags = _lib.HAS_ZI | _lib.WRITE_ZF | _lib.STATE_F64 | self._debug_flags
                d = _lib.SosfiltDesc(channels=C, n_samples=T, n_sections=S, io_dtype=_lib.F32,
                                     coef_dtype=_lib.F64, mode=self.mode, flags=flags,
                                     reserved=0, x_rows=C, x_row_stride=T, y_row_stride=T,
                                     chans_per_filter=C)
                nbytes = L.tsiir_sosfilt_workspace_bytes(d)
                key = stream.cuda_stream
                ws = self._ws.get(key)
                if ws is None or ws.buf.numel() < nbytes:
                    ws = self._ws[key] = _LongWorkspace(self.device, stream, nbytes)
                _lib.check(L.tsiir_sosfilt_forward(d, r.x.data_ptr(), r.y.data_ptr(),
                                                   table.sos(r.stream_id).data_ptr(),
                                                   zi.data_ptr(), zf.data_ptr(),
                                                   ws.buf.data_ptr(), ws.buf.numel(),
                                                   stream.cuda_stream),
                           "tsiir_sosfilt_forward")
                news.append((r, zf))
                for t in (r.x, r.y, zi, zf, table.sos(r.stream_id)):
                    t.record_stream(stream)
            ev = torch.cuda.Event(enable_timing=self.timing)
            ev.record(stream)
            self._stop(ev0, ev, batch)
        return {"event": ev, "slot": None, "batch": batch, "commit": (self._ws[key], news),
                "table": table}

    def _acquire(self, nbytes: int) -> int:
        """A free staging slot of at least nbytes (a new one when none is)."""
        for k in range(len(self._free) - 1, -1, -1):
            i = self._free[k]
            if self._slots[i][0].numel() >= nbytes:
                del self._free[k]
                return i
        size = max(1 << 16, 1 << (max(nbytes, 1) - 1).bit_length())
        self._slots.append((torch.empty(size, dtype=torch.uint8, pin_memory=True),
                            torch.empty(size, dtype=torch.uint8, device=self.device)))
        return len(self._slots) - 1

    def _start(self, stream):
        if not self.timing:
            return None
        ev0 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        return ev0

    def _stop(self, ev0, ev1, batch):
        if self.timing:
            self._timed.append((ev0, ev1, sum(8 * r.x.numel() for r in batch)))

    def collect_timing(self) -> None:
        """Move the finished launches' event times into kernel_ms /
        kernel_bytes (synchronises)."""
        torch.cuda.synchronize(self.device)
        for ev0, ev1, nb in self._timed:
            self.kernel_ms.append(ev0.elapsed_time(ev1))
            self.kernel_bytes.append(nb)
        self._timed.clear()

    def done(self, handle) -> bool:
        """True once the batch completed; long-lane batches then commit their
        new states, or mark their requests failed (states unchanged) when a
        look-back wait of the batch's stream gave up."""
        if not handle["event"].query():
            return False
        if handle.get("slot") is not None:
            self._free.append(handle["slot"])
            handle["slot"] = None
        commit = handle.get("commit")
        if commit is not None:
            ws, news = commit
            failed = ws.error() != 0
            for r, zf in news:
                if failed:
                    r.status = "failed"
                else:
                    handle["table"].commit(r.stream_id, zf)
            if failed:
                # the workspace's error word stays set: start a fresh one
                self._ws = {k: v for k, v in self._ws.items() if v is not ws}
            handle["commit"] = None
        return True


class PersistentExecutor(HipExecutor):
    """HipExecutor whose ragged batches go to a persistent launch through a
    device queue (tsiir_serve_create / submit / poll / destroy): one grid of
    waves keeps taking the submitted chunk channels while it runs, so a
    batch costs a host-side submit instead of a kernel launch.  Long chunks
    of small batches still run the window-parallel forward.  Inputs must be
    complete when submitted: the caller's stream is waited for (host side)
    when it still has work queued."""

    def __init__(self, device: torch.device, n_streams: int = 4, precision: str = "fp32c",
                 capacity: int = 1 << 18, workgroups: int = 2048, idle_us: int = 20000, **kw):
        super().__init__(device, n_streams, precision, **kw)
        q = ctypes.c_void_p()
        with torch.cuda.device(device):
            _lib.check(self._lib.tsiir_serve_create(capacity, workgroups, self.mode, idle_us,
                                                    ctypes.byref(q)), "tsiir_serve_create")
        self._q = q

    def launch(self, batch: list[Request], table: StreamTable, lane: int = 0):
        if (lane == 1 and self.parallel_long and
                sum(r.channels for r in batch) <= self.parallel_long_max_channels):
            return super().launch(batch, table, lane)
        for r in batch:
            if r.y is None:
                r.y = torch.empty_like(r.x)
        cur = torch.cuda.current_stream(self.device)
        if not cur.query():
            cur.synchronize()
        C = np.fromiter((r.channels for r in batch), dtype=np.int64, count=len(batch))
        T = np.fromiter((r.samples for r in batch), dtype=np.int64, count=len(batch))
        S = np.fromiter((table.sections(r.stream_id) for r in batch), dtype=np.int64,
                        count=len(batch))
        n = int(C.sum())
        req = np.repeat(np.arange(len(batch)), C)
        ch = np.arange(n) - np.repeat(np.cumsum(C) - C, C)
        items = np.empty(n, dtype=ITEM_DTYPE)
        xb = np.fromiter((r.x.data_ptr() for r in batch), dtype=np.uint64, count=len(batch))
        yb = np.fromiter((r.y.data_ptr() for r in batch), dtype=np.uint64, count=len(batch))
        tb = np.fromiter((table.table(r.stream_id).data_ptr() for r in batch), dtype=np.uint64,
                         count=len(batch))
        sb = np.fromiter((table.state(r.stream_id).data_ptr() for r in batch), dtype=np.uint64,
                         count=len(batch))
        items["x"] = xb[req] + (4 * ch * T[req]).astype(np.uint64)
        items["y"] = yb[req] + (4 * ch * T[req]).astype(np.uint64)
        items["table"] = tb[req]
        items["state"] = sb[req] + (16 * S[req] * ch).astype(np.uint64)
        items["n_samples"] = T[req]
        items["n_sections"] = S[req]
        items["reserved"] = 0
        first = ctypes.c_uint64()
        t0 = time.perf_counter()
        while True:
            rc = self._lib.tsiir_serve_submit(self._q, items.ctypes.data, n, ctypes.byref(first))
            if rc != _lib.EAGAIN or time.perf_counter() - t0 > 30.0:
                break
            # a slot this batch needs still holds a running item: let it finish
            time.sleep(5e-5)
        _lib.check(rc, "tsiir_serve_submit")
        return {"queue": (first.value, n), "slot": None, "batch": batch, "commit": None,
                "table": table}

    def done(self, handle) -> bool:
        if "queue" not in handle:
            return super().done(handle)
        first, n = handle["queue"]
        flag = ctypes.c_int32()
        _lib.check(self._lib.tsiir_serve_poll(self._q, first, n, ctypes.byref(flag)),
                   "tsiir_serve_poll")
        return bool(flag.value)

    def close(self) -> None:
        """Finish the submitted chunks and release the queue."""
        if self._q:
            _lib.check(self._lib.tsiir_serve_destroy(self._q), "tsiir_serve_destroy")
            self._q = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
        self._cancel: set[int] = set()
        self._cancel_lock = threading.Lock()

    def cancel(self, rid: int) -> None:
        """Cancel request `rid` if it has not been launched yet (safe to call
        from another thread while `run` is going)."""
        with self._cancel_lock:
            self._cancel.add(rid)

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
        if hasattr(self.table, "resident") and not self.table.resident(r.stream_id):
            raise ValueError(f"request {r.rid}: stream {r.stream_id} is evicted; restore it "
                             "first")

    def run(self, requests: list[Request]) -> ServeStats:
        """Serve `requests` (arrival times in seconds after the start) to
        completion.  Python's cyclic garbage collector is paused meanwhile,
        so that none of its pauses lands in a request's latency."""
        was = gc.isenabled()
        gc.disable()
        try:
            return self._run(requests)
        finally:
            if was:
                gc.enable()

    def _run(self, requests: list[Request]) -> ServeStats:
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
            # requests cancelled, or past their deadline, before their launch
            with self._cancel_lock:
                cancel = set(self._cancel)
            if cancel or any(r.deadline is not None for r in ready):
                keep = deque()
                for r in ready:
                    if r.rid in cancel:
                        r.status, r.t_done = "cancelled", now
                    elif r.deadline is not None and now > r.deadline:
                        r.status, r.t_done = "expired", now
                    else:
                        keep.append(r)
                ready = keep
            still = []
            for rec in inflight:
                h, batch, nb, _ = rec
                if self.executor.done(h):
                    tdone = self.clock() - t0
                    for r in batch:
                        r.t_done = tdone
                        if r.status == "pending":
                            r.status = "done"
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
        served = [r for r in reqs if r.status == "done"]
        lat = np.array([r.latency for r in served]) if served else np.zeros(1)
        ss = float(sum(r.channels * r.samples * self.table.sections(r.stream_id)
                       for r in served))
        smp = float(sum(r.channels * r.samples for r in served))
        nbat = len(self.batches)
        return ServeStats(n, wall, ss, smp, lat, nbat, n / max(nbat, 1))


def run_plain(requests: list[Request], sos_by_stream: dict, table_channels: dict,
              device: torch.device, clock: Callable[[], float] = time.perf_counter,
              precision: str = "fp32c") -> ServeStats:
    """Per-request plain calls of the drop-in sosfilt (one call per request
    in arrival order, each from the stream's current zi, waiting for its
    result), the baseline the engine is compared with.  With "fp32c" the
    state is carried in float64 between chunks, as the engine carries it."""
    from ..signal_processing.filter import sosfilt
    reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
    sdt = torch.float64 if precision == "fp32c" else torch.float32
    zi = {sid: torch.zeros((c, sos_by_stream[sid].shape[0], 2), dtype=sdt,
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
        r.status = "done"
    wall = clock() - t0
    lat = np.array([r.latency for r in reqs])
    ss = float(sum(r.channels * r.samples * sos_by_stream[r.stream_id].shape[0] for r in reqs))
    smp = float(sum(r.channels * r.samples for r in reqs))
    return ServeStats(len(reqs), wall, ss, smp, lat, len(reqs), 1.0)
