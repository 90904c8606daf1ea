"""ctypes binding of ``libtsiir.so`` (C ABI declared in include/tsiir.h).

Torch is used only for device memory and streams: tensors hand their device
pointers and the current HIP stream to the C entry points.  There is no
fallback: if the library is missing the import of the filtering ops raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libtsiir.so")

OK, EINVAL, EUNSUPPORTED, ELAUNCH, EWORKSPACE, ESCRATCH, EAGAIN = range(7)
NOALLOC_SCRATCH_BYTES = 64 << 20
F32, F64, BF16 = 0, 1, 2
MODE_PLAIN, MODE_COMPENSATED = 0, 1
HAS_ZI, WRITE_ZF, ZI_BROADCAST, REVERSE = 0x1, 0x2, 0x4, 0x8
PATH_WINDOW, PATH_CHANNEL = 0x10, 0x20
FIR_REVERSE_TAPS = 0x40
FP32_DF2T = 0x80
DEBUG_GIVE_UP = 0x100
STATE_F64 = 0x200
NO_SEGMENT = 0x400
SEGMENT_WARMUP = 4096
PAD_NONE, PAD_ODD, PAD_EVEN, PAD_CONSTANT = 0, 1, 2, 3
# tsiir_sosfiltfilt's bounds (beyond them it returns EUNSUPPORTED)
SOSFILTFILT_MAX_SECTIONS = 8
SOSFILTFILT_MAX_WORK = 16384  # padlen x sections
PADLEN_DEFAULT = -1  # the default padlen, formed on the device
ERR_STEADY_STATE = 4  # workspace error bit: a filter with no steady state

# Every symbol include/tsiir.h declares.
EXPORTED = (
    "tsiir_version",
    "tsiir_status_string",
    "tsiir_sosfilt_workspace_bytes",
    "tsiir_workspace_init",
    "tsiir_workspace_set_error_mailbox",
    "tsiir_sosfilt_forward",
    "tsiir_sosfiltfilt_scratch_bytes",
    "tsiir_sosfiltfilt",
    "tsiir_sosfilt_forward_noalloc",
    "tsiir_lfilter_forward",
    "tsiir_lfilter_workspace_bytes",
    "tsiir_lfilter_forward_ws",
    "tsiir_sosfilt_backward_workspace_bytes",
    "tsiir_sosfilt_backward",
    "tsiir_sosfilt_forward_saved",
    "tsiir_sosfilt_backward_saved_workspace_bytes",
    "tsiir_sosfilt_backward_saved",
    "tsiir_fir_forward",
    "tsiir_fir_fft_workspace_bytes",
    "tsiir_fir_forward_fft",
    "tsiir_fir_backward_taps_workspace_bytes",
    "tsiir_fir_backward_taps",
    "tsiir_sos_frequency_response",
    "tsiir_ba_frequency_response",
    "tsiir_iir_design_sections",
    "tsiir_iir_design_sos",
    "tsiir_iir_design_sos_jacobian",
    "tsiir_serve_table_bytes",
    "tsiir_serve_prepare_filter",
    "tsiir_serve_batch",
    "tsiir_serve_create",
    "tsiir_serve_submit",
    "tsiir_serve_poll",
    "tsiir_serve_destroy",
)


class SosfiltDesc(ctypes.Structure):
    _fields_ = [
        ("channels", ctypes.c_int64),
        ("n_samples", ctypes.c_int64),
        ("n_sections", ctypes.c_int32),
        ("io_dtype", ctypes.c_int32),
        ("coef_dtype", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_int32),
        ("x_rows", ctypes.c_int64),
        ("x_row_stride", ctypes.c_int64),
        ("y_row_stride", ctypes.c_int64),
        ("chans_per_filter", ctypes.c_int64),
    ]


class FirDesc(ctypes.Structure):
    _fields_ = [
        ("channels", ctypes.c_int64),
        ("n_samples", ctypes.c_int64),
        ("x_len", ctypes.c_int64),
        ("n_taps", ctypes.c_int32),
        ("io_dtype", ctypes.c_int32),
        ("shift", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("x_rows", ctypes.c_int64),
        ("x_row_stride", ctypes.c_int64),
        ("y_row_stride", ctypes.c_int64),
        ("chans_per_filter", ctypes.c_int64),
    ]


class IirDesignDesc(ctypes.Structure):
    _fields_ = [
        ("n_filters", ctypes.c_int64),
        ("family", ctypes.c_int32),
        ("band", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("out_dtype", ctypes.c_int32),
        ("param", ctypes.c_double),
    ]


class ServeItem(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_void_p),
        ("y", ctypes.c_void_p),
        ("table", ctypes.c_void_p),
        ("state", ctypes.c_void_p),
        ("n_samples", ctypes.c_int64),
        ("n_sections", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class TsiirError(RuntimeError):
    """A libtsiir entry point returned a non-OK status."""


_lib = None
_lock = threading.Lock()
TORCH_LIB_PATH = os.path.join(_PKG, "_lib", "libtsiir_torch.so")
_ops_loaded = False


def ops():
    """torch.ops.tsiir: the operators registered by libtsiir_torch.so
    (sosfilt_forward / sosfilt_backward with HIP and Meta implementations).
    Loaded once; raises if the library has not been built."""
    global _ops_loaded
    if not _ops_loaded:
        with _lock:
            if not _ops_loaded:
                load()
                if not os.path.exists(TORCH_LIB_PATH):
                    raise ImportError(f"libtsiir_torch.so not found at {TORCH_LIB_PATH}; build "
                                      "it with `python -m torchscience_amd.build`")
                torch.ops.load_library(TORCH_LIB_PATH)
                _ops_loaded = True
    return torch.ops.tsiir


def load(path: str | None = None) -> ctypes.CDLL:
    """Load libtsiir.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"libtsiir.so not found at {p}; build it with `python -m torchscience_amd.build`")
    L = ctypes.CDLL(p)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    dp = ctypes.POINTER(SosfiltDesc)
    L.tsiir_version.restype = ctypes.c_int
    L.tsiir_status_string.restype = ctypes.c_char_p
    L.tsiir_status_string.argtypes = [ctypes.c_int]
    L.tsiir_sosfilt_workspace_bytes.restype = sz
    L.tsiir_sosfilt_workspace_bytes.argtypes = [dp]
    L.tsiir_workspace_init.restype = ctypes.c_int
    L.tsiir_workspace_init.argtypes = [vp, sz, vp]
    L.tsiir_workspace_set_error_mailbox.restype = ctypes.c_int
    L.tsiir_workspace_set_error_mailbox.argtypes = [vp, sz, vp, vp]
    L.tsiir_sosfilt_forward.restype = ctypes.c_int
    L.tsiir_sosfilt_forward.argtypes = [dp, vp, vp, vp, vp, vp, vp, sz, vp]
    L.tsiir_sosfiltfilt_scratch_bytes.restype = sz
    L.tsiir_sosfiltfilt_scratch_bytes.argtypes = [dp, ctypes.c_int64]
    L.tsiir_sosfiltfilt.restype = ctypes.c_int
    L.tsiir_sosfiltfilt.argtypes = [dp, ctypes.c_int, ctypes.c_int64, vp, vp, vp, vp, vp, sz, vp,
                                    sz, vp]
    L.tsiir_sosfilt_forward_noalloc.restype = ctypes.c_int
    L.tsiir_sosfilt_forward_noalloc.argtypes = [dp, vp, vp, vp, vp, vp, vp]
    L.tsiir_lfilter_forward.restype = ctypes.c_int
    L.tsiir_lfilter_forward.argtypes = [dp, vp, vp, vp, vp, vp, vp]
    L.tsiir_lfilter_workspace_bytes.restype = sz
    L.tsiir_lfilter_workspace_bytes.argtypes = [dp]
    L.tsiir_lfilter_forward_ws.restype = ctypes.c_int
    L.tsiir_lfilter_forward_ws.argtypes = [dp, vp, vp, vp, vp, vp, vp, sz, vp]
    L.tsiir_sosfilt_backward_workspace_bytes.restype = sz
    L.tsiir_sosfilt_backward_workspace_bytes.argtypes = [dp]
    L.tsiir_sosfilt_backward.restype = ctypes.c_int
    L.tsiir_sosfilt_backward.argtypes = [dp, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, vp]
    L.tsiir_sosfilt_forward_saved.restype = ctypes.c_int
    L.tsiir_sosfilt_forward_saved.argtypes = [dp, vp, vp, vp, vp, vp, vp, vp, sz, vp]
    L.tsiir_sosfilt_backward_saved_workspace_bytes.restype = sz
    L.tsiir_sosfilt_backward_saved_workspace_bytes.argtypes = [dp]
    L.tsiir_sosfilt_backward_saved.restype = ctypes.c_int
    L.tsiir_sosfilt_backward_saved.argtypes = [dp] + [vp] * 11 + [sz, vp]
    fp = ctypes.POINTER(FirDesc)
    L.tsiir_fir_forward.restype = ctypes.c_int
    L.tsiir_fir_forward.argtypes = [fp, vp, vp, vp, vp, vp, vp]
    L.tsiir_fir_backward_taps_workspace_bytes.restype = sz
    L.tsiir_fir_backward_taps_workspace_bytes.argtypes = [fp]
    L.tsiir_fir_backward_taps.restype = ctypes.c_int
    L.tsiir_fir_backward_taps.argtypes = [fp, vp, vp, vp, vp, sz, vp]
    L.tsiir_fir_fft_workspace_bytes.restype = sz
    L.tsiir_fir_fft_workspace_bytes.argtypes = [fp]
    L.tsiir_fir_forward_fft.restype = ctypes.c_int
    L.tsiir_fir_forward_fft.argtypes = [fp, vp, vp, vp, vp, sz, vp]
    L.tsiir_sos_frequency_response.restype = ctypes.c_int
    L.tsiir_sos_frequency_response.argtypes = [vp, ctypes.c_int32, ctypes.c_int64,
                                               ctypes.c_int32, vp, ctypes.c_int64, vp, vp]
    gp = ctypes.POINTER(IirDesignDesc)
    L.tsiir_iir_design_sections.restype = ctypes.c_int32
    L.tsiir_iir_design_sections.argtypes = [gp]
    L.tsiir_iir_design_sos.restype = ctypes.c_int
    L.tsiir_iir_design_sos.argtypes = [gp, vp, vp, vp]
    L.tsiir_iir_design_sos_jacobian.restype = ctypes.c_int
    L.tsiir_iir_design_sos_jacobian.argtypes = [gp, vp, vp, vp, vp]
    L.tsiir_ba_frequency_response.restype = ctypes.c_int
    L.tsiir_ba_frequency_response.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32,
                                              ctypes.c_int64, vp, ctypes.c_int64, vp, vp]
    L.tsiir_serve_table_bytes.restype = sz
    L.tsiir_serve_table_bytes.argtypes = [ctypes.c_int32]
    L.tsiir_serve_prepare_filter.restype = ctypes.c_int
    L.tsiir_serve_prepare_filter.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp, vp]
    L.tsiir_serve_batch.restype = ctypes.c_int
    L.tsiir_serve_batch.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
    L.tsiir_serve_create.restype = ctypes.c_int
    L.tsiir_serve_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.POINTER(vp)]
    L.tsiir_serve_submit.restype = ctypes.c_int
    L.tsiir_serve_submit.argtypes = [vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]
    L.tsiir_serve_poll.restype = ctypes.c_int
    L.tsiir_serve_poll.argtypes = [vp, ctypes.c_uint64, ctypes.c_int32,
                                   ctypes.POINTER(ctypes.c_int32)]
    L.tsiir_serve_destroy.restype = ctypes.c_int
    L.tsiir_serve_destroy.argtypes = [vp]
    if path is None:
        _lib = L
    return L


def check(status: int, what: str) -> None:
    if status != OK:
        msg = load().tsiir_status_string(status).decode()
        raise TsiirError(f"{what} failed: {msg} (status {status})")


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def io_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return F32
    if dtype == torch.float64:
        return F64
    if dtype == torch.bfloat16:
        return BF16
    raise TypeError(f"unsupported I/O dtype {dtype}")


def error_text(code: int) -> str:
    """What a workspace error code means (the bits of its error word)."""
    if code == ERR_STEADY_STATE:
        return ("sosfiltfilt: a filter has no steady state (a section's system for "
                "sosfilt_zi is singular; error 4): its outputs are NaN")
    return (f"a look-back wait gave up (error {code}): the outputs of the affected "
            "windows are NaN")


def check_errors(device: torch.device | None = None) -> None:
    """Synchronise `device` (or all devices) and raise TsiirError if an
    earlier call failed on the device: a look-back wait gave up, or
    sosfiltfilt met a filter with no steady state (their outputs are NaN).
    Calls also report it themselves, without synchronising, on the next call
    on the same stream."""
    if device is not None:
        torch.cuda.synchronize(device)
    else:
        for i in range(torch.cuda.device_count()):
            torch.cuda.synchronize(i)
    code = int(ops().take_error())
    if code != 0:
        raise TsiirError(error_text(code))
