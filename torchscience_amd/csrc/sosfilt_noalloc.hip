// sosfilt_noalloc.hip - tsiir_sosfilt_forward for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops that
// forbid hipMalloc): the forward's workspace (coefficient tables, scan status
// words, header) lives in a fixed device array of this library instead of a
// caller-sized buffer.
//
// The array `scratch` is a file-scope device variable: each GPU the library
// runs on holds its own copy, at its own address, until the process ends.
// On the host its name is only a stand-in for the symbol, so the per-device
// address comes from hipGetSymbolAddress at the first call on a device and
// is the only pointer the host hands to kernels.
//
// State across calls: the first call on a device zero-fills that device's
// copy with a fill kernel on the caller's stream (so that first call should
// not be under graph capture); afterwards every call advances the header's
// epoch before publishing scan status words, so words left by earlier calls
// never match.  All calls on one GPU share the array: calls that overlap in
// time on one GPU (different streams, or different host threads) race on it
// and must be ordered by the caller.

#include <mutex>

#include "sos_internal.h"
#include "tsiir.h"

namespace tsiir {

constexpr size_t kScratchBytes = TSIIR_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

__global__ void scratch_zero_kernel(uint32_t* p, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride)
    p[i] = 0u;
}

struct ScratchSlot {
  std::once_flag once;
  void* ptr = nullptr;
  tsiir_status status = TSIIR_OK;
};
ScratchSlot g_slots[kMaxDevices];

// The device copy of `scratch` for the current device, zero-filled on
// `stream` by the first call on that device.
tsiir_status scratch_for_device(hipStream_t stream, void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return TSIIR_ELAUNCH;
  ScratchSlot& slot = g_slots[dev];
  std::call_once(slot.once, [&] {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
      slot.status = TSIIR_ELAUNCH;
      return;
    }
    hipLaunchKernelGGL(scratch_zero_kernel, dim3(1024), dim3(256), 0, stream,
                       static_cast<uint32_t*>(p), kScratchBytes / 4);
    if (hipGetLastError() != hipSuccess) {
      slot.status = TSIIR_ELAUNCH;
      return;
    }
    slot.ptr = p;
  });
  if (slot.status != TSIIR_OK) return slot.status;
  *out = slot.ptr;
  return TSIIR_OK;
}

}  // namespace

}  // namespace tsiir

using namespace tsiir;

extern "C" {

tsiir_status tsiir_sosfilt_forward_noalloc(const tsiir_sosfilt_desc* desc, const void* x,
                                           void* y, const void* sos, const void* zi, void* zf,
                                           void* stream) {
  FwdGeometry g;
  if (!fwd_geometry(desc, &g)) return TSIIR_EINVAL;
  if (fwd_workspace_bytes(desc, false) > kScratchBytes) return TSIIR_ESCRATCH;
  void* ws = nullptr;
  const tsiir_status st = scratch_for_device(static_cast<hipStream_t>(stream), &ws);
  if (st != TSIIR_OK) return st;
  return forward_impl(desc, x, y, sos, zi, zf, ws, kScratchBytes, stream, nullptr);
}

}  // extern "C"
