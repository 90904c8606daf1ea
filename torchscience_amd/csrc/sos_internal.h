// sos_internal.h - host-side entry points shared between the kernel files
// of libtsiir (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "tsiir.h"

namespace tsiir {

size_t align_up(size_t v, size_t a);

struct FwdGeometry {
  int W;               // samples per lane
  int64_t n_windows;   // windows of 64*W samples per channel
  int64_t n_filters;
  size_t table_bytes;
  size_t status_bytes;
  size_t fast_off;     // Fast32 tables of the FP32-recurrence kernel
  size_t status_off;   // scan status words
  size_t total;        // workspace bytes
};
// Workspace layout of the forward: header, FP64 tables, Fast32 tables, status
// words.  The tables' offsets depend only on the filter count and the
// section count, so launches over different row geometries of one call (the
// segmented form) can share one table preparation.

bool fwd_geometry(const tsiir_sosfilt_desc* d, FwdGeometry* g);

// Workspace bytes of the forward for this descriptor, including the
// segmented form's buffers where TSIIR_SEGMENTED applies (`saved`: the
// section outputs are kept, which the segmented form does not do).
size_t fwd_workspace_bytes(const tsiir_sosfilt_desc* d, bool saved);

// Coefficient tables (see sos_common.h) for n_rows (filter, section) rows.
// Also advances the workspace epoch by n_epochs (one per following launch
// that publishes scan status words).  With a gate (see FwdArgs::gate) it
// does nothing unless (*gate != 0) == gate_on.
struct WsHeader;
void launch_prep(const void* sos, int coef_dtype, int64_t n_rows, int log2w, bool adjoint,
                 double* tables, WsHeader* hdr, uint32_t n_epochs, hipStream_t stream,
                 const uint32_t* gate = nullptr, uint32_t gate_on = 0);

// tsiir_sosfilt_forward, optionally also writing the outputs of sections
// 0..S-2 to `saved` ((S-1, channels, T) in the I/O type).
// Launch details of a forward that the descriptor does not carry: a device
// word that decides whether its kernels run (see FwdArgs::gate), and the
// two-level row addressing of the segmented form (FwdArgs::seg_n).
// Also, for the FP32-recurrence kernels: tables already built for this call
// (no table preparation is launched; the status words use the epoch
// `epoch_back` launches before the current one), and the segmented form's
// fused warm-up (FwdArgs::warm_fused), with zi in FP64 when zi_f64.
struct FwdGate {
  const uint32_t* word;
  uint32_t on;
  int64_t seg_n = 0;
  int64_t x_seg_len = 0;
  int64_t y_seg_len = 0;
  bool tables_ready = false;
  uint32_t epoch_back = 0;
  bool warm_fused = false;
  bool zi_f64 = false;
};

tsiir_status forward_impl(const tsiir_sosfilt_desc* d, const void* x, void* y, const void* sos,
                          const void* zi, void* zf, void* workspace, size_t workspace_bytes,
                          void* stream, void* saved, const FwdGate* gate = nullptr);

// FP32-recurrence forward (sosfilt_fwd32.hip): unless `tables_ready`, builds
// the FP64 tables of a.tables (for 64-sample lanes) and the Fast32 tables and
// advances the workspace epoch; then filters every full window
// (fwd32_window() samples) of every channel (chained windows).
struct FwdArgs;
struct Fast32;
int64_t fwd32_window();
tsiir_status launch_fwd32_full(const FwdArgs& a, const void* sos, int coef_dtype,
                               int64_t n_filters, Fast32* fast, bool tables_ready,
                               hipStream_t stream);
// The table preparation of launch_fwd32_full alone (advancing the epoch by
// n_epochs), ungated.
void launch_prep32(const void* sos, int coef_dtype, int64_t n_filters, int S, double* tables,
                   Fast32* fast, WsHeader* hdr, uint32_t n_epochs, hipStream_t stream);

}  // namespace tsiir
