// torch_ops.cpp - PyTorch operator registration of the sosfilt path.
//
// The reference registers native operators with TORCH_LIBRARY
// (src/torchscience/csrc/torchscience.cpp:811) and one implementation per
// dispatch key, among them a Meta key for shape inference
// (src/torchscience/csrc/meta/signal_processing/filter.h:14-34,77-92).  Its
// sosfilt itself is Python (_sosfilt.py:13-133, _lfilter.py:153-166); here
// that path becomes these operators:
//
//   tsiir::sosfilt_forward   y, zf, saved section outputs
//   tsiir::sosfilt_backward  dx, dsos, dzi from the saved outputs
//   tsiir::sosfiltfilt       zero-phase filtering (_sosfiltfilt.py:14-173)
//
// with a CUDA-key implementation (HIP devices on ROCm carry PyTorch's CUDA
// dispatch key) that calls the C ABI of libtsiir (include/tsiir.h) on the
// current HIP stream, and a Meta implementation (shapes only) so FakeTensor
// and torch.compile trace through them.  Autograd stays in Python (a
// torch.autograd.Function calling these operators).
//
// The forward's workspace is owned here, one per (device, stream, kind),
// zero-initialised once and registered with a pinned host error mailbox:
// a later call on that stream that finds the mailbox set raises (a
// look-back wait gave up and left NaN outputs), without synchronising.

#include <ATen/ATen.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>
#include <c10/core/DeviceGuard.h>
#include <torch/library.h>

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "tsiir.h"

namespace {

void check_status(tsiir_status st, const char* what) {
  TORCH_CHECK(st == TSIIR_OK, what, " failed: ", tsiir_status_string(st), " (status ",
              static_cast<int>(st), ")");
}

int io_code(at::ScalarType t) {
  switch (t) {
    case at::kFloat: return TSIIR_F32;
    case at::kDouble: return TSIIR_F64;
    case at::kBFloat16: return TSIIR_BF16;
    default: TORCH_CHECK(false, "tsiir: unsupported dtype ", t);
  }
  return TSIIR_F32;
}

at::ScalarType acc_type(at::ScalarType io) { return io == at::kDouble ? at::kDouble : at::kFloat; }

// ---- workspaces ----------------------------------------------------------

struct WsEntry {
  at::Tensor buf;
  at::Tensor mailbox;  // pinned host int32
};
std::mutex g_ws_mu;
std::map<std::tuple<int, hipStream_t, int>, WsEntry> g_ws;

// What a workspace error code means (the bits of the header's error word).
std::string error_text(int code) {
  if (code == static_cast<int>(TSIIR_ERR_STEADY_STATE))
    return "sosfiltfilt: a filter has no steady state (a section's system for sosfilt_zi is "
           "singular; error 4): its outputs are NaN";
  return "a look-back wait gave up (error " + std::to_string(code) +
         "): the outputs of the affected windows are NaN";
}

int mailbox_value(const WsEntry& e) {
  return *static_cast<volatile const int32_t*>(e.mailbox.data_ptr());
}

at::Tensor workspace(const at::Device& dev, hipStream_t stream, int kind, size_t nbytes) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  const auto key = std::make_tuple(static_cast<int>(dev.index()), stream, kind);
  auto it = g_ws.find(key);
  if (it != g_ws.end()) {
    const int code = mailbox_value(it->second);
    if (code != 0) {
      g_ws.erase(it);
      TORCH_CHECK(false, "tsiir: an earlier call on this stream failed: ", error_text(code));
    }
    if (static_cast<size_t>(it->second.buf.numel()) >= nbytes) return it->second.buf;
  }
  size_t size = std::max(nbytes, static_cast<size_t>(1) << 20);
  size = (size + 255) / 256 * 256;
  at::Tensor buf =
      at::empty({static_cast<int64_t>(size)}, at::TensorOptions().dtype(at::kByte).device(dev));
  check_status(tsiir_workspace_init(buf.data_ptr(), size, stream), "tsiir_workspace_init");
  at::Tensor mb = at::zeros({1}, at::TensorOptions().dtype(at::kInt).pinned_memory(true));
  check_status(tsiir_workspace_set_error_mailbox(buf.data_ptr(), size,
                                                 static_cast<uint32_t*>(mb.data_ptr()), stream),
               "tsiir_workspace_set_error_mailbox");
  g_ws[key] = WsEntry{buf, mb};
  return buf;
}

// After the caller has synchronised: the first error a mailbox holds (and
// that workspace is dropped), else 0.
int64_t take_error() {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  for (auto it = g_ws.begin(); it != g_ws.end(); ++it) {
    const int code = mailbox_value(it->second);
    if (code != 0) {
      g_ws.erase(it);
      return code;
    }
  }
  return 0;
}

// ---- routing of FP32 recurrences ----------------------------------------

// Largest pole radius over the sections of sos (..., 6): the roots of
// a0 z^2 + a1 z + a2 (infinite for a0 == 0 or non-finite coefficients).
double max_pole_radius(const at::Tensor& sos) {
  at::Tensor a = sos.detach().to(at::kCPU, at::kDouble).contiguous().reshape({-1, 6});
  const double* p = a.data_ptr<double>();
  double rmax = 0.0;
  for (int64_t i = 0; i < a.size(0); ++i) {
    const double a0 = p[6 * i + 3], a1 = p[6 * i + 4], a2 = p[6 * i + 5];
    double r;
    if (a0 == 0.0 || !std::isfinite(a0) || !std::isfinite(a1) || !std::isfinite(a2)) {
      r = INFINITY;
    } else {
      const double disc = a1 * a1 - 4.0 * a0 * a2;
      if (disc < 0.0) {
        r = std::sqrt(std::fabs(a2 / a0));
      } else {
        const double sq = std::sqrt(disc);
        r = std::max(std::fabs((-a1 + sq) / (2.0 * a0)), std::fabs((-a1 - sq) / (2.0 * a0)));
      }
    }
    rmax = std::max(rmax, r);
  }
  return rmax;
}

// FP32 recurrences of filters with a pole beyond this radius run in DF2T
// order with FP64 carries (TSIIR_FP32_DF2T): measured on 1024 x 65536 and
// 16 x 2^20 FP32 signals against the FP64 oracle, the direct-form-I kernel
// stays below the reference's own FP32 DF2T error for every Butterworth /
// Chebyshev I filter with 1 - r >= 6e-3 and exceeds it (up to 2x: 4.3e-2
// against 2.0e-2 on the order-16 Chebyshev I at cutoff 0.01, 1 - r =
// 2.75e-4) only below 3.1e-3; the DF2T kernel is below the reference's
// error on all of them.
constexpr double kDf2tPoleRadius = 0.995;

// Device coefficients are read once per (storage, version): an entry holds a
// weak reference to the storage it was computed for and is used only while
// that same storage is alive (the weak reference also keeps its address from
// being reused by another storage), so a tensor allocated where a freed one
// lived is read again.  Writes that bypass the version counter (through
// .data or foreign pointers) are not seen; the answer only picks between two
// correct kernels, so that could cost accuracy on such poles, never
// correctness.
struct RouteEntry {
  c10::weak_intrusive_ptr<c10::StorageImpl> storage;
  bool df2t;
};
std::mutex g_route_mu;
std::map<std::tuple<const void*, int64_t, int64_t, int, int>, RouteEntry> g_route;

bool needs_fp32_df2t(const at::Tensor& sos) {
  if (sos.device().is_cpu()) return max_pole_radius(sos) > kDf2tPoleRadius;
  const auto key = std::make_tuple(sos.data_ptr(), sos._version(), sos.numel(),
                                   static_cast<int>(sos.scalar_type()),
                                   static_cast<int>(sos.device().index()));
  const c10::StorageImpl* impl = sos.storage().unsafeGetStorageImpl();
  {
    std::lock_guard<std::mutex> lk(g_route_mu);
    auto it = g_route.find(key);
    if (it != g_route.end()) {
      if (it->second.storage.lock().get() == impl) return it->second.df2t;
      g_route.erase(it);
    }
  }
  const bool r = max_pole_radius(sos) > kDf2tPoleRadius;
  std::lock_guard<std::mutex> lk(g_route_mu);
  if (g_route.size() > 256) g_route.clear();
  g_route.insert_or_assign(key, RouteEntry{sos.storage().getWeakStorageImpl(), r});
  return r;
}

// ---- operators -----------------------------------------------------------

using Fwd = std::tuple<at::Tensor, at::Tensor, at::Tensor>;

constexpr uint32_t kCallerFlags = TSIIR_PATH_WINDOW | TSIIR_PATH_CHANNEL | TSIIR_REVERSE |
                                  TSIIR_DEBUG_GIVE_UP | TSIIR_NO_SEGMENT;

void check_rows(const at::Tensor& x, const at::Tensor& sos, int64_t channels,
                int64_t chans_per_filter) {
  TORCH_CHECK(x.dim() == 2 && x.stride(1) == 1 && (x.size(0) <= 1 || x.stride(0) >= x.size(1)),
              "tsiir: x must be (rows, samples) with unit time stride");
  TORCH_CHECK(sos.dim() == 3 && sos.size(2) == 6 && sos.is_contiguous(),
              "tsiir: sos must be contiguous (n_filters, n_sections, 6)");
  TORCH_CHECK(channels >= 0 && (chans_per_filter >= 1 || channels == 0),
              "tsiir: bad channel counts");
}

// The state type: double for double samples, and when zi comes in double
// with float or bfloat16 samples (TSIIR_STATE_F64: a caller that carries the
// state in FP64 between chunks); else float.
bool state_f64(const at::Tensor& x, const c10::optional<at::Tensor>& zi, int64_t mode) {
  return x.scalar_type() != at::kDouble && zi.has_value() && zi->scalar_type() == at::kDouble;
}

Fwd forward_shapes(const at::Tensor& x, const at::Tensor& sos, int64_t channels, bool want_zf,
                   bool keep_sections, bool f64_state) {
  const int64_t S = sos.size(1), T = x.size(1);
  const auto acc = x.options().dtype(f64_state ? at::kDouble : acc_type(x.scalar_type()));
  at::Tensor y = at::empty({channels, T}, x.options());
  at::Tensor zf = want_zf ? at::empty({channels, S, 2}, acc) : at::empty({0}, acc);
  at::Tensor saved = (keep_sections && S > 1) ? at::empty({S - 1, channels, T}, x.options())
                                              : at::empty({0}, x.options());
  return {y, zf, saved};
}

Fwd sosfilt_forward_meta(const at::Tensor& x, const at::Tensor& sos,
                         const c10::optional<at::Tensor>& zi, int64_t channels,
                         int64_t chans_per_filter, bool want_zf, bool zi_broadcast, int64_t mode,
                         int64_t path_flags, bool keep_sections) {
  check_rows(x, sos, channels, chans_per_filter);
  return forward_shapes(x, sos, channels, want_zf, keep_sections, state_f64(x, zi, mode));
}

tsiir_sosfilt_desc make_desc(int64_t channels, int64_t T, int64_t S, at::ScalarType io,
                             at::ScalarType coef, int64_t mode, uint32_t flags, int64_t x_rows,
                             int64_t x_stride, int64_t cpf) {
  tsiir_sosfilt_desc d{};
  d.channels = channels;
  d.n_samples = T;
  d.n_sections = static_cast<int32_t>(S);
  d.io_dtype = io_code(io);
  d.coef_dtype = io_code(coef);
  d.mode = static_cast<int32_t>(mode);
  d.flags = flags;
  d.reserved = 0;
  d.x_rows = std::max<int64_t>(x_rows, 1);
  d.x_row_stride = x_stride;
  d.y_row_stride = T;
  d.chans_per_filter = std::max<int64_t>(cpf, 1);
  return d;
}

Fwd sosfilt_forward_hip(const at::Tensor& x, const at::Tensor& sos,
                        const c10::optional<at::Tensor>& zi, int64_t channels,
                        int64_t chans_per_filter, bool want_zf, bool zi_broadcast, int64_t mode,
                        int64_t path_flags, bool keep_sections) {
  check_rows(x, sos, channels, chans_per_filter);
  c10::DeviceGuard guard(x.device());
  TORCH_CHECK(sos.device() == x.device(), "tsiir: sos on ", sos.device(), ", x on ", x.device());
  const at::ScalarType io = x.scalar_type();
  const int64_t S = sos.size(1), T = x.size(1);
  const bool f64_state = state_f64(x, zi, mode);
  Fwd out = forward_shapes(x, sos, channels, want_zf, keep_sections, f64_state);
  if (channels == 0) return out;
  at::Tensor& y = std::get<0>(out);
  at::Tensor& zf = std::get<1>(out);
  at::Tensor& saved = std::get<2>(out);
  uint32_t flags = static_cast<uint32_t>(path_flags) & kCallerFlags;
  if (zi.has_value()) {
    TORCH_CHECK(zi->device() == x.device() && zi->is_contiguous() &&
                    (zi->scalar_type() == acc_type(io) || f64_state),
                "tsiir: zi must be contiguous, on x's device, in the compute type or in "
                "double");
    TORCH_CHECK(zi->numel() == (zi_broadcast ? 1 : channels) * S * 2,
                "tsiir: zi must hold ", zi_broadcast ? 1 : channels, " x ", S, " x 2 values");
    flags |= TSIIR_HAS_ZI | (zi_broadcast ? TSIIR_ZI_BROADCAST : 0u);
  }
  if (want_zf) flags |= TSIIR_WRITE_ZF;
  if (f64_state) flags |= TSIIR_STATE_F64;
  if (io != at::kDouble && mode == TSIIR_MODE_PLAIN && needs_fp32_df2t(sos))
    flags |= TSIIR_FP32_DF2T;
  const tsiir_sosfilt_desc d = make_desc(channels, T, S, io, sos.scalar_type(), mode, flags,
                                         x.size(0), x.size(0) > 1 ? x.stride(0) : T,
                                         chans_per_filter);
  const size_t nbytes = tsiir_sosfilt_workspace_bytes(&d);
  TORCH_CHECK(nbytes > 0, "tsiir: invalid sosfilt descriptor");
  hipStream_t stream = at::hip::getCurrentHIPStreamMasqueradingAsCUDA(x.device().index()).stream();
  at::Tensor ws = workspace(x.device(), stream, 0, nbytes);
  const void* zp = zi.has_value() ? zi->data_ptr() : nullptr;
  void* zfp = want_zf ? zf.data_ptr() : nullptr;
  if (keep_sections) {
    check_status(tsiir_sosfilt_forward_saved(&d, x.data_ptr(), y.data_ptr(), sos.data_ptr(), zp,
                                             zfp, S > 1 ? saved.data_ptr() : nullptr,
                                             ws.data_ptr(), static_cast<size_t>(ws.numel()),
                                             stream),
                 "tsiir_sosfilt_forward_saved");
  } else {
    check_status(tsiir_sosfilt_forward(&d, x.data_ptr(), y.data_ptr(), sos.data_ptr(), zp, zfp,
                                       ws.data_ptr(), static_cast<size_t>(ws.numel()), stream),
                 "tsiir_sosfilt_forward");
  }
  return out;
}

using Bwd = std::tuple<at::Tensor, at::Tensor, at::Tensor>;

Bwd backward_shapes(const at::Tensor& dy, const at::Tensor& sos) {
  const int64_t B = dy.size(0), S = sos.size(1);
  return {at::empty_like(dy), at::empty_like(sos),
          at::empty({B, S, 2}, dy.options().dtype(acc_type(dy.scalar_type())))};
}

Bwd sosfilt_backward_meta(const at::Tensor& x, const at::Tensor& y, const at::Tensor& saved,
                          const at::Tensor& sos, const c10::optional<at::Tensor>& zi,
                          const at::Tensor& dy, const c10::optional<at::Tensor>& dzf,
                          int64_t chans_per_filter, bool zi_broadcast) {
  return backward_shapes(dy, sos);
}

Bwd sosfilt_backward_hip(const at::Tensor& x, const at::Tensor& y, const at::Tensor& saved,
                         const at::Tensor& sos, const c10::optional<at::Tensor>& zi,
                         const at::Tensor& dy, const c10::optional<at::Tensor>& dzf,
                         int64_t chans_per_filter, bool zi_broadcast) {
  c10::DeviceGuard guard(dy.device());
  const int64_t B = dy.size(0), T = dy.size(1), S = sos.size(1);
  if (B == 0) return backward_shapes(dy, sos);
  TORCH_CHECK(x.is_contiguous() && y.is_contiguous() && dy.is_contiguous() &&
                  x.sizes() == dy.sizes() && y.sizes() == dy.sizes(),
              "tsiir: x, y and dy must be contiguous (channels, samples) of one shape");
  TORCH_CHECK(S <= 1 || (saved.is_contiguous() && saved.numel() == (S - 1) * B * T),
              "tsiir: saved must hold the outputs of sections 0..S-2");
  TORCH_CHECK(sos.dim() == 3 && sos.size(2) == 6 && sos.is_contiguous(),
              "tsiir: sos must be contiguous (n_filters, n_sections, 6)");
  const at::ScalarType io = dy.scalar_type();
  uint32_t flags = 0;
  if (zi.has_value()) flags |= TSIIR_HAS_ZI | (zi_broadcast ? TSIIR_ZI_BROADCAST : 0u);
  if (dzf.has_value()) {
    TORCH_CHECK(dzf->is_contiguous() && dzf->numel() == B * S * 2 &&
                    dzf->scalar_type() == acc_type(io),
                "tsiir: dzf must be contiguous (channels, n_sections, 2) in the compute type");
    flags |= TSIIR_WRITE_ZF;
  }
  const tsiir_sosfilt_desc d =
      make_desc(B, T, S, io, sos.scalar_type(), TSIIR_MODE_PLAIN, flags, B, T, chans_per_filter);
  const size_t nbytes = tsiir_sosfilt_backward_saved_workspace_bytes(&d);
  TORCH_CHECK(nbytes > 0, "tsiir: invalid sosfilt backward descriptor");
  hipStream_t stream = at::hip::getCurrentHIPStreamMasqueradingAsCUDA(dy.device().index()).stream();
  at::Tensor ws = workspace(dy.device(), stream, 1, nbytes);
  Bwd out = backward_shapes(dy, sos);
  check_status(tsiir_sosfilt_backward_saved(
                   &d, x.data_ptr(), y.data_ptr(), S > 1 ? saved.data_ptr() : nullptr,
                   sos.data_ptr(), zi.has_value() ? zi->data_ptr() : nullptr, dy.data_ptr(),
                   dzf.has_value() ? dzf->data_ptr() : nullptr, std::get<0>(out).data_ptr(),
                   std::get<1>(out).data_ptr(), std::get<2>(out).data_ptr(), ws.data_ptr(),
                   static_cast<size_t>(ws.numel()), stream),
               "tsiir_sosfilt_backward_saved");
  return out;
}

// Zero-phase filtering of rows x (channels, samples) with one filter sos
// (1, S, 6) and its steady state zi_unit (1, S, 2) in double: the C ABI's
// tsiir_sosfiltfilt, the extensions never stored.  zi_unit None and padlen
// -1 form the steady state and the default padlen on the device (no copy of
// sos to the host; a filter without a steady state is reported through the
// stream's error mailbox).  The sosfilt workspace of the stream is shared
// with sosfilt_forward; the first pass's rows go to a scratch tensor of the
// caching allocator.
at::Tensor sosfiltfilt_meta(const at::Tensor& x, const at::Tensor& sos,
                            const c10::optional<at::Tensor>& zi_unit, int64_t padtype,
                            int64_t padlen) {
  check_rows(x, sos, x.size(0), x.size(0));
  return at::empty({x.size(0), x.size(1)}, x.options());
}

at::Tensor sosfiltfilt_hip(const at::Tensor& x, const at::Tensor& sos,
                           const c10::optional<at::Tensor>& zi_unit, int64_t padtype,
                           int64_t padlen) {
  check_rows(x, sos, x.size(0), x.size(0));
  c10::DeviceGuard guard(x.device());
  TORCH_CHECK(sos.device() == x.device(), "tsiir: sos must be on x's device");
  TORCH_CHECK(sos.size(0) == 1, "tsiir: sosfiltfilt takes one filter");
  const int64_t S = sos.size(1), B = x.size(0), T = x.size(1);
  const double* zu = nullptr;
  if (zi_unit.has_value()) {
    TORCH_CHECK(zi_unit->device() == x.device(), "tsiir: zi_unit must be on x's device");
    TORCH_CHECK(zi_unit->scalar_type() == at::kDouble && zi_unit->is_contiguous() &&
                    zi_unit->numel() == S * 2,
                "tsiir: zi_unit must be contiguous double (1, n_sections, 2)");
    zu = zi_unit->data_ptr<double>();
  }
  const at::ScalarType io = x.scalar_type();
  at::Tensor y = at::empty({B, T}, x.options());
  if (B == 0) return y;
  uint32_t flags = 0;
  if (io != at::kDouble && needs_fp32_df2t(sos)) flags |= TSIIR_FP32_DF2T;
  const tsiir_sosfilt_desc d = make_desc(B, T, S, io, sos.scalar_type(), TSIIR_MODE_PLAIN, flags,
                                         B, B > 1 ? x.stride(0) : T, B);
  const size_t nws = tsiir_sosfilt_workspace_bytes(&d);
  const size_t nsc = tsiir_sosfiltfilt_scratch_bytes(&d, padlen);
  TORCH_CHECK(nws > 0 && nsc > 0, "tsiir: invalid sosfiltfilt descriptor");
  hipStream_t stream = at::hip::getCurrentHIPStreamMasqueradingAsCUDA(x.device().index()).stream();
  at::Tensor ws = workspace(x.device(), stream, 0, nws);
  at::Tensor sc = at::empty({static_cast<int64_t>(nsc)}, x.options().dtype(at::kByte));
  check_status(tsiir_sosfiltfilt(&d, static_cast<int>(padtype), padlen, x.data_ptr(),
                                 y.data_ptr(), sos.data_ptr(), zu,
                                 sc.data_ptr(), static_cast<size_t>(sc.numel()), ws.data_ptr(),
                                 static_cast<size_t>(ws.numel()), stream),
               "tsiir_sosfiltfilt");
  return y;
}

}  // namespace

TORCH_LIBRARY(tsiir, m) {
  m.def(
      "sosfilt_forward(Tensor x, Tensor sos, Tensor? zi, int channels, int chans_per_filter, "
      "bool want_zf, bool zi_broadcast, int mode, int path_flags, bool keep_sections) "
      "-> (Tensor y, Tensor zf, Tensor saved)");
  m.def(
      "sosfilt_backward(Tensor x, Tensor y, Tensor saved, Tensor sos, Tensor? zi, Tensor dy, "
      "Tensor? dzf, int chans_per_filter, bool zi_broadcast) -> (Tensor dx, Tensor dsos, "
      "Tensor dzi)");
  m.def("sosfiltfilt(Tensor x, Tensor sos, Tensor? zi_unit, int padtype, int padlen) -> Tensor");
  m.def("take_error() -> int", &take_error);
  m.def("max_pole_radius(Tensor sos) -> float", &max_pole_radius);
}

TORCH_LIBRARY_IMPL(tsiir, CUDA, m) {
  m.impl("sosfilt_forward", &sosfilt_forward_hip);
  m.impl("sosfilt_backward", &sosfilt_backward_hip);
  m.impl("sosfiltfilt", &sosfiltfilt_hip);
}

TORCH_LIBRARY_IMPL(tsiir, Meta, m) {
  m.impl("sosfilt_forward", &sosfilt_forward_meta);
  m.impl("sosfilt_backward", &sosfilt_backward_meta);
  m.impl("sosfiltfilt", &sosfiltfilt_meta);
}
