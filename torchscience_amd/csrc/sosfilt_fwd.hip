// sosfilt_fwd.hip - single-pass SOS cascade over (channels x time) on gfx950.
//
// Replaces the section loop of sosfilt (_sosfilt.py:117-128) and the
// per-sample DF2T loop of lfilter (_lfilter.py:153-166).
//
// Work decomposition
//   One wavefront owns one channel over a window of 64*W samples; lane q owns
//   the W consecutive samples [q*W, (q+1)*W) of that window and keeps them in
//   registers for the whole cascade.  Windows of a channel are chained by a
//   decoupled look-back, so x is read from HBM once and y written once.
//
// Per section s (sections in order, each fed the exact output of s-1):
//   1. zero-state DF2T pass over the lane's W samples (y_zs, end state e_q);
//   2. wave-level Kogge-Stone scan of the affine maps z -> M z + e_q across
//      the 64 lanes, M = A^W, in FP64 (A is the section's 2x2 state matrix);
//   3. publish the window aggregate, look back over earlier windows of the
//      same channel (64 predecessors per round trip, combined with a
//      shuffle-down reduction using A^(64W*2^i)), publish the inclusive state;
//   4. each lane's true incoming state s_q = M^q s_window + exclusive_q, and the
//      zero-input response from s_q is added to y_zs (a 2-term homogeneous
//      DF2T recurrence), which makes the section output exact.
//
// Tickets taken with an atomic counter order the waves (window-major) so a
// wave only ever waits on windows that were issued before it.

#include <stdlib.h>
#include <string.h>

#include "sos_common.h"
#include "sos_internal.h"
#include "tsiir.h"

namespace tsiir {

struct FwdArgs {
  const void* x;
  void* y;
  const void* zi;
  void* zf;
  const double* tables;
  uint64_t* status;
  WsHeader* hdr;
  int64_t B, T;
  int64_t x_rows, x_stride, y_stride, cpf;
  int64_t n_windows;
  void* saved;  // optional (S-1, B, T) section outputs, for the backward
  int S;
  uint32_t flags;
};

// Coefficients of one section, in the compute type.
template <typename TACC>
struct Coef {
  TACC b0, b1, b2, a1, a2;
  __device__ __forceinline__ explicit Coef(const double* tb)
      : b0(static_cast<TACC>(tb[0])), b1(static_cast<TACC>(tb[1])),
        b2(static_cast<TACC>(tb[2])), a1(static_cast<TACC>(tb[3])),
        a2(static_cast<TACC>(tb[4])) {}
};

// Full windows.  Window I/O is LDS-staged (coalesced) when LDS is set and the
// rows are 16-byte aligned, otherwise per-lane 16-byte accesses.
template <typename TIO, typename TACC, int W, bool LDS>
__global__ __launch_bounds__(64) void sos_fwd_kernel(FwdArgs a,
                                                     const double* __restrict__ tables) {
  __shared__ __align__(16) TIO lds[LDS ? LdsStage<TIO, W>::kElems : 8];
  const int lane = threadIdx.x;
  // Window-major order: workgroup i = k * B + c.  Workgroups are dispatched in
  // index order, so a wave only waits on windows dispatched before it; the
  // bounded spin in lookback() ends every wait regardless.
  const uint32_t flag = launch_flag(a.hdr, 0);
  const int64_t k = static_cast<int64_t>(blockIdx.x) / a.B;
  const int64_t c = static_cast<int64_t>(blockIdx.x) % a.B;
  const int64_t win = static_cast<int64_t>(kWave) * W;
  const int64_t t0 = k * win + static_cast<int64_t>(lane) * W;

  const TIO* xr = static_cast<const TIO*>(a.x) + (c % a.x_rows) * a.x_stride;
  TIO* yr = static_cast<TIO*>(a.y) + c * a.y_stride;
  const bool xvec = (reinterpret_cast<uintptr_t>(xr) & 15u) == 0;
  const bool yvec = (reinterpret_cast<uintptr_t>(yr) & 15u) == 0;

  TACC v[W];
  if (LDS && xvec)
    load_window_lds<TIO, TACC, W>(xr + k * win, lds, v);
  else
    load_run<TIO, TACC, W>(xr + t0, xvec, W, v);

  const double* tf = tables + (c / a.cpf) * static_cast<int64_t>(a.S) * kTableStride;
  const bool has_zi = (a.flags & TSIIR_HAS_ZI) != 0;
  // zf comes from here only when the signal ends exactly on this window
  const bool write_zf = (a.flags & TSIIR_WRITE_ZF) != 0 && (t0 + W == a.T);
  const int64_t zi_row = (a.flags & TSIIR_ZI_BROADCAST) ? 0 : c;

  // The zero-input response that corrects section s is applied inside the
  // pass of section s+1 (one sweep over the registers per section).
  TACC pw0 = 0, pw1 = 0, pa1 = 0, pa2 = 0;
  for (int s = 0; s < a.S; ++s) {
    const double* tb = tf + static_cast<int64_t>(s) * kTableStride;
    const Coef<TACC> q(tb);

    // 1. correction of section s-1, then the zero-state DF2T pass of section s
    //    (the update order of _lfilter.py:158-166, with fused multiply-adds)
    TACC z0 = 0, z1 = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const TACC cj = pw0;
      pw0 = pw1 - pa1 * cj;
      pw1 = -pa2 * cj;
      const TACC u = v[j] + cj;
      const TACC yv = q.b0 * u + z0;
      z0 = q.b1 * u + (z1 - q.a1 * yv);
      z1 = q.b2 * u + (-q.a2 * yv);
      v[j] = yv;
    }

    // 2. wave scan of the lanes' end states
    double agg0, agg1, x0, x1;
    wave_scan(tb, lane, static_cast<double>(z0), static_cast<double>(z1), agg0, agg1, x0, x1);

    // 3. publish aggregate, look back, publish inclusive
    uint64_t* st = a.status + ((k * a.B + c) * a.S + s) * kStatusWords;
    publish_pair(st, agg0, agg1, flag, lane);
    double sw0 = 0.0, sw1 = 0.0;
    if (k == 0) {
      if (has_zi) {
        const TACC* zp = static_cast<const TACC*>(a.zi) + (zi_row * a.S + s) * 2;
        sw0 = static_cast<double>(zp[0]);
        sw1 = static_cast<double>(zp[1]);
      }
    } else {
      lookback(a.status, a.B, a.S, a.hdr, k, c, s, flag, tb, 0.0, 0.0, sw0, sw1);
    }
    {
      const Mat2 Mw = load_mat(table_pow(tb, 6));
      publish_pair(st + 4, Mw.a * sw0 + Mw.b * sw1 + agg0, Mw.c * sw0 + Mw.d * sw1 + agg1,
                   flag, lane);
    }

    // 4. the lane's true incoming state; its zero-input response is pending
    double s0, s1;
    lane_incoming(tb, lane, sw0, sw1, x0, x1, s0, s1);
    pw0 = static_cast<TACC>(s0);
    pw1 = static_cast<TACC>(s1);
    pa1 = q.a1;
    pa2 = q.a2;
    if (write_zf) {
      // state after the window's last sample: zero-state part + A^W s_q,
      // A^W s_q = M s_q with M = table power 0
      const Mat2 M = load_mat(table_pow(tb, 0));
      TACC* zp = static_cast<TACC*>(a.zf) + (c * a.S + s) * 2;
      zp[0] = static_cast<TACC>(static_cast<double>(z0) + M.a * s0 + M.b * s1);
      zp[1] = static_cast<TACC>(static_cast<double>(z1) + M.c * s0 + M.d * s1);
    }
    if (a.saved != nullptr && s + 1 < a.S) {
      // make the registers exact now (the pending correction becomes zero)
      // and keep section s's output for the backward pass
#pragma unroll
      for (int j = 0; j < W; ++j) {
        const TACC cj = pw0;
        pw0 = pw1 - pa1 * cj;
        pw1 = -pa2 * cj;
        v[j] += cj;
      }
      pw0 = 0;
      pw1 = 0;
      TIO* sr = static_cast<TIO*>(a.saved) + (static_cast<int64_t>(s) * a.B + c) * a.T;
      const bool svec = (reinterpret_cast<uintptr_t>(sr) & 15u) == 0;
      if (LDS && svec)
        store_window_lds<TIO, TACC, W>(sr + k * win, lds, v);
      else
        store_run<TIO, TACC, W>(sr + t0, svec, W, v);
    }
  }
  {
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const TACC cj = pw0;
      pw0 = pw1 - pa1 * cj;
      pw1 = -pa2 * cj;
      v[j] += cj;
    }
  }

  if (LDS && yvec)
    store_window_lds<TIO, TACC, W>(yr + k * win, lds, v);
  else
    store_run<TIO, TACC, W>(yr + t0, yvec, W, v);
}

// The partial last window of every channel (launched after the full-window
// kernel, one wave per channel).  Its incoming state is the inclusive state
// the full-window kernel published for window k-1 (or zi when k == 0); lane
// states freeze after the last real sample so zf is exact.
template <typename TIO, typename TACC, int W>
__global__ __launch_bounds__(64) void sos_fwd_tail_kernel(FwdArgs a, int64_t k,
                                                          const double* __restrict__ tables) {
  const int lane = threadIdx.x;
  const int64_t c = blockIdx.x;
  const uint32_t flag = launch_flag(a.hdr, 0);
  const int64_t win = static_cast<int64_t>(kWave) * W;
  const int64_t t0 = k * win + static_cast<int64_t>(lane) * W;
  const int64_t rem = a.T - t0;
  const int m = rem <= 0 ? 0 : (rem >= W ? W : static_cast<int>(rem));
  const TIO* xr = static_cast<const TIO*>(a.x) + (c % a.x_rows) * a.x_stride;
  TIO* yr = static_cast<TIO*>(a.y) + c * a.y_stride;
  TACC v[W];
  load_run<TIO, TACC, W>(xr + t0, false, m, v);
  const double* tf = tables + (c / a.cpf) * static_cast<int64_t>(a.S) * kTableStride;
  const bool has_zi = (a.flags & TSIIR_HAS_ZI) != 0;
  const bool write_zf = (a.flags & TSIIR_WRITE_ZF) != 0 && (t0 < a.T) && (t0 + W >= a.T);
  const int64_t zi_row = (a.flags & TSIIR_ZI_BROADCAST) ? 0 : c;
  for (int s = 0; s < a.S; ++s) {
    const double* tb = tf + static_cast<int64_t>(s) * kTableStride;
    const Coef<TACC> q(tb);
    TACC z0 = 0, z1 = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (j < m) {
        const TACC u = v[j];
        const TACC yv = q.b0 * u + z0;
        z0 = (q.b1 * u - q.a1 * yv) + z1;
        z1 = q.b2 * u - q.a2 * yv;
        v[j] = yv;
      }
    }
    double agg0, agg1, x0, x1;
    wave_scan(tb, lane, static_cast<double>(z0), static_cast<double>(z1), agg0, agg1, x0, x1);
    double sw0 = 0.0, sw1 = 0.0;
    if (k == 0) {
      if (has_zi) {
        const TACC* zp = static_cast<const TACC*>(a.zi) + (zi_row * a.S + s) * 2;
        sw0 = static_cast<double>(zp[0]);
        sw1 = static_cast<double>(zp[1]);
      }
    } else {
      const uint64_t* st = a.status + (((k - 1) * a.B + c) * a.S + s) * kStatusWords;
      if (!read_pair(st + 4, flag, sw0, sw1)) {
        if (lane == 0) atomicOr(&a.hdr->error, 2u);
      }
    }
    double s0, s1;
    lane_incoming(tb, lane, sw0, sw1, x0, x1, s0, s1);
    TACC w0 = static_cast<TACC>(s0), w1 = static_cast<TACC>(s1);
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (j < m) {
        const TACC cj = w0;
        v[j] += cj;
        w0 = w1 - q.a1 * cj;
        w1 = -q.a2 * cj;
      }
    }
    if (write_zf) {
      TACC* zp = static_cast<TACC*>(a.zf) + (c * a.S + s) * 2;
      zp[0] = w0 + z0;
      zp[1] = w1 + z1;
    }
    if (a.saved != nullptr && s + 1 < a.S) {
      TIO* sr = static_cast<TIO*>(a.saved) + (static_cast<int64_t>(s) * a.B + c) * a.T;
      store_run<TIO, TACC, W>(sr + t0, false, m, v);
    }
  }
  store_run<TIO, TACC, W>(yr + t0, false, m, v);
}

// Per (filter, section): normalised coefficients and P^(W*2^i), i < kNPow,
// where P is the forward DF2T state matrix A = [[-a1, 1], [-a2, 0]] or, for
// the adjoint recursion, A' = [[-a1, -a2], [1, 0]].
template <typename TC>
__global__ __launch_bounds__(256) void sos_prep_kernel(const TC* sos, int64_t n_rows, int log2w,
                                                       int adjoint, double* tables,
                                                       WsHeader* hdr, uint32_t n_epochs) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0 && hdr != nullptr) hdr->epoch += n_epochs;
  if (i >= n_rows) return;
  const TC* r = sos + i * 6;
  TC b0 = r[0], b1 = r[1], b2 = r[2], a0 = r[3], a1 = r[4], a2 = r[5];
  if (a0 != TC(1)) {
    b0 = b0 / a0;
    b1 = b1 / a0;
    b2 = b2 / a0;
    a1 = a1 / a0;
    a2 = a2 / a0;
  }
  double* t = tables + i * kTableStride;
  t[0] = static_cast<double>(b0);
  t[1] = static_cast<double>(b1);
  t[2] = static_cast<double>(b2);
  t[3] = static_cast<double>(a1);
  t[4] = static_cast<double>(a2);
  t[5] = static_cast<double>(a0);
  t[6] = t[7] = 0.0;
  const double da1 = static_cast<double>(a1), da2 = static_cast<double>(a2);
  Mat2 A = adjoint ? Mat2{-da1, -da2, 1.0, 0.0} : Mat2{-da1, 1.0, -da2, 0.0};
  for (int l = 0; l < log2w; ++l) A = mat_mul(A, A);
  for (int p = 0; p < kNPow; ++p) {
    double* q = t + 8 + 4 * p;
    q[0] = A.a;
    q[1] = A.b;
    q[2] = A.c;
    q[3] = A.d;
    A = mat_mul(A, A);
  }
  for (int p = 8 + 4 * kNPow; p < kTableStride; ++p) t[p] = 0.0;
}

void launch_prep(const void* sos, int coef_dtype, int64_t n_rows, int log2w, bool adjoint,
                 double* tables, WsHeader* hdr, uint32_t n_epochs, hipStream_t stream) {
  const unsigned blocks = static_cast<unsigned>((n_rows + 255) / 256);
  if (coef_dtype == TSIIR_F64)
    hipLaunchKernelGGL(sos_prep_kernel<double>, dim3(blocks), dim3(256), 0, stream,
                       static_cast<const double*>(sos), n_rows, log2w, adjoint ? 1 : 0, tables,
                       hdr, n_epochs);
  else
    hipLaunchKernelGGL(sos_prep_kernel<float>, dim3(blocks), dim3(256), 0, stream,
                       static_cast<const float*>(sos), n_rows, log2w, adjoint ? 1 : 0, tables,
                       hdr, n_epochs);
}

// zf = zi (or zeros) for an empty time axis.
template <typename TACC>
__global__ void copy_state_kernel(const TACC* zi, TACC* zf, int64_t B, int S, int bcast) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= B * S * 2) return;
  const int64_t within = i % (S * 2);
  const int64_t row = i / (S * 2);
  zf[i] = zi ? zi[(bcast ? 0 : row) * S * 2 + within] : TACC(0);
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static inline int samples_per_lane(int io_dtype) { return io_dtype == TSIIR_F64 ? 32 : 64; }


bool fwd_geometry(const tsiir_sosfilt_desc* d, FwdGeometry* g) {
  if (!d || d->channels < 0 || d->n_samples < 0 || d->n_sections < 1) return false;
  if (d->chans_per_filter < 1 || d->x_rows < 1) return false;
  g->W = samples_per_lane(d->io_dtype);
  const int64_t win = static_cast<int64_t>(kWave) * g->W;
  g->n_windows = (d->n_samples + win - 1) / win;
  g->n_filters = d->channels == 0 ? 0 : (d->channels + d->chans_per_filter - 1) / d->chans_per_filter;
  g->table_bytes = align_up(static_cast<size_t>(g->n_filters) * d->n_sections * kTableStride * 8, 256);
  g->status_bytes = static_cast<size_t>(g->n_windows) * d->channels * d->n_sections *
                    kStatusWords * 8;
  g->total = sizeof(WsHeader) + g->table_bytes + g->status_bytes;
  return true;
}

// Window I/O through LDS (coalesced) unless TSIIR_FWD_IO=direct selects the
// per-lane 16-byte accesses (kept for comparison measurements).
static bool fwd_lds_io() {
  static const bool lds = [] {
    const char* e = getenv("TSIIR_FWD_IO");
    return !(e && strcmp(e, "direct") == 0);
  }();
  return lds;
}

template <typename TIO, typename TACC, int W>
static tsiir_status launch_fwd(const tsiir_sosfilt_desc* d, const FwdGeometry& g, const void* x,
                               void* y, const void* sos, const void* zi, void* zf, char* ws,
                               hipStream_t stream, void* saved = nullptr) {
  WsHeader* hdr = reinterpret_cast<WsHeader*>(ws);
  double* tables = reinterpret_cast<double*>(ws + sizeof(WsHeader));
  uint64_t* status = reinterpret_cast<uint64_t*>(ws + sizeof(WsHeader) + g.table_bytes);
  const int64_t n_rows = g.n_filters * d->n_sections;
  const int log2w = (W == 64) ? 6 : 5;
  launch_prep(sos, d->coef_dtype, n_rows, log2w, false, tables, hdr, 1, stream);
  FwdArgs a;
  a.x = x;
  a.y = y;
  a.zi = zi;
  a.zf = zf;
  a.tables = tables;
  a.status = status;
  a.hdr = hdr;
  a.B = d->channels;
  a.T = d->n_samples;
  a.x_rows = d->x_rows;
  a.x_stride = d->x_row_stride;
  a.y_stride = d->y_row_stride;
  a.cpf = d->chans_per_filter;
  a.n_windows = g.n_windows;
  a.S = d->n_sections;
  a.flags = d->flags;
  a.saved = saved;
  const int64_t win = static_cast<int64_t>(kWave) * W;
  const int64_t full = d->n_samples / win;
  const int64_t waves = full * d->channels;
  if (waves > 0) {
    if (fwd_lds_io())
      hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, W, true>), dim3(static_cast<unsigned>(waves)),
                         dim3(kWave), 0, stream, a, a.tables);
    else
      hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, W, false>),
                         dim3(static_cast<unsigned>(waves)), dim3(kWave), 0, stream, a, a.tables);
  }
  if (full * win < d->n_samples)
    hipLaunchKernelGGL((sos_fwd_tail_kernel<TIO, TACC, W>),
                       dim3(static_cast<unsigned>(d->channels)), dim3(kWave), 0, stream, a, full,
                       a.tables);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

tsiir_status forward_impl(const tsiir_sosfilt_desc* d, const void* x, void* y, const void* sos,
                          const void* zi, void* zf, void* workspace, size_t workspace_bytes,
                          void* stream, void* saved) {
  FwdGeometry g;
  if (!fwd_geometry(d, &g)) return TSIIR_EINVAL;
  if (d->io_dtype != TSIIR_F32 && d->io_dtype != TSIIR_F64 && d->io_dtype != TSIIR_BF16)
    return TSIIR_EINVAL;
  if (d->coef_dtype != TSIIR_F32 && d->coef_dtype != TSIIR_F64) return TSIIR_EINVAL;
  if (d->mode != TSIIR_MODE_PLAIN) return TSIIR_EUNSUPPORTED;
  if (d->flags & TSIIR_REVERSE) return TSIIR_EUNSUPPORTED;
  const bool has_zi = (d->flags & TSIIR_HAS_ZI) != 0;
  const bool write_zf = (d->flags & TSIIR_WRITE_ZF) != 0;
  if (has_zi && !zi) return TSIIR_EINVAL;
  if (write_zf && !zf) return TSIIR_EINVAL;
  if (d->channels == 0) return TSIIR_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool f64 = d->io_dtype == TSIIR_F64;
  if (d->n_samples == 0) {
    if (!write_zf) return TSIIR_OK;
    const int64_t n = d->channels * d->n_sections * 2;
    const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
    const int bc = (d->flags & TSIIR_ZI_BROADCAST) ? 1 : 0;
    if (f64)
      hipLaunchKernelGGL(copy_state_kernel<double>, dim3(blocks), dim3(256), 0, st,
                         has_zi ? static_cast<const double*>(zi) : nullptr,
                         static_cast<double*>(zf), d->channels, d->n_sections, bc);
    else
      hipLaunchKernelGGL(copy_state_kernel<float>, dim3(blocks), dim3(256), 0, st,
                         has_zi ? static_cast<const float*>(zi) : nullptr,
                         static_cast<float*>(zf), d->channels, d->n_sections, bc);
    return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
  }
  if (!x || !y || !sos) return TSIIR_EINVAL;
  if (d->x_row_stride < d->n_samples && d->x_rows > 1) return TSIIR_EINVAL;
  if (d->y_row_stride < d->n_samples && d->channels > 1) return TSIIR_EINVAL;
  if (!workspace || workspace_bytes < g.total) return TSIIR_EWORKSPACE;
  if (g.n_windows * d->channels > 0x7fffffffLL) return TSIIR_EUNSUPPORTED;
  char* ws = static_cast<char*>(workspace);
  switch (d->io_dtype) {
    case TSIIR_F32:
      return launch_fwd<float, float, 64>(d, g, x, y, sos, zi, zf, ws, st, saved);
    case TSIIR_BF16:
      return launch_fwd<bf16_t, float, 64>(d, g, x, y, sos, zi, zf, ws, st, saved);
    default:
      return launch_fwd<double, double, 32>(d, g, x, y, sos, zi, zf, ws, st, saved);
  }
}


}  // namespace tsiir

using namespace tsiir;

extern "C" {

int tsiir_version(void) { return TSIIR_ABI_VERSION; }

const char* tsiir_status_string(tsiir_status s) {
  switch (s) {
    case TSIIR_OK: return "ok";
    case TSIIR_EINVAL: return "invalid argument";
    case TSIIR_EUNSUPPORTED: return "unsupported request";
    case TSIIR_ELAUNCH: return "kernel launch failed";
    case TSIIR_EWORKSPACE: return "workspace missing or too small";
  }
  return "unknown status";
}

size_t tsiir_sosfilt_workspace_bytes(const tsiir_sosfilt_desc* d) {
  FwdGeometry g;
  if (!fwd_geometry(d, &g)) return 0;
  return g.total;
}

tsiir_status tsiir_workspace_init(void* workspace, size_t bytes, void* stream) {
  if (!workspace || bytes < sizeof(WsHeader) || (bytes % 4) != 0) return TSIIR_EINVAL;
  const size_t n = bytes / 4;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<uint32_t*>(workspace), n);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

tsiir_status tsiir_sosfilt_forward(const tsiir_sosfilt_desc* d, const void* x, void* y,
                                   const void* sos, const void* zi, void* zf, void* workspace,
                                   size_t workspace_bytes, void* stream) {
  return forward_impl(d, x, y, sos, zi, zf, workspace, workspace_bytes, stream, nullptr);
}

}  // extern "C"
