// sosfiltfilt.hip - zero-phase filtering (tsiir_sosfiltfilt).
//
// The reference (_sosfiltfilt.py:96-173) concatenates [left extension, x,
// right extension], runs sosfilt forward from zi * xp[0], flips, runs it
// again from zi * y1[-1], flips back and keeps the central samples.  Here
// the extensions are never stored.  Filtering a concatenation is filtering
// its pieces with the state carried across, so:
//
//   1. edge kernel, left: filter the left extension from zi_unit * (its
//      first sample); keep only the final state za;
//   2. sosfilt over the rows from za -> y1 (scratch), final state zb;
//   3. edge kernel, right: filter the right extension from zb (outputs y1r,
//      rounded to the I/O type like the reference's tensor), then y1r
//      backward from zi_unit * y1r[-1]; keep only the final state zc;
//   4. sosfilt over y1 in reverse time from zc -> y.
//
// With zi_unit null and padlen < 0 the edge kernels form the steady state
// and the default padlen from the coefficients on the device, so a call
// needs no copy of them to the host; a filter without a steady state gives
// NaN outputs and sets kErrSteadyState in the workspace's error word and
// mailbox instead of failing the call.
//
// The edge kernels run one thread per channel (the extensions are a few
// dozen samples), every section of a sample before the next, in the DF2T
// order of the window kernels (_lfilter.py:158-166 with fused multiply-adds)
// and the working precision of the I/O type; the extension samples are
// formed in the I/O type as the reference's tensor ops form them
// (_sosfiltfilt.py:198-217).  States pass between the steps in FP64.

#include <type_traits>

#include "sos_common.h"
#include "sos_internal.h"
#include "tsiir.h"

namespace tsiir {
namespace {

constexpr int kEdgeMaxS = 8;
constexpr int64_t kEdgeMaxWork = 16384;  // padlen x sections per thread
constexpr int64_t kEdgeLine = 128;       // bytes: y1 rows start on L2 lines
constexpr uint32_t kErrSteadyState = TSIIR_ERR_STEADY_STATE;

struct EdgeArgs {
  const void* x;            // rows, time-ordered
  int64_t x_rows, x_stride;
  int64_t B, T, cpf;
  int S, padtype;
  int64_t padlen;           // < 0: each filter's default (default_padlen)
  int64_t ybuf_stride;      // elements per channel in ybuf
  const void* sos;          // (n_filters, S, 6), coefficient type
  const double* zi_unit;    // (n_filters, S, 2), or null: steady_state() here
  double* zs;               // zi_unit null: the left kernel's steady states for
                            // the right kernel, (n_filters, S, 2)
  WsHeader* hdr;            // error word and mailbox of the call's workspace
  const double* z_in;       // right edge: the state after the pass over the rows
  double* z_out;            // (B, S, 2)
  void* ybuf;               // right edge: its forward outputs, (B, max(padlen, 1))
  const void* y1;           // right edge with padlen 0: the first pass's rows
  int64_t y1_stride;
};

template <typename TIO>
using io_float_t = typename std::conditional<std::is_same<TIO, double>::value, double, float>::type;

template <typename TIO>
__device__ __forceinline__ io_float_t<TIO> rd(const TIO* p, int64_t i) {
  return IOTraits<TIO>::to_f(p[i]);
}
// v rounded to the I/O type (a value a tensor of that type can hold)
template <typename TIO>
__device__ __forceinline__ io_float_t<TIO> to_io(io_float_t<TIO> v) {
  return IOTraits<TIO>::to_f(IOTraits<TIO>::from_f(v));
}

// Sample k of the left (RIGHT = false) or right extension of row xr
// (_sosfiltfilt.py:198-217): odd 2 x[0] - x[padlen - k] / 2 x[T-1] -
// x[T-2-k], even x[padlen - k] / x[T-2-k], constant x[0] / x[T-1].
template <typename TIO, bool RIGHT>
__device__ __forceinline__ io_float_t<TIO> extension(const TIO* xr, const EdgeArgs& e,
                                                     int64_t padlen, int64_t k) {
  const int64_t edge = RIGHT ? e.T - 1 : 0;
  const int64_t mirror = RIGHT ? e.T - 2 - k : padlen - k;
  const io_float_t<TIO> xe = rd(xr, edge);
  if (e.padtype == TSIIR_PAD_ODD) return to_io<TIO>(io_float_t<TIO>(2) * xe - rd(xr, mirror));
  if (e.padtype == TSIIR_PAD_EVEN) return rd(xr, mirror);
  return xe;
}

template <typename TR>
struct Cascade {
  TR b0[kEdgeMaxS], b1[kEdgeMaxS], b2[kEdgeMaxS], a1[kEdgeMaxS], a2[kEdgeMaxS];
  TR z0[kEdgeMaxS], z1[kEdgeMaxS];
  int S;
  // one sample through every section; returns the last section's output
  __device__ __forceinline__ TR step(TR v) {
#pragma unroll
    for (int s = 0; s < kEdgeMaxS; ++s) {
      if (s < S) {
        const TR yv = fma3(b0[s], v, z0[s]);
        z0[s] = b1[s] * v + (z1[s] - a1[s] * yv);
        z1[s] = b2[s] * v + (-a2[s] * yv);
        v = yv;
      }
    }
    return v;
  }
  // z = zi_unit * first (the steady state for a step of height `first`)
  __device__ __forceinline__ void start(const double* zu, TR first) {
#pragma unroll
    for (int s = 0; s < kEdgeMaxS; ++s) {
      if (s < S) {
        z0[s] = static_cast<TR>(zu[2 * s]) * first;
        z1[s] = static_cast<TR>(zu[2 * s + 1]) * first;
      }
    }
  }
  __device__ __forceinline__ void store(double* zo) const {
#pragma unroll
    for (int s = 0; s < kEdgeMaxS; ++s) {
      if (s < S) {
        zo[2 * s] = static_cast<double>(z0[s]);
        zo[2 * s + 1] = static_cast<double>(z1[s]);
      }
    }
  }
};

// The section coefficients normalised by a0 in the coefficient type, as the
// table preparation of the sosfilt kernels normalises them.
template <typename TR, typename TC>
__device__ __forceinline__ void load_coefficients(Cascade<TR>& q, const TC* f) {
#pragma unroll
  for (int s = 0; s < kEdgeMaxS; ++s) {
    q.b0[s] = q.b1[s] = q.b2[s] = q.a1[s] = q.a2[s] = q.z0[s] = q.z1[s] = TR(0);
    if (s < q.S) {
      const TC* r = f + 6 * s;
      TC b0 = r[0], b1 = r[1], b2 = r[2], a0 = r[3], a1 = r[4], a2 = r[5];
      if (a0 != TC(1)) {
        b0 = b0 / a0;
        b1 = b1 / a0;
        b2 = b2 / a0;
        a1 = a1 / a0;
        a2 = a2 / a0;
      }
      q.b0[s] = static_cast<TR>(static_cast<double>(b0));
      q.b1[s] = static_cast<TR>(static_cast<double>(b1));
      q.b2[s] = static_cast<TR>(static_cast<double>(b2));
      q.a1[s] = static_cast<TR>(static_cast<double>(a1));
      q.a2[s] = static_cast<TR>(static_cast<double>(a2));
    }
  }
}

// The reference's default padlen of one filter (_sosfiltfilt.py:100-110):
// 3 * (2 S + 1 - min(sections with b2 == 0, sections with a2 == 0)).
template <typename TC>
__device__ __forceinline__ int64_t default_padlen(const TC* f, int S) {
  int nb = 0, na = 0;
  for (int s = 0; s < S; ++s) {
    nb += f[6 * s + 2] == TC(0) ? 1 : 0;
    na += f[6 * s + 5] == TC(0) ? 1 : 0;
  }
  return 3 * (2 * S + 1 - (nb < na ? nb : na));
}

// sosfilt_zi of one filter in FP64 (_sosfilt_zi.py:55-86 over
// _lfilter_zi.py:100-128): each section's system [[1 + a1, -1], [a2, 1]] zi
// = [b1 - a1 b0, b2 - a2 b0] (normalised by a0) in closed form, scaled by
// the DC gains sum(b) / sum(a) of the sections before it.  No contraction,
// so the states equal sosfilt_zi_host's bit for bit.  False where the
// reference's solve raises: a singular system (det == 0), or, for a0 == 0
// (whose leading zero it strips, leaving [[1 + a2 / a1, -1], [0, 1]] or the
// identity), sum(a) == 0.  An a0 == 0 section otherwise passes: its states
// are not finite here, but neither is any output of a pass that divides by
// a0, the reference's included.
template <typename TC>
__device__ __forceinline__ bool steady_state(const TC* f, int S, double (&zu)[2 * kEdgeMaxS]) {
#pragma clang fp contract(off)
  double scale = 1.0;
  bool ok = true;
#pragma unroll
  for (int s = 0; s < kEdgeMaxS; ++s) {
    if (s < S) {
      const double b0 = f[6 * s], b1 = f[6 * s + 1], b2 = f[6 * s + 2];
      const double a0 = f[6 * s + 3], a1 = f[6 * s + 4], a2 = f[6 * s + 5];
      const double sa = a0 + a1 + a2;
      double n0 = b0, n1 = b1, n2 = b2, m1 = a1, m2 = a2;  // x / 1 == x
      if (a0 != 1.0) {
        n0 = b0 / a0;
        n1 = b1 / a0;
        n2 = b2 / a0;
        m1 = a1 / a0;
        m2 = a2 / a0;
      }
      const double det = (1.0 + m1) + m2;
      ok = ok && (a0 != 0.0 ? det != 0.0 : sa != 0.0);
      const double r0 = n1 - m1 * n0, r1 = n2 - m2 * n0;
      zu[2 * s] = scale * ((r0 + r1) / det);
      zu[2 * s + 1] = scale * (((1.0 + m1) * r1 - m2 * r0) / det);
      scale = scale * ((b0 + b1 + b2) / sa);
    }
  }
  return ok;
}

// Samples are read kEdgeBlk at a time before they are filtered, so the
// loads of a block are in flight together rather than one per step.
constexpr int kEdgeBlk = 8;

// One thread per channel.  RIGHT = false: step 1 above; RIGHT: step 3.
// ONE: a single filter for every channel, so its coefficients, padlen and
// steady state are uniform (scalar loads, scalar loop control).
template <typename TIO, typename TR, typename TC, bool RIGHT, bool ONE>
__global__ __launch_bounds__(kWave) void sos_edge_kernel(EdgeArgs e) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kWave + threadIdx.x;
  if (c >= e.B) return;
  const int64_t f = ONE ? 0 : c / e.cpf;
  Cascade<TR> q;
  q.S = e.S;
  const TC* fc = static_cast<const TC*>(e.sos) + f * e.S * 6;
  load_coefficients<TR, TC>(q, fc);
  const int64_t padlen = e.padlen >= 0 ? e.padlen : default_padlen(fc, e.S);
  double zu[2 * kEdgeMaxS];
  const double* zsrc = e.zi_unit != nullptr ? e.zi_unit : RIGHT ? e.zs : nullptr;
  if (zsrc != nullptr) {
#pragma unroll
    for (int k = 0; k < 2 * kEdgeMaxS; ++k) zu[k] = k < 2 * e.S ? zsrc[f * e.S * 2 + k] : 0.0;
  } else if (!steady_state(fc, e.S, zu)) {
    // reported through the workspace's error word and mailbox; the states
    // (and so every output of the channel) are NaN
    atomicOr(&e.hdr->error, kErrSteadyState);
    if (e.hdr->mailbox != nullptr)
      __hip_atomic_store(e.hdr->mailbox, kErrSteadyState, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
#pragma unroll
    for (int k = 0; k < 2 * kEdgeMaxS; ++k) zu[k] = __builtin_nan("");
  }
  if (!RIGHT && zsrc == nullptr && c % e.cpf == 0) {
#pragma unroll
    for (int k = 0; k < 2 * kEdgeMaxS; ++k)
      if (k < 2 * e.S) e.zs[f * e.S * 2 + k] = zu[k];
  }
  const TIO* xr = static_cast<const TIO*>(e.x) + (c % e.x_rows) * e.x_stride;
  double* zo = e.z_out + c * e.S * 2;
  if constexpr (!RIGHT) {
    q.start(zu, static_cast<TR>(padlen > 0 ? extension<TIO, false>(xr, e, padlen, 0)
                                           : rd(xr, 0)));
    for (int64_t k0 = 0; k0 < padlen; k0 += kEdgeBlk) {
      TR u[kEdgeBlk];
#pragma unroll
      for (int j = 0; j < kEdgeBlk; ++j)
        u[j] = k0 + j < padlen ? static_cast<TR>(extension<TIO, false>(xr, e, padlen, k0 + j))
                               : TR(0);
#pragma unroll
      for (int j = 0; j < kEdgeBlk; ++j)
        if (k0 + j < padlen) q.step(u[j]);
    }
  } else {
    const double* zi = e.z_in + c * e.S * 2;
#pragma unroll
    for (int s = 0; s < kEdgeMaxS; ++s) {
      if (s < q.S) {
        q.z0[s] = static_cast<TR>(zi[2 * s]);
        q.z1[s] = static_cast<TR>(zi[2 * s + 1]);
      }
    }
    TIO* yb = static_cast<TIO*>(e.ybuf) + c * e.ybuf_stride;
    TR last = padlen > 0 ? TR(0)
                           : static_cast<TR>(rd(static_cast<const TIO*>(e.y1) + c * e.y1_stride,
                                                e.T - 1));
    for (int64_t k0 = 0; k0 < padlen; k0 += kEdgeBlk) {
      TR u[kEdgeBlk];
#pragma unroll
      for (int j = 0; j < kEdgeBlk; ++j)
        u[j] = k0 + j < padlen ? static_cast<TR>(extension<TIO, true>(xr, e, padlen, k0 + j))
                               : TR(0);
#pragma unroll
      for (int j = 0; j < kEdgeBlk; ++j) {
        if (k0 + j < padlen) {
          // the output as the reference's tensor holds it (rounded to TIO)
          const TIO o = IOTraits<TIO>::from_f(static_cast<io_float_t<TIO>>(q.step(u[j])));
          yb[k0 + j] = o;
          last = static_cast<TR>(IOTraits<TIO>::to_f(o));
        }
      }
    }
    q.start(zu, last);
    // backward over the stored outputs, a block of them read ahead
    for (int64_t k1 = padlen; k1 > 0; k1 -= kEdgeBlk) {
      TR u[kEdgeBlk];
#pragma unroll
      for (int j = 0; j < kEdgeBlk; ++j)
        u[j] = k1 - 1 - j >= 0 ? static_cast<TR>(rd(yb, k1 - 1 - j)) : TR(0);
#pragma unroll
      for (int j = 0; j < kEdgeBlk; ++j)
        if (k1 - 1 - j >= 0) q.step(u[j]);
    }
  }
  q.store(zo);
}

template <typename TIO, typename TR, bool RIGHT>
void launch_edge(const EdgeArgs& e, int coef_dtype, hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>((e.B + kWave - 1) / kWave)), block(kWave);
  const bool one = e.cpf >= e.B;
  if (coef_dtype == TSIIR_F64) {
    if (one)
      hipLaunchKernelGGL((sos_edge_kernel<TIO, TR, double, RIGHT, true>), grid, block, 0, stream, e);
    else
      hipLaunchKernelGGL((sos_edge_kernel<TIO, TR, double, RIGHT, false>), grid, block, 0, stream, e);
  } else {
    if (one)
      hipLaunchKernelGGL((sos_edge_kernel<TIO, TR, float, RIGHT, true>), grid, block, 0, stream, e);
    else
      hipLaunchKernelGGL((sos_edge_kernel<TIO, TR, float, RIGHT, false>), grid, block, 0, stream, e);
  }
}

template <bool RIGHT>
void launch_edge_io(const EdgeArgs& e, int io_dtype, int coef_dtype, hipStream_t stream) {
  if (io_dtype == TSIIR_F64)
    launch_edge<double, double, RIGHT>(e, coef_dtype, stream);
  else if (io_dtype == TSIIR_BF16)
    launch_edge<bf16_t, float, RIGHT>(e, coef_dtype, stream);
  else
    launch_edge<float, float, RIGHT>(e, coef_dtype, stream);
}

size_t io_bytes(int io_dtype) {
  switch (io_dtype) {
    case TSIIR_F32: return 4;
    case TSIIR_F64: return 8;
    case TSIIR_BF16: return 2;
    default: return 0;
  }
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Scratch: y1 (rows padded to whole L2 lines, so both passes over it take
// the line-aligned kernels), the right extension's outputs, the FP64
// states za, zb, zc and the device-formed steady states zs.
struct ScratchLayout {
  int64_t y1_stride = 0, ybuf_stride = 1;
  size_t y1 = 0, ybuf = 0, za = 0, zb = 0, zc = 0, zs = 0, total = 0;
};

// The largest default padlen of an S-section filter.
int64_t max_default_padlen(int64_t S) { return 3 * (2 * S + 1); }

// padlen < 0: room for any filter's default.
bool scratch_layout(const tsiir_sosfilt_desc* d, int64_t padlen, ScratchLayout* L) {
  const size_t esz = io_bytes(d->io_dtype);
  if (esz == 0 || d->channels < 0 || d->n_samples < 1 || d->n_sections < 1) return false;
  if (padlen < 0) padlen = max_default_padlen(d->n_sections);
  L->ybuf_stride = padlen > 0 ? padlen : 1;
  const int64_t per_line = kEdgeLine / static_cast<int64_t>(esz);
  const size_t B = static_cast<size_t>(d->channels);
  const size_t zs = B * static_cast<size_t>(d->n_sections) * 2 * sizeof(double);
  L->y1_stride = (d->n_samples + per_line - 1) / per_line * per_line;
  size_t off = 0;
  L->y1 = off;
  off = align256(off + B * static_cast<size_t>(L->y1_stride) * esz);
  L->ybuf = off;
  off = align256(off + B * static_cast<size_t>(L->ybuf_stride) * esz);
  L->za = off;
  off = align256(off + zs);
  L->zb = off;
  off = align256(off + zs);
  L->zc = off;
  off = align256(off + zs);
  L->zs = off;
  const size_t n_filters = d->chans_per_filter > 0
                               ? (B + static_cast<size_t>(d->chans_per_filter) - 1) /
                                     static_cast<size_t>(d->chans_per_filter)
                               : 0;
  off = align256(off + n_filters * static_cast<size_t>(d->n_sections) * 2 * sizeof(double));
  L->total = off;
  return true;
}

}  // namespace
}  // namespace tsiir

using namespace tsiir;

extern "C" {

size_t tsiir_sosfiltfilt_scratch_bytes(const tsiir_sosfilt_desc* desc, int64_t padlen) {
  ScratchLayout L;
  if (!desc || !scratch_layout(desc, padlen, &L)) return 0;
  return L.total;
}

tsiir_status tsiir_sosfiltfilt(const tsiir_sosfilt_desc* desc, int padtype, int64_t padlen,
                               const void* x, void* y, const void* sos, const double* zi_unit,
                               void* scratch, size_t scratch_bytes, void* workspace,
                               size_t workspace_bytes, void* stream) {
  if (!desc) return TSIIR_EINVAL;
  const tsiir_sosfilt_desc d = *desc;
  if ((d.flags & ~TSIIR_FP32_DF2T) != 0 || d.mode != TSIIR_MODE_PLAIN) return TSIIR_EINVAL;
  if (padtype < TSIIR_PAD_NONE || padtype > TSIIR_PAD_CONSTANT) return TSIIR_EINVAL;
  if (padlen < 0 && padlen != TSIIR_PADLEN_DEFAULT) return TSIIR_EINVAL;
  if ((padtype == TSIIR_PAD_NONE) != (padlen == 0)) return TSIIR_EINVAL;
  if (d.n_samples < 1 || d.n_sections < 1) return TSIIR_EINVAL;
  // a default padlen has to fit whatever the coefficients are
  const int64_t pad_max = padlen >= 0 ? padlen : max_default_padlen(d.n_sections);
  if (pad_max >= d.n_samples) return TSIIR_EINVAL;
  if (d.chans_per_filter < 1 || d.x_rows < 1) return TSIIR_EINVAL;
  if (d.channels == 0) return TSIIR_OK;
  if (!x || !y || !sos) return TSIIR_EINVAL;
  if (!workspace || workspace_bytes < sizeof(WsHeader)) return TSIIR_EINVAL;
  if (d.n_sections > kEdgeMaxS || pad_max * d.n_sections > kEdgeMaxWork) return TSIIR_EUNSUPPORTED;
  ScratchLayout L;
  if (!scratch_layout(&d, padlen, &L)) return TSIIR_EINVAL;
  if (!scratch || scratch_bytes < L.total) return TSIIR_EWORKSPACE;
  char* sc = static_cast<char*>(scratch);
  void* y1 = sc + L.y1;
  double* za = reinterpret_cast<double*>(sc + L.za);
  double* zb = reinterpret_cast<double*>(sc + L.zb);
  double* zc = reinterpret_cast<double*>(sc + L.zc);
  hipStream_t st = static_cast<hipStream_t>(stream);

  EdgeArgs e{};
  e.x = x;
  e.x_rows = d.x_rows;
  e.x_stride = d.x_rows > 1 ? d.x_row_stride : d.n_samples;
  e.B = d.channels;
  e.T = d.n_samples;
  e.cpf = d.chans_per_filter;
  e.S = d.n_sections;
  e.padtype = padtype;
  e.padlen = padlen;
  e.ybuf_stride = L.ybuf_stride;
  e.sos = sos;
  e.zi_unit = zi_unit;
  e.zs = reinterpret_cast<double*>(sc + L.zs);
  e.hdr = static_cast<WsHeader*>(workspace);
  e.z_out = za;
  e.y1 = y1;
  e.y1_stride = L.y1_stride;
  e.ybuf = sc + L.ybuf;
  launch_edge_io<false>(e, d.io_dtype, d.coef_dtype, st);
  if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;

  const uint32_t keep = d.flags & TSIIR_FP32_DF2T;
  tsiir_sosfilt_desc m = d;
  m.flags = TSIIR_HAS_ZI | TSIIR_WRITE_ZF | TSIIR_STATE_F64 | keep;
  m.y_row_stride = L.y1_stride;
  tsiir_status rc = forward_impl(&m, x, y1, sos, za, zb, workspace, workspace_bytes, stream,
                                 nullptr);
  if (rc != TSIIR_OK) return rc;

  e.z_in = zb;
  e.z_out = zc;
  launch_edge_io<true>(e, d.io_dtype, d.coef_dtype, st);
  if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;

  tsiir_sosfilt_desc r = d;
  r.flags = TSIIR_HAS_ZI | TSIIR_STATE_F64 | TSIIR_REVERSE | keep;
  r.x_rows = d.channels;
  r.x_row_stride = L.y1_stride;
  return forward_impl(&r, y1, y, sos, zc, nullptr, workspace, workspace_bytes, stream, nullptr);
}

}  // extern "C"
