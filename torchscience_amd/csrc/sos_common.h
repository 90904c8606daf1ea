// sos_common.h - shared device helpers for the gfx950 SOS kernels.
//
// Layout constants, the workspace header, coefficient tables and the
// self-validating scan status words used by the decoupled look-back.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "tsiir.h"

namespace tsiir {

constexpr int kWave = 64;        // CDNA wavefront
constexpr int kNPow = 15;         // A^(W*2^i), i = 0..14
constexpr int kGOff = 72;         // doubles before the float zero-input response table
constexpr int kGLen = 65;         // floats in that table: g[0..64]
constexpr int kTableStride = 112; // doubles per (filter, section) table entry (896 B)
static_assert(2 * kGOff + kGLen <= 2 * kTableStride, "table entry too small");
constexpr int kStatusWords = 8;  // 4 words aggregate + 4 words inclusive
constexpr uint32_t kSpinLimit = 1u << 18;

// Workspace header (256 bytes).  `epoch` advances before every launch that
// publishes scan status words (see launch_flag).
struct WsHeader {
  uint32_t seg_gate;   // 1 while the segmented form applies (see seg_check_kernel)
  uint32_t reserved1;
  uint32_t epoch;
  uint32_t error;      // set when a look-back wait gave up (see give_up)
  uint32_t* mailbox;   // caller's device-visible word that also gets it, or null
  // the last shared filter the segmented form's device check judged: its
  // section count, coefficient type and values (as doubles), and the verdict
  uint32_t seg_cache_valid;
  int32_t seg_cache_sections;
  int32_t seg_cache_coef;
  uint32_t seg_cache_verdict;
  uint32_t pad[54];
  double seg_cache_sos[16 * 6];
};
static_assert(sizeof(WsHeader) == 1024, "header size");

// A look-back wait that gives up (its spin limit passed, or the caller's
// TSIIR_DEBUG_GIVE_UP): `bit` goes into the header's error word and the
// caller's mailbox, and the incoming state becomes NaN, so every output that
// depends on the missing state is NaN instead of silently wrong.
__device__ __forceinline__ void give_up(WsHeader* hdr, uint32_t bit, double& sw0, double& sw1) {
  if ((threadIdx.x & (kWave - 1)) == 0) {
    atomicOr(&hdr->error, bit);
    uint32_t* mb = hdr->mailbox;
    if (mb != nullptr) __hip_atomic_store(mb, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  sw0 = __builtin_nan("");
  sw1 = __builtin_nan("");
}
__device__ __forceinline__ bool debug_give_up(uint32_t flags) {
  return (flags & TSIIR_DEBUG_GIVE_UP) != 0;
}

// Element idx of zi / zf ((rows, S, 2) row-major): double with
// TSIIR_STATE_F64, else the compute type TACC.
template <typename TACC>
__device__ __forceinline__ double state_in(const void* zi, uint32_t flags, int64_t idx) {
  if (flags & TSIIR_STATE_F64) return static_cast<const double*>(zi)[idx];
  return static_cast<double>(static_cast<const TACC*>(zi)[idx]);
}
template <typename TACC>
__device__ __forceinline__ void state_out(void* zf, uint32_t flags, int64_t idx, double v) {
  if (flags & TSIIR_STATE_F64)
    static_cast<double*>(zf)[idx] = v;
  else
    static_cast<TACC*>(zf)[idx] = static_cast<TACC>(v);
}

struct Mat2 {
  double a, b, c, d;  // [[a, b], [c, d]]
};

__device__ __forceinline__ Mat2 mat_mul(const Mat2& x, const Mat2& y) {
  return Mat2{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d,
              x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}

// Zero-input response table of a section (float), at kGOff of its entry:
// g[0] = 0, g[1] = 1, g[j] = -a1 g[j-1] - a2 g[j-2].  The DF2T output j
// samples after state (s0, s1) with zero input is g[j+1] s0 + g[j] s1.
__device__ __forceinline__ const float* table_g(const double* t) {
  return reinterpret_cast<const float*>(t + kGOff);
}

// d = a * b + c as one VOP3 FMA with a free destination.  The two-address
// v_fmac form that instruction selection prefers needs a register copy
// whenever the result must land in a loop-carried register, as the sample
// registers of the section sweep do.
__device__ __forceinline__ float fma3(float a, float b, float c) {
  float d;
  asm("v_fma_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
__device__ __forceinline__ double fma3(double a, double b, double c) { return a * b + c; }
// The same with a wave-uniform (SGPR) multiplicand.
__device__ __forceinline__ float fma3s(float a, float b, float c) {
  float d;
  asm("v_fma_f32 %0, %1, %2, %3" : "=v"(d) : "s"(a), "v"(b), "v"(c));
  return d;
}

__device__ __forceinline__ Mat2 load_mat(const double* t) {
  return Mat2{t[0], t[1], t[2], t[3]};
}

// Table entry of one (filter, section):
//   [0..4]  b0 b1 b2 a1 a2 (normalised by a0)
//   [8 + 4i .. 8 + 4i + 3]  A^(W * 2^i) for i < kNPow, where
//   A = [[-a1, 1], [-a2, 0]] is the DF2T state matrix of the section.
__device__ __forceinline__ const double* table_pow(const double* t, int i) {
  return t + 8 + 4 * i;
}

// Table entry of one (filter, section) for the FP32-recurrence forward kernel
// (sosfilt_fwd32.hip), built next to the FP64 entry above.
//  c     DF1 coefficients {b0, b1, b2, -a1, -a2} (normalised by a0).  When
//        `unit` is set (per filter) every row is also divided by its b0, so
//        the section computes y/b0 and the first input tap is 1; p is then
//        the running product of the b0 of sections 0..s (the factor between
//        the true and the computed output of section s) and rp = 1/p.
//        Otherwise b0 stays in c[0] and p = rp = 1.
//  mq    M^q for q < 64 in FP32, M = A^64 the DF2T state matrix over one
//        lane's 64 samples, rows {a, b, c, d} of [[a, b], [c, d]].
struct Fast32 {
  float c[5];
  int unit;
  double p, rp;
  float pad[6];
  float4 mq[64];
};
static_assert(sizeof(Fast32) == 1088, "Fast32 layout");

// Arguments of the forward window kernels.
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
  int io_dtype;
  // When set, the kernel runs only if (*gate != 0) == gate_on: the segmented
  // form and the look-back form of one call are both launched and a device
  // check picks which one works (every kernel of the other exits at once).
  const uint32_t* gate = nullptr;
  uint32_t gate_on = 0;
  // Two-level rows (the segmented form): with seg_n > 0, row r is segment
  // r % seg_n of channel r / seg_n, at (r / seg_n) * stride + (r % seg_n) *
  // seg_len elements; a zero seg_len keeps plain rows (r * stride).
  int64_t seg_n = 0;
  int64_t x_seg_len = 0;
  int64_t y_seg_len = 0;
  // Status words carry the epoch epoch_back launches before the current one
  // (launches after one shared table preparation).
  uint32_t epoch_back = 0;
  // Segmented form, FP32-recurrence chained kernels: the first group of a
  // segment other than the first one processed computes its incoming states
  // itself (a zero-state pass over the TSIIR_SEGMENT_WARMUP samples before
  // it, REVERSE: after it); the first segment reads zi row (row / seg_n).
  int warm_fused = 0;
  // zi's storage type: -1 as TSIIR_STATE_F64 says, else 1 = FP64, 0 = the
  // compute type (TSIIR_STATE_F64 then applies to zf alone).
  int zi_f64 = -1;
};

// Initial-state element idx of zi, and the zi row of channel (row) c.
template <typename TACC>
__device__ __forceinline__ double zi_in(const FwdArgs& a, int64_t idx) {
  const uint32_t f = a.zi_f64 < 0 ? a.flags : (a.zi_f64 ? TSIIR_STATE_F64 : 0u);
  return state_in<TACC>(a.zi, f, idx);
}
__device__ __forceinline__ int64_t zi_row_of(const FwdArgs& a, int64_t c) {
  if (a.flags & TSIIR_ZI_BROADCAST) return 0;
  return a.warm_fused ? c / a.seg_n : c;
}

__device__ __forceinline__ int64_t row_offset(const FwdArgs& a, int64_t r, int64_t stride,
                                              int64_t seg_len) {
  return (a.seg_n > 0 && seg_len > 0) ? (r / a.seg_n) * stride + (r % a.seg_n) * seg_len
                                      : r * stride;
}

__device__ __forceinline__ bool gated_out(const FwdArgs& a) {
  return a.gate != nullptr && ((*a.gate != 0u) != (a.gate_on != 0u));
}

// 64-bit status word = (flag << 32) | 32-bit payload.  A double is stored
// as two words (low half, high half).  Each word is written and read with a
// single 64-bit relaxed atomic at agent scope, so a reader that sees the
// current flag in every word has a consistent value without any fence.
__device__ __forceinline__ uint64_t atomic_ld64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_st64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lanes 0..3 of the calling wave each write one word of (v0, v1).
__device__ __forceinline__ void publish_pair(uint64_t* words, double v0, double v1,
                                             uint32_t flag, int lane) {
  if (lane < 4) {
    const double v = (lane < 2) ? v0 : v1;
    const uint64_t bits = static_cast<uint64_t>(__double_as_longlong(v));
    const uint32_t half = (lane & 1) ? static_cast<uint32_t>(bits >> 32)
                                     : static_cast<uint32_t>(bits);
    atomic_st64(words + lane, (static_cast<uint64_t>(flag) << 32) | half);
  }
}

// Reads a pair published by publish_pair; returns false unless all four
// words carry `flag`.
// The same read in two halves, so that the loads can be issued early and
// checked after unrelated work: load_pair_words, then decode_pair.
struct PairWords {
  uint64_t w[4];
};
__device__ __forceinline__ void load_pair_words(const uint64_t* words, PairWords& p) {
#pragma unroll
  for (int i = 0; i < 4; ++i) p.w[i] = atomic_ld64(words + i);
}
__device__ __forceinline__ bool decode_pair(const PairWords& p, uint32_t flag, double& v0,
                                            double& v1) {
  const uint64_t w0 = p.w[0], w1 = p.w[1], w2 = p.w[2], w3 = p.w[3];
  const bool ok = (static_cast<uint32_t>(w0 >> 32) == flag) &
                  (static_cast<uint32_t>(w1 >> 32) == flag) &
                  (static_cast<uint32_t>(w2 >> 32) == flag) &
                  (static_cast<uint32_t>(w3 >> 32) == flag);
  const uint64_t b0 = (w0 & 0xffffffffull) | ((w1 & 0xffffffffull) << 32);
  const uint64_t b1 = (w2 & 0xffffffffull) | ((w3 & 0xffffffffull) << 32);
  v0 = __longlong_as_double(static_cast<long long>(b0));
  v1 = __longlong_as_double(static_cast<long long>(b1));
  return ok;
}

__device__ __forceinline__ bool read_pair(const uint64_t* words, uint32_t flag,
                                          double& v0, double& v1) {
  const uint64_t w0 = atomic_ld64(words + 0);
  const uint64_t w1 = atomic_ld64(words + 1);
  const uint64_t w2 = atomic_ld64(words + 2);
  const uint64_t w3 = atomic_ld64(words + 3);
  const bool ok = (static_cast<uint32_t>(w0 >> 32) == flag) &
                  (static_cast<uint32_t>(w1 >> 32) == flag) &
                  (static_cast<uint32_t>(w2 >> 32) == flag) &
                  (static_cast<uint32_t>(w3 >> 32) == flag);
  const uint64_t b0 = (w0 & 0xffffffffull) | ((w1 & 0xffffffffull) << 32);
  const uint64_t b1 = (w2 & 0xffffffffull) | ((w3 & 0xffffffffull) << 32);
  v0 = __longlong_as_double(static_cast<long long>(b0));
  v1 = __longlong_as_double(static_cast<long long>(b1));
  return ok;
}

// A double of lane l (wave-uniform l) in every lane.
__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// A double of lane (addr / 4) for every lane (ds_bpermute on both halves).
__device__ __forceinline__ double bpermute_d(int addr, double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, static_cast<int>(b));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// ---- wave-level scan of 2x2 affine maps -----------------------------------
//
// Lane q holds the zero-state end state e_q of its W samples.  With
// M = A^W (table power 0), the inclusive scan gives
//   incl_q = sum_{p <= q} M^(q-p) e_p,
// the window aggregate incl_63 and the exclusive value incl_(q-1).
__device__ __forceinline__ void wave_scan(const double* tb, int lane, double e0, double e1,
                                          double& agg0, double& agg1, double& x0, double& x1) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const int d = 1 << i;
    const int src = ((lane - d) & (kWave - 1)) << 2;  // lanes < d: masked below
    const double n0 = bpermute_d(src, e0);
    const double n1 = bpermute_d(src, e1);
    const Mat2 P = load_mat(table_pow(tb, i));
    if (lane >= d) {
      e0 = fma(P.b, n1, fma(P.a, n0, e0));
      e1 = fma(P.d, n1, fma(P.c, n0, e1));
    }
  }
  agg0 = readlane_d(e0, kWave - 1);
  agg1 = readlane_d(e1, kWave - 1);
  const int src1 = ((lane - 1) & (kWave - 1)) << 2;
  x0 = bpermute_d(src1, e0);
  x1 = bpermute_d(src1, e1);
  if (lane == 0) {
    x0 = 0.0;
    x1 = 0.0;
  }
}

// Incoming state of lane q: M^q s_window + exclusive_q (M^q from the binary
// digits of q and the table powers M^(2^i)).
__device__ __forceinline__ void lane_incoming(const double* tb, int lane, double sw0, double sw1,
                                              double x0, double x1, double& s0, double& s1) {
  s0 = sw0;
  s1 = sw1;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const Mat2 P = load_mat(table_pow(tb, i));
    if ((lane >> i) & 1) {
      const double t0 = P.a * s0 + P.b * s1;
      const double t1 = P.c * s0 + P.d * s1;
      s0 = t0;
      s1 = t1;
    }
  }
  s0 += x0;
  s1 += x1;
}

// Look-back over a span whose power vanishes.  When Mw = A^(span) is exactly
// zero in FP64 (every pole decays below the smallest double over one span,
// as for 32768-sample groups of any section with poles inside radius ~0.97),
// the incoming state of span k is Mw in_(k-1) + agg_(k-1) = agg_(k-1)
// bitwise, and span k-1's inclusive state equals its aggregate: one word of
// the predecessor ends the look-back, whichever of the two is there first.
__device__ __forceinline__ bool span_power_is_zero(const double* tb, int p) {
  const double* q = table_pow(tb, p);
  return q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0 && q[3] == 0.0;
}
// The aggregate is published before the look-back that leads to the
// inclusive state, so its word alone is polled.
__device__ __forceinline__ void lookback_predecessor(const uint64_t* e, WsHeader* hdr,
                                                     uint32_t flag, uint32_t flags,
                                                     double& sw0, double& sw1) {
  if (debug_give_up(flags)) {
    give_up(hdr, 1u, sw0, sw1);
    return;
  }
  uint32_t spins = 0;
  while (!read_pair(e, flag, sw0, sw1)) {
    if (++spins > kSpinLimit) {
      give_up(hdr, 1u, sw0, sw1);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Waits for one inclusive pair (a chained predecessor); gives up as above.
__device__ __forceinline__ void wait_inclusive(const uint64_t* incl, WsHeader* hdr,
                                               uint32_t flag, uint32_t flags, double& sw0,
                                               double& sw1) {
  if (debug_give_up(flags)) {
    give_up(hdr, 1u, sw0, sw1);
    return;
  }
  uint32_t spins = 0;
  while (!read_pair(incl, flag, sw0, sw1)) {
    if (++spins > kSpinLimit) {
      give_up(hdr, 1u, sw0, sw1);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Decoupled look-back: incoming state of window k of (channel c, section s),
// the same value in every lane.  Lane l inspects window k-1-l; the nearest
// window with an inclusive state (or the virtual window -1 holding zi) ends
// the search, windows in between contribute their aggregates:
//   in_k = sum_{l < f} Mw^l agg_(k-1-l) + Mw^f incl_(k-1-f),
// where Mw = A^(window length) is table power `pbase`, combined by a
// shuffle-down reduction with Mw^(2^i) = table power pbase+i.  A wait longer
// than kSpinLimit polls sets hdr->error and gives up, so every wave
// terminates.
__device__ __forceinline__ void lookback(const uint64_t* status, int64_t B, int S, WsHeader* hdr,
                                         int64_t k, int64_t c, int s, uint32_t flag,
                                         const double* tb, int pbase, double zi0, double zi1,
                                         double& sw0, double& sw1, uint32_t flags = 0) {
  const int lane = threadIdx.x & (kWave - 1);
  if (k >= 1 && (debug_give_up(flags) || span_power_is_zero(tb, pbase))) {
    lookback_predecessor(status + (((k - 1) * B + c) * S + s) * kStatusWords, hdr, flag, flags,
                         sw0, sw1);
    return;
  }
  double acc0 = 0.0, acc1 = 0.0;
  Mat2 P{1.0, 0.0, 0.0, 1.0};
  int64_t base = k - 1;
  uint32_t spins = 0;
  while (true) {
    const int64_t j = base - lane;
    int st = 2;
    double v0 = 0.0, v1 = 0.0;
    if (j >= 0) {
      const uint64_t* e = status + ((j * B + c) * S + s) * kStatusWords;
      if (read_pair(e + 4, flag, v0, v1)) {
        st = 2;
      } else if (read_pair(e, flag, v0, v1)) {
        st = 1;
      } else {
        st = 0;
      }
    } else if (j == -1) {
      v0 = zi0;
      v1 = zi1;
    }
    const uint64_t incl = __ballot(st == 2);
    const uint64_t ready = __ballot(st != 0);
    const int f = incl ? (__ffsll(static_cast<unsigned long long>(incl)) - 1) : 64;
    const uint64_t need = (f >= 63) ? ~0ull : ((1ull << (f + 1)) - 1ull);
    if ((ready & need) != need) {
      if (++spins > kSpinLimit) {
        give_up(hdr, 1u, acc0, acc1);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
      continue;
    }
    if (f == 0) {  // the nearest window is already inclusive (the common case)
      const double r0 = __shfl(v0, 0);
      const double r1 = __shfl(v1, 0);
      acc0 += P.a * r0 + P.b * r1;
      acc1 += P.c * r0 + P.d * r1;
      break;
    }
    if (lane > f) {
      v0 = 0.0;
      v1 = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int d = 1 << i;
      const double n0 = __shfl_down(v0, d);
      const double n1 = __shfl_down(v1, d);
      const Mat2 Q = load_mat(table_pow(tb, pbase + i));
      if (lane + d < kWave) {
        v0 += Q.a * n0 + Q.b * n1;
        v1 += Q.c * n0 + Q.d * n1;
      }
    }
    const double r0 = __shfl(v0, 0);
    const double r1 = __shfl(v1, 0);
    acc0 += P.a * r0 + P.b * r1;
    acc1 += P.c * r0 + P.d * r1;
    if (f < 64) break;
    P = mat_mul(P, load_mat(table_pow(tb, pbase + 6)));
    base -= kWave;
  }
  sw0 = acc0;
  sw1 = acc1;
}

// Epoch of the running launch (wave-uniform).  The prep kernel that runs
// before every launch advances hdr->epoch, so status words written by earlier
// launches never carry a current flag.  A call that makes several launches
// reserves one epoch per launch and launch i uses epoch - i.
__device__ __forceinline__ uint32_t launch_flag(const WsHeader* hdr, uint32_t back) {
  return __builtin_amdgcn_readfirstlane(
             __hip_atomic_load(&hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) -
         back;
}

// ---- element I/O ---------------------------------------------------------

struct bf16_t {
  uint16_t bits;
};

template <typename T>
struct IOTraits;

template <>
struct IOTraits<float> {
  using vec_t = float4;
  static constexpr int kVec = 4;
  __device__ static __forceinline__ float to_f(float v) { return v; }
  __device__ static __forceinline__ float from_f(float v) { return v; }
};

template <>
struct IOTraits<double> {
  using vec_t = double2;
  static constexpr int kVec = 2;
  __device__ static __forceinline__ double to_f(double v) { return v; }
  __device__ static __forceinline__ double from_f(double v) { return v; }
};

template <>
struct IOTraits<bf16_t> {
  using vec_t = uint4;
  static constexpr int kVec = 8;
  __device__ static __forceinline__ float to_f(bf16_t v) {
    return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
  }
  // round to nearest even; NaN stays NaN
  __device__ static __forceinline__ bf16_t from_f(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return bf16_t{static_cast<uint16_t>((u >> 16) | 0x40)};
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16_t{static_cast<uint16_t>(u >> 16)};
  }
};

// Load W consecutive elements into registers (converted to TACC).  `vec`
// selects 16-byte loads (row start 16-byte aligned and all W in range);
// otherwise elements at or beyond `m` read as zero.
template <typename TIO, typename TACC, int W>
__device__ __forceinline__ void load_run(const TIO* p, bool vec, int m, TACC (&v)[W]) {
  using Tr = IOTraits<TIO>;
  if (vec) {
    const typename Tr::vec_t* q = reinterpret_cast<const typename Tr::vec_t*>(p);
#pragma unroll
    for (int i = 0; i < W / Tr::kVec; ++i) {
      const typename Tr::vec_t r = q[i];
      const TIO* e = reinterpret_cast<const TIO*>(&r);
#pragma unroll
      for (int j = 0; j < Tr::kVec; ++j) v[i * Tr::kVec + j] = static_cast<TACC>(Tr::to_f(e[j]));
    }
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) v[j] = (j < m) ? static_cast<TACC>(Tr::to_f(p[j])) : TACC(0);
  }
}

// Reverse-order runs: v[j] = p[n_hi - j] for j < m (zero beyond m); with
// `vec`, the W elements [n_hi - W + 1, n_hi] are read as 16-byte vectors
// (n_hi - W + 1 must then be 16-byte aligned).
template <typename TIO, typename TACC, int W>
__device__ __forceinline__ void load_run_rev(const TIO* p, int64_t n_hi, bool vec, int m,
                                             TACC (&v)[W]) {
  using Tr = IOTraits<TIO>;
  if (vec) {
    const typename Tr::vec_t* q = reinterpret_cast<const typename Tr::vec_t*>(p + (n_hi - (W - 1)));
#pragma unroll
    for (int i = 0; i < W / Tr::kVec; ++i) {
      const typename Tr::vec_t r = q[i];
      const TIO* e = reinterpret_cast<const TIO*>(&r);
#pragma unroll
      for (int j = 0; j < Tr::kVec; ++j)
        v[W - 1 - (i * Tr::kVec + j)] = static_cast<TACC>(Tr::to_f(e[j]));
    }
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) v[j] = (j < m) ? static_cast<TACC>(Tr::to_f(p[n_hi - j])) : TACC(0);
  }
}

template <typename TIO, typename TACC, int W>
__device__ __forceinline__ void store_run_rev(TIO* p, int64_t n_hi, bool vec, int m,
                                              const TACC (&v)[W]) {
  using Tr = IOTraits<TIO>;
  if (vec) {
    typename Tr::vec_t* q = reinterpret_cast<typename Tr::vec_t*>(p + (n_hi - (W - 1)));
#pragma unroll
    for (int i = 0; i < W / Tr::kVec; ++i) {
      typename Tr::vec_t r;
      TIO* e = reinterpret_cast<TIO*>(&r);
#pragma unroll
      for (int j = 0; j < Tr::kVec; ++j) e[j] = Tr::from_f(v[W - 1 - (i * Tr::kVec + j)]);
      q[i] = r;
    }
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (j < m) p[n_hi - j] = Tr::from_f(v[j]);
  }
}

// ---- per-lane runs at any element alignment ----------------------------------
//
// A lane's run of W elements that starts r elements past a 16-byte boundary
// (0 < r < kVec) moves as the W / kVec + 1 aligned vectors that cover it:
// each holds at least one element of the run, so none reaches a page the
// run does not touch.  The elements are moved by r in registers, one select
// per element for each bit of r (fold expressions over constant indices, so
// the arrays stay in registers).  Stores write the run's own elements only:
// the first and last vectors element by element.

template <typename T, int S, int N, size_t... J>
__device__ __forceinline__ void run_select_above(T (&f)[N], bool c, std::index_sequence<J...>) {
  ((f[J] = c ? f[J + S] : f[J]), ...);
}
template <typename T, int S, int N, size_t... J>
__device__ __forceinline__ void run_select_below(T (&f)[N], bool c, std::index_sequence<J...>) {
  ((f[N - 1 - J] = c ? f[N - 1 - J - S] : f[N - 1 - J]), ...);
}
// f[j] = f[j + m] (j < N - m), m < KV, KV a power of two <= 8
template <int KV, typename T, int N>
__device__ __forceinline__ void run_move_down(T (&f)[N], int m) {
  if constexpr (KV > 1) run_select_above<T, 1>(f, (m & 1) != 0, std::make_index_sequence<N - 1>{});
  if constexpr (KV > 2) run_select_above<T, 2>(f, (m & 2) != 0, std::make_index_sequence<N - 2>{});
  if constexpr (KV > 4) run_select_above<T, 4>(f, (m & 4) != 0, std::make_index_sequence<N - 4>{});
}
// f[j] = f[j - m] (j >= m), m < KV, KV a power of two <= 8
template <int KV, typename T, int N>
__device__ __forceinline__ void run_move_up(T (&f)[N], int m) {
  if constexpr (KV > 1) run_select_below<T, 1>(f, (m & 1) != 0, std::make_index_sequence<N - 1>{});
  if constexpr (KV > 2) run_select_below<T, 2>(f, (m & 2) != 0, std::make_index_sequence<N - 2>{});
  if constexpr (KV > 4) run_select_below<T, 4>(f, (m & 4) != 0, std::make_index_sequence<N - 4>{});
}

// v[j] = p[j] (REV: v[j] = p[W - 1 - j]), p at any element alignment.
template <typename TIO, typename TACC, int W, bool REV = false>
__device__ __forceinline__ void load_run_any(const TIO* p, TACC (&v)[W]) {
  using Tr = IOTraits<TIO>;
  constexpr int kV = Tr::kVec;
  const int r = static_cast<int>((reinterpret_cast<uintptr_t>(p) & 15u) / sizeof(TIO));
  const typename Tr::vec_t* q = reinterpret_cast<const typename Tr::vec_t*>(p - r);
  TACC f[W + kV];
#pragma unroll
  for (int i = 0; i <= W / kV; ++i) {
    // (with r = 0 the last vector holds no element of the run: not read)
    typename Tr::vec_t rv{};
    if (i < W / kV || r != 0) rv = q[i];
    const TIO* e = reinterpret_cast<const TIO*>(&rv);
#pragma unroll
    for (int j = 0; j < kV; ++j) f[i * kV + j] = static_cast<TACC>(Tr::to_f(e[j]));
  }
  run_move_down<kV>(f, r);
#pragma unroll
  for (int j = 0; j < W; ++j) v[j] = f[REV ? W - 1 - j : j];
}

// p[j] = v[j] (REV: p[W - 1 - j] = v[j]), p at any element alignment.
template <typename TIO, typename TACC, int W, bool REV = false>
__device__ __forceinline__ void store_run_any(TIO* p, const TACC (&v)[W]) {
  using Tr = IOTraits<TIO>;
  constexpr int kV = Tr::kVec;
  const int r = static_cast<int>((reinterpret_cast<uintptr_t>(p) & 15u) / sizeof(TIO));
  typename Tr::vec_t* q = reinterpret_cast<typename Tr::vec_t*>(p - r);
  // h = [kV zeros, the run in time order, kV zeros], moved up by r: h[kV + j]
  // is then element j of the aligned vectors that cover the run
  TACC h[W + 2 * kV];
#pragma unroll
  for (int j = 0; j < kV; ++j) h[j] = h[W + kV + j] = TACC(0);
#pragma unroll
  for (int j = 0; j < W; ++j) h[kV + j] = v[REV ? W - 1 - j : j];
  run_move_up<kV>(h, r);
  TIO* base = p - r;
#pragma unroll
  for (int i = 0; i <= W / kV; ++i) {
    if (i == 0 || i == W / kV) {
#pragma unroll
      for (int j = 0; j < kV; ++j) {
        const int k = i * kV + j;
        if (k >= r && k < W + r) base[k] = Tr::from_f(h[kV + k]);
      }
    } else {
      typename Tr::vec_t rv;
      TIO* e = reinterpret_cast<TIO*>(&rv);
#pragma unroll
      for (int j = 0; j < kV; ++j) e[j] = Tr::from_f(h[kV + i * kV + j]);
      q[i] = rv;
    }
  }
}

// ---- LDS-staged window I/O --------------------------------------------------
//
// A wave's window is 64*W contiguous elements; lane q owns [q*W, (q+1)*W),
// i.e. kNV 16-byte vectors.  The window moves in two rounds, round r carrying
// the half-rows [r*W/2, (r+1)*W/2) of all 64 lanes (kH = kNV/2 vectors per
// lane).  Global side: instruction i, lane l handles row q = i*(64/kH) + l/kH,
// vector t = l % kH of that half-row, so every instruction covers whole runs
// of kH consecutive vectors (fully used cache lines).  LDS side: row q at
// q*(kH+1) vectors (one vector of padding: the 64 lanes reading column t hit
// distinct banks).  Every lane takes part in every LDS access of both rounds,
// so nothing runs under a partial execution mask.
template <typename TIO, int W>
struct LdsStage {
  static constexpr int kVec = IOTraits<TIO>::kVec;      // elements per 16 bytes
  static constexpr int kNV = W / kVec;                  // 16-byte vectors per lane
  static constexpr int kH = kNV / 2;                    // vectors per lane per round
  static constexpr int kRowV = kH + 1;                  // padded row (vectors)
  static constexpr int kElems = kWave * kRowV * kVec;   // buffer size (elements)
  static_assert(kH >= 1 && kWave % kH == 0, "window shape");
  // vector index within the window of (round r, instruction i, lane l)
  __device__ static __forceinline__ int gvec(int r, int i, int l) {
    return (i * (kWave / kH) + l / kH) * kNV + r * kH + l % kH;
  }
  // vector offset in LDS of (instruction i, lane l)
  __device__ static __forceinline__ int lvec(int i, int l) {
    return (i * (kWave / kH) + l / kH) * kRowV + l % kH;
  }
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// LDS rows -> the lane's samples [r*W/2, (r+1)*W/2).  REV: lane q takes
// row 63-q with its elements in reverse order (time-reversed windows).
template <typename TIO, typename TACC, int W, bool REV = false>
__device__ __forceinline__ void lds_rows_to_regs(const TIO* lds, int r, TACC (&v)[W]) {
  using L = LdsStage<TIO, W>;
  using vec_t = typename IOTraits<TIO>::vec_t;
  const int lane = threadIdx.x & (kWave - 1);
  const vec_t* row =
      reinterpret_cast<const vec_t*>(lds) + (REV ? kWave - 1 - lane : lane) * L::kRowV;
#pragma unroll
  for (int t = 0; t < L::kH; ++t) {
    const vec_t rv = row[t];
    const TIO* e = reinterpret_cast<const TIO*>(&rv);
#pragma unroll
    for (int j = 0; j < L::kVec; ++j) {
      const int idx = (r * L::kH + t) * L::kVec + j;
      v[REV ? W - 1 - idx : idx] = static_cast<TACC>(IOTraits<TIO>::to_f(e[j]));
    }
  }
}

// NT: non-temporal global accesses (streamed data that is not re-read).
template <typename TIO, typename TACC, int W, bool REV = false, bool NT = false>
__device__ __forceinline__ void load_window_lds(const TIO* win, TIO* lds, TACC (&v)[W]) {
  using L = LdsStage<TIO, W>;
  const int lane = threadIdx.x & (kWave - 1);
  const u32x4* __restrict__ g = reinterpret_cast<const u32x4*>(win);
  u32x4* lv = reinterpret_cast<u32x4*>(lds);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int i = 0; i < L::kH; ++i) {
      const u32x4* src = g + L::gvec(r, i, lane);
      lv[L::lvec(i, lane)] = NT ? __builtin_nontemporal_load(src) : *src;
    }
    __syncthreads();
    lds_rows_to_regs<TIO, TACC, W, REV>(lds, r, v);
    __syncthreads();
  }
}

// Prefetch form of load_window_lds: issue the window's global loads into
// registers (native 4 x u32 vectors, which stay in VGPRs across the loop
// that carries them), and transpose them through LDS when the window is due.
template <typename TIO, int W>
__device__ __forceinline__ void prefetch_window(const TIO* win,
                                                u32x4 (&pf)[LdsStage<TIO, W>::kNV]) {
  using L = LdsStage<TIO, W>;
  const int lane = threadIdx.x & (kWave - 1);
  const u32x4* __restrict__ g = reinterpret_cast<const u32x4*>(win);
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int i = 0; i < L::kH; ++i) pf[r * L::kH + i] = g[L::gvec(r, i, lane)];
}

template <typename TIO, typename TACC, int W, bool REV = false>
__device__ __forceinline__ void transpose_window_lds(const u32x4 (&pf)[LdsStage<TIO, W>::kNV],
                                                     TIO* lds, TACC (&v)[W]) {
  using L = LdsStage<TIO, W>;
  const int lane = threadIdx.x & (kWave - 1);
  u32x4* lv = reinterpret_cast<u32x4*>(lds);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int i = 0; i < L::kH; ++i) lv[L::lvec(i, lane)] = pf[r * L::kH + i];
    __syncthreads();
    lds_rows_to_regs<TIO, TACC, W, REV>(lds, r, v);
    __syncthreads();
  }
}

template <typename TIO, typename TACC, int W, bool REV = false, bool NT = false>
__device__ __forceinline__ void store_window_lds(TIO* win, TIO* lds, const TACC (&v)[W]) {
  using L = LdsStage<TIO, W>;
  using vec_t = typename IOTraits<TIO>::vec_t;
  const int lane = threadIdx.x & (kWave - 1);
  u32x4* __restrict__ g = reinterpret_cast<u32x4*>(win);
  const u32x4* lv = reinterpret_cast<const u32x4*>(lds);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    vec_t* row = reinterpret_cast<vec_t*>(lds) + (REV ? kWave - 1 - lane : lane) * L::kRowV;
#pragma unroll
    for (int t = 0; t < L::kH; ++t) {
      vec_t rv;
      TIO* e = reinterpret_cast<TIO*>(&rv);
#pragma unroll
      for (int j = 0; j < L::kVec; ++j) {
        const int idx = (r * L::kH + t) * L::kVec + j;
        e[j] = IOTraits<TIO>::from_f(v[REV ? W - 1 - idx : idx]);
      }
      row[t] = rv;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < L::kH; ++i) {
      u32x4* dst = g + L::gvec(r, i, lane);
      if (NT)
        __builtin_nontemporal_store(lv[L::lvec(i, lane)], dst);
      else
        *dst = lv[L::lvec(i, lane)];
    }
    __syncthreads();
  }
}

template <typename TIO, typename TACC, int W>
__device__ __forceinline__ void store_run(TIO* p, bool vec, int m, const TACC (&v)[W]) {
  using Tr = IOTraits<TIO>;
  if (vec) {
    typename Tr::vec_t* q = reinterpret_cast<typename Tr::vec_t*>(p);
#pragma unroll
    for (int i = 0; i < W / Tr::kVec; ++i) {
      typename Tr::vec_t r;
      TIO* e = reinterpret_cast<TIO*>(&r);
#pragma unroll
      for (int j = 0; j < Tr::kVec; ++j) e[j] = Tr::from_f(v[i * Tr::kVec + j]);
      q[i] = r;
    }
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (j < m) p[j] = Tr::from_f(v[j]);
  }
}

}  // namespace tsiir
