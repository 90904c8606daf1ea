// sosfilt_fwd32.hip - forward SOS cascade with FP32 recurrences (FP32 or
// BF16 I/O) over the full 4096-sample windows of every channel, on gfx950.
//
// Replaces the section loop of sosfilt (_sosfilt.py:117-128) and the
// per-sample DF2T loop of lfilter (_lfilter.py:153-166) it calls, for FP32
// compute.  The reference's DF2T update (_lfilter.py:158-166) is used by the
// FP64 kernel in sosfilt_fwd.hip; this kernel computes the same linear
// recurrence in direct form I, which takes one operation fewer per sample.
//
// Work decomposition: one wavefront = one window of 64 lanes x 64
// consecutive samples of one channel, held in registers for the whole
// cascade; a workgroup of NWV waves takes NWV consecutive windows of one
// channel (a group: 4 windows with at least kChainedChannels channels, 8
// with fewer), workgroup i = g * B + c (group-major).  Per section:
//   1. zero-state sweep of the lane's samples, which first adds the pending
//      zero-input response of the previous section (2 FMAs per sample, from
//      the section's float table g) and then runs direct form I with the
//      first input tap normalised to 1 (4 FMAs per sample):
//        y[n] = u[n] + b1' u[n-1] + b2' u[n-2] - a1 y[n-1] - a2 y[n-2].
//      With b' = b / b0 the section outputs y / b0; the factor is carried
//      into the next section, so the whole cascade runs scaled by the
//      running product p_s of the b0 (table entry Fast32::p), which is
//      restored on output, on stored states and on saved section outputs.
//      Filters with a zero b0 or a running product outside [2^-50, 2^50]
//      are not `unit` and keep b0 as a tap (5 FMAs per sample, p = 1).
//   2. FP32 scan over the lanes of the affine maps z -> M z + e_q (e_q the
//      lane's zero-state end state in DF2T coordinates, M = A^64):
//      Kogge-Stone inside each 16-lane row with DPP row shifts, then two
//      DPP row broadcasts that carry the row totals across the wave.
//   3. the group's incoming state, in FP64: zi, or the inclusive state of
//      the group before (many channels: read once for all sections up
//      front; few channels: decoupled look-back), or, for the first group
//      of a segment of the segmented form (sosfilt_fwd.hip), the end states
//      of a zero-state pass over the warm-up window before the segment that
//      one wave of the group runs first; composed through LDS with the wave
//      aggregates of the group; the group's inclusive state is published
//      for the groups after it.
//   4. each lane's incoming state M^q s_wave + exclusive_q (M^q per lane
//      from the table) becomes the pending zero-input response.
// x is read from HBM once and y written once, through LDS with non-temporal
// 16-byte accesses (an unpadded, XOR-swizzled 8 KB staging buffer per wave,
// all 16 loads of a lane in flight at once).  Leftover full windows run one
// per workgroup; partial last windows run in the FP64-carry kernel of
// sosfilt_fwd.hip, which reads the inclusive states published here (true,
// unscaled DF2T states).

#include <math.h>

#include <type_traits>
#include <utility>

#include "sos_common.h"
#include "sos_internal.h"
#include "tsiir.h"

namespace tsiir {

namespace {

constexpr int kW32 = 64;                              // samples per lane
constexpr int64_t kChainedChannels = 1024;  // chained groups from this many channels on
constexpr int64_t kWin32 = static_cast<int64_t>(kWave) * kW32;  // samples per window
constexpr int kLine = 128;  // bytes per L2 cache line

// The lane index, re-read at each use: the empty asm hides that it is
// loop-invariant, so values derived from it (LDS slots, lane offsets) are
// formed where they are needed instead of being hoisted out of the section
// loop and held in registers across the cascade.
__device__ __forceinline__ int lane_now() {
  uint32_t t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return static_cast<int>(t & (kWave - 1));
}

// DPP controls (GFX9 encoding)
constexpr int kDppRowShr = 0x110;      // + shift 1..15, within each row of 16 lanes
constexpr int kDppWaveShl1 = 0x130;    // whole wave, one lane towards lower lanes
constexpr int kDppWaveShr1 = 0x138;    // whole wave, one lane towards higher lanes
constexpr int kDppRowBcast15 = 0x142;  // lane 15 of each row to the next row
constexpr int kDppRowBcast31 = 0x143;  // lane 31 to rows 2 and 3

// v of the DPP source lane; 0 where the source lane is out of range or the
// destination row is masked off.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, true));
}

// e += P * (e of the DPP source lane), P = [[x, y], [z, w]]
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void scan_step(const float4 P, float& e0, float& e1) {
  const float n0 = dpp_f32<CTRL, ROW_MASK>(e0);
  const float n1 = dpp_f32<CTRL, ROW_MASK>(e1);
  e0 = fmaf(P.x, n0, fmaf(P.y, n1, e0));
  e1 = fmaf(P.z, n0, fmaf(P.w, n1, e1));
}

// Inclusive scan of the lanes' end states with zero incoming state,
//   X_q = sum_{p <= q} M^(q-p) e_p,
// in FP32.  Rows of 16 lanes: Kogge-Stone steps 1, 2, 4, 8 with the uniform
// powers M^d.  Across rows: rows 1 and 3 add M^((q mod 16)+1) times the last
// lane of the row before (pa), then rows 2 and 3 add M^(q-31) times lane 31
// (pb), which by then holds the total of rows 0-1.  Returns the exclusive
// value X_(q-1) (0 in lane 0) and the wave total X_63 (uniform).
__device__ __forceinline__ void lane_scan32(const Fast32* __restrict__ f, const float4 pa,
                                            const float4 pb, float e0, float e1, float& x0,
                                            float& x1, float& ag0, float& ag1) {
  scan_step<kDppRowShr + 1, 0xf>(f->mq[1], e0, e1);
  scan_step<kDppRowShr + 2, 0xf>(f->mq[2], e0, e1);
  scan_step<kDppRowShr + 4, 0xf>(f->mq[4], e0, e1);
  scan_step<kDppRowShr + 8, 0xf>(f->mq[8], e0, e1);
  scan_step<kDppRowBcast15, 0xa>(pa, e0, e1);
  scan_step<kDppRowBcast31, 0xc>(pb, e0, e1);
  ag0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e0), kWave - 1));
  ag1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e1), kWave - 1));
  x0 = dpp_f32<kDppWaveShr1, 0xf>(e0);
  x1 = dpp_f32<kDppWaveShr1, 0xf>(e1);
}

// Zero-state sweep of one section over the lane's samples, in three passes
// that each update the sample registers in place (two FMAs per sample each,
// no register copies):
//   correction (PEND), ascending: u[j] = v[j] + g[j+1] w0 + g[j] w1, the
//     previous section's zero-input response;
//   feed-forward, descending: t[j] = u[j] + b1' u[j-1] + b2' u[j-2] (UNIT;
//     b0 u[j] + b1 u[j-1] + b2 u[j-2] otherwise), so u[j-1] and u[j-2] are
//     still in place when t[j] forms;
//   feedback, ascending: y[j] = t[j] - a2 y[j-2] - a1 y[j-1].
// Returns the DF2T state after the last sample n,
//   z0 = b1 u[n] + b2 u[n-1] - a1 y[n] - a2 y[n-1],  z1 = b2 u[n] - a2 y[n].
template <bool UNIT, bool PEND>
__device__ __forceinline__ void df1_sweep(float (&v)[kW32], const Fast32* __restrict__ f,
                                          const float* __restrict__ g, float w0, float w1,
                                          float& z0, float& z1) {
  const float b0 = f->c[0], b1 = f->c[1], b2 = f->c[2], na1 = f->c[3], na2 = f->c[4];
  if constexpr (PEND) {
#pragma unroll
    for (int j = 0; j < kW32; ++j) v[j] = fmaf(g[j + 1], w0, fmaf(g[j], w1, v[j]));
  }
  const float un = v[kW32 - 1], un1 = v[kW32 - 2];
  // the passes stay apart, and the feed-forward pass stays descending:
  // either reordering would need a second register per sample
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int j = kW32 - 1; j >= (UNIT ? 1 : 0); --j) {
    float t = UNIT ? v[j] : b0 * v[j];
    if (j >= 1) t = fmaf(b1, v[j - 1], t);
    if (j >= 2) t = fmaf(b2, v[j - 2], t);
    v[j] = t;
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int j = 1; j < kW32; ++j) {
    float t = v[j];
    if (j >= 2) t = fmaf(na2, v[j - 2], t);
    v[j] = fmaf(na1, v[j - 1], t);
  }
  z0 = fmaf(b1, un, fmaf(b2, un1, fmaf(na1, v[kW32 - 1], na2 * v[kW32 - 2])));
  z1 = fmaf(b2, un, na2 * v[kW32 - 1]);
}

// v = scale * (v + g[j+1] w0 + g[j] w1); SCALE = false: scale == 1.
template <bool SCALE>
__device__ __forceinline__ void apply_corr(float (&v)[kW32], const float* __restrict__ g,
                                           float w0, float w1, float scale) {
  if constexpr (SCALE) {
    w0 *= scale;
    w1 *= scale;
  }
#pragma unroll
  for (int j = 0; j < kW32; ++j) {
    const float b = SCALE ? v[j] * scale : v[j];
    v[j] = fmaf(g[j + 1], w0, fmaf(g[j], w1, b));
  }
}

// ---- window I/O through LDS ------------------------------------------------
//
// A window is 64 rows (lanes) of kNV 16-byte vectors.  It moves in two
// rounds of half-rows (kH vectors per lane per round).  Global side:
// instruction i, lane l handles row q = i * (64 / kH) + l / kH, vector
// t = l % kH of the half-row, so every instruction covers whole 128-byte
// lines.  LDS side: half-row q at q * kH vectors, vector t stored at slot
// t ^ (q mod kH), so a column read by 8 consecutive lanes hits 8 distinct
// slots (no bank conflicts) without padding: the buffer is 64 * kH vectors
// (8 KB for FP32 samples).
template <typename TIO>
struct Stage32 {
  static constexpr int kVec = IOTraits<TIO>::kVec;  // elements per 16 bytes
  static constexpr int kNV = kW32 / kVec;           // vectors per lane
  static constexpr int kH = kNV / 2;                // vectors per lane per round
  static constexpr int kBytes = kWave * kH * 16;
  static_assert(kWave % kH == 0 && (kH & (kH - 1)) == 0, "window shape");
  __device__ static __forceinline__ int grow(int i, int l) { return i * (kWave / kH) + l / kH; }
  __device__ static __forceinline__ int gvec(int r, int i, int l) {
    return grow(i, l) * kNV + r * kH + l % kH;
  }
  __device__ static __forceinline__ int slot(int q, int t) { return q * kH + (t ^ (q & (kH - 1))); }
};

// Wave-local ordering of the staging buffer: each wave stages through its
// own LDS region, and a wave's LDS accesses complete in issue order, so a
// wait for its outstanding LDS operations (and a compiler barrier) suffices.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The window's 16-byte vectors as this lane loads them (both rounds), not
// waited for.
template <typename TIO>
struct RawWin {
  u32x4 r[2][Stage32<TIO>::kH];
};
template <typename TIO>
__device__ __forceinline__ void issue_window32(const TIO* win, RawWin<TIO>& raw) {
  using L = Stage32<TIO>;
  const int lane = lane_now();
  const u32x4* __restrict__ gp = reinterpret_cast<const u32x4*>(win);
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int i = 0; i < L::kH; ++i)
      raw.r[r][i] = __builtin_nontemporal_load(gp + L::gvec(r, i, lane));
}

// Loaded vectors -> registers through the wave's staging buffer: lane q gets
// row q (REV: row 63 - q, elements reversed).  The raw vectors of a round
// are dead once staged, so the samples take their registers.
template <typename TIO, bool REV>
__device__ __forceinline__ void stage_window32(const RawWin<TIO>& raw, u32x4* lds,
                                               float (&v)[kW32]) {
  using L = Stage32<TIO>;
  using vec_t = typename IOTraits<TIO>::vec_t;
  const int lane = lane_now();
  const int q = REV ? kWave - 1 - lane : lane;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int i = 0; i < L::kH; ++i) lds[L::slot(L::grow(i, lane), lane % L::kH)] = raw.r[r][i];
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < L::kH; ++t) {
      const u32x4 rw = lds[L::slot(q, t)];
      const vec_t rv = *reinterpret_cast<const vec_t*>(&rw);
      const TIO* e = reinterpret_cast<const TIO*>(&rv);
#pragma unroll
      for (int j = 0; j < L::kVec; ++j) {
        const int idx = (r * L::kH + t) * L::kVec + j;
        v[REV ? kW32 - 1 - idx : idx] = IOTraits<TIO>::to_f(e[j]);
      }
    }
    wave_lds_sync();
  }
}

// Window -> registers, both rounds' loads in flight at once.
template <typename TIO, bool REV>
__device__ __forceinline__ void load_window32(const TIO* win, u32x4* lds, float (&v)[kW32]) {
  RawWin<TIO> raw;
  issue_window32<TIO>(win, raw);
  stage_window32<TIO, REV>(raw, lds, v);
}

// *p = x; NT: a streaming (non-temporal) store.  Streaming stores suit
// whole 128-byte lines: two of them that write the two parts of one line reach
// memory as partial lines (a window 16 bytes off a line boundary measured 2x
// slower), where regular stores meet in the L2 first.
template <bool NT, typename T>
__device__ __forceinline__ void put(T* p, const T& x) {
  if constexpr (NT)
    __builtin_nontemporal_store(x, p);
  else
    *p = x;
}

// Registers -> window, each value multiplied by `scale`.  NT: streaming
// stores (every window of the launch starts on a 128-byte line).  LEAN: one 16-byte vector in flight at a time (scheduling barriers between
// them), for stores in the middle of the cascade, where the sample registers
// stay live and a batch of staged vectors would need registers of its own.
// ELEM: the staged vectors go out element by element (a window that is not
// 16-byte aligned), still one coalesced store instruction per element slot.
// HEAD: the window's first 16-byte vector is written from element head_skip
// on (the elements before it belong to the window before; see
// store_window32_any).
template <typename TIO, bool REV, bool NT, bool LEAN = false, bool ELEM = false,
          bool HEAD = false>
__device__ __forceinline__ void store_window32(TIO* win, u32x4* lds, const float (&v)[kW32],
                                               float scale, int head_skip = 0) {
  using L = Stage32<TIO>;
  using vec_t = typename IOTraits<TIO>::vec_t;
  const int lane = lane_now();
  u32x4* __restrict__ gp = reinterpret_cast<u32x4*>(win);
  const int q = REV ? kWave - 1 - lane : lane;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
#pragma unroll
    for (int t = 0; t < L::kH; ++t) {
      vec_t rv;
      TIO* e = reinterpret_cast<TIO*>(&rv);
#pragma unroll
      for (int j = 0; j < L::kVec; ++j) {
        const int idx = (r * L::kH + t) * L::kVec + j;
        e[j] = IOTraits<TIO>::from_f(v[REV ? kW32 - 1 - idx : idx] * scale);
      }
      lds[L::slot(q, t)] = *reinterpret_cast<const u32x4*>(&rv);
      if constexpr (LEAN) __builtin_amdgcn_sched_barrier(0);
    }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < L::kH; ++i) {
      const u32x4 o = lds[L::slot(L::grow(i, lane), lane % L::kH)];
      // element bits (FP32: 32, BF16: 16)
      using raw_t = typename std::conditional<sizeof(TIO) == 2, uint16_t, uint32_t>::type;
      if (HEAD && r == 0 && i == 0 && lane == 0) {
        // vector 0 of the window (round 0, instruction 0, lane 0), in part
        const raw_t* oe = reinterpret_cast<const raw_t*>(&o);
        raw_t* dst = reinterpret_cast<raw_t*>(win);
#pragma unroll
        for (int k = 1; k < L::kVec; ++k)
          if (k >= head_skip) put<NT>(dst + k, oe[k]);
      } else if constexpr (ELEM) {
        const raw_t* oe = reinterpret_cast<const raw_t*>(&o);
        raw_t* dst = reinterpret_cast<raw_t*>(win + static_cast<int64_t>(L::gvec(r, i, lane)) *
                                                        L::kVec);
#pragma unroll
        for (int k = 0; k < L::kVec; ++k) put<NT>(dst + k, oe[k]);
      } else {
        put<NT>(gp + L::gvec(r, i, lane), o);
      }
      if constexpr (LEAN) __builtin_amdgcn_sched_barrier(0);
    }
    wave_lds_sync();
  }
}

// Direct per-lane store of the lane's samples: forward, p[j] = v[j]; REVERSE,
// p[-j] = v[j] (p = the lane's last time).
template <typename TIO>
__device__ __forceinline__ void store_lane32(TIO* p, bool rev, const float (&v)[kW32],
                                             float scale) {
#pragma unroll
  for (int j = 0; j < kW32; ++j) p[rev ? -j : j] = IOTraits<TIO>::from_f(v[j] * scale);
}

// ---- windows that do not start on a 16-byte boundary -------------------------
//
// A window (time-ordered start `win`) r elements past a 16-byte boundary (0 <
// r < kVec) is staged as the aligned run that starts r elements earlier, base
// = win - r, so every 16-byte access stays aligned and coalesced; each lane
// then moves its samples by r in registers and exchanges the r samples at
// its edge with the neighbouring lane (one DPP wave shift per sample).  The
// edge lane takes them from the aligned vector at base + kWin32 instead,
// which holds the window's last r elements.  Every vector read holds at least
// one element of the window (the staged run's first vector holds win[0], the
// extra one win[kWin32 - 1]), and an aligned 16-byte vector never crosses a
// page, so no read can fault beyond the caller's row.  Stores write only the
// window's own elements: the staged run's first vector in part, the last r
// elements one by one.
//
// Forward (lane q holds times q*64 .. q*64+63 of the window): the staged lane
// holds A[j] = win[q*64 + j - r], so v[j] = A[j + r] and the top r come from
// lane q+1's A[0 .. r-1] (lane 63: the extra vector).  REV (lane q holds the
// window's times 4095 - q*64 - j): A[j] = time 4095 - q*64 - j - r, so v[j] =
// A[j - r] and the bottom r come from lane q-1's A[64-r .. 63] (lane 0: the
// extra vector, reversed).

// The moves by the runtime r (wave-uniform, 0 <= r < KV) are branch-free:
// one select per element for each bit of r, over the lane's samples extended
// by KV neighbouring ones.  Every index of the selects is a constant of a
// fold expression: written as a loop, a select between two elements of the
// array is turned into a load through a selected address before the loop
// unrolls, and the array then stays in scratch memory.

// f[j] = c ? f[j + S] : f[j], j ascending
template <int S, int N, size_t... J>
__device__ __forceinline__ void select_from_above(float (&f)[N], bool c,
                                                  std::index_sequence<J...>) {
  ((f[J] = c ? f[J + S] : f[J]), ...);
}
// f[j] = c ? f[j - S] : f[j], j descending
template <int S, int N, size_t... J>
__device__ __forceinline__ void select_from_below(float (&f)[N], bool c,
                                                  std::index_sequence<J...>) {
  ((f[N - 1 - J] = c ? f[N - 1 - J - S] : f[N - 1 - J]), ...);
}
// f[j] = f[j + m] for j < N - m (0 <= m < KV <= 8, KV a power of two)
template <int KV, int N>
__device__ __forceinline__ void move_down(float (&f)[N], int m) {
  if constexpr (KV > 1) select_from_above<1>(f, (m & 1) != 0, std::make_index_sequence<N - 1>{});
  if constexpr (KV > 2) select_from_above<2>(f, (m & 2) != 0, std::make_index_sequence<N - 2>{});
  if constexpr (KV > 4) select_from_above<4>(f, (m & 4) != 0, std::make_index_sequence<N - 4>{});
}
// f[j] = f[j - m] for j >= m (0 <= m < KV <= 8, KV a power of two)
template <int KV, int N>
__device__ __forceinline__ void move_up(float (&f)[N], int m) {
  if constexpr (KV > 1) select_from_below<1>(f, (m & 1) != 0, std::make_index_sequence<N - 1>{});
  if constexpr (KV > 2) select_from_below<2>(f, (m & 2) != 0, std::make_index_sequence<N - 2>{});
  if constexpr (KV > 4) select_from_below<4>(f, (m & 4) != 0, std::make_index_sequence<N - 4>{});
}

// v[j] = (v ++ ext)[j + r]; ext[k] = the KV samples after v[kW32 - 1].
template <int KV>
__device__ __forceinline__ void shift_down_any(float (&v)[kW32], const float (&ext)[8], int r) {
  float f[kW32 + KV];
#pragma unroll
  for (int j = 0; j < kW32; ++j) f[j] = v[j];
#pragma unroll
  for (int k = 0; k < KV; ++k) f[kW32 + k] = ext[k];
  move_down<KV>(f, r);
#pragma unroll
  for (int j = 0; j < kW32; ++j) v[j] = f[j];
}
// v[j] = (ext ++ v)[KV + j - r]; ext[k] = the KV samples before v[0], in
// order (ext[KV - 1] is next to v[0]).
template <int KV>
__device__ __forceinline__ void shift_up_any(float (&v)[kW32], const float (&ext)[8], int r) {
  float f[KV + kW32];
#pragma unroll
  for (int k = 0; k < KV; ++k) f[k] = ext[k];
#pragma unroll
  for (int j = 0; j < kW32; ++j) f[KV + j] = v[j];
  move_up<KV>(f, r);
#pragma unroll
  for (int j = 0; j < kW32; ++j) v[j] = f[KV + j];
}

// v of the lane a wave shift (CTRL) names, evaluated in every lane before the
// caller selects: left to itself the compiler may run the DPP move only in
// the lanes that keep its result, and a source lane outside EXEC reads as 0.
template <int CTRL>
__device__ __forceinline__ float dpp_all_lanes(float v) {
  float d = dpp_f32<CTRL, 0xf>(v);
  asm volatile("" : "+v"(d));
  return d;
}

// The KV samples after this lane's v (lane q+1's v[0 .. KV-1]), or tail[] in
// lane 63.
template <int KV>
__device__ __forceinline__ void next_lane_head(const float (&v)[kW32], const float (&tail)[8],
                                               float (&ext)[8]) {
  const bool top = lane_now() == kWave - 1;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k < KV) {
      const float d = dpp_all_lanes<kDppWaveShl1>(v[k]);
      ext[k] = top ? tail[k] : d;
    } else {
      ext[k] = 0.0f;
    }
}
// The KV samples before this lane's v (lane q-1's v[kW32-KV .. kW32-1]), or
// head[] in lane 0.
template <int KV>
__device__ __forceinline__ void prev_lane_tail(const float (&v)[kW32], const float (&head)[8],
                                               float (&ext)[8]) {
  const bool bottom = lane_now() == 0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k < KV) {
      const float d = dpp_all_lanes<kDppWaveShr1>(v[kW32 - KV + k]);
      ext[k] = bottom ? head[k] : d;
    } else {
      ext[k] = 0.0f;
    }
}

// Elements (converted) of the aligned vector at p.
template <typename TIO>
__device__ __forceinline__ void load_vec_f(const TIO* p, float (&out)[8]) {
  using vec_t = typename IOTraits<TIO>::vec_t;
  const vec_t rv = *reinterpret_cast<const vec_t*>(p);
  const TIO* e = reinterpret_cast<const TIO*>(&rv);
#pragma unroll
  for (int k = 0; k < 8; ++k) out[k] = k < IOTraits<TIO>::kVec ? IOTraits<TIO>::to_f(e[k]) : 0.0f;
}

// Window -> registers at any element alignment of `win` (its time-ordered
// start).
template <typename TIO, bool REV>
__device__ __forceinline__ void load_window32_any(const TIO* win, u32x4* lds, float (&v)[kW32]) {
  constexpr int kV = IOTraits<TIO>::kVec;
  const int r = static_cast<int>((reinterpret_cast<uintptr_t>(win) & 15u) / sizeof(TIO));
  if (r == 0) {
    load_window32<TIO, REV>(win, lds, v);
    return;
  }
  const TIO* base = win - r;
  load_window32<TIO, REV>(base, lds, v);
  float x[8];
  load_vec_f<TIO>(base + kWin32, x);  // win[kWin32 - r .. kWin32 - 1] in x[0 .. r-1]
  float ext[8];
  if constexpr (REV) {
    // lane 0: the samples before v[0] are the window's last times, x
    // reversed (only head[kV - r ..] are used)
    float head[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) head[k] = k < kV ? x[kV - 1 - k] : 0.0f;
    prev_lane_tail<kV>(v, head, ext);
    shift_up_any<kV>(v, ext, r);
  } else {
    next_lane_head<kV>(v, x, ext);
    shift_down_any<kV>(v, ext, r);
  }
}

// Registers -> window at any element alignment of `win` (time-ordered start);
// v is consumed.
template <typename TIO, bool REV>
__device__ __forceinline__ void store_window32_any(TIO* win, u32x4* lds, float (&v)[kW32]) {
  constexpr int kV = IOTraits<TIO>::kVec;
  const int r = static_cast<int>((reinterpret_cast<uintptr_t>(win) & 15u) / sizeof(TIO));
  if (r == 0) {
    store_window32<TIO, REV, false>(win, lds, v, 1.0f);
    return;
  }
  TIO* base = win - r;
  // the window's last r elements, win[kWin32 - r + i], from the edge lane
  // (forward: lane 63's top samples v[64 - r + i]; REV: lane 0's bottom
  // samples, v[r - 1 - i]), before the move below overwrites them
  if (lane_now() == (REV ? 0 : kWave - 1)) {
    // g[k] = forward v[kW32 - kV + k], REV v[kV - 1 - k]: win[kWin32 - r + i]
    // is g[kV - r + i]
    float g[2 * kV];
#pragma unroll
    for (int k = 0; k < kV; ++k) {
      g[k] = REV ? v[kV - 1 - k] : v[kW32 - kV + k];
      g[kV + k] = 0.0f;
    }
    move_down<kV>(g, kV - r);
#pragma unroll
    for (int i = 0; i < kV; ++i)
      if (i < r) win[kWin32 - r + i] = IOTraits<TIO>::from_f(g[i]);
  }
  // aligned layout: forward A[j] = v[j - r] (lane 0's first r belong to the
  // window before and are not stored), REV A[j] = v[j + r] (lane 63's last
  // r likewise)
  const float zero[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  float ext[8];
  if constexpr (REV) {
    next_lane_head<kV>(v, zero, ext);
    shift_down_any<kV>(v, ext, r);
  } else {
    prev_lane_tail<kV>(v, zero, ext);
    shift_up_any<kV>(v, ext, r);
  }
  store_window32<TIO, REV, false, false, false, true>(base, lds, v, 1.0f, r);
}

// Per-wave state shared by the section steps.  A workgroup of NWV waves
// filters NWV consecutive windows of one channel (a group); wave v takes
// window k = kg + v.
template <int NWV>
struct Win32 {
  int64_t kg, c, t0w;  // first window of the group, channel, first sample of this wave
  int wv;              // wave within the group
  const Fast32* ff;    // the channel's filter, section 0
  const double* tf;    // its FP64 table entries
  int64_t zi_row;
  bool has_zi, write_zf;
  uint32_t flag;       // epoch of this launch
  // chained form: the group's incoming state of every section, read up front
  // (zi or zeros for the first group, else window kg-1's inclusive state)
  const double* prev;  // (S, 2)
  const int* prev_ok;  // (S): all words carried the current flag
  double* wagg;        // (2, NWV, 2): wave aggregates, double-buffered by section
  double* sgrp;        // (2, 2): the group's incoming state (look-back form)
  u32x4* stage;        // this wave's staging buffer
};

// log2 of the windows per group (1, 2, 4 or 8): a group spans A^(4096 NWV),
// table power 6 + log2(NWV).
template <int NWV>
constexpr int kLog2Nwv = NWV == 8 ? 3 : (NWV == 4 ? 2 : (NWV == 2 ? 1 : 0));
// Predecessor groups inspected per look-back round.  The status words are
// read with coherent loads that go past the XCD's L2; their number, more than
// their latency, bounds the look-back form (long signals, few channels):
// 4 per round measured faster than 8, 16 or 64.
constexpr int kLbLanes = 4;
constexpr int kLbLog2 = 2;
static_assert((1 << kLbLog2) == kLbLanes, "look-back lanes");

// Decoupled look-back over the groups before group gi of channel c (NWV
// windows each; a group's status words sit at the index of its last window):
// the incoming state of group gi, the same value in every lane.  Per round,
// lane l < kLbLanes inspects group gi-1-l; the nearest group with an
// inclusive state (or the virtual group -1 holding zi) ends the search,
// groups in between contribute their aggregates,
//   in = sum_{l < f} Mg^l agg_(gi-1-l) + Mg^f incl_(gi-1-f),
// Mg = A^(4096 NWV), combined by a shuffle-down reduction with the powers
// Mg^(2^i); without an inclusive state in the round the next kLbLanes groups
// follow, scaled by Mg^kLbLanes.  A wait longer than kSpinLimit polls sets
// the error word and gives up, so every wave terminates.
template <int NWV>
__device__ __forceinline__ void lookback_groups(const uint64_t* status, int64_t B, int S,
                                                WsHeader* hdr, int64_t gi, int64_t c, int s,
                                                uint32_t flag, uint32_t flags, const double* tb,
                                                double zi0, double zi1, double& sw0,
                                                double& sw1) {
  constexpr int kPbase = 6 + kLog2Nwv<NWV>;
  const int lane = lane_now();
  if (gi >= 1 && (debug_give_up(flags) || span_power_is_zero(tb, kPbase))) {
    lookback_predecessor(status + ((((gi - 1) * NWV + NWV - 1) * B + c) * S + s) * kStatusWords,
                         hdr, flag, flags, sw0, sw1);
    return;
  }
  double acc0 = 0.0, acc1 = 0.0;
  Mat2 P{1.0, 0.0, 0.0, 1.0};
  int64_t base = gi - 1;
  uint32_t spins = 0;
  while (true) {
    const int64_t j = base - lane;
    int st = 2;
    double v0 = 0.0, v1 = 0.0;
    if (lane >= kLbLanes) {
      st = 1;  // outside this round: ready, contributes nothing
    } else if (j >= 0) {
      const uint64_t* e = status + (((j * NWV + NWV - 1) * B + c) * S + s) * kStatusWords;
      if (read_pair(e + 4, flag, v0, v1))
        st = 2;
      else if (read_pair(e, flag, v0, v1))
        st = 1;
      else
        st = 0;
    } else if (j == -1) {
      v0 = zi0;
      v1 = zi1;
    } else {
      st = 1;  // before the first group: nothing
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
    if (lane > f) {
      v0 = 0.0;
      v1 = 0.0;
    }
#pragma unroll
    for (int i = 0; i < kLbLog2; ++i) {
      const int d = 1 << i;
      const double n0 = __shfl_down(v0, d);
      const double n1 = __shfl_down(v1, d);
      const Mat2 Q = load_mat(table_pow(tb, kPbase + i));
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
    P = mat_mul(P, load_mat(table_pow(tb, kPbase + kLbLog2)));
    base -= kLbLanes;
  }
  sw0 = acc0;
  sw1 = acc1;
}

// The lane's powers of M for section s: pm[0] = M^((q mod 16)+1) and
// pm[1] = M^(max(q-31, 0)) for the scan, pm[2] = M^q for the incoming state.
// Loaded one section ahead so their latency hides behind the sweep.
struct LanePow {
  float4 pm[3];
};
__device__ __forceinline__ LanePow lane_pow(const float4* mq) {
  const int lane = lane_now();
  return LanePow{{mq[(lane & 15) + 1], mq[lane > 31 ? lane - 31 : 0], mq[lane]}};
}

// Steps 2-4 of section s after its sweep: lane scan, composition of the
// wave aggregates of the group with the group's incoming state (FP64),
// publication of the group's inclusive state, and the lane's incoming state
// (pw0, pw1) in the scaled units of the section, which becomes the pending
// zero-input response.
// Look-back form, wave 0: the predecessor group's aggregate words of a
// section whose group-span power is exactly zero (then they hold the group's
// incoming state, see lookback_predecessor), loaded before the section's
// sweep so that the round trip overlaps it.
struct GroupPrefetch {
  PairWords words;
  bool valid;
};
template <int NWV, bool LB>
__device__ __forceinline__ GroupPrefetch prefetch_group_in(const FwdArgs& a,
                                                           const Win32<NWV>& w, int s) {
  GroupPrefetch pf;
  pf.valid = false;
  if constexpr (LB) {
    const double* tb = w.tf + static_cast<int64_t>(s) * kTableStride;
    if (w.wv == 0 && w.kg > 0 && span_power_is_zero(tb, 6 + kLog2Nwv<NWV>)) {
      load_pair_words(a.status + (((w.kg - 1) * a.B + w.c) * a.S + s) * kStatusWords,
                      pf.words);
      pf.valid = true;
    }
  }
  return pf;
}

template <int NWV, bool LB>
__device__ __forceinline__ void section_carry(const FwdArgs& a, const Win32<NWV>& w, int s,
                                              const LanePow& lp, float z0, float z1, float& pw0,
                                              float& pw1, const GroupPrefetch& pf) {
  const int lane = lane_now();
  const int S = a.S;
  const uint32_t flag = w.flag;
  const Fast32* fs = w.ff + s;
  const double* tb = w.tf + static_cast<int64_t>(s) * kTableStride;
  const float4 pa = lp.pm[0], pb = lp.pm[1], pq = lp.pm[2];

  // lane scan (FP32, zero incoming state)
  float x0, x1, ag0, ag1;
  lane_scan32(fs, pa, pb, z0, z1, x0, x1, ag0, ag1);
  const double P = fs->p;
  const double A0 = P * static_cast<double>(ag0);  // window aggregate, true units
  const double A1 = P * static_cast<double>(ag1);
  double* wagg = w.wagg + (s & 1) * NWV * 2;
  if (NWV > 1) {
    if (lane == 0) {
      wagg[2 * w.wv] = A0;
      wagg[2 * w.wv + 1] = A1;
    }
    __syncthreads();
  }

  // the group's incoming state (true units, FP64): zi, or the inclusive
  // state window kg-1 published (chained), or the look-back's result
  double sw0 = 0.0, sw1 = 0.0;
  const Mat2 Mw = load_mat(table_pow(tb, 6));  // A^4096, one window
  if constexpr (LB) {
    // wave 0 publishes the group's aggregate, resolves its incoming state by
    // the look-back, publishes its inclusive state and hands the incoming
    // state to the other waves through LDS
    double* sg = w.sgrp + (s & 1) * 2;
    if (w.wv == 0) {
      double g0 = A0, g1 = A1;
      if (NWV > 1) {
        g0 = wagg[0];
        g1 = wagg[1];
#pragma unroll
        for (int j = 1; j < NWV; ++j) {
          const double n0 = fma(Mw.a, g0, fma(Mw.b, g1, wagg[2 * j]));
          const double n1 = fma(Mw.c, g0, fma(Mw.d, g1, wagg[2 * j + 1]));
          g0 = n0;
          g1 = n1;
        }
      }
      uint64_t* st = a.status + (((w.kg + NWV - 1) * a.B + w.c) * S + s) * kStatusWords;
      publish_pair(st, g0, g1, flag, lane);
      double zi0 = 0.0, zi1 = 0.0;
      if (w.has_zi) {
        const int64_t zi_i = (w.zi_row * S + s) * 2;
        zi0 = zi_in<float>(a, zi_i);
        zi1 = zi_in<float>(a, zi_i + 1);
      }
      sw0 = zi0;
      sw1 = zi1;
      if (w.kg > 0 && !(pf.valid && !debug_give_up(a.flags) &&
                        decode_pair(pf.words, flag, sw0, sw1)))
        lookback_groups<NWV>(a.status, a.B, S, a.hdr, w.kg / NWV, w.c, s, flag, a.flags, tb,
                             zi0, zi1, sw0, sw1);
      const Mat2 Mg = load_mat(table_pow(tb, 6 + kLog2Nwv<NWV>));
      const double i0 = fma(Mg.a, sw0, fma(Mg.b, sw1, g0));
      const double i1 = fma(Mg.c, sw0, fma(Mg.d, sw1, g1));
      publish_pair(st + 4, i0, i1, flag, lane);
      if (w.write_zf && lane == 0) {
        state_out<float>(a.zf, a.flags, (w.c * S + s) * 2, i0);
        state_out<float>(a.zf, a.flags, (w.c * S + s) * 2 + 1, i1);
      }
      if (NWV > 1 && lane == 0) {
        sg[0] = sw0;
        sg[1] = sw1;
      }
    }
    if (NWV > 1) {
      __syncthreads();
      sw0 = sg[0];
      sw1 = sg[1];
    }
  } else {
    sw0 = w.prev[2 * s];
    sw1 = w.prev[2 * s + 1];
    if (!w.prev_ok[s]) {
      const uint64_t* sp = a.status + (((w.kg - 1) * a.B + w.c) * S + s) * kStatusWords;
      wait_inclusive(sp + 4, a.hdr, flag, a.flags, sw0, sw1);
    }
  }
  // this wave's incoming state: the group's, carried over the windows of
  // the waves before it
#pragma unroll
  for (int j = 0; j < NWV - 1; ++j) {
    if (j < w.wv) {
      const double n0 = fma(Mw.a, sw0, fma(Mw.b, sw1, wagg[2 * j]));
      const double n1 = fma(Mw.c, sw0, fma(Mw.d, sw1, wagg[2 * j + 1]));
      sw0 = n0;
      sw1 = n1;
    }
  }
  if (!LB && w.wv == NWV - 1) {
    // the last wave publishes the group's inclusive state (window kg+NWV-1)
    const double i0 = fma(Mw.a, sw0, fma(Mw.b, sw1, A0));
    const double i1 = fma(Mw.c, sw0, fma(Mw.d, sw1, A1));
    uint64_t* st = a.status + (((w.kg + NWV - 1) * a.B + w.c) * S + s) * kStatusWords;
    publish_pair(st + 4, i0, i1, flag, lane);
    if (w.write_zf && lane == 0) {
      state_out<float>(a.zf, a.flags, (w.c * S + s) * 2, i0);
      state_out<float>(a.zf, a.flags, (w.c * S + s) * 2 + 1, i1);
    }
  }

  // the lane's incoming state, scaled units
  const double rp = fs->rp;
  const float sws0 = static_cast<float>(sw0 * rp), sws1 = static_cast<float>(sw1 * rp);
  pw0 = fmaf(pq.x, sws0, fmaf(pq.y, sws1, x0));
  pw1 = fmaf(pq.z, sws0, fmaf(pq.w, sws1, x1));
}

// Section s's exact output in true units, for the backward pass: applies the
// pending correction now (which clears it) and stores p_s * v.
template <typename TIO, int NWV, int SAVED>
__device__ __forceinline__ void keep_section(const FwdArgs& a, const Win32<NWV>& w, int s,
                                             float (&v)[kW32], float& pw0, float& pw1) {
  const double* tb = w.tf + static_cast<int64_t>(s) * kTableStride;
  apply_corr<false>(v, table_g(tb), pw0, pw1, 1.0f);
  pw0 = 0.0f;
  pw1 = 0.0f;
  // LDS-staged like the window output (per-lane stores at a 256-byte stride
  // measured 7% slower on the training step)
  TIO* sw = static_cast<TIO*>(a.saved) + (static_cast<int64_t>(s) * a.B + w.c) * a.T + w.t0w;
  const float scale = static_cast<float>(w.ff[s].p);
  // (the alignment of every saved row is known at launch: one store form per
  // kernel keeps the register allocation of the cascade intact)
  store_window32<TIO, false, SAVED == 1, true, SAVED == 2>(sw, w.stage, v, scale);
}

// All sections of the window (section 0 has no pending correction, the
// others take the previous section's); on return v holds the cascade output
// in scaled units with the last section's correction (pw0, pw1) pending
// (finish_cascade applies it).  SAVED (nonzero): the outputs of
// sections 0..S-2 are also stored, with 1: streaming stores of 16-byte
// vectors (every saved window starts on a 128-byte line), 3: regular stores
// of 16-byte vectors (on 16-byte boundaries), 2: element stores.
template <typename TIO, int SAVED, bool UNIT, int NWV, bool LB>
__device__ __forceinline__ void run_cascade(const FwdArgs& a, const Win32<NWV>& w,
                                            float (&v)[kW32], float& pw0, float& pw1) {
  const int S = a.S;
  float z0, z1;
  pw0 = 0.0f;
  pw1 = 0.0f;
  LanePow lp = lane_pow(w.ff->mq);
  GroupPrefetch pf = prefetch_group_in<NWV, LB>(a, w, 0);
  df1_sweep<UNIT, false>(v, w.ff, table_g(w.tf), 0.0f, 0.0f, z0, z1);
  section_carry<NWV, LB>(a, w, 0, lp, z0, z1, pw0, pw1, pf);
  if (SAVED && S > 1) keep_section<TIO, NWV, SAVED>(a, w, 0, v, pw0, pw1);
  for (int s = 1; s < S; ++s) {
    lp = lane_pow(w.ff[s].mq);
    pf = prefetch_group_in<NWV, LB>(a, w, s);
    const float* pg = table_g(w.tf + static_cast<int64_t>(s - 1) * kTableStride);
    df1_sweep<UNIT, true>(v, w.ff + s, pg, pw0, pw1, z0, z1);
    section_carry<NWV, LB>(a, w, s, lp, z0, z1, pw0, pw1, pf);
    if (SAVED && s + 1 < S) keep_section<TIO, NWV, SAVED>(a, w, s, v, pw0, pw1);
  }
}

// The last section's zero-input response table, and the factor that
// restores true units (UNIT: the running b0 product).
template <int NWV>
__device__ __forceinline__ const float* last_g(const FwdArgs& a, const Win32<NWV>& w) {
  return table_g(w.tf + static_cast<int64_t>(a.S - 1) * kTableStride);
}
template <bool UNIT, int NWV>
__device__ __forceinline__ float out_scale(const FwdArgs& a, const Win32<NWV>& w) {
  return UNIT ? static_cast<float>(w.ff[a.S - 1].p) : 1.0f;
}

// v = the cascade output in true units.
template <bool UNIT, int NWV>
__device__ __forceinline__ void finish_cascade(const FwdArgs& a, const Win32<NWV>& w,
                                               float (&v)[kW32], float pw0, float pw1) {
  apply_corr<UNIT>(v, last_g(a, w), pw0, pw1, out_scale<UNIT>(a, w));
}

// One window per wave: load, cascade, store.  ANY = false: every window of
// x starts on a 16-byte boundary and every window of y on a 128-byte line
// (streaming stores).  ANY: windows at any element alignment, regular stores
// (the moves of load_window32_any / store_window32_any take registers, so
// aligned launches use the plain form).
template <typename TIO, bool REV, int SAVED, bool UNIT, int NWV, bool LB, bool ANY>
__device__ __forceinline__ void window32(const FwdArgs& a, Win32<NWV>& w, u32x4* stage) {
  const TIO* xr =
      static_cast<const TIO*>(a.x) + row_offset(a, w.c % a.x_rows, a.x_stride, a.x_seg_len);
  TIO* yr = static_cast<TIO*>(a.y) + row_offset(a, w.c, a.y_stride, a.y_seg_len);
  // REV: processing index t is time T-1-t, so the window is the time-ordered
  // run [T - t0w - 4096, T - t0w); either way at any element alignment
  const int64_t tw = REV ? a.T - w.t0w - kWin32 : w.t0w;
  float v[kW32];
  if constexpr (ANY) {
    load_window32_any<TIO, REV>(xr + tw, stage, v);
  } else {
    // the launch checked that every window is aligned; the per-lane form
    // below stays as the register allocation of these kernels was measured
    // with it
    const bool xvec = (reinterpret_cast<uintptr_t>(xr + tw) & 15u) == 0;
    const int64_t t0 = w.t0w + static_cast<int64_t>(lane_now()) * kW32;
    if (xvec)
      load_window32<TIO, REV>(xr + tw, stage, v);
    else if (REV)
      load_run_rev<TIO, float, kW32>(xr, a.T - 1 - t0, false, kW32, v);
    else
      load_run<TIO, float, kW32>(xr + t0, false, kW32, v);
  }
  float pw0, pw1;
  run_cascade<TIO, SAVED, UNIT, NWV, LB>(a, w, v, pw0, pw1);
  finish_cascade<UNIT>(a, w, v, pw0, pw1);
  if constexpr (ANY) {
    store_window32_any<TIO, REV>(yr + tw, stage, v);
  } else {
    const bool yvec = (reinterpret_cast<uintptr_t>(yr + tw) & 15u) == 0;
    const int64_t t0 = w.t0w + static_cast<int64_t>(lane_now()) * kW32;
    if (yvec)
      store_window32<TIO, REV, true>(yr + tw, stage, v, 1.0f);
    else if (REV)
      store_lane32<TIO>(yr + (a.T - 1 - t0), true, v, 1.0f);
    else
      store_lane32<TIO>(yr + t0, false, v, 1.0f);
  }
}

// Chained form, wave 0: the group's incoming state of every section into
// LDS: zi (or zeros) for the first group, else the inclusive states window
// kg-1 published; prev_ok = 0 where a word did not carry the current flag
// yet (section_carry then waits for it).
template <int NWV>
__device__ __forceinline__ void read_group_in(const FwdArgs& a, const Win32<NWV>& w,
                                              double* prev, int* prev_ok) {
  const int lane = lane_now();
  const int S = a.S;
  if (w.wv != 0 || lane >= S) return;
  double p0 = 0.0, p1 = 0.0;
  int ok = 1;
  if (w.kg > 0) {
    const uint64_t* st = a.status + (((w.kg - 1) * a.B + w.c) * S + lane) * kStatusWords;
    ok = (read_pair(st + 4, w.flag, p0, p1) && !debug_give_up(a.flags)) ? 1 : 0;
  } else if (w.has_zi) {
    const int64_t zi_i = (w.zi_row * S + lane) * 2;
    p0 = zi_in<float>(a, zi_i);
    p1 = zi_in<float>(a, zi_i + 1);
  }
  prev_ok[lane] = ok;
  prev[2 * lane] = p0;
  prev[2 * lane + 1] = p1;
}

// Segmented form with the fused warm-up: does the group start a segment
// whose incoming states come from a warm-up (every segment but the first one
// processed: k = 0 forward, k = seg_n - 1 along decreasing time)?
template <bool REV, int NWV>
__device__ __forceinline__ bool warm_group(const FwdArgs& a, const Win32<NWV>& w) {
  if (!a.warm_fused || w.kg != 0) return false;
  const int64_t k = w.c % a.seg_n;
  return k != (REV ? a.seg_n - 1 : 0);
}

// The fused warm-up, one wave of such a group: the cascade from
// zero state over the kWin32 samples before the segment (REV: the samples
// after it, along decreasing time), whose end states stand for the
// segment's incoming states because the filter forgets its state within
// that span (checked on the device before the launch).  The same zero-state
// window steps as run_cascade; the end state of section s is the window's
// aggregate in true units, written to prev[] (prev_ok set) where the chained
// form reads the group's incoming states.
// Section s's end state of the warm-up window (its aggregate, zero incoming
// state) in true units, as the group's incoming state of section s.
__device__ __forceinline__ void warm_record(const Fast32* fs, int s, float ag0, float ag1,
                                            double* prev, int* prev_ok) {
  if (lane_now() == 0) {
    prev[2 * s] = fs->p * static_cast<double>(ag0);
    prev[2 * s + 1] = fs->p * static_cast<double>(ag1);
    prev_ok[s] = 1;
  }
}

template <typename TIO, bool REV, bool UNIT, int NWV, bool ANY>
__device__ __forceinline__ void warm_states(const FwdArgs& a, const Win32<NWV>& w, double* prev,
                                            int* prev_ok) {
  const TIO* xr =
      static_cast<const TIO*>(a.x) + row_offset(a, w.c % a.x_rows, a.x_stride, a.x_seg_len);
  // REV: processing indices -kWin32 .. -1 are times T + kWin32 - 1 .. T
  float v[kW32];
  if constexpr (ANY)
    load_window32_any<TIO, REV>(xr + (REV ? a.T : -kWin32), w.stage, v);
  else
    load_window32<TIO, REV>(xr + (REV ? a.T : -kWin32), w.stage, v);
  // section 0 peeled, as in run_cascade
  float z0, z1, x0, x1, ag0, ag1;
  LanePow lp = lane_pow(w.ff->mq);
  df1_sweep<UNIT, false>(v, w.ff, table_g(w.tf), 0.0f, 0.0f, z0, z1);
  lane_scan32(w.ff, lp.pm[0], lp.pm[1], z0, z1, x0, x1, ag0, ag1);
  warm_record(w.ff, 0, ag0, ag1, prev, prev_ok);
  for (int s = 1; s < a.S; ++s) {
    const Fast32* fs = w.ff + s;
    lp = lane_pow(fs->mq);
    // the lane's incoming state of section s-1 with a zero window state is
    // its exclusive scan value: the pending correction
    df1_sweep<UNIT, true>(v, fs, table_g(w.tf + static_cast<int64_t>(s - 1) * kTableStride), x0,
                          x1, z0, z1);
    lane_scan32(fs, lp.pm[0], lp.pm[1], z0, z1, x0, x1, ag0, ag1);
    warm_record(fs, s, ag0, ag1, prev, prev_ok);
  }
}

// Full windows of every channel, NWV consecutive windows (a group) per
// workgroup: workgroup i = g * B + c takes windows k0 + NWV g .. k0 + NWV g +
// NWV - 1 of channel c.  Chained (!LB): a group waits for the inclusive
// states the group before it published (dispatched B workgroups earlier,
// normally finished), read for all sections up front.  LB: every section, the
// group publishes its aggregate and resolves its incoming state by a
// decoupled look-back over the groups before it (few channels: those run at
// the same time).  Inside a group the wave aggregates are composed through
// LDS every section.
// Four-window groups make each workgroup's HBM traffic 64 KB contiguous per
// direction, which streams at the copy rate where single windows scattered
// over the channels do not.  The shared memory is sized to 10 KB per wave so
// that 4 waves share a SIMD (5 measured slower).  The two forms (first input
// tap 1 with the running b0 product, or b0 kept as a tap) branch before
// anything is loaded, so no window data is live across the branch.  REV:
// along decreasing time (TSIIR_REVERSE); ANY: windows at any element
// alignment (fwd32_windows_aligned false); SAVED: keep the section outputs
// (forward_saved), 1 where every saved row is 16-byte aligned, 2 where not.
template <typename TIO, bool REV, int SAVED, int NWV, bool LB, bool ANY>
__global__ __launch_bounds__(64 * NWV) void sos_fwd32_kernel(FwdArgs a,
                                                             const Fast32* __restrict__ fast,
                                                             const double* __restrict__ tables,
                                                             int64_t k0) {
  constexpr int kStageV = Stage32<TIO>::kBytes / 16;
  constexpr int kSmemV = 10240 * NWV / 16;
  constexpr int kAuxV = (2 * kWave * 8 + kWave * 4 + 2 * NWV * 2 * 8 + 2 * 2 * 8) / 16;
  static_assert(NWV * kStageV + kAuxV <= kSmemV, "shared memory layout");
  if (gated_out(a)) return;
  __shared__ u32x4 smem[kSmemV];
  const int S = a.S;
  Win32<NWV> w;
  w.wv = threadIdx.x >> 6;
  w.kg = k0 + static_cast<int64_t>(blockIdx.x) / a.B * NWV;
  w.c = static_cast<int64_t>(blockIdx.x) % a.B;
  const int64_t fi = w.c / a.cpf;
  w.ff = fast + fi * S;
  w.tf = tables + fi * static_cast<int64_t>(S) * kTableStride;
  w.t0w = (w.kg + w.wv) * kWin32;
  w.has_zi = (a.flags & TSIIR_HAS_ZI) != 0;
  w.zi_row = zi_row_of(a, w.c);
  w.write_zf = (a.flags & TSIIR_WRITE_ZF) != 0 && (w.kg + NWV) * kWin32 == a.T;
  double* prev = reinterpret_cast<double*>(smem + NWV * kStageV);
  int* prev_ok = reinterpret_cast<int*>(prev + 2 * kWave);
  w.prev = prev;
  w.prev_ok = prev_ok;
  w.wagg = reinterpret_cast<double*>(prev_ok + kWave);
  w.sgrp = w.wagg + 2 * NWV * 2;
  w.flag = launch_flag(a.hdr, a.epoch_back);
  u32x4* stage = smem + w.wv * kStageV;
  w.stage = stage;
  if (!LB) {
    if (warm_group<REV, NWV>(a, w)) {
      // one wave per group runs the warm-up; which one rotates with the
      // workgroup so that the groups resident on a CU spread it over its SIMDs
      if (w.wv == static_cast<int>(blockIdx.x % NWV)) {
        if (w.ff->unit != 0)
          warm_states<TIO, REV, true, NWV, ANY>(a, w, prev, prev_ok);
        else
          warm_states<TIO, REV, false, NWV, ANY>(a, w, prev, prev_ok);
      }
    } else {
      read_group_in<NWV>(a, w, prev, prev_ok);
    }
  }
  __syncthreads();
  if (w.ff->unit != 0)
    window32<TIO, REV, SAVED, true, NWV, LB, ANY>(a, w, stage);
  else
    window32<TIO, REV, SAVED, false, NWV, LB, ANY>(a, w, stage);
}

// Sixteen lanes per (filter, section) row, four rows per wave: both the FP64
// table entry the kernels of sosfilt_fwd.hip read for 64-sample lanes (the
// layout sos_prep_kernel writes) and the Fast32 entry (see sos_common.h).
// Lane l < 15 forms the table power A^(64 2^l) by repeated squaring (the
// operation sequence of the one-thread prep).  Lane l covers entries
// j = l + 16k (k < 4) of the other two tables, so each store instruction
// writes one contiguous run per row: the zero-input response from the pair
// (g[l+1], g[l]) = C^l (1, 0) (C the companion matrix, binary powering),
// stepped by C^16; M^l by binary powering of M = A^64, stepped by M^16.
// Coefficients are
// normalised by a0 in the coefficient type, then promoted.  The first thread
// advances the workspace epoch by n_epochs (one per following launch that
// publishes status words).
constexpr int kPrepLanes = 16;
constexpr int kPrepRowsPerBlock = 256 / kPrepLanes;

template <typename TC>
__global__ __launch_bounds__(256) void sos_prep32_kernel(const TC* __restrict__ sos,
                                                         int64_t n_rows, int S,
                                                         double* __restrict__ tables,
                                                         Fast32* __restrict__ fast,
                                                         WsHeader* hdr, uint32_t n_epochs,
                                                         const uint32_t* gate, uint32_t gate_on) {
  if (gate != nullptr && ((*gate != 0u) != (gate_on != 0u))) return;
  const int64_t i =
      static_cast<int64_t>(blockIdx.x) * kPrepRowsPerBlock + (threadIdx.x / kPrepLanes);
  const int l = threadIdx.x % kPrepLanes;
  if (blockIdx.x == 0 && threadIdx.x == 0 && hdr != nullptr) hdr->epoch += n_epochs;
  if (i >= n_rows) return;
  const int64_t f = i / S;
  const int s = static_cast<int>(i % S);
  bool unit = true;
  double run = 1.0, ps = 1.0;
  for (int r = 0; r < S; ++r) {
    const TC* row = sos + (f * S + r) * 6;
    TC b0 = row[0];
    const TC a0 = row[3];
    if (a0 != TC(1)) b0 = b0 / a0;
    run *= static_cast<double>(b0);
    const double m = fabs(run);
    if (!(m >= 0x1p-50 && m <= 0x1p50)) unit = false;  // also zero, inf, NaN
    if (r == s) ps = run;
  }
  const TC* row = sos + i * 6;
  TC b0 = row[0], b1 = row[1], b2 = row[2], a0 = row[3], a1 = row[4], a2 = row[5];
  if (a0 != TC(1)) {
    b0 = b0 / a0;
    b1 = b1 / a0;
    b2 = b2 / a0;
    a1 = a1 / a0;
    a2 = a2 / a0;
  }
  const double db0 = b0, db1 = b1, db2 = b2, da1 = a1, da2 = a2;

  // FP64 table entry
  double* t = tables + i * kTableStride;
  const Mat2 A{-da1, 1.0, -da2, 0.0};
  if (l == 0) {
    t[0] = db0;
    t[1] = db1;
    t[2] = db2;
    t[3] = da1;
    t[4] = da2;
    t[5] = static_cast<double>(a0);
    t[6] = t[7] = 0.0;
  }
  static_assert(kNPow <= kPrepLanes, "one table power per lane");
  if (l < kNPow) {  // A^(64 * 2^l)
    Mat2 Q = A;
    for (int k = 0; k < 6 + l; ++k) Q = mat_mul(Q, Q);
    double* q = t + 8 + 4 * l;
    q[0] = Q.a;
    q[1] = Q.b;
    q[2] = Q.c;
    q[3] = Q.d;
  }
  // zero-input response g[0..64] (floats from double kGOff), zeros around it
  const Mat2 C{-da1, -da2, 1.0, 0.0};
  Mat2 R{1.0, 0.0, 0.0, 1.0}, Pw = C;
#pragma unroll
  for (int bit = 0; bit < 4; ++bit) {
    if ((l >> bit) & 1) R = mat_mul(R, Pw);
    Pw = mat_mul(Pw, Pw);
  }
  // Pw = C^16; (gn, gm) = (g[j+1], g[j]) for j = l + 16k
  float* g = reinterpret_cast<float*>(t + kGOff);
  double gn = R.a, gm = R.c;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    g[l + kPrepLanes * k] = static_cast<float>(gm);
    const double n1 = Pw.a * gn + Pw.b * gm;
    const double m1 = Pw.c * gn + Pw.d * gm;
    gn = n1;
    gm = m1;
  }
  if (l == 0) {  // g[64] (now in gm of lane 0), then the table's pad
    g[kWave] = static_cast<float>(gm);
    g[kWave + 1] = 0.0f;
  }
  if (l < kGOff - (8 + 4 * kNPow)) t[8 + 4 * kNPow + l] = 0.0;
  constexpr int kGEnd = kGOff + (kGLen + 1) / 2;  // first double after g
  static_assert(kTableStride - kGEnd <= kPrepLanes, "tail zeros");
  if (kGEnd + l < kTableStride) t[kGEnd + l] = 0.0;

  // Fast32 entry
  Fast32* o = fast + i;
  if (l == 0) {
    if (unit) {
      o->c[0] = 1.0f;
      o->c[1] = static_cast<float>(db1 / db0);
      o->c[2] = static_cast<float>(db2 / db0);
      o->p = ps;
      o->rp = 1.0 / ps;
    } else {
      o->c[0] = static_cast<float>(db0);
      o->c[1] = static_cast<float>(db1);
      o->c[2] = static_cast<float>(db2);
      o->p = 1.0;
      o->rp = 1.0;
    }
    o->c[3] = static_cast<float>(-da1);
    o->c[4] = static_cast<float>(-da2);
    o->unit = unit ? 1 : 0;
    for (int p = 0; p < 6; ++p) o->pad[p] = 0.0f;
  }
  Mat2 M = A;
  for (int k = 0; k < 6; ++k) M = mat_mul(M, M);  // A^64
  Mat2 Q{1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int bit = 0; bit < 4; ++bit) {
    if ((l >> bit) & 1) Q = mat_mul(Q, M);
    M = mat_mul(M, M);
  }
  // M = (A^64)^16
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    o->mq[l + kPrepLanes * k] = make_float4(static_cast<float>(Q.a), static_cast<float>(Q.b),
                                            static_cast<float>(Q.c), static_cast<float>(Q.d));
    if (k < 3) Q = mat_mul(Q, M);
  }
}

}  // namespace

int64_t fwd32_window() { return kWin32; }

// Can the launch take the plain (aligned, streaming-store) kernels: does every
// window (and every warm-up window of the segmented form) start on a 16-byte
// boundary in x and on a 128-byte line in y and in the saved rows?  Windows
// start at a row's offset plus a multiple of kWin32 samples, REV at the row's
// end minus such a multiple, so this is a question about the row bases and T.
static bool fwd32_windows_aligned(const FwdArgs& a, size_t esz, bool rev) {
  auto al = [&](int64_t elems, int64_t bytes) {
    return (elems * static_cast<int64_t>(esz)) % bytes == 0;
  };
  auto ptr_al = [](const void* p, uintptr_t bytes) {
    return (reinterpret_cast<uintptr_t>(p) % bytes) == 0;
  };
  if (!ptr_al(a.x, 16) || !ptr_al(a.y, kLine)) return false;
  if ((a.x_rows > 1 || a.seg_n > 0) && !al(a.x_stride, 16)) return false;
  if ((a.B > 1 || a.seg_n > 0) && !al(a.y_stride, kLine)) return false;
  if (a.seg_n > 0 && (!al(a.x_seg_len, 16) || !al(a.y_seg_len, kLine))) return false;
  if (rev && !al(a.T, kLine)) return false;  // x and y windows end-anchored
  // saved rows (s * B + c) * T, forward only
  if (a.saved != nullptr && (!ptr_al(a.saved, kLine) || !al(a.T, kLine))) return false;
  return true;
}

template <typename TIO, bool REV, int NWV, bool LB>
static void launch_fwd32_io(const FwdArgs& a, const Fast32* fast, int64_t groups, int64_t k0,
                            hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>(groups)), block(kWave * NWV);
  const bool any = !fwd32_windows_aligned(a, sizeof(TIO), REV);
#define TSIIR_LAUNCH_FWD32(SAVED, ANY)                                                       \
  hipLaunchKernelGGL((sos_fwd32_kernel<TIO, REV, SAVED, NWV, LB, ANY>), grid, block, 0, stream, \
                     a, fast, a.tables, k0)
  if (a.saved == nullptr) {
    if (any)
      TSIIR_LAUNCH_FWD32(0, true);
    else
      TSIIR_LAUNCH_FWD32(0, false);
  } else if constexpr (!REV) {  // saved section outputs: forward only
    // saved rows (s * B + c) * T: on 128-byte lines (the plain kernels
    // only), on 16-byte boundaries, or neither
    const bool saved_vec = (reinterpret_cast<uintptr_t>(a.saved) & 15u) == 0 &&
                           (a.T * static_cast<int64_t>(sizeof(TIO))) % 16 == 0;
    if (!any)
      TSIIR_LAUNCH_FWD32(1, false);
    else if (saved_vec)
      TSIIR_LAUNCH_FWD32(3, true);
    else
      TSIIR_LAUNCH_FWD32(2, true);
  }
#undef TSIIR_LAUNCH_FWD32
}

// Many channels: groups of four windows, chained, then the leftover full
// windows one per workgroup, chained to the last group.  Fewer than
// kChainedChannels channels (a group's predecessor runs at the same time):
// groups of eight windows with the look-back (half the look-backs per
// sample of four-window groups; 0.32 against 0.43 ms on 8 x 2^24 samples),
// then the leftover full windows one per workgroup, also with the look-back
// (a chain of them would run one after another).
template <typename TIO, bool REV>
static void launch_fwd32_windows(const FwdArgs& a, const Fast32* fast, hipStream_t stream) {
  const int64_t full = a.T / kWin32;
  if (a.B >= kChainedChannels) {
    const int64_t n4 = full / 4;
    if (n4 > 0) launch_fwd32_io<TIO, REV, 4, false>(a, fast, n4 * a.B, 0, stream);
    if (full > 4 * n4)
      launch_fwd32_io<TIO, REV, 1, false>(a, fast, (full - 4 * n4) * a.B, 4 * n4, stream);
  } else {
    const int64_t n8 = full / 8;
    if (n8 > 0) launch_fwd32_io<TIO, REV, 8, true>(a, fast, n8 * a.B, 0, stream);
    if (full > 8 * n8)
      launch_fwd32_io<TIO, REV, 1, true>(a, fast, (full - 8 * n8) * a.B, 8 * n8, stream);
  }
}

static void prep32(const void* sos, int coef_dtype, int64_t n_filters, int S, double* tables,
                   Fast32* fast, WsHeader* hdr, uint32_t n_epochs, const uint32_t* gate,
                   uint32_t gate_on, hipStream_t stream) {
  const int64_t n_rows = n_filters * S;
  const unsigned pblocks =
      static_cast<unsigned>((n_rows + kPrepRowsPerBlock - 1) / kPrepRowsPerBlock);
  if (coef_dtype == TSIIR_F64)
    hipLaunchKernelGGL(sos_prep32_kernel<double>, dim3(pblocks), dim3(256), 0, stream,
                       static_cast<const double*>(sos), n_rows, S, tables, fast, hdr, n_epochs,
                       gate, gate_on);
  else
    hipLaunchKernelGGL(sos_prep32_kernel<float>, dim3(pblocks), dim3(256), 0, stream,
                       static_cast<const float*>(sos), n_rows, S, tables, fast, hdr, n_epochs,
                       gate, gate_on);
}

void launch_prep32(const void* sos, int coef_dtype, int64_t n_filters, int S, double* tables,
                   Fast32* fast, WsHeader* hdr, uint32_t n_epochs, hipStream_t stream) {
  prep32(sos, coef_dtype, n_filters, S, tables, fast, hdr, n_epochs, nullptr, 0u, stream);
}

tsiir_status launch_fwd32_full(const FwdArgs& a, const void* sos, int coef_dtype,
                               int64_t n_filters, Fast32* fast, bool tables_ready,
                               hipStream_t stream) {
  if (!tables_ready)
    prep32(sos, coef_dtype, n_filters, a.S, const_cast<double*>(a.tables), fast, a.hdr, 1u,
           a.gate, a.gate_on, stream);
  const bool rev = (a.flags & TSIIR_REVERSE) != 0;
  if (a.io_dtype == TSIIR_BF16) {
    if (rev)
      launch_fwd32_windows<bf16_t, true>(a, fast, stream);
    else
      launch_fwd32_windows<bf16_t, false>(a, fast, stream);
  } else {
    if (rev)
      launch_fwd32_windows<float, true>(a, fast, stream);
    else
      launch_fwd32_windows<float, false>(a, fast, stream);
  }
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // namespace tsiir
