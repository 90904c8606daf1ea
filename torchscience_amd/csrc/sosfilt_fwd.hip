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

#include <cmath>
#include <stdlib.h>
#include <type_traits>
#include <string.h>

#include "sos_common.h"
#include "sos_internal.h"
#include "tsiir.h"

namespace tsiir {

// Window shape policy.  Many channels: one wave per window and chained
// windows (a window's predecessor was dispatched a full window-layer earlier
// and is normally complete).  Few channels (long signals): four waves per
// window and the decoupled look-back, which keeps the number of windows in
// flight per channel, and so the look-back depth, small.
struct FwdPolicy {
  int nw;
  bool chained;
};
static bool fwd_lds_io();
// Full windows of FP32 recurrences (FP32 or BF16 I/O, plain mode) run in the
// FP32-recurrence kernel of sosfilt_fwd32.hip, unless TSIIR_FWD32=0 selects
// this file's kernels (kept for comparison measurements).
static bool fwd32_enabled() {
  static const bool on = [] {
    const char* e = getenv("TSIIR_FWD32");
    return !(e && strcmp(e, "0") == 0);
  }();
  return on;
}
static inline bool use_fwd32(const tsiir_sosfilt_desc* d) {
  return (d->io_dtype == TSIIR_F32 || d->io_dtype == TSIIR_BF16) &&
         d->mode == TSIIR_MODE_PLAIN && d->n_sections <= kWave &&
         (d->flags & TSIIR_FP32_DF2T) == 0 && fwd32_enabled() && fwd_lds_io();
}
static inline FwdPolicy fwd_policy(const tsiir_sosfilt_desc* d) {
  if (d->channels >= 1024 && d->n_sections <= 64) return FwdPolicy{1, true};
  if (use_fwd32(d)) return FwdPolicy{1, false};  // status words per 4096-sample window
  return FwdPolicy{4, false};
}

// Coefficients of one section, in the compute type.
template <typename TACC>
struct Coef {
  TACC b0, b1, b2, a1, a2;
  __device__ __forceinline__ explicit Coef(const double* tb)
      : b0(static_cast<TACC>(tb[0])), b1(static_cast<TACC>(tb[1])),
        b2(static_cast<TACC>(tb[2])), a1(static_cast<TACC>(tb[3])),
        a2(static_cast<TACC>(tb[4])) {}
};

// v[j] += zero-input response j samples after state (w0, w1): from the float
// table g (TABG) or by the DF2T homogeneous recurrence with (a1, a2).
template <typename TACC, typename TR, int W, bool TABG>
__device__ __forceinline__ void apply_pending(TACC (&v)[W], const float* g, TR w0, TR w1, TR a1,
                                              TR a2) {
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if constexpr (TABG) {
      v[j] = static_cast<TACC>(fmaf(g[j + 1], w0, fmaf(g[j], w1, static_cast<TR>(v[j]))));
    } else {
      const TR cj = w0;
      w0 = w1 - a1 * cj;
      w1 = -a2 * cj;
      v[j] = static_cast<TACC>(static_cast<TR>(v[j]) + cj);
    }
  }
}

// One workgroup = NW waves = one window of NW*64*W samples of one channel;
// wave w owns the sub-window [w*64*W, (w+1)*64*W) and lane q of it the W
// samples [q*W, (q+1)*W) of that sub-window.
//
// TAIL = false: the full windows of the signal, workgroup i = k * B + c
// (window-major), decoupled look-back over earlier windows.  Workgroups are
// dispatched in index order, so a look-back only waits on windows dispatched
// before it; the bounded spin in lookback() ends every wait regardless.
// TAIL = true: the partial last window of every channel (workgroup i =
// channel i), launched after the full windows; its incoming state is the
// inclusive state published for window k-1, samples at or after T are
// masked and the lane states freeze after the last sample so zf is exact.
//
// CHAINED (NW == 1, full windows): instead of the decoupled look-back, the
// wave takes the inclusive states of window k-1 (normally complete: with many
// channels, window k-1 of a channel was dispatched one window-layer earlier),
// reading every section's state in one round trip up front.  The window's
// incoming state is then folded into lane 0 before the wave scan, so the scan
// yields every lane's incoming state directly and only the inclusive state
// is published.
template <typename TIO, typename TACC, typename TR, int W, int NW, bool LDS, bool TAIL,
          bool CHAINED>
__global__ __launch_bounds__(64 * NW) void sos_fwd_kernel(FwdArgs a, int64_t k_tail,
                                                          const double* __restrict__ tables) {
  constexpr int kLog2NW = NW == 1 ? 0 : NW == 2 ? 1 : NW == 4 ? 2 : 3;
  static_assert((1 << kLog2NW) == NW, "NW must be 1, 2, 4 or 8");
  static_assert(!CHAINED || (NW == 1 && !TAIL), "chained windows are single-wave");
  static_assert(W < kGLen, "the zero-input response table covers 64 samples");
  if (gated_out(a)) return;
  __shared__ __align__(16) TIO lds[(LDS && !TAIL) ? NW * LdsStage<TIO, W>::kElems : 8];
  __shared__ double wagg[NW][2];  // wave aggregates of the current section
  __shared__ double wsin[2];      // incoming state of the window
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x >> 6;
  const uint32_t flag = launch_flag(a.hdr, a.epoch_back);
  const int64_t k = TAIL ? k_tail : static_cast<int64_t>(blockIdx.x) / a.B;
  const int64_t c = TAIL ? static_cast<int64_t>(blockIdx.x)
                         : static_cast<int64_t>(blockIdx.x) % a.B;
  const int64_t wwin = static_cast<int64_t>(kWave) * W;
  const int64_t t0w = k * NW * wwin + wv * wwin;  // first sample of this wave
  const int64_t t0 = t0w + static_cast<int64_t>(lane) * W;
  int m = W;
  if (TAIL) {
    const int64_t rem = a.T - t0;
    m = rem <= 0 ? 0 : (rem >= W ? W : static_cast<int>(rem));
  }

  const TIO* xr =
      static_cast<const TIO*>(a.x) + row_offset(a, c % a.x_rows, a.x_stride, a.x_seg_len);
  TIO* yr = static_cast<TIO*>(a.y) + row_offset(a, c, a.y_stride, a.y_seg_len);
  // REVERSE: processing index t maps to time T-1-t (input and output), so
  // the wave's samples are the time-ordered run ending at T-1-t0w.
  const bool rev = (a.flags & TSIIR_REVERSE) != 0;
  const bool tvec = !rev || (a.T % IOTraits<TIO>::kVec) == 0;
  // a lane whose W samples all lie before T reads and writes them as 16-byte
  // vectors (its run starts at a multiple of W, so the row's alignment is its
  // own); in the TAIL variant the lane that holds sample T-1 goes element by
  // element and the lanes after it do nothing
  const bool whole = !TAIL || m == W;
  const bool xvec = whole && tvec && (reinterpret_cast<uintptr_t>(xr) & 15u) == 0;
  const bool yvec = whole && tvec && (reinterpret_cast<uintptr_t>(yr) & 15u) == 0;
  // the partial window's whole runs at any alignment (load_run_any); the
  // full-window variants keep element runs there, as the register moves
  // would cost them 20-40 VGPRs and an occupancy step
  const bool any_run = TAIL && whole;
  const int64_t t_rev = a.T - t0w - wwin;  // first time of the wave's run (REVERSE)
  TIO* my_lds = lds + ((LDS && !TAIL) ? wv * LdsStage<TIO, W>::kElems : 0);

  TACC v[W];
  // (partial window: whole runs of rows that are not 16-byte aligned move
  // as aligned vectors, the lane that holds sample T-1 element by element)
  if (rev) {
    if (LDS && !TAIL && xvec)
      load_window_lds<TIO, TACC, W, true>(xr + t_rev, my_lds, v);
    else if (any_run && !xvec)
      load_run_any<TIO, TACC, W, true>(xr + (a.T - t0 - W), v);
    else
      load_run_rev<TIO, TACC, W>(xr, a.T - 1 - t0, xvec, m, v);
  } else if (LDS && !TAIL && xvec) {
    load_window_lds<TIO, TACC, W>(xr + t0w, my_lds, v);
  } else if (any_run && !xvec) {
    load_run_any<TIO, TACC, W>(xr + t0, v);
  } else {
    load_run<TIO, TACC, W>(xr + t0, xvec, m, v);
  }

  const double* tf = tables + (c / a.cpf) * static_cast<int64_t>(a.S) * kTableStride;
  const bool has_zi = (a.flags & TSIIR_HAS_ZI) != 0;
  const bool write_zf = (a.flags & TSIIR_WRITE_ZF) != 0 &&
                        (TAIL ? (t0 < a.T && t0 + W >= a.T) : (t0 + W == a.T));
  const int64_t zi_row = zi_row_of(a, c);

  // CHAINED and TAIL: lane l < S fetches window k-1's inclusive state of
  // section l (one round trip for all sections)
  double pv0 = 0.0, pv1 = 0.0;
  bool pok = true;
  if ((CHAINED || TAIL) && k > 0 && lane < a.S) {
    const uint64_t* st = a.status + (((k - 1) * a.B + c) * a.S + lane) * kStatusWords;
    pok = read_pair(st + 4, flag, pv0, pv1);
  }

  // The zero-input response that corrects section s is applied inside the
  // pass of section s+1 (one sweep over the registers per section).  FP32
  // recurrences read it from the section's float table g (two FMAs per
  // sample: the table is uniform across lanes and scalar-loaded); FP64
  // recurrences run its 2-term recurrence.  The TAIL variant applies each
  // correction at once, so none is pending in its sweep.
  constexpr bool kTabG = std::is_same<TR, float>::value;
  TR pw0 = 0, pw1 = 0, pa1 = 0, pa2 = 0;
  const float* pg = table_g(tf);  // g of the section whose correction is pending
  for (int s = 0; s < a.S; ++s) {
    const double* tb = tf + static_cast<int64_t>(s) * kTableStride;
    const Coef<TR> q(tb);

    // 1. correction of section s-1, then the zero-state DF2T pass of section s
    //    (the update order of _lfilter.py:158-166, with fused multiply-adds)
    TR z0 = 0, z1 = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (!TAIL || j < m) {
        TR u = static_cast<TR>(v[j]);
        if constexpr (TAIL) {
        } else if constexpr (kTabG) {
          u = fma3s(pg[j + 1], pw0, fma3s(pg[j], pw1, u));
        } else {
          const TR cj = pw0;
          pw0 = pw1 - pa1 * cj;
          pw1 = -pa2 * cj;
          u += cj;
        }
        const TR yv = fma3(q.b0, u, z0);
        z0 = q.b1 * u + (z1 - q.a1 * yv);
        z1 = q.b2 * u + (-q.a2 * yv);
        v[j] = static_cast<TACC>(yv);
      }
    }

    double s0, s1;  // the lane's true incoming state
    if (CHAINED) {
      // window incoming state: zi (window 0) or window k-1's inclusive state
      double sw0 = 0.0, sw1 = 0.0;
      if (k == 0) {
        if (has_zi) {
          const int64_t zi_i = (zi_row * a.S + s) * 2;
          sw0 = zi_in<TACC>(a, zi_i);
          sw1 = zi_in<TACC>(a, zi_i + 1);
        }
      } else {
        sw0 = __shfl(pv0, s);
        sw1 = __shfl(pv1, s);
        if (!__shfl(static_cast<int>(pok), s) || debug_give_up(a.flags)) {
          const uint64_t* st = a.status + (((k - 1) * a.B + c) * a.S + s) * kStatusWords;
          wait_inclusive(st + 4, a.hdr, flag, a.flags, sw0, sw1);
        }
      }
      // lane 0 carries the window's incoming state: e_0 += M sw
      double e0 = static_cast<double>(z0), e1 = static_cast<double>(z1);
      if (lane == 0) {
        const Mat2 M = load_mat(table_pow(tb, 0));
        e0 += M.a * sw0 + M.b * sw1;
        e1 += M.c * sw0 + M.d * sw1;
      }
      double incl0, incl1;
      wave_scan(tb, lane, e0, e1, incl0, incl1, s0, s1);
      if (lane == 0) {
        s0 = sw0;
        s1 = sw1;
      }
      uint64_t* st = a.status + ((k * a.B + c) * a.S + s) * kStatusWords;
      publish_pair(st + 4, incl0, incl1, flag, lane);
    } else {
    // 2. wave scan of the lanes' end states, then the waves of the window
    double agg0, agg1, x0, x1;
    wave_scan(tb, lane, static_cast<double>(z0), static_cast<double>(z1), agg0, agg1, x0, x1);
    if (NW > 1) {
      if (lane == 0) {
        wagg[wv][0] = agg0;
        wagg[wv][1] = agg1;
      }
      __syncthreads();
    }
    const Mat2 Mw = load_mat(table_pow(tb, 6));  // A^(64W): one wave's sub-window
    double e0 = 0.0, e1 = 0.0, g0 = agg0, g1 = agg1;
    if (NW > 1) {
      g0 = 0.0;
      g1 = 0.0;
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        if (j == wv) {
          e0 = g0;
          e1 = g1;
        }
        const double n0 = Mw.a * g0 + Mw.b * g1 + wagg[j][0];
        const double n1 = Mw.c * g0 + Mw.d * g1 + wagg[j][1];
        g0 = n0;
        g1 = n1;
      }
    }

    // 3. window incoming state: zi for window 0, else the look-back
    if (wv == 0) {
      double zi0 = 0.0, zi1 = 0.0;
      if (has_zi) {
        const int64_t zi_i = (zi_row * a.S + s) * 2;
        zi0 = zi_in<TACC>(a, zi_i);
        zi1 = zi_in<TACC>(a, zi_i + 1);
      }
      double sw0 = zi0, sw1 = zi1;
      if (!TAIL) {
        uint64_t* st = a.status + ((k * a.B + c) * a.S + s) * kStatusWords;
        publish_pair(st, g0, g1, flag, lane);
        if (k > 0)
          lookback(a.status, a.B, a.S, a.hdr, k, c, s, flag, tb, 6 + kLog2NW, zi0, zi1, sw0,
                   sw1, a.flags);
        const Mat2 Mwin = load_mat(table_pow(tb, 6 + kLog2NW));
        publish_pair(st + 4, Mwin.a * sw0 + Mwin.b * sw1 + g0, Mwin.c * sw0 + Mwin.d * sw1 + g1,
                     flag, lane);
      } else if (k > 0) {
        sw0 = __shfl(pv0, s);
        sw1 = __shfl(pv1, s);
        if (!__shfl(static_cast<int>(pok), s)) give_up(a.hdr, 2u, sw0, sw1);
      }
      if (NW > 1 && lane == 0) {
        wsin[0] = sw0;
        wsin[1] = sw1;
      }
      if (NW == 1) {
        e0 = sw0;
        e1 = sw1;
      }
    }
    double sv0, sv1;
    if (NW > 1) {
      __syncthreads();
      // wave incoming state: Mw^wv s_window + exclusive offset of the wave
      double t0v = wsin[0], t1v = wsin[1];
#pragma unroll
      for (int i = 0; i < kLog2NW; ++i) {
        const Mat2 P = load_mat(table_pow(tb, 6 + i));
        if ((wv >> i) & 1) {
          const double u0 = P.a * t0v + P.b * t1v;
          const double u1 = P.c * t0v + P.d * t1v;
          t0v = u0;
          t1v = u1;
        }
      }
      sv0 = t0v + e0;
      sv1 = t1v + e1;
    } else {
      sv0 = e0;
      sv1 = e1;
    }
    lane_incoming(tb, lane, sv0, sv1, x0, x1, s0, s1);
    }

    // 4. the zero-input response from the lane's incoming state is pending
    pw0 = static_cast<TR>(s0);
    pw1 = static_cast<TR>(s1);
    pa1 = q.a1;
    pa2 = q.a2;
    pg = table_g(tb);
    if (TAIL) {
      // apply now, freezing the homogeneous state after the last sample
      TR w0 = pw0, w1 = pw1;
#pragma unroll
      for (int j = 0; j < W; ++j) {
        if (j < m) {
          const TR cj = w0;
          v[j] = static_cast<TACC>(static_cast<TR>(v[j]) + cj);
          w0 = w1 - q.a1 * cj;
          w1 = -q.a2 * cj;
        }
      }
      pw0 = 0;
      pw1 = 0;
      if (write_zf) {
        const int64_t zf_i = (c * a.S + s) * 2;
        state_out<TACC>(a.zf, a.flags, zf_i, static_cast<double>(w0 + z0));
        state_out<TACC>(a.zf, a.flags, zf_i + 1, static_cast<double>(w1 + z1));
      }
    } else if (write_zf) {
      // state after the window's last sample: zero-state part + A^W s_q
      const Mat2 M = load_mat(table_pow(tb, 0));
      const int64_t zf_i = (c * a.S + s) * 2;
      state_out<TACC>(a.zf, a.flags, zf_i, static_cast<double>(z0) + M.a * s0 + M.b * s1);
      state_out<TACC>(a.zf, a.flags, zf_i + 1, static_cast<double>(z1) + M.c * s0 + M.d * s1);
    }
    if (a.saved != nullptr && s + 1 < a.S) {
      // make the registers exact now (the pending correction becomes zero)
      // and keep section s's output for the backward pass
      if (!TAIL) {
        apply_pending<TACC, TR, W, kTabG>(v, pg, pw0, pw1, pa1, pa2);
        pw0 = 0;
        pw1 = 0;
      }
      TIO* sr = static_cast<TIO*>(a.saved) + (static_cast<int64_t>(s) * a.B + c) * a.T;
      const bool svec = whole && (reinterpret_cast<uintptr_t>(sr) & 15u) == 0;
      if (LDS && !TAIL && svec)
        store_window_lds<TIO, TACC, W>(sr + t0w, my_lds, v);
      else if (any_run && !svec)
        store_run_any<TIO, TACC, W>(sr + t0, v);
      else
        store_run<TIO, TACC, W>(sr + t0, svec, m, v);
    }
    if (NW > 1) __syncthreads();  // wagg / wsin are reused by the next section
  }
  if (!TAIL) apply_pending<TACC, TR, W, kTabG>(v, pg, pw0, pw1, pa1, pa2);

  if (rev) {
    if (LDS && !TAIL && yvec)
      store_window_lds<TIO, TACC, W, true>(yr + t_rev, my_lds, v);
    else if (any_run && !yvec)
      store_run_any<TIO, TACC, W, true>(yr + (a.T - t0 - W), v);
    else
      store_run_rev<TIO, TACC, W>(yr, a.T - 1 - t0, yvec, m, v);
  } else if (LDS && !TAIL && yvec) {
    store_window_lds<TIO, TACC, W>(yr + t0w, my_lds, v);
  } else if (any_run && !yvec) {
    store_run_any<TIO, TACC, W>(yr + t0, v);
  } else {
    store_run<TIO, TACC, W>(yr + t0, yvec, m, v);
  }
}

// ---- channel-persistent form ----------------------------------------------
//
// One wave per channel walks the channel's windows in order.  The incoming
// state of every section at the next window is the inclusive state the wave
// scan produced for the current one, so it stays in registers (lane s holds
// section s's) and no wave waits on another.  The next window's input is
// loaded into registers while the current window is filtered, so with as
// many channels as the GPU holds waves the HBM traffic overlaps the VALU
// work instead of arriving in bursts at window boundaries.

// One window of the channel-persistent form.  FULL: 64*W samples, the
// corrections deferred into the next section's sweep.  !FULL: the partial
// last window (lane q has m samples), corrections applied at once, and the
// state after the last sample taken from lane `owner`.  (ws0, ws1): lane s
// holds section s's state at the window start; on return, at the window end.
template <typename TACC, typename TR, int W, bool FULL>
__device__ __forceinline__ void chan_window(TACC (&v)[W], const double* __restrict__ tf, int S,
                                            int m, int owner, double& ws0, double& ws1) {
  constexpr bool kTabG = std::is_same<TR, float>::value;
  const int lane = threadIdx.x & (kWave - 1);
  TR pw0 = 0, pw1 = 0, pa1 = 0, pa2 = 0;
  const float* pg = table_g(tf);
  for (int s = 0; s < S; ++s) {
    const double* tb = tf + static_cast<int64_t>(s) * kTableStride;
    const Coef<TR> q(tb);
    TR z0 = 0, z1 = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (FULL || j < m) {
        TR u = static_cast<TR>(v[j]);
        if constexpr (FULL && kTabG) {
          u = fma3s(pg[j + 1], pw0, fma3s(pg[j], pw1, u));
        } else if constexpr (FULL) {
          const TR cj = pw0;
          pw0 = pw1 - pa1 * cj;
          pw1 = -pa2 * cj;
          u += cj;
        }
        const TR yv = fma3(q.b0, u, z0);
        z0 = q.b1 * u + (z1 - q.a1 * yv);
        z1 = q.b2 * u + (-q.a2 * yv);
        v[j] = static_cast<TACC>(yv);
      }
    }
    // the window's incoming state enters through lane 0 before the scan
    const double sw0 = readlane_d(ws0, s), sw1 = readlane_d(ws1, s);
    double e0 = static_cast<double>(z0), e1 = static_cast<double>(z1);
    if (lane == 0) {
      const Mat2 M = load_mat(table_pow(tb, 0));
      e0 = fma(M.b, sw1, fma(M.a, sw0, e0));
      e1 = fma(M.d, sw1, fma(M.c, sw0, e1));
    }
    double incl0, incl1, s0, s1;
    wave_scan(tb, lane, e0, e1, incl0, incl1, s0, s1);
    if (lane == 0) {
      s0 = sw0;
      s1 = sw1;
    }
    if constexpr (FULL) {
      if (lane == s) {
        ws0 = incl0;
        ws1 = incl1;
      }
      pw0 = static_cast<TR>(s0);
      pw1 = static_cast<TR>(s1);
      pa1 = q.a1;
      pa2 = q.a2;
      pg = table_g(tb);
    } else {
      TR w0 = static_cast<TR>(s0), w1 = static_cast<TR>(s1);
#pragma unroll
      for (int j = 0; j < W; ++j) {
        if (j < m) {
          const TR cj = w0;
          v[j] = static_cast<TACC>(static_cast<TR>(v[j]) + cj);
          w0 = w1 - q.a1 * cj;
          w1 = -q.a2 * cj;
        }
      }
      const double f0 = readlane_d(static_cast<double>(w0) + static_cast<double>(z0), owner);
      const double f1 = readlane_d(static_cast<double>(w1) + static_cast<double>(z1), owner);
      if (lane == s) {
        ws0 = f0;
        ws1 = f1;
      }
    }
  }
  if constexpr (FULL) apply_pending<TACC, TR, W, kTabG>(v, pg, pw0, pw1, pa1, pa2);
}

template <typename TIO, typename TACC, typename TR, int W>
__global__ __launch_bounds__(64) void sos_chan_kernel(FwdArgs a,
                                                      const double* __restrict__ tables) {
  using L = LdsStage<TIO, W>;
  static_assert(W < kGLen, "the zero-input response table covers 64 samples");
  __shared__ __align__(16) TIO lds[L::kElems];
  const int lane = threadIdx.x;
  const int64_t c = blockIdx.x;
  const int S = a.S;
  const int64_t T = a.T;
  const TIO* xr = static_cast<const TIO*>(a.x) + (c % a.x_rows) * a.x_stride;
  TIO* yr = static_cast<TIO*>(a.y) + c * a.y_stride;
  const double* tf = tables + (c / a.cpf) * static_cast<int64_t>(S) * kTableStride;
  const int64_t win = static_cast<int64_t>(kWave) * W;
  const int64_t nfull = T / win;
  const bool vec = ((reinterpret_cast<uintptr_t>(xr) | reinterpret_cast<uintptr_t>(yr)) & 15u) == 0;

  double ws0 = 0.0, ws1 = 0.0;  // lane s < S: section s's state
  if ((a.flags & TSIIR_HAS_ZI) && lane < S) {
    const int64_t zi_i = (zi_row_of(a, c) * S + lane) * 2;
    ws0 = zi_in<TACC>(a, zi_i);
    ws1 = zi_in<TACC>(a, zi_i + 1);
  }

  u32x4 pf[L::kNV];
  if (vec && nfull > 0) prefetch_window<TIO, W>(xr, pf);
  for (int64_t k = 0; k < nfull; ++k) {
    const int64_t k0 = k * win;
    TACC v[W];
    if (vec) {
      transpose_window_lds<TIO, TACC, W>(pf, lds, v);
      if (k + 1 < nfull) prefetch_window<TIO, W>(xr + k0 + win, pf);
    } else {
      load_run<TIO, TACC, W>(xr + k0 + lane * W, false, W, v);
    }
    chan_window<TACC, TR, W, true>(v, tf, S, W, 0, ws0, ws1);
    if (vec)
      store_window_lds<TIO, TACC, W>(yr + k0, lds, v);
    else
      store_run<TIO, TACC, W>(yr + k0 + lane * W, false, W, v);
  }
  if (nfull * win < T) {
    const int64_t k0 = nfull * win;
    const int64_t t0 = k0 + static_cast<int64_t>(lane) * W;
    const int64_t rem = T - t0;
    const int m = rem <= 0 ? 0 : (rem >= W ? W : static_cast<int>(rem));
    const int owner = static_cast<int>((T - 1 - k0) / W);
    TACC v[W];
    load_run<TIO, TACC, W>(xr + (m > 0 ? t0 : 0), false, m, v);
    chan_window<TACC, TR, W, false>(v, tf, S, m, owner, ws0, ws1);
    store_run<TIO, TACC, W>(yr + (m > 0 ? t0 : 0), false, m, v);
  }
  if ((a.flags & TSIIR_WRITE_ZF) && lane < S) {
    const int64_t zf_i = (c * S + lane) * 2;
    state_out<TACC>(a.zf, a.flags, zf_i, ws0);
    state_out<TACC>(a.zf, a.flags, zf_i + 1, ws1);
  }
}

// Per (filter, section): normalised coefficients and P^(W*2^i), i < kNPow,
// where P is the forward DF2T state matrix A = [[-a1, 1], [-a2, 0]] or, for
// the adjoint recursion, A' = [[-a1, -a2], [1, 0]]; then the float
// zero-input response table g of the forward DF2T (table_g).
template <typename TC>
__global__ __launch_bounds__(256) void sos_prep_kernel(const TC* sos, int64_t n_rows, int log2w,
                                                       int adjoint, double* tables,
                                                       WsHeader* hdr, uint32_t n_epochs,
                                                       const uint32_t* gate, uint32_t gate_on) {
  if (gate != nullptr && ((*gate != 0u) != (gate_on != 0u))) return;
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
  // zero-input response of the forward DF2T, computed in FP64
  float* g = reinterpret_cast<float*>(t + kGOff);
  double gm2 = 0.0, gm1 = 1.0;
  g[0] = 0.0f;
  g[1] = 1.0f;
  for (int j = 2; j < kGLen; ++j) {
    const double gj = -da1 * gm1 - da2 * gm2;
    g[j] = static_cast<float>(gj);
    gm2 = gm1;
    gm1 = gj;
  }
}

void launch_prep(const void* sos, int coef_dtype, int64_t n_rows, int log2w, bool adjoint,
                 double* tables, WsHeader* hdr, uint32_t n_epochs, hipStream_t stream,
                 const uint32_t* gate, uint32_t gate_on) {
  const unsigned blocks = static_cast<unsigned>((n_rows + 255) / 256);
  if (coef_dtype == TSIIR_F64)
    hipLaunchKernelGGL(sos_prep_kernel<double>, dim3(blocks), dim3(256), 0, stream,
                       static_cast<const double*>(sos), n_rows, log2w, adjoint ? 1 : 0, tables,
                       hdr, n_epochs, gate, gate_on);
  else
    hipLaunchKernelGGL(sos_prep_kernel<float>, dim3(blocks), dim3(256), 0, stream,
                       static_cast<const float*>(sos), n_rows, log2w, adjoint ? 1 : 0, tables,
                       hdr, n_epochs, gate, gate_on);
}

// Short rows: one thread per channel filters its whole row, every section
// of a sample before the next sample, with every section's state in
// registers (at most SMAX sections: the loops unroll and test s < S).  The
// update is the window kernels' DF2T (the order of _lfilter.py:158-166 with
// fused multiply-adds) in TR, on the coefficients normalised as the table
// preparation normalises them (by a0, in the coefficient type).  Samples
// are loaded kShortBlk at a time, one block ahead of the block being
// filtered.  No workspace is used.
constexpr int kShortBlk = 8;
constexpr int64_t kShortMaxT = 512;        // rows up to this length
constexpr int64_t kShortMaxWork = 4096;    // and samples x sections up to this
constexpr int kShortMaxS = 8;              // sections held in registers
template <typename TIO, typename TACC, typename TR, typename TC, int SMAX>
__global__ __launch_bounds__(kWave) void sos_short_kernel(FwdArgs a, const TC* __restrict__ sos) {
  if (gated_out(a)) return;
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kWave + threadIdx.x;
  if (c >= a.B) return;
  using io_f = typename std::conditional<std::is_same<TIO, double>::value, double, float>::type;
  const int S = a.S;
  const TC* f = sos + (c / a.cpf) * static_cast<int64_t>(S) * 6;
  const bool has_zi = (a.flags & TSIIR_HAS_ZI) != 0;
  const int64_t zi_row = zi_row_of(a, c);
  TR b0[SMAX], b1[SMAX], b2[SMAX], a1[SMAX], a2[SMAX], z0[SMAX], z1[SMAX];
#pragma unroll
  for (int s = 0; s < SMAX; ++s) {
    b0[s] = b1[s] = b2[s] = a1[s] = a2[s] = z0[s] = z1[s] = TR(0);
    if (s < S) {
      const TC* r = f + 6 * s;
      TC cb0 = r[0], cb1 = r[1], cb2 = r[2], ca0 = r[3], ca1 = r[4], ca2 = r[5];
      if (ca0 != TC(1)) {
        cb0 = cb0 / ca0;
        cb1 = cb1 / ca0;
        cb2 = cb2 / ca0;
        ca1 = ca1 / ca0;
        ca2 = ca2 / ca0;
      }
      b0[s] = static_cast<TR>(static_cast<double>(cb0));
      b1[s] = static_cast<TR>(static_cast<double>(cb1));
      b2[s] = static_cast<TR>(static_cast<double>(cb2));
      a1[s] = static_cast<TR>(static_cast<double>(ca1));
      a2[s] = static_cast<TR>(static_cast<double>(ca2));
      if (has_zi) {
        z0[s] = static_cast<TR>(zi_in<TACC>(a, (zi_row * S + s) * 2));
        z1[s] = static_cast<TR>(zi_in<TACC>(a, (zi_row * S + s) * 2 + 1));
      }
    }
  }
  const TIO* xr =
      static_cast<const TIO*>(a.x) + row_offset(a, c % a.x_rows, a.x_stride, a.x_seg_len);
  TIO* yr = static_cast<TIO*>(a.y) + row_offset(a, c, a.y_stride, a.y_seg_len);
  const bool rev = (a.flags & TSIIR_REVERSE) != 0;  // processing index t is time T-1-t
  const int64_t T = a.T;
  // block t0 of the row (samples at or after T read as 0)
  auto load_block = [&](int64_t t0, TR (&u)[kShortBlk]) {
#pragma unroll
    for (int j = 0; j < kShortBlk; ++j) {
      const int64_t t = t0 + j;
      u[j] = t < T ? static_cast<TR>(IOTraits<TIO>::to_f(xr[rev ? T - 1 - t : t])) : TR(0);
    }
  };
  TR nxt[kShortBlk];
  load_block(0, nxt);
  for (int64_t t0 = 0; t0 < T; t0 += kShortBlk) {
    TR u[kShortBlk];
#pragma unroll
    for (int j = 0; j < kShortBlk; ++j) u[j] = nxt[j];
    // the next block's loads are in flight while this block is filtered
    if (t0 + kShortBlk < T) load_block(t0 + kShortBlk, nxt);
#pragma unroll
    for (int j = 0; j < kShortBlk; ++j) {
      TR v = u[j];
#pragma unroll
      for (int s = 0; s < SMAX; ++s) {
        if (s < S) {
          const TR yv = fma3(b0[s], v, z0[s]);
          z0[s] = b1[s] * v + (z1[s] - a1[s] * yv);
          z1[s] = b2[s] * v + (-a2[s] * yv);
          v = yv;
        }
      }
      u[j] = v;
      if (t0 + j + 1 == T) break;  // the states stay at the last sample
    }
#pragma unroll
    for (int j = 0; j < kShortBlk; ++j) {
      const int64_t t = t0 + j;
      if (t < T) yr[rev ? T - 1 - t : t] = IOTraits<TIO>::from_f(static_cast<io_f>(u[j]));
    }
  }
  if (a.flags & TSIIR_WRITE_ZF) {
#pragma unroll
    for (int s = 0; s < SMAX; ++s) {
      if (s < S) {
        state_out<TACC>(a.zf, a.flags, (c * S + s) * 2, static_cast<double>(z0[s]));
        state_out<TACC>(a.zf, a.flags, (c * S + s) * 2 + 1, static_cast<double>(z1[s]));
      }
    }
  }
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

// Segmented rows: row r = c * nseg + k is segment k (in time) of channel c.
// Its incoming state is channel c's zi (or zeros) for the segment processed
// first (k = 0; with `rev`, along decreasing time, k = nseg - 1), else the
// warm-up end state computed for it: row r - 1's entry (rev: row r's, whose
// warm-up ran over the start of segment r + 1).  FP64 throughout; runs only
// while the gate is set.
template <typename TS>
__global__ void seg_state_in_kernel(const uint32_t* gate, const double* __restrict__ warm,
                                    const TS* __restrict__ zi, int bcast,
                                    double* __restrict__ zin, int64_t rows, int64_t nseg, int S,
                                    int rev) {
  if (*gate == 0u) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t per = static_cast<int64_t>(S) * 2;
  if (i >= rows * per) return;
  const int64_t r = i / per, within = i % per;
  const int64_t k = r % nseg, c = r / nseg;
  double v;
  if (k != (rev ? nseg - 1 : 0))
    v = warm[(rev ? r : r - 1) * per + within];
  else
    v = zi ? static_cast<double>(zi[(bcast ? 0 : c) * per + within]) : 0.0;
  zin[i] = v;
}

// zf of channel c = the final state of its segment processed last (the
// last in time; with `rev` the first), in the caller's state type.
template <typename TS>
__global__ void seg_state_out_kernel(const uint32_t* gate, const double* __restrict__ zall,
                                     TS* __restrict__ zf, int64_t channels, int64_t nseg, int S,
                                     int rev) {
  if (*gate == 0u) return;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t per = static_cast<int64_t>(S) * 2;
  if (i >= channels * per) return;
  const int64_t c = i / per, within = i % per;
  zf[i] = static_cast<TS>(zall[(c * nseg + (rev ? 0 : nseg - 1)) * per + within]);
}

__global__ void seg_gate_init_kernel(WsHeader* hdr) { hdr->seg_gate = 1u; }

// One wave per filter: Phi, the 2S x 2S state-transition matrix of the
// cascade in DF2T form (one zero-input step), squared `squarings` times in
// FP64.  The filter forgets its state within 2^squarings samples when the
// infinity norm of some Phi^(2^q) falls below 2^-64 (every higher power is
// then smaller still); any other outcome (a norm that stays larger, is not
// finite, or a zero a0) is a no.  S <= 16.
//
// `single` (one filter for every channel, one workgroup): the header keeps
// the last filter judged (its section count, coefficient type and values,
// compared bit for bit) with its verdict, so a call with the same filter
// reads the verdict instead of recomputing it, and the kernel sets the gate
// itself.  Otherwise (one workgroup per filter) a no clears the gate, which
// seg_gate_init_kernel set beforehand.
constexpr int kSegCheckThreads = kWave;
template <typename TC>
__global__ __launch_bounds__(kSegCheckThreads) void seg_check_kernel(const TC* __restrict__ sos,
                                                                   int S, int squarings,
                                                                   int coef_code, int single,
                                                                   WsHeader* hdr) {
  __shared__ double P[32 * 32];
  __shared__ double Q[32 * 32];
  const int n = 2 * S;
  const int t = threadIdx.x;
  const TC* sf = sos + static_cast<int64_t>(blockIdx.x) * S * 6;
  if (single) {
    bool same = hdr->seg_cache_valid != 0u && hdr->seg_cache_sections == S &&
                hdr->seg_cache_coef == coef_code;
    for (int i = t; i < 6 * S; i += kSegCheckThreads)
      same = same && __double_as_longlong(hdr->seg_cache_sos[i]) ==
                         __double_as_longlong(static_cast<double>(sf[i]));
    if (__syncthreads_and(same)) {
      if (t == 0) hdr->seg_gate = hdr->seg_cache_verdict;
      return;
    }
  }
  // lane j < n: the cascade's zero-input step from unit state e_j, its
  // state vector kept in row j of Q
  bool ok = true;
  if (t < n) {
    for (int i = 0; i < n; ++i) Q[t * n + i] = (i == t) ? 1.0 : 0.0;
    double u = 0.0;
    for (int s = 0; s < S; ++s) {
      const double a0 = static_cast<double>(sf[6 * s + 3]);
      if (!(a0 != 0.0) || !isfinite(a0)) ok = false;
      const double b0 = static_cast<double>(sf[6 * s]) / a0;
      const double b1 = static_cast<double>(sf[6 * s + 1]) / a0;
      const double b2 = static_cast<double>(sf[6 * s + 2]) / a0;
      const double a1 = static_cast<double>(sf[6 * s + 4]) / a0;
      const double a2 = static_cast<double>(sf[6 * s + 5]) / a0;
      const double yv = b0 * u + Q[t * n + 2 * s];
      const double n0 = b1 * u - a1 * yv + Q[t * n + 2 * s + 1];
      const double n1 = b2 * u - a2 * yv;
      Q[t * n + 2 * s] = n0;
      Q[t * n + 2 * s + 1] = n1;
      u = yv;
    }
  }
  int state = __syncthreads_or(!ok) ? 2 : 0;  // 0 running, 1 small, 2 failed
  for (int e = t; e < n * n; e += kSegCheckThreads) P[e] = Q[(e % n) * n + e / n];
  __syncthreads();
  for (int q = 0; q <= squarings && state == 0; ++q) {
    double row = 0.0;
    if (t < n)
      for (int j = 0; j < n; ++j) row += fabs(P[t * n + j]);
    // the infinity norm: the largest row sum (a wave reduction; a NaN row
    // sum makes the norm test fail through `fin`)
    double m = row;
    int fin = isfinite(row) ? 1 : 0;
    for (int off = 32; off >= 1; off >>= 1) {
      m = fmax(m, __shfl_xor(m, off));
      fin &= __shfl_xor(fin, off);
    }
    if (!fin) state = 2;
    else if (m < 0x1p-64) state = 1;
    if (state != 0 || q == squarings) break;
    for (int e = t; e < n * n; e += kSegCheckThreads) {
      const int i = e / n, j = e % n;
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc = fma(P[i * n + k], P[k * n + j], acc);
      Q[e] = acc;
    }
    __syncthreads();
    for (int e = t; e < n * n; e += kSegCheckThreads) P[e] = Q[e];
    __syncthreads();
  }
  const uint32_t verdict = state == 1 ? 1u : 0u;
  if (single) {
    for (int i = t; i < 6 * S; i += kSegCheckThreads)
      hdr->seg_cache_sos[i] = static_cast<double>(sf[i]);
    if (t == 0) {
      hdr->seg_cache_sections = S;
      hdr->seg_cache_coef = coef_code;
      hdr->seg_cache_verdict = verdict;
      hdr->seg_cache_valid = 1u;
      hdr->seg_gate = verdict;
    }
  } else if (t == 0 && verdict == 0u) {
    __hip_atomic_fetch_and(&hdr->seg_gate, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void set_mailbox_kernel(WsHeader* hdr, uint32_t* mailbox) { hdr->mailbox = mailbox; }

__global__ void fill_u32_kernel(uint32_t* p, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Samples per lane: 64 for FP32 recurrences, 32 when the recurrence (and so
// the register footprint of its temporaries) is FP64.  The compensated form
// (FP32 samples, FP64 updates) takes 64 in its chained single-wave windows
// (many channels: 227 VGPRs, two waves per SIMD, half the scans per sample;
// 7.39 against 8.28 ms on 2048 x 2^20 samples x 8 sections) and 32 in the
// four-wave look-back windows.  The launch takes W from this geometry.
static inline int samples_per_lane(const tsiir_sosfilt_desc* d) {
  if (d->io_dtype == TSIIR_F64) return 32;
  if (d->mode == TSIIR_MODE_COMPENSATED) return fwd_policy(d).chained ? 64 : 32;
  return 64;
}


bool fwd_geometry(const tsiir_sosfilt_desc* d, FwdGeometry* g) {
  if (!d || d->channels < 0 || d->n_samples < 0 || d->n_sections < 1) return false;
  if (d->chans_per_filter < 1 || d->x_rows < 1) return false;
  g->W = samples_per_lane(d);
  const int64_t win = static_cast<int64_t>(kWave) * g->W * fwd_policy(d).nw;
  g->n_windows = (d->n_samples + win - 1) / win;
  g->n_filters = d->channels == 0 ? 0 : (d->channels + d->chans_per_filter - 1) / d->chans_per_filter;
  g->table_bytes = align_up(static_cast<size_t>(g->n_filters) * d->n_sections * kTableStride * 8, 256);
  g->status_bytes = static_cast<size_t>(g->n_windows) * d->channels * d->n_sections *
                    kStatusWords * 8;
  g->fast_off = align_up(sizeof(WsHeader) + g->table_bytes, 256);
  const size_t fast_bytes =
      use_fwd32(d) ? static_cast<size_t>(g->n_filters) * d->n_sections * sizeof(Fast32) : 0;
  g->status_off = align_up(g->fast_off + fast_bytes, 256);
  g->total = g->status_off + g->status_bytes;
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

template <typename TIO, typename TACC, typename TR, int W>
static tsiir_status launch_fwd(const tsiir_sosfilt_desc* d, const FwdGeometry& g, const void* x,
                               void* y, const void* sos, const void* zi, void* zf, char* ws,
                               hipStream_t stream, void* saved, const FwdGate* gate) {
  WsHeader* hdr = reinterpret_cast<WsHeader*>(ws);
  double* tables = reinterpret_cast<double*>(ws + sizeof(WsHeader));
  uint64_t* status = reinterpret_cast<uint64_t*>(ws + g.status_off);
  const int64_t n_rows = g.n_filters * d->n_sections;
  const int log2w = (W == 64) ? 6 : 5;
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
  a.io_dtype = d->io_dtype;
  a.gate = gate ? gate->word : nullptr;
  a.gate_on = gate ? gate->on : 0u;
  a.seg_n = gate ? gate->seg_n : 0;
  a.x_seg_len = gate ? gate->x_seg_len : 0;
  a.y_seg_len = gate ? gate->y_seg_len : 0;
  a.epoch_back = gate ? gate->epoch_back : 0u;
  a.warm_fused = (gate && gate->warm_fused) ? 1 : 0;
  a.zi_f64 = (gate && gate->warm_fused) ? (gate->zi_f64 ? 1 : 0) : -1;
  const FwdPolicy pol = fwd_policy(d);
  if constexpr (std::is_same<TR, float>::value && W == 64) {
    if (use_fwd32(d)) {
      // one fused launch builds both table sets (and advances the epoch)
      Fast32* fast = reinterpret_cast<Fast32*>(ws + g.fast_off);
      const tsiir_status rc = launch_fwd32_full(a, sos, d->coef_dtype, g.n_filters, fast,
                                                gate != nullptr && gate->tables_ready, stream);
      if (rc != TSIIR_OK) return rc;
      const int64_t full32 = d->n_samples / fwd32_window();
      if (full32 * fwd32_window() < d->n_samples)
        hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, TR, W, 1, false, true, false>),
                           dim3(static_cast<unsigned>(d->channels)), dim3(kWave), 0, stream, a,
                           full32, a.tables);
      return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
    }
  }
  launch_prep(sos, d->coef_dtype, n_rows, log2w, false, tables, hdr, 1, stream, a.gate,
              a.gate_on);
  const int64_t win = static_cast<int64_t>(kWave) * W * pol.nw;
  const int64_t full = d->n_samples / win;
  const int64_t groups = full * d->channels;
  const unsigned grid = static_cast<unsigned>(groups);
  const bool lds = fwd_lds_io();
  if (pol.nw == 1) {
    const dim3 block(kWave);
    if (groups > 0) {
      if (lds)
        hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, TR, W, 1, true, false, true>), dim3(grid),
                           block, 0, stream, a, int64_t(0), a.tables);
      else
        hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, TR, W, 1, false, false, true>),
                           dim3(grid), block, 0, stream, a, int64_t(0), a.tables);
    }
    if (full * win < d->n_samples)
      hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, TR, W, 1, false, true, false>),
                         dim3(static_cast<unsigned>(d->channels)), block, 0, stream, a, full,
                         a.tables);
  } else {
    const dim3 block(kWave * 4);
    if (groups > 0) {
      if (lds)
        hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, TR, W, 4, true, false, false>),
                           dim3(grid), block, 0, stream, a, int64_t(0), a.tables);
      else
        hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, TR, W, 4, false, false, false>),
                           dim3(grid), block, 0, stream, a, int64_t(0), a.tables);
    }
    if (full * win < d->n_samples)
      hipLaunchKernelGGL((sos_fwd_kernel<TIO, TACC, TR, W, 4, false, true, false>),
                         dim3(static_cast<unsigned>(d->channels)), block, 0, stream, a, full,
                         a.tables);
  }
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

// Channel-persistent launch: prep (tables for its W) + one wave per channel.
template <typename TIO, typename TACC, typename TR, int W>
static tsiir_status launch_chan(const tsiir_sosfilt_desc* d, const FwdGeometry& g,
                                const void* x, void* y, const void* sos, const void* zi,
                                void* zf, char* ws, hipStream_t stream) {
  double* tables = reinterpret_cast<double*>(ws + sizeof(WsHeader));
  const int64_t n_rows = g.n_filters * d->n_sections;
  launch_prep(sos, d->coef_dtype, n_rows, W == 64 ? 6 : (W == 32 ? 5 : 4), false, tables,
              nullptr, 0, stream);
  FwdArgs a{};
  a.x = x;
  a.y = y;
  a.zi = zi;
  a.zf = zf;
  a.tables = tables;
  a.B = d->channels;
  a.T = d->n_samples;
  a.x_rows = d->x_rows;
  a.x_stride = d->x_row_stride;
  a.y_stride = d->y_row_stride;
  a.cpf = d->chans_per_filter;
  a.S = d->n_sections;
  a.flags = d->flags;
  hipLaunchKernelGGL((sos_chan_kernel<TIO, TACC, TR, W>),
                     dim3(static_cast<unsigned>(d->channels)), dim3(kWave), 0, stream, a,
                     a.tables);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

// The window-parallel form is the default: at steady state it matched or
// beat the channel-persistent form on every measured shape (its 4 waves per
// SIMD hide the scan latency better than the 2 that the prefetch registers
// leave).  TSIIR_PATH_CHANNEL in the flags (or TSIIR_FWD_PATH=chan, for
// comparison measurements) selects the channel-persistent form when no
// section outputs are kept for a backward pass; TSIIR_CHAN_W=32 or 64 picks
// its FP32 samples per lane.
static bool use_chan(const tsiir_sosfilt_desc* d, const void* saved) {
  static const int forced = [] {
    const char* e = getenv("TSIIR_FWD_PATH");
    if (e && strcmp(e, "window") == 0) return 0;
    if (e && strcmp(e, "chan") == 0) return 1;
    return -1;
  }();
  if (saved != nullptr || d->n_sections > kWave) return false;
  if (d->flags & (TSIIR_PATH_WINDOW | TSIIR_REVERSE)) return false;
  if (d->flags & TSIIR_PATH_CHANNEL) return true;
  return forced == 1;
}
static int chan_w_f32() {
  static const int w = [] {
    const char* e = getenv("TSIIR_CHAN_W");
    return (e && strcmp(e, "32") == 0) ? 32 : 64;
  }();
  return w;
}

// ---- segmented form ----------------------------------------------------------
//
// Few channels of long rows wait at every section of every group for the
// look-back.  A filter that forgets its state within TSIIR_SEGMENT_WARMUP
// samples does not need it: each segment of a row can start from the state
// that a warm-up pass from zero state over the last TSIIR_SEGMENT_WARMUP
// samples before it reaches (the neglected part is Phi^4096 times the true
// state, below 2^-64 of it), and then every segment is an independent row of
// the many-channel kernels, chained only within itself.  Whether the filter
// forgets is decided on the device (seg_check_kernel), so both forms are
// launched and the gate lets exactly one of them run.

constexpr int64_t kSegRows = 1024;      // rows from which the forward chains its windows
constexpr int64_t kSegMinLen = 16384;   // a four-window group of the FP32 kernel
constexpr int kSegSquarings = 12;       // 2^12 = TSIIR_SEGMENT_WARMUP
static_assert((int64_t{1} << kSegSquarings) == TSIIR_SEGMENT_WARMUP, "warm-up length");

struct SegPlan {
  int64_t lseg = 0, nseg = 0, rows = 0, rem = 0;
  bool rev = false;
  tsiir_sosfilt_desc warm{}, main{}, tail{};
  size_t base = 0, ywarm_off = 0, warm_off = 0, zin_off = 0, zall_off = 0, ztin_off = 0,
         ztout_off = 0, total = 0;
};

static size_t io_bytes(int32_t dt) { return dt == TSIIR_F64 ? 8 : dt == TSIIR_BF16 ? 2 : 4; }

// Each channel's row splits into nseg segments of lseg samples (the longest
// power of two that still gives the many-channel kernels at least kSegRows
// of them) and a remainder of rem < lseg samples, which runs last (after
// the last segment; with REVERSE, time runs backwards and the remainder is
// the start of the row) from the state the segments end in.  False where
// the form does not apply.
static bool seg_plan(const tsiir_sosfilt_desc* d, bool saved, SegPlan* p) {
  if ((d->flags & TSIIR_NO_SEGMENT) || saved || use_chan(d, nullptr)) return false;
  if (d->flags & (TSIIR_PATH_CHANNEL | TSIIR_DEBUG_GIVE_UP)) return false;
  if (d->n_sections > 16) return false;
  const int64_t B = d->channels, T = d->n_samples;
  if (B < 1 || B >= kSegRows || T < 2 * kSegMinLen || d->x_rows != B) return false;
  const int64_t nmin = (kSegRows + B - 1) / B;
  int64_t lseg = 0;
  for (int64_t l = kSegMinLen; l <= T / 2; l *= 2)
    if (T / l >= nmin) lseg = l;
  if (lseg == 0) return false;
  p->lseg = lseg;
  p->nseg = T / lseg;
  p->rem = T - p->nseg * lseg;
  p->rows = B * p->nseg;
  p->rev = (d->flags & TSIIR_REVERSE) != 0;
  const bool want_zf = (d->flags & TSIIR_WRITE_ZF) != 0;
  const bool shared = d->chans_per_filter >= B;
  // channel strides (any value serves one channel)
  const int64_t xs = B > 1 ? d->x_row_stride : T, ys = B > 1 ? d->y_row_stride : T;
  const uint32_t keep = d->flags & (TSIIR_FP32_DF2T | TSIIR_PATH_WINDOW | TSIIR_REVERSE);
  // the warm-up rows: one per segment, over the TSIIR_SEGMENT_WARMUP samples
  // before it (rev: after it, so there is none after the very last segment)
  const int64_t wrows = p->rev ? p->rows - 1 : p->rows;
  tsiir_sosfilt_desc w = *d;
  w.channels = wrows;
  w.n_samples = TSIIR_SEGMENT_WARMUP;
  w.x_rows = wrows;
  w.x_row_stride = xs;
  w.y_row_stride = TSIIR_SEGMENT_WARMUP;
  w.chans_per_filter = shared ? wrows : d->chans_per_filter * p->nseg;
  w.flags = keep | TSIIR_WRITE_ZF | TSIIR_STATE_F64;
  tsiir_sosfilt_desc m = *d;
  m.channels = p->rows;
  m.n_samples = lseg;
  m.x_rows = p->rows;
  m.x_row_stride = xs;
  m.y_row_stride = ys;
  m.chans_per_filter = shared ? p->rows : d->chans_per_filter * p->nseg;
  m.flags = keep | TSIIR_HAS_ZI | TSIIR_STATE_F64 |
            ((want_zf || p->rem > 0) ? TSIIR_WRITE_ZF : 0u);
  tsiir_sosfilt_desc t = *d;
  t.n_samples = p->rem;
  t.x_row_stride = xs;
  t.y_row_stride = ys;
  t.flags = keep | TSIIR_HAS_ZI | TSIIR_STATE_F64 | (want_zf ? TSIIR_WRITE_ZF : 0u);
  p->warm = w;
  p->main = m;
  p->tail = t;
  FwdGeometry gw, gm, gt;
  if (!fwd_geometry(&w, &gw) || !fwd_geometry(&m, &gm) || !fwd_geometry(&t, &gt)) return false;
  size_t base = gw.total > gm.total ? gw.total : gm.total;
  if (p->rem > 0 && gt.total > base) base = gt.total;
  const size_t states = static_cast<size_t>(p->rows) * d->n_sections * 2 * sizeof(double);
  const size_t cstates = static_cast<size_t>(B) * d->n_sections * 2 * sizeof(double);
  p->base = align_up(base, 256);
  p->ywarm_off = p->base;
  p->warm_off = align_up(p->ywarm_off + static_cast<size_t>(p->rows) * TSIIR_SEGMENT_WARMUP *
                                            io_bytes(d->io_dtype), 256);
  p->zin_off = align_up(p->warm_off + states, 256);
  p->zall_off = align_up(p->zin_off + states, 256);
  p->ztin_off = align_up(p->zall_off + states, 256);
  p->ztout_off = align_up(p->ztin_off + cstates, 256);
  p->total = p->ztout_off + cstates;
  return true;
}

// Workspace bytes of the forward for this descriptor: where the segmented
// form applies, enough for it and for the look-back form (one of them runs,
// one after the other on the stream).
size_t fwd_workspace_bytes(const tsiir_sosfilt_desc* d, bool saved) {
  FwdGeometry g;
  if (!fwd_geometry(d, &g)) return 0;
  SegPlan p;
  if (seg_plan(d, saved, &p)) return p.total > g.total ? p.total : g.total;
  return g.total;
}

// The device check (gate); the warm-up pass over the TSIIR_SEGMENT_WARMUP
// samples before every segment (its end states are the segments' incoming
// states); the incoming states assembled; the main pass over all segments
// as two-level rows; then the remainder of each row from the state its
// segments end in, and zf.  Every launch after the check runs only while
// the gate is set.
// FP32-recurrence forms (the direct-form-I kernels) instead build the
// tables once for every launch of the call, the look-back form's included
// (`fused` returns true), and fuse the warm-up into the main pass: the
// first group of each segment runs it itself (FwdArgs::warm_fused), so the
// warm-up pass, the state assembly and two table preparations drop out.
static tsiir_status forward_segmented(const tsiir_sosfilt_desc* d, const SegPlan& p,
                                      const void* x, void* y, const void* sos, const void* zi,
                                      void* zf, char* ws, hipStream_t st, int64_t n_filters,
                                      bool* fused) {
  const int S = d->n_sections;
  WsHeader* hdr = reinterpret_cast<WsHeader*>(ws);
  const uint32_t* gate = &hdr->seg_gate;
  const int single = n_filters == 1 ? 1 : 0;
  if (!single) hipLaunchKernelGGL(seg_gate_init_kernel, dim3(1), dim3(1), 0, st, hdr);
  if (d->coef_dtype == TSIIR_F64)
    hipLaunchKernelGGL(seg_check_kernel<double>, dim3(static_cast<unsigned>(n_filters)),
                       dim3(kSegCheckThreads), 0, st, static_cast<const double*>(sos), S,
                       kSegSquarings, d->coef_dtype, single, hdr);
  else
    hipLaunchKernelGGL(seg_check_kernel<float>, dim3(static_cast<unsigned>(n_filters)),
                       dim3(kSegCheckThreads), 0, st, static_cast<const float*>(sos), S,
                       kSegSquarings, d->coef_dtype, single, hdr);
  if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
  const int64_t esz = static_cast<int64_t>(io_bytes(d->io_dtype));
  double* warm = reinterpret_cast<double*>(ws + p.warm_off);
  double* zin = reinterpret_cast<double*>(ws + p.zin_off);
  double* zall = reinterpret_cast<double*>(ws + p.zall_off);
  double* ztin = reinterpret_cast<double*>(ws + p.ztin_off);
  double* ztout = reinterpret_cast<double*>(ws + p.ztout_off);
  const char* xc = static_cast<const char*>(x);
  char* yc = static_cast<char*>(y);
  // the segments start at sample 0 (rev: at the remainder's end)
  const int64_t seg0 = p.rev ? p.rem : 0;
  const bool want_zf = (d->flags & TSIIR_WRITE_ZF) != 0;
  const bool sd = d->io_dtype == TSIIR_F64 || (d->flags & TSIIR_STATE_F64);
  const int rev = p.rev ? 1 : 0;
  tsiir_status rc = TSIIR_OK;
  *fused = use_fwd32(d);
  if (*fused) {
    FwdGeometry gm;
    if (!fwd_geometry(&p.main, &gm) || gm.n_filters != n_filters) return TSIIR_EINVAL;
    launch_prep32(sos, d->coef_dtype, n_filters, S, reinterpret_cast<double*>(ws + sizeof(WsHeader)),
                  reinterpret_cast<Fast32*>(ws + gm.fast_off), hdr, 1u, st);
    tsiir_sosfilt_desc m = p.main;
    m.flags = (m.flags & ~TSIIR_HAS_ZI) | (d->flags & (TSIIR_HAS_ZI | TSIIR_ZI_BROADCAST));
    FwdGate on_main{gate, 1u, p.nseg, p.lseg, p.lseg};
    on_main.tables_ready = true;
    on_main.warm_fused = true;
    on_main.zi_f64 = sd;
    rc = forward_impl(&m, xc + seg0 * esz, yc + seg0 * esz, sos, zi,
                      (want_zf || p.rem > 0) ? zall : nullptr, ws, p.base, st, nullptr, &on_main);
    if (rc != TSIIR_OK) return rc;
  } else {
    // warm-up rows start TSIIR_SEGMENT_WARMUP samples before each segment's
    // end (rev: at the start of the segment after it)
    const int64_t wofs = seg0 + (p.rev ? p.lseg : p.lseg - TSIIR_SEGMENT_WARMUP);
    const FwdGate on_warm{gate, 1u, p.nseg, p.lseg, 0};
    rc = forward_impl(&p.warm, xc + wofs * esz, ws + p.ywarm_off, sos, nullptr, warm, ws, p.base,
                      st, nullptr, &on_warm);
    if (rc != TSIIR_OK) return rc;
    const bool has_zi = (d->flags & TSIIR_HAS_ZI) != 0;
    const int bc = (d->flags & TSIIR_ZI_BROADCAST) ? 1 : 0;
    const int64_t n_in = p.rows * S * 2;
    const unsigned bi = static_cast<unsigned>((n_in + 255) / 256);
    if (sd)
      hipLaunchKernelGGL(seg_state_in_kernel<double>, dim3(bi), dim3(256), 0, st, gate, warm,
                         has_zi ? static_cast<const double*>(zi) : nullptr, bc, zin, p.rows,
                         p.nseg, S, rev);
    else
      hipLaunchKernelGGL(seg_state_in_kernel<float>, dim3(bi), dim3(256), 0, st, gate, warm,
                         has_zi ? static_cast<const float*>(zi) : nullptr, bc, zin, p.rows,
                         p.nseg, S, rev);
    if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
    const FwdGate on_main{gate, 1u, p.nseg, p.lseg, p.lseg};
    rc = forward_impl(&p.main, xc + seg0 * esz, yc + seg0 * esz, sos, zin,
                      (want_zf || p.rem > 0) ? zall : nullptr, ws, p.base, st, nullptr, &on_main);
    if (rc != TSIIR_OK) return rc;
  }
  const int64_t n_out = d->channels * S * 2;
  const unsigned bo = static_cast<unsigned>((n_out + 255) / 256);
  const double* zlast = zall;  // the state the segments end in, per channel
  int64_t zn = p.nseg;
  int zrev = rev;
  if (p.rem > 0) {
    // the remainder: after the segments (rev: the start of the row, last)
    hipLaunchKernelGGL(seg_state_out_kernel<double>, dim3(bo), dim3(256), 0, st, gate, zall,
                       ztin, d->channels, p.nseg, S, rev);
    if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
    const int64_t t0 = p.rev ? 0 : p.nseg * p.lseg;
    const FwdGate on_tail{gate, 1u, 0, 0, 0};
    rc = forward_impl(&p.tail, xc + t0 * esz, yc + t0 * esz, sos, ztin, want_zf ? ztout : nullptr,
                      ws, p.base, st, nullptr, &on_tail);
    if (rc != TSIIR_OK) return rc;
    zlast = ztout;
    zn = 1;
    zrev = 0;
  }
  if (!want_zf) return TSIIR_OK;
  if (sd)
    hipLaunchKernelGGL(seg_state_out_kernel<double>, dim3(bo), dim3(256), 0, st, gate, zlast,
                       static_cast<double*>(zf), d->channels, zn, S, zrev);
  else
    hipLaunchKernelGGL(seg_state_out_kernel<float>, dim3(bo), dim3(256), 0, st, gate, zlast,
                       static_cast<float*>(zf), d->channels, zn, S, zrev);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

// Short rows (and no section outputs kept): the thread-per-channel kernel.
// A row of 39 samples took 75-85 µs in the one-wave-per-channel window form
// over 4096 channels (one lane of each wave busy); the forms chosen by
// TSIIR_PATH_WINDOW / TSIIR_PATH_CHANNEL and the look-back tests keep theirs.
static bool use_short(const tsiir_sosfilt_desc* d, const void* saved) {
  if (saved != nullptr || d->n_sections > kShortMaxS) return false;
  if (d->flags & (TSIIR_PATH_WINDOW | TSIIR_PATH_CHANNEL | TSIIR_DEBUG_GIVE_UP)) return false;
  return d->n_samples <= kShortMaxT && d->n_samples * d->n_sections <= kShortMaxWork;
}

template <typename TIO, typename TACC, typename TR>
static tsiir_status launch_short(const tsiir_sosfilt_desc* d, const void* x, void* y,
                                 const void* sos, const void* zi, void* zf, hipStream_t stream,
                                 const FwdGate* gate) {
  FwdArgs a;
  a.x = x;
  a.y = y;
  a.zi = zi;
  a.zf = zf;
  a.B = d->channels;
  a.T = d->n_samples;
  a.x_rows = d->x_rows;
  a.x_stride = d->x_row_stride;
  a.y_stride = d->y_row_stride;
  a.cpf = d->chans_per_filter;
  a.S = d->n_sections;
  a.flags = d->flags;
  a.io_dtype = d->io_dtype;
  a.gate = gate ? gate->word : nullptr;
  a.gate_on = gate ? gate->on : 0u;
  a.seg_n = gate ? gate->seg_n : 0;
  a.x_seg_len = gate ? gate->x_seg_len : 0;
  a.y_seg_len = gate ? gate->y_seg_len : 0;
  a.warm_fused = (gate && gate->warm_fused) ? 1 : 0;
  a.zi_f64 = (gate && gate->warm_fused) ? (gate->zi_f64 ? 1 : 0) : -1;
  const dim3 grid(static_cast<unsigned>((d->channels + kWave - 1) / kWave)), block(kWave);
  if (d->coef_dtype == TSIIR_F64)
    hipLaunchKernelGGL((sos_short_kernel<TIO, TACC, TR, double, kShortMaxS>), grid, block, 0,
                       stream, a, static_cast<const double*>(sos));
  else
    hipLaunchKernelGGL((sos_short_kernel<TIO, TACC, TR, float, kShortMaxS>), grid, block, 0,
                       stream, a, static_cast<const float*>(sos));
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

tsiir_status forward_impl(const tsiir_sosfilt_desc* d, const void* x, void* y, const void* sos,
                          const void* zi, void* zf, void* workspace, size_t workspace_bytes,
                          void* stream, void* saved, const FwdGate* gate) {
  FwdGeometry g;
  if (!fwd_geometry(d, &g)) return TSIIR_EINVAL;
  if (d->io_dtype != TSIIR_F32 && d->io_dtype != TSIIR_F64 && d->io_dtype != TSIIR_BF16)
    return TSIIR_EINVAL;
  if (d->coef_dtype != TSIIR_F32 && d->coef_dtype != TSIIR_F64) return TSIIR_EINVAL;
  if (d->mode != TSIIR_MODE_PLAIN && d->mode != TSIIR_MODE_COMPENSATED) return TSIIR_EINVAL;
  if ((d->flags & TSIIR_REVERSE) && (saved != nullptr || (d->flags & TSIIR_PATH_CHANNEL)))
    return TSIIR_EUNSUPPORTED;
  if ((d->flags & TSIIR_PATH_WINDOW) && (d->flags & TSIIR_PATH_CHANNEL)) return TSIIR_EINVAL;
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
    if (f64 || (d->flags & TSIIR_STATE_F64))
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
  FwdGate off{nullptr, 0u};
  SegPlan sp;
  if (gate == nullptr && seg_plan(d, saved != nullptr, &sp)) {
    if (workspace_bytes < sp.total) return TSIIR_EWORKSPACE;
    bool fused = false;
    const tsiir_status rc =
        forward_segmented(d, sp, x, y, sos, zi, zf, ws, st, g.n_filters, &fused);
    if (rc != TSIIR_OK) return rc;
    // the look-back form runs only where the check cleared the gate (on the
    // tables the segmented form built, for the FP32-recurrence forms)
    off = FwdGate{&reinterpret_cast<WsHeader*>(ws)->seg_gate, 0u};
    off.tables_ready = fused;
    gate = &off;
  }
  const bool comp = d->mode == TSIIR_MODE_COMPENSATED;
  if (use_short(d, saved)) {
    switch (d->io_dtype) {
      case TSIIR_F32:
        return comp ? launch_short<float, float, double>(d, x, y, sos, zi, zf, st, gate)
                    : launch_short<float, float, float>(d, x, y, sos, zi, zf, st, gate);
      case TSIIR_BF16:
        return comp ? launch_short<bf16_t, float, double>(d, x, y, sos, zi, zf, st, gate)
                    : launch_short<bf16_t, float, float>(d, x, y, sos, zi, zf, st, gate);
      default:
        return launch_short<double, double, double>(d, x, y, sos, zi, zf, st, gate);
    }
  }
  if (use_chan(d, saved)) {
    const bool w64 = chan_w_f32() == 64;
    switch (d->io_dtype) {
      case TSIIR_F32:
        if (comp) return launch_chan<float, float, double, 32>(d, g, x, y, sos, zi, zf, ws, st);
        return w64 ? launch_chan<float, float, float, 64>(d, g, x, y, sos, zi, zf, ws, st)
                   : launch_chan<float, float, float, 32>(d, g, x, y, sos, zi, zf, ws, st);
      case TSIIR_BF16:
        if (comp) return launch_chan<bf16_t, float, double, 32>(d, g, x, y, sos, zi, zf, ws, st);
        return w64 ? launch_chan<bf16_t, float, float, 64>(d, g, x, y, sos, zi, zf, ws, st)
                   : launch_chan<bf16_t, float, float, 32>(d, g, x, y, sos, zi, zf, ws, st);
      default:
        return launch_chan<double, double, double, 32>(d, g, x, y, sos, zi, zf, ws, st);
    }
  }
  switch (d->io_dtype) {
    case TSIIR_F32:
      if (!comp) return launch_fwd<float, float, float, 64>(d, g, x, y, sos, zi, zf, ws, st, saved, gate);
      return g.W == 64
                 ? launch_fwd<float, float, double, 64>(d, g, x, y, sos, zi, zf, ws, st, saved, gate)
                 : launch_fwd<float, float, double, 32>(d, g, x, y, sos, zi, zf, ws, st, saved, gate);
    case TSIIR_BF16:
      if (!comp)
        return launch_fwd<bf16_t, float, float, 64>(d, g, x, y, sos, zi, zf, ws, st, saved, gate);
      return g.W == 64
                 ? launch_fwd<bf16_t, float, double, 64>(d, g, x, y, sos, zi, zf, ws, st, saved, gate)
                 : launch_fwd<bf16_t, float, double, 32>(d, g, x, y, sos, zi, zf, ws, st, saved, gate);
    default:
      return launch_fwd<double, double, double, 32>(d, g, x, y, sos, zi, zf, ws, st, saved, gate);
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
    case TSIIR_EAGAIN: return "serving queue full: poll and submit again";
    case TSIIR_ESCRATCH:
      return "request needs more workspace than the fixed scratch of the no-allocation entry point";
  }
  return "unknown status";
}

size_t tsiir_sosfilt_workspace_bytes(const tsiir_sosfilt_desc* d) {
  return fwd_workspace_bytes(d, false);
}

tsiir_status tsiir_workspace_init(void* workspace, size_t bytes, void* stream) {
  if (!workspace || bytes < sizeof(WsHeader) || (bytes % 4) != 0) return TSIIR_EINVAL;
  const size_t n = bytes / 4;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<uint32_t*>(workspace), n);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

tsiir_status tsiir_workspace_set_error_mailbox(void* workspace, size_t workspace_bytes,
                                               uint32_t* mailbox, void* stream) {
  if (!workspace || workspace_bytes < sizeof(WsHeader)) return TSIIR_EINVAL;
  uint32_t* dev = mailbox;
  if (mailbox != nullptr) {
    // pinned host memory: its device-side address; device memory: as given
    void* dp = nullptr;
    if (hipHostGetDevicePointer(&dp, mailbox, 0) == hipSuccess && dp != nullptr)
      dev = static_cast<uint32_t*>(dp);
    else
      (void)hipGetLastError();
  }
  hipLaunchKernelGGL(set_mailbox_kernel, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream),
                     static_cast<WsHeader*>(workspace), dev);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

tsiir_status tsiir_sosfilt_forward(const tsiir_sosfilt_desc* d, const void* x, void* y,
                                   const void* sos, const void* zi, void* zf, void* workspace,
                                   size_t workspace_bytes, void* stream) {
  return forward_impl(d, x, y, sos, zi, zf, workspace, workspace_bytes, stream, nullptr);
}

tsiir_status tsiir_sosfilt_forward_saved(const tsiir_sosfilt_desc* desc, const void* x, void* y,
                                         const void* sos, const void* zi, void* zf, void* saved,
                                         void* workspace, size_t workspace_bytes, void* stream) {
  if (!desc) return TSIIR_EINVAL;
  if (desc->n_sections > 1 && !saved) return TSIIR_EINVAL;
  if (desc->flags & TSIIR_PATH_CHANNEL) return TSIIR_EINVAL;
  return forward_impl(desc, x, y, sos, zi, zf, workspace, workspace_bytes, stream,
                      desc->n_sections > 1 ? saved : nullptr);
}

}  // extern "C"
