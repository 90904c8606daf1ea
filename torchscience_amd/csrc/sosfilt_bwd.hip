// sosfilt_bwd.hip - hand-written adjoint of the SOS cascade on gfx950.
//
// Replaces the autograd tape the reference records through the per-sample
// DF2T loop (_lfilter.py:153-166 inside _sosfilt.py:117-128).
//
// For one section with input u, output v (v[n] = b0 u[n] + z0[n-1], ...),
// output gradient g and final-state gradient dzf, the adjoint variable
//   gamma[n] = g[n] - a1 gamma[n+1] - a2 gamma[n+2],
//   (gamma[T], gamma[T+1]) = (dzf0, dzf1)
// is the time-reversed all-pole filter of g.  From it:
//   du[n]  = b0 gamma[n] + b1 gamma[n+1] + b2 gamma[n+2]
//   db_k   = sum_n gamma[n+k] u[n]            (k = 0, 1, 2)
//   da_k   = -sum_n gamma[n+k] v[n]           (k = 1, 2)
//   dzi    = (gamma[0], gamma[1]).
// The kernel runs the reverse recursion with the same single-pass machinery
// as the forward (lane sub-chunks, wave scan, decoupled look-back, here over
// windows taken from the end of the signal), with the state
// (gamma[n+1], gamma[n+2]) and matrix A' = [[-a1, -a2], [1, 0]], so every
// boundary value the FIR part and dzi need is an exact state.  Coefficient
// sums are accumulated per lane in the compute type, reduced per wave in
// FP64, and summed per (filter, section) by a final reduction kernel, which
// also maps them back through the a0 normalisation (zero a0 gradient when
// a0 == 1, as the reference's untaken branch at _lfilter.py:96 gives).
//
// Sections are processed last to first; the section inputs and outputs come
// from one forward recompute that keeps every section's output.

#include <stdlib.h>
#include <string.h>

#include <type_traits>

#include "sos_common.h"
#include "sos_internal.h"
#include "tsiir.h"

namespace tsiir {

struct AdjArgs {
  const void* g;
  void* du;  // may alias g
  const void* u;
  const void* v;
  const void* dzf;
  void* dzi;
  double* partial;  // (n_windows, B, 5) for this section
  const double* tables;
  uint64_t* status;
  WsHeader* hdr;
  int64_t B, T, cpf, n_windows;
  int S, s;
};

template <typename T, int W>
__device__ __forceinline__ void load_rev(const T* row, int64_t n_hi, bool vec, int m, T (&r)[W]) {
  if (vec) {
    const T* p = row + (n_hi - (W - 1));
    T tmp[W];
    using Tr = IOTraits<T>;
    const typename Tr::vec_t* q = reinterpret_cast<const typename Tr::vec_t*>(p);
#pragma unroll
    for (int i = 0; i < W / Tr::kVec; ++i) {
      const typename Tr::vec_t rv = q[i];
      const T* e = reinterpret_cast<const T*>(&rv);
#pragma unroll
      for (int j = 0; j < Tr::kVec; ++j) tmp[i * Tr::kVec + j] = e[j];
    }
#pragma unroll
    for (int j = 0; j < W; ++j) r[j] = tmp[W - 1 - j];
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) r[j] = (j < m) ? row[n_hi - j] : T(0);
  }
}

template <typename T, int W>
__device__ __forceinline__ void store_rev(T* row, int64_t n_hi, bool vec, int m, const T (&r)[W]) {
  if (vec) {
    T* p = row + (n_hi - (W - 1));
    using Tr = IOTraits<T>;
    typename Tr::vec_t* q = reinterpret_cast<typename Tr::vec_t*>(p);
#pragma unroll
    for (int i = 0; i < W / Tr::kVec; ++i) {
      typename Tr::vec_t rv;
      T* e = reinterpret_cast<T*>(&rv);
#pragma unroll
      for (int j = 0; j < Tr::kVec; ++j) e[j] = r[W - 1 - (i * Tr::kVec + j)];
      q[i] = rv;
    }
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (j < m) row[n_hi - j] = r[j];
  }
}

// One window of one channel per wave.  TAIL = false: the full windows of a
// launch (window-major workgroup order, decoupled look-back, publishes
// status).  TAIL = true: the partial last window (earliest samples) of every
// channel, one wave per channel, launched after the full windows; it reads
// the inclusive state of window k-1 and masks samples before time 0.
template <typename T, int W, bool TAIL>
__global__ __launch_bounds__(64) void sos_adjoint_kernel(AdjArgs a, int64_t k_tail,
                                                         const double* __restrict__ tables) {
  const int lane = threadIdx.x;
  // this launch's epoch is epoch - s (the adjoint prep reserved one per section)
  const uint32_t flag = launch_flag(a.hdr, static_cast<uint32_t>(a.s));
  const int64_t k = TAIL ? k_tail : static_cast<int64_t>(blockIdx.x) / a.B;  // from the end
  const int64_t c = TAIL ? static_cast<int64_t>(blockIdx.x)
                         : static_cast<int64_t>(blockIdx.x) % a.B;
  const int64_t win = static_cast<int64_t>(kWave) * W;
  const int64_t r0 = k * win + static_cast<int64_t>(lane) * W;
  const int64_t rem = a.T - r0;
  const int m = TAIL ? (rem <= 0 ? 0 : (rem >= W ? W : static_cast<int>(rem))) : W;
  const int64_t n_hi = a.T - 1 - r0;  // time of the lane's first processed sample

  const T* grow = static_cast<const T*>(a.g) + c * a.T;
  T* durow = static_cast<T*>(a.du) + c * a.T;
  const T* urow = static_cast<const T*>(a.u) + c * a.T;
  const T* vrow = static_cast<const T*>(a.v) + c * a.T;
  constexpr int kVec = IOTraits<T>::kVec;
  const bool vec = !TAIL && (a.T % kVec == 0) &&
                   ((reinterpret_cast<uintptr_t>(grow) & 15u) == 0) &&
                   ((reinterpret_cast<uintptr_t>(durow) & 15u) == 0) &&
                   ((reinterpret_cast<uintptr_t>(urow) & 15u) == 0) &&
                   ((reinterpret_cast<uintptr_t>(vrow) & 15u) == 0);

  const double* tb = tables + ((c / a.cpf) * a.S + a.s) * static_cast<int64_t>(kTableStride);
  const T b0 = static_cast<T>(tb[0]);
  const T b1 = static_cast<T>(tb[1]);
  const T b2 = static_cast<T>(tb[2]);
  const T a1 = static_cast<T>(tb[3]);
  const T a2 = static_cast<T>(tb[4]);

  T gz[W];
  load_rev<T, W>(grow, n_hi, vec, m, gz);

  // zero-state reverse recursion (state (gamma[n+1], gamma[n+2]))
  T p1 = 0, p2 = 0;
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (!TAIL || j < m) {
      const T gam = gz[j] - a1 * p1 - a2 * p2;
      p2 = p1;
      p1 = gam;
      gz[j] = gam;
    }
  }

  double agg0, agg1, x0, x1;
  wave_scan(tb, lane, static_cast<double>(p1), static_cast<double>(p2), agg0, agg1, x0, x1);
  // dzf is the inclusive state of the virtual window -1 (after the end)
  double in0 = 0.0, in1 = 0.0;
  if (a.dzf != nullptr) {
    const T* dp = static_cast<const T*>(a.dzf) + (c * a.S + a.s) * 2;
    in0 = static_cast<double>(dp[0]);
    in1 = static_cast<double>(dp[1]);
  }
  double sw0 = in0, sw1 = in1;
  if (!TAIL) {
    uint64_t* st = a.status + (k * a.B + c) * kStatusWords;
    publish_pair(st, agg0, agg1, flag, lane);
    if (k > 0) lookback(a.status, a.B, 1, a.hdr, k, c, 0, flag, tb, 6, in0, in1, sw0, sw1);
    const Mat2 Mw = load_mat(table_pow(tb, 6));
    publish_pair(st + 4, Mw.a * sw0 + Mw.b * sw1 + agg0, Mw.c * sw0 + Mw.d * sw1 + agg1, flag,
                 lane);
  } else if (k > 0) {
    const uint64_t* st = a.status + ((k - 1) * a.B + c) * kStatusWords;
    if (!read_pair(st + 4, flag, sw0, sw1)) give_up(a.hdr, 2u, sw0, sw1);
  }
  double s0, s1;
  lane_incoming(tb, lane, sw0, sw1, x0, x1, s0, s1);

  // exact gamma: add the zero-input response from (gamma[n+1], gamma[n+2])
  {
    T q1 = static_cast<T>(s0), q2 = static_cast<T>(s1);
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (!TAIL || j < m) {
        const T h = -a1 * q1 - a2 * q2;
        q2 = q1;
        q1 = h;
        gz[j] += h;
      }
    }
  }

  T ur[W], vr[W];
  load_rev<T, W>(urow, n_hi, vec, m, ur);
  load_rev<T, W>(vrow, n_hi, vec, m, vr);
  T acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0, acc4 = 0;
  p1 = static_cast<T>(s0);
  p2 = static_cast<T>(s1);
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (!TAIL || j < m) {
      const T gam = gz[j];
      const T uu = ur[j];
      const T vv = vr[j];
      gz[j] = b0 * gam + b1 * p1 + b2 * p2;
      acc0 += gam * uu;
      acc1 += p1 * uu;
      acc2 += p2 * uu;
      acc3 += p1 * vv;
      acc4 += p2 * vv;
      p2 = p1;
      p1 = gam;
    }
  }
  if (a.dzi != nullptr && m > 0 && n_hi - (m - 1) == 0) {
    T* dp = static_cast<T*>(a.dzi) + (c * a.S + a.s) * 2;
    dp[0] = p1;
    dp[1] = p2;
  }
  store_rev<T, W>(durow, n_hi, vec, m, gz);

  double r[5] = {static_cast<double>(acc0), static_cast<double>(acc1), static_cast<double>(acc2),
                 static_cast<double>(acc3), static_cast<double>(acc4)};
#pragma unroll
  for (int i = 0; i < 5; ++i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r[i] += __shfl_xor(r[i], off);
  }
  if (lane < 5) {
    double val = r[0];
#pragma unroll
    for (int i = 1; i < 5; ++i)
      if (lane == i) val = r[i];
    a.partial[(k * a.B + c) * 5 + lane] = val;
  }
}

// ---- fused adjoint: every section in one pass --------------------------
//
// The same recursion as sos_adjoint_kernel, for all sections of the cascade
// in registers, mirroring the chained forward: one wave per window (windows
// taken from the end of the signal, window-major workgroup order), lane q
// owning reverse indices [q*W, (q+1)*W) of its window, sections last to
// first.  The correction that makes section p's gamma exact (its zero-input
// response from the lane's incoming state) is applied in the sweep that runs
// section p-1's zero-state recursion; that sweep also forms du_p (the output
// gradient of section p-1), accumulates section p's coefficient sums and,
// for the lane holding time 0, writes dzi_p.  Each window's incoming state of
// section s is the inclusive state window k-1 published for s (read for all
// sections up front), folded into lane 0 before the wave scan.  Inputs: dy,
// the forward output y, the section outputs saved by the forward
// ((S-1, B, T)) and x; output dx.  HBM traffic: S+2 reads and one write of
// (B, T) instead of four per section.

struct AdjFArgs {
  const void* dy;
  const void* y;
  const void* x;
  const void* saved;  // (S-1, B, T): output of section s at [s]
  const void* dzf;
  void* dx;
  void* dzi;
  double* partial;    // (S, n_windows, B, 5)
  uint64_t* status;   // (n_windows, B, S) x kStatusWords
  WsHeader* hdr;
  int64_t B, T, cpf, n_windows;
  int S;
};

// Row of the output of section s (s = -1: x, the input of section 0).
template <typename T>
__device__ __forceinline__ const T* sig_row(const AdjFArgs& a, int s, int64_t c) {
  if (s < 0) return static_cast<const T*>(a.x) + c * a.T;
  if (s == a.S - 1) return static_cast<const T*>(a.y) + c * a.T;
  return static_cast<const T*>(a.saved) + (static_cast<int64_t>(s) * a.B + c) * a.T;
}

template <typename T, int W, bool TAIL>
__global__ __launch_bounds__(64) void sos_adjoint_fused_kernel(AdjFArgs a, int64_t k_tail,
                                                               const double* __restrict__ tables) {
  constexpr bool kTabG = std::is_same<T, float>::value;
  static_assert(W + 1 < kGLen, "zero-input table covers the lane");
  __shared__ __align__(16) T lds[TAIL ? 8 : LdsStage<T, W>::kElems];
  const int lane = threadIdx.x;
  const uint32_t flag = launch_flag(a.hdr, 0);
  const int64_t k = TAIL ? k_tail : static_cast<int64_t>(blockIdx.x) / a.B;  // from the end
  const int64_t c = TAIL ? static_cast<int64_t>(blockIdx.x)
                         : static_cast<int64_t>(blockIdx.x) % a.B;
  const int S = a.S;
  const int64_t win = static_cast<int64_t>(kWave) * W;
  const int64_t r0 = k * win + static_cast<int64_t>(lane) * W;
  int m = W;
  if (TAIL) {
    const int64_t rem = a.T - r0;
    m = rem <= 0 ? 0 : (rem >= W ? W : static_cast<int>(rem));
  }
  const int64_t n_hi = a.T - 1 - r0;      // time of the lane's first sample
  const int64_t t_lo = a.T - (k + 1) * win;  // first time of a full window
  constexpr int kVec = IOTraits<T>::kVec;
  const bool vec =
      !TAIL && (a.T % kVec == 0) &&
      (((reinterpret_cast<uintptr_t>(a.dy) | reinterpret_cast<uintptr_t>(a.y) |
         reinterpret_cast<uintptr_t>(a.x) | reinterpret_cast<uintptr_t>(a.saved) |
         reinterpret_cast<uintptr_t>(a.dx)) & 15u) == 0);
  const bool owns_t0 = a.dzi != nullptr && m > 0 && n_hi - (m - 1) == 0;
  const double* tf = tables + (c / a.cpf) * static_cast<int64_t>(S) * kTableStride;

  // window k-1's inclusive states, lane l < S fetching section l's
  double pv0 = 0.0, pv1 = 0.0;
  bool pok = true;
  if (k > 0 && lane < S) {
    const uint64_t* st = a.status + (((k - 1) * a.B + c) * S + lane) * kStatusWords;
    pok = read_pair(st + 4, flag, pv0, pv1);
  }

  // vector path: y (the output of the last section) is loaded with dy, and
  // the input window of the next pending section is loaded while the current
  // section sweeps, so a window waits for HBM once instead of once per section
  T G[W], Up[W], Vp[W];
  u32x4 pf[LdsStage<T, W>::kNV];
  if (vec) {
    if (S >= 2) prefetch_window<T, W>(sig_row<T>(a, S - 1, c) + t_lo, pf);
    load_window_lds<T, T, W, true>(static_cast<const T*>(a.dy) + c * a.T + t_lo, lds, G);
    if (S >= 2) {
      transpose_window_lds<T, T, W, true>(pf, lds, Vp);
      prefetch_window<T, W>(sig_row<T>(a, S - 2, c) + t_lo, pf);
    }
  } else {
    load_rev<T, W>(static_cast<const T*>(a.dy) + c * a.T, n_hi, false, m, G);
  }

  // pending section p = s + 1: lane incoming state, coefficients, table
  T q1 = 0, q2 = 0;
  T pb0 = 0, pb1 = 0, pb2 = 0, pa1 = 0, pa2 = 0;
  const float* pg = table_g(tf);
  for (int s = S - 1; s >= -1; --s) {
    const bool pending = s < S - 1;
    if (pending) {
      // u_p = output of section s, v_p = output of section p = s + 1
      if (s == S - 2) {
        if (!vec) load_rev<T, W>(sig_row<T>(a, S - 1, c), n_hi, false, m, Vp);
      } else {
#pragma unroll
        for (int j = 0; j < W; ++j) Vp[j] = Up[j];
      }
      if (vec) {
        transpose_window_lds<T, T, W, true>(pf, lds, Up);
        if (s >= 0) prefetch_window<T, W>(sig_row<T>(a, s - 1, c) + t_lo, pf);
      } else {
        load_rev<T, W>(sig_row<T>(a, s, c), n_hi, false, m, Up);
      }
    }
    const double* tb = tf + static_cast<int64_t>(s < 0 ? 0 : s) * kTableStride;
    const T a1 = static_cast<T>(tb[3]), a2 = static_cast<T>(tb[4]);
    const T q2n = -pa2 * q2;
    T h1 = q1, h2 = q2;     // homogeneous recursion (FP64 form)
    T pp1 = q1, pp2 = q2;   // exact gamma_p at n+1, n+2
    T p1 = 0, p2 = 0;       // zero-state recursion of section s
    T acc0 = 0, acc1 = 0, acc3 = 0, acc2 = 0, acc4 = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (!TAIL || j < m) {
        T gcur = G[j];
        if (pending) {
          T gam;
          if constexpr (kTabG) {
            gam = fma3s(pg[j + 2], q1, fma3s(pg[j + 1], q2n, gcur));
          } else {
            const T h = -pa1 * h1 - pa2 * h2;
            h2 = h1;
            h1 = h;
            gam = gcur + h;
          }
          const T uu = Up[j], vv = Vp[j];
          acc0 += gam * uu;
          acc1 += pp1 * uu;
          acc2 += pp2 * uu;
          acc3 += pp1 * vv;
          acc4 += pp2 * vv;
          gcur = pb0 * gam + pb1 * pp1 + pb2 * pp2;
          pp2 = pp1;
          pp1 = gam;
        }
        if (s >= 0) {
          const T gz = gcur - a1 * p1 - a2 * p2;
          p2 = p1;
          p1 = gz;
          gcur = gz;
        }
        G[j] = gcur;
      }
    }
    if (pending) {
      const int p = s + 1;
      if (owns_t0) {
        T* dp = static_cast<T*>(a.dzi) + (c * S + p) * 2;
        dp[0] = pp1;
        dp[1] = pp2;
      }
      double r[5] = {static_cast<double>(acc0), static_cast<double>(acc1),
                     static_cast<double>(acc2), static_cast<double>(acc3),
                     static_cast<double>(acc4)};
#pragma unroll
      for (int i = 0; i < 5; ++i) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) r[i] += bpermute_d((lane ^ off) << 2, r[i]);
      }
      if (lane < 5) {
        double val = r[0];
#pragma unroll
        for (int i = 1; i < 5; ++i)
          if (lane == i) val = r[i];
        a.partial[((static_cast<int64_t>(p) * a.n_windows + k) * a.B + c) * 5 + lane] = val;
      }
    }
    if (s < 0) break;

    // window incoming state of section s: dzf (window 0) or window k-1's
    double sw0 = 0.0, sw1 = 0.0;
    if (k == 0) {
      if (a.dzf != nullptr) {
        const T* dp = static_cast<const T*>(a.dzf) + (c * S + s) * 2;
        sw0 = static_cast<double>(dp[0]);
        sw1 = static_cast<double>(dp[1]);
      }
    } else {
      sw0 = readlane_d(pv0, s);
      sw1 = readlane_d(pv1, s);
      if (!__builtin_amdgcn_readlane(static_cast<int>(pok), s)) {
        const uint64_t* st = a.status + (((k - 1) * a.B + c) * S + s) * kStatusWords;
        wait_inclusive(st + 4, a.hdr, flag, 0u, sw0, sw1);
      }
    }
    double e0 = static_cast<double>(p1), e1 = static_cast<double>(p2);
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
    if (!TAIL) {
      uint64_t* st = a.status + ((k * a.B + c) * S + s) * kStatusWords;
      publish_pair(st + 4, incl0, incl1, flag, lane);
    }
    q1 = static_cast<T>(s0);
    q2 = static_cast<T>(s1);
    pb0 = static_cast<T>(tb[0]);
    pb1 = static_cast<T>(tb[1]);
    pb2 = static_cast<T>(tb[2]);
    pa1 = a1;
    pa2 = a2;
    pg = table_g(tb);
  }
  if (vec)
    store_window_lds<T, T, W, true>(static_cast<T*>(a.dx) + c * a.T + t_lo, lds, G);
  else
    store_rev<T, W>(static_cast<T*>(a.dx) + c * a.T, n_hi, false, m, G);
}

// dsos[f, s, :] from the per-(window, channel) partial sums: one block of NT
// threads per (filter, section).  Thread t sums channels t, t + NT, ... of
// every window (no index division), four windows' loads in flight at a
// time, then a shared-memory tree; the order of the sums is fixed, so the
// result is deterministic.
template <typename TC, int NT>
__global__ __launch_bounds__(NT) void dsos_reduce_kernel(const double* partial,
                                                         const double* tables, TC* dsos,
                                                         int S, int64_t B, int64_t cpf,
                                                         int64_t n_windows) {
  const int64_t fs = blockIdx.x;  // f * S + s
  const int64_t f = fs / S;
  const int s = static_cast<int>(fs % S);
  const int64_t c_lo = f * cpf;
  const int64_t c_hi = (f + 1) * cpf < B ? (f + 1) * cpf : B;
  const double* ps = partial + static_cast<int64_t>(s) * n_windows * B * 5;
  double acc[5] = {0, 0, 0, 0, 0};
  for (int64_t c = c_lo + threadIdx.x; c < c_hi; c += NT) {
    int64_t kk = 0;
    for (; kk + 4 <= n_windows; kk += 4) {
      double v[4][5];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < 5; ++i) v[u][i] = ps[((kk + u) * B + c) * 5 + i];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < 5; ++i) acc[i] += v[u][i];
    }
    for (; kk < n_windows; ++kk)
#pragma unroll
      for (int i = 0; i < 5; ++i) acc[i] += ps[(kk * B + c) * 5 + i];
  }
  __shared__ double sh[5][NT];
#pragma unroll
  for (int i = 0; i < 5; ++i) sh[i][threadIdx.x] = acc[i];
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) {
#pragma unroll
      for (int i = 0; i < 5; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double* t = tables + fs * kTableStride;
    const double db0 = sh[0][0], db1 = sh[1][0], db2 = sh[2][0];
    const double da1 = -sh[3][0], da2 = -sh[4][0];
    const double a0 = t[5];
    double out[6];
    if (a0 == 1.0) {
      out[0] = db0;
      out[1] = db1;
      out[2] = db2;
      out[3] = 0.0;
      out[4] = da1;
      out[5] = da2;
    } else {
      out[0] = db0 / a0;
      out[1] = db1 / a0;
      out[2] = db2 / a0;
      out[3] = -(db0 * t[0] + db1 * t[1] + db2 * t[2] + da1 * t[3] + da2 * t[4]) / a0;
      out[4] = da1 / a0;
      out[5] = da2 / a0;
    }
    for (int i = 0; i < 6; ++i) dsos[fs * 6 + i] = static_cast<TC>(out[i]);
  }
}

// Block size by channels per filter: wide blocks for shared filters, narrow
// ones for per-channel filters (many rows, one channel each).
template <typename TC>
static void launch_dsos_reduce(const double* partial, const double* tables, TC* dsos, int S,
                               int64_t B, int64_t cpf, int64_t n_windows, int64_t n_rows,
                               hipStream_t st) {
  const dim3 grid(static_cast<unsigned>(n_rows));
  if (cpf >= 1024)
    hipLaunchKernelGGL((dsos_reduce_kernel<TC, 1024>), grid, dim3(1024), 0, st, partial, tables,
                       dsos, S, B, cpf, n_windows);
  else if (cpf >= 128)
    hipLaunchKernelGGL((dsos_reduce_kernel<TC, 256>), grid, dim3(256), 0, st, partial, tables,
                       dsos, S, B, cpf, n_windows);
  else
    hipLaunchKernelGGL((dsos_reduce_kernel<TC, 64>), grid, dim3(64), 0, st, partial, tables, dsos,
                       S, B, cpf, n_windows);
}

template <typename T>
__global__ void fill_kernel(T* p, int64_t n, T v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ void copy_kernel(const T* s, T* d, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

struct BwdLayout {
  FwdGeometry fwd;
  int W;
  int64_t n_windows;
  bool fused;
  size_t off_tables, off_status, off_saved, off_partial, total;
};

// Workspace: forward workspace | adjoint tables | scan status | [section
// outputs of the forward recompute] | coefficient partial sums.  The fused
// adjoint (all sections in one pass, chained windows) serves at least 1024
// channels, like the chained forward; fewer channels use one launch per
// section with the decoupled look-back.
// FP32 samples per lane of the fused adjoint: 32 unless TSIIR_ADJ_W=16 (kept
// for comparison measurements).
static int adj_w_f32() {
  static const int w = [] {
    const char* e = getenv("TSIIR_ADJ_W");
    return (e && strcmp(e, "16") == 0) ? 16 : 32;
  }();
  return w;
}

static bool bwd_layout(const tsiir_sosfilt_desc* d, BwdLayout* L, bool recompute) {
  if (!fwd_geometry(d, &L->fwd)) return false;
  const bool f64 = d->io_dtype == TSIIR_F64;
  L->fused = d->channels >= 1024 && d->n_sections <= kWave;
  L->W = f64 ? 16 : (L->fused ? adj_w_f32() : 32);
  const int64_t win = static_cast<int64_t>(kWave) * L->W;
  L->n_windows = (d->n_samples + win - 1) / win;
  const size_t esz = f64 ? 8 : 4;
  L->off_tables = align_up(L->fwd.total, 256);
  L->off_status = L->off_tables + L->fwd.table_bytes;
  const size_t status_sets = L->fused ? static_cast<size_t>(d->n_sections) : 1;
  L->off_saved = align_up(L->off_status + static_cast<size_t>(L->n_windows) * d->channels *
                                              status_sets * kStatusWords * 8,
                          256);
  const size_t saved_bytes =
      recompute ? static_cast<size_t>(d->n_sections) * d->channels * d->n_samples * esz : 0;
  L->off_partial = align_up(L->off_saved + saved_bytes, 256);
  L->total = L->off_partial +
             static_cast<size_t>(d->n_sections) * L->n_windows * d->channels * 5 * 8;
  return true;
}

// y: forward output, saved: (S-1, B, T) section outputs.  Both null: the
// forward is recomputed into the workspace first.
template <typename T>
static tsiir_status backward_typed(const tsiir_sosfilt_desc* d, const BwdLayout& L, const void* x,
                                   const void* y, const void* saved, const void* sos,
                                   const void* zi, const void* dy, const void* dzf, void* dx,
                                   void* dsos, void* dzi, char* ws, hipStream_t st) {
  const int64_t B = d->channels, T_ = d->n_samples;
  const int S = d->n_sections;
  const int64_t BT = B * T_;
  double* adj_tables = reinterpret_cast<double*>(ws + L.off_tables);
  uint64_t* adj_status = reinterpret_cast<uint64_t*>(ws + L.off_status);
  double* partial = reinterpret_cast<double*>(ws + L.off_partial);

  if (y == nullptr) {
    // 1. forward recompute, keeping every section's output (y -> [S-1])
    T* sv = reinterpret_cast<T*>(ws + L.off_saved);
    tsiir_sosfilt_desc fd = *d;
    fd.flags = d->flags & (TSIIR_HAS_ZI | TSIIR_ZI_BROADCAST);
    fd.x_rows = B;
    fd.x_row_stride = T_;
    fd.y_row_stride = T_;
    tsiir_status rc = forward_impl(&fd, x, sv + static_cast<int64_t>(S - 1) * BT, sos, zi,
                                   nullptr, ws, L.fwd.total, st, sv);
    if (rc != TSIIR_OK) return rc;
    y = sv + static_cast<int64_t>(S - 1) * BT;
    saved = sv;
  }
  const T* yv = static_cast<const T*>(y);
  const T* sv = static_cast<const T*>(saved);

  // 2. adjoint tables (one epoch per launch set: 1 fused, S per section)
  const int64_t n_rows = L.fwd.n_filters * S;
  launch_prep(sos, d->coef_dtype, n_rows, L.W == 32 ? 5 : 4, true, adj_tables,
              reinterpret_cast<WsHeader*>(ws), static_cast<uint32_t>(L.fused ? 1 : S), st);
  const int64_t win = static_cast<int64_t>(kWave) * L.W;
  const int64_t full = T_ / win;
  const unsigned waves = static_cast<unsigned>(full * B);
  const bool tail = full * win < T_;

  if (L.fused) {
    // 3. all sections in one pass
    AdjFArgs a;
    a.dy = dy;
    a.y = yv;
    a.x = x;
    a.saved = sv;
    a.dzf = dzf;
    a.dx = dx;
    a.dzi = dzi;
    a.partial = partial;
    a.status = adj_status;
    a.hdr = reinterpret_cast<WsHeader*>(ws);
    a.B = B;
    a.T = T_;
    a.cpf = d->chans_per_filter;
    a.n_windows = L.n_windows;
    a.S = S;
    if (L.W == 32) {
      if (waves) hipLaunchKernelGGL((sos_adjoint_fused_kernel<T, 32, false>), dim3(waves),
                                    dim3(kWave), 0, st, a, int64_t(0), adj_tables);
      if (tail) hipLaunchKernelGGL((sos_adjoint_fused_kernel<T, 32, true>),
                                   dim3(static_cast<unsigned>(B)), dim3(kWave), 0, st, a, full,
                                   adj_tables);
    } else {
      if (waves) hipLaunchKernelGGL((sos_adjoint_fused_kernel<T, 16, false>), dim3(waves),
                                    dim3(kWave), 0, st, a, int64_t(0), adj_tables);
      if (tail) hipLaunchKernelGGL((sos_adjoint_fused_kernel<T, 16, true>),
                                   dim3(static_cast<unsigned>(B)), dim3(kWave), 0, st, a, full,
                                   adj_tables);
    }
    if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
  } else {
    // 3. sections last to first
    for (int s = S - 1; s >= 0; --s) {
      AdjArgs a;
      a.g = (s == S - 1) ? dy : dx;
      a.du = dx;
      a.u = (s == 0) ? x : static_cast<const void*>(sv + static_cast<int64_t>(s - 1) * BT);
      a.v = (s == S - 1) ? static_cast<const void*>(yv)
                         : static_cast<const void*>(sv + static_cast<int64_t>(s) * BT);
      a.dzf = dzf;
      a.dzi = dzi;
      a.partial = partial + static_cast<int64_t>(s) * L.n_windows * B * 5;
      a.tables = adj_tables;
      a.status = adj_status;
      a.hdr = reinterpret_cast<WsHeader*>(ws);
      a.B = B;
      a.T = T_;
      a.cpf = d->chans_per_filter;
      a.n_windows = L.n_windows;
      a.S = S;
      a.s = s;
      if (L.W == 32) {
        if (waves) hipLaunchKernelGGL((sos_adjoint_kernel<T, 32, false>), dim3(waves),
                                      dim3(kWave), 0, st, a, int64_t(0), adj_tables);
        if (tail) hipLaunchKernelGGL((sos_adjoint_kernel<T, 32, true>),
                                     dim3(static_cast<unsigned>(B)), dim3(kWave), 0, st, a, full,
                                     adj_tables);
      } else {
        if (waves) hipLaunchKernelGGL((sos_adjoint_kernel<T, 16, false>), dim3(waves),
                                      dim3(kWave), 0, st, a, int64_t(0), adj_tables);
        if (tail) hipLaunchKernelGGL((sos_adjoint_kernel<T, 16, true>),
                                     dim3(static_cast<unsigned>(B)), dim3(kWave), 0, st, a, full,
                                     adj_tables);
      }
      if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
    }
  }

  // 4. coefficient gradients
  if (dsos != nullptr) {
    if (d->coef_dtype == TSIIR_F64)
      launch_dsos_reduce<double>(partial, adj_tables, static_cast<double*>(dsos), S, B,
                                 d->chans_per_filter, L.n_windows, n_rows, st);
    else
      launch_dsos_reduce<float>(partial, adj_tables, static_cast<float*>(dsos), S, B,
                                d->chans_per_filter, L.n_windows, n_rows, st);
  }
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

template <typename T>
static tsiir_status backward_empty(const tsiir_sosfilt_desc* d, const void* dzf, void* dsos,
                                   void* dzi, int64_t n_filters, hipStream_t st) {
  const int64_t nz = d->channels * d->n_sections * 2;
  const int64_t ns = n_filters * d->n_sections * 6;
  if (dzi) {
    if (dzf)
      hipLaunchKernelGGL(copy_kernel<T>, dim3(static_cast<unsigned>((nz + 255) / 256)),
                         dim3(256), 0, st, static_cast<const T*>(dzf), static_cast<T*>(dzi), nz);
    else
      hipLaunchKernelGGL(fill_kernel<T>, dim3(static_cast<unsigned>((nz + 255) / 256)),
                         dim3(256), 0, st, static_cast<T*>(dzi), nz, T(0));
  }
  if (dsos) {
    if (d->coef_dtype == TSIIR_F64)
      hipLaunchKernelGGL(fill_kernel<double>, dim3(static_cast<unsigned>((ns + 255) / 256)),
                         dim3(256), 0, st, static_cast<double*>(dsos), ns, 0.0);
    else
      hipLaunchKernelGGL(fill_kernel<float>, dim3(static_cast<unsigned>((ns + 255) / 256)),
                         dim3(256), 0, st, static_cast<float*>(dsos), ns, 0.0f);
  }
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // namespace tsiir

using namespace tsiir;

extern "C" {

size_t tsiir_sosfilt_backward_workspace_bytes(const tsiir_sosfilt_desc* d) {
  BwdLayout L;
  if (!bwd_layout(d, &L, true)) return 0;
  return L.total;
}

size_t tsiir_sosfilt_backward_saved_workspace_bytes(const tsiir_sosfilt_desc* d) {
  BwdLayout L;
  if (!bwd_layout(d, &L, false)) return 0;
  return L.total;
}

static tsiir_status backward_entry(const tsiir_sosfilt_desc* d, const void* x, const void* y,
                                   const void* saved, const void* sos, const void* zi,
                                   const void* dy, const void* dzf, void* dx, void* dsos,
                                   void* dzi, void* workspace, size_t workspace_bytes,
                                   void* stream) {
  BwdLayout L;
  if (!bwd_layout(d, &L, y == nullptr)) return TSIIR_EINVAL;
  if (d->io_dtype != TSIIR_F32 && d->io_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  if (d->coef_dtype != TSIIR_F32 && d->coef_dtype != TSIIR_F64) return TSIIR_EINVAL;
  if (d->mode != TSIIR_MODE_PLAIN) return TSIIR_EUNSUPPORTED;
  if (d->x_rows != d->channels) return TSIIR_EUNSUPPORTED;
  if ((d->flags & TSIIR_HAS_ZI) && !zi) return TSIIR_EINVAL;
  if (d->channels == 0) return TSIIR_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool f64 = d->io_dtype == TSIIR_F64;
  if (d->n_samples == 0)
    return f64 ? backward_empty<double>(d, dzf, dsos, dzi, L.fwd.n_filters, st)
               : backward_empty<float>(d, dzf, dsos, dzi, L.fwd.n_filters, st);
  if (!x || !sos || !dy || !dx) return TSIIR_EINVAL;
  if (!workspace || workspace_bytes < L.total) return TSIIR_EWORKSPACE;
  if (L.n_windows * d->channels > 0x7fffffffLL) return TSIIR_EUNSUPPORTED;
  char* ws = static_cast<char*>(workspace);
  return f64 ? backward_typed<double>(d, L, x, y, saved, sos, zi, dy, dzf, dx, dsos, dzi, ws, st)
             : backward_typed<float>(d, L, x, y, saved, sos, zi, dy, dzf, dx, dsos, dzi, ws, st);
}

tsiir_status tsiir_sosfilt_backward(const tsiir_sosfilt_desc* d, const void* x, const void* sos,
                                    const void* zi, const void* dy, const void* dzf, void* dx,
                                    void* dsos, void* dzi, void* workspace,
                                    size_t workspace_bytes, void* stream) {
  return backward_entry(d, x, nullptr, nullptr, sos, zi, dy, dzf, dx, dsos, dzi, workspace,
                        workspace_bytes, stream);
}

tsiir_status tsiir_sosfilt_backward_saved(const tsiir_sosfilt_desc* d, const void* x,
                                          const void* y, const void* saved, const void* sos,
                                          const void* zi, const void* dy, const void* dzf,
                                          void* dx, void* dsos, void* dzi, void* workspace,
                                          size_t workspace_bytes, void* stream) {
  if (!y || (d && d->n_sections > 1 && !saved)) return TSIIR_EINVAL;
  return backward_entry(d, x, y, saved, sos, zi, dy, dzf, dx, dsos, dzi, workspace,
                        workspace_bytes, stream);
}

}  // extern "C"
