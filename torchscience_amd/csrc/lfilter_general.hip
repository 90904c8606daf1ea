// lfilter_general.hip - DF2T of arbitrary order N (1..64), one lane per
// channel, serial in time.
//
// Replaces _lfilter.py:153-166 for filters of order > 2 (long FIR taps,
// direct-form IIR polynomials).  Coefficients are normalised by a[0] when
// a[0] != 1 (_lfilter.py:96-98) and were zero-padded to N+1 by the caller
// (_lfilter.py:114-117).  The state lives in registers for N <= 16 and in
// per-lane scratch above that.  Each lane streams its row in blocks of 16
// samples (8 for orders 9..16), the next block's loads issued before the
// current block's recurrence, so the serial chain does not wait on memory
// every sample.  TSIIR_REVERSE runs along decreasing time (processing index
// n is time T-1-n, for input and output alike): the second pass of filtfilt
// without flipped copies.

#include "sos_common.h"
#include "tsiir.h"

namespace tsiir {

// samples per lane per block: fewer for deep states, whose unrolled
// recurrence would not fit the registers
template <int NMAX>
constexpr int kGenBlk = NMAX <= 8 ? 16 : (NMAX <= 16 ? 8 : 1);

// A row along processing index n: p[n], or with REV p[-n], p then being the
// row's last element (time T-1).
template <typename P, bool REV>
struct RowRef {
  P* p;
  __device__ __forceinline__ P& operator[](int64_t n) const { return p[REV ? -n : n]; }
};
template <bool REV, typename P>
__device__ __forceinline__ RowRef<P, REV> row_ref(P* row, int64_t T) {
  return RowRef<P, REV>{REV ? row + (T - 1) : row};
}

template <typename TACC, int NMAX>
__device__ inline void df2t_step(const TACC (&bn)[NMAX + 1], const TACC (&an)[NMAX + 1],
                                 int N, TACC (&z)[NMAX], TACC xn, TACC& yn) {
  yn = bn[0] * xn + z[0];
#pragma unroll
  for (int k = 0; k < NMAX - 1; ++k)
    if (k < N - 1) z[k] = (bn[k + 1] * xn - an[k + 1] * yn) + z[k + 1];
#pragma unroll
  for (int k = 0; k < NMAX; ++k)
    if (k == N - 1) z[k] = bn[N] * xn - an[N] * yn;
}

// df2t_step at order NMAX over coefficients zero-padded past N, with no
// test of N: the states past N - 1 stay zero, and the value of every other
// state and output is the same (up to the sign of a zero: z[N - 1] gains a
// "+ 0").  With N a run-time value the tests of df2t_step cost more than
// the padded terms (branches and state moves around each update).
template <typename TACC, int NMAX>
__device__ __forceinline__ void df2t_step_padded(const TACC (&bn)[NMAX + 1],
                                                 const TACC (&an)[NMAX + 1], TACC (&z)[NMAX],
                                                 TACC xn, TACC& yn) {
  yn = bn[0] * xn + z[0];
#pragma unroll
  for (int k = 0; k < NMAX - 1; ++k) z[k] = (bn[k + 1] * xn - an[k + 1] * yn) + z[k + 1];
  z[NMAX - 1] = bn[NMAX] * xn - an[NMAX] * yn;
}

template <typename TACC, int NMAX>
__device__ inline void load_coef(const TACC* __restrict__ coef, int N, TACC (&bn)[NMAX + 1],
                                 TACC (&an)[NMAX + 1]) {
  const TACC a0 = coef[N + 1];
  const bool norm = a0 != TACC(1);
#pragma unroll
  for (int i = 0; i <= NMAX; ++i) {
    if (i <= N) {
      bn[i] = norm ? coef[i] / a0 : coef[i];
      an[i] = norm ? coef[N + 1 + i] / a0 : coef[N + 1 + i];
    } else {
      bn[i] = 0;
      an[i] = 0;
    }
  }
}

// DF2T over xr[n0, n1) from state z (updated in place); outputs to yr when
// OUT.  Whole blocks of BLK samples have the next block's loads in flight
// while the current block runs through the recurrence.
template <typename TACC, int NMAX, bool OUT, typename XR, typename YR>
__device__ inline void run_span(const TACC (&bn)[NMAX + 1], const TACC (&an)[NMAX + 1], int N,
                                TACC (&z)[NMAX], XR xr, YR yr, int64_t n0, int64_t n1) {
  constexpr int BLK = kGenBlk<NMAX>;
  const int64_t nb = BLK > 1 ? n0 + (n1 - n0) / BLK * BLK : n0;
  if constexpr (BLK > 1) {
    TACC cur[BLK], nxt[BLK];
    if (nb > n0) {
#pragma unroll
      for (int i = 0; i < BLK; ++i) cur[i] = xr[n0 + i];
    }
    for (int64_t m0 = n0; m0 < nb; m0 += BLK) {
      if (m0 + BLK < nb) {
#pragma unroll
        for (int i = 0; i < BLK; ++i) nxt[i] = xr[m0 + BLK + i];
      }
#pragma unroll
      for (int i = 0; i < BLK; ++i) {
        TACC yn;
        df2t_step<TACC, NMAX>(bn, an, N, z, cur[i], yn);
        if (OUT) yr[m0 + i] = yn;
      }
#pragma unroll
      for (int i = 0; i < BLK; ++i) cur[i] = nxt[i];
    }
  }
  for (int64_t n = nb; n < n1; ++n) {
    TACC yn;
    df2t_step<TACC, NMAX>(bn, an, N, z, xr[n], yn);
    if (OUT) yr[n] = yn;
  }
}

// Serial form: one lane per channel walks the whole row.
template <typename TACC, int NMAX, bool REV>
__global__ __launch_bounds__(64) void lfilter_general_kernel(
    const TACC* __restrict__ coef, int N, const TACC* __restrict__ x, TACC* __restrict__ y,
    int64_t B, int64_t T, int64_t x_rows, int64_t x_stride, int64_t y_stride,
    const TACC* __restrict__ zi, TACC* __restrict__ zf, uint32_t flags) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= B) return;
  TACC bn[NMAX + 1], an[NMAX + 1], z[NMAX];
  load_coef<TACC, NMAX>(coef, N, bn, an);
  const int64_t zrow = (flags & TSIIR_ZI_BROADCAST) ? 0 : c;
#pragma unroll
  for (int k = 0; k < NMAX; ++k)
    z[k] = (k < N && (flags & TSIIR_HAS_ZI)) ? zi[zrow * N + k] : TACC(0);
  const auto xr = row_ref<REV>(x + (c % x_rows) * x_stride, T);
  const auto yr = row_ref<REV>(y + c * y_stride, T);
  if constexpr (kGenBlk<NMAX> > 1) {
    run_span<TACC, NMAX, true>(bn, an, N, z, xr, yr, 0, T);
  } else {
    // deep states: the loop written out here (through run_span it spills)
    for (int64_t n = 0; n < T; ++n) {
      const TACC xn = xr[n];
      const TACC yn = bn[0] * xn + z[0];
#pragma unroll
      for (int k = 0; k < NMAX - 1; ++k)
        if (k < N - 1) z[k] = (bn[k + 1] * xn - an[k + 1] * yn) + z[k + 1];
#pragma unroll
      for (int k = 0; k < NMAX; ++k)
        if (k == N - 1) z[k] = bn[N] * xn - an[N] * yn;
      yr[n] = yn;
    }
  }
  if (flags & TSIIR_WRITE_ZF) {
#pragma unroll
    for (int k = 0; k < NMAX; ++k)
      if (k < N) zf[c * N + k] = z[k];
  }
}

// Chunked form.  Workgroup (c, g) owns chunks 64 g .. 64 g + 63 of channel
// c, lane l chunk j = 64 g + l, samples [j L, min((j + 1) L, T)), L =
// kGenChunk.  First pass (SECOND false): filter the chunk from zero state
// and store the final state (FP64) in ws[(c M + j) N ..].  The scan kernel
// replaces those with the chunk start states.  Second pass: filter the
// chunk from its start state, writing y (and zf from the last chunk).
// Full chunks advance together in blocks of BLK samples.  Each block's
// inputs come through an LDS tile: the wave loads the 64 chunks' BLK-sample
// segments as 16-byte vectors (an instruction covers 64 / (BLK / V)
// segments rather than 64 scattered addresses) and each lane reads its own
// row of the tile; the next block's vectors are in flight while the current
// block runs through the recurrence.  In the second pass the outputs go out
// the same way.  Along decreasing time (REV) a segment is a vector run of
// memory read or written in reverse order.  Rows that are not 16-byte
// aligned (or, with REV, whose length is not a multiple of V) take
// per-lane element accesses instead.  A partial last chunk runs after that,
// sample by sample.
constexpr int kGenChunk = 256;

template <typename TACC, int NMAX, bool SECOND, bool REV>
__global__ __launch_bounds__(64) void lfilter_chunk_kernel(
    const TACC* __restrict__ coef, int N, const TACC* __restrict__ x, TACC* __restrict__ y,
    int64_t T, int64_t M, int64_t G, int64_t x_rows, int64_t x_stride, int64_t y_stride,
    double* __restrict__ ws, TACC* __restrict__ zf, uint32_t flags) {
  constexpr int BLK = kGenBlk<NMAX>;
  static_assert(BLK > 1 && kGenChunk % BLK == 0, "chunked form needs register blocks");
  // tile rows of BLK samples padded to a multiple of 4 (16-byte vectors)
  constexpr int TS = BLK + 4;
  constexpr int V = 16 / sizeof(TACC);            // samples per 16-byte vector
  constexpr int VPR = BLK / V;                    // vectors per tile row
  constexpr int NB = kGenChunk / BLK;             // blocks per chunk
  static_assert(BLK % V == 0 && (TS * sizeof(TACC)) % 16 == 0, "tile rows of whole vectors");
  __shared__ __align__(16) TACC tile[64 * TS];
  const int l = threadIdx.x;
  const int64_t c = blockIdx.x / G;
  const int64_t g = blockIdx.x - c * G;
  const int64_t j = g * 64 + l;
  const bool valid = j < M;
  const int64_t n0 = j * kGenChunk;
  const int64_t n1 = valid ? (n0 + kGenChunk < T ? n0 + kGenChunk : T) : n0;
  const bool full = valid && n1 - n0 == kGenChunk;
  TACC bn[NMAX + 1], an[NMAX + 1], z[NMAX];
  load_coef<TACC, NMAX>(coef, N, bn, an);
  double* st = ws + (c * M + j) * N;
#pragma unroll
  for (int k = 0; k < NMAX; ++k)
    z[k] = (SECOND && valid && k < N) ? static_cast<TACC>(st[k]) : TACC(0);
  const TACC* xrow = x + (c % x_rows) * x_stride;
  TACC* yrow = y + c * y_stride;
  const auto xr = row_ref<REV>(xrow, T);
  const auto yr = row_ref<REV>(yrow, T);
  // vector I/O: every segment of a full chunk starts on a 16-byte boundary
  // (processing index n is memory n, or with REV memory T - 1 - n)
  const bool t_ok = !REV || T % V == 0;
  const bool xv_ok = t_ok && (reinterpret_cast<uintptr_t>(xrow) & 15) == 0;
  const bool yv_ok =
      t_ok && ((reinterpret_cast<uintptr_t>(yrow) | (y_stride * sizeof(TACC))) & 15) == 0;
  // a tile row holds a block in memory order: the memory start of its
  // vector `part` for chunk js's block sb (processing indices m .. m + BLK
  // - 1, with REV the memory run T - m - BLK .. T - m - 1)
  auto seg = [&](int64_t js, int sb, int part) __attribute__((always_inline)) -> int64_t {
    const int64_t m = js * kGenChunk + sb * BLK;
    return REV ? T - m - BLK + part * V : m + part * V;
  };
  u32x4 xin[VPR];
  // the wave's loads of block sb: vector e = l + 64 q is vector e % VPR of
  // chunk 64 g + e / VPR (only full chunks), one 16-byte load or, off a
  // 16-byte boundary, V element loads
  auto load_block = [&](int sb) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < VPR; ++q) {
      const int e = l + 64 * q;
      const int64_t js = g * 64 + e / VPR;
      if ((js + 1) * kGenChunk <= T) {
        const TACC* src = xrow + seg(js, sb, e % VPR);
        if (xv_ok) {
          xin[q] = *reinterpret_cast<const u32x4*>(src);
        } else {
          TACC v[V];
#pragma unroll
          for (int i = 0; i < V; ++i) v[i] = src[i];
          __builtin_memcpy(&xin[q], v, sizeof(v));
        }
      }
    }
  };
  // block sb's vectors into the tile in memory order, then each lane's row
  // out in processing order
  auto take_block = [&](TACC (&cur)[BLK]) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < VPR; ++q) {
      const int e = l + 64 * q;
      const int64_t js = g * 64 + e / VPR;
      if ((js + 1) * kGenChunk <= T)
        *reinterpret_cast<u32x4*>(&tile[(e / VPR) * TS + (e % VPR) * V]) = xin[q];
    }
    __syncthreads();
    if (full) {
#pragma unroll
      for (int i = 0; i < BLK; ++i) cur[i] = tile[l * TS + (REV ? BLK - 1 - i : i)];
    }
    __syncthreads();
  };
  TACC cur[BLK];
  load_block(0);
  take_block(cur);
  for (int sb = 0; sb < NB; ++sb) {
    if (sb + 1 < NB) load_block(sb + 1);
    TACC yv[BLK];
    if (full) {
#pragma unroll
      for (int i = 0; i < BLK; ++i) df2t_step_padded<TACC, NMAX>(bn, an, z, cur[i], yv[i]);
    }
    if constexpr (SECOND) {
      if (full) {
#pragma unroll
        for (int i = 0; i < BLK; ++i) tile[l * TS + (REV && yv_ok ? BLK - 1 - i : i)] = yv[i];
      }
      __syncthreads();
      // vector e = l + 64 q: source lane e / VPR, its samples from V (e %
      // VPR); with yv_ok the tile row holds them in memory order
#pragma unroll
      for (int q = 0; q < VPR; ++q) {
        const int e = l + 64 * q;
        const int src = e / VPR, part = e % VPR;
        const int64_t js = g * 64 + src;
        if ((js + 1) * kGenChunk > T) continue;
        if (yv_ok) {
          *reinterpret_cast<u32x4*>(yrow + seg(js, sb, part)) =
              *reinterpret_cast<const u32x4*>(tile + src * TS + part * V);
        } else {
          const int64_t k0 = js * kGenChunk + sb * BLK + part * V;
          const TACC* from = tile + src * TS + part * V;
#pragma unroll
          for (int v = 0; v < V; ++v) yr[k0 + v] = from[v];
        }
      }
      __syncthreads();
    }
    if (sb + 1 < NB) take_block(cur);
  }
  if (valid && !full) {
    for (int64_t n = n0; n < n1; ++n) {
      TACC yn;
      df2t_step_padded<TACC, NMAX>(bn, an, z, xr[n], yn);
      if (SECOND) yr[n] = yn;
    }
  }
  if (!valid) return;
  if (!SECOND) {
#pragma unroll
    for (int k = 0; k < NMAX; ++k)
      if (k < N) st[k] = static_cast<double>(z[k]);
  } else if ((flags & TSIIR_WRITE_ZF) && j == M - 1) {
#pragma unroll
    for (int k = 0; k < NMAX; ++k)
      if (k < N) zf[c * N + k] = z[k];
  }
}

// Chunk start states in FP64: S_0 = zi (or 0), S_{j+1} = A^L S_j + s_j with
// s_j the zero-state final state of chunk j and A the DF2T state matrix
// (A[r][0] = -a_{r+1}, A[r][r+1] = 1, a normalised by a0).  Lane c < N
// forms column c of P = A^L by L products with the companion matrix (O(N)
// each): repeated squaring loses the accuracy of this non-normal matrix
// (relative error of A^256 for an order-16 Butterworth lowpass above 1,
// against 4e-10 this way).  Then NMAX lanes per channel walk its chunks,
// lane r holding row r of P and component r of the state: each step takes
// the state's components from the channel's lanes (NMAX shuffles) and forms
// S_{j+1}[r] = s_j[r] + sum_m P[r][m] S_j[m] in the order m = 0, 1, ...
// (64 / NMAX channels per wave, the chunk states loaded kScanGroup chunks
// ahead, a channel's lanes reading and writing one contiguous run).
template <typename TACC, int NMAX>
__global__ __launch_bounds__(64) void lfilter_scan_kernel(const TACC* __restrict__ coef, int N,
                                                          int64_t B, int64_t M,
                                                          double* __restrict__ ws,
                                                          const TACC* __restrict__ zi,
                                                          uint32_t flags) {
  __shared__ double P[NMAX * NMAX], na[NMAX];
  const int l = threadIdx.x;
  const double a0 = static_cast<double>(coef[N + 1]);
  const bool norm = a0 != 1.0;
  if (l < N) {
    const double ar = static_cast<double>(coef[N + 2 + l]);
    na[l] = -(norm ? ar / a0 : ar);
  }
  __syncthreads();
  if (l < N) {
    // column l of A^L: e_l through L steps of v <- A v,
    // (A v)[r] = -a_{r+1} v[0] + v[r+1]
    double v[NMAX];
#pragma unroll
    for (int k = 0; k < NMAX; ++k) v[k] = k == l ? 1.0 : 0.0;
    for (int it = 0; it < kGenChunk; ++it) {
      const double v0 = v[0];
#pragma unroll
      for (int r = 0; r < NMAX; ++r)
        if (r < N) v[r] = na[r] * v0 + (r + 1 < N ? v[r + 1] : 0.0);
    }
#pragma unroll
    for (int r = 0; r < NMAX; ++r)
      if (r < N) P[r * N + l] = v[r];
  }
  __syncthreads();
  constexpr int LPC = NMAX;                       // lanes per channel
  const int r = l % LPC;
  const int64_t c = (static_cast<int64_t>(blockIdx.x) * blockDim.x + l) / LPC;
  const bool live = c < B;
  const int base = l - r;                         // the channel's first lane
  double Pr[NMAX];
#pragma unroll
  for (int m = 0; m < NMAX; ++m) Pr[m] = (r < N && m < N) ? P[r * N + m] : 0.0;
  const int64_t zrow = (flags & TSIIR_ZI_BROADCAST) ? 0 : c;
  double* w = ws + (live ? c : 0) * M * N;
  const bool mine = live && r < N;                // this lane's component is stored
  double S = (mine && (flags & TSIIR_HAS_ZI)) ? static_cast<double>(zi[zrow * N + r]) : 0.0;
  constexpr int kScanGroup = 8;
  double cur[kScanGroup], nxt[kScanGroup];
#pragma unroll
  for (int q = 0; q < kScanGroup; ++q) cur[q] = (mine && q < M) ? w[q * N + r] : 0.0;
  for (int64_t j0 = 0; j0 < M; j0 += kScanGroup) {
#pragma unroll
    for (int q = 0; q < kScanGroup; ++q)
      nxt[q] = (mine && j0 + kScanGroup + q < M) ? w[(j0 + kScanGroup + q) * N + r] : 0.0;
#pragma unroll
    for (int q = 0; q < kScanGroup; ++q) {
      if (j0 + q < M) {
        if (mine) w[(j0 + q) * N + r] = S;
        double acc = cur[q];
#pragma unroll
        for (int m = 0; m < NMAX; ++m) acc += Pr[m] * __shfl(S, base + m);
        S = acc;
      }
    }
#pragma unroll
    for (int q = 0; q < kScanGroup; ++q) cur[q] = nxt[q];
  }
}

// chunks per channel of the chunked form, or 0 when the serial form is used
// (order above 16, fewer than two chunks, or no workspace)
constexpr int kGenChunkMaxOrder = 16;

int64_t general_chunks(const tsiir_sosfilt_desc* d) {
  if (d->n_sections > kGenChunkMaxOrder) return 0;
  const int64_t M = (d->n_samples + kGenChunk - 1) / kGenChunk;
  return M >= 2 ? M : 0;
}

template <typename TACC, bool REV>
static tsiir_status launch_general(const tsiir_sosfilt_desc* d, const void* coef, const void* x,
                                   void* y, const void* zi, void* zf, void* ws, hipStream_t st) {
  const int N = d->n_sections;
  const int64_t M = ws ? general_chunks(d) : 0;
  if (M > 0) {
    const int64_t G = (M + 63) / 64;   // workgroups per channel
    if (d->channels * G > 0x7fffffff) return TSIIR_EUNSUPPORTED;
    const unsigned cblocks = static_cast<unsigned>(d->channels * G);
    // the scan: N rounded up to the kernel's order (4, 8 or 16) lanes per channel
    const int64_t lpc = N <= 4 ? 4 : (N <= 8 ? 8 : 16);
    const unsigned sblocks = static_cast<unsigned>((d->channels * lpc + 63) / 64);
    auto go = [&](auto first, auto scan, auto second) {
      hipLaunchKernelGGL(first, dim3(cblocks), dim3(64), 0, st, static_cast<const TACC*>(coef),
                         N, static_cast<const TACC*>(x), static_cast<TACC*>(y),
                         d->n_samples, M, G, d->x_rows, d->x_row_stride, d->y_row_stride,
                         static_cast<double*>(ws), static_cast<TACC*>(zf), d->flags);
      hipLaunchKernelGGL(scan, dim3(sblocks), dim3(64), 0, st, static_cast<const TACC*>(coef), N,
                         d->channels, M, static_cast<double*>(ws),
                         static_cast<const TACC*>(zi), d->flags);
      hipLaunchKernelGGL(second, dim3(cblocks), dim3(64), 0, st,
                         static_cast<const TACC*>(coef), N, static_cast<const TACC*>(x),
                         static_cast<TACC*>(y), d->n_samples, M, G, d->x_rows,
                         d->x_row_stride, d->y_row_stride, static_cast<double*>(ws),
                         static_cast<TACC*>(zf), d->flags);
    };
    if (N <= 4)
      go(lfilter_chunk_kernel<TACC, 4, false, REV>, lfilter_scan_kernel<TACC, 4>,
         lfilter_chunk_kernel<TACC, 4, true, REV>);
    else if (N <= 8)
      go(lfilter_chunk_kernel<TACC, 8, false, REV>, lfilter_scan_kernel<TACC, 8>,
         lfilter_chunk_kernel<TACC, 8, true, REV>);
    else
      go(lfilter_chunk_kernel<TACC, 16, false, REV>, lfilter_scan_kernel<TACC, 16>,
         lfilter_chunk_kernel<TACC, 16, true, REV>);
    return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
  }
  const unsigned blocks = static_cast<unsigned>((d->channels + 63) / 64);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64), 0, st, static_cast<const TACC*>(coef), N,
                       static_cast<const TACC*>(x), static_cast<TACC*>(y), d->channels,
                       d->n_samples, d->x_rows, d->x_row_stride, d->y_row_stride,
                       static_cast<const TACC*>(zi), static_cast<TACC*>(zf), d->flags);
  };
  if (N <= 4)
    go(lfilter_general_kernel<TACC, 4, REV>);
  else if (N <= 8)
    go(lfilter_general_kernel<TACC, 8, REV>);
  else if (N <= 16)
    go(lfilter_general_kernel<TACC, 16, REV>);
  else
    go(lfilter_general_kernel<TACC, 64, REV>);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // namespace tsiir

using namespace tsiir;

static tsiir_status launch_general_io(const tsiir_sosfilt_desc* d, const void* coef,
                                      const void* x, void* y, const void* zi, void* zf, void* ws,
                                      hipStream_t st) {
  const bool rev = (d->flags & TSIIR_REVERSE) != 0;
  if (d->io_dtype == TSIIR_F64)
    return rev ? launch_general<double, true>(d, coef, x, y, zi, zf, ws, st)
               : launch_general<double, false>(d, coef, x, y, zi, zf, ws, st);
  return rev ? launch_general<float, true>(d, coef, x, y, zi, zf, ws, st)
             : launch_general<float, false>(d, coef, x, y, zi, zf, ws, st);
}

static tsiir_status check_general(const tsiir_sosfilt_desc* d, const void* coef, const void* x,
                                  void* y, const void* zi, void* zf) {
  if (!d || d->channels < 0 || d->n_samples < 0) return TSIIR_EINVAL;
  if (d->n_sections < 1 || d->n_sections > 64) return TSIIR_EUNSUPPORTED;
  if (d->io_dtype != d->coef_dtype) return TSIIR_EINVAL;
  if (d->io_dtype != TSIIR_F32 && d->io_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  if (d->x_rows < 1) return TSIIR_EINVAL;
  if ((d->flags & TSIIR_HAS_ZI) && !zi) return TSIIR_EINVAL;
  if ((d->flags & TSIIR_WRITE_ZF) && !zf) return TSIIR_EINVAL;
  if (d->channels > 0 && (!coef || (d->n_samples > 0 && (!x || !y)))) return TSIIR_EINVAL;
  return TSIIR_OK;
}

extern "C" tsiir_status tsiir_lfilter_forward(const tsiir_sosfilt_desc* d, const void* coef,
                                              const void* x, void* y, const void* zi, void* zf,
                                              void* stream) {
  const tsiir_status s = check_general(d, coef, x, y, zi, zf);
  if (s != TSIIR_OK || d->channels == 0) return s;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return launch_general_io(d, coef, x, y, zi, zf, nullptr, st);
}

extern "C" size_t tsiir_lfilter_workspace_bytes(const tsiir_sosfilt_desc* d) {
  if (!d || d->channels <= 0 || d->n_samples <= 0 || d->n_sections < 1) return 0;
  return static_cast<size_t>(d->channels) * general_chunks(d) * d->n_sections * sizeof(double);
}

extern "C" tsiir_status tsiir_lfilter_forward_ws(const tsiir_sosfilt_desc* d, const void* coef,
                                                 const void* x, void* y, const void* zi,
                                                 void* zf, void* workspace,
                                                 size_t workspace_bytes, void* stream) {
  const tsiir_status s = check_general(d, coef, x, y, zi, zf);
  if (s != TSIIR_OK || d->channels == 0) return s;
  const size_t need = tsiir_lfilter_workspace_bytes(d);
  if (need > 0 && (!workspace || workspace_bytes < need)) return TSIIR_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  void* ws = need > 0 ? workspace : nullptr;
  return launch_general_io(d, coef, x, y, zi, zf, ws, st);
}
