// fir.hip - direct-form FIR filtering and its tap gradient.
//
// Serves every FIR use of the filtering path with one kernel:
//  * lfilter(b, a, x, zi) with a single denominator coefficient
//    (_lfilter.py:96-166 with a = [a0]: the DF2T recurrence degenerates to
//    y[n] = sum_j b_j x[n-j] + zi[n] (n < N), zf[k] = sum_{m>k} b_m x[T+k-m]
//    + zi[T+k] (T+k < N)),
//  * batched_filter_apply(filter_format="fir") (_batched.py:533-561): the
//    (filter, signal) pairs are channels of one launch,
//  * fftfilt (_fftfilt.py:9-99): the 'same' part of the full convolution,
//    y[n] = sum_j b_j x[n + (K-1)/2 - j], computed directly.
// In general y[n] = sum_{j<K} h[j] x[n + shift - j], x zero outside
// [0, x_len).  With zf requested the outputs n in [T, T+N) go to zf, which
// is exactly the DF2T final state.
//
// Work decomposition: one workgroup of 256 lanes produces a tile of 2048
// consecutive outputs of one channel.  The tile's input window (tile + taps
// of one chunk) is staged in LDS once; each lane owns 8 consecutive outputs
// and walks the taps 8 at a time, reading one aligned 8-sample vector of the
// window from LDS and 8 taps (FP32: a broadcast LDS vector; FP64: through
// the scalar cache into SGPRs) per step, both prefetched a step ahead, and
// doing 64 FMAs on them (a sliding register window).  Filters longer than
// one chunk (1024 taps) restage.
// The kernel is HBM-bound for short filters and VALU-bound from about 10
// FLOP/byte (40 FP32 taps) on.
//
// The tap gradient dtaps[f][j] = sum_c sum_n g[c][n] x[c][n + shift - j]
// (the correlation of the output gradient with the input, summed over the
// channels of filter f) runs as a grid of (filter, part) workgroups that walk
// time tiles of their filter's channels, lanes split into 8-tap groups x time
// slices, with FP64 partial sums per part reduced in a fixed order.

#include <hip/hip_runtime.h>

#include "tsiir.h"

namespace tsiir {
namespace {

constexpr int kFirThreads = 256;
constexpr int kFirR = 8;                        // outputs per lane, taps per step
constexpr int kFirTile = kFirThreads * kFirR;   // outputs per workgroup
constexpr int kFirChunk = 1024;                 // taps per LDS pass

__device__ inline void lds_read8(const float* p, float (&v)[8]) {
  const float4 a = reinterpret_cast<const float4*>(p)[0];
  const float4 b = reinterpret_cast<const float4*>(p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

__device__ inline void lds_read8(const double* p, double (&v)[8]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double2 a = reinterpret_cast<const double2*>(p)[i];
    v[2 * i] = a.x;
    v[2 * i + 1] = a.y;
  }
}

__device__ inline void store8(float* p, const float (&v)[8]) {
  reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
  reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}

__device__ inline void store8(double* p, const double (&v)[8]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    reinterpret_cast<double2*>(p)[i] = make_double2(v[2 * i], v[2 * i + 1]);
}

template <typename T>
struct FirArgs {
  const T* taps;      // (n_filters, K)
  const T* x;
  T* y;
  const T* zi;        // (channels or 1, N) or null
  T* zf;              // (channels, N) or null
  int64_t T_out;      // outputs per channel that go to y
  int64_t x_len;
  int64_t x_rows, x_stride, y_stride, cpf;
  int64_t tiles;      // workgroups per channel
  int32_t K;
  int32_t N;          // state length K - 1
  int32_t shift;
  uint32_t flags;
};

// 64 FMAs of one 8-tap step: acc[r] += h[q] * w[r - q + 7], w = lo ++ hi.
template <typename T>
__device__ __forceinline__ void fir_step(T (&acc)[kFirR], const T (&hq)[kFirR],
                                         const T (&lo)[kFirR], const T (&hi)[kFirR]) {
#pragma unroll
  for (int q = 0; q < kFirR; ++q) {
#pragma unroll
    for (int r = 0; r < kFirR; ++r) {
      const int u = r - q + kFirR - 1;
      acc[r] = fma(hq[q], u < kFirR ? lo[u] : hi[u - kFirR], acc[r]);
    }
  }
}

// Taps jg .. jg + 7 of the (possibly reversed) filter; needs jg + 8 <= K.
// The address is uniform, so these are contiguous scalar loads.
template <typename T, bool REV>
__device__ __forceinline__ void load_taps(const T* __restrict__ h, int K, int jg,
                                          T (&hq)[kFirR]) {
  const T* p = REV ? h + (K - kFirR - jg) : h + jg;
#pragma unroll
  for (int q = 0; q < kFirR; ++q) hq[q] = p[REV ? kFirR - 1 - q : q];
}

// Taps jg .. jg + 7, zero from K on (the last, partial block).
template <typename T, bool REV>
__device__ __forceinline__ void load_taps_tail(const T* __restrict__ h, int K, int jg,
                                               T (&hq)[kFirR]) {
#pragma unroll
  for (int q = 0; q < kFirR; ++q) {
    const int j = jg + q;
    hq[q] = j < K ? h[REV ? K - 1 - j : j] : T(0);
  }
}

// Taps through LDS (broadcast reads) for FP32, through the scalar cache into
// SGPRs for FP64: FP64 window reads already take half of the LDS bandwidth
// at the FMA rate, while FP32 runs faster with both operands from LDS.
template <typename T>
constexpr bool kTapsInLds = sizeof(T) == 4;

template <typename T, bool REV>
__global__ __launch_bounds__(kFirThreads) void fir_kernel(FirArgs<T> a) {
  constexpr int KC = kFirChunk;
  constexpr bool TL = kTapsInLds<T>;
  // xs[kFirR + i] holds x[mbase + i]; the first kFirR slots absorb the
  // window prefetch past the last step.  hs: the chunk's taps, zero-padded.
  __shared__ __align__(16) T xs[kFirR + kFirTile + KC];
  __shared__ __align__(16) T hs[TL ? KC + kFirR : 1];
  const int64_t blk = blockIdx.x;
  const int64_t c = blk / a.tiles;
  const int64_t n0 = (blk - c * a.tiles) * kFirTile;
  const T* __restrict__ h = a.taps + (c / a.cpf) * a.K;
  const T* __restrict__ xr = a.x + (c % a.x_rows) * a.x_stride;
  const int t = threadIdx.x;
  T acc[kFirR];
#pragma unroll
  for (int r = 0; r < kFirR; ++r) acc[r] = T(0);

  for (int jc = 0; jc < a.K; jc += KC) {
    const int kc = min(KC, a.K - jc);
    const int kcp = (kc + kFirR - 1) & ~(kFirR - 1);
    // output n0 + 8t + r with tap jc + jl + q reads x[mbase + 8t + r + kcp - 1
    // - jl - q]
    const int64_t mbase = n0 + a.shift - jc - (kcp - 1);
    if (jc) __syncthreads();
    if (mbase >= 0 && mbase + kFirTile + kcp <= a.x_len) {
      // interior window: all loads in flight before the first LDS store
      T v[kFirR];
#pragma unroll
      for (int k = 0; k < kFirR; ++k) v[k] = xr[mbase + t + k * kFirThreads];
      for (int i = kFirTile + t; i < kFirTile + kcp; i += kFirThreads)
        xs[kFirR + i] = xr[mbase + i];
#pragma unroll
      for (int k = 0; k < kFirR; ++k) xs[kFirR + t + k * kFirThreads] = v[k];
    } else {
      for (int i = t; i < kFirTile + kcp; i += kFirThreads) {
        const int64_t m = mbase + i;
        xs[kFirR + i] = (m >= 0 && m < a.x_len) ? xr[m] : T(0);
      }
    }
    if (TL) {
      for (int i = t; i < kcp + kFirR; i += kFirThreads) {
        const int j = jc + i;
        hs[i] = (i < kc) ? h[REV ? a.K - 1 - j : j] : T(0);
      }
    }
    __syncthreads();
    // Step k uses the window xw - 8k (hi = 8 samples above it, lo = at it)
    // and taps jc + 8k ..; three window buffers and two tap buffers rotate
    // through a 6-step unrolled body, so no register moves are needed, and
    // each step prefetches the next step's window and taps.  From LDS every
    // step is a full (zero-padded) block.  From SGPRs full 8-tap blocks run
    // there (the tap prefetch index is clamped into the filter; what it
    // fetches past the last full block is not used) and a partial last block
    // runs as one more step that reads its window afresh.
    const T* xw = xs + kFirR + t * kFirR + kcp - kFirR;
    const int nsteps = TL ? kcp / kFirR : kc / kFirR;
    const int jmax = a.K - kFirR;   // >= 0 whenever a full block exists
    if (nsteps > 0) {
      T W0[kFirR], W1[kFirR], W2[kFirR], H0[kFirR], H1[kFirR];
      lds_read8(xw + kFirR, W0);
      lds_read8(xw, W1);
      if (TL)
        lds_read8(hs, H0);
      else
        load_taps<T, REV>(h, a.K, jc, H0);
      for (int k = 0; k < nsteps; k += 6) {
#define FIR_STEP(HI, LO, NEXT, HQ, HN, i)                                          \
        if (k + (i) < nsteps) {                                                    \
          lds_read8(xw - kFirR * ((i) + 1), NEXT);                                 \
          if (TL)                                                                  \
            lds_read8(hs + kFirR * (k + (i) + 1), HN);                             \
          else                                                                     \
            load_taps<T, REV>(h, a.K, min(jc + kFirR * (k + (i) + 1), jmax), HN);  \
          fir_step(acc, HQ, LO, HI);                                               \
        }
        FIR_STEP(W0, W1, W2, H0, H1, 0)
        FIR_STEP(W1, W2, W0, H1, H0, 1)
        FIR_STEP(W2, W0, W1, H0, H1, 2)
        FIR_STEP(W0, W1, W2, H1, H0, 3)
        FIR_STEP(W1, W2, W0, H0, H1, 4)
        FIR_STEP(W2, W0, W1, H1, H0, 5)
#undef FIR_STEP
        xw -= 6 * kFirR;
      }
    }
    if (!TL && kc > nsteps * kFirR) {
      const T* xt = xs + kFirR + t * kFirR + kcp - kFirR * (nsteps + 1);
      T lo[kFirR], hi[kFirR], hq[kFirR];
      lds_read8(xt + kFirR, hi);
      lds_read8(xt, lo);
      load_taps_tail<T, REV>(h, a.K, jc + kFirR * nsteps, hq);
      fir_step(acc, hq, lo, hi);
    }
  }

  const int64_t nb = n0 + static_cast<int64_t>(t) * kFirR;
  if ((a.flags & TSIIR_HAS_ZI) && nb < a.N) {
    const T* z = a.zi + ((a.flags & TSIIR_ZI_BROADCAST) ? 0 : c) * a.N;
#pragma unroll
    for (int r = 0; r < kFirR; ++r)
      if (nb + r < a.N) acc[r] += z[nb + r];
  }
  T* yr = a.y + c * a.y_stride;
  if (nb + kFirR <= a.T_out && (reinterpret_cast<uintptr_t>(yr) & 15) == 0) {
    store8(yr + nb, acc);
  } else {
#pragma unroll
    for (int r = 0; r < kFirR; ++r) {
      const int64_t n = nb + r;
      if (n < a.T_out)
        yr[n] = acc[r];
      else if ((a.flags & TSIIR_WRITE_ZF) && n < a.T_out + a.N)
        a.zf[c * a.N + (n - a.T_out)] = acc[r];
    }
  }
}

// ---- tap gradient -----------------------------------------------------------

constexpr int kDtTile = 2048;     // gradient samples per staged tile
constexpr int kDtTaps = 2048;     // taps per pass (256 groups of 8)

template <typename T>
struct DtapsArgs {
  const T* x;
  const T* g;         // (channels, G), row stride g_stride
  double* part;       // (n_filters, parts, K)
  int64_t G;          // gradient samples per channel
  int64_t x_len;
  int64_t x_rows, x_stride, g_stride, cpf;
  int64_t tiles;      // time tiles per channel
  int32_t K;
  int32_t shift;
  int32_t parts;
};

// Workgroup (p, f) sums g[c][n] x[c][n + shift - j] over the items
// (channel of f, time tile) p, p + parts, ... for every tap j of one pass.
// Lane t: tap group tg = t % ng (taps jb + 8tg .. +7), time slice t / ng.
template <typename T>
__global__ __launch_bounds__(kFirThreads) void fir_dtaps_kernel(DtapsArgs<T> a) {
  __shared__ __align__(16) T gs[kDtTile];
  __shared__ __align__(16) T xs[kDtTile + kDtTaps];
  const int p = blockIdx.x;
  const int64_t f = blockIdx.y;
  const int t = threadIdx.x;
  const int64_t items = a.cpf * a.tiles;
  for (int jb = 0; jb < a.K; jb += kDtTaps) {
    const int kp = ((min(kDtTaps, a.K - jb)) + 7) & ~7;
    const int ng = kp / 8;                    // tap groups (<= 256)
    int ns = 1;                               // time slices: power of two
    while (ns * 2 * ng <= kFirThreads) ns *= 2;
    const int L = kDtTile / ns;               // samples per slice (>= 8)
    const int tg = t % ng, sl = t / ng;
    const bool active = sl < ns;
    double sum[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) sum[q] = 0.0;
    for (int64_t it = p; it < items; it += a.parts) {
      const int64_t c = f * a.cpf + it / a.tiles;
      const int64_t n0 = (it % a.tiles) * kDtTile;
      const T* __restrict__ gr = a.g + c * a.g_stride;
      const T* __restrict__ xr = a.x + (c % a.x_rows) * a.x_stride;
      // xs[i] holds x[mbase + i]; (n0 + i_n, tap jb + j) reads
      // xs[i_n + kp - 1 - j].
      const int64_t mbase = n0 + a.shift - jb - (kp - 1);
      __syncthreads();
      for (int i = t; i < kDtTile; i += kFirThreads) {
        const int64_t n = n0 + i;
        gs[i] = n < a.G ? gr[n] : T(0);
      }
      for (int i = t; i < kDtTile + kp; i += kFirThreads) {
        const int64_t m = mbase + i;
        xs[i] = (m >= 0 && m < a.x_len) ? xr[m] : T(0);
      }
      __syncthreads();
      if (active) {
        T s[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] = T(0);
        // n-block at i_n = sl*L + nbk: window xs[B0 .. B0 + 15] with
        // B0 = i_n - 8tg + kp - 8; (r, q) reads u = r - q + 7.
        int B0 = sl * L - 8 * tg + kp - 8;
        T lo[8], hi[8];
        lds_read8(xs + B0, lo);
        for (int nbk = 0; nbk < L; nbk += 8) {
          lds_read8(xs + B0 + 8, hi);
          T gv[8];
          lds_read8(gs + sl * L + nbk, gv);
#pragma unroll
          for (int r = 0; r < 8; ++r) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              const int u = r - q + 7;
              s[q] = fma(gv[r], u < 8 ? lo[u] : hi[u - 8], s[q]);
            }
          }
#pragma unroll
          for (int i = 0; i < 8; ++i) lo[i] = hi[i];
          B0 += 8;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) sum[q] += static_cast<double>(s[q]);
      }
    }
    // reduce the slices of each tap group in slice order
    __syncthreads();
    double* red = reinterpret_cast<double*>(xs);   // ns * kp doubles fit
    if (active) {
#pragma unroll
      for (int q = 0; q < 8; ++q) red[sl * kp + 8 * tg + q] = sum[q];
    }
    __syncthreads();
    for (int j = t; j < kp; j += kFirThreads) {
      double v = 0.0;
      for (int s2 = 0; s2 < ns; ++s2) v += red[s2 * kp + j];
      if (jb + j < a.K) a.part[(f * a.parts + p) * a.K + jb + j] = v;
    }
  }
}

template <typename T>
__global__ void fir_dtaps_reduce_kernel(const double* __restrict__ part, T* __restrict__ dtaps,
                                        int64_t n_filters, int32_t parts, int32_t K) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_filters * K) return;
  const int64_t f = i / K, j = i % K;
  double v = 0.0;
  for (int p = 0; p < parts; ++p) v += part[(f * parts + p) * K + j];
  dtaps[i] = static_cast<T>(v);
}

tsiir_status check_fir(const tsiir_fir_desc* d) {
  if (!d || d->channels < 0 || d->n_samples < 0 || d->x_len < 0) return TSIIR_EINVAL;
  if (d->n_taps < 1 || d->n_taps > (1 << 20)) return TSIIR_EUNSUPPORTED;
  if (d->io_dtype != TSIIR_F32 && d->io_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  if (d->x_rows < 1 || d->chans_per_filter < 1) return TSIIR_EINVAL;
  if (d->x_row_stride < 0 || d->y_row_stride < 0) return TSIIR_EINVAL;
  return TSIIR_OK;
}

int dtaps_parts(const tsiir_fir_desc* d) {
  const int64_t nf = (d->channels + d->chans_per_filter - 1) / d->chans_per_filter;
  const int64_t tiles = (d->n_samples + kDtTile - 1) / kDtTile;
  const int64_t items = d->chans_per_filter * tiles;
  int64_t parts = (4096 + nf - 1) / (nf > 0 ? nf : 1);
  if (parts > items) parts = items;
  if (parts < 1) parts = 1;
  return static_cast<int>(parts);
}

template <typename T>
tsiir_status launch_fir(const tsiir_fir_desc* d, const void* taps, const void* x, void* y,
                        const void* zi, void* zf, hipStream_t st) {
  FirArgs<T> a;
  a.taps = static_cast<const T*>(taps);
  a.x = static_cast<const T*>(x);
  a.y = static_cast<T*>(y);
  a.zi = static_cast<const T*>(zi);
  a.zf = static_cast<T*>(zf);
  a.T_out = d->n_samples;
  a.x_len = d->x_len;
  a.x_rows = d->x_rows;
  a.x_stride = d->x_row_stride;
  a.y_stride = d->y_row_stride;
  a.cpf = d->chans_per_filter;
  a.K = d->n_taps;
  a.N = d->n_taps - 1;
  a.shift = d->shift;
  a.flags = d->flags;
  const int64_t outs = d->n_samples + ((d->flags & TSIIR_WRITE_ZF) ? a.N : 0);
  a.tiles = (outs + kFirTile - 1) / kFirTile;
  const int64_t blocks = d->channels * a.tiles;
  if (blocks == 0) return TSIIR_OK;
  if (blocks > 0x7fffffff) return TSIIR_EUNSUPPORTED;
  if (d->flags & TSIIR_FIR_REVERSE_TAPS)
    hipLaunchKernelGGL((fir_kernel<T, true>), dim3(static_cast<unsigned>(blocks)),
                       dim3(kFirThreads), 0, st, a);
  else
    hipLaunchKernelGGL((fir_kernel<T, false>), dim3(static_cast<unsigned>(blocks)),
                       dim3(kFirThreads), 0, st, a);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

template <typename T>
tsiir_status launch_dtaps(const tsiir_fir_desc* d, const void* x, const void* g, void* dtaps,
                          void* ws, hipStream_t st) {
  DtapsArgs<T> a;
  a.x = static_cast<const T*>(x);
  a.g = static_cast<const T*>(g);
  a.part = static_cast<double*>(ws);
  a.G = d->n_samples;
  a.x_len = d->x_len;
  a.x_rows = d->x_rows;
  a.x_stride = d->x_row_stride;
  a.g_stride = d->y_row_stride;
  a.cpf = d->chans_per_filter;
  a.tiles = (d->n_samples + kDtTile - 1) / kDtTile;
  a.K = d->n_taps;
  a.shift = d->shift;
  a.parts = dtaps_parts(d);
  const int64_t nf = d->channels / d->chans_per_filter;
  if (nf > 65535) return TSIIR_EUNSUPPORTED;
  hipLaunchKernelGGL(fir_dtaps_kernel<T>, dim3(a.parts, static_cast<unsigned>(nf)),
                     dim3(kFirThreads), 0, st, a);
  if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
  const int64_t n = nf * d->n_taps;
  hipLaunchKernelGGL(fir_dtaps_reduce_kernel<T>, dim3(static_cast<unsigned>((n + 255) / 256)),
                     dim3(256), 0, st, a.part, static_cast<T*>(dtaps), nf, a.parts, a.K);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // namespace
}  // namespace tsiir

using namespace tsiir;

extern "C" tsiir_status tsiir_fir_forward(const tsiir_fir_desc* d, const void* taps,
                                          const void* x, void* y, const void* zi, void* zf,
                                          void* stream) {
  const tsiir_status s = check_fir(d);
  if (s != TSIIR_OK) return s;
  const bool state = d->flags & (TSIIR_HAS_ZI | TSIIR_WRITE_ZF);
  if (state && d->shift != 0) return TSIIR_EINVAL;
  if ((d->flags & TSIIR_HAS_ZI) && !zi) return TSIIR_EINVAL;
  if ((d->flags & TSIIR_WRITE_ZF) && d->n_taps > 1 && !zf) return TSIIR_EINVAL;
  if (d->channels == 0) return TSIIR_OK;
  if (!taps || (d->n_samples > 0 && !y) || (d->x_len > 0 && !x)) return TSIIR_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->io_dtype == TSIIR_F64) return launch_fir<double>(d, taps, x, y, zi, zf, st);
  return launch_fir<float>(d, taps, x, y, zi, zf, st);
}

extern "C" size_t tsiir_fir_backward_taps_workspace_bytes(const tsiir_fir_desc* d) {
  if (check_fir(d) != TSIIR_OK || d->channels == 0) return 0;
  const int64_t nf = d->channels / d->chans_per_filter;
  return static_cast<size_t>(nf) * dtaps_parts(d) * d->n_taps * sizeof(double);
}

extern "C" tsiir_status tsiir_fir_backward_taps(const tsiir_fir_desc* d, const void* x,
                                                const void* g, void* dtaps, void* workspace,
                                                size_t workspace_bytes, void* stream) {
  const tsiir_status s = check_fir(d);
  if (s != TSIIR_OK) return s;
  if (d->channels % d->chans_per_filter != 0) return TSIIR_EINVAL;
  if (d->flags & TSIIR_FIR_REVERSE_TAPS) return TSIIR_EINVAL;
  if (d->channels == 0) return TSIIR_OK;
  if (!dtaps || (d->n_samples > 0 && !g) || (d->x_len > 0 && !x)) return TSIIR_EINVAL;
  if (!workspace || workspace_bytes < tsiir_fir_backward_taps_workspace_bytes(d))
    return TSIIR_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->io_dtype == TSIIR_F64) return launch_dtaps<double>(d, x, g, dtaps, workspace, st);
  return launch_dtaps<float>(d, x, g, dtaps, workspace, st);
}
