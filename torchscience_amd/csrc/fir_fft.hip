// fir_fft.hip - FIR filtering by FFT (overlap-save) for long filters, one filter
// per group of chans_per_filter consecutive channels.
//
// Computes the same y[n] = sum_{j<K} h[j] x[n + shift - j] (x zero outside
// [0, x_len)) as the direct kernel, the way fftfilt (_fftfilt.py:60-92)
// does: through a discrete Fourier transform, here block by block
// (overlap-save) with a 4096-point transform held in LDS, so that the cost
// per output no longer grows with K.
//
// Work decomposition: one workgroup of 256 lanes filters TWO consecutive
// blocks of L = 4096 - K + 1 outputs of one channel at once: the first block
// is the real part and the second the imaginary part of one complex input
// window (a channel with an odd block count leaves its last block alone,
// with a zero imaginary part).  The taps are real, so the circular
// convolution keeps the two parts apart and no post-processing of a real
// transform is needed.  Both parts belong to one channel, so rounding
// shared through the transform, and a NaN or Inf in the input, never reach
// another channel (they reach at most the two blocks of the window).
//
// The transform: six radix-4 stages, decimation in frequency (natural order
// in, base-4 digit-reversed order out), fused in pairs into three radix-16
// passes (16 points per lane in registers, one LDS round trip per pass).
// The transform of the taps is made by the same passes, so it sits in the
// same permuted order and the pointwise product needs no reordering.  The
// inverse runs the mirrored stages (twiddle by the conjugate, then the
// inverse butterfly) in reverse order, which undoes each forward stage
// exactly; 1/4096 is folded into the taps' transform.  The last forward
// pass, the product and the first inverse pass share one register set.
// LDS holds one padding point after every 16 so that the lanes of the
// 16-point-per-lane pass fall on distinct banks.
//
// Twiddles: a table of e^{-2 pi i k / 4096} for k < 1024, formed in FP64
// with sincospi and rounded once to the I/O type; a butterfly loads w^j and
// forms w^2j and w^3j by complex products (one load instead of three).

#include <hip/hip_runtime.h>

#include "tsiir.h"

namespace tsiir {
namespace {

constexpr int kFftN = 4096;
constexpr int kFftThreads = 256;
constexpr int kFftTw = kFftN / 4;
constexpr int kFftLds = kFftN + kFftN / 16;
constexpr int kFftMaxTaps = kFftN / 2;

template <typename T>
struct CxOf;
template <>
struct CxOf<float> {
  using type = float2;
};
template <>
struct CxOf<double> {
  using type = double2;
};

__device__ inline int pad(int i) { return i + (i >> 4); }

template <typename C>
__device__ inline C cmul(C a, C w) {
  return C{a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x};
}

// a times the conjugate of w
template <typename C>
__device__ inline C cmulc(C a, C w) {
  return C{a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y};
}

// 4-point DFT in place: forward uses e^{-2 pi i / 4} = -i, the inverse +i.
template <typename C, bool INV>
__device__ inline void dft4(C& a0, C& a1, C& a2, C& a3) {
  const C t0{a0.x + a2.x, a0.y + a2.y};
  const C t1{a0.x - a2.x, a0.y - a2.y};
  const C t2{a1.x + a3.x, a1.y + a3.y};
  const C t3{a1.x - a3.x, a1.y - a3.y};
  a0 = C{t0.x + t2.x, t0.y + t2.y};
  a2 = C{t0.x - t2.x, t0.y - t2.y};
  if (INV) {
    a1 = C{t1.x - t3.y, t1.y + t3.x};
    a3 = C{t1.x + t3.y, t1.y - t3.x};
  } else {
    a1 = C{t1.x + t3.y, t1.y - t3.x};
    a3 = C{t1.x - t3.y, t1.y + t3.x};
  }
}

// The 16 points of lane t in the pass whose first stage has span Q (second
// stage span Q / 4): G * 4Q + j0 + m1 * Q + m2 * Q / 4 with j0 = t mod Q/4,
// G = t div Q/4.
template <int Q>
__device__ inline int point(int t, int m1, int m2) {
  constexpr int Q4 = Q / 4;
  return (t / Q4) * 4 * Q + (t % Q4) + m1 * Q + m2 * Q4;
}

template <int Q, typename C>
__device__ inline void lds_load(const C* lds, int t, C (&v)[4][4]) {
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1)
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) v[m1][m2] = lds[pad(point<Q>(t, m1, m2))];
}

template <int Q, typename C>
__device__ inline void lds_store(C* lds, int t, const C (&v)[4][4]) {
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1)
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) lds[pad(point<Q>(t, m1, m2))] = v[m1][m2];
}

// w^1, w^2, w^3 from w = tw[k]
template <typename C>
__device__ inline void powers(const C* __restrict__ tw, int k, C (&w)[4]) {
  w[1] = tw[k];
  w[2] = cmul(w[1], w[1]);
  w[3] = cmul(w[1], w[2]);
}

// Forward stages of span Q then Q / 4 on the lane's 16 points.  Stage of
// span q: butterfly over the 4 points j + m q of a group, then point m times
// e^{-2 pi i m j / (4q)} = (tw[j N / (4q)])^m.
template <int Q, typename C>
__device__ inline void fwd16(C (&v)[4][4], const C* __restrict__ tw, int t) {
  constexpr int Q4 = Q / 4;
  const int j0 = t % Q4;
  C w[4];
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) {
    dft4<C, false>(v[0][m2], v[1][m2], v[2][m2], v[3][m2]);
    powers(tw, (j0 + m2 * Q4) * (kFftN / (4 * Q)), w);
#pragma unroll
    for (int m1 = 1; m1 < 4; ++m1) v[m1][m2] = cmul(v[m1][m2], w[m1]);
  }
  if (Q4 > 1) powers(tw, j0 * (kFftN / Q), w);
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1) {
    dft4<C, false>(v[m1][0], v[m1][1], v[m1][2], v[m1][3]);
    if (Q4 > 1) {
#pragma unroll
      for (int m2 = 1; m2 < 4; ++m2) v[m1][m2] = cmul(v[m1][m2], w[m2]);
    }
  }
}

// Exact mirror of fwd16 (times 16): undo the span-Q/4 stage, then span Q.
template <int Q, typename C>
__device__ inline void inv16(C (&v)[4][4], const C* __restrict__ tw, int t) {
  constexpr int Q4 = Q / 4;
  const int j0 = t % Q4;
  C w[4];
  if (Q4 > 1) powers(tw, j0 * (kFftN / Q), w);
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1) {
    if (Q4 > 1) {
#pragma unroll
      for (int m2 = 1; m2 < 4; ++m2) v[m1][m2] = cmulc(v[m1][m2], w[m2]);
    }
    dft4<C, true>(v[m1][0], v[m1][1], v[m1][2], v[m1][3]);
  }
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) {
    powers(tw, (j0 + m2 * Q4) * (kFftN / (4 * Q)), w);
#pragma unroll
    for (int m1 = 1; m1 < 4; ++m1) v[m1][m2] = cmulc(v[m1][m2], w[m1]);
    dft4<C, true>(v[0][m2], v[1][m2], v[2][m2], v[3][m2]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void fft_twiddle_kernel(typename CxOf<T>::type* tw) {
  using C = typename CxOf<T>::type;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kFftTw) return;
  double s, c;
  sincospi(2.0 * k / kFftN, &s, &c);
  tw[k] = C{static_cast<T>(c), static_cast<T>(-s)};
}

// Transform of the taps (zero-padded to N, reversed with REV, scaled by
// 1 / N), left in the forward passes' permuted order and stored lane-major:
// H[k * 256 + t] is point k = m1 * 4 + m2 of lane t in the last forward
// pass, so the filtering kernel's loads of it are coalesced.  One workgroup
// per filter (blockIdx.x).
template <typename T>
__global__ __launch_bounds__(256) void fft_taps_kernel(const T* __restrict__ taps, int K,
                                                       int rev,
                                                       const typename CxOf<T>::type* __restrict__ tw,
                                                       typename CxOf<T>::type* __restrict__ H) {
  using C = typename CxOf<T>::type;
  __shared__ C lds[kFftLds];
  const int t = threadIdx.x;
  taps += static_cast<int64_t>(blockIdx.x) * K;
  H += static_cast<int64_t>(blockIdx.x) * kFftN;
  const T scale = static_cast<T>(1.0 / kFftN);
  for (int r = 0; r < kFftN / kFftThreads; ++r) {
    const int i = t + r * kFftThreads;
    const T h = i < K ? taps[rev ? K - 1 - i : i] * scale : T(0);
    lds[pad(i)] = C{h, T(0)};
  }
  __syncthreads();
  C v[4][4];
  lds_load<1024>(lds, t, v);
  fwd16<1024>(v, tw, t);
  lds_store<1024>(lds, t, v);
  __syncthreads();
  lds_load<64>(lds, t, v);
  fwd16<64>(v, tw, t);
  lds_store<64>(lds, t, v);
  __syncthreads();
  lds_load<4>(lds, t, v);
  fwd16<4>(v, tw, t);
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1)
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) H[(m1 * 4 + m2) * kFftThreads + t] = v[m1][m2];
}

template <typename T>
struct FftArgs {
  const T* x;
  T* y;
  const typename CxOf<T>::type* tw;
  const typename CxOf<T>::type* H;
  int64_t channels;
  int64_t T_out;
  int64_t x_len;
  int64_t x_rows;
  int64_t x_stride;
  int64_t y_stride;
  int64_t shift;
  int64_t blocks;  // output blocks per channel
  int64_t pairs;   // block pairs per channel, ceil(blocks / 2)
  int64_t cpf;     // channels per filter (<= channels)
  int K;
  int L;           // outputs per block, N - K + 1
};

template <typename T>
__global__ __launch_bounds__(256) void fir_fft_kernel(FftArgs<T> a) {
  using C = typename CxOf<T>::type;
  __shared__ C lds[kFftLds];
  const int t = threadIdx.x;
  const int64_t c0 = blockIdx.x / a.pairs;         // channel
  const int64_t bp = blockIdx.x - c0 * a.pairs;     // its block pair
  const int64_t f = c0 / a.cpf;
  const bool has1 = 2 * bp + 1 < a.blocks;
  const C* __restrict__ H = a.H + f * kFftN;
  const T* __restrict__ x0 = a.x + (c0 % a.x_rows) * a.x_stride;
  const int64_t n0 = 2 * bp * a.L;                  // first output of the real block
  const int64_t src0 = n0 + a.shift - (a.K - 1);
  // the first pass's points t + m1 * 1024 + m2 * 256 are coalesced in
  // memory: load them straight into registers
  C v[4][4];
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1)
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) {
      const int64_t s = src0 + point<1024>(t, m1, m2);
      const int64_t s1 = s + a.L;                 // the same point of the next block
      const bool in = s >= 0 && s < a.x_len;
      const bool in1 = has1 && s1 >= 0 && s1 < a.x_len;
      v[m1][m2] = C{in ? x0[s] : T(0), in1 ? x0[s1] : T(0)};
    }
  fwd16<1024>(v, a.tw, t);
  lds_store<1024>(lds, t, v);
  __syncthreads();
  lds_load<64>(lds, t, v);
  fwd16<64>(v, a.tw, t);
  lds_store<64>(lds, t, v);
  __syncthreads();
  lds_load<4>(lds, t, v);
  fwd16<4>(v, a.tw, t);
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1)
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2)
      v[m1][m2] = cmul(v[m1][m2], H[(m1 * 4 + m2) * kFftThreads + t]);
  inv16<4>(v, a.tw, t);
  lds_store<4>(lds, t, v);
  __syncthreads();
  lds_load<64>(lds, t, v);
  inv16<64>(v, a.tw, t);
  lds_store<64>(lds, t, v);
  __syncthreads();
  lds_load<1024>(lds, t, v);
  inv16<1024>(v, a.tw, t);
  // circular outputs i >= K - 1 are the linear ones: y[n0 + i - (K - 1)]
  T* __restrict__ y0 = a.y + c0 * a.y_stride;
#pragma unroll
  for (int m1 = 0; m1 < 4; ++m1)
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) {
      const int i = point<1024>(t, m1, m2);
      const int64_t n = n0 + i - (a.K - 1);
      if (i >= a.K - 1) {
        if (n < a.T_out) y0[n] = v[m1][m2].x;
        if (has1 && n + a.L < a.T_out) y0[n + a.L] = v[m1][m2].y;
      }
    }
}

tsiir_status check_fft(const tsiir_fir_desc* d) {
  if (!d || d->channels < 0 || d->n_samples < 0 || d->x_len < 0) return TSIIR_EINVAL;
  if (d->io_dtype != TSIIR_F32 && d->io_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  if (d->x_rows < 1 || d->chans_per_filter < 1) return TSIIR_EINVAL;
  if (d->x_row_stride < 0 || d->y_row_stride < 0) return TSIIR_EINVAL;
  if (d->flags & (TSIIR_HAS_ZI | TSIIR_WRITE_ZF)) return TSIIR_EINVAL;
  if (d->chans_per_filter < d->channels && d->channels % d->chans_per_filter != 0)
    return TSIIR_EINVAL;
  if (d->n_taps < 1 || d->n_taps > kFftMaxTaps) return TSIIR_EUNSUPPORTED;
  return TSIIR_OK;
}

int64_t fft_filters(const tsiir_fir_desc* d) {
  if (d->channels == 0) return 1;
  return d->chans_per_filter >= d->channels ? 1 : d->channels / d->chans_per_filter;
}

size_t fft_ws_bytes(const tsiir_fir_desc* d) {
  const size_t c = d->io_dtype == TSIIR_F64 ? sizeof(double2) : sizeof(float2);
  return (kFftTw + static_cast<size_t>(fft_filters(d)) * kFftN) * c;
}

template <typename T>
tsiir_status launch_fft(const tsiir_fir_desc* d, const void* taps, const void* x, void* y,
                        void* ws, hipStream_t st) {
  using C = typename CxOf<T>::type;
  C* tw = static_cast<C*>(ws);
  C* H = tw + kFftTw;
  FftArgs<T> a;
  a.x = static_cast<const T*>(x);
  a.y = static_cast<T*>(y);
  a.tw = tw;
  a.H = H;
  a.channels = d->channels;
  a.T_out = d->n_samples;
  a.x_len = d->x_len;
  a.x_rows = d->x_rows;
  a.x_stride = d->x_row_stride;
  a.y_stride = d->y_row_stride;
  a.shift = d->shift;
  a.K = d->n_taps;
  a.L = kFftN - d->n_taps + 1;
  a.blocks = (d->n_samples + a.L - 1) / a.L;
  a.pairs = (a.blocks + 1) / 2;
  const int64_t nf = fft_filters(d);
  a.cpf = d->channels / nf;
  const int64_t grid = d->channels * a.pairs;
  if (grid > 0x7fffffff || nf > 0x7fffffff) return TSIIR_EUNSUPPORTED;
  hipLaunchKernelGGL(fft_twiddle_kernel<T>, dim3((kFftTw + 255) / 256), dim3(256), 0, st, tw);
  if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
  const int rev = (d->flags & TSIIR_FIR_REVERSE_TAPS) ? 1 : 0;
  hipLaunchKernelGGL(fft_taps_kernel<T>, dim3(static_cast<unsigned>(nf)), dim3(kFftThreads), 0,
                     st,
                     static_cast<const T*>(taps), d->n_taps, rev, tw, H);
  if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
  hipLaunchKernelGGL(fir_fft_kernel<T>, dim3(static_cast<unsigned>(grid)), dim3(kFftThreads), 0,
                     st, a);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // namespace
}  // namespace tsiir

using namespace tsiir;

extern "C" size_t tsiir_fir_fft_workspace_bytes(const tsiir_fir_desc* d) {
  if (check_fft(d) != TSIIR_OK) return 0;
  return fft_ws_bytes(d);
}

extern "C" tsiir_status tsiir_fir_forward_fft(const tsiir_fir_desc* d, const void* taps,
                                              const void* x, void* y, void* workspace,
                                              size_t workspace_bytes, void* stream) {
  const tsiir_status s = check_fft(d);
  if (s != TSIIR_OK) return s;
  if (d->channels == 0 || d->n_samples == 0) return TSIIR_OK;
  if (!taps || !y || (d->x_len > 0 && !x)) return TSIIR_EINVAL;
  if (!workspace || workspace_bytes < fft_ws_bytes(d)) return TSIIR_EWORKSPACE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->io_dtype == TSIIR_F64) return launch_fft<double>(d, taps, x, y, workspace, st);
  return launch_fft<float>(d, taps, x, y, workspace, st);
}
