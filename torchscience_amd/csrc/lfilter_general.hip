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
// every sample.

#include "sos_common.h"
#include "tsiir.h"

namespace tsiir {

// samples per lane per block: fewer for deep states, whose unrolled
// recurrence would not fit the registers
template <int NMAX>
constexpr int kGenBlk = NMAX <= 8 ? 16 : (NMAX <= 16 ? 8 : 1);

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

template <typename TACC, int NMAX>
__global__ __launch_bounds__(64) void lfilter_general_kernel(
    const TACC* __restrict__ coef, int N, const TACC* __restrict__ x, TACC* __restrict__ y,
    int64_t B, int64_t T, int64_t x_rows, int64_t x_stride, int64_t y_stride,
    const TACC* __restrict__ zi, TACC* __restrict__ zf, uint32_t flags) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= B) return;
  TACC bn[NMAX + 1], an[NMAX + 1], z[NMAX];
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
  const int64_t zrow = (flags & TSIIR_ZI_BROADCAST) ? 0 : c;
#pragma unroll
  for (int k = 0; k < NMAX; ++k)
    z[k] = (k < N && (flags & TSIIR_HAS_ZI)) ? zi[zrow * N + k] : TACC(0);
  const TACC* xr = x + (c % x_rows) * x_stride;
  TACC* yr = y + c * y_stride;
  // Whole blocks of BLK samples: the next block's loads are in flight while
  // the current block runs through the recurrence.
  constexpr int BLK = kGenBlk<NMAX>;
  const int64_t Tb = BLK > 1 ? T / BLK * BLK : 0;
  if constexpr (BLK > 1) {
    TACC cur[BLK], nxt[BLK];
    if (Tb > 0) {
#pragma unroll
      for (int i = 0; i < BLK; ++i) cur[i] = xr[i];
    }
    for (int64_t n0 = 0; n0 < Tb; n0 += BLK) {
      if (n0 + BLK < Tb) {
#pragma unroll
        for (int i = 0; i < BLK; ++i) nxt[i] = xr[n0 + BLK + i];
      }
#pragma unroll
      for (int i = 0; i < BLK; ++i) {
        TACC yn;
        df2t_step<TACC, NMAX>(bn, an, N, z, cur[i], yn);
        yr[n0 + i] = yn;
      }
#pragma unroll
      for (int i = 0; i < BLK; ++i) cur[i] = nxt[i];
    }
  }
  // the remaining samples one at a time (the same update as df2t_step,
  // written out: through the helper the order-64 form spills)
  for (int64_t n = Tb; n < T; ++n) {
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
  if (flags & TSIIR_WRITE_ZF) {
#pragma unroll
    for (int k = 0; k < NMAX; ++k)
      if (k < N) zf[c * N + k] = z[k];
  }
}

template <typename TACC>
static tsiir_status launch_general(const tsiir_sosfilt_desc* d, const void* coef, const void* x,
                                   void* y, const void* zi, void* zf, hipStream_t st) {
  const int N = d->n_sections;
  const unsigned blocks = static_cast<unsigned>((d->channels + 63) / 64);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(64), 0, st, static_cast<const TACC*>(coef), N,
                       static_cast<const TACC*>(x), static_cast<TACC*>(y), d->channels,
                       d->n_samples, d->x_rows, d->x_row_stride, d->y_row_stride,
                       static_cast<const TACC*>(zi), static_cast<TACC*>(zf), d->flags);
  };
  if (N <= 4)
    go(lfilter_general_kernel<TACC, 4>);
  else if (N <= 8)
    go(lfilter_general_kernel<TACC, 8>);
  else if (N <= 16)
    go(lfilter_general_kernel<TACC, 16>);
  else
    go(lfilter_general_kernel<TACC, 64>);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // namespace tsiir

using namespace tsiir;

extern "C" tsiir_status tsiir_lfilter_forward(const tsiir_sosfilt_desc* d, const void* coef,
                                              const void* x, void* y, const void* zi, void* zf,
                                              void* stream) {
  if (!d || d->channels < 0 || d->n_samples < 0) return TSIIR_EINVAL;
  if (d->n_sections < 1 || d->n_sections > 64) return TSIIR_EUNSUPPORTED;
  if (d->io_dtype != d->coef_dtype) return TSIIR_EINVAL;
  if (d->io_dtype != TSIIR_F32 && d->io_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  if (d->x_rows < 1) return TSIIR_EINVAL;
  if ((d->flags & TSIIR_HAS_ZI) && !zi) return TSIIR_EINVAL;
  if ((d->flags & TSIIR_WRITE_ZF) && !zf) return TSIIR_EINVAL;
  if (d->channels == 0) return TSIIR_OK;
  if (!coef || (d->n_samples > 0 && (!x || !y))) return TSIIR_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d->io_dtype == TSIIR_F64) return launch_general<double>(d, coef, x, y, zi, zf, st);
  return launch_general<float>(d, coef, x, y, zi, zf, st);
}
