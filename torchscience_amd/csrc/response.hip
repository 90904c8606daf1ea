// response.hip - frequency response of cascades of second-order sections.
//
// H(e^{jw}) = prod_s (b0 + b1 z^-1 + b2 z^-2) / (a0 + a1 z^-1 + a2 z^-2) at
// z = e^{jw}, in FP64 complex arithmetic, for every (filter, frequency)
// pair of a batch: one lane per pair, z^-1 = cos w - j sin w and z^-2 from
// sincos(2w) (not by squaring).  Replaces the per-section loop of
// frequency_response_sos (_frequency_response_sos.py:128-137), which
// evaluates the same product with complex torch ops; a0 is used as given
// (that function only validates it).
//
// Also B(z^-1) / A(z^-1) for direct-form coefficients, both polynomials by
// Horner's rule in z^-1 from the highest power down (the evaluation order of
// _frequency_response.py:121-147), for frequency_response and
// frequency_response_fir (_frequency_response.py:104-116,
// _frequency_response_fir.py:66-79).

#include <hip/hip_runtime.h>

#include "tsiir.h"

namespace tsiir {
namespace {

template <typename TC>
__global__ __launch_bounds__(256) void sos_freq_kernel(const TC* __restrict__ sos,
                                                       int64_t n_filters, int n_sections,
                                                       const double* __restrict__ w,
                                                       int64_t n_freq, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_filters * n_freq) return;
  const int64_t f = i / n_freq, k = i - f * n_freq;
  double s1, c1, s2, c2;
  sincos(w[k], &s1, &c1);
  sincos(2.0 * w[k], &s2, &c2);
  // z^-1 = c1 - j s1, z^-2 = c2 - j s2
  double hr = 1.0, hi = 0.0;
  const TC* p = sos + f * n_sections * 6;
  for (int s = 0; s < n_sections; ++s, p += 6) {
    const double b0 = p[0], b1 = p[1], b2 = p[2], a0 = p[3], a1 = p[4], a2 = p[5];
    const double nr = b0 + b1 * c1 + b2 * c2, ni = -(b1 * s1 + b2 * s2);
    const double dr = a0 + a1 * c1 + a2 * c2, di = -(a1 * s1 + a2 * s2);
    // q = n / d
    const double dd = dr * dr + di * di;
    const double qr = (nr * dr + ni * di) / dd, qi = (ni * dr - nr * di) / dd;
    const double tr = hr * qr - hi * qi;
    hi = hr * qi + hi * qr;
    hr = tr;
  }
  out[2 * i] = hr;
  out[2 * i + 1] = hi;
}

// An empty polynomial evaluates to 1, as in _frequency_response.py:139-140.
__device__ inline void horner(const double* __restrict__ c, int n, double zr, double zi,
                              double& pr, double& pi) {
  pr = n == 0 ? 1.0 : 0.0;
  pi = 0.0;
  for (int k = n - 1; k >= 0; --k) {
    const double tr = pr * zr - pi * zi + c[k];
    pi = pr * zi + pi * zr;
    pr = tr;
  }
}

__global__ __launch_bounds__(256) void ba_freq_kernel(const double* __restrict__ b, int nb,
                                                      const double* __restrict__ a, int na,
                                                      int64_t n_filters,
                                                      const double* __restrict__ w,
                                                      int64_t n_freq, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_filters * n_freq) return;
  const int64_t f = i / n_freq, k = i - f * n_freq;
  double s, c;
  sincos(w[k], &s, &c);
  // z^-1 = c - j s
  double nr, ni, dr, di;
  horner(b + f * nb, nb, c, -s, nr, ni);
  horner(a + f * na, na, c, -s, dr, di);
  const double dd = dr * dr + di * di;
  out[2 * i] = (nr * dr + ni * di) / dd;
  out[2 * i + 1] = (ni * dr - nr * di) / dd;
}

}  // namespace
}  // namespace tsiir

using namespace tsiir;

extern "C" tsiir_status tsiir_sos_frequency_response(const void* sos, int32_t coef_dtype,
                                                     int64_t n_filters, int32_t n_sections,
                                                     const double* w, int64_t n_freq,
                                                     double* out, void* stream) {
  if (n_filters < 0 || n_sections < 0 || n_freq < 0) return TSIIR_EINVAL;
  if (coef_dtype != TSIIR_F32 && coef_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  const int64_t n = n_filters * n_freq;
  if (n == 0) return TSIIR_OK;
  if (!w || !out || (n_sections > 0 && !sos)) return TSIIR_EINVAL;
  if ((n + 255) / 256 > 0x7fffffff) return TSIIR_EUNSUPPORTED;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>((n + 255) / 256));
  if (coef_dtype == TSIIR_F64)
    hipLaunchKernelGGL(sos_freq_kernel<double>, grid, dim3(256), 0, st,
                       static_cast<const double*>(sos), n_filters, n_sections, w, n_freq, out);
  else
    hipLaunchKernelGGL(sos_freq_kernel<float>, grid, dim3(256), 0, st,
                       static_cast<const float*>(sos), n_filters, n_sections, w, n_freq, out);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

extern "C" tsiir_status tsiir_ba_frequency_response(const double* b, int32_t nb,
                                                    const double* a, int32_t na,
                                                    int64_t n_filters, const double* w,
                                                    int64_t n_freq, double* out,
                                                    void* stream) {
  if (n_filters < 0 || n_freq < 0 || nb < 0 || na < 0) return TSIIR_EINVAL;
  const int64_t n = n_filters * n_freq;
  if (n == 0) return TSIIR_OK;
  if (!w || !out || (nb > 0 && !b) || (na > 0 && !a)) return TSIIR_EINVAL;
  if ((n + 255) / 256 > 0x7fffffff) return TSIIR_EUNSUPPORTED;
  hipLaunchKernelGGL(ba_freq_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), b, nb, a, na, n_filters, w, n_freq, out);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}
