// design.hip - batched FP64 IIR design (Butterworth, Chebyshev I / II) into
// second-order sections, one lane per filter.
//
// Replaces the per-filter Python loop of batched_butterworth_design /
// batched_chebyshev_type_1_design / batched_chebyshev_type_2_design
// (_batched.py:22-344), each iteration of which runs the reference's design
// chain on a handful of tensors:
//  * analog prototype: Butterworth poles exp(j pi (2k+N+1) / 2N), gain 1
//    (_butterworth_prototype.py:84-91); Chebyshev I poles -sinh(a) sin(t_k)
//    + j cosh(a) cos(t_k), gain Re prod(-p) (/ sqrt(1 + eps^2) for even N)
//    (_chebyshev_type_1_prototype.py:103-124); Chebyshev II poles 1 / (the
//    type I poles for eps = 1 / sqrt(10^(As/10) - 1)), zeros j / cos(t_k)
//    (the middle k skipped for odd N), gain Re(prod(-p) / prod(-z))
//    (_chebyshev_type_2_prototype.py:108-142);
//  * pre-warp 4 tan(pi f / 2) (_butterworth_design.py:118);
//  * lowpass -> lowpass / highpass / bandpass / bandstop
//    (_lowpass_to_*_zpk.py; the bandstop map keeps the gain unchanged, as
//    _lowpass_to_bandstop_zpk.py:103 does);
//  * bilinear map at fs = 2 (_bilinear_transform_zpk.py:57-87);
//  * zpk -> sos with the reference's separation of real values (|imag| <
//    1e-6 |real| + 1e-10, or |x| < 1e-6) and positive-imaginary members of
//    conjugate pairs in order of appearance, its pairing (complex poles
//    first, with complex zeros, else two real zeros, else -1; then real pole
//    pairs; then a lone real pole) and its gain split |k|^(1/S) over the
//    sections, a negative gain negating b0 of the first section
//    (_zpk_to_sos.py:44-197).
// Everything is computed in FP64 (the reference builds its prototypes at
// torch's default dtype unless dtype is given).  Complex square roots take
// the principal branch with the sign of a zero imaginary part respected, as
// torch.sqrt does.

#include <hip/hip_runtime.h>

#include "tsiir.h"

namespace tsiir {
namespace {

constexpr int kMaxDesignOrder = 32;
constexpr int kMaxRoots = 2 * kMaxDesignOrder;

struct Cx {
  double r, i;
};

__device__ inline Cx cx(double r, double i = 0.0) { return Cx{r, i}; }
__device__ inline Cx operator+(Cx a, Cx b) { return {a.r + b.r, a.i + b.i}; }
__device__ inline Cx operator-(Cx a, Cx b) { return {a.r - b.r, a.i - b.i}; }
__device__ inline Cx operator-(Cx a) { return {-a.r, -a.i}; }
__device__ inline Cx operator*(Cx a, Cx b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ inline Cx operator*(double s, Cx a) { return {s * a.r, s * a.i}; }
__device__ inline Cx operator/(Cx a, Cx b) {
  // Smith's method
  if (fabs(b.r) >= fabs(b.i)) {
    const double q = b.i / b.r, d = b.r + b.i * q;
    return {(a.r + a.i * q) / d, (a.i - a.r * q) / d};
  }
  const double q = b.r / b.i, d = b.r * q + b.i;
  return {(a.r * q + a.i) / d, (a.i * q - a.r) / d};
}
__device__ inline double cabs(Cx a) { return hypot(a.r, a.i); }
// principal square root; a zero imaginary part keeps its sign
__device__ inline Cx csqrt(Cx a) {
  if (a.r == 0.0 && a.i == 0.0) return {0.0, a.i};
  const double t = sqrt((fabs(a.r) + cabs(a)) * 0.5);
  if (a.r >= 0.0) return {t, a.i / (2.0 * t)};
  return {fabs(a.i) / (2.0 * t), copysign(t, a.i)};
}

struct Zpk {
  Cx z[kMaxRoots];
  Cx p[kMaxRoots];
  int nz, np;
  double k;
};

__device__ Cx prod_neg(const Cx* v, int n) {
  Cx acc = cx(1.0);
  for (int i = 0; i < n; ++i) acc = acc * (-v[i]);
  return acc;
}

__device__ void prototype(int family, int N, double param, Zpk& s) {
  const double pi = 3.141592653589793;
  s.nz = 0;
  s.np = N;
  if (family == TSIIR_DESIGN_BUTTERWORTH) {
    for (int k = 0; k < N; ++k) {
      double sn, cs;
      sincos(pi * (2 * k + N + 1) / (2.0 * N), &sn, &cs);
      s.p[k] = cx(cs, sn);
    }
    s.k = 1.0;
    return;
  }
  const double eps = family == TSIIR_DESIGN_CHEBYSHEV1
                         ? sqrt(pow(10.0, param / 10.0) - 1.0)
                         : 1.0 / sqrt(pow(10.0, param / 10.0) - 1.0);
  const double a = asinh(1.0 / eps) / N;
  const double sha = sinh(a), cha = cosh(a);
  for (int k = 0; k < N; ++k) {
    double sn, cs;
    sincos(pi * (2 * k + 1) / (2.0 * N), &sn, &cs);
    const Cx p1 = cx(-sha * sn, cha * cs);
    s.p[k] = family == TSIIR_DESIGN_CHEBYSHEV1 ? p1 : cx(1.0) / p1;
  }
  if (family == TSIIR_DESIGN_CHEBYSHEV1) {
    double g = prod_neg(s.p, N).r;
    if (N % 2 == 0) g = g / sqrt(1.0 + eps * eps);
    s.k = g;
    return;
  }
  const int mid = N / 2;
  for (int k = 0; k < N; ++k) {
    if (N % 2 == 1 && k == mid) continue;
    s.z[s.nz++] = cx(0.0, 1.0 / cos(pi * (2 * k + 1) / (2.0 * N)));
  }
  const Cx num = prod_neg(s.p, s.np);
  s.k = s.nz > 0 ? (num / prod_neg(s.z, s.nz)).r : num.r;
}

// s -> band transform of roots v[0..n): out = [h + d ..., h - d ...] with h
// = num(v) and d = sqrt(h^2 - w0^2); num(v) = bw v / 2 (bandpass) or
// bw / (2 v) (bandstop).
__device__ void band_roots(const Cx* v, int n, double bw, double w0sq, bool stop, Cx* out) {
  for (int i = 0; i < n; ++i) {
    const Cx h = stop ? cx(bw) / (2.0 * v[i]) : (bw / 2.0) * v[i];
    const Cx d = csqrt(h * h - cx(w0sq));
    out[i] = h + d;
    out[n + i] = h - d;
  }
}

__device__ void transform(int band, double w0, double w1, Zpk& s) {
  const int deg = s.np - s.nz;
  if (band == TSIIR_BAND_LOWPASS) {
    for (int i = 0; i < s.nz; ++i) s.z[i] = w0 * s.z[i];
    for (int i = 0; i < s.np; ++i) s.p[i] = w0 * s.p[i];
    s.k *= pow(w0, static_cast<double>(deg));
    return;
  }
  if (band == TSIIR_BAND_HIGHPASS) {
    const Cx pz = s.nz > 0 ? prod_neg(s.z, s.nz) : cx(1.0);
    const Cx pp = prod_neg(s.p, s.np);
    for (int i = 0; i < s.np; ++i) s.p[i] = cx(w0) / s.p[i];
    for (int i = 0; i < s.nz; ++i) s.z[i] = cx(w0) / s.z[i];
    for (int i = 0; i < deg; ++i) s.z[s.nz + i] = cx(0.0);
    s.nz += deg;
    s.k *= (pz / pp).r;
    return;
  }
  const double bw = w1 - w0;
  const double wc = sqrt(w0 * w1);
  const double wcsq = wc * wc;
  const bool stop = band == TSIIR_BAND_BANDSTOP;
  Cx tmp[kMaxRoots];
  band_roots(s.p, s.np, bw, wcsq, stop, tmp);
  for (int i = 0; i < 2 * s.np; ++i) s.p[i] = tmp[i];
  band_roots(s.z, s.nz, bw, wcsq, stop, tmp);
  for (int i = 0; i < 2 * s.nz; ++i) s.z[i] = tmp[i];
  const int nz2 = 2 * s.nz;
  s.np *= 2;
  if (stop) {
    for (int i = 0; i < deg; ++i) {
      s.z[nz2 + 2 * i] = cx(0.0, wc);
      s.z[nz2 + 2 * i + 1] = cx(0.0, -wc);
    }
    s.nz = nz2 + 2 * deg;
  } else {
    for (int i = 0; i < deg; ++i) s.z[nz2 + i] = cx(0.0);
    s.nz = nz2 + deg;
    s.k *= pow(bw, static_cast<double>(deg));
  }
}

__device__ void bilinear(Zpk& s) {
  const double fs2 = 4.0;
  const int deg = s.np - s.nz;
  Cx num = cx(1.0), den = cx(1.0);
  for (int i = 0; i < s.nz; ++i) num = num * (cx(fs2) - s.z[i]);
  for (int i = 0; i < s.np; ++i) den = den * (cx(fs2) - s.p[i]);
  for (int i = 0; i < s.np; ++i)
    s.p[i] = (cx(1.0) + (1.0 / fs2) * s.p[i]) / (cx(1.0) - (1.0 / fs2) * s.p[i]);
  for (int i = 0; i < s.nz; ++i)
    s.z[i] = (cx(1.0) + (1.0 / fs2) * s.z[i]) / (cx(1.0) - (1.0 / fs2) * s.z[i]);
  for (int i = 0; i < deg; ++i) s.z[s.nz + i] = cx(-1.0);
  s.nz += deg;
  s.k *= (num / den).r;
}

// real values (as doubles) and positive-imaginary complex values, in order
__device__ void separate(const Cx* v, int n, double* re, int& nre, Cx* cp, int& ncp) {
  nre = ncp = 0;
  for (int i = 0; i < n; ++i) {
    const double tol = 1e-6 * fabs(v[i].r) + 1e-10;
    const bool is_real = fabs(v[i].i) < tol || cabs(v[i]) < 1e-6;
    if (is_real)
      re[nre++] = v[i].r;
    else if (v[i].i > 0.0)
      cp[ncp++] = v[i];
  }
}

template <typename TO>
__device__ void zpk_to_sos(Zpk& s, TO* out) {
  const int n_sec = (s.np + 1) / 2;
  while (s.nz < s.np) s.z[s.nz++] = cx(-1.0);
  double rp[kMaxRoots], rz[kMaxRoots];
  Cx cp[kMaxRoots], cz[kMaxRoots];
  int nrp, ncp, nrz, ncz;
  separate(s.p, s.np, rp, nrp, cp, ncp);
  separate(s.z, s.nz, rz, nrz, cz, ncz);
  double sos[kMaxDesignOrder][6];
  int ns = 0, irz = 0, icz = 0, irp = 0;
  for (int ip = 0; ip < ncp && ns < kMaxDesignOrder; ++ip) {
    const Cx p1 = cp[ip];
    double b1, b2;
    if (icz < ncz) {
      const Cx z1 = cz[icz++];
      b1 = -2.0 * z1.r;
      b2 = z1.r * z1.r + z1.i * z1.i;
    } else if (nrz - irz >= 2) {
      const double z1 = rz[irz], z2 = rz[irz + 1];
      irz += 2;
      b1 = -(z1 + z2);
      b2 = z1 * z2;
    } else {
      b1 = 2.0;   // z1 = -1
      b2 = 1.0;
    }
    double* r = sos[ns++];
    r[0] = 1.0; r[1] = b1; r[2] = b2;
    r[3] = 1.0; r[4] = -2.0 * p1.r; r[5] = p1.r * p1.r + p1.i * p1.i;
  }
  while (nrp - irp >= 2 && ns < kMaxDesignOrder) {
    const double p1 = rp[irp], p2 = rp[irp + 1];
    irp += 2;
    double z1 = -1.0, z2 = -1.0;
    if (nrz - irz >= 2) {
      z1 = rz[irz];
      z2 = rz[irz + 1];
      irz += 2;
    }
    double* r = sos[ns++];
    r[0] = 1.0; r[1] = -(z1 + z2); r[2] = z1 * z2;
    r[3] = 1.0; r[4] = -(p1 + p2); r[5] = p1 * p2;
  }
  if (nrp - irp == 1 && ns < kMaxDesignOrder) {
    const double p1 = rp[irp];
    const double z1 = nrz - irz >= 1 ? rz[irz] : -1.0;
    double* r = sos[ns++];
    r[0] = 1.0; r[1] = -z1; r[2] = 0.0;
    r[3] = 1.0; r[4] = -p1; r[5] = 0.0;
  }
  const double g = pow(fabs(s.k), 1.0 / n_sec);
  for (int i = 0; i < n_sec; ++i) {
    for (int c = 0; c < 6; ++c) {
      // sections the pairing did not fill (only for degenerate inputs) stay
      // an identity section
      double v = i < ns ? sos[i][c] : (c == 0 || c == 3 ? 1.0 : 0.0);
      if (c < 3) v *= g;
      if (c == 0 && i == 0 && s.k < 0.0) v = -v;
      out[i * 6 + c] = static_cast<TO>(v);
    }
  }
}

template <typename TO>
__global__ __launch_bounds__(64) void iir_design_kernel(tsiir_iir_design_desc d,
                                                        const double* __restrict__ cutoffs,
                                                        TO* __restrict__ sos) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= d.n_filters) return;
  const bool two = d.band == TSIIR_BAND_BANDPASS || d.band == TSIIR_BAND_BANDSTOP;
  const double pi = 3.141592653589793;
  const double c0 = cutoffs[two ? 2 * f : f];
  const double c1 = two ? cutoffs[2 * f + 1] : 0.0;
  Zpk s;
  prototype(d.family, d.order, d.param, s);
  transform(d.band, 4.0 * tan(pi * c0 / 2.0), two ? 4.0 * tan(pi * c1 / 2.0) : 0.0, s);
  bilinear(s);
  const int n_sec = (s.np + 1) / 2;
  zpk_to_sos(s, sos + f * n_sec * 6);
}

}  // namespace
}  // namespace tsiir

using namespace tsiir;

extern "C" int32_t tsiir_iir_design_sections(const tsiir_iir_design_desc* d) {
  if (!d || d->order < 1 || d->order > kMaxDesignOrder) return 0;
  const bool two = d->band == TSIIR_BAND_BANDPASS || d->band == TSIIR_BAND_BANDSTOP;
  return two ? d->order : (d->order + 1) / 2;
}

extern "C" tsiir_status tsiir_iir_design_sos(const tsiir_iir_design_desc* d,
                                             const double* cutoffs, void* sos, void* stream) {
  if (!d || d->n_filters < 0) return TSIIR_EINVAL;
  if (d->order < 1 || d->order > kMaxDesignOrder) return TSIIR_EUNSUPPORTED;
  if (d->family < TSIIR_DESIGN_BUTTERWORTH || d->family > TSIIR_DESIGN_CHEBYSHEV2)
    return TSIIR_EINVAL;
  if (d->band < TSIIR_BAND_LOWPASS || d->band > TSIIR_BAND_BANDSTOP) return TSIIR_EINVAL;
  if (d->family != TSIIR_DESIGN_BUTTERWORTH && !(d->param > 0.0)) return TSIIR_EINVAL;
  if (d->out_dtype != TSIIR_F32 && d->out_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  if (d->n_filters == 0) return TSIIR_OK;
  if (!cutoffs || !sos) return TSIIR_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>((d->n_filters + 63) / 64));
  if (d->out_dtype == TSIIR_F64)
    hipLaunchKernelGGL(iir_design_kernel<double>, grid, dim3(64), 0, st, *d, cutoffs,
                       static_cast<double*>(sos));
  else
    hipLaunchKernelGGL(iir_design_kernel<float>, grid, dim3(64), 0, st, *d, cutoffs,
                       static_cast<float*>(sos));
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}
