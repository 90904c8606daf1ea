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
//
// Every quantity that depends on the cutoffs carries its derivatives with
// respect to them (forward-mode differentiation, type D), so one launch can
// also write the Jacobian d sos / d cutoffs that the reference's autograd
// provides through its tensor ops.  The pairing decisions depend on values
// only, as in the reference.

#include <hip/hip_runtime.h>

#include "tsiir.h"

namespace tsiir {
namespace {

constexpr int kMaxDesignOrder = 32;
constexpr int kMaxRoots = 2 * kMaxDesignOrder;

// value and its derivatives with respect to the filter's two cutoffs
struct D {
  double v, d0, d1;
  __device__ D(double x = 0.0) : v(x), d0(0.0), d1(0.0) {}
  __device__ D(double x, double a, double b) : v(x), d0(a), d1(b) {}
};
__device__ inline D operator+(D a, D b) { return D(a.v + b.v, a.d0 + b.d0, a.d1 + b.d1); }
__device__ inline D operator-(D a, D b) { return D(a.v - b.v, a.d0 - b.d0, a.d1 - b.d1); }
__device__ inline D operator-(D a) { return D(-a.v, -a.d0, -a.d1); }
__device__ inline D operator*(D a, D b) {
  return D(a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1);
}
__device__ inline D operator/(D a, D b) {
  const double q = a.v / b.v;
  return D(q, (a.d0 - q * b.d0) / b.v, (a.d1 - q * b.d1) / b.v);
}
__device__ inline D dsqrt(D a) {
  const double r = sqrt(a.v);
  const double h = r > 0.0 ? 0.5 / r : 0.0;
  return D(r, a.d0 * h, a.d1 * h);
}
__device__ inline D dtan(D a) {
  const double t = tan(a.v);
  const double k = 1.0 + t * t;
  return D(t, a.d0 * k, a.d1 * k);
}
// a^e for a > 0 (a == 0 gives 0 with zero derivatives)
__device__ inline D dpow(D a, double e) {
  const double p = pow(a.v, e);
  const double k = a.v != 0.0 ? e * p / a.v : 0.0;
  return D(p, a.d0 * k, a.d1 * k);
}
__device__ inline D dabs(D a) { return a.v < 0.0 ? -a : a; }

struct Cx {
  D r, i;
};

__device__ inline Cx cx(D r, D i = D(0.0)) { return Cx{r, i}; }
__device__ inline Cx operator+(Cx a, Cx b) { return {a.r + b.r, a.i + b.i}; }
__device__ inline Cx operator-(Cx a, Cx b) { return {a.r - b.r, a.i - b.i}; }
__device__ inline Cx operator-(Cx a) { return {-a.r, -a.i}; }
__device__ inline Cx operator*(Cx a, Cx b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ inline Cx operator*(D s, Cx a) { return {s * a.r, s * a.i}; }
__device__ inline Cx operator/(Cx a, Cx b) {
  // Smith's method
  if (fabs(b.r.v) >= fabs(b.i.v)) {
    const D q = b.i / b.r, d = b.r + b.i * q;
    return {(a.r + a.i * q) / d, (a.i - a.r * q) / d};
  }
  const D q = b.r / b.i, d = b.r * q + b.i;
  return {(a.r * q + a.i) / d, (a.i * q - a.r) / d};
}
__device__ inline D cabs(Cx a) {
  const double m = hypot(a.r.v, a.i.v);
  if (m == 0.0) return D(0.0);
  return D(m, (a.r.v * a.r.d0 + a.i.v * a.i.d0) / m, (a.r.v * a.r.d1 + a.i.v * a.i.d1) / m);
}
// principal square root; a zero imaginary part keeps its sign
__device__ inline Cx csqrt(Cx a) {
  if (a.r.v == 0.0 && a.i.v == 0.0) return {D(0.0), a.i};
  const D t = dsqrt((dabs(a.r) + cabs(a)) * D(0.5));
  if (a.r.v >= 0.0) return {t, a.i / (D(2.0) * t)};
  return {dabs(a.i) / (D(2.0) * t), signbit(a.i.v) ? -t : t};
}

struct Zpk {
  Cx z[kMaxRoots];
  Cx p[kMaxRoots];
  int nz, np;
  D k;
};

__device__ Cx prod_neg(const Cx* v, int n) {
  Cx acc = cx(D(1.0));
  for (int i = 0; i < n; ++i) acc = acc * (-v[i]);
  return acc;
}

// analog prototype (independent of the cutoffs)
__device__ void prototype(int family, int N, double param, Zpk& s) {
  const double pi = 3.141592653589793;
  s.nz = 0;
  s.np = N;
  if (family == TSIIR_DESIGN_BUTTERWORTH) {
    for (int k = 0; k < N; ++k) {
      double sn, cs;
      sincos(pi * (2 * k + N + 1) / (2.0 * N), &sn, &cs);
      s.p[k] = cx(D(cs), D(sn));
    }
    s.k = D(1.0);
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
    const Cx p1 = cx(D(-sha * sn), D(cha * cs));
    s.p[k] = family == TSIIR_DESIGN_CHEBYSHEV1 ? p1 : cx(D(1.0)) / p1;
  }
  if (family == TSIIR_DESIGN_CHEBYSHEV1) {
    double g = prod_neg(s.p, N).r.v;
    if (N % 2 == 0) g = g / sqrt(1.0 + eps * eps);
    s.k = D(g);
    return;
  }
  const int mid = N / 2;
  for (int k = 0; k < N; ++k) {
    if (N % 2 == 1 && k == mid) continue;
    s.z[s.nz++] = cx(D(0.0), D(1.0 / cos(pi * (2 * k + 1) / (2.0 * N))));
  }
  const Cx num = prod_neg(s.p, s.np);
  s.k = s.nz > 0 ? (num / prod_neg(s.z, s.nz)).r : num.r;
}

// s -> band transform of roots v[0..n): out = [h + d ..., h - d ...] with
// d = sqrt(h^2 - wc^2) and h = bw v / 2 (bandpass) or bw / (2 v) (bandstop)
__device__ void band_roots(const Cx* v, int n, D bw, D wcsq, bool stop, Cx* out) {
  for (int i = 0; i < n; ++i) {
    const Cx h = stop ? cx(bw) / (D(2.0) * v[i]) : (bw / D(2.0)) * v[i];
    const Cx d = csqrt(h * h - cx(wcsq));
    out[i] = h + d;
    out[n + i] = h - d;
  }
}

__device__ void transform(int band, D w0, D w1, Zpk& s) {
  const int deg = s.np - s.nz;
  if (band == TSIIR_BAND_LOWPASS) {
    for (int i = 0; i < s.nz; ++i) s.z[i] = w0 * s.z[i];
    for (int i = 0; i < s.np; ++i) s.p[i] = w0 * s.p[i];
    s.k = s.k * dpow(w0, static_cast<double>(deg));
    return;
  }
  if (band == TSIIR_BAND_HIGHPASS) {
    const Cx pz = s.nz > 0 ? prod_neg(s.z, s.nz) : cx(D(1.0));
    const Cx pp = prod_neg(s.p, s.np);
    for (int i = 0; i < s.np; ++i) s.p[i] = cx(w0) / s.p[i];
    for (int i = 0; i < s.nz; ++i) s.z[i] = cx(w0) / s.z[i];
    for (int i = 0; i < deg; ++i) s.z[s.nz + i] = cx(D(0.0));
    s.nz += deg;
    s.k = s.k * (pz / pp).r;
    return;
  }
  const D bw = w1 - w0;
  const D wc = dsqrt(w0 * w1);
  const D wcsq = wc * wc;
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
      s.z[nz2 + 2 * i] = cx(D(0.0), wc);
      s.z[nz2 + 2 * i + 1] = cx(D(0.0), -wc);
    }
    s.nz = nz2 + 2 * deg;
  } else {
    for (int i = 0; i < deg; ++i) s.z[nz2 + i] = cx(D(0.0));
    s.nz = nz2 + deg;
    s.k = s.k * dpow(bw, static_cast<double>(deg));
  }
}

__device__ void bilinear(Zpk& s) {
  const Cx fs2 = cx(D(4.0));
  const D q = D(0.25);    // 1 / fs2
  const int deg = s.np - s.nz;
  Cx num = cx(D(1.0)), den = cx(D(1.0));
  for (int i = 0; i < s.nz; ++i) num = num * (fs2 - s.z[i]);
  for (int i = 0; i < s.np; ++i) den = den * (fs2 - s.p[i]);
  for (int i = 0; i < s.np; ++i)
    s.p[i] = (cx(D(1.0)) + q * s.p[i]) / (cx(D(1.0)) - q * s.p[i]);
  for (int i = 0; i < s.nz; ++i)
    s.z[i] = (cx(D(1.0)) + q * s.z[i]) / (cx(D(1.0)) - q * s.z[i]);
  for (int i = 0; i < deg; ++i) s.z[s.nz + i] = cx(D(-1.0));
  s.nz += deg;
  s.k = s.k * (num / den).r;
}

// indices of the real values and of the positive-imaginary complex values,
// in order (the classification uses values only)
__device__ void separate(const Cx* v, int n, unsigned char* re, int& nre, unsigned char* cp,
                         int& ncp) {
  nre = ncp = 0;
  for (int i = 0; i < n; ++i) {
    const double tol = 1e-6 * fabs(v[i].r.v) + 1e-10;
    const bool is_real = fabs(v[i].i.v) < tol || hypot(v[i].r.v, v[i].i.v) < 1e-6;
    if (is_real)
      re[nre++] = static_cast<unsigned char>(i);
    else if (v[i].i.v > 0.0)
      cp[ncp++] = static_cast<unsigned char>(i);
  }
}

template <typename TO>
struct SosOut {
  TO* sos;          // (n_sec, 6)
  double* jac;      // (n_sec, 6, 2) or null
  D g;              // per-section gain |k|^(1/n_sec)
  bool neg;         // k < 0: negate b0 of the first section
  __device__ void row(int i, D b0, D b1, D b2, D a1, D a2) const {
    D v[6] = {b0 * g, b1 * g, b2 * g, D(1.0), a1, a2};
    if (i == 0 && neg) v[0] = -v[0];
    for (int c = 0; c < 6; ++c) {
      sos[i * 6 + c] = static_cast<TO>(v[c].v);
      if (jac) {
        jac[(i * 6 + c) * 2] = v[c].d0;
        jac[(i * 6 + c) * 2 + 1] = v[c].d1;
      }
    }
  }
};

template <typename TO>
__device__ void zpk_to_sos(Zpk& s, TO* sos, double* jac) {
  const int n_sec = (s.np + 1) / 2;
  while (s.nz < s.np) s.z[s.nz++] = cx(D(-1.0));
  unsigned char rp[kMaxRoots], rz[kMaxRoots], cp[kMaxRoots], cz[kMaxRoots];
  int nrp, ncp, nrz, ncz;
  separate(s.p, s.np, rp, nrp, cp, ncp);
  separate(s.z, s.nz, rz, nrz, cz, ncz);
  SosOut<TO> o{sos, jac, dpow(dabs(s.k), 1.0 / n_sec), s.k.v < 0.0};
  int ns = 0, irz = 0, icz = 0, irp = 0;
  for (int ip = 0; ip < ncp && ns < n_sec; ++ip) {
    const Cx p1 = s.p[cp[ip]];
    D b1, b2;
    if (icz < ncz) {
      const Cx z1 = s.z[cz[icz++]];
      b1 = D(-2.0) * z1.r;
      b2 = z1.r * z1.r + z1.i * z1.i;
    } else if (nrz - irz >= 2) {
      const D z1 = s.z[rz[irz]].r, z2 = s.z[rz[irz + 1]].r;
      irz += 2;
      b1 = -(z1 + z2);
      b2 = z1 * z2;
    } else {
      b1 = D(2.0);   // z1 = -1
      b2 = D(1.0);
    }
    o.row(ns++, D(1.0), b1, b2, D(-2.0) * p1.r, p1.r * p1.r + p1.i * p1.i);
  }
  while (nrp - irp >= 2 && ns < n_sec) {
    const D p1 = s.p[rp[irp]].r, p2 = s.p[rp[irp + 1]].r;
    irp += 2;
    D z1 = D(-1.0), z2 = D(-1.0);
    if (nrz - irz >= 2) {
      z1 = s.z[rz[irz]].r;
      z2 = s.z[rz[irz + 1]].r;
      irz += 2;
    }
    o.row(ns++, D(1.0), -(z1 + z2), z1 * z2, -(p1 + p2), p1 * p2);
  }
  if (nrp - irp == 1 && ns < n_sec) {
    const D p1 = s.p[rp[irp]].r;
    const D z1 = nrz - irz >= 1 ? s.z[rz[irz]].r : D(-1.0);
    o.row(ns++, D(1.0), -z1, D(0.0), -p1, D(0.0));
  }
  // sections the pairing did not fill (degenerate inputs only): identity
  while (ns < n_sec) o.row(ns++, D(1.0), D(0.0), D(0.0), D(0.0), D(0.0));
}

template <typename TO>
__global__ __launch_bounds__(64) void iir_design_kernel(tsiir_iir_design_desc d,
                                                        const double* __restrict__ cutoffs,
                                                        TO* __restrict__ sos,
                                                        double* __restrict__ jac) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= d.n_filters) return;
  const bool two = d.band == TSIIR_BAND_BANDPASS || d.band == TSIIR_BAND_BANDSTOP;
  const double pi = 3.141592653589793;
  const D c0(cutoffs[two ? 2 * f : f], 1.0, 0.0);
  const D c1(two ? cutoffs[2 * f + 1] : 0.0, 0.0, 1.0);
  Zpk s;
  prototype(d.family, d.order, d.param, s);
  const D w0 = D(4.0) * dtan(D(pi) * c0 / D(2.0));
  const D w1 = two ? D(4.0) * dtan(D(pi) * c1 / D(2.0)) : D(0.0);
  transform(d.band, w0, w1, s);
  bilinear(s);
  const int n_sec = (s.np + 1) / 2;
  zpk_to_sos(s, sos + f * n_sec * 6, jac ? jac + f * n_sec * 12 : nullptr);
}

tsiir_status check_design(const tsiir_iir_design_desc* d) {
  if (!d || d->n_filters < 0) return TSIIR_EINVAL;
  if (d->order < 1 || d->order > kMaxDesignOrder) return TSIIR_EUNSUPPORTED;
  if (d->family < TSIIR_DESIGN_BUTTERWORTH || d->family > TSIIR_DESIGN_CHEBYSHEV2)
    return TSIIR_EINVAL;
  if (d->band < TSIIR_BAND_LOWPASS || d->band > TSIIR_BAND_BANDSTOP) return TSIIR_EINVAL;
  if (d->family != TSIIR_DESIGN_BUTTERWORTH && !(d->param > 0.0)) return TSIIR_EINVAL;
  if (d->out_dtype != TSIIR_F32 && d->out_dtype != TSIIR_F64) return TSIIR_EUNSUPPORTED;
  return TSIIR_OK;
}

tsiir_status launch_design(const tsiir_iir_design_desc* d, const double* cutoffs, void* sos,
                           double* jac, void* stream) {
  const tsiir_status s = check_design(d);
  if (s != TSIIR_OK) return s;
  if (d->n_filters == 0) return TSIIR_OK;
  if (!cutoffs || !sos) return TSIIR_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>((d->n_filters + 63) / 64));
  if (d->out_dtype == TSIIR_F64)
    hipLaunchKernelGGL(iir_design_kernel<double>, grid, dim3(64), 0, st, *d, cutoffs,
                       static_cast<double*>(sos), jac);
  else
    hipLaunchKernelGGL(iir_design_kernel<float>, grid, dim3(64), 0, st, *d, cutoffs,
                       static_cast<float*>(sos), jac);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
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
  return launch_design(d, cutoffs, sos, nullptr, stream);
}

extern "C" tsiir_status tsiir_iir_design_sos_jacobian(const tsiir_iir_design_desc* d,
                                                      const double* cutoffs, void* sos,
                                                      double* jacobian, void* stream) {
  if (!jacobian && d && d->n_filters > 0) return TSIIR_EINVAL;
  return launch_design(d, cutoffs, sos, jacobian, stream);
}
