/*
 * oracle.c - FP64 CPU restatement of torchscience's IIR filtering path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU kernels are
 * compared against (and the CPU baseline bench.py times).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never calls it.
 *
 * What it restates (reference paths relative to the reference checkout):
 *   - the DF2T per-sample update of lfilter,
 *       src/torchscience/signal_processing/filter/_lfilter.py:153-166:
 *         y_n = b[0]*x_n + z[0]
 *         z[k] = b[k+1]*x_n - a[k+1]*y_n + z[k+1]      (k < N-1)
 *         z[N-1] = b[N]*x_n - a[N]*y_n
 *     evaluated left to right with one rounding per operation, as the
 *     separate ATen ops do (build with -ffp-contract=off);
 *   - the normalisation "if a[0] != 1: b /= a[0]; a /= a[0]" (_lfilter.py:96-98)
 *     and the zero padding of b / a to a common length (_lfilter.py:114-117);
 *   - the section cascade of sosfilt, section 0 first, each section fed the
 *     previous section's output (_sosfilt.py:117-128).
 * Axis moves, broadcasting and dtype promotion live in oracle.py.
 *
 * Parity pins (tests/test_oracle.py): scipy.signal.sosfilt / lfilter golden
 * vectors in tests/golden, the biquad closed-form impulse response, FIR
 * (a = [1]) against numpy.convolve, the Toeplitz/Markov brute force on tiny
 * inputs, zf -> zi continuation and the step-input steady state of
 * sosfilt_zi.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* One general-order DF2T run over a single channel.  b, a already
 * normalised and padded to n = order + 1 coefficients.  z (length order) is
 * the state, updated in place. */
static void df2t_channel(const double* b, const double* a, int order,
                         const double* x, double* y, int64_t T, double* z) {
  if (order == 0) {
    for (int64_t n = 0; n < T; ++n) y[n] = b[0] * x[n];
    return;
  }
  for (int64_t n = 0; n < T; ++n) {
    const double xn = x[n];
    const double t0 = b[0] * xn;
    const double yn = t0 + z[0];
    for (int k = 0; k < order - 1; ++k) {
      const double p = b[k + 1] * xn;
      const double q = a[k + 1] * yn;
      const double d = p - q;
      z[k] = d + z[k + 1];
    }
    {
      const double p = b[order] * xn;
      const double q = a[order] * yn;
      z[order - 1] = p - q;
    }
    y[n] = yn;
  }
}

/* Normalise (b, a) as _lfilter.py:96-117 does; writes n = max(nb, na)
 * coefficients into bn / an and returns the order n - 1. */
static int normalise(const double* b, int nb, const double* a, int na,
                     double* bn, double* an) {
  const int n = nb > na ? nb : na;
  for (int i = 0; i < n; ++i) {
    bn[i] = i < nb ? b[i] : 0.0;
    an[i] = i < na ? a[i] : 0.0;
  }
  if (an[0] != 1.0) {
    const double a0 = an[0];
    for (int i = 0; i < nb; ++i) bn[i] = b[i] / a0;
    for (int i = 0; i < na; ++i) an[i] = a[i] / a0;
  }
  return n - 1;
}

/* lfilter over B channels of T samples (rows contiguous).
 * zi: (B, order) or NULL (zeros); zf: (B, order) or NULL. */
void oracle_lfilter_f64(const double* b, int nb, const double* a, int na,
                        const double* x, double* y, int64_t B, int64_t T,
                        const double* zi, double* zf, int n_threads) {
  const int n = nb > na ? nb : na;
  double* bn = (double*)malloc(sizeof(double) * (size_t)n);
  double* an = (double*)malloc(sizeof(double) * (size_t)n);
  const int order = normalise(b, nb, a, na, bn, an);
#pragma omp parallel for schedule(static) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t c = 0; c < B; ++c) {
    double z[64];
    for (int k = 0; k < order; ++k) z[k] = zi ? zi[c * order + k] : 0.0;
    df2t_channel(bn, an, order, x + c * T, y + c * T, T, z);
    if (zf)
      for (int k = 0; k < order; ++k) zf[c * order + k] = z[k];
  }
  free(bn);
  free(an);
}

/* sosfilt over B channels of T samples (rows contiguous), in place on y.
 * sos: (n_filters, S, 6); channel c uses filter c / chans_per_filter.
 * zi / zf: (B, S, 2) or NULL.  x and y may alias. */
void oracle_sosfilt_f64(const double* sos, int S, int64_t chans_per_filter,
                        const double* x, double* y, int64_t B, int64_t T,
                        const double* zi, double* zf, int n_threads) {
#pragma omp parallel for schedule(static) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t c = 0; c < B; ++c) {
    const double* f = sos + (c / chans_per_filter) * (int64_t)S * 6;
    double* yc = y + c * T;
    if (yc != x + c * T) memcpy(yc, x + c * T, sizeof(double) * (size_t)T);
    for (int s = 0; s < S; ++s) {
      double bn[3], an[3], z[2];
      const int order = normalise(f + 6 * s, 3, f + 6 * s + 3, 3, bn, an);
      z[0] = zi ? zi[(c * S + s) * 2 + 0] : 0.0;
      z[1] = zi ? zi[(c * S + s) * 2 + 1] : 0.0;
      df2t_channel(bn, an, order, yc, yc, T, z);
      if (zf) {
        zf[(c * S + s) * 2 + 0] = z[0];
        zf[(c * S + s) * 2 + 1] = z[1];
      }
    }
  }
}

/* FP32-I/O variant for the CPU baseline: float rows in and out, FP64
 * arithmetic, state held per channel.  Streams one channel at a time so the
 * working set stays in cache. */
void oracle_sosfilt_f32io(const double* sos, int S, int64_t chans_per_filter,
                          const float* x, float* y, int64_t B, int64_t T,
                          int n_threads) {
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1)
  {
    double* buf = (double*)malloc(sizeof(double) * (size_t)(T > 0 ? T : 1));
#pragma omp for schedule(static)
    for (int64_t c = 0; c < B; ++c) {
      const double* f = sos + (c / chans_per_filter) * (int64_t)S * 6;
      for (int64_t n = 0; n < T; ++n) buf[n] = (double)x[c * T + n];
      for (int s = 0; s < S; ++s) {
        double bn[3], an[3], z[2] = {0.0, 0.0};
        const int order = normalise(f + 6 * s, 3, f + 6 * s + 3, 3, bn, an);
        df2t_channel(bn, an, order, buf, buf, T, z);
      }
      for (int64_t n = 0; n < T; ++n) y[c * T + n] = (float)buf[n];
    }
    free(buf);
  }
}

/* FP32 restatement of the reference's FP32 behaviour: with float32 sos and
 * x the reference runs every operation of _lfilter.py:96-98 (the a0
 * normalisation) and :153-166 (the DF2T update) in float32, one rounding per
 * ATen op, sections in order (_sosfilt.py:117-128).  Same layout as
 * oracle_sosfilt_f64; zi / zf (B, S, 2) float or NULL. */
static void df2t_channel_f32(float b0, float b1, float b2, float a1, float a2,
                             float* v, int64_t T, float* z) {
  float z0 = z[0], z1 = z[1];
  for (int64_t n = 0; n < T; ++n) {
    const float xn = v[n];
    const float t0 = b0 * xn;
    const float yn = t0 + z0;
    const float p1 = b1 * xn;
    const float q1 = a1 * yn;
    const float d1 = p1 - q1;
    z0 = d1 + z1;
    const float p2 = b2 * xn;
    const float q2 = a2 * yn;
    z1 = p2 - q2;
    v[n] = yn;
  }
  z[0] = z0;
  z[1] = z1;
}

void oracle_sosfilt_f32(const float* sos, int S, int64_t chans_per_filter,
                        const float* x, float* y, int64_t B, int64_t T,
                        const float* zi, float* zf, int n_threads) {
#pragma omp parallel for schedule(static) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t c = 0; c < B; ++c) {
    const float* f = sos + (c / chans_per_filter) * (int64_t)S * 6;
    float* yc = y + c * T;
    if (yc != x + c * T) memcpy(yc, x + c * T, sizeof(float) * (size_t)T);
    for (int s = 0; s < S; ++s) {
      float b0 = f[6 * s], b1 = f[6 * s + 1], b2 = f[6 * s + 2];
      float a0 = f[6 * s + 3], a1 = f[6 * s + 4], a2 = f[6 * s + 5];
      if (a0 != 1.0f) {
        b0 = b0 / a0;
        b1 = b1 / a0;
        b2 = b2 / a0;
        a1 = a1 / a0;
        a2 = a2 / a0;
      }
      float z[2];
      z[0] = zi ? zi[(c * S + s) * 2 + 0] : 0.0f;
      z[1] = zi ? zi[(c * S + s) * 2 + 1] : 0.0f;
      df2t_channel_f32(b0, b1, b2, a1, a2, yc, T, z);
      if (zf) {
        zf[(c * S + s) * 2 + 0] = z[0];
        zf[(c * S + s) * 2 + 1] = z[1];
      }
    }
  }
}

int oracle_abi_version(void) { return 1; }
