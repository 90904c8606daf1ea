// serve.hip - ragged batches of streaming sosfilt chunks (one launch per batch).
//
// Serving form of the sosfilt path (_sosfilt.py:13-133 with zi/zf): each
// request is a chunk of a persistent stream, filtered from the stream's
// carried state, which the kernel updates in place.  A batch mixes requests
// of different channel counts, lengths and section counts.
//
// Work item = one channel of one request = one wavefront.  The wave walks the
// chunk in windows of 64*W samples; inside a window it uses the same
// per-section machinery as the forward kernel (fused correction + zero-state
// pass, FP64 wave scan), with the window's incoming state known exactly from
// the previous window (kept in LDS), so items never wait on each other.  An
// item's result therefore does not depend on what else is in the batch.

#include "sos_common.h"
#include "sos_internal.h"
#include "tsiir.h"

namespace tsiir {

constexpr int kServeMaxSections = 32;

// One window of an item.  FULL: 64*W samples, LDS-staged I/O when aligned,
// corrections deferred into the next section's sweep.  !FULL: the chunk's
// partial last window (masked), corrections applied at once, and the state
// after the chunk's last sample taken from the lane that holds it.
template <typename TR, int W, bool FULL>
__device__ __forceinline__ void serve_window(const float* xr, float* yr, const double* tf, int S,
                                             int64_t k0, int64_t T, bool aligned, float* lds,
                                             double (*sst)[2]) {
  const int lane = threadIdx.x;
  const int64_t t0 = k0 + static_cast<int64_t>(lane) * W;
  int m = W;
  if (!FULL) {
    const int64_t rem = T - t0;
    m = rem <= 0 ? 0 : (rem >= W ? W : static_cast<int>(rem));
  }
  const bool owns_end = !FULL && t0 < T && t0 + W >= T;
  float v[W];
  if (FULL && aligned)
    load_window_lds<float, float, W>(xr + k0, lds, v);
  else
    load_run<float, float, W>(xr + t0, false, m, v);

  TR pw0 = 0, pw1 = 0, pa1 = 0, pa2 = 0;
  for (int s = 0; s < S; ++s) {
    const double* tb = tf + static_cast<int64_t>(s) * kTableStride;
    const TR b0 = static_cast<TR>(tb[0]), b1 = static_cast<TR>(tb[1]);
    const TR b2 = static_cast<TR>(tb[2]), a1 = static_cast<TR>(tb[3]);
    const TR a2 = static_cast<TR>(tb[4]);
    TR z0 = 0, z1 = 0;
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (FULL || j < m) {
        const TR cj = pw0;
        pw0 = pw1 - pa1 * cj;
        pw1 = -pa2 * cj;
        const TR u = static_cast<TR>(v[j]) + cj;
        const TR yv = b0 * u + z0;
        z0 = b1 * u + (z1 - a1 * yv);
        z1 = b2 * u + (-a2 * yv);
        v[j] = static_cast<float>(yv);
      }
    }
    // chained fold: lane 0 carries the window's incoming state
    const double sw0 = sst[s][0], sw1 = sst[s][1];
    double e0 = static_cast<double>(z0), e1 = static_cast<double>(z1);
    if (lane == 0) {
      const Mat2 M = load_mat(table_pow(tb, 0));
      e0 += M.a * sw0 + M.b * sw1;
      e1 += M.c * sw0 + M.d * sw1;
    }
    double incl0, incl1, s0, s1;
    wave_scan(tb, lane, e0, e1, incl0, incl1, s0, s1);
    if (lane == 0) {
      s0 = sw0;
      s1 = sw1;
    }
    __syncthreads();  // every lane has read sst[s]
    if (FULL) {
      pw0 = static_cast<TR>(s0);
      pw1 = static_cast<TR>(s1);
      pa1 = a1;
      pa2 = a2;
      if (lane == 0) {
        sst[s][0] = incl0;
        sst[s][1] = incl1;
      }
    } else {
      TR w0 = static_cast<TR>(s0), w1 = static_cast<TR>(s1);
#pragma unroll
      for (int j = 0; j < W; ++j) {
        if (j < m) {
          const TR cj = w0;
          v[j] = static_cast<float>(static_cast<TR>(v[j]) + cj);
          w0 = w1 - a1 * cj;
          w1 = -a2 * cj;
        }
      }
      if (owns_end) {
        sst[s][0] = static_cast<double>(w0) + static_cast<double>(z0);
        sst[s][1] = static_cast<double>(w1) + static_cast<double>(z1);
      }
    }
    __syncthreads();
  }
  if (FULL) {
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const TR cj = pw0;
      pw0 = pw1 - pa1 * cj;
      pw1 = -pa2 * cj;
      v[j] = static_cast<float>(static_cast<TR>(v[j]) + cj);
    }
  }
  if (FULL && aligned)
    store_window_lds<float, float, W>(yr + k0, lds, v);
  else
    store_run<float, float, W>(yr + t0, false, m, v);
}

template <typename TR, int W>
__global__ __launch_bounds__(64) void sos_serve_kernel(const tsiir_serve_item* __restrict__ items) {
  __shared__ __align__(16) float lds[LdsStage<float, W>::kElems];
  __shared__ double sst[kServeMaxSections][2];
  const int lane = threadIdx.x;
  const tsiir_serve_item it = items[blockIdx.x];
  const int S = it.n_sections;
  const int64_t T = it.n_samples;
  const float* xr = static_cast<const float*>(it.x);
  float* yr = static_cast<float*>(it.y);
  double* state = it.state;
  if (lane < 2 * S) sst[lane >> 1][lane & 1] = state[lane];
  __syncthreads();
  const int64_t win = static_cast<int64_t>(kWave) * W;
  const bool aligned =
      ((reinterpret_cast<uintptr_t>(xr) | reinterpret_cast<uintptr_t>(yr)) & 15u) == 0;
  int64_t k0 = 0;
  for (; k0 + win <= T; k0 += win)
    serve_window<TR, W, true>(xr, yr, it.table, S, k0, T, aligned, lds, sst);
  if (k0 < T) serve_window<TR, W, false>(xr, yr, it.table, S, k0, T, aligned, lds, sst);
  __syncthreads();
  if (lane < 2 * S) state[lane] = sst[lane >> 1][lane & 1];
}

}  // namespace tsiir

using namespace tsiir;

extern "C" {

size_t tsiir_serve_table_bytes(int32_t n_sections) {
  if (n_sections < 1 || n_sections > kServeMaxSections) return 0;
  return static_cast<size_t>(n_sections) * kTableStride * sizeof(double);
}

tsiir_status tsiir_serve_prepare_filter(const void* sos, int32_t coef_dtype, int32_t n_sections,
                                        double* table, void* stream) {
  if (!sos || !table || n_sections < 1 || n_sections > kServeMaxSections) return TSIIR_EINVAL;
  if (coef_dtype != TSIIR_F32 && coef_dtype != TSIIR_F64) return TSIIR_EINVAL;
  launch_prep(sos, coef_dtype, n_sections, 5, false, table, nullptr, 0,
              static_cast<hipStream_t>(stream));
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

tsiir_status tsiir_serve_batch(const tsiir_serve_item* items, int32_t n_items, int32_t mode,
                               void* stream) {
  if (n_items < 0 || (n_items > 0 && !items)) return TSIIR_EINVAL;
  if (mode != TSIIR_MODE_PLAIN && mode != TSIIR_MODE_COMPENSATED) return TSIIR_EINVAL;
  if (n_items == 0) return TSIIR_OK;
  if (mode == TSIIR_MODE_COMPENSATED)
    hipLaunchKernelGGL((sos_serve_kernel<double, 32>), dim3(static_cast<unsigned>(n_items)),
                       dim3(kWave), 0, static_cast<hipStream_t>(stream), items);
  else
    hipLaunchKernelGGL((sos_serve_kernel<float, 32>), dim3(static_cast<unsigned>(n_items)),
                       dim3(kWave), 0, static_cast<hipStream_t>(stream), items);
  return hipGetLastError() == hipSuccess ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // extern "C"
