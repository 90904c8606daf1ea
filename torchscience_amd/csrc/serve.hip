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

#include <cstring>
#include <mutex>

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

// One item (one channel of one chunk) walked by the calling wave, window
// by window, from its FP64 state, which it overwrites with the final state.
template <typename TR, int W>
__device__ __forceinline__ void serve_item(const tsiir_serve_item& it, float* lds,
                                           double (*sst)[2]) {
  const int lane = threadIdx.x;
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
  __syncthreads();
}

template <typename TR, int W>
__global__ __launch_bounds__(64) void sos_serve_kernel(const tsiir_serve_item* __restrict__ items) {
  __shared__ __align__(16) float lds[LdsStage<float, W>::kElems];
  __shared__ double sst[kServeMaxSections][2];
  serve_item<TR, W>(items[blockIdx.x], lds, sst);
}

// ---- persistent serving launch ---------------------------------------------
//
// A fixed grid of single-wave workgroups that takes items from a queue the
// host appends to while the launch runs.  The ring of items, the count of
// submitted items (tail), the stop flag and the per-slot completion words
// live in host memory the device reads and writes coherently; the count of
// claimed items (head) lives in device memory.  A wave claims the next
// submitted item with a compare-and-swap on head, walks it (serve_item),
// makes its outputs and state visible system-wide, and writes the item's
// ticket + 1 into the slot's completion word.  Every wave leaves the loop
// when no item is waiting and either the host has set stop or it has found
// nothing to do for idle_limit polls in a row, so the launch always drains;
// the host launches it again when it submits work after that.

struct ServeQueueCtl {
  uint64_t tail;   // items submitted (tickets 0 .. tail-1)
  uint32_t stop;   // the host's request to leave once nothing is waiting
  uint32_t pad;
};

// Device-memory side of the queue: the claim counter, and the doorbell
// wave's copies of tail and stop, so that the waiting waves poll device
// memory and only one wave reads host memory.
struct ServeDevCtl {
  uint64_t head;         // tickets claimed
  uint64_t tail_mirror;  // ServeQueueCtl::tail as the doorbell last read it
  uint32_t stop_mirror;
  uint32_t live;         // workers still in the loop (set by the host per launch)
};

constexpr uint64_t kNoTicket = ~0ull;
// Tickets a wave claims at once (at most kWave / 6 so that one load per
// lane brings in the whole group of items): fewer claims on the shared
// counter, one round trip to host memory and one system-scope release per
// group instead of per item.
constexpr uint64_t kServeClaim = 8;
// (Claims that shrink to one ticket while the backlog is short, so that a
// small batch's items run on as many waves, lowered the median latency of
// the 20,000 requests/s workload from 0.49 to 0.30-0.36 ms but raised its
// p99 from 1.5 to 7-26 ms, one system-scope release per item being the
// cost; always one ticket cut the capacity from 142 to 37 G s*s/s.)
// Longest idle pause of a worker, in s_sleep(127) units (~3.4 us each).
// Shorter pauses did not lower the median latency of the 20,000 requests/s
// workload (0.50 ms at caps 1, 4 and 16) and raised its p99 (17.8, 5.7 and
// 1.5 ms): idle workers polling the claim counter slow the busy ones.
constexpr int kServeMaxPause = 16;
constexpr int kItemWords = static_cast<int>(sizeof(tsiir_serve_item) / 8);
static_assert(kServeClaim * kItemWords <= kWave, "one item word per lane");

__device__ __forceinline__ uint64_t sys_load64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint64_t dev_load64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup 0 is the doorbell: it copies the host's tail and stop into
// device memory about every microsecond, and leaves once every worker has
// left.  The other workgroups claim groups of tickets against the copy and
// walk their items; a worker leaves once nothing is waiting and stop is set
// or it has been idle for idle_ticks of the 100 MHz real-time clock
// (polling with a growing pause, up to ~55 us).  So the launch ends idle_ticks
// after the last work, and the host launches it again for new work.
template <typename TR, int W>
__global__ __launch_bounds__(64) void sos_serve_persistent_kernel(
    const ServeQueueCtl* ctl, const tsiir_serve_item* ring, uint64_t* done, ServeDevCtl* dctl,
    uint64_t capacity, uint64_t idle_ticks) {
  __shared__ __align__(16) float lds[LdsStage<float, W>::kElems];
  __shared__ double sst[kServeMaxSections][2];
  __shared__ uint64_t s_items[kServeClaim * kItemWords];
  __shared__ uint64_t s_ticket;
  __shared__ uint32_t s_count, s_stop;
  const int lane = threadIdx.x;
  if (blockIdx.x == 0) {
    uint64_t seen = ~0ull;
    while (true) {
      if (lane == 0) {
        const uint64_t tail =
            __hip_atomic_load(&ctl->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        const uint32_t stop =
            __hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (tail != seen) {
          __hip_atomic_store(&dctl->tail_mirror, tail, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
          seen = tail;
        }
        __hip_atomic_store(&dctl->stop_mirror, stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_stop = __hip_atomic_load(&dctl->live, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u
                     ? 1u
                     : 0u;
      }
      __syncthreads();
      const uint32_t leave = s_stop;
      __syncthreads();
      if (leave) return;
      __builtin_amdgcn_s_sleep(32);
    }
  }
  uint64_t last = wall_clock64();
  int pause = 1;
  while (true) {
    if (lane == 0) {
      uint64_t ticket = kNoTicket;
      uint32_t count = 0;
      const uint64_t tail =
          __hip_atomic_load(&dctl->tail_mirror, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      uint64_t h = dev_load64(&dctl->head);
      while (h < tail) {
        const uint64_t k = tail - h < kServeClaim ? tail - h : kServeClaim;
        if (__hip_atomic_compare_exchange_strong(&dctl->head, &h, h + k, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          ticket = h;
          count = static_cast<uint32_t>(k);
          break;
        }
      }
      s_ticket = ticket;
      s_count = count;
      s_stop = ticket == kNoTicket ? __hip_atomic_load(&dctl->stop_mirror, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                                   : 0u;
    }
    __syncthreads();
    const uint64_t ticket = s_ticket;
    const uint32_t count = s_count;
    const uint32_t stop = s_stop;
    if (ticket != kNoTicket) {
      // the group's items, one 8-byte word per lane, read past every cache
      // (the slots are reused)
      if (lane < static_cast<int>(count) * kItemWords) {
        const int j = lane / kItemWords, i = lane % kItemWords;
        const uint64_t* w = reinterpret_cast<const uint64_t*>(ring + (ticket + j) % capacity);
        s_items[lane] = sys_load64(w + i);
      }
      __syncthreads();
      for (uint32_t j = 0; j < count; ++j) {
        tsiir_serve_item it;
        uint64_t* iw = reinterpret_cast<uint64_t*>(&it);
#pragma unroll
        for (int i = 0; i < kItemWords; ++i) iw[i] = s_items[j * kItemWords + i];
        serve_item<TR, W>(it, lds, sst);
      }
      // outputs and states of the group visible to every XCD and the host
      // (one write-back of this XCD's L2 per group), then the completions
      __threadfence_system();
      __syncthreads();
      if (lane < static_cast<int>(count))
        __hip_atomic_store(done + (ticket + lane) % capacity, ticket + lane + 1,
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      last = wall_clock64();
      pause = 1;
    } else {
      __syncthreads();
      if (stop != 0u || wall_clock64() - last > idle_ticks) break;
      for (int p = 0; p < pause; ++p) __builtin_amdgcn_s_sleep(127);
      if (pause < kServeMaxPause) pause *= 2;
    }
  }
  if (lane == 0) __hip_atomic_fetch_sub(&dctl->live, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// ---- persistent serving queue (host side) ----------------------------------

struct tsiir_serve_queue {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ended = nullptr;    // recorded after the latest launch
  bool launched = false;
  ServeQueueCtl* ctl = nullptr;  // host memory, coherent with the device
  tsiir_serve_item* ring = nullptr;
  uint64_t* done = nullptr;
  ServeQueueCtl* ctl_dev = nullptr;  // the same memory, as the device addresses it
  tsiir_serve_item* ring_dev = nullptr;
  uint64_t* done_dev = nullptr;
  ServeDevCtl* dctl = nullptr;   // device memory
  uint64_t capacity = 0;
  uint64_t tail = 0;             // host copy of ctl->tail
  int32_t workgroups = 0;
  int32_t mode = TSIIR_MODE_COMPENSATED;
  uint64_t idle_ticks = 0;       // of the device's 100 MHz real-time clock
  std::mutex mu;
};

namespace {

template <typename T>
bool host_alloc(T** host, T** dev, size_t bytes) {
  if (hipHostMalloc(reinterpret_cast<void**>(host), bytes, hipHostMallocCoherent) != hipSuccess)
    return false;
  memset(*host, 0, bytes);
  void* d = nullptr;
  if (hipHostGetDevicePointer(&d, *host, 0) != hipSuccess) return false;
  *dev = static_cast<T*>(d);
  return true;
}

// Launch the persistent grid unless a launch is still running.
tsiir_status ensure_running(tsiir_serve_queue* q) {
  if (q->launched && hipEventQuery(q->ended) == hipErrorNotReady) return TSIIR_OK;
  (void)hipGetLastError();
  // workgroup 0 is the doorbell, the others serve items
  const dim3 grid(static_cast<unsigned>(q->workgroups) + 1u), block(kWave);
  if (hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(&q->dctl->live),
                        static_cast<int>(q->workgroups), 1, q->stream) != hipSuccess)
    return TSIIR_ELAUNCH;
  if (q->mode == TSIIR_MODE_COMPENSATED)
    hipLaunchKernelGGL((sos_serve_persistent_kernel<double, 32>), grid, block, 0, q->stream,
                       q->ctl_dev, q->ring_dev, q->done_dev, q->dctl, q->capacity, q->idle_ticks);
  else
    hipLaunchKernelGGL((sos_serve_persistent_kernel<float, 32>), grid, block, 0, q->stream,
                       q->ctl_dev, q->ring_dev, q->done_dev, q->dctl, q->capacity, q->idle_ticks);
  if (hipGetLastError() != hipSuccess) return TSIIR_ELAUNCH;
  if (hipEventRecord(q->ended, q->stream) != hipSuccess) return TSIIR_ELAUNCH;
  q->launched = true;
  return TSIIR_OK;
}

void free_queue(tsiir_serve_queue* q) {
  if (q->dctl) (void)hipFree(q->dctl);
  if (q->ctl) (void)hipHostFree(q->ctl);
  if (q->ring) (void)hipHostFree(q->ring);
  if (q->done) (void)hipHostFree(q->done);
  if (q->ended) (void)hipEventDestroy(q->ended);
  if (q->stream) (void)hipStreamDestroy(q->stream);
  delete q;
}

}  // namespace

extern "C" {

tsiir_status tsiir_serve_create(int32_t capacity, int32_t workgroups, int32_t mode,
                                int32_t idle_us, tsiir_serve_queue** out) {
  if (!out || capacity < 1 || workgroups < 1 || idle_us < 1) return TSIIR_EINVAL;
  if (mode != TSIIR_MODE_PLAIN && mode != TSIIR_MODE_COMPENSATED) return TSIIR_EINVAL;
  *out = nullptr;
  tsiir_serve_queue* q = new tsiir_serve_queue();
  q->capacity = static_cast<uint64_t>(capacity);
  q->workgroups = workgroups;
  q->mode = mode;
  q->idle_ticks = static_cast<uint64_t>(idle_us) * 100;  // 100 MHz real-time clock
  bool ok = hipGetDevice(&q->device) == hipSuccess &&
            hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&q->ended, hipEventDisableTiming) == hipSuccess &&
            host_alloc(&q->ctl, &q->ctl_dev, sizeof(ServeQueueCtl)) &&
            host_alloc(&q->ring, &q->ring_dev, q->capacity * sizeof(tsiir_serve_item)) &&
            host_alloc(&q->done, &q->done_dev, q->capacity * sizeof(uint64_t)) &&
            hipMalloc(reinterpret_cast<void**>(&q->dctl), sizeof(ServeDevCtl)) == hipSuccess &&
            hipMemset(q->dctl, 0, sizeof(ServeDevCtl)) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    free_queue(q);
    return TSIIR_ELAUNCH;
  }
  *out = q;
  return TSIIR_OK;
}

tsiir_status tsiir_serve_submit(tsiir_serve_queue* q, const tsiir_serve_item* items,
                                int32_t n_items, uint64_t* first_ticket) {
  if (!q || n_items < 0 || (n_items > 0 && !items)) return TSIIR_EINVAL;
  std::lock_guard<std::mutex> lk(q->mu);
  if (first_ticket) *first_ticket = q->tail;
  if (n_items == 0) return TSIIR_OK;
  if (static_cast<uint64_t>(n_items) > q->capacity) return TSIIR_EAGAIN;
  for (int32_t i = 0; i < n_items; ++i) {
    if (items[i].n_sections < 1 || items[i].n_sections > kServeMaxSections ||
        items[i].n_samples < 0 || !items[i].table || !items[i].state ||
        (items[i].n_samples > 0 && (!items[i].x || !items[i].y)))
      return TSIIR_EINVAL;
  }
  // every slot to be written must have finished its previous ticket
  for (int32_t i = 0; i < n_items; ++i) {
    const uint64_t t = q->tail + static_cast<uint64_t>(i);
    if (t >= q->capacity) {
      const uint64_t prev = t - q->capacity;
      const uint64_t d = __atomic_load_n(q->done + t % q->capacity, __ATOMIC_ACQUIRE);
      if (d != prev + 1) return TSIIR_EAGAIN;
    }
  }
  for (int32_t i = 0; i < n_items; ++i) q->ring[(q->tail + i) % q->capacity] = items[i];
  q->tail += static_cast<uint64_t>(n_items);
  __atomic_store_n(&q->ctl->tail, q->tail, __ATOMIC_RELEASE);
  return ensure_running(q);
}

tsiir_status tsiir_serve_poll(tsiir_serve_queue* q, uint64_t first_ticket, int32_t n_items,
                              int32_t* all_done) {
  if (!q || !all_done || n_items < 0) return TSIIR_EINVAL;
  std::lock_guard<std::mutex> lk(q->mu);
  if (first_ticket + static_cast<uint64_t>(n_items) > q->tail) return TSIIR_EINVAL;
  int32_t ok = 1;
  for (int32_t i = 0; i < n_items && ok; ++i) {
    const uint64_t t = first_ticket + static_cast<uint64_t>(i);
    if (t + q->capacity < q->tail) continue;  // its slot was reused: done long ago
    ok = __atomic_load_n(q->done + t % q->capacity, __ATOMIC_ACQUIRE) == t + 1;
  }
  *all_done = ok;
  // items submitted as the launch ran out of idle polls: launch again
  return ok ? TSIIR_OK : ensure_running(q);
}

tsiir_status tsiir_serve_destroy(tsiir_serve_queue* q) {
  if (!q) return TSIIR_EINVAL;
  __atomic_store_n(&q->ctl->stop, 1u, __ATOMIC_RELEASE);
  const bool ok = hipStreamSynchronize(q->stream) == hipSuccess;
  free_queue(q);
  return ok ? TSIIR_OK : TSIIR_ELAUNCH;
}

}  // extern "C"
