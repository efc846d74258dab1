// Global atomic throughput of the device for the patterns the pass kernels
// use (include/tsdf_calib.h).  Every lane's addresses come from a hash of
// (lane, step), so no two lanes of a launch meet on a line by construction
// of the pattern except where the pattern says so (SAME_RET32), and the
// returned values feed a sum that is stored only under a condition the
// kernel cannot meet, so no atomic is dropped as unused.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/tsdf_calib.h"

namespace {

constexpr int kThreads = 256;
constexpr int64_t kTableWords64 = int64_t(1) << 24;  // 128 MiB of 64-bit counters
constexpr int kLineWords64 = 16;                     // 128-byte lines
constexpr int64_t kLines = kTableWords64 / kLineWords64;

thread_local char g_err[256] = "";
int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return -1;
}

__device__ __forceinline__ uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}

__global__ void __launch_bounds__(kThreads) k_cal_scatter_ret64(unsigned long long* table, int steps,
                                                                 unsigned long long* sink) {
  const uint32_t lane = blockIdx.x * kThreads + threadIdx.x;
  unsigned long long acc = 0;
  for (int s = 0; s < steps; ++s) {
    const uint32_t line = mix(lane * 0x9E3779B9u + (uint32_t)s) & (uint32_t)(kLines - 1);
    acc += atomicAdd(&table[(int64_t)line * kLineWords64], 1ull);
  }
  if (acc == ~0ull) *sink = acc;
}

__global__ void __launch_bounds__(kThreads) k_cal_scatter_noret32(unsigned int* table, int steps) {
  const uint32_t lane = blockIdx.x * kThreads + threadIdx.x;
  for (int s = 0; s < steps; ++s) {
    const uint32_t line = mix(lane * 0x9E3779B9u + (uint32_t)s) & (uint32_t)(kLines - 1);
    atomicAdd(&table[(int64_t)line * (2 * kLineWords64)], 1u);
  }
}

__global__ void __launch_bounds__(kThreads) k_cal_coalesced_ret32(unsigned int* table, int steps, unsigned int* sink) {
  const uint32_t wave = (blockIdx.x * kThreads + threadIdx.x) >> 6;
  const uint32_t l = threadIdx.x & 63;
  unsigned int acc = 0;
  for (int s = 0; s < steps; ++s) {
    // 256-byte spans (two lines) at hashed places; 64 words per span
    const uint32_t span = mix(wave * 0x9E3779B9u + (uint32_t)s) & (uint32_t)(kLines / 2 - 1);
    acc += atomicAdd(&table[(int64_t)span * 64 + l], 1u);
  }
  if (acc == ~0u) *sink = acc;
}

__global__ void __launch_bounds__(kThreads) k_cal_same_ret32(unsigned int* word, int steps, unsigned int* sink) {
  if ((threadIdx.x & 63) != 0) return;
  unsigned int acc = 0;
  for (int s = 0; s < steps; ++s) acc += atomicAdd(word, 1u);
  if (acc == ~0u) *sink = acc;
}

#define HIPOK(x)                                                                  \
  do {                                                                            \
    const hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) return fail("%s: %s", #x, hipGetErrorString(e_));       \
  } while (0)

struct Resources {
  void* table = nullptr;
  void* sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Resources() {
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    if (st) (void)hipStreamDestroy(st);
    if (table) (void)hipFree(table);
    if (sink) (void)hipFree(sink);
  }
};

int launch(int pattern, int workgroups, int steps, Resources& R) {
  switch (pattern) {
    case TSDF_CAL_SCATTER_RET64:
      hipLaunchKernelGGL(k_cal_scatter_ret64, dim3(workgroups), dim3(kThreads), 0, R.st,
                         static_cast<unsigned long long*>(R.table), steps, static_cast<unsigned long long*>(R.sink));
      break;
    case TSDF_CAL_SCATTER_NORET32:
      hipLaunchKernelGGL(k_cal_scatter_noret32, dim3(workgroups), dim3(kThreads), 0, R.st,
                         static_cast<unsigned int*>(R.table), steps);
      break;
    case TSDF_CAL_COALESCED_RET32:
      hipLaunchKernelGGL(k_cal_coalesced_ret32, dim3(workgroups), dim3(kThreads), 0, R.st,
                         static_cast<unsigned int*>(R.table), steps, static_cast<unsigned int*>(R.sink));
      break;
    case TSDF_CAL_SAME_RET32:
      hipLaunchKernelGGL(k_cal_same_ret32, dim3(workgroups), dim3(kThreads), 0, R.st,
                         static_cast<unsigned int*>(R.table), steps, static_cast<unsigned int*>(R.sink));
      break;
    default:
      return fail("unknown pattern %d", pattern);
  }
  HIPOK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

const char* tsdf_calib_last_error(void) { return g_err; }

int tsdf_calib_atomics(int device, int pattern, int workgroups, int steps, int64_t* ops, double* seconds) {
  if (!ops || !seconds) return fail("null output");
  if (pattern < TSDF_CAL_SCATTER_RET64 || pattern > TSDF_CAL_SAME_RET32) return fail("unknown pattern %d", pattern);
  if (workgroups < 1 || workgroups > (1 << 20) || steps < 1 || steps > (1 << 16))
    return fail("workgroups must be 1..2^20 and steps 1..2^16");
  HIPOK(hipSetDevice(device));
  Resources R;
  HIPOK(hipMalloc(&R.table, kTableWords64 * sizeof(unsigned long long)));
  HIPOK(hipMalloc(&R.sink, sizeof(unsigned long long)));
  HIPOK(hipStreamCreate(&R.st));
  HIPOK(hipEventCreate(&R.ev[0]));
  HIPOK(hipEventCreate(&R.ev[1]));
  HIPOK(hipMemsetAsync(R.table, 0, kTableWords64 * sizeof(unsigned long long), R.st));
  if (launch(pattern, workgroups, steps, R)) return -1;  // untimed: first-touch and code load
  HIPOK(hipEventRecord(R.ev[0], R.st));
  if (launch(pattern, workgroups, steps, R)) return -1;
  HIPOK(hipEventRecord(R.ev[1], R.st));
  HIPOK(hipEventSynchronize(R.ev[1]));
  float ms = 0.f;
  HIPOK(hipEventElapsedTime(&ms, R.ev[0], R.ev[1]));
  const int64_t lanes = (int64_t)workgroups * kThreads / (pattern == TSDF_CAL_SAME_RET32 ? 64 : 1);
  *ops = lanes * steps;
  *seconds = ms * 1e-3;
  return 0;
}

}  // extern "C"
