// user_kernels.hip -- test kernels written against the public device API
// (include/heterosync_device.h) the way a user kernel would be: the
// reference's microbenchmark loops (hip/syncPrims/uvm/main.hip.cpp:104-149,
// 542-584, 712-857, 861-930) with the API's locks, semaphores and barriers
// and a plain-load/store critical section that increments each line in
// place, so the final storage is count-exact.  The GPU test compares it with
// the oracle.  Host entry point: hsu_run (C ABI, for ctypes).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "heterosync_device.h"

namespace {

// accessData (main.hip.cpp:60-72) with the copy replaced by an in-place
// increment of lines 1..L of the thread's chain; plain loads and stores.
// tid0: the first thread of the WG-sized region (region index x WG size).
__device__ inline void cs_inc(int32_t *s, long tid0, int L) {
  const long tid = tid0 + threadIdx.x;
  int32_t *col = s + (tid / 16) * (L + 1) * 16 + threadIdx.x % 16;
  for (int n = 1; n <= L; ++n) col[n * 16] += 1;
}

// accessData_semRd (main.hip.cpp:89-100): loads summed into LDS.
__device__ inline void cs_read(const int32_t *s, long tid0, int L, volatile int32_t *sink) {
  const long tid = tid0 + threadIdx.x;
  const int32_t *col = s + (tid / 16) * (L + 1) * 16 + threadIdx.x % 16;
  int32_t acc = 0;
  for (int n = 0; n < L; ++n) acc += col[n * 16];
  sink[threadIdx.x] += acc;
}

struct Args {
  int32_t *storage;
  uint32_t *state;  // lock / semaphore block, or barrier words
  uint32_t *abort;
  uint64_t ticks;
  int L, iters, num_cu, kind;
};

constexpr int kMaxWg = 256;  // WGs of 64 (the reference's) up to 256 threads

// kind 0 spin, 1 EBO, 3 ticket; every WG works on region 0 (main.hip.cpp:771)
__global__ void __launch_bounds__(kMaxWg) k_mutex(Args a) {
  hs_guard g = hs_guard_begin(a.abort, a.ticks);
  for (int i = 0; i < a.iters; ++i) {
    uint32_t ticket = 0;
    const bool ok = a.kind == 0   ? hs_mutex_spin_lock(a.state, g)
                    : a.kind == 1 ? hs_mutex_ebo_lock(a.state, g)
                                  : hs_mutex_fa_lock(a.state, &ticket, g);
    if (!ok) return;
    cs_inc(a.storage, 0, a.L);
    if (a.kind == 0) hs_mutex_spin_unlock(a.state);
    else if (a.kind == 1) hs_mutex_ebo_unlock(a.state);
    else hs_mutex_fa_unlock(a.state, ticket);
  }
}

// kind 0 spin, 1 EBO; one writer WG per group (main.hip.cpp:883), which
// walks regions 0..perCU-1; readers read region perCU_WGID
__global__ void __launch_bounds__(kMaxWg) k_sem(Args a) {
  __shared__ int32_t sink[kMaxWg];
  hs_guard g = hs_guard_begin(a.abort, a.ticks);
  const int b = blockIdx.x, nw = gridDim.x;
  const int nab = nw < a.num_cu ? nw : a.num_cu, per = (nw + nab - 1) / nab;
  const int cu = b % nab, pw = b / nab;
  const bool writer = pw == cu % per;
  sink[threadIdx.x] = 0;
  for (int i = 0; i < a.iters; ++i) {
    const bool ok = a.kind == 0 ? hs_semaphore_wait(a.state, writer, g)
                                : hs_semaphore_ebo_wait(a.state, writer, g);
    if (!ok) return;
    if (writer) {
      for (int j = 0; j < per; ++j) cs_inc(a.storage, (long)j * blockDim.x, a.L);
    } else {
      cs_read(a.storage, (long)pw * blockDim.x, a.L, sink);
    }
    hs_semaphore_post(a.state, writer);
  }
}

// the atomic tree barrier kernel (main.hip.cpp:104-149): own region, barrier,
// next region
__global__ void __launch_bounds__(kMaxWg) k_tree(Args a) {
  hs_guard g = hs_guard_begin(a.abort, a.ticks);
  const hs_tree_barrier bar = hs_tree_barrier_make(a.state, gridDim.x, a.num_cu);
  int cur = blockIdx.x;
  for (int i = 0; i < a.iters; ++i) {
    cs_inc(a.storage, (long)cur * blockDim.x, a.L);
    if (!hs_tree_barrier_join(bar, g)) return;
    cur = (cur + 1) % gridDim.x;
  }
}

// the lock-free tree barrier kernel (main.hip.cpp:542-584)
__global__ void __launch_bounds__(kMaxWg) k_lf(Args a) {
  hs_guard g = hs_guard_begin(a.abort, a.ticks);
  const hs_lf_barrier bar = hs_lf_barrier_make(a.state, gridDim.x, a.num_cu);
  hs_lf_epoch e = {0, 0};
  int cur = blockIdx.x;
  for (int i = 0; i < a.iters; ++i) {
    cs_inc(a.storage, (long)cur * blockDim.x, a.L);
    if (!hs_lf_barrier_join(bar, e, g)) return;
    cur = (cur + 1) % gridDim.x;
  }
}

#define CK(x)                       \
  do {                              \
    if ((x) != hipSuccess) return -2; \
  } while (0)

}  // namespace

// test: 0 mutex (kind 0 spin, 1 EBO, 3 ticket), 1 semaphore (kind 0 spin,
// 1 EBO; sem_count its size), 2 atomic tree barrier, 3 lock-free barrier.
// wg: threads per WG (64 to 256, a multiple of 64).  storage: n words, in
// and out.  Returns 0, -1 bad arguments, -2 HIP error,
// -4 the WGs cannot all be resident, -5 the watchdog fired.  *episodes:
// barrier episodes the device counted.
extern "C" int hsu_run(int test, int kind, int sem_count, int num_wgs, int num_cu, int L,
                       int iters, int wg, int32_t *storage, long n, double timeout_s,
                       uint32_t *episodes) {
  if (test < 0 || test > 3 || num_wgs < 1 || num_cu < 1 || L < 0 || iters < 0 || !storage ||
      wg < 64 || wg > kMaxWg || wg % 64)
    return -1;
  const void *kern = test == 0 ? (const void *)k_mutex
                     : test == 1 ? (const void *)k_sem
                     : test == 2 ? (const void *)k_tree
                                 : (const void *)k_lf;
  int dev = 0, per_cu = 0, ncu_dev = 0, khz = 0;
  CK(hipGetDevice(&dev));
  CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, wg, 0));
  CK(hipDeviceGetAttribute(&ncu_dev, hipDeviceAttributeMultiprocessorCount, dev));
  CK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (test > 0 && num_wgs > per_cu * ncu_dev) return -4;
  size_t words = HS_LOCK_WORDS;
  if (test == 2) words = hs_tree_barrier_words(num_wgs, num_cu);
  if (test == 3) words = hs_lf_barrier_words(num_wgs, num_cu);
  std::vector<uint32_t> init(words, 0);
  if (test == 1) hs_semaphore_init_words(init.data(), (uint32_t)sem_count);
  Args a;
  std::memset(&a, 0, sizeof a);
  CK(hipMalloc(&a.storage, n * 4));
  CK(hipMalloc(&a.state, words * 4));
  CK(hipMalloc(&a.abort, 256));
  CK(hipMemcpy(a.storage, storage, n * 4, hipMemcpyHostToDevice));
  CK(hipMemcpy(a.state, init.data(), words * 4, hipMemcpyHostToDevice));
  CK(hipMemset(a.abort, 0, 256));
  a.ticks = (uint64_t)(timeout_s * khz * 1000.0);
  a.L = L;
  a.iters = iters;
  a.num_cu = num_cu;
  a.kind = kind;
  hipLaunchKernelGGL((void (*)(Args))kern, dim3(num_wgs), dim3(wg), 0, 0, a);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  uint32_t ab = 0;
  CK(hipMemcpy(storage, a.storage, n * 4, hipMemcpyDeviceToHost));
  CK(hipMemcpy(&ab, a.abort, 4, hipMemcpyDeviceToHost));
  if (episodes) {
    *episodes = 0;
    if (test >= 2) CK(hipMemcpy(episodes, a.state + words - 32, 4, hipMemcpyDeviceToHost));
  }
  CK(hipFree(a.storage));
  CK(hipFree(a.state));
  CK(hipFree(a.abort));
  return ab ? -5 : 0;
}
