// hs_device.h -- gfx950 (CDNA4, wave64) device primitives shared by every
// kernel of the synchronization hot path.
//
// Memory-model policy: every word that another workgroup reads or writes is
// accessed with an agent-scope atomic, with the ordering chosen per access:
// polls are relaxed and followed by one acquire fence once they succeed,
// arrivals are release (or acq_rel) RMWs, hand-offs are release stores.
// Agent scope is the smallest scope that is correct here: on MI355X the
// workgroups of one kernel sit on 8 XCDs with separate L2s, so the
// reference's __threadfence_block() for inter-WG words (HMS:74, HBA:98)
// is not enough.  Lanes of one WG synchronise through LDS + __syncthreads.
//
// Every wait loop is bounded by a watchdog: after `deadline` (wall-clock
// ticks) or once the shared abort word is set, the waiter gives up, sets the
// abort word and the kernel drains.  A bug therefore ends a launch with an
// error instead of hanging the GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HS_AGENT __HIP_MEMORY_SCOPE_AGENT
#define HS_WG __HIP_MEMORY_SCOPE_WORKGROUP

// Lock and counter words live in their own 128-byte lines: index helpers.
#define HS_LINE_WORDS 32

namespace hs {

__device__ __forceinline__ uint32_t ld_rlx(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, HS_AGENT);
}
__device__ __forceinline__ uint32_t ld_acq(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, HS_AGENT);
}
__device__ __forceinline__ void st_rlx(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, HS_AGENT);
}
__device__ __forceinline__ void st_rel(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, HS_AGENT);
}
// A flag store that stays in this XCD's L2: a plain store (no cache-policy
// bits), left dirty in the L2 of the storing CU's XCD, where agent-scope
// loads from the same XCD return it without a memory round trip.  Only for
// words that, within one launch, are stored and loaded by WGs of one XCD
// alone: a word also stored by another XCD could leave a stale dirty copy
// here.
__device__ __forceinline__ void st_xcd(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ uint32_t add_rlx(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, HS_AGENT);
}
// CUDA/HIP atomicInc: returns the old value and stores old + 1, or 0 once
// old reaches `limit`, so a counter of n arrivals wraps itself at the nth.
__device__ __forceinline__ uint32_t inc_wrap_rlx(uint32_t *p, uint32_t limit) {
  return __builtin_amdgcn_atomic_inc32(p, limit, __ATOMIC_RELAXED, "agent");
}
// Bookkeeping count bumped after a release: a non-returning add, so the
// releaser does not wait on it, and increments from successive releasers
// commute, so the total is exact once the kernel ends.
__device__ __forceinline__ void count_rlx(uint32_t *p) {
  (void)__hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, HS_AGENT);
}
__device__ __forceinline__ void count_rlx_n(uint32_t *p, uint32_t n) {
  (void)__hip_atomic_fetch_add(p, n, __ATOMIC_RELAXED, HS_AGENT);
}
__device__ __forceinline__ uint32_t add_rel(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, HS_AGENT);
}
__device__ __forceinline__ uint32_t add_acqrel(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, HS_AGENT);
}
// CAS returning the old value (CUDA atomicCAS convention).
__device__ __forceinline__ uint32_t cas_acq(uint32_t *p, uint32_t expected,
                                            uint32_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQUIRE,
                                       __ATOMIC_RELAXED, HS_AGENT);
  return expected;
}
__device__ __forceinline__ uint32_t cas_rlx(uint32_t *p, uint32_t expected,
                                            uint32_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED, HS_AGENT);
  return expected;
}
__device__ __forceinline__ void fence_acq() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
__device__ __forceinline__ void fence_rel() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

// s_sleep takes an immediate; sleep `n` quanta (~64 clocks each) by binary
// decomposition (the role of sleepFunc, HL:22-69).  s_sleep is a scalar
// instruction and ignores the exec mask, so the count is taken from the
// first active lane: a count the compiler keeps in a vector register (one
// derived from a loaded value) would otherwise run every s_sleep of the
// decomposition.  Callers pass a wave-uniform count.
__device__ __forceinline__ void sleep_quanta(uint32_t n) {
  n = __builtin_amdgcn_readfirstlane(n);
  while (n >= 127) {
    __builtin_amdgcn_s_sleep(127);
    n -= 127;
  }
  if (n & 64) __builtin_amdgcn_s_sleep(64);
  if (n & 32) __builtin_amdgcn_s_sleep(32);
  if (n & 16) __builtin_amdgcn_s_sleep(16);
  if (n & 8) __builtin_amdgcn_s_sleep(8);
  if (n & 4) __builtin_amdgcn_s_sleep(4);
  if (n & 2) __builtin_amdgcn_s_sleep(2);
  if (n & 1) __builtin_amdgcn_s_sleep(1);
}

// Watchdog carried by every waiter.
struct Guard {
  uint32_t *abort;     // device word, nonzero = give up
  uint64_t deadline;   // wall_clock64() ticks
  uint32_t polls;

  __device__ __forceinline__ void init(uint32_t *abort_word, uint64_t ticks) {
    abort = abort_word;
    deadline = wall_clock64() + ticks;
    polls = 0;
  }
  // Call once per failed poll; false means "stop waiting".
  __device__ __forceinline__ bool ok() {
    if ((++polls & 31u) != 0) return true;
    if (ld_rlx(abort) != 0) return false;
    if (wall_clock64() > deadline) {
      st_rlx(abort, 1);
      return false;
    }
    return true;
  }
  __device__ __forceinline__ bool aborted() const { return ld_rlx(abort) != 0; }
};

__device__ __forceinline__ bool is_master() {
  return threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0;
}

}  // namespace hs
