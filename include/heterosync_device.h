/*
 * heterosync_device.h -- device-side API of the synchronization hot path,
 * for user kernels (HIP, gfx950, wave64).  Header-only: include it from
 * device code built with `hipcc --offload-arch=gfx950 -I<repo>/include`.
 *
 * Reference device interfaces replaced (paths relative to the reference
 * tree, hip/syncPrims/uvm/):
 *   hs_mutex_spin_lock/unlock      hipMutexSpinLock/Unlock(+Local)
 *                                  hipLocksMutexSpin.h:15-97
 *   hs_mutex_ebo_lock/unlock       hipMutexEBOLock/Unlock(+Local)
 *                                  hipLocksMutexEBO.h:14-124
 *   hs_mutex_fa_lock/unlock        hipMutexFALock/Unlock(+Local)
 *                                  hipLocksMutexFA.h:14-144
 *   hs_semaphore_wait/post         hipSemaphoreSpinWait/Post(+Local)
 *                                  hipLocksSemaphoreSpin.h:148-428
 *   hs_semaphore_ebo_wait          hipSemaphoreEBOWait hipLocksSemaphoreEBO.h:148-249
 *   hs_semaphore_init_words        hipSemaphoreCreateSpin hipLocksSemaphoreSpin.h:7-30
 *   hs_tree_barrier_join           joinBarrier_helper hipLocksBarrierAtomic.h:156-181
 *   hs_lf_barrier_join             joinLFBarrier_helper hipLocksBarrierFast.h:318-333
 *
 * Conventions (as the reference's): every call is workgroup-collective --
 * every thread of the WG makes it with the same arguments, because it
 * contains WG barriers -- and thread (0,0,0) does the inter-WG atomics.
 * The reference's "Local" variants are the same calls on a per-CU (or
 * per-group) state block: pass that block.
 *
 * Differences from the reference, by design:
 *   - state is plain device memory the caller owns (hipMalloc), addressed
 *     by pointer instead of a handle into a process-global buffer: a lock
 *     or semaphore is HS_LOCK_WORDS words (its words on their own 128-byte
 *     lines), a barrier the size its *_words() function gives.  Zero-filled
 *     memory is an unlocked mutex and an idle barrier; a semaphore's block
 *     is set by hs_semaphore_init_words (host or device);
 *   - ordering follows the HIP memory model at agent scope: acquires are
 *     acquire RMWs or polls followed by an acquire fence, releases are
 *     release stores or RMWs after a WG barrier, so the critical section may
 *     use plain loads and stores (the reference's Local variants fence at
 *     workgroup scope only, which does not order accesses of WGs on other
 *     XCDs);
 *   - every wait is bounded by a watchdog (hs_guard): a call returns false
 *     once the guard's deadline passes or its abort word is set (and sets
 *     the abort word), so a protocol error ends the kernel instead of
 *     hanging the GPU.  A kernel should return when a call returns false.
 *   - a semaphore keeps count and writerWaiting in one word (one CAS per
 *     transition, no lock word); the state machine is the reference's:
 *     writers need the full count and take it, readers need
 *     !writerWaiting && count > 1 and take one.
 */
#ifndef HETEROSYNC_DEVICE_H
#define HETEROSYNC_DEVICE_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "hs_detail/hs_syncprims_dev.h"

#define HS_LOCK_WORDS 128

/* ---- watchdog ------------------------------------------------------------ */

typedef hs::Guard hs_guard;

/* abort_word: a device word shared by the kernel's WGs (zero at launch);
 * timeout_ticks: wall_clock64() ticks, i.e. seconds x the device's wall
 * clock rate (hipDeviceAttributeWallClockRate kHz x 1000; 100 MHz on
 * MI355X). */
__device__ inline hs_guard hs_guard_begin(uint32_t *abort_word, uint64_t timeout_ticks) {
  hs_guard g;
  g.init(abort_word, timeout_ticks);
  return g;
}

/* ---- mutexes (state: HS_LOCK_WORDS words, zero = unlocked) -------------- */

namespace hs_api {
template <int KIND>
__device__ inline bool lock(uint32_t *lk, uint32_t *ticket, const hs::Backoff &bo, hs_guard &g) {
  __shared__ int s_ok;
  __shared__ uint32_t s_ticket;
  if (hs::is_master()) {
    uint32_t t = 0;
    s_ok = hs::mutex_lock<KIND, false>(lk, nullptr, 0, t, bo, g);
    s_ticket = t;
  }
  __syncthreads();
  const bool ok = s_ok != 0;
  if (ticket) *ticket = s_ticket;
  __syncthreads();  // s_ok / s_ticket are read before a later call rewrites them
  return ok;
}

template <int KIND>
__device__ inline void unlock(uint32_t *lk, uint32_t ticket) {
  __syncthreads();  // every thread's critical-section accesses come first
  if (hs::is_master()) hs::mutex_unlock<KIND, false>(lk, nullptr, 0, ticket);
}

__device__ inline hs::Backoff backoff(uint32_t bmin, uint32_t bmax) {
  hs::Backoff b;
  b.min_q = bmin ? bmin : 1;
  b.max_q = bmax ? bmax : 1023;
  b.poll_q = 1;
  return b;
}
}  // namespace hs_api

/* Test-and-test-and-set spin lock. */
__device__ inline bool hs_mutex_spin_lock(uint32_t *lock, hs_guard &g) {
  return hs_api::lock<hs::kSpin>(lock, nullptr, hs_api::backoff(0, 0), g);
}
__device__ inline void hs_mutex_spin_unlock(uint32_t *lock) {
  hs_api::unlock<hs::kSpin>(lock, 0);
}

/* Spin lock with exponential s_sleep backoff from backoff_min up to the cap
 * backoff_max (2^k - 1); 0 selects the reference's 1 and 1023. */
__device__ inline bool hs_mutex_ebo_lock(uint32_t *lock, hs_guard &g, uint32_t backoff_min = 0,
                                         uint32_t backoff_max = 0) {
  return hs_api::lock<hs::kEBO>(lock, nullptr, hs_api::backoff(backoff_min, backoff_max), g);
}
__device__ inline void hs_mutex_ebo_unlock(uint32_t *lock) {
  hs_api::unlock<hs::kEBO>(lock, 0);
}

/* Ticket (fetch-and-add) lock: *ticket receives the WG's ticket, which the
 * matching unlock takes back. */
__device__ inline bool hs_mutex_fa_lock(uint32_t *lock, uint32_t *ticket, hs_guard &g) {
  return hs_api::lock<hs::kFA>(lock, ticket, hs_api::backoff(0, 0), g);
}
__device__ inline void hs_mutex_fa_unlock(uint32_t *lock, uint32_t ticket) {
  hs_api::unlock<hs::kFA>(lock, ticket);
}

/* ---- reader/writer semaphores (state: HS_LOCK_WORDS words) -------------- */

/* A semaphore of `count` (the reference's semaphore size: spinSem1/2/10/120
 * are 1, 2, 10, 120): its state holds count + 1 with a writer waiting
 * (hipLocksSemaphoreSpin.h:15-28).  Fill a zeroed HS_LOCK_WORDS block. */
__host__ __device__ inline void hs_semaphore_init_words(uint32_t *block, uint32_t count) {
  block[hs::kLkHead] = (count + 1) | hs::kSemWaitBit;
  block[hs::kLkMax] = count + 1;
}

namespace hs_api {
template <bool EBO>
__device__ inline bool sem_wait(uint32_t *sem, bool writer, const hs::Backoff &bo, hs_guard &g) {
  __shared__ int s_ok;
  if (hs::is_master()) s_ok = hs::sem_wait<EBO, false, false>(sem, writer, bo, g);
  __syncthreads();
  const bool ok = s_ok != 0;
  __syncthreads();
  return ok;
}
}  // namespace hs_api

__device__ inline bool hs_semaphore_wait(uint32_t *sem, bool is_writer, hs_guard &g) {
  return hs_api::sem_wait<false>(sem, is_writer, hs_api::backoff(0, 0), g);
}
/* The same wait with the reference's EBO schedule: writers back off
 * exponentially, readers linearly, both capped at backoff_max. */
__device__ inline bool hs_semaphore_ebo_wait(uint32_t *sem, bool is_writer, hs_guard &g,
                                             uint32_t backoff_min = 0, uint32_t backoff_max = 0) {
  return hs_api::sem_wait<true>(sem, is_writer, hs_api::backoff(backoff_min, backoff_max), g);
}
/* A writer returns the whole count, a reader one. */
__device__ inline void hs_semaphore_post(uint32_t *sem, bool is_writer) {
  __syncthreads();  // every thread's accesses inside come first
  if (hs::is_master())
    hs::sem_post<false, false>(sem, is_writer ? hs::ld_rlx(sem + hs::kLkMax) : 1u);
}

/* ---- atomic tree barrier -------------------------------------------------- */

/* Two levels, as joinBarrier_helper: the min(num_wgs, num_cu) groups of WGs
 * b with the same b % groups arrive on their group's counter, the group's
 * last arriver on the device counter.  Every WG of the grid joins every
 * episode, and all of them must be co-resident. */
typedef hs::TreeBar hs_tree_barrier;

__host__ __device__ inline size_t hs_tree_barrier_words(int num_wgs, int num_cu) {
  const int groups = num_wgs < num_cu ? num_wgs : num_cu;
  return (size_t)hs::kBarWords * (1 + groups) + 32;
}
/* words: hs_tree_barrier_words() zero-filled words. */
__host__ __device__ inline hs_tree_barrier hs_tree_barrier_make(uint32_t *words, int num_wgs,
                                                                int num_cu) {
  const int groups = num_wgs < num_cu ? num_wgs : num_cu;
  hs_tree_barrier t;
  t.global = words;
  t.local = words + hs::kBarWords;
  t.episodes = words + (size_t)hs::kBarWords * (1 + groups);
  t.nab = groups;
  t.per_cu = (num_wgs + groups - 1) / groups;
  return t;
}
__device__ inline bool hs_tree_barrier_join(const hs_tree_barrier &b, hs_guard &g) {
  __shared__ int s_ok;
  __syncthreads();
  if (hs::is_master())
    s_ok = hs::atomic_tree_arrive(b, (int)(blockIdx.x % (unsigned)b.nab), false, 1, g);
  __syncthreads();
  const bool ok = s_ok != 0;
  __syncthreads();
  return ok;
}
/* Completed episodes (the device releaser bumps it once per episode). */
__device__ inline uint32_t hs_tree_barrier_episodes(const hs_tree_barrier &b) {
  return hs::ld_rlx(b.episodes);
}

/* ---- lock-free (flag) tree barrier --------------------------------------- */

/* As joinLFBarrier_helper: every WG raises an in-flag; each group's first WG
 * gathers its members' flags, WG 0 the groups' flags, and releases flow back
 * through out-flags.  Flags carry episode numbers, so nothing is reset:
 * each WG counts its episodes in an hs_lf_epoch ({0, 0} at the first episode
 * on zero-filled flags). */
typedef hs::FlagBar hs_lf_barrier;
struct hs_lf_epoch {
  uint32_t local, global;
};

__host__ __device__ inline size_t hs_lf_barrier_words(int num_wgs, int num_cu) {
  const int groups = num_wgs < num_cu ? num_wgs : num_cu;
  const int per = (num_wgs + groups - 1) / groups;
  return 2 * (size_t)groups + 2 * (size_t)groups * per + 32;
}
__host__ __device__ inline hs_lf_barrier hs_lf_barrier_make(uint32_t *words, int num_wgs,
                                                            int num_cu) {
  const int groups = num_wgs < num_cu ? num_wgs : num_cu;
  const int per = (num_wgs + groups - 1) / groups;
  hs_lf_barrier f;
  f.gin = words;
  f.gout = words + groups;
  f.lin = words + 2 * (size_t)groups;
  f.lout = f.lin + (size_t)groups * per;
  f.episodes = f.lout + (size_t)groups * per;
  f.nab = groups;
  f.per_cu = per;
  return f;
}
__device__ inline bool hs_lf_barrier_join(const hs_lf_barrier &b, hs_lf_epoch &e, hs_guard &g) {
  __shared__ int s_flag;
  const int grp = (int)(blockIdx.x % (unsigned)b.nab), q = (int)(blockIdx.x / (unsigned)b.nab);
  ++e.local;
  ++e.global;
  const bool ok = hs::flag_tree_barrier(b, grp, q, e.local, e.global, false, 1, g, &s_flag);
  __syncthreads();
  return ok;
}

#endif /* HETEROSYNC_DEVICE_H */
