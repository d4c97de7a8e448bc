// hs_syncprims_dev.h -- device code of the syncPrims hot path: the
// critical-section data walk, the four centralized mutexes, the reader/
// writer spin semaphores, and the atomic and lock-free tree barriers.
//
// Shared by the plain per-instance kernels (hs_syncprims.hip) and the
// batched persistent kernel (hs_batch.hip).  All functions are
// workgroup-collective: every thread of the WG calls them with uniform
// arguments; the WG's thread 0 performs the inter-WG atomics and the
// outcome is broadcast through LDS.
#pragma once

#include "hs_device.h"

namespace hs {

constexpr int kWordsPerLine = 16;  // oracle layout constant (HM:12)

// Lock block: 128 words (512 B) per lock instance.
//   word  0 : lock / head / ticket        (line 0, polled by spin/EBO)
//   word  1 : semaphore count              (line 0, only under the lock)
//   word  2 : semaphore writerWaiting      (line 0, only under the lock)
//   word  3 : semaphore max count
//   word 32 : tail / turn                 (line 1, polled by FA)
//   word 64 : writer / CS entry counter   (line 2, plain, under the lock)
//   word 96 : reader entry counter        (line 3, atomic)
constexpr int kLockWords = 128;
constexpr int kLkHead = 0, kLkCount = 1, kLkWaiting = 2, kLkMax = 3;
constexpr int kLkTail = 32, kLkCs = 64, kLkRd = 96;

// Barrier block: 64 words per level instance (count @0, generation @32).
constexpr int kBarWords = 64;

enum MutexKind { kSpin = 0, kEBO = 1, kSleep = 2, kFA = 3 };

struct Backoff {
  uint32_t min_q, max_q;  // s_sleep quanta; the reference's 1 .. 1023
  uint32_t poll_q;        // quanta between plain polls
};

// ---------------------------------------------------------------------------
// Critical-section data walk (accessData, HM:60-72): shift a chain of L+1
// 16-word lines up by one line.  MODE 1 ("inc") increments lines 1..L in
// place instead, which makes every lost or duplicated pass visible.
// Loads are batched 16 lines at a time; within one thread all addresses
// are distinct and no other thread touches them, so the reordering keeps
// the reference's semantics exactly.
// N source lines starting at `lo`: N independent loads, then N stores.
template <int MODE, int N>
__device__ __forceinline__ void shift_chunk(int32_t *__restrict__ col, int lo) {
  int32_t v[N];
#pragma unroll
  for (int k = 0; k < N; ++k)
    v[k] = col[(lo + k + (MODE == 0 ? 0 : 1)) * kWordsPerLine];
#pragma unroll
  for (int k = N - 1; k >= 0; --k)
    col[(lo + k + 1) * kWordsPerLine] = MODE == 0 ? v[k] : v[k] + 1;
}

template <int MODE>
__device__ __forceinline__ void access_data(int32_t *__restrict__ s,
                                            long base_line, int off, int L) {
  int32_t *col = s + base_line * kWordsPerLine + off;
  int hi = L;  // source lines [0, hi) remain, processed from the top down
  for (; hi >= 16; hi -= 16) shift_chunk<MODE, 16>(col, hi - 16);
  switch (hi) {
    case 15: shift_chunk<MODE, 15>(col, 0); break;
    case 14: shift_chunk<MODE, 14>(col, 0); break;
    case 13: shift_chunk<MODE, 13>(col, 0); break;
    case 12: shift_chunk<MODE, 12>(col, 0); break;
    case 11: shift_chunk<MODE, 11>(col, 0); break;
    case 10: shift_chunk<MODE, 10>(col, 0); break;
    case 9: shift_chunk<MODE, 9>(col, 0); break;
    case 8: shift_chunk<MODE, 8>(col, 0); break;
    case 7: shift_chunk<MODE, 7>(col, 0); break;
    case 6: shift_chunk<MODE, 6>(col, 0); break;
    case 5: shift_chunk<MODE, 5>(col, 0); break;
    case 4: shift_chunk<MODE, 4>(col, 0); break;
    case 3: shift_chunk<MODE, 3>(col, 0); break;
    case 2: shift_chunk<MODE, 2>(col, 0); break;
    case 1: shift_chunk<MODE, 1>(col, 0); break;
    default: break;
  }
}

// Lane `t` of a WG whose data region starts at global thread id `tid0`.
template <int MODE>
__device__ __forceinline__ void wg_access(int32_t *s, long tid0, int L) {
  const long tid = tid0 + threadIdx.x;
  access_data<MODE>(s, (tid / kWordsPerLine) * (L + 1),
                    threadIdx.x % kWordsPerLine, L);
}

// Semaphore reader body (accessData_semRd, HM:89-100): L loads, summed
// into LDS so they are not dead.
__device__ __forceinline__ void wg_read(const int32_t *s, long tid0, int L,
                                        volatile int32_t *sink) {
  const long tid = tid0 + threadIdx.x;
  const int32_t *col = s + (tid / kWordsPerLine) * (L + 1) * kWordsPerLine +
                       threadIdx.x % kWordsPerLine;
  int32_t acc = 0;
  for (int n = L - 1; n >= 0; --n) acc += col[n * kWordsPerLine];
  sink[threadIdx.x] += acc;
}

// ---------------------------------------------------------------------------
// Wait helpers (thread-level; return false once the guard gives up).
__device__ __forceinline__ bool spin_until_eq(const uint32_t *p, uint32_t v,
                                              uint32_t poll_q, Guard &g) {
  while (ld_rlx(p) != v) {
    if (!g.ok()) return false;
    sleep_quanta(poll_q);
  }
  fence_acq();
  return true;
}
__device__ __forceinline__ bool spin_until_ne(const uint32_t *p, uint32_t v,
                                              uint32_t poll_q, Guard &g) {
  while (ld_rlx(p) == v) {
    if (!g.ok()) return false;
    sleep_quanta(poll_q);
  }
  fence_acq();
  return true;
}

// ---------------------------------------------------------------------------
// Mutexes (master-thread parts).  `lk` is the lock block.
//   spin  : test-and-test-and-set on word 0 (HMS:15-52, CAS 0->1 / store 0)
//   EBO   : same with capped exponential s_sleep backoff (HME:14-66)
//   FA    : ticket lock, ticket = word 0, turn = word 32 (HMF:14-75)
//   sleep : array queue lock over a ring of slots (HMZ:22-123)
template <int KIND>
__device__ __forceinline__ bool mutex_lock(uint32_t *lk, uint32_t *ring,
                                           int ring_stride, uint32_t &ticket,
                                           const Backoff &bo, Guard &g) {
  if (KIND == kSpin || KIND == kEBO) {
    uint32_t b = bo.min_q;
    for (;;) {
      if (ld_rlx(lk + kLkHead) == 0 && cas_acq(lk + kLkHead, 0, 1) == 0)
        return true;
      if (!g.ok()) return false;
      if (KIND == kEBO) {
        sleep_quanta(b);
        b = ((b << 1) + 1) & bo.max_q;
        if (b < bo.min_q) b = bo.min_q;
      } else {
        sleep_quanta(bo.poll_q);
      }
    }
  } else if (KIND == kFA) {
    ticket = add_rlx(lk + kLkHead, 1);
    // proportional wait: the further back in the queue, the longer the nap
    for (;;) {
      const uint32_t turn = ld_rlx(lk + kLkTail);
      if (turn == ticket) break;
      if (!g.ok()) return false;
      const uint32_t ahead = ticket - turn;
      sleep_quanta(bo.poll_q * (ahead < 8 ? ahead : 8));
    }
    fence_acq();
    return true;
  } else {  // kSleep
    ticket = add_rlx(lk + kLkTail, 1);
    uint32_t *slot = ring + (long)ticket * ring_stride;
    uint32_t b = bo.min_q;
    while (ld_rlx(slot) != 1u) {
      if (!g.ok()) return false;
      sleep_quanta(b);
      b = ((b << 1) + 1) & bo.max_q;
      if (b < bo.min_q) b = bo.min_q;
    }
    fence_acq();
    return true;
  }
}

template <int KIND>
__device__ __forceinline__ void mutex_unlock(uint32_t *lk, uint32_t *ring,
                                             int ring_stride, uint32_t ticket) {
  if (KIND == kSpin || KIND == kEBO) {
    st_rel(lk + kLkHead, 0);
  } else if (KIND == kFA) {
    st_rel(lk + kLkTail, ticket + 1);
  } else {
    st_rlx(ring + (long)ticket * ring_stride, 0xffffffffu);
    st_rel(ring + (long)(ticket + 1) * ring_stride, 1u);
  }
}

// ---------------------------------------------------------------------------
// Reader/writer semaphore (HSS:32-232, HSE:32-249) over one record:
// {lock, count, writerWaiting, max}.  Writers need count == max and take
// the whole count; readers need !writerWaiting && count > 1 and take one.
// Returns 1 = entered, 0 = not yet, -1 = aborted.
__device__ __forceinline__ int sem_try_wait(uint32_t *lk, bool writer,
                                            Guard &g) {
  if (!(ld_rlx(lk + kLkHead) == 0 && cas_acq(lk + kLkHead, 0, 1) == 0))
    return g.ok() ? 0 : -1;
  const uint32_t cnt = ld_rlx(lk + kLkCount);
  const uint32_t mx = ld_rlx(lk + kLkMax);
  bool ok;
  if (writer) {
    ok = cnt == mx;
    if (ok) {
      st_rlx(lk + kLkCount, cnt - mx);
      st_rlx(lk + kLkWaiting, 0);
    } else {
      st_rlx(lk + kLkWaiting, 1);
    }
  } else {
    ok = ld_rlx(lk + kLkWaiting) == 0 && cnt > 1;
    if (ok) st_rlx(lk + kLkCount, cnt - 1);
  }
  st_rel(lk + kLkHead, 0);
  if (ok) {
    fence_acq();  // the CS data written by the previous holder
    return 1;
  }
  return g.ok() ? 0 : -1;
}

template <bool EBO>
__device__ __forceinline__ bool sem_wait(uint32_t *lk, bool writer,
                                         const Backoff &bo, Guard &g) {
  uint32_t b = bo.min_q;
  for (;;) {
    const int r = sem_try_wait(lk, writer, g);
    if (r > 0) return true;
    if (r < 0) return false;
    if (EBO) {
      sleep_quanta(b);
      // writers back off exponentially, readers linearly (HSE:177-181),
      // both capped at max_q
      b = writer ? (((b << 1) + 1) & bo.max_q) : b + 5;
      if (b > bo.max_q) b = bo.max_q;
      if (b < bo.min_q) b = bo.min_q;
    } else {
      sleep_quanta(bo.poll_q);
    }
  }
}

__device__ __forceinline__ bool sem_post(uint32_t *lk, bool writer,
                                         const Backoff &bo, Guard &g) {
  uint32_t lin = 0;
  for (;;) {
    if (ld_rlx(lk + kLkHead) == 0 && cas_acq(lk + kLkHead, 0, 1) == 0) break;
    if (!g.ok()) return false;
    lin = (lin + 1) % 31;  // linear backoff of HSS:202-203
    sleep_quanta(lin > bo.poll_q ? bo.poll_q : lin);
  }
  const uint32_t cnt = ld_rlx(lk + kLkCount);
  st_rlx(lk + kLkCount, writer ? cnt + ld_rlx(lk + kLkMax) : cnt + 1);
  st_rel(lk + kLkHead, 0);  // also publishes the writer's CS data
  return true;
}

// ---------------------------------------------------------------------------
// Atomic tree barrier (HBA:7-181), generation-counting form.
// Level 1: the perCU WGs of group g (blockIdx % nab == g, all on one XCD
// because nab is a multiple of 8) arrive on the group counter; the LAST
// arriver climbs to level 2 (global counter over the nab groups), and on
// the way down bumps the group generation that its peers poll.  One RMW
// per WG per level, read-only polling, no reset race: the generation word
// replaces the reference's pair of alternating counters (HBA:81-84).
struct TreeBar {
  uint32_t *global;  // kBarWords: count @0, gen @32
  uint32_t *local;   // nab * kBarWords
  uint32_t *episodes;  // plain counter, bumped by the global releaser
  int nab, per_cu;
};

__device__ __forceinline__ bool atomic_global_arrive(const TreeBar &t,
                                                     uint32_t poll_q,
                                                     Guard &g) {
  uint32_t *cnt = t.global, *gen = t.global + 32;
  const uint32_t gg = ld_rlx(gen);
  if (add_acqrel(cnt, 1) == (uint32_t)t.nab - 1) {
    st_rlx(cnt, 0);
    st_rlx(t.episodes, ld_rlx(t.episodes) + 1);
    st_rel(gen, gg + 1);
    return true;
  }
  return spin_until_ne(gen, gg, poll_q, g);
}

// Master-thread part; local_only = the local-exchange barrier (HM:352-353).
__device__ __forceinline__ bool atomic_tree_arrive(const TreeBar &t, int grp,
                                                   bool local_only,
                                                   uint32_t poll_q, Guard &g) {
  if (t.per_cu > 1) {
    uint32_t *cnt = t.local + (long)grp * kBarWords, *gen = cnt + 32;
    const uint32_t lg = ld_rlx(gen);
    if (add_acqrel(cnt, 1) == (uint32_t)t.per_cu - 1) {
      st_rlx(cnt, 0);
      bool ok = true;
      if (!local_only) ok = atomic_global_arrive(t, poll_q, g);
      st_rel(gen, lg + 1);
      return ok;
    }
    return spin_until_ne(gen, lg, poll_q, g);
  }
  return local_only ? true : atomic_global_arrive(t, poll_q, g);
}

// ---------------------------------------------------------------------------
// Lock-free (flag) tree barrier (HBF:11-333), epoch-flag form.  Each WG
// publishes an in-flag; the group's master WG polls its members' flags with
// one lane per flag, the global master WG (group 0's master) polls the
// nab group flags lane-parallel (a 256-group scan is one load per lane), and
// releases flow back down through per-WG out-flags.  Flags carry the
// epoch number instead of 0/1, so nothing has to be reset (HBF:110-123).
struct FlagBar {
  uint32_t *gin, *gout;  // nab words each
  uint32_t *lin, *lout;  // nab * per_cu words each, group-major
  uint32_t *episodes;
  int nab, per_cu;
};

// Whole-WG call.  grp = blockIdx % nab, q = blockIdx / nab.
__device__ __forceinline__ bool flag_tree_barrier(const FlagBar &f, int grp,
                                                  int q, uint32_t e_local,
                                                  uint32_t e_global,
                                                  bool local_only,
                                                  uint32_t poll_q, Guard &g,
                                                  int *s_flag) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  __syncthreads();
  if (f.per_cu > 1) {
    uint32_t *lin = f.lin + (long)grp * f.per_cu;
    uint32_t *lout = f.lout + (long)grp * f.per_cu;
    if (q != 0) {
      if (tid == 0) {
        st_rel(lin + q, e_local);
        *s_flag = spin_until_eq(lout + q, e_local, poll_q, g);
      }
      __syncthreads();
      return *s_flag != 0;
    }
    bool ok = true;
    for (int m = 1 + tid; m < f.per_cu; m += nthr)
      ok = ok && spin_until_eq(lin + m, e_local, poll_q, g);
    if (!__syncthreads_and(ok)) return false;
  }
  if (!local_only) {
    if (grp != 0) {
      if (tid == 0) {
        st_rel(f.gin + grp, e_global);
        *s_flag = spin_until_eq(f.gout + grp, e_global, poll_q, g);
      }
      __syncthreads();
      if (!*s_flag) return false;
    } else {
      bool ok = true;
      for (int m = 1 + tid; m < f.nab; m += nthr)
        ok = ok && spin_until_eq(f.gin + m, e_global, poll_q, g);
      if (!__syncthreads_and(ok)) return false;
      if (tid == 0) st_rlx(f.episodes, ld_rlx(f.episodes) + 1);
      for (int m = 1 + tid; m < f.nab; m += nthr) st_rel(f.gout + m, e_global);
    }
  }
  if (f.per_cu > 1) {
    uint32_t *lout = f.lout + (long)grp * f.per_cu;
    __syncthreads();
    for (int m = 1 + tid; m < f.per_cu; m += nthr) st_rel(lout + m, e_local);
  }
  return true;
}

}  // namespace hs
