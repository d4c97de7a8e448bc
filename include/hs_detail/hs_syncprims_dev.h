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
//   word  0 : lock / head / ticket        (line 0, polled by spin/EBO);
//             semaphore state: count | writerWaiting << 16
//   word  3 : semaphore max count
//   word 32 : tail / turn                 (line 1, polled by FA)
//   word 48 : posts in flight             (line 1, priority semaphores)
//   word 64 : writer / CS entry counter   (line 2, plain, under the lock)
//   word 96 : reader entry counter        (line 3, atomic)
constexpr int kLockWords = 128;
constexpr int kLkHead = 0, kLkMax = 3;
constexpr int kLkTail = 32, kLkCs = 64, kLkRd = 96;
// priority semaphores: number of posts in flight (HSS:26, 591-595), on the
// tail's line, which semaphores do not otherwise use
constexpr int kLkPrio = 48;

// Barrier block: 64 words per level instance (count @0, generation @32).
constexpr int kBarWords = 64;

// Waits until every memory access this wave issued has completed.
__device__ __forceinline__ void wait_mem() { __builtin_amdgcn_s_waitcnt(0); }

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
// SC1 = true: loads and stores are agent-scope relaxed atomics
// (global_load/store ... sc1): loads bypass the CU's L1 and stores are
// written through, so the data is at the device coherence point once the
// wave's memory counter drains.  Used by the XCD-aware barriers.
template <bool SC1>
__device__ __forceinline__ int32_t ld_data(const int32_t *p) {
  if (SC1) return __hip_atomic_load(p, __ATOMIC_RELAXED, HS_AGENT);
  return *p;
}

// N source lines of a thread's column starting at `lo`: N independent
// loads (MODE 0: lines lo..lo+N-1; MODE 1: lines lo+1..lo+N), and the
// stores of lines lo+1..lo+N (copied, or incremented in place).  N is a
// compile-time constant so the accesses are straight-line code that the
// compiler keeps in flight together.
template <int MODE, int N, bool SC1>
__device__ __forceinline__ void chunk_ld(const int32_t *col, int lo, int32_t (&v)[16]) {
#pragma unroll
  for (int k = 0; k < N; ++k)
    v[k] = ld_data<SC1>(col + (lo + k + (MODE == 0 ? 0 : 1)) * kWordsPerLine);
}
template <int MODE, int N, bool SC1>
__device__ __forceinline__ void chunk_st(int32_t *col, int lo, const int32_t (&v)[16]) {
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    const int32_t val = MODE == 0 ? v[k] : v[k] + 1;
    if (SC1)  // write-through to the coherence point: nothing left dirty in L2
      __hip_atomic_store(col + (lo + k + 1) * kWordsPerLine, val, __ATOMIC_RELAXED, HS_AGENT);
    else
      col[(lo + k + 1) * kWordsPerLine] = val;
  }
}

// The bottom chunk [0, r), r < 16 lines: loads or stores by a switch over
// the compile-time sizes.
template <int MODE, bool SC1, bool STORE>
__device__ __forceinline__ void chunk_tail(int32_t *col, int r, int32_t (&v)[16]) {
  switch (r) {
#define HS_TAIL(n)                                          \
  case n:                                                   \
    if (STORE) chunk_st<MODE, n, SC1>(col, 0, v);           \
    else chunk_ld<MODE, n, SC1>(col, 0, v);                 \
    break;
    HS_TAIL(1) HS_TAIL(2) HS_TAIL(3) HS_TAIL(4) HS_TAIL(5) HS_TAIL(6) HS_TAIL(7) HS_TAIL(8)
    HS_TAIL(9) HS_TAIL(10) HS_TAIL(11) HS_TAIL(12) HS_TAIL(13) HS_TAIL(14) HS_TAIL(15)
#undef HS_TAIL
    default: break;
  }
}

// The walk goes top-down: full chunks of 16 source lines, then the bottom
// chunk of L % 16.  A chunk's loads are issued before the chunk above it is
// stored: they read lines that chunk does not write (it stores lines
// lo+1..lo+16, the chunk below reads lines under lo), and the lowest line a
// chunk reads is overwritten only by the stores of the chunk below, issued
// after that chunk's own loads have returned (vector memory operations
// complete in order).  The memory counter counts loads and stores in issue
// order, so waiting for a chunk's loads never waits for stores issued after
// them: each chunk costs one load round trip instead of a load and a store.
// SC1: agent-scope loads and stores; SC1_ST (default SC1) the stores' scope
// alone (SC1 loads with plain stores leave the lines dirty in this XCD's L2).
template <int MODE, bool SC1 = false, bool SC1_ST = SC1>
__device__ __forceinline__ void access_data(int32_t *__restrict__ s,
                                            long base_line, int off, int L) {
  if (L <= 0) return;
  int32_t *col = s + base_line * kWordsPerLine + off;
  const int full = L / 16, r = L % 16;
  int32_t a[16], b[16];
  if (full == 0) {
    chunk_tail<MODE, SC1, false>(col, r, a);
    chunk_tail<MODE, SC1_ST, true>(col, r, a);
    return;
  }
  int lo = L - 16;
  chunk_ld<MODE, 16, SC1>(col, lo, a);
  for (int f = 1; f < full; ++f) {
    chunk_ld<MODE, 16, SC1>(col, lo - 16, b);
    chunk_st<MODE, 16, SC1_ST>(col, lo, a);
    lo -= 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = b[k];
  }
  if (r > 0) chunk_tail<MODE, SC1, false>(col, r, b);
  chunk_st<MODE, 16, SC1_ST>(col, lo, a);
  if (r > 0) chunk_tail<MODE, SC1_ST, true>(col, r, b);
}

// Lane `t` of a WG whose data region starts at global thread id `tid0`.
template <int MODE, bool SC1 = false>
__device__ __forceinline__ void wg_access(int32_t *s, long tid0, int L) {
  const long tid = tid0 + threadIdx.x;
  access_data<MODE, SC1>(s, (tid / kWordsPerLine) * (L + 1),
                         threadIdx.x % kWordsPerLine, L);
}

// The same pass done by the whole WG with 16-byte accesses.  A WG's region
// is the contiguous chains of its blockDim threads (blockDim is a multiple
// of 64, so the region starts on a chain), and each chain is L+1 lines of
// 64 B.  The WG moves it band by band, top band first (at most BAND source
// lines, i.e. BAND/4 pieces per thread), with all loads of a band complete
// before any of its stores -- exactly the per-thread top-down walk's result,
// since a band's stores only overwrite lines its own loads have already
// read.  Each band costs a load and a store round trip, so a long walk
// (numLdSt = 100) takes taller bands: BAND = 64 moves it in two.
// SC1: agent-scope accesses (L1 bypassed, stores written through).
// keep_l2 (with SC1): the region's next access is by a WG on this XCD, which
// stores the region again.  Plain stores then leave the lines dirty in this
// XCD's L2, where that WG's agent-scope loads find them (a plain store stays
// in its XCD's L2 and agent-scope loads on that XCD return it).  The last
// store before a region's next access moves to another XCD must not use
// keep_l2: an agent-scope store over a dirty line leaves no stale copy for a
// later return of the region to this XCD.  Both facts are measured by the
// synchronization probe (visibility and line-return sections).
typedef uint32_t hs_v4u __attribute__((ext_vector_type(4)));
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer descriptor, gfx9 family

template <int MODE, bool SC1 = false, int BAND = 16, bool INC_BAR = false>
__device__ __forceinline__ void wg_move(int32_t *s, long tid0, int L, bool keep_l2 = false) {
  static_assert(BAND % 16 == 0 && BAND >= 16 && BAND <= 64, "band of 16..64 lines");
  constexpr int kItems = BAND / 4;  // 16-byte pieces per thread per band
  if (L <= 0) return;
  const int nt = (int)blockDim.x, t = (int)threadIdx.x;
  const int nc = nt / kWordsPerLine;  // chains in the region
  int32_t *base = s + (tid0 / kWordsPerLine) * (long)(L + 1) * kWordsPerLine;
  const int bytes = nc * (L + 1) * kWordsPerLine * 4;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, bytes, kRsrcWord3);
  constexpr int kAux = SC1 ? 16 : 0;  // cache policy: sc1
  for (int hi = L; hi > 0; hi -= BAND) {  // source lines [lo, hi)
    const int lo = hi > BAND ? hi - BAND : 0, nb = hi - lo, items = nc * nb * 4;
    hs_v4u v[kItems];
    int dst[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
      const int q = t + k * nt;
      dst[k] = -1;
      if (q < items) {
        const int part = q & 3, line = lo + (q >> 2) % nb, chain = (q >> 2) / nb;
        dst[k] = ((chain * (L + 1) + line + 1) * kWordsPerLine + part * 4) * 4;
        const int src = MODE == 0 ? dst[k] - kWordsPerLine * 4 : dst[k];
        v[k] = __builtin_amdgcn_raw_buffer_load_b128(rs, src, 0, kAux);
      }
    }
    // Copy mode: other lanes store into lines this band still reads.  Inc
    // mode (INC_BAR) takes the same barrier where two waves touch one
    // 128-byte line (see inc_band_barrier).
    if (MODE == 0 || INC_BAR) {
      wait_mem();
      __syncthreads();
    }
    if (SC1 && !keep_l2) {
#pragma unroll
      for (int k = 0; k < kItems; ++k)
        if (dst[k] >= 0)
          __builtin_amdgcn_raw_buffer_store_b128(MODE == 0 ? v[k] : v[k] + 1u, rs, dst[k], 0, kAux);
    } else {
#pragma unroll
      for (int k = 0; k < kItems; ++k)
        if (dst[k] >= 0)
          __builtin_amdgcn_raw_buffer_store_b128(MODE == 0 ? v[k] : v[k] + 1u, rs, dst[k], 0, 0);
    }
  }
}

// Whether an inc-mode band walk of WGs of `wg` threads should drain and
// barrier between its loads and stores (wg_move's INC_BAR).  Each wave holds
// 16 consecutive lines of a band's chain-major order, so a wave boundary
// falls inside a chain unless the band's lines per chain divide 16; one
// wave's stores then interleave with another wave's loads of a shared
// 128-byte line; a walk of several bands also keeps the next band's loads
// behind the stores of the one above.  Measured at 1024 x 256 (XCD-aware
// atomicTreeBarrUniq, one box): numLdSt 10, 3.90 us per episode without the
// barrier and 3.41 with it (copy mode 3.34); numLdSt 100, 34.0 without and
// 32.8 with it (copy 32.6-32.8); numLdSt 1 and 4, 2.64 and 2.75 without it,
// 2.72 and 2.84 with it (on another box all within 1-2%).  The choice is per
// launch and compile-time inside the walk: a run-time test between the
// loads and the stores measured slower than either form.  The mutex and
// semaphore walks (one band of 32 lines, the reference's 64-thread WGs)
// always drain first: without it an inc-mode faMutex hand-off took 2.05 us
// against 1.57 us in copy mode (32 WGs x 64, numLdSt 10); with it, 1.57.
__host__ __device__ inline bool inc_band_barrier(int wg, int L, int band = 16) {
  if (wg <= 64 || L <= 0) return false;
  return L > band || band % L != 0;
}

// The data pass of a barrier episode: wg_move, with the inc-mode band
// barrier chosen per launch (a uniform kernel argument).
template <int MODE, bool SC1>
__device__ __forceinline__ void wg_pass(bool inc_bar, int32_t *s, long tid0, int L,
                                        bool keep_l2 = false) {
  if (MODE == 1 && inc_bar) wg_move<1, SC1, 16, true>(s, tid0, L, keep_l2);
  else wg_move<MODE, SC1, 16, false>(s, tid0, L, keep_l2);
}

// Lines a reader loads before it adds any of them.
constexpr int kReadBatch = 32;

// Semaphore reader body (accessData_semRd, HM:89-100): L loads, summed
// into LDS so they are not dead.
template <bool SC1 = false>
__device__ __forceinline__ void wg_read(const int32_t *s, long tid0, int L,
                                        volatile int32_t *sink) {
  const long tid = tid0 + threadIdx.x;
  const int32_t *col = s + (tid / kWordsPerLine) * (L + 1) * kWordsPerLine +
                       threadIdx.x % kWordsPerLine;
  // independent loads, issued 32 at a time before any is added (a load-add
  // loop over agent-scope loads waits out one round trip per line; eight at
  // a time waited out one per eight lines)
  int32_t acc = 0;
  for (int n0 = 0; n0 < L; n0 += kReadBatch) {
    int32_t v[kReadBatch];
#pragma unroll
    for (int u = 0; u < kReadBatch; ++u)
      v[u] = n0 + u < L ? ld_data<SC1>(col + (n0 + u) * kWordsPerLine) : 0;
#pragma unroll
    for (int u = 0; u < kReadBatch; ++u) acc += v[u];
  }
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
// AGENT: the critical section's data accesses are agent-scope (L1 bypassed,
// stores written through) and every wave drains them before the WG releases
// the lock, so acquiring and releasing need no cache fence.
template <int KIND, bool AGENT = false>
__device__ __forceinline__ bool mutex_lock(uint32_t *lk, uint32_t *ring,
                                           int ring_stride, uint32_t &ticket,
                                           const Backoff &bo, Guard &g) {
  if (KIND == kSpin || KIND == kEBO) {
    uint32_t b = bo.min_q;
    for (;;) {
      if (ld_rlx(lk + kLkHead) == 0 &&
          (AGENT ? cas_rlx(lk + kLkHead, 0, 1) : cas_acq(lk + kLkHead, 0, 1)) == 0)
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
    if (!AGENT) fence_acq();
    return true;
  } else {  // kSleep
    ticket = add_rlx(lk + kLkTail, 1);
    uint32_t *slot = ring + (long)ticket * ring_stride;
    uint32_t b = bo.min_q;
    for (;;) {
      const uint32_t v = ld_rlx(slot);
      const uint32_t done = ld_rlx(lk + kLkCs);  // entries completed (read alongside)
      if (v == 1u) break;
      if (!g.ok()) return false;
      // the reference's exponential schedule (HMS:52-56), capped by this
      // waiter's place in the queue: it polls a slot of its own, so a sleep
      // longer than the hand-offs ahead of it only delays its own turn
      const uint32_t ahead = ticket - done;
      const uint32_t cap = bo.poll_q * (ahead < 8 ? ahead : 8);
      sleep_quanta(b < cap ? b : cap);
      b = ((b << 1) + 1) & bo.max_q;
      if (b < bo.min_q) b = bo.min_q;
    }
    if (!AGENT) fence_acq();
    return true;
  }
}

// Master thread, after every wave's critical-section accesses (AGENT: and
// their completion) and a WG barrier.
template <int KIND, bool AGENT = false>
__device__ __forceinline__ void mutex_unlock(uint32_t *lk, uint32_t *ring,
                                             int ring_stride, uint32_t ticket) {
  if (AGENT) wait_mem();  // this wave's own stores (the CS counter) first
  if (KIND == kSpin || KIND == kEBO) {
    if (AGENT) st_rlx(lk + kLkHead, 0); else st_rel(lk + kLkHead, 0);
  } else if (KIND == kFA) {
    if (AGENT) st_rlx(lk + kLkTail, ticket + 1); else st_rel(lk + kLkTail, ticket + 1);
  } else {
    st_rlx(ring + (long)ticket * ring_stride, 0xffffffffu);
    if (AGENT) st_rlx(ring + (long)(ticket + 1) * ring_stride, 1u);
    else st_rel(ring + (long)(ticket + 1) * ring_stride, 1u);
  }
}

// ---------------------------------------------------------------------------
// Reader/writer semaphore (HSS:32-232, HSE:32-249).  The reference keeps
// {count, writerWaiting} behind a CAS lock word and updates them with plain
// accesses; here both live in ONE atomic word (count in bits 0-15,
// writerWaiting in bit 16) and every transition is a single CAS, so a failed
// attempt costs one RMW instead of a lock hand-off with a release/acquire
// pair.  The state machine is the reference's: writers need count == max and
// take the whole count, clearing writerWaiting; a failing writer sets
// writerWaiting; readers need !writerWaiting && count > 1 and take one;
// Post adds max (writer) or 1 (reader).
constexpr uint32_t kSemWaitBit = 1u << 16;

__device__ __forceinline__ uint32_t sem_init_word(uint32_t max) { return max | kSemWaitBit; }

// Returns 1 = entered, 0 = not yet, -1 = aborted.  AGENT as for mutexes.
template <bool AGENT = false>
__device__ __forceinline__ int sem_try_wait(uint32_t *lk, bool writer, Guard &g) {
  const uint32_t mx = ld_rlx(lk + kLkMax);
  uint32_t s = ld_rlx(lk + kLkHead);
  for (;;) {
    const uint32_t cnt = s & 0xffffu;
    const bool waiting = (s & kSemWaitBit) != 0;
    uint32_t want;
    bool enter;
    if (writer) {
      enter = cnt == mx;
      if (enter) want = 0;                       // count -= max, writerWaiting = 0
      else if (!waiting) want = s | kSemWaitBit;  // announce the waiting writer
      else return g.ok() ? 0 : -1;
    } else {
      enter = !waiting && cnt > 1;
      if (!enter) return g.ok() ? 0 : -1;
      want = s - 1;
    }
    const uint32_t old =
        (enter && !AGENT) ? cas_acq(lk + kLkHead, s, want) : cas_rlx(lk + kLkHead, s, want);
    if (old == s) return enter ? 1 : (g.ok() ? 0 : -1);
    s = old;  // raced with another update: re-evaluate
    if (!g.ok()) return -1;
  }
}

// Priority variants (HSS:430-470): an entry attempt first waits until no
// post is in flight, backing off exponentially while one is.
__device__ __forceinline__ bool sem_yield_to_posts(const uint32_t *lk, const Backoff &bo,
                                                   Guard &g) {
  uint32_t b = bo.min_q;
  while (ld_rlx(lk + kLkPrio) != 0) {
    if (!g.ok()) return false;
    sleep_quanta(b);
    b = ((b << 1) + 1) & bo.max_q;
    if (b < bo.min_q) b = bo.min_q;
  }
  return true;
}

template <bool EBO, bool PRIO = false, bool AGENT = false>
__device__ __forceinline__ bool sem_wait(uint32_t *lk, bool writer,
                                         const Backoff &bo, Guard &g) {
  uint32_t b = bo.min_q;
  for (;;) {
    if (PRIO && !sem_yield_to_posts(lk, bo, g)) return false;
    const int r = sem_try_wait<AGENT>(lk, writer, g);
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

// Post: one release RMW (also publishes a writer's CS data) that returns
// `amount` to the count: the semaphore's max for a writer, 1 for a reader.
// The caller reads max once (it never changes after initialisation), so
// the release waits on no load.  The priority form brackets it with the
// in-flight count that entering WGs yield to (HSS:591-595, 639-640).
template <bool PRIO = false, bool AGENT = false>
__device__ __forceinline__ void sem_post(uint32_t *lk, uint32_t amount) {
  if (PRIO) add_rlx(lk + kLkPrio, 1u);
  if (AGENT) {
    wait_mem();  // this wave's own stores (the CS counters) first
    add_rlx(lk + kLkHead, amount);
  } else {
    add_rel(lk + kLkHead, amount);
  }
  if (PRIO) add_rlx(lk + kLkPrio, 0xffffffffu);
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
    st_rel(gen, gg + 1);
    count_rlx(t.episodes);
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
      for (int m = 1 + tid; m < f.nab; m += nthr) st_rel(f.gout + m, e_global);
      if (tid == 0) count_rlx(f.episodes);
    }
  }
  if (f.per_cu > 1) {
    uint32_t *lout = f.lout + (long)grp * f.per_cu;
    __syncthreads();
    for (int m = 1 + tid; m < f.per_cu; m += nthr) st_rel(lout + m, e_local);
  }
  return true;
}

// ---------------------------------------------------------------------------
// XCD-aware tree barriers.
//
// MI355X has 8 XCDs, each with its own L2; agent-scope release/acquire
// (buffer_wbl2 sc1 / buffer_inv sc1) costs ~56 ns and these operations are
// serialised per XCD, so a barrier where each of 1024 WGs releases and
// acquires spends ~20 us per episode in cache maintenance alone.  Workgroups
// are dispatched round-robin over XCDs (WG b on XCD b % 8, verified at kernel
// start through the XCC_ID hardware register), so every reference group
// (blockIdx % NUM_CU) lies inside one XCD.  The barrier therefore has three
// levels -- group, XCD, device -- and only the XCD leader touches the caches:
//   * below the XCD leader, arrivals are relaxed atomics issued after the
//     arriving WG's stores have completed to the shared L2 (s_waitcnt 0),
//     and waiters read shared data with L1-bypassing agent-scope loads;
//   * the XCD leader writes the XCD's L2 back (release) before arriving at
//     the device level and invalidates it (acquire) before releasing its XCD.
// Per episode this is 2 cache operations per XCD instead of 2-3 per WG.

__device__ __forceinline__ bool poll_ne(const uint32_t *p, uint32_t v,
                                        uint32_t poll_q, Guard &g) {
  while (ld_rlx(p) == v) {
    if (!g.ok()) return false;
    sleep_quanta(poll_q);
  }
  return true;
}
__device__ __forceinline__ bool poll_eq(const uint32_t *p, uint32_t v,
                                        uint32_t poll_q, Guard &g) {
  while (ld_rlx(p) != v) {
    if (!g.ok()) return false;
    sleep_quanta(poll_q);
  }
  return true;
}

// Release waits by one lane per WG: two polls in flight.  Each check waits
// only for the older load while the newer one is already on its way, so a
// word that changes is seen about half a memory round trip sooner than by a
// load-check-reload loop.  Loads of one wave return in order, so the checks
// see the word's values in time order.  (Gathers, where many lanes of the
// gathering WG poll at once, keep one load per lane.)
__device__ __forceinline__ bool poll_eq_pipe(const uint32_t *p, uint32_t v,
                                             uint32_t poll_q, Guard &g) {
  uint32_t older = ld_rlx(p);
  for (;;) {
    sleep_quanta(poll_q);
    const uint32_t newer = ld_rlx(p);
    if (older == v) return true;
    if (!g.ok()) return false;
    sleep_quanta(poll_q);
    older = ld_rlx(p);
    if (newer == v) return true;
    if (!g.ok()) return false;
  }
}

__device__ __forceinline__ uint32_t xcc_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 0xfu;  // HW_REG_XCC_ID
}

__device__ __forceinline__ int groups_on_xcd(int nab, int nxcd, int x) {
  return nab / nxcd + (x < nab % nxcd ? 1 : 0);
}

// The XCD-aware barriers split each XCD's WGs into sub-groups of `sub`
// (a power of two the host picks from the grid, sub_group_size()).
struct XcdTree {
  uint32_t *glob;      // kBarWords: count @0, gen @32 (one arrival per sub-counter)
  uint32_t *sub;       // nxcd * sub_max * kBarWords: sub-counters, XCD-major
  uint32_t *grp;       // nab * kBarWords: count @0, gen @32 (local exchange)
  uint32_t *rel;       // nxcd * kBarWords: per-XCD copy of the device generation
  uint32_t *episodes;
  int nab, per_cu, nxcd;
  int sub_max, n_sub;  // sub-counters per XCD (at most) and in all
  int sub_wgs;         // WGs per sub-counter
};

// A WG's fixed place in the XCD-aware tree, computed once per kernel so
// that the divisions locating it stay off every arrival's path.
struct XcdSeat {
  uint32_t *sc;       // its sub-counter
  uint32_t *rel;      // its XCD's release word
  uint32_t sub_last;  // arrivals at that sub-counter - 1
};
__device__ __forceinline__ XcdSeat xcd_seat(const XcdTree &t, int grp, int q) {
  const int per = t.per_cu;
  const int x = grp % t.nxcd;
  const int nwx = groups_on_xcd(t.nab, t.nxcd, x) * per;  // WGs on this XCD
  const int wx = (grp / t.nxcd) * per + q;
  const int sw = t.sub_wgs, k = wx / sw;
  const int ck = nwx - k * sw < sw ? nwx - k * sw : sw;
  return XcdSeat{t.sub + (long)(x * t.sub_max + k) * kBarWords, t.rel + (long)x * kBarWords,
                 (uint32_t)ck - 1};
}

// Master thread; every wave of the WG has run wait_mem() + __syncthreads().
// A global arrival climbs two levels of relaxed atomicInc RMWs: the WG's
// sub-counter (up to sub_wgs WGs of one XCD), then the device counter (one
// arrival per sub-counter).  The last arriver of a level is the one that
// reads the level's size - 1; its RMW wraps the counter back to 0, so no
// reset store exists for a WG of the next episode to overtake (it arrives
// only after the release, which follows every counter's wrap).  The device releaser writes one release word per XCD and
// every WG of that XCD polls its word, so the release is a single hop.
// `seat` locates the WG's sub-counter and release word (xcd_seat).
// `epoch` is the number of barriers of this kind completed once
// this one is: the value the release word (global) or the group's
// generation (local_only) must reach, so nothing is read before arriving.
// The WGs' data accesses are agent-scope where they leave the XCD and
// complete before they arrive, so no level needs a cache fence.
__device__ __forceinline__ bool xcd_tree_arrive(const XcdTree &t, const XcdSeat &seat, int grp,
                                                bool local_only, uint32_t epoch,
                                                uint32_t poll_q, Guard &g) {
  const uint32_t per = (uint32_t)t.per_cu;
  if (local_only) {  // group-level barrier (local exchange, HM:352-353)
    if (per <= 1) return true;
    uint32_t *cnt = t.grp + (long)grp * kBarWords, *gen = cnt + 32;
    if (inc_wrap_rlx(cnt, per - 1) != per - 1) return poll_eq_pipe(gen, epoch, poll_q, g);
    // a group lies inside one XCD: its generation stays in that XCD's L2
    st_xcd(gen, epoch);
    return true;
  }
  uint32_t *rel = seat.rel;
  if (inc_wrap_rlx(seat.sc, seat.sub_last) != seat.sub_last) return poll_eq_pipe(rel, epoch, poll_q, g);
  const uint32_t nsub = (uint32_t)t.n_sub;
  if (inc_wrap_rlx(t.glob, nsub - 1) != nsub - 1) return poll_eq_pipe(rel, epoch, poll_q, g);
  const uint32_t nx = (uint32_t)(t.nab < t.nxcd ? t.nab : t.nxcd);
  for (uint32_t y = 0; y < nx; ++y) st_rlx(t.rel + (long)y * kBarWords, epoch);
  st_rlx(t.glob + 32, epoch);
  count_rlx(t.episodes);
  return true;
}

// Flag i of an array lives at word i * kFlagStride: one 128-byte line per
// flag, so the WGs that store flags and the gatherers that poll them never
// share a line.
constexpr int kFlagStride = HS_LINE_WORDS;

struct XcdFlags {
  uint32_t *gout;        // nxcd: release word per XCD
  uint32_t *din;         // nxcd * sub_max: sub-group leaders' arrivals (device level)
  uint32_t *sin;         // nxcd * sub_max * sub_wgs: members' arrivals (sub level)
  uint32_t *lin, *lout;  // nab * per_cu: group level (local exchange), group-major;
                         // lout stays 0 (the group barrier is all-to-all)
  uint32_t *episodes;
  int nab, per_cu, nxcd;
  int sub_max;           // sub-groups per XCD (at most), as for XcdTree
  int sub_wgs;           // WGs per sub-group
  int a2a;               // XCD form: one sub-group per XCD, leaders all-to-all
};

// Poll until the word has reached `v` (wrapping compare), for a flag that
// its writer can move one epoch past `v` before this poll has seen `v`.
__device__ __forceinline__ bool poll_ge(const uint32_t *p, uint32_t v,
                                        uint32_t poll_q, Guard &g) {
  while ((int32_t)(ld_rlx(p) - v) < 0) {
    if (!g.ok()) return false;
    sleep_quanta(poll_q);
  }
  return true;
}

// Most flags one lane of a gathering WG polls at once (gather_eq's K).
constexpr int kGatherMax = 8;

// A lane's share of a gather: the flags first + i * step below end, i < K,
// at base + j * stride, polled until every one equals v.  All K loads are
// issued together each pass (a lane with fewer flags re-reads its first),
// so a pass costs one round trip whatever K is, with no per-flag branches.
template <int K>
__device__ __forceinline__ bool gather_eq(const uint32_t *base, long stride, int first, int step,
                                          int end, uint32_t v, uint32_t poll_q, Guard &g) {
  if (first >= end) return true;
  const uint32_t *p[K];
  bool valid[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int j = first + i * step;
    valid[i] = j < end;
    p[i] = base + (long)(valid[i] ? j : first) * stride;
  }
  for (;;) {
    uint32_t w[K];
#pragma unroll
    for (int i = 0; i < K; ++i) w[i] = ld_rlx(p[i]);
    bool all = true;
#pragma unroll
    for (int i = 0; i < K; ++i) all = all && (!valid[i] || w[i] == v);
    if (all) return true;
    if (!g.ok()) return false;
    sleep_quanta(poll_q);
  }
}

// Whole-WG call.  el = group-level epoch, eg = episode (device level).
// The global barrier gathers arrival flags on two levels, as the tree
// barrier counts them: the WGs of one XCD form sub-groups of up to
// sub_wgs (numbered as in xcd_tree_arrive), and each leader gathers its
// members' flags one per lane.  Then one of two forms (the host picks):
//   * XCD form (a2a; one sub-group per XCD, when the leader has a lane for
//     each of its XCD's other WGs): the 8 XCD leaders meet all-to-all --
//     each publishes its XCD's arrival and gathers the other leaders' --
//     and each releases its own XCD.  One cross-XCD hop per episode;
//   * device form: the device master (leader of sub-group 0) gathers every
//     sub-group leader's flag, one per lane, and writes one release word
//     per XCD (gout[x]) for every XCD.  Two cross-XCD hops.
// Every WG of XCD x waits on gout[x].  The local-exchange barrier is
// all-to-all: each member stores its in-flag and polls the other members'.
// Flags and data are agent-scope loads and drained stores, so no level
// needs a cache fence.  A flag that, within one launch, is stored and
// polled only inside one XCD (the sub-group and group flags; the release
// words of the XCD form) is stored plainly and stays dirty in that XCD's
// L2, where the agent-scope polls find it without a memory round trip
// (lfTreeBarrUniq at 1024 x 256: 210k -> 241k episodes/s, 253k with the
// XCD form).  Cross-XCD flags are written through.
__device__ __forceinline__ bool xcd_flag_barrier(const XcdFlags &f, int grp, int q,
                                                 uint32_t el, uint32_t eg,
                                                 bool local_only, uint32_t poll_q,
                                                 Guard &g, int *s_flag) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  constexpr long S = kFlagStride;
  wait_mem();
  __syncthreads();
  if (local_only) {  // all-to-all within the group: one hop, no master
    if (f.per_cu <= 1) return true;
    // Each member stores its own flag and polls the others', one per lane.
    // Epochs only grow, and a member stores its next one only after the
    // global barrier in between, which every member reaches after its
    // polls here, so each poll sees exactly `el` before it moves on.
    uint32_t *lin = f.lin + (long)grp * f.per_cu * S;
    if (tid == 0) st_xcd(lin + q * S, el);
    bool ok = true;
    for (int m = tid; m < f.per_cu; m += nthr)
      if (m != q) ok = ok && poll_eq(lin + m * S, el, poll_q, g);
    return __syncthreads_and(ok) != 0;
  }
  const int x = grp % f.nxcd;
  const uint32_t *rel = f.gout + x * S;
  const int per = f.per_cu;
  const int nwx = groups_on_xcd(f.nab, f.nxcd, x) * per;  // WGs on this XCD
  const int wx = (grp / f.nxcd) * per + q;
  const int sw = f.sub_wgs, k = wx / sw, m = wx % sw;
  const int ck = nwx - k * sw < sw ? nwx - k * sw : sw;
  const long sid = (long)x * f.sub_max + k;
  uint32_t *sin = f.sin + sid * sw * S;
  if (m != 0) {  // member: arrive at the sub level, wait for the release
    if (tid == 0) {
      st_xcd(sin + m * S, eg);
      // XCD form: the release word is kept in this XCD's L2, and one poll
      // in flight finds it sooner (lfTreeBarrUniqLocalExch 6.34 -> 6.21 us
      // per episode at 1024 x 256); the device form's release comes through
      // memory, where two in flight do
      *s_flag = f.a2a ? poll_eq(rel, eg, poll_q, g) : poll_eq_pipe(rel, eg, poll_q, g);
    }
    __syncthreads();
    return *s_flag != 0;
  }
  // sub-group leader: gather the members, flags 1 + tid + i * nthr per lane
  const int kpl = (ck - 1 + nthr - 1) / nthr;  // flags per lane, WG-uniform
  bool ok;
  if (kpl <= 1) ok = gather_eq<1>(sin, S, 1 + tid, nthr, ck, eg, poll_q, g);
  else if (kpl <= 2) ok = gather_eq<2>(sin, S, 1 + tid, nthr, ck, eg, poll_q, g);
  else if (kpl <= 4) ok = gather_eq<4>(sin, S, 1 + tid, nthr, ck, eg, poll_q, g);
  else if (kpl <= kGatherMax) ok = gather_eq<kGatherMax>(sin, S, 1 + tid, nthr, ck, eg, poll_q, g);
  else {
    ok = true;
    for (int j = 1 + tid; j < ck; j += nthr) ok = ok && poll_eq(sin + j * S, eg, poll_q, g);
  }
  if (!__syncthreads_and(ok)) return false;
  if (f.a2a) {
    // XCD leader: publish the XCD's arrival (written through: other XCDs
    // poll it), gather the other leaders', release its own XCD.  A leader
    // can publish the next episode before a slower leader has seen this
    // one (it needs only its own XCD's arrivals), never two episodes ahead,
    // so the gather accepts any epoch >= eg.
    const int nx = f.nab < f.nxcd ? f.nab : f.nxcd;
    if (tid == 0) st_rlx(f.din + (long)x * S, eg);
    ok = true;
    for (int y = tid; y < nx; y += nthr)
      if (y != x) ok = ok && poll_ge(f.din + (long)y * S, eg, poll_q, g);
    if (!__syncthreads_and(ok)) return false;
    if (tid == 0) {
      st_xcd(f.gout + (long)x * S, eg);  // stored and polled on this XCD only
      if (x == 0) count_rlx(f.episodes);
    }
    return true;
  }
  if (sid != 0) {  // arrive at the device level, wait for the release
    if (tid == 0) {
      st_rlx(f.din + sid * S, eg);
      *s_flag = poll_eq_pipe(rel, eg, poll_q, g);
    }
    __syncthreads();
    return *s_flag != 0;
  }
  // device master: gather every other sub-group leader, one flag per lane,
  // then release every XCD
  const int nsub = f.nxcd * f.sub_max;
  ok = true;
  for (int y = 1 + tid; y < nsub; y += nthr) {
    const int xy = y / f.sub_max, ky = y % f.sub_max;
    if (ky * sw < groups_on_xcd(f.nab, f.nxcd, xy) * per)  // sub-group exists
      ok = ok && poll_eq(f.din + (long)y * S, eg, poll_q, g);
  }
  if (!__syncthreads_and(ok)) return false;
  const int nx = f.nab < f.nxcd ? f.nab : f.nxcd;
  for (int y = tid; y < nx; y += nthr) st_rlx(f.gout + (long)y * S, eg);
  if (tid == 0) count_rlx(f.episodes);
  return true;
}

// ---------------------------------------------------------------------------
// Flat (single-level) barrier: the "naive" sense-reversing barriers
// (hipLocksBarrierAtomicSRB.h:19-64, 227-271).  `units` is the number of
// arrivals per episode: one per WG, or one per wave when every thread joins
// (NaiveAll: one RMW per wave instead of one per lane).  Arrivals are the
// reference's atomicInc with limit units - 1: the last arrival of an
// episode wraps the counter back to 0 in the same RMW, so there is no reset
// store for a fast WG of the next episode to overtake (it arrives only
// after it has seen the release, and finds the counter at 0).  `epoch` is
// the number of barriers of this kind completed once this one is: the last
// arriver publishes it in the generation word, and the others wait for it.
//   blk[0] arrival count, blk[32] generation.
// AGENT: the participants' data accesses are agent-scope and drained before
// they arrive, so the arrival and the generation need no fences.
template <bool AGENT = false>
__device__ __forceinline__ bool flat_arrive(uint32_t *blk, uint32_t units, uint32_t epoch,
                                            uint32_t *episodes, uint32_t poll_q, Guard &g) {
  uint32_t *cnt = blk, *gen = blk + 32;
  const uint32_t old = AGENT
      ? __builtin_amdgcn_atomic_inc32(cnt, units - 1, __ATOMIC_RELAXED, "agent")
      : __builtin_amdgcn_atomic_inc32(cnt, units - 1, __ATOMIC_ACQ_REL, "agent");
  if (old == units - 1) {
    if (AGENT) st_rlx(gen, epoch); else st_rel(gen, epoch);
    if (episodes) count_rlx(episodes);
    return true;
  }
  return AGENT ? poll_eq_pipe(gen, epoch, poll_q, g) : spin_until_eq(gen, epoch, poll_q, g);
}

}  // namespace hs
