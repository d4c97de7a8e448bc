// hs_batch.hip -- batched persistent form: many independent mutex /
// semaphore instances (uneven WG counts, iteration counts and numLdSt) run in
// ONE persistent launch per repeat, the repeats captured in a hipGraph.
//
// Each instance keeps the plain form's semantics exactly (the kernels of
// hip/syncPrims/uvm/main.hip.cpp:645-930, 1147-1214): numWGs "virtual" WGs
// of 64 lanes, each doing numCSIters {acquire; critical-section walk;
// release}.  A fixed grid of co-resident workers of 1, 2 or 4 waves
// multiplexes the virtual WGs (static assignment: round robin, or, with the
// mutex cohort, each spin / EBO mutex instance in consecutive slots).  A
// worker never blocks on one
// virtual WG: it makes one non-blocking attempt (TTAS, ticket check, queue
// slot check or semaphore CAS) and moves on, so a spinning reader can never
// starve the writer it waits for, whatever the residency.  Backoff becomes
// "skip this virtual WG for N rounds".
// Ordering is the plain form's fence-free one: critical-section data goes
// through agent-scope accesses (L1 bypassed, stores written through), every
// lane drains them before the barrier that precedes a release, and the lock
// and semaphore words are relaxed atomics, so no acquire or release needs a
// cache fence (the fenced protocol's release wrote back up to 25 KB of dirty
// L2 per critical section).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hs_host.h"
#include "../../include/hs_detail/hs_syncprims_dev.h"

using namespace hs;

namespace {

constexpr int kMaxVpw = 64;  // virtual WGs per worker

struct BInst {
  int32_t *storage;
  uint32_t *lk;
  uint32_t *ring;
  int ring_stride, prim, nw, iters, L, nab, per, vbase;
  uint32_t sem_max;  // semaphores: the count a writer takes and returns
};

struct BArgs {
  const BInst *inst;
  int ninst, nv, vpw;
  const int *vmap;  // worker w's slot j holds virtual WG vmap[w * vpw + j] (-1: none)
  int cohort;       // spin / EBO mutexes: most critical sections per lock hold
  uint32_t *ctl;
  uint64_t ticks;
  Backoff bo;
};

enum { kIdle = 0, kWaiting = 1, kDone = 2 };

__device__ __forceinline__ int find_inst(const BInst *in, int n, int v) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (in[mid].vbase <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Critical section of one 64-thread virtual WG run by a worker of 64 * S
// lanes.  Lane t works on column t % 64 of the virtual WG (the reference's
// thread t % 64, main.hip.cpp:127-128) and on slice t / 64 of its chain's
// L lines.  S = 1: the per-lane top-down chunk walk.  S > 1, inc mode: the
// slices are disjoint line ranges, each incremented in place.  S > 1, copy
// mode: the chain moves band by band, top band first, each band 16 * S
// source lines (16 per slice); every load of a band completes before any of
// its stores (WG barrier), and a band's stores (lines lo+1..hi) never touch
// the lines the band below still has to read (lines under lo).
// Loads are agent-scope.  Stores are agent-scope (written through), or with
// KEEP plain: the lines stay dirty in this XCD's L2, where the next critical
// section on this worker finds them (agent-scope loads return a line this
// XCD's L2 holds dirty).  The last section before the instance's state goes
// to another worker stores at agent scope: an agent-scope store over a dirty
// line leaves no stale copy behind (the barrier walk's keep_l2 relies on the
// same two facts).
template <int MODE, int S, bool KEEP = false>
__device__ __forceinline__ void vwg_access(int32_t *s, long tid0, int L) {
  const int col = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const long tid = tid0 + col;
  const long base = (tid / kWordsPerLine) * (L + 1);
  const int off = col % kWordsPerLine;
  if (S == 1) {
    access_data<MODE, true, !KEEP>(s, base, off, L);
    return;
  }
  if (MODE == 1) {
    const int per = (L + S - 1) / S, lo = sl * per, hi = lo + per < L ? lo + per : L;
    if (hi > lo) access_data<1, true, !KEEP>(s, base + lo, off, hi - lo);
    return;
  }
  int32_t *c = s + base * kWordsPerLine + off;
  for (int top = L; top > 0; top -= 16 * S) {
    const int blo = top > 16 * S ? top - 16 * S : 0;
    const int lo = blo + 16 * sl;
    const int n = top - lo < 16 ? top - lo : 16;  // <= 0: no lines in this slice
    int32_t v[16];
    if (n == 16) chunk_ld<0, 16, true>(c, lo, v);
    else if (n > 0) chunk_tail<0, true, false>(c + (long)lo * kWordsPerLine, n, v);
    wait_mem();
    __syncthreads();
    if (n == 16) chunk_st<0, 16, !KEEP>(c, lo, v);
    else if (n > 0) chunk_tail<0, !KEEP, true>(c + (long)lo * kWordsPerLine, n, v);
  }
}

// Semaphore reader body (accessData_semRd, main.hip.cpp:89-100) over the
// same column / slice split: the lane's lines, loaded kReadBatch at a time,
// summed into LDS so the loads are not dead.
template <int S>
__device__ __forceinline__ void vwg_read(const int32_t *s, long tid0, int L,
                                         volatile int32_t *sink) {
  const int col = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const long tid = tid0 + col;
  const int32_t *c = s + (tid / kWordsPerLine) * (L + 1) * kWordsPerLine + col % kWordsPerLine;
  const int per = (L + S - 1) / S, lo = sl * per, hi = lo + per < L ? lo + per : L;
  int32_t acc = 0;
  for (int n0 = lo; n0 < hi; n0 += kReadBatch) {
    int32_t v[kReadBatch];
#pragma unroll
    for (int u = 0; u < kReadBatch; ++u)
      v[u] = n0 + u < hi ? ld_data<true>(c + (n0 + u) * kWordsPerLine) : 0;
#pragma unroll
    for (int u = 0; u < kReadBatch; ++u) acc += v[u];
  }
  sink[threadIdx.x] += acc;
}

// MODE 0: the reference's copy chain; MODE 1: in-place increment of lines
// 1..L, so a lost, doubled or overlapping critical section changes the
// final storage (with a copy chain every pass past the L-th leaves the same
// array).  S: waves per worker WG; all of them share each critical section.
//
// A pass over the worker's virtual WGs makes one attempt for each, in one of
// two ways (the autotuner picks):
//   PAR 0, serial: lane 0 attempts the virtual WGs one after another (a
//     test-and-test-and-set, a ticket or queue-slot check, a semaphore
//     CAS), one memory round trip per virtual WG;
//   PAR 1, probed: lane j of the first wave first probes virtual WG j's lock
//     state (one load per lane, all in flight together: one round trip for
//     the pass), taking a ticket or queue slot where the lock needs one.  A
//     ticket whose turn has come, or a queue slot holding the hand-off, is
//     the lock: that virtual WG enters.  A free lock word or an enterable
//     semaphore word is only a hint: lane 0 then makes the CAS.  Failed
//     probes of EBO instances count as failed attempts for the backoff.
// Measured on the seeded 64-instance mix (inc mode, one launch): probed
// 1.10 s with 256-lane workers against serial 1.50 s with 64-lane ones (each
// form's best), but probed is slower with 64-lane workers (2.5-3.3 s against
// 1.5-2.2 s): every pass of many narrow workers loads every lock word.
template <int MODE, int S, bool PAR>
__global__ void __launch_bounds__(64 * S) k_batch(BArgs a) {
  __shared__ int s_inst[kMaxVpw], s_b[kMaxVpw], s_iter[kMaxVpw], s_phase[kMaxVpw];
  __shared__ uint32_t s_ticket[kMaxVpw], s_skip[kMaxVpw], s_bo[kMaxVpw];
  __shared__ int s_writer[kMaxVpw];  // mutex WG, or a semaphore's writer (main.hip.cpp:883)
  __shared__ BInst s_desc[kMaxVpw];  // each slot's instance: no global load on an attempt
  __shared__ int s_remaining, s_enter, s_ok, s_hold_n;
  __shared__ unsigned long long s_hold_m;
  __shared__ unsigned long long s_cand, s_sure;
  __shared__ int32_t sink[64 * S];
  const int lane = threadIdx.x, w = blockIdx.x;
  Guard g;
  g.init(a.ctl, a.ticks);
  sink[lane] = 0;
  if (lane == 0) {
    int rem = 0;
    for (int j = 0; j < a.vpw; ++j) {
      const int v = a.vmap[w * a.vpw + j];
      if (v >= 0 && v < a.nv) {
        const int i = find_inst(a.inst, a.ninst, v);
        s_inst[j] = i;
        s_b[j] = v - a.inst[i].vbase;
        s_desc[j] = a.inst[i];
        const BInst &I0 = s_desc[j];
        s_writer[j] = I0.prim < 8 || (s_b[j] / I0.nab) == (s_b[j] % I0.nab) % I0.per;
        s_phase[j] = a.inst[i].iters > 0 ? kIdle : kDone;
        rem += a.inst[i].iters > 0;
      } else {
        s_phase[j] = kDone;
        s_writer[j] = 0;
      }
      s_iter[j] = 0;
      s_skip[j] = 0;
      s_bo[j] = a.bo.min_q;
    }
    s_remaining = rem;
    s_ok = !g.aborted();
  }
  __syncthreads();
  // the probing lane's own virtual WG, fixed for the kernel
  int my_p = -1;
  bool my_writer = false;
  uint32_t my_max = 0;
  uint32_t *my_lk = nullptr, *my_ring = nullptr;
  int my_stride = 0;
  if (lane < a.vpw && s_phase[lane] != kDone) {
    const BInst &I = a.inst[s_inst[lane]];
    my_p = I.prim;
    my_lk = I.lk;
    my_ring = I.ring;
    my_stride = I.ring_stride;
    my_max = I.sem_max;
    const int b = s_b[lane], grp = b % I.nab, pw = b / I.nab;
    my_writer = my_p < 8 || pw == grp % I.per;  // main.hip.cpp:883
  }
  while (s_remaining > 0 && s_ok) {
    // ---- probe: one lane per virtual WG (serial: every live one is a candidate)
    if (!PAR && lane < 64) {
      const unsigned long long live = __ballot(lane < a.vpw && s_phase[lane] != kDone);
      if (lane == 0) {
        s_cand = live;
        s_sure = 0;
      }
    }
    if (PAR && lane < 64) {
      int maybe = 0, sure = 0;
      bool probed = false;
      const int j = lane;
      if (my_p >= 0 && s_phase[j] != kDone) {
        probed = s_skip[j] == 0;
        if (!probed) {
          --s_skip[j];
        } else if (my_p == 4 || my_p == 5) {  // spin / EBO: the lock word is free
          maybe = ld_rlx(my_lk + kLkHead) == 0;
        } else if (my_p == 7) {  // ticket: take one, then compare the turn
          if (s_phase[j] == kIdle) {
            s_ticket[j] = add_rlx(my_lk + kLkHead, 1);
            s_phase[j] = kWaiting;
          }
          sure = ld_rlx(my_lk + kLkTail) == s_ticket[j];
        } else if (my_p == 6) {  // queue lock: take a slot, then check it
          if (s_phase[j] == kIdle) {
            s_ticket[j] = add_rlx(my_lk + kLkTail, 1);
            s_phase[j] = kWaiting;
          }
          sure = ld_rlx(my_ring + (long)s_ticket[j] * my_stride) == 1u;
        } else {  // semaphore: could this WG enter, or announce a waiting writer?
          const uint32_t st = ld_rlx(my_lk + kLkHead);
          const uint32_t cnt = st & 0xffffu;
          const bool waiting = (st & kSemWaitBit) != 0;
          maybe = my_writer ? (cnt == my_max || !waiting) : (!waiting && cnt > 1);
        }
        // a busy probe is a failed attempt for the EBO schedules
        if (probed && !maybe && !sure && (my_p == 5 || my_p >= 12)) {
          s_skip[j] = s_bo[j] >> 2;
          uint32_t nb = (my_p == 5 || my_writer) ? ((s_bo[j] << 1) + 1) & a.bo.max_q
                                                 : s_bo[j] + 5;
          s_bo[j] = nb > a.bo.max_q ? a.bo.max_q : nb;
        }
      }
      const unsigned long long mc = __ballot(maybe || sure), ms = __ballot(sure);
      if (lane == 0) {
        s_cand = mc;
        s_sure = ms;
      }
    }
    __syncthreads();
    unsigned long long cand = s_cand;
    const unsigned long long sure = s_sure;
    int progress = 0;
    while (cand) {
      // virtual WGs whose turn has come (ticket / queue lock) first: their
      // instance waits for them, whichever worker they are on
      const unsigned long long first = cand & sure;
      int j = (int)__builtin_ctzll(first ? first : cand);
      cand &= ~(1ull << j);
      if (s_phase[j] == kDone) continue;  // finished by a cohort hand-off this pass
      const BInst &I = s_desc[j];
      const int p = I.prim;
      const bool sem = p >= 8;
      const int b = s_b[j];
      const int grp = b % I.nab, pw = b / I.nab;
      const bool writer = !sem || pw == grp % I.per;  // main.hip.cpp:883
      if (lane == 0) {
        int enter = (sure >> j) & 1;
        if (PAR) {
          if (!enter && (p == 4 || p == 5)) {
            enter = cas_rlx(I.lk + kLkHead, 0, 1) == 0;
          } else if (!enter && sem) {
            const int r = sem_try_wait<true>(I.lk, writer, g);
            if (r < 0) s_ok = 0;
            enter = r > 0;
          }
        } else if (s_skip[j] > 0) {
          --s_skip[j];
        } else if (!sem) {
          uint32_t *lk = I.lk;
          if (p == 4 || p == 5) {  // spin / EBO: one test-and-test-and-set
            enter = ld_rlx(lk + kLkHead) == 0 && cas_rlx(lk + kLkHead, 0, 1) == 0;
            if (!enter && p == 5) {
              s_skip[j] = s_bo[j] >> 2;
              s_bo[j] = ((s_bo[j] << 1) + 1) & a.bo.max_q;
            }
          } else if (p == 7) {  // ticket: take one, then check the turn
            if (s_phase[j] == kIdle) {
              s_ticket[j] = add_rlx(lk + kLkHead, 1);
              s_phase[j] = kWaiting;
            }
            enter = ld_rlx(lk + kLkTail) == s_ticket[j];
          } else {  // p == 6, queue lock: take a slot, then check it
            if (s_phase[j] == kIdle) {
              s_ticket[j] = add_rlx(lk + kLkTail, 1);
              s_phase[j] = kWaiting;
            }
            enter = ld_rlx(I.ring + (long)s_ticket[j] * I.ring_stride) == 1u;
          }
        } else {
          const int r = sem_try_wait<true>(I.lk, writer, g);
          if (r < 0) s_ok = 0;
          enter = r > 0;
          if (!enter && p >= 12) {  // EBO semaphore
            s_skip[j] = s_bo[j] >> 2;
            s_bo[j] = writer ? ((s_bo[j] << 1) + 1) & a.bo.max_q : s_bo[j] + 5;
            if (s_bo[j] > a.bo.max_q) s_bo[j] = a.bo.max_q;
          }
        }
        s_enter = enter;
      }
      __syncthreads();
      if (!s_enter) continue;
      // Virtual WG j holds the lock or has entered the semaphore.  Its
      // successor is chosen before the critical section, when it is one of
      // this worker's virtual WGs of the same instance (only this worker
      // changes their state), so that the section can leave its lines in
      // this XCD's L2 for it:
      //   * spin / EBO mutex, or a semaphore writer (cohort hand-off, up to
      //     `cohort` sections per hold): a waiting WG that needs the same
      //     exclusive hold takes it over without the release and the next
      //     attempt -- the outcome of that WG winning right after the
      //     release; mutual exclusion and the per-WG iteration counts are
      //     unchanged;
      //   * ticket / queue lock: the lock is released as always, and the
      //     owner of the next ticket, if it is here, enters at once instead
      //     of at the worker's next probe (its turn has come; the order
      //     stays the lock's).
      uint32_t held = 0;  // exclusive entries of this hold, counted at its end
      // the chain stays within one instance: its descriptor is read once
      const BInst H = s_desc[j];
      if (a.cohort > 1 && (H.prim == 4 || H.prim == 5 || (H.prim >= 8 && s_writer[j] != 0))) {
        // Cohort hold.  Every critical section of a mutex instance, and
        // every writer section of a semaphore, walks the same lines, so the
        // hold runs its sections back to back: as many as this worker's
        // waiting WGs of the instance still owe, up to `cohort`, each a full
        // walk of the lines (loads and stores), all but the last leaving
        // them in this XCD's L2.  A lane touches the same words in every
        // section (copy mode over several waves excepted, which drains and
        // synchronizes between sections), so no barrier separates them.
        // The sections are then credited to those WGs round robin from j,
        // and the lock (or the semaphore's count) is released once.
        if (lane < 64) {
          const bool want = lane < a.vpw && s_phase[lane] != kDone &&
                            s_inst[lane] == s_inst[j] && s_writer[lane] != 0;
          const unsigned long long m = __ballot(want);
          if (lane == 0) {
            uint32_t owed = 0;
            for (unsigned long long r = m; r; r &= r - 1)
              owed += (uint32_t)(H.iters - s_iter[__builtin_ctzll(r)]);
            s_hold_n = (int)(owed < (uint32_t)a.cohort ? owed : (uint32_t)a.cohort);
            s_hold_m = m;
          }
        }
        __syncthreads();
        const int n = s_hold_n;
        const int nreg = H.prim >= 8 ? H.per : 1;  // a semaphore writer walks perCU regions
        for (int c = 0; c < n; ++c) {
          for (int k = 0; k < nreg; ++k) {
            if (c + 1 < n) vwg_access<MODE, S, true>(H.storage, (long)k * 64, H.L);
            else vwg_access<MODE, S, false>(H.storage, (long)k * 64, H.L);
          }
          if (MODE == 0 && S > 1) {
            wait_mem();
            __syncthreads();
          }
        }
        wait_mem();  // every lane's accesses complete before the release
        __syncthreads();
        if (lane == 0) {
          unsigned long long m = s_hold_m;
          int k = j;
          for (int c = 0; c < n; ++c) {
            s_bo[k] = a.bo.min_q;
            s_skip[k] = 0;
            if (++s_iter[k] == H.iters) {
              s_phase[k] = kDone;
              --s_remaining;
              m &= ~(1ull << k);
            } else {
              s_phase[k] = kIdle;
            }
            if (!m) break;
            const unsigned long long hi = k < 63 ? m & (~0ull << (k + 1)) : 0;
            k = (int)__builtin_ctzll(hi ? hi : m);  // the next waiting WG, cyclically
          }
          if (H.prim < 8) mutex_unlock<kSpin, true>(H.lk, nullptr, 0, 0);
          else sem_post<false, true>(H.lk, H.sem_max);
          count_rlx_n(H.lk + kLkCs, (uint32_t)n);
        }
        progress = 1;
        __syncthreads();
        continue;
      }
      for (int chain = 1;; ++chain) {
        const int hp = H.prim;
        const bool hsem = hp >= 8;
        const int hb = s_b[j];
        const int hpw = hb / H.nab;
        const bool hwriter = s_writer[j] != 0;
        const bool fifo = hp == 6 || hp == 7;
        const bool excl = hp == 4 || hp == 5 || (hsem && hwriter);
        int nxt = -1;
        if ((excl && chain < a.cohort) || fifo) {
          if (lane < 64) {
            bool want = lane < a.vpw && lane != j && s_inst[lane] == s_inst[j];
            want = want && (fifo ? s_phase[lane] == kWaiting && s_ticket[lane] == s_ticket[j] + 1
                                 : s_phase[lane] != kDone && s_writer[lane] != 0);
            const unsigned long long m = __ballot(want);
            if (lane == 0) {
              // the next one after j, cyclically
              const unsigned long long hi = j < 63 ? m & (~0ull << (j + 1)) : 0;
              s_enter = m == 0 ? -1 : (int)__builtin_ctzll(hi ? hi : m);
            }
          }
          __syncthreads();
          nxt = s_enter;
        }
        // ---- critical section of virtual WG hb of instance H ------------
        if (!hsem) {
          if (nxt >= 0) vwg_access<MODE, S, true>(H.storage, 0, H.L);
          else vwg_access<MODE, S, false>(H.storage, 0, H.L);
        } else if (hwriter) {
          for (int k = 0; k < H.per; ++k) {
            if (nxt >= 0) vwg_access<MODE, S, true>(H.storage, (long)k * 64, H.L);
            else vwg_access<MODE, S, false>(H.storage, (long)k * 64, H.L);
          }
        } else {
          vwg_read<S>(H.storage, (long)hpw * 64, H.L, sink);
        }
        // Every lane's CS accesses complete before the release.  Before a
        // section that stays on this worker the drain is only needed where
        // another wave reads the lines next (copy mode's bands over several
        // waves); elsewhere each word is next touched by the same lane,
        // after its own store.
        if (nxt < 0 || fifo || (MODE == 0 && S > 1)) wait_mem();
        __syncthreads();
        if (lane == 0) {
          uint32_t *lk = H.lk;
          // entry counts, off the hand-off path: one add per hold (a
          // returning RMW in every section would make the next section's
          // drain wait out its round trip)
          if (!hsem) {
            if (hp == 4 || hp == 5) {
              if (nxt < 0) mutex_unlock<kSpin, true>(lk, nullptr, 0, 0);
            } else if (hp == 7) {
              mutex_unlock<kFA, true>(lk, nullptr, 0, s_ticket[j]);
            } else {
              mutex_unlock<kSleep, true>(lk, H.ring, H.ring_stride, s_ticket[j]);
            }
            ++held;
            if (nxt < 0) count_rlx_n(lk + kLkCs, held);
          } else if (hwriter) {
            ++held;
            if (nxt < 0) {
              sem_post<false, true>(lk, H.sem_max);
              count_rlx_n(lk + kLkCs, held);
            }
          } else {
            sem_post<false, true>(lk, 1u);
            count_rlx_n(lk + kLkRd, 1);
          }
          s_bo[j] = a.bo.min_q;
          s_skip[j] = 0;
          s_phase[j] = ++s_iter[j] == H.iters ? kDone : kIdle;
          if (s_phase[j] == kDone) --s_remaining;
        }
        progress = 1;
        __syncthreads();
        if (nxt < 0) break;
        j = nxt;
      }
    }
    if (lane == 0 && !progress) {
      if (!g.ok()) s_ok = 0;
      sleep_quanta(a.bo.poll_q);
    }
    __syncthreads();
  }
}

__global__ void k_batch_init(uint32_t *lk, uint32_t sem_max, uint32_t *ring, long ring_len,
                             int stride) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && sem_max) {
    lk[kLkHead] = sem_init_word(sem_max);
    lk[kLkMax] = sem_max;
  }
  if (ring)
    for (long k = i; k < ring_len; k += (long)gridDim.x * blockDim.x)
      ring[k * stride] = k == 0 ? 1u : 0xffffffffu;
}

uint32_t sem_max_of(int p) {
  static const uint32_t kN[4] = {1, 2, 10, 120};
  return p >= 8 && p <= 15 ? kN[(p - 8) % 4] + 1 : 0;
}

struct Batch {
  DevArena ar;
  std::vector<BInst> h;
  std::vector<long> words;
  BInst *d_inst = nullptr;
  uint32_t *d_ctl = nullptr;
  int nv = 0;
};

int batch_prepare(hs_ctx *c, const hs_batch_instance *in, int n, int reps, Batch &B,
                  int32_t *const *storages) {
  const int ncu = c->num_cu;
  B.h.resize(n);
  B.words.resize(n);
  int rc;
  for (int i = 0; i < n; ++i) {
    const int p = in[i].prim, nw = in[i].num_wgs;
    if (p < 4 || p > 15 || nw < 1 || nw > c->max_wgs || in[i].num_cs_iters < 0 ||
        in[i].num_ldst < 0)
      return HS_EINVAL;
    BInst &I = B.h[i];
    std::memset(&I, 0, sizeof I);
    I.prim = p;
    I.nw = nw;
    I.iters = in[i].num_cs_iters;
    I.L = in[i].num_ldst;
    I.nab = std::min(nw, ncu);
    I.per = (nw + I.nab - 1) / I.nab;
    I.vbase = B.nv;
    B.nv += nw;
    B.words[i] = hs_storage_words(p, nw, ncu, 64, I.L);
    if ((rc = B.ar.alloc(&I.storage, B.words[i]))) return rc;
    if ((rc = B.ar.alloc(&I.lk, kLockWords))) return rc;
    HS_HIP(hipMemsetAsync(I.lk, 0, kLockWords * 4, c->stream));
    long ring_len = 0;
    if (p == 6) {
      ring_len = (long)nw * I.iters * reps + 1;
      I.ring_stride = ring_len <= (1L << 18) ? HS_LINE_WORDS : 1;
      if ((rc = B.ar.alloc(&I.ring, (size_t)ring_len * I.ring_stride))) return rc;
    }
    const uint32_t sm = sem_max_of(p);
    I.sem_max = sm;
    if (sm || I.ring) {
      const int blocks = (int)std::min<long>((ring_len + 255) / 256 + 1, 1024);
      k_batch_init<<<blocks, 256, 0, c->stream>>>(I.lk, sm, I.ring, ring_len, I.ring_stride);
      HS_HIP(hipGetLastError());
    }
    if (storages) {
      HS_HIP(hipMemcpyAsync(I.storage, storages[i], B.words[i] * 4, hipMemcpyHostToDevice,
                            c->stream));
    } else {
      std::vector<int32_t> init(B.words[i]);
      for (long k = 0; k < B.words[i]; ++k) init[k] = (int32_t)k;
      HS_HIP(hipMemcpyAsync(I.storage, init.data(), B.words[i] * 4, hipMemcpyHostToDevice,
                            c->stream));
      HS_HIP(hipStreamSynchronize(c->stream));
    }
  }
  if ((rc = B.ar.alloc(&B.d_inst, n))) return rc;
  if ((rc = B.ar.alloc(&B.d_ctl, 64))) return rc;
  HS_HIP(hipMemcpyAsync(B.d_inst, B.h.data(), sizeof(BInst) * n, hipMemcpyHostToDevice,
                        c->stream));
  HS_HIP(hipMemsetAsync(B.d_ctl, 0, 256, c->stream));
  HS_HIP(hipStreamSynchronize(c->stream));
  return HS_OK;
}

using BatchFn = void (*)(BArgs);

// Worker width in waves: 64, 128 or 256 lanes per worker WG (0 = 64).
int worker_waves(const hs_batch_cfg *cfg) {
  const int w = cfg->wg_size > 0 ? cfg->wg_size : 64;
  return w == 64 ? 1 : w == 128 ? 2 : w == 256 ? 4 : 0;
}

template <bool PAR>
BatchFn pick_batch_t(int mode, int waves) {
  switch (waves) {
    case 1: return mode ? k_batch<1, 1, PAR> : k_batch<0, 1, PAR>;
    case 2: return mode ? k_batch<1, 2, PAR> : k_batch<0, 2, PAR>;
    case 4: return mode ? k_batch<1, 4, PAR> : k_batch<0, 4, PAR>;
    default: return nullptr;
  }
}

BatchFn pick_batch(const hs_batch_cfg *cfg) {
  const int waves = worker_waves(cfg);
  return cfg->probe ? pick_batch_t<true>(cfg->mode, waves) : pick_batch_t<false>(cfg->mode, waves);
}

// Places the virtual WGs on the workers' slots (worker w's slot j holds
// map[w * vpw + j]): round robin (worker v % P, slot v / P), or, with the
// cohort, each mutex instance's WGs in consecutive slots, filling one
// worker after another.  A mutex serializes its instance wherever its WGs
// run, and in blocks a spin / EBO lock holder finds the instance's other
// waiters on its own worker, and a worker's WGs take consecutive tickets of
// a ticket or queue lock (one wave-wide fetch-add), so the next ticket's
// owner is usually local.  Semaphores whose readers enter one or two at a
// time go in blocks as well (their writers hand over in cohorts); the wider
// ones stay round robin over the free slots, since their readers run
// concurrently only on different workers.
// The blocks fill the workers XCD by XCD (w = x, x + nxcd, x + 2 nxcd, ...:
// the workers a round-robin dispatch puts on one XCD), so an instance's
// workers share an XCD unless it straddles the end of one XCD's slots, and
// its lock hand-offs and the lines they pass stay within one L2 and one
// XCD's path to memory.  On the seeded 64-instance mix: 311-319 ms per
// launch against 343-368 ms with the blocks over consecutive workers (one
// per XCD in turn).  Packing the instances onto the XCDs by weight
// (heaviest first, each on the least loaded XCD that holds it) ran 429-436
// ms.
std::vector<int> place_vwgs(const Batch &B, int P, int vpw, bool blocks, int nxcd) {
  std::vector<int> map((size_t)P * vpw, -1);
  if (!blocks) {
    for (int v = 0; v < B.nv; ++v) map[(size_t)(v % P) * vpw + v / P] = v;
    return map;
  }
  std::vector<int> fill(P, 0), order;
  if (nxcd < 1 || nxcd > P) nxcd = 1;
  for (int x = 0; x < nxcd; ++x)
    for (int w = x; w < P; w += nxcd) order.push_back(w);
  int wi = 0;
  // in blocks: every mutex, and the semaphores whose readers enter one or
  // two at a time (spinSem1/2, spinSemEBO1/2), whose writers hand over in
  // cohorts like the mutexes
  auto blocked = [](int p) { return p <= 7 || ((p - 8) % 4) < 2; };
  for (const BInst &I : B.h) {
    if (!blocked(I.prim)) continue;
    for (int k = 0; k < I.nw; ++k) {
      while (wi < P && fill[order[wi]] == vpw) ++wi;
      if (wi == P) return {};
      const int w = order[wi];
      map[(size_t)w * vpw + fill[w]++] = I.vbase + k;
    }
  }
  int rr = 0;
  for (const BInst &I : B.h) {
    if (blocked(I.prim)) continue;
    for (int k = 0; k < I.nw; ++k) {
      for (int tries = 0; fill[rr] == vpw; ++tries) {
        if (tries == P) return {};
        rr = (rr + 1) % P;
      }
      map[(size_t)rr * vpw + fill[rr]++] = I.vbase + k;
      rr = (rr + 1) % P;
    }
  }
  return map;
}

// Runs `reps` persistent launches (optionally as one captured graph) and
// returns the ROI in ms.
int batch_launch(hs_ctx *c, Batch &B, int n, const hs_batch_cfg *cfg, int reps, int workers,
                 double *roi_ms, double *kern_ms) {
  const int S = worker_waves(cfg);
  const BatchFn kfn = pick_batch(cfg);
  if (!kfn) return HS_EINVAL;
  BArgs a;
  a.inst = B.d_inst;
  a.ninst = n;
  a.nv = B.nv;
  const int P = std::max(1, std::min(workers, B.nv));
  a.vpw = (B.nv + P - 1) / P;
  if (a.vpw > kMaxVpw) return HS_EINVAL;
  int rc;
  {
    const std::vector<int> map = place_vwgs(B, P, a.vpw, cfg->cohort > 1, num_xcds(c));
    if (map.empty()) return HS_EINVAL;
    int *d_map = nullptr;
    if ((rc = B.ar.alloc(&d_map, map.size()))) return rc;
    HS_HIP(hipMemcpy(d_map, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
    a.vmap = d_map;
  }
  a.cohort = cfg->cohort > 1 ? cfg->cohort : 1;
  a.ctl = B.d_ctl;
  a.ticks = timeout_ticks(c, cfg->timeout_s);
  a.bo.min_q = cfg->backoff_min > 0 ? cfg->backoff_min : 1;
  a.bo.max_q = cfg->backoff_max > 0 ? cfg->backoff_max : 1023;
  a.bo.poll_q = cfg->poll_sleep < 0 ? 0 : (cfg->poll_sleep > 0 ? cfg->poll_sleep : 1);
  EventSet ev;
  if ((rc = ev.create(2))) return rc;
  struct Graph {  // destroyed on every return path
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~Graph() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
    }
  } gr;
  if (cfg->use_graph) {
    HS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(kfn, dim3(P), dim3(64 * S), 0, c->stream, a);
    HS_HIP(hipStreamEndCapture(c->stream, &gr.g));
    HS_HIP(hipGraphInstantiate(&gr.x, gr.g, nullptr, nullptr, 0));
  }
  HS_HIP(hipDeviceSynchronize());
  HS_HIP(hipEventRecord(ev[0], c->stream));
  if (cfg->use_graph) {
    HS_HIP(hipGraphLaunch(gr.x, c->stream));
  } else {
    for (int r = 0; r < reps; ++r) {
      hipLaunchKernelGGL(kfn, dim3(P), dim3(64 * S), 0, c->stream, a);
      HS_HIP(hipGetLastError());
    }
  }
  HS_HIP(hipEventRecord(ev[1], c->stream));
  HS_HIP(hipEventSynchronize(ev[1]));
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  *roi_ms = ms;
  if (kern_ms) *kern_ms = ms;
  return HS_OK;
}

// Worker WGs of the persistent grid: cfg->persistent_wgs, or 8 waves per CU
// by default, doubled while a worker would carry more than kMaxVpw virtual
// WGs.
int default_workers(const hs_ctx *c, const hs_batch_cfg *cfg, int nv) {
  const int S = std::max(1, worker_waves(cfg));
  int P = cfg->persistent_wgs > 0 ? cfg->persistent_wgs : std::max(1, 8 * c->num_cu / S);
  while ((nv + P - 1) / P > kMaxVpw) P *= 2;
  return P;
}

// Refuses a grid whose workers cannot all be resident at once.
int check_workers(hs_ctx *c, const hs_batch_cfg *cfg, int P, int nv) {
  const int S = worker_waves(cfg);
  int fit = 0, rc;
  if ((rc = coresident_wgs(c, (const void *)pick_batch(cfg), 64 * S, &fit))) return rc;
  return fit < std::min(P, nv) ? HS_EOCCUPANCY : HS_OK;
}

}  // namespace

extern "C" int hs_batch_run(hs_ctx *c, const hs_batch_instance *in, int n,
                            const hs_batch_cfg *cfg, int32_t *const *storages, int64_t *entries,
                            hs_batch_result *res) {
  if (!c || !in || n < 1 || !cfg || !storages || !entries || !res) return HS_EINVAL;
  if (worker_waves(cfg) == 0) return HS_EINVAL;  // workers of 64, 128 or 256 lanes
  if (cfg->probe != 0 && cfg->probe != 1) return HS_EINVAL;
  if (cfg->mode != 0 && cfg->mode != 1) return HS_EINVAL;
  if (cfg->cohort < 0) return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  const int reps = cfg->reps > 0 ? cfg->reps : 10;
  HS_HIP(hipSetDevice(c->device));
  Batch B;
  int rc = batch_prepare(c, in, n, reps, B, storages);
  if (rc) return rc;
  const int P = default_workers(c, cfg, B.nv);
  if ((rc = check_workers(c, cfg, P, B.nv))) return rc;
  double roi = 0, kern = 0;
  if ((rc = batch_launch(c, B, n, cfg, reps, P, &roi, &kern))) return rc;
  res->roi_ms = roi;
  res->kernel_ms = kern;
  uint32_t ctl[64];
  HS_HIP(hipMemcpy(ctl, B.d_ctl, sizeof ctl, hipMemcpyDeviceToHost));
  bool ok = true;
  for (int i = 0; i < n; ++i) {
    const BInst &I = B.h[i];
    HS_HIP(hipMemcpy(storages[i], I.storage, B.words[i] * 4, hipMemcpyDeviceToHost));
    uint32_t lk[kLockWords];
    HS_HIP(hipMemcpy(lk, I.lk, sizeof lk, hipMemcpyDeviceToHost));
    entries[i] = (int64_t)lk[kLkCs] + lk[kLkRd];
    res->acquires += entries[i];
    const uint64_t acq = (uint64_t)I.nw * I.iters * reps;
    const int p = I.prim;
    if (p == 4 || p == 5) ok = ok && lk[kLkHead] == 0;
    if (p == 7) ok = ok && lk[kLkHead] == (uint32_t)acq && lk[kLkTail] == (uint32_t)acq;
    if (p == 6) ok = ok && lk[kLkTail] == (uint32_t)acq;
    if (p >= 8)  // an instance nobody entered keeps its initial word
      ok = ok && lk[kLkHead] == (acq ? sem_max_of(p) : sem_max_of(p) | kSemWaitBit);
    ok = ok && (uint64_t)entries[i] == acq;
  }
  res->all_ok = ok && ctl[0] == 0;
  return ctl[0] ? HS_ETIMEOUT : HS_OK;
}

extern "C" int hs_batch_autotune(hs_ctx *c, const hs_batch_instance *in, int n,
                                 hs_batch_cfg *cfg, double *best_ms) {
  if (!c || !in || n < 1 || !cfg) return HS_EINVAL;
  if (cfg->mode != 0 && cfg->mode != 1) return HS_EINVAL;
  HS_HIP(hipSetDevice(c->device));
  // calibration copy: iterations capped so each trial takes milliseconds.
  // The cohort is timed on the batch itself (iterations capped at 2000, so
  // that tuning a very long batch stays bounded): its holds grow with the
  // sections each WG still owes, and no short copy ranked it right on the
  // seeded 64-instance mix (up to 1000 iterations), where a cohort of 8 ran
  // 395-436 ms per launch against 311-319 ms for 64: capped at 20 the two
  // timed alike, capped at 200 the cohort of 8 won (144-152 ms).
  std::vector<hs_batch_instance> cal(in, in + n), full(in, in + n);
  for (auto &x : cal) x.num_cs_iters = std::min(x.num_cs_iters, 20);
  for (auto &x : full) x.num_cs_iters = std::min(x.num_cs_iters, 2000);
  double best = 1e30;
  hs_batch_cfg best_cfg = *cfg;
  best_cfg.wg_size = 64;
  best_cfg.backoff_min = 1;
  best_cfg.probe = 0;
  best_cfg.cohort = 0;
  int best_wpc = 8;
  // One calibration launch of t; *ms < 0 when its workers cannot all be resident.
  auto trial_on = [&](const std::vector<hs_batch_instance> &set, const hs_batch_cfg &t,
                      double *ms) -> int {
    Batch B;
    int rc = batch_prepare(c, set.data(), n, 1, B, nullptr);
    if (rc) return rc;
    const int P = default_workers(c, &t, B.nv);
    rc = check_workers(c, &t, P, B.nv);
    if (rc == HS_EOCCUPANCY) {
      *ms = -1;
      return HS_OK;
    }
    if (rc) return rc;
    if ((rc = batch_launch(c, B, n, &t, 1, P, ms, nullptr))) return rc;
    uint32_t ab = 0;
    HS_HIP(hipMemcpy(&ab, B.d_ctl, 4, hipMemcpyDeviceToHost));
    return ab ? HS_ETIMEOUT : HS_OK;
  };
  auto trial = [&](const hs_batch_cfg &t, double *ms) { return trial_on(cal, t, ms); };
  // stage 1, 64-lane workers starting backoff at 1 quantum (the reference's
  // NUM_WIS_PER_WG and backoff start, main.hip.cpp:9, hipLocksMutexEBO.h:26):
  // waves per CU x poll interval x EBO cap
  static const int kWavesPerCu[] = {2, 4, 8, 16};
  static const int kPoll[] = {0, 1, 4};
  static const int kBackoffMax[] = {63, 255, 1023};
  for (int wpc : kWavesPerCu)
    for (int poll : kPoll)
      for (int bmax : kBackoffMax) {
        hs_batch_cfg t = best_cfg;
        t.persistent_wgs = wpc * c->num_cu;
        t.poll_sleep = poll == 0 ? -1 : poll;
        t.backoff_max = bmax;
        t.reps = 1;
        double ms = 0;
        const int rc = trial(t, &ms);
        if (rc) return rc;
        if (ms >= 0 && ms < best) {
          best = ms;
          best_cfg = t;
          best_wpc = wpc;
        }
      }
  // stage 2, at the same waves per CU: worker width (lanes sharing each
  // critical section) x backoff start I_min x attempt form (serial, probed)
  static const int kWidth[] = {64, 128, 256};
  static const int kBackoffMin[] = {1, 4, 16};
  const hs_batch_cfg base = best_cfg;
  for (int probe : {0, 1})
  for (int wd : kWidth)
    for (int bmin : kBackoffMin) {
      if (!probe && wd == 64 && bmin == 1) continue;  // stage 1's winner
      hs_batch_cfg t = base;
      t.probe = probe;
      t.wg_size = wd;
      t.persistent_wgs = std::max(1, best_wpc * c->num_cu / (wd / 64));
      t.backoff_min = bmin;
      double ms = 0;
      const int rc = trial(t, &ms);
      if (rc) return rc;
      if (ms >= 0 && ms < best) {
        best = ms;
        best_cfg = t;
      }
    }
  // last, the spin / EBO mutex cohort at the winning knobs, on the batch
  // itself
  {
    const hs_batch_cfg base2 = best_cfg;
    best = 1e30;
    for (int coh : {0, 8, 64}) {
      hs_batch_cfg t = base2;
      t.cohort = coh;
      double ms = 0;
      const int rc = trial_on(full, t, &ms);
      if (rc) return rc;
      if (ms >= 0 && ms < best) {
        best = ms;
        best_cfg = t;
      }
    }
  }
  best_cfg.reps = cfg->reps;
  *cfg = best_cfg;
  if (best_ms) *best_ms = best;
  return HS_OK;
}
