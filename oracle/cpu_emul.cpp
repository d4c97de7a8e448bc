// cpu_emul.cpp -- std::thread + std::atomic emulation of the syncPrims
// microbenchmarks.  TEST INFRASTRUCTURE / CPU BASELINE ONLY: it is the timed
// host baseline next to the GPU numbers and a concurrent cross-check of the
// closed-form golden; the shipped library never calls it.
//
// One host thread per emulated workgroup; each thread loops over its lanes
// for the critical-section data accesses (accessData, HM:60-72).  Lock
// algorithms follow the reference's device code:
//   spin mutex      HMS:15-52      EBO mutex     HME:14-66
//   ticket (FA)     HMF:14-75      sleep (queue) HMZ:22-123
//   semaphores      HSS:32-232, HSE:32-249 (single record per semaphore)
//   atomic tree     HBA:7-181      LF tree       HBF:11-333
// (HM = hip/syncPrims/uvm/main.hip.cpp, H* = hip/syncPrims/uvm/hipLocks*.h)
// Waiting uses std::atomic::wait so that more emulated WGs than host cores
// still make progress.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "hs_oracle.h"

namespace {

constexpr int kWordsPerLine = 16;

struct alignas(128) Word {
  std::atomic<uint32_t> v{0};
};

inline void access_data(int32_t *s, long base_line, int off, int L, int mode) {
  for (int n = L - 1; n >= 0; --n) {
    long w = (base_line + n + 1) * kWordsPerLine + off;
    if (mode == 0)
      s[w] = s[(base_line + n) * kWordsPerLine + off];
    else
      s[w] += 1;
  }
}

// The whole WG's accesses: thread t of the WG covers global thread id tid.
inline void wg_access(int32_t *s, long region_tid0, int wg, int L, int mode) {
  for (int t = 0; t < wg; ++t) {
    long tid = region_tid0 + t;
    access_data(s, (tid / kWordsPerLine) * (L + 1), t % kWordsPerLine, L, mode);
  }
}

inline void backoff_spin(int n) {
  for (int i = 0; i < n; ++i) std::this_thread::yield();
}

// Generation barrier over n participants (futex wait).
struct GenBarrier {
  Word count, gen;
  uint32_t n = 1;
  void arrive(const std::function<void()> *inner = nullptr);
};


void GenBarrier::arrive(const std::function<void()> *inner) {
  uint32_t g = gen.v.load(std::memory_order_acquire);
  if (count.v.fetch_add(1, std::memory_order_acq_rel) == n - 1) {
    count.v.store(0, std::memory_order_relaxed);
    if (inner) (*inner)();
    gen.v.store(g + 1, std::memory_order_release);
    gen.v.notify_all();
  } else {
    while (gen.v.load(std::memory_order_acquire) == g) gen.v.wait(g);
  }
}

int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace

extern "C" double hso_cpu_run(int prim, int L, int num_wgs, int iters,
                              int reps, int wg, int num_cu, int mode,
                              int32_t *s, long n, int64_t *cs_count) {
  const long N = hso_num_storage_locs(prim, num_wgs, num_cu, wg, L);
  if (N < 0 || N != n) return -1.0;
  for (long i = 0; i < n; ++i) s[i] = (int32_t)i;
  const int nab = num_wgs < num_cu ? num_wgs : num_cu;
  const int perCU = ceil_div(num_wgs, nab);
  const bool srb = prim >= 37 && prim <= 42;  // sense-reversing family
  const bool barrier = prim < 4 || srb;
  const bool lx = prim == 1 || prim == 3 || prim == 38 || prim == 40 || prim == 42;
  const bool mutex = prim >= 4 && prim <= 7;
  /* global semaphores, including both priority families (16-19 spin,
   * 32-35 EBO; HM:2408-2470 dispatch both to the global kernels) */
  const bool prio = (prim >= 16 && prim <= 19) || (prim >= 32 && prim <= 35);
  const bool sem = (prim >= 8 && prim <= 19) || (prim >= 32 && prim <= 35);
  const bool umutex = prim >= 20 && prim <= 23;
  const bool usem = prim >= 24 && prim <= 31;
  if (!(barrier || mutex || sem || umutex || usem)) return -1.0;
  if (barrier && num_wgs > num_cu && num_wgs % num_cu) return -1.0;
  if (usem && num_wgs % nab) return -1.0;

  static const int kSemMax[4] = {1, 2, 10, 120};
  const int kind = mutex ? prim - 4 : umutex ? prim - 20 : 0;
  const int sem_n = (sem || usem) ? kSemMax[(prim - 8) % 4] : 0;
  const bool sem_ebo = (prim >= 12 && prim <= 15) || (usem && prim >= 28) || prim >= 32;

  // lock state: index 0 = global, 1 + c = per group c
  const int nlocks = 1 + num_cu;
  std::vector<Word> head(nlocks), tail(nlocks);
  std::vector<std::vector<Word>> ring(nlocks);
  if (kind == 2) {  // sleep mutex ring, HI:69-76 (size from the workload)
    const long total = (long)num_wgs * iters * reps + 2;
    for (int l = 0; l < (umutex ? nlocks : 1); ++l) {
      ring[l] = std::vector<Word>(total);
      ring[l][0].v.store(1);
      for (long i = 1; i < total; ++i) ring[l][i].v.store((uint32_t)-1);
    }
  }
  struct SemRec {
    Word lock, count, waiting, prio;
    uint32_t max;
  };
  std::vector<SemRec> sems(nlocks);
  for (auto &r : sems) {
    r.count.v.store(sem_n + 1);
    r.max = sem_n + 1;
    r.waiting.v.store(1);  // HSS:23: writer goes first
  }
  std::atomic<int64_t> cs{0};

  // barrier state
  GenBarrier global;
  global.n = nab;
  // SRB: flat barriers over every WG (the NaiveAll forms count threads, which
  // orders exactly the same WGs); 37/38 are a tree only above 4 WGs per group
  // (hipLocksBarrierAtomicSRB.h:186-225)
  GenBarrier flat, flat_local;
  flat.n = flat_local.n = num_wgs;
  const bool srb_tree = (prim == 37 || prim == 38) && perCU > 4;
  std::vector<GenBarrier> local(nab);
  for (int c = 0; c < nab; ++c) local[c].n = perCU;
  // LF barrier: epoch flags
  std::vector<Word> in_flag(num_wgs), out_flag(num_wgs), gin(nab), gout(nab);

  auto lf_barrier = [&](int b, uint32_t epoch) {
    const int c = b % nab, p = b / nab;
    auto wait_eq = [](Word &w, uint32_t e) {
      for (uint32_t v; (v = w.v.load(std::memory_order_acquire)) != e;) w.v.wait(v);
    };
    if (perCU > 1) {
      in_flag[b].v.store(epoch, std::memory_order_release);
      in_flag[b].v.notify_all();
      if (p == 0) {
        for (int q = 1; q < perCU; ++q) wait_eq(in_flag[q * nab + c], epoch);
      } else {
        wait_eq(out_flag[b], epoch);
        return;
      }
    }
    // group leader joins the global flag barrier (HBF:185-220)
    gin[c].v.store(epoch, std::memory_order_release);
    gin[c].v.notify_all();
    if (c == 0) {
      for (int q = 1; q < nab; ++q) wait_eq(gin[q], epoch);
      for (int q = 1; q < nab; ++q) {
        gout[q].v.store(epoch, std::memory_order_release);
        gout[q].v.notify_all();
      }
    } else {
      wait_eq(gout[c], epoch);
    }
    if (perCU > 1)
      for (int q = 1; q < perCU; ++q) {
        out_flag[q * nab + c].v.store(epoch, std::memory_order_release);
        out_flag[q * nab + c].v.notify_all();
      }
  };

  auto atomic_barrier = [&](int b) {
    const int c = b % nab;
    if (perCU > 1) {
      std::function<void()> up = [&] { global.arrive(); };
      local[c].arrive(&up);
    } else {
      global.arrive();
    }
  };

  auto worker = [&](int b) {
    if (barrier) {
      int cur = b;
      for (int rep = 0; rep < reps; ++rep)
        for (int i = 0; i < iters; ++i) {
          wg_access(s, (long)cur * wg, wg, L, mode);
          if (lx && perCU > 1) {
            // local exchange: local barrier then the neighbour's region;
            // NaiveSRB's exchange stride is the grid, i.e. its own region
            // (main.hip.cpp:472-474, 1568)
            if (prim == 40 || prim == 42)
              flat_local.arrive();
            else
              local[b % nab].arrive();
            if (prim != 40) cur = (cur + nab) % num_wgs;
            wg_access(s, (long)cur * wg, wg, L, mode);
          }
          if (prim <= 1 || (srb && srb_tree))
            atomic_barrier(b);
          else if (srb)
            flat.arrive();
          else
            lf_barrier(b, (uint32_t)(rep * iters + i + 1));
          cur = (cur + 1) % num_wgs;
        }
      return;
    }
    if (mutex || umutex) {
      const int l = umutex ? 1 + (b % num_cu) : 0;
      const long tid0 = umutex ? (long)(b % num_cu) * wg : 0;
      for (int k = 0; k < iters * reps; ++k) {
        uint32_t ticket = 0;
        if (kind == 0 || kind == 1) {  // spin / EBO: CAS 0 -> 1
          int bo = 1;
          for (uint32_t z = 0; !head[l].v.compare_exchange_weak(
                   z, 1, std::memory_order_acquire); z = 0)
            if (kind == 1) {
              backoff_spin(bo);
              bo = ((bo << 1) + 1) & 1023;
            }
        } else if (kind == 3) {  // ticket
          ticket = head[l].v.fetch_add(1, std::memory_order_relaxed);
          for (uint32_t t; (t = tail[l].v.load(std::memory_order_acquire)) != ticket;)
            tail[l].v.wait(t);
        } else {  // queue: wait for own slot
          ticket = tail[l].v.fetch_add(1, std::memory_order_relaxed);
          auto &slot = ring[l][ticket].v;
          for (uint32_t t; (t = slot.load(std::memory_order_acquire)) != 1;) slot.wait(t);
        }
        wg_access(s, tid0, wg, L, mode);
        cs.fetch_add(1, std::memory_order_relaxed);
        if (kind == 0 || kind == 1) {
          head[l].v.store(0, std::memory_order_release);
        } else if (kind == 3) {
          tail[l].v.store(ticket + 1, std::memory_order_release);
          tail[l].v.notify_all();
        } else {
          ring[l][ticket].v.store((uint32_t)-1, std::memory_order_relaxed);
          ring[l][ticket + 1].v.store(1, std::memory_order_release);
          ring[l][ticket + 1].v.notify_all();
        }
      }
      return;
    }
    // semaphores
    const int ncu = nab;
    const int c = b % ncu, p = b / ncu;
    const bool writer = (p == c % perCU);
    SemRec &r = sems[usem ? 1 + c : 0];
    for (int k = 0; k < iters * reps; ++k) {
      int bo = 1;
      for (;;) {  // TryWait, HSS:32-141
        if (prio)  // yield to posts in flight, HSS:464-469
          while (r.prio.v.load(std::memory_order_relaxed) != 0) std::this_thread::yield();
        uint32_t z = 0;
        if (!r.lock.v.compare_exchange_strong(z, 1, std::memory_order_acquire)) {
          std::this_thread::yield();
          continue;
        }
        uint32_t cnt = r.count.v.load(std::memory_order_relaxed);
        bool ok = writer ? cnt == r.max
                         : (r.waiting.v.load(std::memory_order_relaxed) == 0 && cnt > 1);
        if (ok) {
          if (writer) {
            r.count.v.store(cnt - r.max, std::memory_order_relaxed);
            r.waiting.v.store(0, std::memory_order_relaxed);
          } else {
            r.count.v.store(cnt - 1, std::memory_order_relaxed);
          }
          r.lock.v.store(0, std::memory_order_release);
          break;
        }
        if (writer) r.waiting.v.store(1, std::memory_order_relaxed);
        r.lock.v.store(0, std::memory_order_release);
        backoff_spin(sem_ebo ? bo : 1);
        if (sem_ebo) bo = writer ? ((bo << 1) + 1) & 1023 : bo + 5;
      }
      if (writer) {
        if (usem) {
          int cur = b;
          for (int j = 0; j < perCU; ++j) {
            wg_access(s, (long)cur * wg, wg, L, mode);
            cur = (cur + ncu) % num_wgs;
          }
        } else {
          for (int j = 0; j < perCU; ++j) wg_access(s, (long)j * wg, wg, L, mode);
        }
        cs.fetch_add(1, std::memory_order_relaxed);
      } else {
        // reader: loads only
        int32_t acc = 0;
        const long tid0 = usem ? (long)b * wg : (long)p * wg;
        for (int t = 0; t < wg; ++t)
          for (int q = L - 1; q >= 0; --q)
            acc += s[((tid0 + t) / kWordsPerLine * (L + 1) + q) * kWordsPerLine +
                     t % kWordsPerLine];
        volatile int32_t sink = acc;
        (void)sink;
      }
      if (prio) r.prio.v.fetch_add(1, std::memory_order_relaxed);  // HSS:595
      for (;;) {  // Post, HSS:160-232
        uint32_t z = 0;
        if (r.lock.v.compare_exchange_strong(z, 1, std::memory_order_acquire)) break;
        std::this_thread::yield();
      }
      uint32_t cnt = r.count.v.load(std::memory_order_relaxed);
      r.count.v.store(writer ? cnt + r.max : cnt + 1, std::memory_order_relaxed);
      r.lock.v.store(0, std::memory_order_release);
      if (prio) r.prio.v.fetch_sub(1, std::memory_order_relaxed);  // HSS:640
    }
  };

  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  th.reserve(num_wgs);
  for (int b = 0; b < num_wgs; ++b) th.emplace_back(worker, b);
  for (auto &t : th) t.join();
  auto t1 = std::chrono::steady_clock::now();
  if (cs_count) *cs_count = cs.load();
  return std::chrono::duration<double>(t1 - t0).count();
}

extern "C" double hso_cpu_barrier_run(int prim, int L, int num_wgs, int iters, int reps,
                                      int wg, int num_cu, int mode, int threads,
                                      int32_t *s, long n) {
  const long N = hso_num_storage_locs(prim, num_wgs, num_cu, wg, L);
  if (N < 0 || N != n || prim < 0 || prim > 3 || threads < 1) return -1.0;
  if (num_wgs > num_cu && num_wgs % num_cu) return -1.0;
  for (long i = 0; i < n; ++i) s[i] = (int32_t)i;
  const int nab = num_wgs < num_cu ? num_wgs : num_cu;
  const int perCU = ceil_div(num_wgs, nab);
  const bool lx = (prim == 1 || prim == 3) && perCU > 1;
  const int T = threads < num_wgs ? threads : num_wgs;
  GenBarrier bar;
  bar.n = (uint32_t)T;
  auto worker = [&](int t) {
    const int b0 = (int)((long)num_wgs * t / T), b1 = (int)((long)num_wgs * (t + 1) / T);
    std::vector<int> cur(b1 - b0);
    for (int b = b0; b < b1; ++b) cur[b - b0] = b;
    for (int e = 0; e < iters * reps; ++e) {
      for (int b = b0; b < b1; ++b) wg_access(s, (long)cur[b - b0] * wg, wg, L, mode);
      if (lx) {  // local barrier (here: all threads), then the neighbour's region
        bar.arrive();
        for (int b = b0; b < b1; ++b) {
          cur[b - b0] = (cur[b - b0] + nab) % num_wgs;
          wg_access(s, (long)cur[b - b0] * wg, wg, L, mode);
        }
      }
      bar.arrive();
      for (int b = b0; b < b1; ++b) cur[b - b0] = (cur[b - b0] + 1) % num_wgs;
    }
  };
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) th.emplace_back(worker, t);
  for (auto &x : th) x.join();
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}
