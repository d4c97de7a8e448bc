// hs_uts.hip -- unbalanced tree search (UTS) with work stealing, gfx950.
//
// Two tree families:
//   * the reference's binary tree (cuda/uts/uvm/uts.cu:290-339): host-built
//     arrays value/child_a/child_b, leaf <=> child_a == 0; results are the
//     leaf count and leaf-value sum (uts.cu:412-438);
//   * geometric SHA-1 trees (the public UTS algorithm, no implementation in
//     the reference): node state = 20-byte SHA-1 digest, root =
//     SHA1(16 zero bytes || seed), child i = SHA1(parent || i), children
//     count from the parent's digest through integer thresholds
//     (geometric, fixed shape, b0, depth limit).
//
// Engine (one wavefront per workgroup):
//   * every WG owns a ring deque in HBM; only the owner pushes and pops, so
//     the hot loop needs no atomics;
//   * a batch pops up to 64 parents (one per lane), prefix-sums their child
//     counts across the wave and then generates the children 64 at a time,
//     one SHA-1 compression per lane per round (no divergence from uneven
//     fan-out);
//   * load balance: WGs donate the oldest half of their deque to shared
//     pools while other WGs are hungry and take from pools when empty;
//   * termination and per-tree completion: credit counting.  Each tree has
//     a device counter; nodes in pools are always credited; a WG holding
//     nodes of a tree keeps at least one credit for it and returns its
//     credits when it holds no more, so the counter reaches zero exactly
//     when the tree is finished, and the WG that takes it to zero knows.
//
// The serving loop (hs_uts_serve) runs the same engine persistently over
// many trees ("slots"): requests are admitted from a host queue as slots
// and pool memory free up, and finished trees are reported back while
// other trees are still running.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "hs_host.h"
#include "hs_device.h"

using namespace hs;

namespace {

constexpr int kWave = 64;
constexpr int kMaxSlots = 64;
constexpr int kMaxThr = 128;

// ---------------------------------------------------------------------------
// SHA-1 (FIPS 180-4) of one padded 64-byte block.
__device__ __forceinline__ uint32_t rotl(uint32_t x, int n) {
  return __builtin_amdgcn_alignbit(x, x, 32 - n);
}

__device__ __forceinline__ void sha1_block(uint32_t h[5], uint32_t w[16]) {
  uint32_t a = 0x67452301u, b = 0xefcdab89u, c = 0x98badcfeu, d = 0x10325476u,
           e = 0xc3d2e1f0u;
#pragma unroll
  for (int t = 0; t < 80; ++t) {
    uint32_t wt;
    if (t < 16) {
      wt = w[t];
    } else {
      wt = rotl(w[(t - 3) & 15] ^ w[(t - 8) & 15] ^ w[(t - 14) & 15] ^ w[t & 15], 1);
      w[t & 15] = wt;
    }
    uint32_t f, k;
    if (t < 20) {
      f = (b & c) | (~b & d);
      k = 0x5a827999u;
    } else if (t < 40) {
      f = b ^ c ^ d;
      k = 0x6ed9eba1u;
    } else if (t < 60) {
      f = (b & c) | (b & d) | (c & d);
      k = 0x8f1bbcdcu;
    } else {
      f = b ^ c ^ d;
      k = 0xca62c1d6u;
    }
    const uint32_t tmp = rotl(a, 5) + f + e + k + wt;
    e = d;
    d = c;
    c = rotl(b, 30);
    b = a;
    a = tmp;
  }
  h[0] = 0x67452301u + a;
  h[1] = 0xefcdab89u + b;
  h[2] = 0x98badcfeu + c;
  h[3] = 0x10325476u + d;
  h[4] = 0xc3d2e1f0u + e;
}

// child i of a parent digest: SHA1(parent(20 B) || i(4 B, big endian))
__device__ __forceinline__ void uts_spawn(const uint32_t p[5], uint32_t i, uint32_t out[5]) {
  uint32_t w[16] = {p[0], p[1], p[2], p[3], p[4], i, 0x80000000u, 0, 0, 0, 0, 0, 0, 0, 0, 192u};
  sha1_block(out, w);
}

// root: SHA1(16 zero bytes || seed(4 B, big endian))
__device__ __forceinline__ void uts_root(uint32_t seed, uint32_t out[5]) {
  uint32_t w[16] = {0, 0, 0, 0, seed, 0x80000000u, 0, 0, 0, 0, 0, 0, 0, 0, 0, 160u};
  sha1_block(out, w);
}

// ---------------------------------------------------------------------------
// Records.  Geometric: 5 digest words + meta (depth | slot << 16).
// Reference tree: node index + meta (slot << 16).
template <int TREE>
struct Rec;
template <>
struct Rec<0> {
  static constexpr int W = 2;
};
template <>
struct Rec<1> {
  static constexpr int W = 6;
};

struct UtsArgs {
  // deques: num_blocks rings of dq_cap records
  uint32_t *deq;
  int dq_log2;
  // pools
  uint32_t *pool_rec;   // np * pc records
  uint32_t *pool_meta;  // np * 64 words: lock @0, count @32
  int np, pc;
  // per-slot device state (kMaxSlots each, 64-bit counters)
  uint32_t *credit;     // kMaxSlots * 32 words (own line each)
  unsigned long long *nodes, *leaves, *leafsum;
  uint32_t *maxd;
  uint32_t *slot_gen_mx;
  // thresholds (geometric)
  const uint32_t *thr;
  int nthr;
  // reference tree
  const uint32_t *value;
  const int32_t *child_a, *child_b;
  // control
  uint32_t *hungry;  // idle WG count
  uint32_t *ctl;     // [0] abort, [32] pool overflow
  uint64_t ticks;
  // serving (null when not serving)
  uint32_t *live;          // number of live (admitted, unfinished) trees
  uint32_t *admitted;      // requests taken from the inbox (device)
  // pinned host memory, system scope:
  const uint32_t *inbox;   // per request: slot, seed, gen_mx, pad
  const uint32_t *submitted;
  uint32_t *resp;          // per slot: 8 words (nodes lo/hi, leaves lo/hi, maxd, done seq)
  const uint32_t *stop;
  // optional per-WG statistics (null: off), kStatWords each:
  // busy ticks, idle ticks, batches, nodes, donations, donated, steals, stolen
  unsigned long long *stats;
};

constexpr int kStatWords = 8;

__device__ __forceinline__ uint32_t meta_slot(uint32_t m) { return m >> 16; }
__device__ __forceinline__ uint32_t meta_depth(uint32_t m) { return m & 0xffffu; }

// wave-wide exclusive prefix sum of v (lanes 0..63)
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t *total) {
  const int lane = threadIdx.x & 63;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  *total = __shfl(x, 63, 64);
  return x - v;
}

// Pool lock (master lane).
__device__ __forceinline__ bool pool_lock(uint32_t *lk, Guard &g) {
  for (;;) {
    if (ld_rlx(lk) == 0 && cas_acq(lk, 0, 1) == 0) return true;
    if (!g.ok()) return false;
    sleep_quanta(2);
  }
}

// Records are stored structure-of-arrays: word w of entry i of a deque or
// pool with capacity `cap` is base[w * cap + i], so lane-parallel record
// accesses are contiguous.

template <int TREE, bool SERVE>
__global__ void __launch_bounds__(64) k_uts(UtsArgs a) {
  constexpr int W = Rec<TREE>::W;
  __shared__ uint32_t s_pst[W * kWave];   // parent records of the batch
  __shared__ uint32_t s_off[kWave + 1];   // exclusive scan of child counts
  __shared__ int s_hold[kMaxSlots];       // nodes of each slot held by this WG
  __shared__ int s_cred[kMaxSlots];       // credits of each slot held by this WG
  __shared__ unsigned long long s_nodes[kMaxSlots], s_leaves[kMaxSlots], s_lsum[kMaxSlots];
  __shared__ uint32_t s_maxd[kMaxSlots];
  __shared__ uint32_t s_thr[kMaxThr];
  const int lane = threadIdx.x;
  const long dq_cap = 1L << a.dq_log2, dq_mask = dq_cap - 1;
  uint32_t *deq = a.deq + (long)blockIdx.x * dq_cap * W;
  Guard g;
  g.init(a.ctl, a.ticks);
  for (int s = lane; s < kMaxSlots; s += kWave) {
    s_hold[s] = 0;
    s_cred[s] = 0;
    s_nodes[s] = s_leaves[s] = s_lsum[s] = 0;
    s_maxd[s] = 0;
  }
  for (int k = lane; k < kMaxThr; k += kWave) s_thr[k] = k < a.nthr ? a.thr[k] : 0xffffffffu;
  long top = 0, bot = 0;  // deque holds [bot, top)
  bool hungry = false;
  uint32_t batches = 0;
  unsigned long long st_busy = 0, st_idle = 0, st_nodes = 0, st_don = 0, st_donr = 0,
                     st_steal = 0, st_stolen = 0;
  unsigned long long t_mark = a.stats ? wall_clock64() : 0;
  // folds the time since the last mark into busy or idle
  auto tick = [&](bool busy) {
    if (!a.stats) return;
    const unsigned long long now = wall_clock64();
    (busy ? st_busy : st_idle) += now - t_mark;
    t_mark = now;
  };
  auto flush_stats = [&]() {
    if (!a.stats || lane != 0) return;
    unsigned long long *o = a.stats + (long)blockIdx.x * kStatWords;
    o[0] = st_busy;
    o[1] = st_idle;
    o[2] = batches;
    o[3] = st_nodes;
    o[4] = st_don;
    o[5] = st_donr;
    o[6] = st_steal;
    o[7] = st_stolen;
  };
  const int home = blockIdx.x % a.np;
  __syncthreads();

  // Returns the slot credits / counts of every slot this WG no longer
  // holds; the WG that brings a tree's credit to zero finishes the tree.
  auto settle = [&]() {
    for (int s = lane; s < kMaxSlots; s += kWave) {
      if (s_cred[s] == 0 || s_hold[s] != 0) continue;  // still holding: keep credits
      // counts first, then the credits (release), so the finisher sees all
      if (s_nodes[s]) atomicAdd(a.nodes + s, s_nodes[s]);
      if (s_leaves[s]) atomicAdd(a.leaves + s, s_leaves[s]);
      if (s_lsum[s]) atomicAdd(a.leafsum + s, s_lsum[s]);
      if (s_maxd[s]) atomicMax(a.maxd + s, s_maxd[s]);
      s_nodes[s] = s_leaves[s] = s_lsum[s] = 0;
      s_maxd[s] = 0;
      const uint32_t c = (uint32_t)s_cred[s];
      s_cred[s] = 0;
      const uint32_t old = __hip_atomic_fetch_sub(a.credit + (long)s * HS_LINE_WORDS, c,
                                                  __ATOMIC_ACQ_REL, HS_AGENT);
      if (old == c && SERVE) {  // tree finished: this lane publishes it
        const unsigned long long nn = __hip_atomic_load(a.nodes + s, __ATOMIC_RELAXED, HS_AGENT);
        const unsigned long long ll = __hip_atomic_load(a.leaves + s, __ATOMIC_RELAXED, HS_AGENT);
        uint32_t *r = a.resp + (long)s * 8;
        const int sys = __HIP_MEMORY_SCOPE_SYSTEM;
        __hip_atomic_store(r + 0, (uint32_t)nn, __ATOMIC_RELAXED, sys);
        __hip_atomic_store(r + 1, (uint32_t)(nn >> 32), __ATOMIC_RELAXED, sys);
        __hip_atomic_store(r + 2, (uint32_t)ll, __ATOMIC_RELAXED, sys);
        __hip_atomic_store(r + 3, (uint32_t)(ll >> 32), __ATOMIC_RELAXED, sys);
        __hip_atomic_store(r + 4, ld_rlx(a.maxd + s), __ATOMIC_RELAXED, sys);
        __hip_atomic_fetch_add(r + 5, 1u, __ATOMIC_RELEASE, sys);  // completion sequence
        add_rlx(a.live, 0xffffffffu);
      }
    }
    __syncthreads();
  };

  for (;;) {
    const long held = top - bot;
    if (held > 0) {
      // ---- pop up to 64 parents (newest first: depth-first order) -------
      const int n = held < kWave ? (int)held : kWave;
      uint32_t nc = 0, meta = 0, slot = 0, dep = 0, lval = 0;
      if (lane < n) {
        const long at = (top - 1 - lane) & dq_mask;
        uint32_t r[W];
#pragma unroll
        for (int w = 0; w < W; ++w) r[w] = deq[w * dq_cap + at];
#pragma unroll
        for (int w = 0; w < W; ++w) s_pst[lane * W + w] = r[w];
        meta = r[W - 1];
        slot = meta_slot(meta);
        if (TREE == 1) {
          const uint32_t depth = meta_depth(meta);
          const uint32_t h = r[4] & 0x7fffffffu;
          if (depth < a.slot_gen_mx[slot] || depth == 0) {
            int lo = 0, hi = a.nthr;  // count thresholds <= h
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (s_thr[mid] <= h) lo = mid + 1; else hi = mid;
            }
            nc = lo;
          }
          dep = depth;
        } else {
          const int32_t ca = a.child_a[r[0]];
          nc = ca == 0 ? 0 : 2;
          if (ca == 0) lval = a.value[r[0]];
        }
      }
      // per-tree accounting: one lane when the batch is all one tree
      {
        const uint32_t slot0 = __shfl(slot, 0, 64);
        const bool active = lane < n;
        if (__all(!active || slot == slot0)) {
          const unsigned long long leafmask = __ballot(active && nc == 0);
          uint32_t dmax = active ? dep : 0;
          unsigned long long lsum = active ? lval : 0;
          int dnc = active ? (int)nc - 1 : 0;
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) {
            const uint32_t od = __shfl_xor(dmax, o, 64);
            dmax = od > dmax ? od : dmax;
            lsum += __shfl_xor(lsum, o, 64);
            dnc += __shfl_xor(dnc, o, 64);
          }
          if (lane == 0) {
            s_nodes[slot0] += (unsigned long long)n;
            s_leaves[slot0] += (unsigned long long)__popcll(leafmask);
            s_hold[slot0] += dnc;
            if (TREE == 1 && dmax > s_maxd[slot0]) s_maxd[slot0] = dmax;
            if (TREE == 0) s_lsum[slot0] += lsum;
          }
        } else if (active) {
          atomicAdd(&s_nodes[slot], 1ull);
          if (nc == 0) atomicAdd(&s_leaves[slot], 1ull);
          atomicAdd(&s_hold[slot], (int)nc - 1);
          if (TREE == 1) atomicMax(&s_maxd[slot], dep);
          if (TREE == 0 && lval) atomicAdd(&s_lsum[slot], (unsigned long long)lval);
        }
      }
      top -= n;
      uint32_t total;
      const uint32_t off = wave_excl_scan(lane < n ? nc : 0, &total);
      s_off[lane] = off;
      if (lane == 63) s_off[64] = off + (lane < n ? nc : 0);
      __syncthreads();
      if (top + total - bot > dq_cap) {  // would overflow the ring
        if (lane == 0) {
          st_rlx(a.ctl + 32, 1);
          st_rlx(a.ctl, 1);
        }
        return;
      }
      // ---- generate the children, 64 per round ---------------------------
      for (uint32_t base = 0; base < total; base += kWave) {
        const uint32_t c = base + lane;
        if (c < total) {
          int lo = 0, hi = n - 1;  // parent p: s_off[p] <= c < s_off[p+1]
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= c) lo = mid; else hi = mid - 1;
          }
          const uint32_t *pr = &s_pst[lo * W];
          const long at = (top + c) & dq_mask;
          if (TREE == 1) {
            const uint32_t i = c - s_off[lo];
            uint32_t ch[5];
            const uint32_t par[5] = {pr[0], pr[1], pr[2], pr[3], pr[4]};
            uts_spawn(par, i, ch);
#pragma unroll
            for (int w = 0; w < 5; ++w) deq[w * dq_cap + at] = ch[w];
            deq[5 * dq_cap + at] = pr[5] + 1;  // depth + 1, same slot
          } else {
            const uint32_t node = pr[0];
            deq[at] = (c - s_off[lo]) == 0 ? (uint32_t)a.child_a[node] : (uint32_t)a.child_b[node];
            deq[dq_cap + at] = pr[1];
          }
        }
      }
      top += total;
      ++batches;
      st_nodes += n;
      __syncthreads();
      // ---- donate the oldest half while others are hungry ----------------
      if ((batches & 3) == 0 && top - bot >= 2 * kWave) {
        const uint32_t hg = lane == 0 ? ld_rlx(a.hungry) : 0;
        if (__shfl(hg, 0, 64) != 0) {
          const long k0 = (top - bot) / 2;
          const int p = (home + batches / 4) % a.np;
          uint32_t *meta_p = a.pool_meta + (long)p * 64;
          __shared__ int s_k, s_base;
          if (lane == 0) {
            int k = -1, basei = 0;  // -1: lock not taken
            if (pool_lock(meta_p, g)) {
              const uint32_t cnt = ld_rlx(meta_p + 32);
              k = (int)std::min<long>(k0, (long)a.pc - cnt);
              if (k < 0) k = 0;
              basei = (int)cnt;
              st_rlx(meta_p + 32, cnt + (uint32_t)k);
            }
            s_k = k;
            s_base = basei;
          }
          __syncthreads();
          const int k = s_k;
          if (k > 0) {
            uint32_t *pool = a.pool_rec + (long)p * a.pc * W;
            for (int i = lane; i < k; i += kWave) {
              const long at = (bot + i) & dq_mask;
              const uint32_t sl = meta_slot(deq[(W - 1) * dq_cap + at]);
#pragma unroll
              for (int w = 0; w < W; ++w) pool[(long)w * a.pc + s_base + i] = deq[w * dq_cap + at];
              atomicAdd(&s_hold[sl], -1);
              atomicAdd(&s_cred[sl], -1);  // transferred to the pool (may go < 0)
            }
            bot += k;
            ++st_don;
            st_donr += k;
            __syncthreads();
            // keep >= 1 credit for every slot still held or with counts not
            // yet flushed (settle flushes them under that credit); buy the rest
            for (int s = lane; s < kMaxSlots; s += kWave) {
              const int want = (s_hold[s] > 0 || s_nodes[s] != 0) ? 1 : 0;
              if (s_cred[s] < want) {
                __hip_atomic_fetch_add(a.credit + (long)s * HS_LINE_WORDS,
                                       (uint32_t)(want - s_cred[s]), __ATOMIC_RELAXED, HS_AGENT);
                s_cred[s] = want;
              }
            }
            __syncthreads();
            if (lane == 0) {
              fence_rel();  // pool records and credits before the unlock
              st_rel(meta_p, 0);
            }
          } else if (lane == 0 && k == 0) {
            st_rel(meta_p, 0);  // pool full: nothing moved
          }
          __syncthreads();
        }
      }
      if ((batches & 63) == 0 && g.aborted()) {
        flush_stats();
        return;
      }
      tick(true);
      continue;
    }

    // ---- no local work: settle, then look for more -------------------------
    settle();
    __shared__ int s_got;
    if (lane == 0) {
      s_got = 0;
      for (int t = 0; t < a.np && !s_got; ++t) {
        const int p = (home + t) % a.np;
        uint32_t *meta_p = a.pool_meta + (long)p * 64;
        if (ld_rlx(meta_p + 32) == 0) continue;
        if (!pool_lock(meta_p, g)) break;
        const uint32_t cnt = ld_rlx(meta_p + 32);
        const int k = (int)std::min<uint32_t>(cnt, (uint32_t)(dq_cap / 4));
        if (k > 0) {
          st_rlx(meta_p + 32, cnt - (uint32_t)k);
          s_got = k;
          s_pst[0] = (uint32_t)p;
          s_pst[1] = cnt - (uint32_t)k;
        } else {
          st_rel(meta_p, 0);
        }
      }
    }
    __syncthreads();
    const int got = s_got;
    if (got > 0) {
      const int p = (int)s_pst[0];
      const long from = s_pst[1];
      const uint32_t *pool = a.pool_rec + (long)p * a.pc * W;
      __syncthreads();
      for (int i = lane; i < got; i += kWave) {
        const long at = (top + i) & dq_mask;
        uint32_t last = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          last = pool[(long)w * a.pc + from + i];
          deq[w * dq_cap + at] = last;
        }
        const uint32_t sl = meta_slot(last);
        atomicAdd(&s_hold[sl], 1);
        atomicAdd(&s_cred[sl], 1);  // pool records carry their credits
      }
      top += got;
      ++st_steal;
      st_stolen += got;
      __syncthreads();
      if (lane == 0) st_rel(a.pool_meta + (long)p * 64, 0);
      if (hungry && lane == 0) add_rlx(a.hungry, 0xffffffffu);
      hungry = false;
      continue;
    }
    if (SERVE) {  // admit the next submitted request, if any
      __shared__ int s_adm;
      if (lane == 0) {
        s_adm = -1;
        const uint32_t sub = __hip_atomic_load(a.submitted, __ATOMIC_ACQUIRE,
                                               __HIP_MEMORY_SCOPE_SYSTEM);
        const uint32_t adm = ld_rlx(a.admitted);
        if (adm < sub && cas_rlx(a.admitted, adm, adm + 1) == adm) {
          const uint32_t *q = a.inbox + (long)adm * 4;
          const int sys = __HIP_MEMORY_SCOPE_SYSTEM;
          const uint32_t slot = __hip_atomic_load(q + 0, __ATOMIC_RELAXED, sys);
          const uint32_t seed = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, sys);
          const uint32_t gmx = __hip_atomic_load(q + 2, __ATOMIC_RELAXED, sys);
          __hip_atomic_store(a.nodes + slot, 0ull, __ATOMIC_RELAXED, HS_AGENT);
          __hip_atomic_store(a.leaves + slot, 0ull, __ATOMIC_RELAXED, HS_AGENT);
          __hip_atomic_store(a.leafsum + slot, 0ull, __ATOMIC_RELAXED, HS_AGENT);
          st_rlx(a.maxd + slot, 0);
          st_rlx(a.slot_gen_mx + slot, gmx);
          st_rlx(a.credit + (long)slot * HS_LINE_WORDS, 1);
          add_rlx(a.live, 1);
          s_pst[0] = slot;
          s_pst[1] = seed;
          s_adm = 1;
        }
      }
      __syncthreads();
      if (s_adm > 0) {
        const uint32_t slot = s_pst[0];
        if (lane == 0 && TREE == 1) {
          uint32_t st[5];
          uts_root(s_pst[1], st);
          const long at = top & dq_mask;
#pragma unroll
          for (int w = 0; w < 5; ++w) deq[w * dq_cap + at] = st[w];
          deq[(W - 1) * dq_cap + at] = slot << 16;
          s_hold[slot] = 1;
          s_cred[slot] = 1;
        }
        top += 1;
        __syncthreads();
        if (hungry && lane == 0) add_rlx(a.hungry, 0xffffffffu);
        hungry = false;
        continue;
      }
    }
    if (!hungry && lane == 0) add_rlx(a.hungry, 1);
    hungry = true;
    // finished?  single tree: every credit returned; serving: host stop
    __shared__ int s_exit;
    if (lane == 0) {
      bool fin;
      if (SERVE)
        fin = __hip_atomic_load(a.stop, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) != 0 &&
              ld_rlx(a.live) == 0;
      else
        fin = ld_rlx(a.credit) == 0;
      if (!fin && !g.ok()) fin = true;
      if (!fin) sleep_quanta(8);
      s_exit = fin;
    }
    __syncthreads();
    tick(false);
    if (s_exit) {
      flush_stats();
      return;
    }
  }
}

__global__ void k_root_digest(uint32_t seed, uint32_t *out) {
  uint32_t st[5];
  uts_root(seed, st);
  for (int i = 0; i < 5; ++i) out[i] = st[i];
}

}  // namespace

// ===========================================================================
// host side
namespace {

// Thresholds: children(h) = #{k : h >= thr[k]} equals
// floor(log(1 - h/2^31) / log(1 - 1/(1+b0))) for every 31-bit h (the
// classic UTS formula, evaluated here in double precision and located by
// bisection, so the device count is the formula's count bit for bit).
int geo_thresholds(double b0, std::vector<uint32_t> &thr) {
  auto children = [&](int64_t h) {
    const double p = 1.0 / (1.0 + b0);
    const double u = (double)h / 2147483648.0;
    const int nc = (int)std::floor(std::log(1.0 - u) / std::log(1.0 - p));
    return nc > 100 ? 100 : nc;  // MAXNUMCHILDREN
  };
  thr.clear();
  for (int k = 1; k <= 100; ++k) {
    int64_t lo = 0, hi = 2147483648LL;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (children(mid) >= k) hi = mid; else lo = mid + 1;
    }
    if (lo >= 2147483648LL) break;
    thr.push_back((uint32_t)lo);
  }
  return (int)thr.size();
}

struct Engine {
  DevArena ar;
  UtsArgs a;
  int blocks = 0, W = 0;
  uint32_t *d_thr = nullptr;

  // tree: 0 reference, 1 geometric
  int setup(hs_ctx *c, int tree, int num_blocks, int dq_log2, int np, int pc, double b0) {
    std::memset(&a, 0, sizeof a);
    blocks = num_blocks > 0 ? num_blocks : 4 * c->num_cu;
    W = tree == 1 ? 6 : 2;
    a.dq_log2 = dq_log2;
    a.np = np;
    a.pc = pc;
    int rc;
    if ((rc = ar.alloc(&a.deq, ((size_t)blocks << dq_log2) * W))) return rc;
    if ((rc = ar.alloc(&a.pool_rec, (size_t)np * pc * W))) return rc;
    if ((rc = ar.alloc(&a.pool_meta, (size_t)np * 64))) return rc;
    if ((rc = ar.alloc(&a.credit, (size_t)kMaxSlots * HS_LINE_WORDS))) return rc;
    if ((rc = ar.alloc(&a.nodes, kMaxSlots))) return rc;
    if ((rc = ar.alloc(&a.leaves, kMaxSlots))) return rc;
    if ((rc = ar.alloc(&a.leafsum, kMaxSlots))) return rc;
    if ((rc = ar.alloc(&a.maxd, kMaxSlots))) return rc;
    if ((rc = ar.alloc(&a.slot_gen_mx, kMaxSlots))) return rc;
    if ((rc = ar.alloc(&d_thr, kMaxThr))) return rc;
    if ((rc = ar.alloc(&a.hungry, 64))) return rc;
    if ((rc = ar.alloc(&a.ctl, 64))) return rc;
    if ((rc = ar.alloc(&a.live, 64))) return rc;
    a.admitted = a.live + 32;
    HS_HIP(hipMemset(a.pool_meta, 0, (size_t)np * 64 * 4));
    HS_HIP(hipMemset(a.credit, 0, (size_t)kMaxSlots * HS_LINE_WORDS * 4));
    HS_HIP(hipMemset(a.nodes, 0, kMaxSlots * 8));
    HS_HIP(hipMemset(a.leaves, 0, kMaxSlots * 8));
    HS_HIP(hipMemset(a.leafsum, 0, kMaxSlots * 8));
    HS_HIP(hipMemset(a.maxd, 0, kMaxSlots * 4));
    HS_HIP(hipMemset(a.slot_gen_mx, 0, kMaxSlots * 4));
    HS_HIP(hipMemset(a.hungry, 0, 256));
    HS_HIP(hipMemset(a.ctl, 0, 256));
    HS_HIP(hipMemset(a.live, 0, 256));
    if (tree == 1) {
      std::vector<uint32_t> thr;
      a.nthr = geo_thresholds(b0, thr);
      if (a.nthr <= 0 || a.nthr > kMaxThr) return HS_EINVAL;
      HS_HIP(hipMemcpy(d_thr, thr.data(), thr.size() * 4, hipMemcpyHostToDevice));
    }
    a.thr = d_thr;
    a.ticks = timeout_ticks(c, 0);
    return HS_OK;
  }

  // Puts one record (W words) into pool 0 with its credit, as a root.
  int seed_root(const uint32_t *rec, int slot) {
    for (int w = 0; w < W; ++w)  // entry 0 of pool 0, word-major
      HS_HIP(hipMemcpy(a.pool_rec + (size_t)w * a.pc, rec + w, 4, hipMemcpyHostToDevice));
    const uint32_t one = 1;
    HS_HIP(hipMemcpy(a.pool_meta + 32, &one, 4, hipMemcpyHostToDevice));
    HS_HIP(hipMemcpy(a.credit + (size_t)slot * HS_LINE_WORDS, &one, 4, hipMemcpyHostToDevice));
    return HS_OK;
  }

  int read_slot(int slot, hs_uts_result *r) {
    unsigned long long n = 0, l = 0, ls = 0;
    uint32_t md = 0, ctl[64];
    HS_HIP(hipMemcpy(&n, a.nodes + slot, 8, hipMemcpyDeviceToHost));
    HS_HIP(hipMemcpy(&l, a.leaves + slot, 8, hipMemcpyDeviceToHost));
    HS_HIP(hipMemcpy(&ls, a.leafsum + slot, 8, hipMemcpyDeviceToHost));
    HS_HIP(hipMemcpy(&md, a.maxd + slot, 4, hipMemcpyDeviceToHost));
    HS_HIP(hipMemcpy(ctl, a.ctl, sizeof ctl, hipMemcpyDeviceToHost));
    r->nodes = n;
    r->leaves = l;
    r->leaf_sum = ls;
    r->max_depth = (int)md;
    r->aborted = ctl[0] != 0;
    if (ctl[32]) return HS_ENOMEM;  // a deque ring overflowed
    return ctl[0] ? HS_ETIMEOUT : HS_OK;
  }
};

// HS_UTS_STATS=1 in the environment: per-WG statistics of the next run,
// summarised on stderr (busy/idle time, batch fill, donations, steals).
int stats_begin(Engine &e) {
  const char *v = std::getenv("HS_UTS_STATS");
  if (!v || !*v || *v == '0') return HS_OK;
  int rc;
  if ((rc = e.ar.alloc(&e.a.stats, (size_t)e.blocks * kStatWords))) return rc;
  HS_HIP(hipMemset(e.a.stats, 0, (size_t)e.blocks * kStatWords * 8));
  return HS_OK;
}

int stats_report(Engine &e, double kernel_ms) {
  if (!e.a.stats) return HS_OK;
  std::vector<unsigned long long> st((size_t)e.blocks * kStatWords);
  HS_HIP(hipMemcpy(st.data(), e.a.stats, st.size() * 8, hipMemcpyDeviceToHost));
  unsigned long long sum[kStatWords] = {};
  std::vector<unsigned long long> nodes(e.blocks);
  for (int b = 0; b < e.blocks; ++b) {
    for (int k = 0; k < kStatWords; ++k) sum[k] += st[(size_t)b * kStatWords + k];
    nodes[b] = st[(size_t)b * kStatWords + 3];
  }
  std::sort(nodes.begin(), nodes.end());
  const double busy = (double)sum[0], idle = (double)sum[1];
  std::fprintf(stderr,
               "uts stats: %d WGs, %.3f ms: busy %.1f%% (%.3f ms/WG), batches %llu, "
               "nodes/batch %.1f, donations %llu (%.1f rec), steals %llu (%.1f rec), "
               "nodes/WG min %llu med %llu max %llu\n",
               e.blocks, kernel_ms, 100.0 * busy / std::max(1.0, busy + idle),
               busy / 1e5 / e.blocks, sum[2], (double)sum[3] / std::max(1ull, sum[2]), sum[4],
               (double)sum[5] / std::max(1ull, sum[4]), sum[6],
               (double)sum[7] / std::max(1ull, sum[6]), nodes.front(), nodes[nodes.size() / 2],
               nodes.back());
  return HS_OK;
}

template <int TREE>
int run_single(hs_ctx *c, Engine &e, hs_uts_result *res) {
  int src = stats_begin(e);
  if (src) return src;
  hipEvent_t e0, e1;
  HS_HIP(hipEventCreate(&e0));
  HS_HIP(hipEventCreate(&e1));
  HS_HIP(hipDeviceSynchronize());
  HS_HIP(hipEventRecord(e0, c->stream));
  hipLaunchKernelGGL((k_uts<TREE, false>), dim3(e.blocks), dim3(kWave), 0, c->stream, e.a);
  HS_HIP(hipGetLastError());
  HS_HIP(hipEventRecord(e1, c->stream));
  HS_HIP(hipEventSynchronize(e1));
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->kernel_ms = ms;
  if ((src = stats_report(e, ms))) return src;
  return e.read_slot(0, res);
}

}  // namespace

extern "C" int hs_uts_ref_run(hs_ctx *c, int num_nodes, const uint32_t *value,
                              const int32_t *child_a, const int32_t *child_b, int num_blocks,
                              hs_uts_result *res) {
  if (!c || !value || !child_a || !child_b || !res || num_nodes < 1 || num_blocks < 0)
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  for (int i = 0; i < num_nodes; ++i)  // binary tree, children in range
    if (child_a[i] < 0 || child_a[i] >= num_nodes || child_b[i] < 0 || child_b[i] >= num_nodes ||
        (child_a[i] == 0) != (child_b[i] == 0))
      return HS_EINVAL;
  HS_HIP(hipSetDevice(c->device));
  Engine e;
  int rc = e.setup(c, 0, num_blocks, 14, 64, std::max(1024, num_nodes), 0);
  if (rc) return rc;
  uint32_t *d_val;
  int32_t *d_a, *d_b;
  if ((rc = e.ar.alloc(&d_val, num_nodes))) return rc;
  if ((rc = e.ar.alloc(&d_a, num_nodes))) return rc;
  if ((rc = e.ar.alloc(&d_b, num_nodes))) return rc;
  HS_HIP(hipMemcpy(d_val, value, (size_t)num_nodes * 4, hipMemcpyHostToDevice));
  HS_HIP(hipMemcpy(d_a, child_a, (size_t)num_nodes * 4, hipMemcpyHostToDevice));
  HS_HIP(hipMemcpy(d_b, child_b, (size_t)num_nodes * 4, hipMemcpyHostToDevice));
  e.a.value = d_val;
  e.a.child_a = d_a;
  e.a.child_b = d_b;
  const uint32_t root[2] = {0, 0};
  if ((rc = e.seed_root(root, 0))) return rc;
  return run_single<0>(c, e, res);
}

extern "C" int hs_uts_geo_run(hs_ctx *c, int seed, double b0, int gen_mx, int num_blocks,
                              hs_uts_result *res) {
  if (!c || !res || !(b0 > 0 && b0 < 1000) || gen_mx < 0 || gen_mx > 60000 || num_blocks < 0)
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  HS_HIP(hipSetDevice(c->device));
  Engine e;
  int rc = e.setup(c, 1, num_blocks, 16, 64, 1 << 16, b0);
  if (rc) return rc;
  const uint32_t g = (uint32_t)gen_mx;
  HS_HIP(hipMemcpy(e.a.slot_gen_mx, &g, 4, hipMemcpyHostToDevice));
  // root digest: SHA1(16 zero bytes || seed) computed on the device
  uint32_t root[6];
  {
    uint32_t *d_root;
    if ((rc = e.ar.alloc(&d_root, 8))) return rc;
    hipLaunchKernelGGL(k_root_digest, dim3(1), dim3(1), 0, c->stream, (uint32_t)seed, d_root);
    HS_HIP(hipGetLastError());
    HS_HIP(hipMemcpyAsync(root, d_root, 20, hipMemcpyDeviceToHost, c->stream));
    HS_HIP(hipStreamSynchronize(c->stream));
  }
  root[5] = 0;  // depth 0, slot 0
  if ((rc = e.seed_root(root, 0))) return rc;
  return run_single<1>(c, e, res);
}

extern "C" int hs_uts_serve(hs_ctx *c, const hs_uts_request *req, int n, const hs_serve_cfg *cfg,
                            hs_uts_response *resp, hs_serve_summary *sum) {
  if (!c || !req || !resp || !sum || n < 1 || !cfg) return HS_EINVAL;
  const double b0 = req[0].b0;
  for (int i = 0; i < n; ++i)
    if (req[i].b0 != b0 || req[i].gen_mx < 0 || req[i].gen_mx > 60000 || req[i].arrival_s < 0)
      return HS_EINVAL;  // one shape (b0) per serving session
  std::memset(sum, 0, sizeof *sum);
  std::memset(resp, 0, sizeof(*resp) * n);
  const int slots = cfg->slots > 0 ? std::min(cfg->slots, kMaxSlots) : 8;
  HS_HIP(hipSetDevice(c->device));
  Engine e;
  int rc = e.setup(c, 1, cfg->num_blocks, 16, 64, 1 << 16, b0);
  if (rc) return rc;
  if (cfg->timeout_s > 0) e.a.ticks = timeout_ticks(c, cfg->timeout_s);
  // pinned, host-coherent request inbox / response board / stop word
  uint32_t *h_inbox = nullptr, *h_ctl = nullptr, *h_resp = nullptr;
  HS_HIP(hipHostMalloc((void **)&h_inbox, (size_t)n * 16, hipHostMallocCoherent | hipHostMallocMapped));
  HS_HIP(hipHostMalloc((void **)&h_ctl, 256, hipHostMallocCoherent | hipHostMallocMapped));
  HS_HIP(hipHostMalloc((void **)&h_resp, (size_t)kMaxSlots * 32, hipHostMallocCoherent | hipHostMallocMapped));
  std::memset(h_inbox, 0, (size_t)n * 16);
  std::memset(h_ctl, 0, 256);
  std::memset(h_resp, 0, (size_t)kMaxSlots * 32);
  struct Pinned {
    void *p[3];
    ~Pinned() { for (void *q : p) (void)hipHostFree(q); }
  } pinned{{h_inbox, h_ctl, h_resp}};
  volatile uint32_t *submitted = h_ctl, *stop = h_ctl + 32;
  e.a.inbox = h_inbox;
  e.a.submitted = h_ctl;
  e.a.stop = h_ctl + 32;
  e.a.resp = h_resp;

  using clk = std::chrono::steady_clock;
  std::vector<int> slot_req(slots, -1);
  std::vector<uint32_t> slot_seq(slots, 0);
  std::vector<double> lat(n, 0);
  int next = 0, done = 0, live = 0;
  const auto t0 = clk::now();
  hipLaunchKernelGGL((k_uts<1, true>), dim3(e.blocks), dim3(kWave), 0, c->stream, e.a);
  if (hipGetLastError() != hipSuccess) return HS_EHIP;
  auto now_s = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
  const double deadline = cfg->timeout_s > 0 ? cfg->timeout_s : 600.0;
  rc = HS_OK;
  while (done < n) {
    const double t = now_s();
    // admission: arrived requests go to free slots
    while (next < n && req[next].arrival_s <= t && live < slots) {
      int s = 0;
      while (slot_req[s] >= 0) ++s;
      slot_req[s] = next;
      uint32_t *q = h_inbox + (size_t)next * 4;
      q[0] = (uint32_t)s;
      q[1] = (uint32_t)req[next].seed;
      q[2] = (uint32_t)req[next].gen_mx;
      resp[next].admit_s = t;
      __atomic_store_n(submitted, (uint32_t)(next + 1), __ATOMIC_RELEASE);
      ++next;
      ++live;
    }
    // completions
    for (int s = 0; s < slots; ++s) {
      if (slot_req[s] < 0) continue;
      const uint32_t seq = __atomic_load_n(h_resp + (size_t)s * 8 + 5, __ATOMIC_ACQUIRE);
      if (seq == slot_seq[s]) continue;
      slot_seq[s] = seq;
      const int r = slot_req[s];
      const volatile uint32_t *w = h_resp + (size_t)s * 8;
      resp[r].nodes = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
      resp[r].leaves = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
      resp[r].max_depth = (int)w[4];
      resp[r].done_s = now_s();
      lat[r] = resp[r].done_s - req[r].arrival_s;
      sum->total_nodes += resp[r].nodes;
      slot_req[s] = -1;
      --live;
      ++done;
    }
    if (now_s() > deadline) {
      rc = HS_ETIMEOUT;
      break;
    }
    if (hipStreamQuery(c->stream) == hipSuccess) {  // kernel ended early (watchdog)
      if (done < n) rc = HS_ETIMEOUT;
      break;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  __atomic_store_n(stop, 1u, __ATOMIC_RELEASE);
  HS_HIP(hipStreamSynchronize(c->stream));
  sum->wall_s = now_s();
  uint32_t ctl[64];
  HS_HIP(hipMemcpy(ctl, e.a.ctl, sizeof ctl, hipMemcpyDeviceToHost));
  if (rc == HS_OK && ctl[32]) rc = HS_ENOMEM;
  if (rc == HS_OK && ctl[0]) rc = HS_ETIMEOUT;
  std::vector<double> v(lat.begin(), lat.begin() + done);
  std::sort(v.begin(), v.end());
  if (!v.empty()) {
    sum->p50_s = v[(size_t)(0.50 * (v.size() - 1) + 0.5)];
    sum->p99_s = v[(size_t)(0.99 * (v.size() - 1) + 0.5)];
  }
  return rc;
}
