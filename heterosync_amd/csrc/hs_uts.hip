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
//   * load balance by direct hand-off: an idle WG appends its id to a
//     device-wide request ring; a busy WG with at least two batches of
//     work claims the oldest request and writes the oldest half of its
//     deque into that WG's private inbox, then releases the record count
//     on the receiver's own line.  Receivers poll only their own line and
//     every request is served exactly once, by one donor;
//   * termination and per-tree completion: credit counting.  Each tree has
//     a device counter; records in flight to an inbox carry one credit
//     each; a WG holding nodes of a tree keeps at least one credit for it
//     and returns its credits when it holds no more, so the counter reaches
//     zero exactly when the tree is finished, and the WG that takes it to
//     zero knows.
//
// The serving loop (hs_uts_serve) runs the same engine persistently over
// many trees ("slots"): requests are admitted from a host queue as slots
// free up, and finished trees are reported back while other trees are
// still running.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "hs_host.h"
#include "../../include/hs_detail/hs_device.h"

using namespace hs;

namespace {

constexpr int kWave = 64;
constexpr int kMaxSlots = 64;
constexpr int kMaxThr = 128;
constexpr long kGrow = 1024;  // deque size below which a WG pops oldest first
constexpr int kDefaultDeque = 1 << 16;  // records per WG deque by default
constexpr int kMinDeque = 64, kMaxDeque = 1 << 24;

// ---------------------------------------------------------------------------
// SHA-1 (FIPS 180-4) of one padded 64-byte block.  The round functions and
// the message-schedule XORs use gfx950's three-input logic op, v_bitop3_b32,
// whose 8-bit immediate is the truth table over inputs (x, y, z) weighted
// 0xF0, 0xCC, 0xAA: Ch = 0xCA, parity = 0x96, Maj = 0xE8.
__device__ __forceinline__ uint32_t rotl(uint32_t x, int n) {
  return __builtin_amdgcn_alignbit(x, x, 32 - n);
}

__device__ __forceinline__ uint32_t lop3_ch(uint32_t x, uint32_t y, uint32_t z) {
  return __builtin_amdgcn_bitop3_b32(x, y, z, 0xCA);  // (x & y) | (~x & z)
}
__device__ __forceinline__ uint32_t lop3_xor(uint32_t x, uint32_t y, uint32_t z) {
  return __builtin_amdgcn_bitop3_b32(x, y, z, 0x96);  // x ^ y ^ z
}
__device__ __forceinline__ uint32_t lop3_maj(uint32_t x, uint32_t y, uint32_t z) {
  return __builtin_amdgcn_bitop3_b32(x, y, z, 0xE8);  // majority
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
      wt = rotl(lop3_xor(w[(t - 3) & 15], w[(t - 8) & 15], w[(t - 14) & 15]) ^ w[t & 15], 1);
      w[t & 15] = wt;
    }
    uint32_t f, k;
    if (t < 20) {
      f = lop3_ch(b, c, d);
      k = 0x5a827999u;
    } else if (t < 40) {
      f = lop3_xor(b, c, d);
      k = 0x6ed9eba1u;
    } else if (t < 60) {
      f = lop3_maj(b, c, d);
      k = 0x8f1bbcdcu;
    } else {
      f = lop3_xor(b, c, d);
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

// Child i of a parent digest p: SHA1(p (20 B) || i (4 B, big endian)), one
// padded block w = {p0..p4, i, 0x80000000, 0 x 8, 192}.  Split in two: the
// part that depends on the parent alone is computed once per parent and
// shared by its children (about 45 of the block's ~600 instructions).
//  - Rounds 0-4 read only p0..p4.  Round 5 adds w5 = i to a sum of the
//    round-4 state: P0 = rotl(a,5) + Ch(b,c,d) + e + K, so a6 = P0 + i, and
//    b6..e6 = a5, rotl(b5,30), c5, d5 (P1..P4).
//  - Of the schedule, w16, w17, w18, w20, w23 and w26 do not involve w5 (by
//    w[t] = rotl1(w[t-3] ^ w[t-8] ^ w[t-14] ^ w[t-16]) over the constant
//    words); w19 = rotl1(w16 ^ w11 ^ w5 ^ w3) is rotl1(P8 ^ i).  Every other
//    w[t], t >= 21, reads only w5..w15 and schedule words, so a child needs
//    P0..P11 and i, never p itself.
constexpr int kPre = 12;
// A parent's LDS record for its children: the prefix, then the child meta
// (depth + 1, same slot), the leaf threshold (a child is a leaf iff its draw
// is below it) and the index of its first child in the batch; 16 words, so
// a child finds it with one shift and four 16-byte loads.
constexpr int kPreStride = 16;
__device__ __forceinline__ void sha1_prefix(const uint32_t p[5], uint32_t P[kPre]) {
  uint32_t a = 0x67452301u, b = 0xefcdab89u, c = 0x98badcfeu, d = 0x10325476u,
           e = 0xc3d2e1f0u;
#pragma unroll
  for (int t = 0; t < 5; ++t) {
    const uint32_t tmp = rotl(a, 5) + lop3_ch(b, c, d) + e + 0x5a827999u + p[t];
    e = d;
    d = c;
    c = rotl(b, 30);
    b = a;
    a = tmp;
  }
  P[0] = rotl(a, 5) + lop3_ch(b, c, d) + e + 0x5a827999u;
  P[1] = a;
  P[2] = rotl(b, 30);
  P[3] = c;
  P[4] = d;
  // the message with w5 = 0: the words below do not read it
  uint32_t m[27] = {p[0], p[1], p[2], p[3], p[4], 0, 0x80000000u, 0, 0, 0, 0, 0, 0, 0, 0, 192u};
#pragma unroll
  for (int t = 16; t < 27; ++t) m[t] = rotl(lop3_xor(m[t - 3], m[t - 8], m[t - 14]) ^ m[t - 16], 1);
  P[5] = m[16];
  P[6] = m[17];
  P[7] = m[18];
  P[8] = lop3_xor(m[16], m[11], m[3]);  // w19 without its w5 term, before the rotate
  P[9] = m[20];
  P[10] = m[23];
  P[11] = m[26];
}

__device__ __forceinline__ void sha1_child(const uint32_t P[kPre], uint32_t i, uint32_t h[5]) {
  // message words 6..15 (w0..w4 are only read through P); the ring slot of
  // w[t] is t & 15
  uint32_t w[16] = {0, 0, 0, 0, 0, i, 0x80000000u, 0, 0, 0, 0, 0, 0, 0, 0, 192u};
  uint32_t a = P[0] + i, b = P[1], c = P[2], d = P[3], e = P[4];
#pragma unroll
  for (int t = 6; t < 80; ++t) {
    uint32_t wt;
    if (t < 16) {
      wt = w[t];
    } else {
      if (t == 16) wt = P[5];
      else if (t == 17) wt = P[6];
      else if (t == 18) wt = P[7];
      else if (t == 19) wt = rotl(P[8] ^ i, 1);
      else if (t == 20) wt = P[9];
      else if (t == 23) wt = P[10];
      else if (t == 26) wt = P[11];
      else wt = rotl(lop3_xor(w[(t - 3) & 15], w[(t - 8) & 15], w[(t - 14) & 15]) ^ w[t & 15], 1);
      w[t & 15] = wt;
    }
    uint32_t f, k;
    if (t < 20) {
      f = lop3_ch(b, c, d);
      k = 0x5a827999u;
    } else if (t < 40) {
      f = lop3_xor(b, c, d);
      k = 0x6ed9eba1u;
    } else if (t < 60) {
      f = lop3_maj(b, c, d);
      k = 0x8f1bbcdcu;
    } else {
      f = lop3_xor(b, c, d);
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
  // deques: num_blocks rings of 2^dq_log2 records
  uint32_t *deq;
  int dq_log2;
  // Bounded deques (set per run).  A WG never holds more than dq_cap
  // records.  A one-record-at-a-time depth-first walk from any record adds
  // at most depth_limit * (max_children - 1) records, so every WG keeps that
  // reserve free: a batch, a delivery or a breadth-first pop that would
  // leave more than dq_soft = dq_cap - reserve records is cut to what fits,
  // down to a single depth-first pop, which the reserve always covers.  No
  // deque overflows, so nothing aborts.  box_lim (<= dq_soft) caps a
  // donation, so a delivery to an empty deque always fits.
  int dq_cap, dq_soft, box_lim;
  // work requests: nrq rings; ring r has its head and tail in one 64-bit
  // word at rq[r * 32] (head low) and rq_mask + 1 entries at
  // rq_ring[r * (rq_mask + 1)] holding requester ids + 1 (0 = not yet
  // written).  WG b requests on ring b % nrq; donors visit the rings in turn.
  uint32_t *rq;
  uint32_t *rq_ring;
  uint32_t rq_mask;
  int nrq;
  // direct hand-off: per WG a record count on its own line and an inbox of
  // wbox_cap records (blocked by 64 like the deques)
  uint32_t *wbox_n;
  uint32_t *wbox;
  int wbox_cap;
  // per-slot device state (kMaxSlots each, 64-bit counters)
  uint32_t *credit;     // kMaxSlots * 32 words (own line each)
  unsigned long long *nodes, *leaves, *leafsum;
  uint32_t *maxd;
  uint32_t *slot_gen_mx;
  // thresholds (geometric)
  const uint32_t *thr;
  int nthr;
  float inv_log2q;  // 1 / log2(1 - 1/(1+b0)): the count's float estimate
  // reference tree
  const uint32_t *value;
  const int32_t *child_a, *child_b;
  // control
  uint32_t *ctl;     // [0] abort, [32] deque overflow (a bug: never set),
                     // [40] batches cut by the deque bound, [48] the most
                     // records any WG held
  uint64_t ticks;
  // serving (null when not serving)
  uint32_t *live;          // number of live (admitted, unfinished) trees
  uint32_t *admitted;      // requests taken from the inbox (device)
  // pinned host memory, system scope; only the dispatcher WG reads it:
  const uint32_t *inbox;   // per request: slot, seed, gen_mx, pad
  const uint32_t *submitted;
  uint32_t *resp;          // per slot: 8 words (nodes lo/hi, leaves lo/hi, maxd, done seq)
  const uint32_t *stop;
  // device copies kept by the dispatcher: requests (4 words each), the
  // submitted count (dev_ctl[0]) and the stop word (dev_ctl[32])
  uint32_t *dev_inbox;
  uint32_t *dev_ctl;
  // host worker threads (single-tree runs, hs_uts_*_run_cpu): client c's
  // mailbox is cpu_mail + c * kMailWords and its box cpu_box + c *
  // cpu_box_cap * W (pinned, host-coherent); WG c * cpu_stride serves it
  uint32_t *cpu_mail, *cpu_box;
  int ncpu, cpu_stride, cpu_box_cap;
  // optional per-WG statistics (null: off), kStatWords each:
  // busy ticks, idle ticks, batches, nodes, donations, donated, steals,
  // stolen, pop ticks, generate ticks, donate ticks, generate rounds
  unsigned long long *stats;
};

constexpr int kStatWords = 12;

// A host worker thread's mailbox (pinned host memory, one per thread).  The
// host sets kMailWant when idle; its serving WG, seeing kMailWant with an
// empty box, writes k records into the box and then k into kMailBox; the
// host expands them depth-first, publishes its cumulative counts (nodes,
// leaves, leaf sum, max depth) and returned credits, bumps kMailSeq, and
// clears kMailBox.  The serving WG adds each new total's increment to the
// tree's device counters and subtracts the returned credits.  Each word has
// one writer per transition, so no atomic RMW crosses the host link.
constexpr int kMailWords = 64;
constexpr int kMailWant = 0, kMailBox = 16, kMailSeq = 32;
constexpr int kMailNodes = 33, kMailLeaves = 35, kMailLsum = 37, kMailCred = 39, kMailMaxd = 40;

// Most children one record can have: the geometric count is capped by the
// threshold table (at most 100, the UTS MAXNUMCHILDREN); the reference tree
// is binary.
template <int TREE>
__device__ __forceinline__ int max_children(const UtsArgs &a) {
  return TREE == 1 ? a.nthr : 2;
}

__device__ __forceinline__ uint32_t meta_slot(uint32_t m) { return m >> 16; }
__device__ __forceinline__ uint32_t meta_depth(uint32_t m) { return m & 0xffffu; }

// wave-wide exclusive prefix sum of v (lanes 0..63)
// Wave64 inclusive scans by DPP (GFX9 data-parallel primitives, no LDS
// crossbar): a Kogge-Stone scan inside each 16-lane row (row_shr 1, 2, 4,
// 8; lanes shifted in from outside the row read the identity), then row 0's
// total (row_bcast:15) into rows 1 and 3, and rows 0-1's total
// (row_bcast:31) into rows 2 and 3.  Every lane of the wave must be active.
constexpr int kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <bool MAX>
__device__ __forceinline__ uint32_t dpp_step(uint32_t x, uint32_t y) {
  return MAX ? (x > y ? x : y) : x + y;
}
template <bool MAX>
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int id = 0;  // identity of + and of unsigned max
  x = dpp_step<MAX>(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 1, 0xF, 0xF, false));
  x = dpp_step<MAX>(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 2, 0xF, 0xF, false));
  x = dpp_step<MAX>(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 4, 0xF, 0xF, false));
  x = dpp_step<MAX>(x, __builtin_amdgcn_update_dpp(id, x, kDppRowShr + 8, 0xF, 0xF, false));
  x = dpp_step<MAX>(x, __builtin_amdgcn_update_dpp(id, x, kDppRowBcast15, 0xA, 0xF, false));
  x = dpp_step<MAX>(x, __builtin_amdgcn_update_dpp(id, x, kDppRowBcast31, 0xC, 0xF, false));
  return x;
}

// Exclusive prefix sum over the wave; *total = the sum of all lanes.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t *total) {
  const uint32_t x = wave_incl_scan<false>(v);
  *total = __builtin_amdgcn_readlane(x, 63);
  return x - v;
}

// Adds d to the held-node and credit counts of each valid lane's slot;
// one LDS update when the whole wave holds records of a single slot.
__device__ __forceinline__ void slot_count_add(int *hold, int *cred, uint32_t sl, bool valid,
                                               int d) {
  const uint32_t sl0 = __builtin_amdgcn_readlane(sl, 0);
  const unsigned long long vm = __ballot(valid);
  if (__all(!valid || sl == sl0)) {
    if ((threadIdx.x & 63) == 0 && vm) {
      const int c = __popcll(vm) * d;
      hold[sl0] += c;
      cred[sl0] += c;
    }
  } else if (valid) {
    atomicAdd(&hold[sl], d);
    atomicAdd(&cred[sl], d);
  }
}

// Deques are blocked by 64 records: the W words of record i sit in block
// i / 64, word-major inside the block, so a batch of 64 consecutive records
// is W contiguous 256-byte runs and a WG's hot region is a few KB.  Each
// WG's deque is offset by one extra line from the previous one.
constexpr long kDqSkew = 32;
template <int W>
__device__ __forceinline__ long dq_idx(long at, int w) {
  return ((at >> 6) * W + w) * 64 + (at & 63);
}

// Serving: the one WG that reads host memory.  It copies newly submitted
// requests into device memory and publishes their count, and forwards the
// host's stop word; the workers poll only these device copies, so host
// memory sees one poller however many WGs are idle.  Whole wave.
__device__ void dispatch(const UtsArgs &a, Guard &g) {
  const int lane = threadIdx.x;
  const int sys = __HIP_MEMORY_SCOPE_SYSTEM;
  uint32_t seen = 0;
  for (;;) {
    const uint32_t sub =
        __builtin_amdgcn_readlane(
            lane == 0 ? __hip_atomic_load(a.submitted, __ATOMIC_ACQUIRE, sys) : 0u, 0);
    if (sub != seen) {
      for (uint32_t i = seen + lane; i < sub; i += kWave)
        for (int w = 0; w < 3; ++w)
          st_rlx(a.dev_inbox + (long)i * 4 + w,
                 __hip_atomic_load(a.inbox + (long)i * 4 + w, __ATOMIC_RELAXED, sys));
      __syncthreads();
      if (lane == 0) st_rel(a.dev_ctl, sub);
      seen = sub;
    }
    int stop = 0;
    if (lane == 0) {
      stop = __hip_atomic_load(a.stop, __ATOMIC_ACQUIRE, sys) != 0;
      if (stop) st_rel(a.dev_ctl + 32, 1u);
      if (!g.ok()) stop = 1;
      if (!stop) sleep_quanta(16);
    }
    if (__builtin_amdgcn_readlane(stop, 0)) return;
  }
}

// The work-stealing worker (one wave per WG); the kernels below wrap it.
// STATS: per-WG timing and hand-off counters (a separate instantiation, so
// the kernel that runs by default spends no registers on them).
template <int TREE, bool SERVE, bool STATS, bool CPU = false>
__device__ __forceinline__ void uts_worker(const UtsArgs &a) {
  constexpr int W = Rec<TREE>::W;
  // parent records of the batch (the reference tree; the serving
  // dispatcher's admission uses two words of it)
  __shared__ uint32_t s_pst[TREE == 1 ? 2 : W * kWave];
  __shared__ uint32_t s_top[W * kWave];   // newest children, word-major, not in the deque
  __shared__ __align__(16) uint32_t s_pre[TREE == 1 ? kPreStride * kWave : 1];
  __shared__ uint32_t s_off[kWave + 1];   // exclusive scan of child counts
  __shared__ uint32_t s_first[kWave];     // per round: parent + 1 at its first child
  __shared__ int s_hold[kMaxSlots];       // nodes of each slot held by this WG
  __shared__ int s_cred[kMaxSlots];       // credits of each slot held by this WG
  __shared__ unsigned long long s_nodes[kMaxSlots], s_leaves[kMaxSlots], s_lsum[kMaxSlots];
  __shared__ uint32_t s_maxd[kMaxSlots];
  __shared__ uint32_t s_thr[kMaxThr];
  __shared__ uint32_t s_gmx[kMaxSlots];   // depth limit of each slot's tree
  const int lane = threadIdx.x;
  const long dq_ring = 1L << a.dq_log2, dq_mask = dq_ring - 1;
  const long dq_cap = a.dq_cap, dq_soft = a.dq_soft;
  uint32_t *deq = a.deq + (long)blockIdx.x * (dq_ring * W + kDqSkew);
  Guard g;
  g.init(a.ctl, a.ticks);
  if (SERVE && blockIdx.x == gridDim.x - 1) {  // the dispatcher, not a worker
    dispatch(a, g);
    return;
  }
  for (int s = lane; s < kMaxSlots; s += kWave) {
    s_hold[s] = 0;
    s_cred[s] = 0;
    s_nodes[s] = s_leaves[s] = s_lsum[s] = 0;
    s_maxd[s] = 0;
  }
  for (int k = lane; k < kMaxThr; k += kWave) s_thr[k] = k < a.nthr ? a.thr[k] : 0xffffffffu;
  // a slot's limit is written before its root exists; a WG gets records of a
  // slot only by admitting the tree or through a delivery (acquire), and
  // refreshes this copy on both
  for (int k = lane; k < kMaxSlots; k += kWave) s_gmx[k] = ld_rlx(a.slot_gen_mx + k);
  long top = 0, bot = 0;  // deque holds [bot, top)
  bool requested = false;  // this WG's id is in the request ring
  uint32_t batches = 0;
  uint32_t idle_q = 1;  // idle poll interval (sleep quanta), doubles to 32
  uint32_t idle_iter = 0, donate_checks = 0;
  const int me = blockIdx.x;
  uint32_t *my_n = a.wbox_n + (long)me * HS_LINE_WORDS;
  uint32_t cut_batches = 0, high_water = 0;  // deque bound: batches cut, most held
  unsigned long long st_busy = 0, st_idle = 0, st_nodes = 0, st_don = 0, st_donr = 0,
                     st_steal = 0, st_stolen = 0, st_tpop = 0, st_tgen = 0, st_tdon = 0,
                     st_rounds = 0;
  // phase clock: adds the time since the last call to `acc`
  unsigned long long t_ph = 0;
  auto phase = [&](unsigned long long &acc) {
    if (!STATS) return;
    const unsigned long long now = wall_clock64();
    acc += now - t_ph;
    t_ph = now;
  };
  unsigned long long t_mark = STATS ? wall_clock64() : 0;
  // folds the time since the last mark into busy or idle
  auto tick = [&](bool busy) {
    if (!STATS) return;
    const unsigned long long now = wall_clock64();
    (busy ? st_busy : st_idle) += now - t_mark;
    t_mark = now;
  };
  auto flush_stats = [&]() {
    if (lane == 0) {
      if (cut_batches) add_rlx(a.ctl + 40, cut_batches);
      if (high_water) __hip_atomic_fetch_max(a.ctl + 48, high_water, __ATOMIC_RELAXED, HS_AGENT);
    }
    if (!STATS || lane != 0) return;
    unsigned long long *o = a.stats + (long)blockIdx.x * kStatWords;
    o[0] = st_busy;
    o[1] = st_idle;
    o[2] = batches;
    o[3] = st_nodes;
    o[4] = st_don;
    o[5] = st_donr;
    o[6] = st_steal;
    o[7] = st_stolen;
    o[8] = st_tpop;
    o[9] = st_tgen;
    o[10] = st_tdon;
    o[11] = st_rounds;
  };
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

  // The newest lds_n (<= 64) records of the deque -- the last batch's last
  // children -- are held only in s_top: a depth-first pop takes exactly
  // them, so they never make the round trip through HBM.  flush_top() stores
  // them to their deque slots [top - lds_n, top) when the deque itself must
  // be complete (a breadth-first pop, a delivery appended at the top).
  int lds_n = 0;
  auto flush_top = [&]() {
    if (lds_n == 0) return;
    if (lane < lds_n) {
      const long at = (top - lds_n + lane) & dq_mask;
#pragma unroll
      for (int w = 0; w < W; ++w) deq[dq_idx<W>(at, w)] = s_top[w * kWave + lane];
    }
    lds_n = 0;
  };

  // Takes the k records a donor (or the host, for a root) left in this WG's
  // inbox; their credits come with them.  A delivery that would take the
  // deque past dq_soft waits in the inbox (false) until the WG has worked
  // its deque down; k <= box_lim <= dq_soft, so an empty deque always
  // takes it.  Whole wave.
  auto absorb = [&](int k) -> bool {
    if (top - bot + k > dq_soft) return false;
    flush_top();  // the delivery lands above the current top
    fence_acq();  // the donor released the records with the count
    if (SERVE)
      for (int s2 = lane; s2 < kMaxSlots; s2 += kWave) s_gmx[s2] = ld_rlx(a.slot_gen_mx + s2);
    const uint32_t *box = a.wbox + (long)me * a.wbox_cap * W;
    for (int i0 = 0; i0 < k; i0 += kWave) {
      const int i = i0 + lane;
      uint32_t sl = 0;
      if (i < k) {
        const long at = (top + i) & dq_mask;
        uint32_t v[W];
#pragma unroll
        for (int w = 0; w < W; ++w) v[w] = box[dq_idx<W>(i, w)];
#pragma unroll
        for (int w = 0; w < W; ++w) deq[dq_idx<W>(at, w)] = v[w];
        sl = meta_slot(v[W - 1]);
      }
      slot_count_add(s_hold, s_cred, sl, i < k, 1);
    }
    top += k;
    ++st_steal;
    st_stolen += k;
    idle_q = 1;
    __syncthreads();
    if (lane == 0) st_rlx(my_n, 0);  // ordered before the next request (release)
    requested = false;
    return true;
  };

  // Claims the oldest work request, if any, and hands that WG the oldest
  // half of this deque (at most one inbox).  Whole wave.
  auto donate = [&]() {
    int v = -1;
    if (lane == 0) {
      const int r = (me + (int)(donate_checks++)) & (a.nrq - 1);
      uint32_t *ctl = a.rq + (long)r * HS_LINE_WORDS;
      const uint64_t ht = __hip_atomic_load(reinterpret_cast<uint64_t *>(ctl),
                                            __ATOMIC_RELAXED, HS_AGENT);
      const uint32_t h = (uint32_t)ht, t = (uint32_t)(ht >> 32);
      if (h != t && cas_rlx(ctl, h, h + 1) == h) {
        uint32_t *ent = a.rq_ring + (long)r * (a.rq_mask + 1) + (h & a.rq_mask);
        uint32_t id;
        while ((id = ld_rlx(ent)) == 0 && g.ok()) {
        }  // the requester writes its id right after taking the slot
        if (id != 0) {
          st_rlx(ent, 0);
          v = (int)id - 1;
        }
      }
    }
    v = __builtin_amdgcn_readlane(v, 0);
    if (v < 0) return;
    const int k = (int)std::min<long>((top - bot) / 2, (long)a.box_lim);
    uint32_t *box = a.wbox + (long)v * a.wbox_cap * W;
    for (int i0 = 0; i0 < k; i0 += kWave) {
      const int i = i0 + lane;
      uint32_t sl = 0;
      if (i < k) {
        const long at = (bot + i) & dq_mask;
        uint32_t r[W];
#pragma unroll
        for (int w = 0; w < W; ++w) r[w] = deq[dq_idx<W>(at, w)];
#pragma unroll
        for (int w = 0; w < W; ++w) box[dq_idx<W>(i, w)] = r[w];
        sl = meta_slot(r[W - 1]);
      }
      // each record leaves with one credit (s_cred may go < 0 here)
      slot_count_add(s_hold, s_cred, sl, i < k, -1);
    }
    bot += k;
    ++st_don;
    st_donr += k;
    __syncthreads();
    // keep >= 1 credit for every slot still held or with counts not yet
    // flushed (settle flushes them under that credit); buy the rest
    for (int s = lane; s < kMaxSlots; s += kWave) {
      const int want = (s_hold[s] > 0 || s_nodes[s] != 0) ? 1 : 0;
      if (s_cred[s] < want) {
        __hip_atomic_fetch_add(a.credit + (long)s * HS_LINE_WORDS, (uint32_t)(want - s_cred[s]),
                               __ATOMIC_RELAXED, HS_AGENT);
        s_cred[s] = want;
      }
    }
    __syncthreads();
    if (lane == 0) st_rel(a.wbox_n + (long)v * HS_LINE_WORDS, (uint32_t)k);
  };

  // ---- host workers (CPU): the serving WG's side of the mailbox ------------
  const int cpu_client = (CPU && me % a.cpu_stride == 0) ? me / a.cpu_stride : -1;
  const bool serves_cpu = CPU && cpu_client >= 0 && cpu_client < a.ncpu;
  uint32_t *mail = serves_cpu ? a.cpu_mail + (long)cpu_client * kMailWords : nullptr;
  unsigned long long ap_nodes = 0, ap_leaves = 0, ap_lsum = 0;  // totals applied so far
  uint32_t ap_cred = 0, ap_seq = 0;
  constexpr int kSys = __HIP_MEMORY_SCOPE_SYSTEM;
  auto sys_ld = [&](const uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, kSys); };
  // Applies the host's new totals: counts first, then the credits (release),
  // as settle() does.  Whole wave; lane 0 works.
  auto cpu_collect = [&]() {
    if (!serves_cpu) return;
    if (lane == 0) {
      const uint32_t seq = sys_ld(mail + kMailSeq);
      if (seq != ap_seq) {
        auto u64 = [&](int k) {
          return (unsigned long long)__hip_atomic_load(mail + k, __ATOMIC_RELAXED, kSys) |
                 ((unsigned long long)__hip_atomic_load(mail + k + 1, __ATOMIC_RELAXED, kSys)
                  << 32);
        };
        const unsigned long long nn = u64(kMailNodes), ll = u64(kMailLeaves), ls = u64(kMailLsum);
        const uint32_t cr = __hip_atomic_load(mail + kMailCred, __ATOMIC_RELAXED, kSys);
        const uint32_t md = __hip_atomic_load(mail + kMailMaxd, __ATOMIC_RELAXED, kSys);
        if (nn != ap_nodes) atomicAdd(a.nodes, nn - ap_nodes);
        if (ll != ap_leaves) atomicAdd(a.leaves, ll - ap_leaves);
        if (ls != ap_lsum) atomicAdd(a.leafsum, ls - ap_lsum);
        if (md) atomicMax(a.maxd, md);
        if (cr != ap_cred)
          __hip_atomic_fetch_sub(a.credit, cr - ap_cred, __ATOMIC_ACQ_REL, HS_AGENT);
        ap_nodes = nn;
        ap_leaves = ll;
        ap_lsum = ls;
        ap_cred = cr;
        ap_seq = seq;
      }
    }
  };
  // Hands the host worker up to 64 of the newest records when it asks and
  // its box is empty; they leave with one credit each.  A host core expands
  // about a million nodes a second, and the tree cannot finish while it
  // holds records, so it gets only small subtrees: geometric records one
  // level above the depth limit (their children are leaves), and at most 16
  // reference-tree records (newest, so deepest).  Whole wave.
  auto cpu_deliver = [&]() {
    if (!serves_cpu) return;
    int k = 0;
    if (lane == 0 && sys_ld(mail + kMailWant) != 0 && sys_ld(mail + kMailBox) == 0)
      k = (int)std::min<long>(std::min<long>(top - bot - 2 * kWave, TREE == 1 ? kWave : 16),
                              (long)a.cpu_box_cap);
    k = __builtin_amdgcn_readlane(k, 0);
    if (k <= 0) return;
    flush_top();  // the newest records go; the deque must hold them
    if (TREE == 1) {  // the newest run of records at depth >= limit - 1
      const uint32_t lim = s_gmx[0];
      const bool ok_rec =
          lane < k && meta_depth(deq[dq_idx<W>((top - 1 - lane) & dq_mask, W - 1)]) + 1 >= lim;
      const unsigned long long bad = __ballot(lane < k && !ok_rec);
      if (bad) k = (int)__builtin_ctzll(bad);
      if (k <= 0) return;
    }
    uint32_t *box = a.cpu_box + (long)cpu_client * a.cpu_box_cap * W;
    uint32_t sl = 0;
    const bool mine = lane < k;
    if (mine) {
      const long at = (top - k + lane) & dq_mask;
      uint32_t r[W];
#pragma unroll
      for (int w = 0; w < W; ++w) r[w] = deq[dq_idx<W>(at, w)];
#pragma unroll
      for (int w = 0; w < W; ++w)
        __hip_atomic_store(box + dq_idx<W>(lane, w), r[w], __ATOMIC_RELAXED, kSys);
      sl = meta_slot(r[W - 1]);
    }
    slot_count_add(s_hold, s_cred, sl, mine, -1);
    top -= k;
    __syncthreads();
    for (int s2 = lane; s2 < kMaxSlots; s2 += kWave) {  // as donate(): keep >= 1 credit
      const int want = (s_hold[s2] > 0 || s_nodes[s2] != 0) ? 1 : 0;
      if (s_cred[s2] < want) {
        __hip_atomic_fetch_add(a.credit + (long)s2 * HS_LINE_WORDS,
                               (uint32_t)(want - s_cred[s2]), __ATOMIC_RELAXED, HS_AGENT);
        s_cred[s2] = want;
      }
    }
    __syncthreads();
    // the records (this wave's stores), then the count: a system-scope release
    if (lane == 0) __hip_atomic_store(mail + kMailBox, (uint32_t)k, __ATOMIC_RELEASE, kSys);
  };

  for (;;) {
    const long held = top - bot;
    if (held > 0) {
      if (STATS) t_ph = wall_clock64();
      // a delivery for a request still outstanding, read now, used after
      const uint32_t box_n = (lane == 0 && requested) ? ld_rlx(my_n) : 0;
      // ---- pop up to 64 parents ------------------------------------------
      // Newest first (depth-first, bounded memory) once the deque holds a
      // donatable frontier; oldest first (breadth-first) while it is
      // smaller, so that it grows into one: the frontier is what donations
      // are cut from.
      int n = held < kWave ? (int)held : kWave;
      // breadth-first only while a whole batch of the widest fan-out stays
      // under the deque bound, so a breadth-first batch is never cut
      const bool bfs = held < kGrow && held + (long)kWave * max_children<TREE>(a) <= dq_soft;
      if (bfs) flush_top();  // pops from the bottom; this batch refills s_top
      uint32_t nc = 0, meta = 0, slot = 0, dep = 0, lval = 0, cthr = 0;
      if (lane < n) {
        const long at = (bfs ? bot + lane : top - 1 - lane) & dq_mask;
        uint32_t r[W];
        if (lane < lds_n) {  // depth-first: the i-th newest is s_top slot lds_n-1-i
#pragma unroll
          for (int w = 0; w < W; ++w) r[w] = s_top[w * kWave + (lds_n - 1 - lane)];
        } else {
#pragma unroll
          for (int w = 0; w < W; ++w) r[w] = deq[dq_idx<W>(at, w)];
        }
        if (TREE == 1) {  // the children read only the prefix record
          uint32_t P[kPre];
          sha1_prefix(r, P);
#pragma unroll
          for (int j = 0; j < kPre; ++j) s_pre[lane * kPreStride + j] = P[j];
        } else {
#pragma unroll
          for (int w = 0; w < W; ++w) s_pst[lane * W + w] = r[w];
        }
        meta = r[W - 1];
        slot = meta_slot(meta);
        if (TREE == 1) {
          const uint32_t depth = meta_depth(meta);
          const uint32_t h = r[4] & 0x7fffffffu;
          if (depth < s_gmx[slot] || depth == 0) {
            // #thresholds <= h: start from the closed form
            // floor(log(1 - h/2^31) / log(1 - p)) in float, then step to
            // the exact count on the integer table (usually 0-1 steps)
            const float u = (float)h * (1.0f / 2147483648.0f);
            const float est = __log2f(1.0f - u) * a.inv_log2q;
            int k = est < 0.0f ? 0 : (est > (float)a.nthr ? a.nthr : (int)est);
            while (k > 0 && s_thr[k - 1] > h) --k;
            while (k < a.nthr && s_thr[k] <= h) ++k;
            nc = k;
          }
          dep = depth;
          // a child at the depth limit is a leaf whatever its draw, and no
          // draw (< 2^31) reaches 0xffffffff
          cthr = depth + 1 < s_gmx[slot] ? s_thr[0] : 0xffffffffu;
        } else {
          const int32_t ca = a.child_a[r[0]];
          nc = ca == 0 ? 0 : 2;
          if (ca == 0) lval = a.value[r[0]];
        }
      }
      // child offsets (exclusive scan) and the batch total
      uint32_t total;
      uint32_t off = wave_excl_scan(lane < n ? nc : 0, &total);
      // The children are generated one per lane in rounds of 64, and a
      // round costs the same whatever number of lanes it fills.  A batch
      // whose children overflow a multiple of 64 is cut back to the parents
      // whose children fill the whole rounds; the others stay in the deque
      // for the next batch.  The last round of a batch ran 17 % idle on
      // average before (1.20 rounds per batch of 63.9 nodes, depth-15 tree)
      // and the tree took 185.0 ms; cut, 178.6 ms (depth 13: 19.7 -> 18.4).
      const bool bound_cut = !bfs && held - n + (long)total > dq_soft;
      const bool fill = total > (uint32_t)kWave && (total & (kWave - 1)) != 0;
      if (bound_cut || fill) {
        // Deque bound: pop only the newest parents whose children keep the
        // deque at or under dq_soft (a prefix of the lanes), at least one.
        // A single depth-first pop always fits: every record's subtree,
        // walked one record at a time, fits in the reserve above dq_soft.
        // Both cuts take the first lane whose prefix breaks either, so the
        // prefix kept also keeps the deque bound.
        const uint32_t lim = total & ~(uint32_t)(kWave - 1);
        const bool over = lane < n && ((!bfs && held - (lane + 1) + (long)(off + nc) > dq_soft) ||
                                       (fill && off + nc > lim));
        const unsigned long long bad = __ballot(over);
        int fit = bad ? (int)__builtin_ctzll(bad) : n;
        if (fit == 0) fit = 1;
        // s_top records this batch no longer pops go back to their deque
        // slots [top - lds_n, top - fit)
        if (lane < lds_n - fit) {
          const long at = (top - lds_n + lane) & dq_mask;
#pragma unroll
          for (int w = 0; w < W; ++w) deq[dq_idx<W>(at, w)] = s_top[w * kWave + lane];
        }
        n = fit;
        off = wave_excl_scan(lane < n ? nc : 0, &total);
        if (bound_cut) ++cut_batches;
      }
      if (TREE == 1) {
        s_pre[lane * kPreStride + kPre] = meta + 1;
        s_pre[lane * kPreStride + kPre + 1] = cthr;
        s_pre[lane * kPreStride + kPre + 2] = off;
        s_pre[lane * kPreStride + kPre + 3] = 0;
      } else {
        s_off[lane] = off;
        if (lane == 63) s_off[64] = off + (lane < n ? nc : 0);
      }
      // per-tree accounting: one lane when the batch is all one tree.  The
      // popped parents leave the held count; the reference tree's children
      // enter it here (total), a geometric tree's after they are generated
      // (only the ones that are not leaves are stored and held).  A child
      // is one deeper than its parent, so a geometric batch's depth maximum
      // takes dep + 1 for every parent with children: its leaves are covered
      // without looking at them, the others are counted again when popped.
      const uint32_t slot0 = __builtin_amdgcn_readlane(slot, 0);
      const bool batch_uni = __all(lane >= n || slot == slot0);
      {
        const bool active = lane < n;
        if (batch_uni) {
          const unsigned long long leafmask = __ballot(active && nc == 0);
          uint32_t dmax = 0;
          unsigned long long lsum = active ? lval : 0;
          if (TREE == 1) {
            dmax = __builtin_amdgcn_readlane(
                wave_incl_scan<true>(active ? dep + (nc > 0 ? 1u : 0u) : 0u), 63);
          } else {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
          }
          if (lane == 0) {
            s_nodes[slot0] += (unsigned long long)n;
            s_leaves[slot0] += (unsigned long long)__popcll(leafmask);
            s_hold[slot0] += (TREE == 1 ? 0 : (int)total) - n;
            if (TREE == 1 && dmax > s_maxd[slot0]) s_maxd[slot0] = dmax;
            if (TREE == 0) s_lsum[slot0] += lsum;
          }
        } else if (active) {
          atomicAdd(&s_nodes[slot], 1ull);
          if (nc == 0) atomicAdd(&s_leaves[slot], 1ull);
          atomicAdd(&s_hold[slot], (TREE == 1 ? 0 : (int)nc) - 1);
          if (TREE == 1) atomicMax(&s_maxd[slot], dep + (nc > 0 ? 1u : 0u));
          if (TREE == 0 && lval) atomicAdd(&s_lsum[slot], (unsigned long long)lval);
        }
      }
      if (bfs) bot += n; else top -= n;
      lds_n = 0;  // popped, or written back to the deque above
      __syncthreads();
      if (top + total - bot > dq_cap) {  // the bound failed: never taken
        if (lane == 0) {
          st_rlx(a.ctl + 32, 1);
          st_rlx(a.ctl, 1);
        }
        return;
      }
      phase(st_tpop);
      // ---- generate the children, 64 per round ---------------------------
      // Reference tree: the newest min(total, 64) children stay in s_top
      // (slot c - keep0), the others go to the deque.  Geometric tree: a
      // child's own child count is known from its digest as soon as it is
      // generated, and a leaf (at the depth limit, or drawn below the first
      // threshold: count 0) is counted at once and never stored, so the
      // deque and the batches hold only records with children (about a
      // fifth of the nodes at b0 = 4).  The stored children are packed: the
      // first 64 into s_top, the rest into the deque under them (their
      // order within a batch does not matter), s_top slot k being deque
      // position top - lds_n + k as always.
      const uint32_t keep = total < (uint32_t)kWave ? total : (uint32_t)kWave;
      const uint32_t keep0 = total - keep;
      uint32_t carry = 0;  // parent of the last child of the previous round
      uint32_t stored = 0;  // geometric: children stored so far (wave-uniform)
      for (uint32_t base = 0; base < total; base += kWave) {
        ++st_rounds;
        // Parent of child c: the last parent with children whose first child
        // is <= c.  Each such parent marks its first child's position in this
        // round (parents with children start at distinct offsets), and an
        // inclusive max-scan carries the marks forward; positions before the
        // round's first mark belong to the previous round's last parent.
        s_first[lane] = 0;
        __syncthreads();
        if (lane < n && nc > 0 && off >= base && off < base + (uint32_t)kWave)
          s_first[off - base] = (uint32_t)lane + 1;
        __syncthreads();
        const uint32_t mark = wave_incl_scan<true>(s_first[lane]);
        const int lo = mark ? (int)mark - 1 : (int)carry;
        carry = (uint32_t)__builtin_amdgcn_readlane(lo, 63);
        __syncthreads();  // s_first is read before the next round clears it
        const uint32_t c = base + lane;
        const bool valid = c < total;
        if (TREE == 1) {
          // every lane hashes (lo is a parent of this batch either way; a
          // lane past the batch's children hashes a value nobody keeps)
          uint32_t P[kPreStride];
          const uint4 *pp = reinterpret_cast<const uint4 *>(&s_pre[lo * kPreStride]);
#pragma unroll
          for (int j = 0; j < kPreStride / 4; ++j) {
            const uint4 t = pp[j];
            P[4 * j] = t.x;
            P[4 * j + 1] = t.y;
            P[4 * j + 2] = t.z;
            P[4 * j + 3] = t.w;
          }
          uint32_t ch[5];
          sha1_child(P, c - P[kPre + 2], ch);
          const uint32_t cmeta = P[kPre];
          const bool inner = valid && (ch[4] & 0x7fffffffu) >= P[kPre + 1];
          const unsigned long long im = __ballot(inner);
          const uint32_t q = stored + __builtin_amdgcn_mbcnt_hi(
                                          (uint32_t)(im >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)im, 0u));
          const uint32_t nst = (uint32_t)__popcll(im);
          if (inner && q < (uint32_t)kWave) {
#pragma unroll
            for (int w = 0; w < 5; ++w) s_top[w * kWave + q] = ch[w];
            s_top[5 * kWave + q] = cmeta;
          }
          if (stored + nst > (uint32_t)kWave) {  // wave-uniform, and rare: past the LDS top
            if (inner && q >= (uint32_t)kWave) {
              const long at = (top + (q - kWave)) & dq_mask;
#pragma unroll
              for (int w = 0; w < 5; ++w) deq[dq_idx<W>(at, w)] = ch[w];
              deq[dq_idx<W>(at, 5)] = cmeta;
            }
          }
          // the stored children are held, the leaves done: counted after the
          // batch when it is all one tree, else here, per child
          if (!batch_uni && valid) {
            const uint32_t csl = meta_slot(cmeta);
            if (inner) {
              atomicAdd(&s_hold[csl], 1);
            } else {
              atomicAdd(&s_nodes[csl], 1ull);
              atomicAdd(&s_leaves[csl], 1ull);
            }
          }
          stored += nst;
        } else if (valid) {
          const uint32_t *pr = &s_pst[lo * W];
          const long at = (top + c) & dq_mask;
          const uint32_t node = pr[0];
          const uint32_t kid =
              (c - s_off[lo]) == 0 ? (uint32_t)a.child_a[node] : (uint32_t)a.child_b[node];
          if (c >= keep0) {
            s_top[c - keep0] = kid;
            s_top[kWave + (c - keep0)] = pr[1];
          } else {
            deq[dq_idx<W>(at, 0)] = kid;
            deq[dq_idx<W>(at, 1)] = pr[1];
          }
        }
      }
      if (TREE == 1) {
        if (batch_uni && lane == 0) {
          s_hold[slot0] += (int)stored;
          s_nodes[slot0] += (unsigned long long)(total - stored);
          s_leaves[slot0] += (unsigned long long)(total - stored);
        }
        st_nodes += total - stored;
        top += stored;
        lds_n = stored < (uint32_t)kWave ? (int)stored : kWave;
      } else {
        top += total;
        lds_n = (int)keep;
      }
      if (top - bot > (long)high_water) high_water = (uint32_t)(top - bot);
      ++batches;
      st_nodes += n;
      __syncthreads();
      phase(st_tgen);
      // ---- hand-offs: a late delivery in, the oldest half out ---------------
      {
        const int k_in = (int)__builtin_amdgcn_readlane(box_n, 0);
        if (k_in > 0) absorb(k_in);  // waits in the inbox while it does not fit
      }
      if ((batches & 1) == 0 && top - bot >= 2 * kWave) donate();
      if (CPU && (batches & 7) == 0 && top - bot > 3 * kWave) {
        cpu_collect();
        cpu_deliver();
      }
      // return the credits of trees this WG no longer holds, so that a tree
      // finishes while its last holders work on other trees
      if (SERVE && (batches & 7) == 0) settle();
      phase(st_tdon);
      if ((batches & 63) == 0 && g.aborted()) {
        flush_stats();
        return;
      }
      tick(true);
      continue;
    }

    // ---- no local work: settle, then look for more -------------------------
    settle();
    cpu_collect();  // a host worker's returned credits may end the tree
    {
      const int k_in = (int)__builtin_amdgcn_readlane(lane == 0 ? ld_rlx(my_n) : 0u, 0);
      if (k_in > 0 && absorb(k_in)) continue;  // an empty deque takes any delivery
    }
    // admit the next submitted request, if any.  The request counter lives
    // in host memory: each idle WG looks every 4th pass, staggered
    if (SERVE && ((idle_iter + (uint32_t)me) & 3) == 0) {
      __shared__ int s_adm;
      if (lane == 0) {
        s_adm = -1;
        const uint32_t sub = ld_acq(a.dev_ctl);
        const uint32_t adm = ld_rlx(a.admitted);
        if (adm < sub && cas_rlx(a.admitted, adm, adm + 1) == adm) {
          const uint32_t *q = a.dev_inbox + (long)adm * 4;
          const uint32_t slot = ld_rlx(q + 0);
          const uint32_t seed = ld_rlx(q + 1);
          const uint32_t gmx = ld_rlx(q + 2);
          __hip_atomic_store(a.nodes + slot, 0ull, __ATOMIC_RELAXED, HS_AGENT);
          __hip_atomic_store(a.leaves + slot, 0ull, __ATOMIC_RELAXED, HS_AGENT);
          __hip_atomic_store(a.leafsum + slot, 0ull, __ATOMIC_RELAXED, HS_AGENT);
          st_rlx(a.maxd + slot, 0);
          st_rlx(a.slot_gen_mx + slot, gmx);
          s_gmx[slot] = gmx;
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
          for (int w = 0; w < 5; ++w) deq[dq_idx<W>(at, w)] = st[w];
          deq[dq_idx<W>(at, W - 1)] = slot << 16;
          s_hold[slot] = 1;
          s_cred[slot] = 1;
        }
        top += 1;
        __syncthreads();
        continue;
      }
    }
    // post one work request per idle period: take a ring slot (release:
    // the cleared inbox count and the inbox reads come first), then write
    // this WG's id into it
    if (!requested) {
      if (lane == 0) {
        const int r = me & (a.nrq - 1);
        const uint32_t t = __hip_atomic_fetch_add(a.rq + (long)r * HS_LINE_WORDS + 1, 1u,
                                                  __ATOMIC_RELEASE, HS_AGENT);
        st_rlx(a.rq_ring + (long)r * (a.rq_mask + 1) + (t & a.rq_mask), (uint32_t)me + 1);
      }
      requested = true;
    }
    // finished?  single tree: every credit returned; serving: host stop
    __shared__ int s_exit;
    if (lane == 0) {
      // the termination words are shared by every idle WG: look every 8th pass
      bool fin = false;
      if ((idle_iter & 7) == 0) {
        if (SERVE)
          fin = ld_acq(a.dev_ctl + 32) != 0 && ld_rlx(a.live) == 0;
        else
          fin = ld_rlx(a.credit) == 0;
      }
      if (!fin && !g.ok()) fin = true;
      if (!fin) sleep_quanta(idle_q);
      s_exit = fin;
    }
    __syncthreads();
    tick(false);
    idle_q = idle_q >= 32 ? 32 : 2 * idle_q;
    ++idle_iter;
    if (s_exit) {
      flush_stats();
      return;
    }
  }
}

// One tree per launch.  The register budget admits four waves per SIMD (at
// most 128 VGPRs; the geometric kernel needs 86, no spills, and 9.7 KB of
// LDS, 16 WGs per CU); big trees run three per SIMD by default, whose record
// loads hide behind the SHA-1 work of the other waves.
template <int TREE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_uts(UtsArgs a) {
  uts_worker<TREE, false, false>(a);
}

// The same with per-WG statistics (HS_UTS_STATS=1).
template <int TREE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_uts_stats(UtsArgs a) {
  uts_worker<TREE, false, true>(a);
}

// The persistent serving loop keeps its full register allocation (about 150
// VGPRs: three waves per SIMD).  Held to 128 (four per SIMD, a few spills) it
// served the 256-tree depth 6-14 mix no faster (61.0-61.1 against 61.4-62.1
// G nodes/s on one MI355X).
__global__ void __launch_bounds__(64) k_uts_serve(UtsArgs a) { uts_worker<1, true, false>(a); }

// One tree with host worker threads taking part (hs_uts_*_run_cpu).
template <int TREE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_uts_cpu(UtsArgs a) {
  uts_worker<TREE, false, false, true>(a);
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

// Device state of one engine.  The per-run words sit in one block that a
// single fill clears: inbox counts, request rings, credits, the results span
// (nodes, leaves, leaf sums, max depths, control words: copied back in one
// piece), depth limits and the serving counters.
struct Engine {
  DevArena ar;
  UtsArgs a;
  int tree = -1, blocks = 0, W = 0, dq_log2 = 0, cap = 0;
  double b0 = 0;
  uint32_t *d_thr = nullptr;
  std::vector<uint32_t> h_thr;  // the same thresholds, for host workers
  uint32_t *state = nullptr;
  size_t state_words = 0;
  uint32_t *results = nullptr;  // kResultWords words inside `state`
  static constexpr size_t kResultWords = 3 * 2 * kMaxSlots + kMaxSlots + 64;

  bool matches(int t, int nb, int cap_, double b) const {
    return tree == t && blocks == nb && cap == cap_ && (t == 0 || b0 == b);
  }

  // tree: 0 reference, 1 geometric.  cap_: records per WG deque (the ring
  // is the next power of two).
  int setup(hs_ctx *c, int tree_, int num_blocks, int cap_, double b0_) {
    std::memset(&a, 0, sizeof a);
    if (cap_ < kMinDeque || cap_ > kMaxDeque) return HS_EINVAL;
    tree = tree_;
    cap = cap_;
    dq_log2 = 0;
    while ((1L << dq_log2) < cap) ++dq_log2;
    b0 = b0_;
    blocks = num_blocks > 0 ? num_blocks : 8 * c->num_cu;
    W = tree == 1 ? 6 : 2;
    a.dq_log2 = dq_log2;
    a.dq_cap = cap;
    // an inbox holds at most a quarter deque
    a.wbox_cap = (int)std::min<long>(4096, cap / 4);
    // request rings: up to 32; every WG has at most one request outstanding
    // on its ring, and a claimed entry may be cleared late: twice the ring's
    // WG count keeps a new request off it
    a.nrq = 1;
    while (a.nrq < 32 && a.nrq * 8 < blocks) a.nrq <<= 1;
    const uint32_t per_ring = (uint32_t)((blocks + a.nrq - 1) / a.nrq);
    uint32_t ring = 1;
    while (ring < 2u * per_ring) ring <<= 1;
    a.rq_mask = ring - 1;
    // state block layout, every region on a 256-byte boundary
    auto up = [](size_t w) { return (w + 63) & ~(size_t)63; };
    const size_t o_wbn = 0;
    const size_t o_rq = up(o_wbn + (size_t)blocks * HS_LINE_WORDS);
    const size_t o_ring = up(o_rq + (size_t)a.nrq * HS_LINE_WORDS);
    const size_t o_cred = up(o_ring + (size_t)a.nrq * ring);
    const size_t o_res = up(o_cred + (size_t)kMaxSlots * HS_LINE_WORDS);
    const size_t o_gmx = up(o_res + kResultWords);
    const size_t o_live = up(o_gmx + kMaxSlots);
    state_words = up(o_live + 64);
    int rc;
    if ((rc = ar.alloc(&a.deq, (size_t)blocks * (((size_t)W << dq_log2) + kDqSkew)))) return rc;
    if ((rc = ar.alloc(&a.wbox, (size_t)blocks * a.wbox_cap * W))) return rc;
    if ((rc = ar.alloc(&state, state_words))) return rc;
    if ((rc = ar.alloc(&d_thr, kMaxThr))) return rc;
    a.wbox_n = state + o_wbn;
    a.rq = state + o_rq;
    a.rq_ring = state + o_ring;
    a.credit = state + o_cred;
    results = state + o_res;
    a.nodes = reinterpret_cast<unsigned long long *>(results);
    a.leaves = a.nodes + kMaxSlots;
    a.leafsum = a.leaves + kMaxSlots;
    a.maxd = results + 6 * kMaxSlots;
    a.ctl = a.maxd + kMaxSlots;
    a.slot_gen_mx = state + o_gmx;
    a.live = state + o_live;
    a.admitted = a.live + 32;
    if (tree == 1) {
      a.inv_log2q = (float)(1.0 / std::log2(1.0 - 1.0 / (1.0 + b0)));
      a.nthr = geo_thresholds(b0, h_thr);
      if (a.nthr <= 0 || a.nthr > kMaxThr) return HS_EINVAL;
      HS_HIP(hipMemcpy(d_thr, h_thr.data(), h_thr.size() * 4, hipMemcpyHostToDevice));
    }
    a.thr = d_thr;
    a.ticks = timeout_ticks(c, 0);
    HS_HIP(hipMemset(state, 0, state_words * 4));
    return HS_OK;
  }

  // The deque bound of a run whose trees go at most `depth` levels below
  // their roots: a one-record-at-a-time depth-first walk from any record
  // adds at most depth * (max_children - 1) records, and every WG keeps
  // that much of its deque free.  HS_EINVAL when the deque cannot hold even
  // that reserve.
  int bound(long depth) {
    const long kids = tree == 1 ? a.nthr : 2;
    const long soft = (long)cap - depth * (kids - 1);
    if (soft < 1) return HS_EINVAL;
    a.dq_soft = (int)soft;
    a.box_lim = (int)std::min<long>(a.wbox_cap, soft);
    return HS_OK;
  }

  // Clears the per-run state on the stream (never inside a captured graph).
  int reset(hipStream_t s) {
    HS_HIP(hipMemsetAsync(state, 0, state_words * 4, s));
    return HS_OK;
  }

  // Copies the results span back and fills r for one slot.
  int read_slot(hipStream_t s, int slot, hs_uts_result *r) {
    uint32_t res[kResultWords];
    HS_HIP(hipMemcpyAsync(res, results, sizeof res, hipMemcpyDeviceToHost, s));
    HS_HIP(hipStreamSynchronize(s));
    auto u64 = [&](int k) {
      return (unsigned long long)res[2 * k] | ((unsigned long long)res[2 * k + 1] << 32);
    };
    const uint32_t *ctl = res + 7 * kMaxSlots;
    r->nodes = u64(slot);
    r->leaves = u64(kMaxSlots + slot);
    r->leaf_sum = u64(2 * kMaxSlots + slot);
    r->max_depth = (int)res[6 * kMaxSlots + slot];
    r->aborted = ctl[0] != 0;
    if (ctl[32]) return HS_ENOMEM;  // a deque ring overflowed (the bound failed)
    return ctl[0] ? HS_ETIMEOUT : HS_OK;
  }

  // One geometric tree: 8 one-wave WGs per CU (two per SIMD), or 12 (three
  // per SIMD, if that many can be resident) when the tree is expected to be
  // big.  Measured on MI355X (b0 = 4), kernel time at 2048 / 3072 / 4096
  // WGs: depth 15 (7.85e9 nodes) 138.0 / 129.7 / 129.7 ms, depth 14
  // (1.96e9) 37.4 / 36.6 / 38.5-39.6 ms, depth 13 (4.9e8) 11.5 / 12.6 /
  // 13.9 ms, depth 12 4.6 / 6.4 / 6.7 ms: below about 1e9 nodes more WGs
  // only add idle polling.
  static int default_blocks(hs_ctx *c, double b0, int gen_mx) {
    if (std::pow(b0, (double)gen_mx) < 2.5e8) return 8 * c->num_cu;
    int fit = 0;
    if (coresident_wgs(c, (const void *)k_uts<1>, 64, &fit) != HS_OK || fit <= 0)
      return 8 * c->num_cu;
    return std::min(fit, 12 * c->num_cu);
  }
};

// Root of a run, delivered to WG 0's inbox with its credit: the geometric
// root digest SHA1(16 zero bytes || seed) at depth 0, or node 0 of the
// reference tree.  One wave.
__global__ void k_seed_root(UtsArgs a, int tree, uint32_t seed, uint32_t gen_mx) {
  if (threadIdx.x != 0) return;
  if (tree == 1) {
    uint32_t st[5];
    uts_root(seed, st);
    for (int w = 0; w < 5; ++w) a.wbox[w * 64] = st[w];  // record 0: word w at w * 64
    a.wbox[5 * 64] = 0;                                  // depth 0, slot 0
  } else {
    a.wbox[0] = 0;   // node 0
    a.wbox[64] = 0;  // slot 0
  }
  a.slot_gen_mx[0] = gen_mx;
  a.credit[0] = 1;
  a.wbox_n[0] = 1;
}

// HS_UTS_STATS=1 in the environment: per-WG statistics of the next run,
// summarised on stderr (busy/idle time, batch fill, hand-offs).
int stats_begin(Engine &e, hipStream_t s) {
  const char *v = std::getenv("HS_UTS_STATS");
  if (!v || !*v || *v == '0') {
    e.a.stats = nullptr;
    return HS_OK;
  }
  static_assert(sizeof(unsigned long long) == 8, "stats words");
  if (!e.a.stats) {
    int rc;
    if ((rc = e.ar.alloc(&e.a.stats, (size_t)e.blocks * kStatWords))) return rc;
  }
  HS_HIP(hipMemsetAsync(e.a.stats, 0, (size_t)e.blocks * kStatWords * 8, s));
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
               "nodes/WG min %llu med %llu max %llu; per batch: pop %.2f us, generate %.2f us "
               "(%.2f rounds), donate %.2f us\n",
               e.blocks, kernel_ms, 100.0 * busy / std::max(1.0, busy + idle),
               busy / 1e5 / e.blocks, sum[2], (double)sum[3] / std::max(1ull, sum[2]), sum[4],
               (double)sum[5] / std::max(1ull, sum[4]), sum[6],
               (double)sum[7] / std::max(1ull, sum[6]), nodes.front(), nodes[nodes.size() / 2],
               nodes.back(), sum[8] / 100.0 / std::max(1ull, sum[2]),
               sum[9] / 100.0 / std::max(1ull, sum[2]), (double)sum[11] / std::max(1ull, sum[2]),
               sum[10] / 100.0 / std::max(1ull, sum[2]));
  return HS_OK;
}

// ---- host worker threads ----------------------------------------------------

// SHA-1 (FIPS 180-4) of one padded 64-byte block on the host, with the word
// layout of the device's sha1_block.
inline uint32_t host_rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

void host_sha1_block(uint32_t h[5], uint32_t w[16]) {
  uint32_t a = 0x67452301u, b = 0xefcdab89u, c = 0x98badcfeu, d = 0x10325476u, e = 0xc3d2e1f0u;
  for (int t = 0; t < 80; ++t) {
    uint32_t wt;
    if (t < 16) {
      wt = w[t];
    } else {
      wt = host_rotl(w[(t - 3) & 15] ^ w[(t - 8) & 15] ^ w[(t - 14) & 15] ^ w[t & 15], 1);
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
    const uint32_t tmp = host_rotl(a, 5) + f + e + k + wt;
    e = d;
    d = c;
    c = host_rotl(b, 30);
    b = a;
    a = tmp;
  }
  h[0] = 0x67452301u + a;
  h[1] = 0xefcdab89u + b;
  h[2] = 0x98badcfeu + c;
  h[3] = 0x10325476u + d;
  h[4] = 0xc3d2e1f0u + e;
}

// What a host worker needs to expand a record.
struct HostTree {
  int tree = 1;
  const uint32_t *thr = nullptr;  // geometric: child-count thresholds
  int nthr = 0;
  uint32_t gen_mx = 0;
  const uint32_t *value = nullptr;  // reference tree arrays
  const int32_t *child_a = nullptr, *child_b = nullptr;
};

struct HostTotals {
  uint64_t nodes = 0, leaves = 0, lsum = 0;
  uint32_t maxd = 0;
};

// Depth-first search of one record's subtree on the host: the device's
// expansion (the same SHA-1 children and threshold counts, or the
// reference tree's arrays), counted the same way.
void host_expand(const HostTree &t, const uint32_t *rec, HostTotals &tot,
                 std::vector<uint32_t> &stack) {
  const int W = t.tree == 1 ? 6 : 2;
  stack.assign(rec, rec + W);
  while (!stack.empty()) {
    uint32_t r[6];
    std::copy(stack.end() - W, stack.end(), r);
    stack.resize(stack.size() - W);
    ++tot.nodes;
    if (t.tree == 1) {
      const uint32_t meta = r[5], depth = meta & 0xffffu;
      tot.maxd = std::max(tot.maxd, depth);
      const uint32_t h = r[4] & 0x7fffffffu;
      const int nc = (depth < t.gen_mx || depth == 0)
                         ? (int)(std::upper_bound(t.thr, t.thr + t.nthr, h) - t.thr)
                         : 0;
      if (nc == 0) ++tot.leaves;
      for (int i = 0; i < nc; ++i) {  // child i = SHA1(parent || i)
        uint32_t w[16] = {r[0], r[1], r[2], r[3], r[4], (uint32_t)i, 0x80000000u, 0, 0, 0, 0, 0,
                          0, 0, 0, 192u};
        uint32_t ch[5];
        host_sha1_block(ch, w);
        stack.insert(stack.end(), ch, ch + 5);
        stack.push_back(meta + 1);
      }
    } else {
      const uint32_t node = r[0];
      if (t.child_a[node] == 0) {
        ++tot.leaves;
        tot.lsum += t.value[node];
      } else {
        stack.push_back((uint32_t)t.child_a[node]);
        stack.push_back(r[1]);
        stack.push_back((uint32_t)t.child_b[node]);
        stack.push_back(r[1]);
      }
    }
  }
}

// One host worker: asks for records, expands them, publishes its totals.
void host_worker(uint32_t *mail, const uint32_t *box, int W, const HostTree &t,
                 const std::atomic<int> *stop, hs_uts_cpu_result *acc, std::mutex *mu) {
  HostTotals tot;
  uint32_t cred = 0, seq = 0;
  uint64_t deliveries = 0;
  double busy = 0;
  std::vector<uint32_t> stack, rec(W);
  auto put = [&](int k, uint32_t v, int order) { __atomic_store_n(mail + k, v, order); };
  put(kMailWant, 1, __ATOMIC_RELEASE);
  while (!stop->load(std::memory_order_acquire)) {
    const uint32_t k = __atomic_load_n(mail + kMailBox, __ATOMIC_ACQUIRE);
    if (k == 0) {
      std::this_thread::yield();
      continue;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < k; ++i) {
      for (int w = 0; w < W; ++w) rec[w] = box[((i >> 6) * W + w) * 64 + (i & 63)];
      host_expand(t, rec.data(), tot, stack);
    }
    cred += k;
    put(kMailNodes, (uint32_t)tot.nodes, __ATOMIC_RELAXED);
    put(kMailNodes + 1, (uint32_t)(tot.nodes >> 32), __ATOMIC_RELAXED);
    put(kMailLeaves, (uint32_t)tot.leaves, __ATOMIC_RELAXED);
    put(kMailLeaves + 1, (uint32_t)(tot.leaves >> 32), __ATOMIC_RELAXED);
    put(kMailLsum, (uint32_t)tot.lsum, __ATOMIC_RELAXED);
    put(kMailLsum + 1, (uint32_t)(tot.lsum >> 32), __ATOMIC_RELAXED);
    put(kMailCred, cred, __ATOMIC_RELAXED);
    put(kMailMaxd, tot.maxd, __ATOMIC_RELAXED);
    put(kMailSeq, ++seq, __ATOMIC_RELEASE);  // the totals, then their sequence
    put(kMailBox, 0, __ATOMIC_RELEASE);      // the box may be refilled
    ++deliveries;
    busy += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  std::lock_guard<std::mutex> lk(*mu);
  acc->nodes += tot.nodes;
  acc->deliveries += deliveries;
  acc->busy_s += busy;
}

// One run: clear the state, deliver the root, run the kernel, read slot 0.
// depth: how many levels the tree goes below its root (the deque bound).
// cpu_threads > 0: that many host workers take part (k_uts_cpu); `host`
// describes the tree to them.
template <int TREE>
int run_single(hs_ctx *c, Engine &e, uint32_t seed, uint32_t gen_mx, long depth,
               hs_uts_result *res, int cpu_threads = 0, HostTree host = HostTree(),
               hs_uts_cpu_result *cpu = nullptr) {
  int rc;
  if ((rc = e.bound(depth))) return rc;
  if ((rc = e.reset(c->stream))) return rc;
  if ((rc = stats_begin(e, c->stream))) return rc;
  e.a.ncpu = 0;
  e.a.cpu_mail = e.a.cpu_box = nullptr;
  struct Pinned {  // host workers' mailboxes and boxes
    void *p[2] = {nullptr, nullptr};
    ~Pinned() { for (void *q : p) if (q) (void)hipHostFree(q); }
  } pinned;
  const int box_cap = kWave;
  if (cpu_threads > 0) {
    const unsigned kPinFlags = hipHostMallocCoherent | hipHostMallocMapped;
    HS_HIP(hipHostMalloc(&pinned.p[0], (size_t)cpu_threads * kMailWords * 4, kPinFlags));
    HS_HIP(hipHostMalloc(&pinned.p[1], (size_t)cpu_threads * box_cap * e.W * 4, kPinFlags));
    std::memset(pinned.p[0], 0, (size_t)cpu_threads * kMailWords * 4);
    e.a.cpu_mail = static_cast<uint32_t *>(pinned.p[0]);
    e.a.cpu_box = static_cast<uint32_t *>(pinned.p[1]);
    e.a.ncpu = cpu_threads;
    e.a.cpu_stride = std::max(1, e.blocks / cpu_threads);
    e.a.cpu_box_cap = box_cap;
  }
  hipLaunchKernelGGL(k_seed_root, dim3(1), dim3(kWave), 0, c->stream, e.a, TREE, seed, gen_mx);
  HS_HIP(hipGetLastError());
  EventSet ev;
  if ((rc = ev.create(2))) return rc;
  std::atomic<int> stop{0};
  std::mutex mu;
  hs_uts_cpu_result dummy;  // the workers' totals when the caller wants none
  std::memset(&dummy, 0, sizeof dummy);
  std::vector<std::thread> workers;  // joined below before any return
  if (cpu_threads > 0) {
    HS_HIP(hipStreamSynchronize(c->stream));
    if (cpu) std::memset(cpu, 0, sizeof *cpu);
    for (int t = 0; t < cpu_threads; ++t)
      workers.emplace_back(host_worker, e.a.cpu_mail + (size_t)t * kMailWords,
                           e.a.cpu_box + (size_t)t * box_cap * e.W, e.W, std::cref(host), &stop,
                           cpu ? cpu : &dummy, &mu);
  }
  HS_HIP(hipEventRecord(ev[0], c->stream));
  if (cpu_threads > 0)
    hipLaunchKernelGGL((k_uts_cpu<TREE>), dim3(e.blocks), dim3(kWave), 0, c->stream, e.a);
  else if (e.a.stats)
    hipLaunchKernelGGL((k_uts_stats<TREE>), dim3(e.blocks), dim3(kWave), 0, c->stream, e.a);
  else
    hipLaunchKernelGGL((k_uts<TREE>), dim3(e.blocks), dim3(kWave), 0, c->stream, e.a);
  const hipError_t le = hipGetLastError();
  const hipError_t re = hipEventRecord(ev[1], c->stream);
  const hipError_t se = hipEventSynchronize(ev[1]);
  stop.store(1, std::memory_order_release);
  for (auto &w : workers) w.join();
  if (cpu) cpu->threads = cpu_threads;
  HS_HIP(le);
  HS_HIP(re);
  HS_HIP(se);
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  res->kernel_ms = ms;
  if ((rc = stats_report(e, ms))) return rc;
  return e.read_slot(c->stream, 0, res);
}

// The geometric engines a context keeps: the two most recently used grids.
// The default grid depends on the tree's expected size (all resident WGs
// for large trees, 8 per CU otherwise), so a stream of mixed trees
// alternates between two grids; one cached engine would be freed and
// rebuilt (GBs of deques) at every alternation.
struct EngineCache {
  static constexpr int kEntries = 2;
  Engine *e[kEntries] = {};
  ~EngineCache() {
    for (Engine *p : e) delete p;
  }
  // The engine for (nb, b0), most recently used first; built on a miss,
  // replacing the least recently used.  nullptr and *rc on failure.
  Engine *get(hs_ctx *c, int nb, double b0, int *rc) {
    int i = 0;
    while (i < kEntries && !(e[i] && e[i]->matches(1, nb, kDefaultDeque, b0))) ++i;
    Engine *hit = nullptr;
    if (i < kEntries) {
      hit = e[i];
    } else {
      i = kEntries - 1;
      delete e[i];
      e[i] = nullptr;
      hit = new Engine;
      if ((*rc = hit->setup(c, 1, nb, kDefaultDeque, b0))) {
        delete hit;
        return nullptr;
      }
    }
    for (; i > 0; --i) e[i] = e[i - 1];
    e[0] = hit;
    return hit;
  }
};

void free_engines(void *p) { delete static_cast<EngineCache *>(p); }

}  // namespace

namespace {

int uts_ref(hs_ctx *c, int num_nodes, const uint32_t *value, const int32_t *child_a,
            const int32_t *child_b, int num_blocks, int cpu_threads, hs_uts_result *res,
            hs_uts_cpu_result *cpu) {
  if (!c || !value || !child_a || !child_b || !res || num_nodes < 1 || num_blocks < 0 ||
      cpu_threads < 0)
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  for (int i = 0; i < num_nodes; ++i)  // binary tree, children in range
    if (child_a[i] < 0 || child_a[i] >= num_nodes || child_b[i] < 0 || child_b[i] >= num_nodes ||
        (child_a[i] == 0) != (child_b[i] == 0))
      return HS_EINVAL;
  // the tree's height bounds the deques (binary: one extra record per level)
  long height = 0;
  {
    std::vector<int32_t> depth(num_nodes, -1), queue(1, 0);
    depth[0] = 0;
    for (size_t q = 0; q < queue.size(); ++q) {
      const int v = queue[q];
      if (child_a[v] == 0) continue;
      for (const int ch : {child_a[v], child_b[v]}) {
        if (depth[ch] >= 0) return HS_EINVAL;  // not a tree
        depth[ch] = depth[v] + 1;
        height = std::max<long>(height, depth[ch]);
        queue.push_back(ch);
      }
    }
  }
  HS_HIP(hipSetDevice(c->device));
  Engine e;
  const int cap = (int)std::min<long>(kMaxDeque, std::max<long>(1L << 14, height + 4096));
  int rc = e.setup(c, 0, num_blocks, cap, 0);
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
  if (cpu_threads > e.blocks) return HS_EINVAL;  // one serving WG per host worker
  HostTree host;
  host.tree = 0;
  host.value = value;
  host.child_a = child_a;
  host.child_b = child_b;
  return run_single<0>(c, e, 0, 0, height, res, cpu_threads, host, cpu);
}

// The geometric engines of a context are kept between calls (same grid,
// deque size and b0; the two most recent grids), so a call costs a state
// fill, the root delivery, the kernel and one small copy back.
int uts_geo(hs_ctx *c, int seed, double b0, int gen_mx, int num_blocks, int cpu_threads,
            hs_uts_result *res, hs_uts_cpu_result *cpu) {
  if (!c || !res || !(b0 > 0 && b0 < 1000) || gen_mx < 0 || gen_mx > 60000 || num_blocks < 0 ||
      cpu_threads < 0)
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  HS_HIP(hipSetDevice(c->device));
  const int nb = num_blocks > 0 ? num_blocks : Engine::default_blocks(c, b0, gen_mx);
  if (!c->uts_cache) {
    c->uts_cache = new EngineCache;
    c->uts_cache_free = free_engines;
  }
  int rc = HS_OK;
  Engine *e = static_cast<EngineCache *>(c->uts_cache)->get(c, nb, b0, &rc);
  if (!e) return rc;
  if (cpu_threads > e->blocks) return HS_EINVAL;  // one serving WG per host worker
  HostTree host;
  host.tree = 1;
  host.thr = e->h_thr.data();
  host.nthr = (int)e->h_thr.size();
  host.gen_mx = (uint32_t)gen_mx;
  return run_single<1>(c, *e, (uint32_t)seed, (uint32_t)gen_mx, gen_mx, res, cpu_threads, host,
                       cpu);
}

}  // namespace

extern "C" int hs_uts_ref_run(hs_ctx *c, int num_nodes, const uint32_t *value,
                              const int32_t *child_a, const int32_t *child_b, int num_blocks,
                              hs_uts_result *res) {
  return uts_ref(c, num_nodes, value, child_a, child_b, num_blocks, 0, res, nullptr);
}

extern "C" int hs_uts_ref_run_cpu(hs_ctx *c, int num_nodes, const uint32_t *value,
                                  const int32_t *child_a, const int32_t *child_b, int num_blocks,
                                  int cpu_threads, hs_uts_result *res, hs_uts_cpu_result *cpu) {
  return uts_ref(c, num_nodes, value, child_a, child_b, num_blocks, cpu_threads, res, cpu);
}

extern "C" int hs_uts_geo_run(hs_ctx *c, int seed, double b0, int gen_mx, int num_blocks,
                              hs_uts_result *res) {
  return uts_geo(c, seed, b0, gen_mx, num_blocks, 0, res, nullptr);
}

extern "C" int hs_uts_geo_run_cpu(hs_ctx *c, int seed, double b0, int gen_mx, int num_blocks,
                                  int cpu_threads, hs_uts_result *res, hs_uts_cpu_result *cpu) {
  return uts_geo(c, seed, b0, gen_mx, num_blocks, cpu_threads, res, cpu);
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
  if (cfg->num_blocks < 0 || cfg->deque_nodes < 0) return HS_EINVAL;
  HS_HIP(hipSetDevice(c->device));
  int rc, fit = 0;
  // persistent: every WG, the dispatcher (the grid's last) included, must be
  // resident.  The default is every WG that can be (three waves per SIMD at
  // the kernel's register use): on the 256-tree depth 6-14 mix it served
  // 38-43 G nodes/s against 35-37 G with 8 WGs per CU.
  if ((rc = coresident_wgs(c, (const void *)k_uts_serve, 64, &fit))) return rc;
  const int blocks = cfg->num_blocks > 0 ? cfg->num_blocks : fit;
  if (blocks < 2) return HS_EINVAL;  // the dispatcher and at least one worker
  if (blocks > fit) return HS_EOCCUPANCY;
  Engine e;
  const int cap = cfg->deque_nodes > 0 ? cfg->deque_nodes : kDefaultDeque;
  if ((rc = e.setup(c, 1, blocks, cap, b0))) return rc;
  {  // every WG keeps the reserve of the deepest tree of the session
    long dmax = 0;
    for (int i = 0; i < n; ++i) dmax = std::max<long>(dmax, req[i].gen_mx);
    if ((rc = e.bound(dmax))) return rc;
  }
  if (cfg->timeout_s > 0) e.a.ticks = timeout_ticks(c, cfg->timeout_s);
  // pinned, host-coherent request inbox / response board / stop word
  struct Pinned {
    void *p[3] = {nullptr, nullptr, nullptr};
    ~Pinned() { for (void *q : p) if (q) (void)hipHostFree(q); }
  } pinned;
  const unsigned kPinFlags = hipHostMallocCoherent | hipHostMallocMapped;
  HS_HIP(hipHostMalloc(&pinned.p[0], (size_t)n * 16, kPinFlags));
  HS_HIP(hipHostMalloc(&pinned.p[1], 256, kPinFlags));
  HS_HIP(hipHostMalloc(&pinned.p[2], (size_t)kMaxSlots * 32, kPinFlags));
  uint32_t *h_inbox = static_cast<uint32_t *>(pinned.p[0]);
  uint32_t *h_ctl = static_cast<uint32_t *>(pinned.p[1]);
  uint32_t *h_resp = static_cast<uint32_t *>(pinned.p[2]);
  std::memset(h_inbox, 0, (size_t)n * 16);
  std::memset(h_ctl, 0, 256);
  std::memset(h_resp, 0, (size_t)kMaxSlots * 32);
  volatile uint32_t *submitted = h_ctl, *stop = h_ctl + 32;
  e.a.inbox = h_inbox;
  e.a.submitted = h_ctl;
  e.a.stop = h_ctl + 32;
  // device copies of the requests and control words (one dispatcher WG,
  // the grid's last, keeps them; the others are workers)
  if ((rc = e.ar.alloc(&e.a.dev_inbox, (size_t)n * 4))) return rc;
  if ((rc = e.ar.alloc(&e.a.dev_ctl, 64))) return rc;
  HS_HIP(hipMemset(e.a.dev_inbox, 0, (size_t)n * 16));
  HS_HIP(hipMemset(e.a.dev_ctl, 0, 256));
  e.a.resp = h_resp;

  using clk = std::chrono::steady_clock;
  std::vector<int> slot_req(slots, -1);
  std::vector<uint32_t> slot_seq(slots, 0);
  std::vector<double> lat(n, 0);
  int next = 0, done = 0, live = 0;
  const auto t0 = clk::now();
  hipLaunchKernelGGL(k_uts_serve, dim3(e.blocks), dim3(kWave), 0, c->stream, e.a);
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
    std::this_thread::sleep_for(std::chrono::microseconds(5));
  }
  __atomic_store_n(stop, 1u, __ATOMIC_RELEASE);
  HS_HIP(hipStreamSynchronize(c->stream));
  sum->wall_s = now_s();
  uint32_t ctl[64];
  HS_HIP(hipMemcpy(ctl, e.a.ctl, sizeof ctl, hipMemcpyDeviceToHost));
  sum->cut_batches = ctl[40];
  sum->deque_high_water = ctl[48];
  sum->deque_soft = (uint32_t)e.a.dq_soft;
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
