// hs_syncprims.hip -- the plain per-instance form of the syncPrims
// microbenchmarks (hip/syncPrims/uvm/main.hip.cpp): one kernel launch per
// repeat, each followed by a synchronize, the whole sequence timed as the
// region of interest (main.hip.cpp:2354-2511).
//
// Kernels (one template instance per primitive family):
//   k_mutex<KIND, UNIQ, MODE, AGENT>  spin/EBO/sleep/FA mutex, global or per-CU
//                               (main.hip.cpp:645-857)
//   k_sem<EBO, UNIQ, MODE, PRIO, AGENT>  spin/EBO reader-writer semaphores
//                               (main.hip.cpp:861-930, 1005-1073,
//                                1147-1214, 1285-1352)
//   k_barrier<FLAG, LX, MODE>   atomic / lock-free tree barrier, with and
//                               without local exchange
//                               (main.hip.cpp:104-149, 315-373, 542-643)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "hs_host.h"
#include "../../include/hs_detail/hs_syncprims_dev.h"

using namespace hs;

namespace {

struct SPArgs {
  int32_t *storage;
  uint32_t *ctl;  // [0] abort word
  uint32_t *locks;
  uint32_t *ring;
  long ring_len;
  int ring_stride;
  TreeBar tb;
  FlagBar fb;
  XcdTree xt;
  XcdFlags xf;
  int L, iters, num_cu, nab, per_cu, nxcd;
  uint32_t e_global0, e_local0;
  Backoff bo;
  uint64_t ticks;
  uint32_t *xcd_map;  // per launch: physical XCD of each logical XCD (b % nxcd)
  uint32_t *flat;     // sense-reversing family: main block, local block
  int inc_bar;        // inc-mode data pass: barrier between loads and stores
};

// AGENT: critical-section data through agent-scope accesses, drained by
// every wave before the WG barrier that precedes the release, so the lock
// and semaphore words need no release/acquire fences (the strict form keeps
// them and plain data accesses).
template <int KIND, bool UNIQ, int MODE, bool AGENT>
__global__ void __launch_bounds__(1024) k_mutex(SPArgs a) {
  __shared__ int s_ok;
  Guard g;
  g.init(a.ctl, a.ticks);
  const int lock = UNIQ ? (int)(blockIdx.x % a.num_cu) : 0;
  uint32_t *lk = a.locks + (long)lock * kLockWords;
  uint32_t *ring = a.ring ? a.ring + (long)lock * a.ring_len * a.ring_stride : nullptr;
  // all WGs share region 0, or the per-CU region (main.hip.cpp:717, 742)
  const long tid0 = UNIQ ? (long)lock * blockDim.x : 0;
  uint32_t ticket = 0;
  if (is_master()) s_ok = !g.aborted();
  __syncthreads();
  if (!s_ok) return;
  for (int i = 0; i < a.iters; ++i) {
    if (is_master()) s_ok = mutex_lock<KIND, AGENT>(lk, ring, a.ring_stride, ticket, a.bo, g);
    __syncthreads();
    if (!s_ok) return;
    wg_move<MODE, AGENT, 32, true>(a.storage, tid0, a.L);
    if (AGENT) wait_mem();
    __syncthreads();
    if (is_master()) {
      mutex_unlock<KIND, AGENT>(lk, ring, a.ring_stride, ticket);
      add_rlx(lk + kLkCs, 1);  // entry count, off the hand-off path (no return)
    }
  }
}

template <bool EBO, bool UNIQ, int MODE, bool PRIO, bool AGENT>
__global__ void __launch_bounds__(1024) k_sem(SPArgs a) {
  __shared__ int s_ok;
  __shared__ int32_t sink[1024];
  Guard g;
  g.init(a.ctl, a.ticks);
  const int b = blockIdx.x, grid = gridDim.x, nab = a.nab, per = a.per_cu;
  const int cu = b % nab, pw = b / nab;
  const bool writer = pw == cu % per;  // main.hip.cpp:883
  uint32_t *lk = a.locks + (UNIQ ? (long)cu * kLockWords : 0);
  sink[threadIdx.x] = 0;
  uint32_t post_amount = 1u;  // a writer returns the whole count (max)
  if (is_master()) {
    s_ok = !g.aborted();
    if (writer) post_amount = ld_rlx(lk + kLkMax);
  }
  __syncthreads();
  if (!s_ok) return;
  for (int i = 0; i < a.iters; ++i) {
    if (is_master()) s_ok = sem_wait<EBO, PRIO, AGENT>(lk, writer, a.bo, g);
    __syncthreads();
    if (!s_ok) return;
    if (writer) {
      if (UNIQ) {  // main.hip.cpp:1040-1058: own region, then +numCU WGs
        int cur = b;
        for (int j = 0; j < per; ++j) {
          wg_move<MODE, AGENT, 32, true>(a.storage, (long)cur * blockDim.x, a.L);
          cur = (cur + nab) % grid;
        }
      } else {  // main.hip.cpp:903-915: regions 0..perCU-1
        for (int j = 0; j < per; ++j)
          wg_move<MODE, AGENT, 32, true>(a.storage, (long)j * blockDim.x, a.L);
      }
    } else {
      wg_read<AGENT>(a.storage, (long)(UNIQ ? b : pw) * blockDim.x, a.L, sink);
    }
    if (AGENT) wait_mem();
    __syncthreads();
    if (is_master()) {
      sem_post<PRIO, AGENT>(lk, post_amount);
      add_rlx(lk + (writer ? kLkCs : kLkRd), 1);  // entry counts, off the hand-off path
    }
  }
}

// Which WG touches which region, in the XCD-aware form.  The final storage
// depends only on each region's sequence of passes, not on which WG makes
// them, so the data walk is laid out for the XCDs' L2s: a pass whose region
// is next touched by a WG on the same XCD stores with keep_l2, and only a
// pass whose region moves to another XCD stores at agent scope.  Logical XCD
// r = b % nxcd (the placement check makes it one physical XCD); j = b / nxcd
// numbers the WGs of one XCD.
//   - Without local exchange, region cur is next touched by the WG with the
//     next lower data id.  Data ids are numbered XCD by XCD, d = r *
//     (grid / nxcd) + j, so only the first id of each XCD hands its region to
//     another XCD.
//   - NaiveSRB's exchange re-visits the WG's own region: the first pass
//     keeps, the second follows the rule above.
//   - The group exchange (stride nab) hands the first region to d - nab and
//     the second to d - 1 (next episode).  With gx = nab / nxcd, the ids
//     d = nab * (j / gx) + gx * r + j % gx put d - nab on the same XCD always
//     and d - 1 unless d % gx == 0.  The barrier groups follow the data ids
//     (members d % nab, index d / nab, as the reference's exchange needs),
//     renumbered so that group g lies on logical XCD g % nxcd.
// Agent-scope stores over lines left dirty by another CU measured slow when
// every WG makes them (a walk that did so ran 2x slower), so each walk keeps
// them to a few WGs per XCD.  Kept stores rely on the placement check (WGs
// with equal b % nxcd share an XCD), for the writers and for the WGs that
// read after them; `placed` is the same for the whole grid.  In the strict
// form data ids and groups are the reference's (d = b, group b % nab).
struct DataPlan {
  int id;          // data id: the first region this WG touches
  int grp, q;      // barrier group and index within it
  bool keep_first, keep_second;
  bool placed;     // some WG of the grid keeps stores in L2
};
__device__ __forceinline__ DataPlan data_plan(bool xcd, bool lx, bool self_exchange, int b,
                                              int grid, int nab, int nx) {
  DataPlan d{b, b % nab, b / nab, false, false, false};
  if (!xcd || nx <= 1 || grid % nx) return d;
  const int r = b % nx, j = b / nx;
  if (!lx || self_exchange) {
    const int per_x = grid / nx;
    d.id = r * per_x + j;
    const bool next_here = d.id % per_x != 0;
    d.keep_first = lx ? true : next_here;
    d.keep_second = lx && next_here;
    d.placed = lx || per_x > 1;
  } else if (nab % nx == 0 && grid % nab == 0) {
    const int gx = nab / nx;
    d.id = nab * (j / gx) + gx * r + j % gx;
    const int gd = d.id % nab;  // data group: members gd + k * nab, all on XCD r
    d.grp = (gd % gx) * nx + gd / gx;
    d.q = d.id / nab;
    d.keep_first = true;
    d.keep_second = d.id % gx != 0;
    d.placed = true;
  }
  return d;
}

template <bool FLAG, bool LX, int MODE, bool XCD>
__global__ void __launch_bounds__(1024) k_barrier(SPArgs a) {
  __shared__ int s_ok, s_flag;
  Guard g;
  g.init(a.ctl, a.ticks);
  const int b = blockIdx.x, grid = gridDim.x, nab = a.nab;
  const DataPlan dp = data_plan(XCD, LX && a.per_cu > 1, false, b, grid, nab, a.nxcd);
  const int grp = dp.grp, q = dp.q;  // cuID, perCU_WGID (HM:122-132)
  const XcdSeat seat = (XCD && !FLAG) ? xcd_seat(a.xt, grp, q) : XcdSeat{};
  int cur = dp.id;
  uint32_t eg = a.e_global0, el = a.e_local0;
  if (is_master()) {
    s_ok = !g.aborted();
    // The XCD-aware protocol needs every WG with the same b % nxcd on the
    // same XCD (round-robin dispatch; the starting XCD varies per launch).
    if (XCD) {
      const uint32_t me = xcc_id();
      const uint32_t old = cas_rlx(a.xcd_map + b % a.nxcd, 0xffffffffu, me);
      if (old != 0xffffffffu && old != me) {
        st_rlx(a.ctl + 32, 1);
        st_rlx(a.ctl, 1);
        s_ok = 0;
      }
    }
  }
  __syncthreads();
  if (!s_ok) return;
  for (int i = 0; i < a.iters; ++i) {
    wg_pass<MODE, XCD>(a.inc_bar, a.storage, (long)cur * blockDim.x, a.L, dp.keep_first);
    if (LX && a.per_cu > 1) {
      bool ok;
      if (FLAG) {
        ok = XCD ? xcd_flag_barrier(a.xf, grp, q, ++el, 0, true, a.bo.poll_q, g, &s_flag)
                 : flag_tree_barrier(a.fb, grp, q, ++el, 0, true, a.bo.poll_q, g, &s_flag);
      } else {
        if (XCD) wait_mem();
        __syncthreads();
        if (is_master())
          s_ok = XCD ? xcd_tree_arrive(a.xt, seat, grp, true, a.e_global0 + i + 1, a.bo.poll_q, g)
                     : atomic_tree_arrive(a.tb, grp, true, a.bo.poll_q, g);
        __syncthreads();
        ok = s_ok;
      }
      if (!ok) return;
      cur = (cur + nab) % grid;  // HM:355-358
      wg_pass<MODE, XCD>(a.inc_bar, a.storage, (long)cur * blockDim.x, a.L, dp.keep_second);
    }
    bool ok;
    if (FLAG) {
      ++el;
      ++eg;
      ok = XCD ? xcd_flag_barrier(a.xf, grp, q, el, eg, false, a.bo.poll_q, g, &s_flag)
               : flag_tree_barrier(a.fb, grp, q, el, eg, false, a.bo.poll_q, g, &s_flag);
    } else {
      if (XCD) wait_mem();
      __syncthreads();
      if (is_master())
        s_ok = XCD ? xcd_tree_arrive(a.xt, seat, grp, false, a.e_global0 + i + 1, a.bo.poll_q, g)
                   : atomic_tree_arrive(a.tb, grp, false, a.bo.poll_q, g);
      __syncthreads();
      ok = s_ok;
    }
    if (!ok) return;
    cur = (cur + 1) % grid;  // HM:145-147
  }
}

// Sense-reversing barrier family (main.hip.cpp:195-539, ids 37-42).
//   VAR 0 SRB / 1 SRB + local exchange: tree (group -> device) when more
//     than 4 WGs share a group, else flat over the grid
//     (joinBarrier_helperSRB, hipLocksBarrierAtomicSRB.h:186-225);
//   VAR 2 NaiveSRB / 3 + local exchange: flat over the grid's WGs;
//   VAR 4 NaiveAllSRB / 5 + local exchange: flat over every thread
//     (one arrival of weight 64 per wave).
// The local-exchange barrier is the per-group barrier (VAR 1) or a second
// flat barrier (VAR 3, 5).  VAR 3 re-visits its own region, because its
// exchange stride is the grid (main.hip.cpp:472-474, 1568); VAR 1 and 5
// step by the group count (main.hip.cpp:411-413, 513-515).  XCD selects
// agent-scope data accesses with fence-free arrivals: the XCD-aware tree
// (which checks the WGs' placement) or the flat barrier (which does not).
template <int VAR, int MODE, bool XCD>
__global__ void __launch_bounds__(1024) k_srb(SPArgs a) {
  __shared__ int s_ok;
  Guard g;
  g.init(a.ctl, a.ticks);
  const int b = blockIdx.x, grid = gridDim.x, nab = a.nab;
  const bool lx = (VAR & 1) && a.per_cu > 1;
  const bool all = VAR >= 4;
  const bool tree = VAR <= 1 && a.per_cu > 4;
  const uint32_t units = all ? (uint32_t)grid * (blockDim.x / 64) : (uint32_t)grid;  // flat arrivals
  const bool wave_lead = (threadIdx.x & 63) == 0;
  const DataPlan dp = data_plan(XCD, lx, VAR == 3, b, grid, nab, a.nxcd);
  const int grp = dp.grp;
  const XcdSeat seat = (XCD && tree) ? xcd_seat(a.xt, grp, dp.q) : XcdSeat{};
  int cur = dp.id;
  if (is_master()) {
    s_ok = !g.aborted();
    // the XCD tree and the L2-kept data walk depend on where WGs run
    if (XCD && (tree || dp.placed)) {
      const uint32_t me = xcc_id();
      const uint32_t old = cas_rlx(a.xcd_map + b % a.nxcd, 0xffffffffu, me);
      if (old != 0xffffffffu && old != me) {
        st_rlx(a.ctl + 32, 1);
        st_rlx(a.ctl, 1);
        s_ok = 0;
      }
    }
  }
  __syncthreads();
  if (!s_ok) return;
  uint32_t epoch = a.e_global0;  // barriers of each kind completed (tree and flat forms)
  auto barrier = [&](bool local) -> bool {
    wait_mem();
    __syncthreads();
    bool ok = true;
    if ((local && VAR == 1) || (!local && tree)) {
      if (is_master())
        ok = XCD ? xcd_tree_arrive(a.xt, seat, grp, local, epoch, a.bo.poll_q, g)
                 : atomic_tree_arrive(a.tb, grp, local, a.bo.poll_q, g);
    } else {
      uint32_t *blk = a.flat + (local ? kBarWords : 0);
      uint32_t *ep = local ? nullptr : a.tb.episodes;
      if (all ? wave_lead : is_master())
        ok = flat_arrive<XCD>(blk, units, epoch, ep, a.bo.poll_q, g);
    }
    return __syncthreads_and(ok) != 0;
  };
  for (int i = 0; i < a.iters; ++i) {
    epoch = a.e_global0 + i + 1;
    wg_pass<MODE, XCD>(a.inc_bar, a.storage, (long)cur * blockDim.x, a.L, dp.keep_first);
    if (lx) {
      if (!barrier(true)) return;
      if (VAR != 3) cur = (cur + nab) % grid;
      wg_pass<MODE, XCD>(a.inc_bar, a.storage, (long)cur * blockDim.x, a.L, dp.keep_second);
    }
    if (!barrier(false)) return;
    cur = (cur + 1) % grid;
  }
}

// Initialises lock blocks and (sleep mutex) ring slots.
__global__ void k_fill(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_init_locks(uint32_t *locks, int nlocks, uint32_t sem_max,
                             uint32_t *ring, long ring_len, int stride) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nlocks && sem_max) {
    uint32_t *lk = locks + i * kLockWords;
    lk[kLkHead] = sem_init_word(sem_max);  // count = max, writerWaiting = 1 (HSS:15-28)
    lk[kLkMax] = sem_max;
  }
  if (ring) {
    const long total = (long)nlocks * ring_len;
    for (long k = i; k < total; k += (long)gridDim.x * blockDim.x)
      ring[k * stride] = (k % ring_len == 0) ? 1u : 0xffffffffu;  // HI:69-76
  }
}

// WGs per sub-group of the XCD-aware barriers.  The counter tree costs
// about (sub + arrivals at the device counter) same-address RMWs in series,
// and the device counter takes grid / sub arrivals, so sub near the square
// root of the grid balances the two levels (MI355X: 16 best at 256 and 512
// WGs, 32 at 1024 and 2048).  The flag barrier's gathers poll one flag per
// lane, which favours smaller sub-groups: the square root of half the grid
// (16 at 1024 WGs, 32 at 2048).  A power of two in [16, 64].
int sub_group_size(int grid, bool flags) {
  const double r = std::sqrt(flags ? grid / 2.0 : (double)grid);
  int s = 16;
  while (s < 64 && 2 * s <= r) s *= 2;
  return s;
}

// Whether the XCD-aware lock-free barrier takes its XCD form (one
// sub-group per XCD, the XCD leaders all-to-all, each releasing its own
// XCD): when the leader's lanes can gather the XCD's other WGs' flags at up
// to kGatherMax flags per lane, all of a lane's loads in flight together.
// Measured per episode (numLdSt 10; device form / XCD form): 256 x 64
// 3.54 / 3.43 us, 512 x 256 3.76 / 3.35, 1024 x 256 4.16 / 3.85, 512 x 512
// 4.41 / 4.01, 256 x 1024 4.79 / 4.38, 1024 x 64 (two flags per lane) 3.83 /
// 3.57, 2048 x 64 (four) 4.02 / 3.86, 4096 x 64 (eight) 5.72 / 4.55.
bool lf_xcd_form(int nab, int per, int nxcd, int wg) {
  const int nwx = (nab / nxcd + (nab % nxcd ? 1 : 0)) * per;  // most WGs on one XCD
  return nwx - 1 <= kGatherMax * wg;
}

bool is_barrier(int p) { return p < 4 || (p >= 37 && p <= 42); }
bool is_mutex(int p) { return (p >= 4 && p <= 7) || (p >= 20 && p <= 23); }
bool is_sem(int p) { return (p >= 8 && p <= 19) || (p >= 24 && p <= 35); }
// per-group locks and regions: the uniq mutexes and uniq semaphores.  The
// priority EBO semaphores 32-35 are dispatched to the global kernel
// (main.hip.cpp:2455-2470 -> invokeEBOSemaphorePriority) although their
// storage is sized like the uniq semaphores' (main.hip.cpp:2109-2128).
bool is_uniq(int p) { return p >= 20 && p <= 31; }
bool is_prio(int p) { return (p >= 16 && p <= 19) || (p >= 32 && p <= 35); }

using KernelFn = void (*)(SPArgs);

template <int MODE, bool XCD>
KernelFn pick_kernel(int p) {
  switch (p) {
    case 0: return k_barrier<false, false, MODE, XCD>;
    case 1: return k_barrier<false, true, MODE, XCD>;
    case 2: return k_barrier<true, false, MODE, XCD>;
    case 3: return k_barrier<true, true, MODE, XCD>;
    case 4: return k_mutex<kSpin, false, MODE, XCD>;
    case 5: return k_mutex<kEBO, false, MODE, XCD>;
    case 6: return k_mutex<kSleep, false, MODE, XCD>;
    case 7: return k_mutex<kFA, false, MODE, XCD>;
    case 20: return k_mutex<kSpin, true, MODE, XCD>;
    case 21: return k_mutex<kEBO, true, MODE, XCD>;
    case 22: return k_mutex<kSleep, true, MODE, XCD>;
    case 23: return k_mutex<kFA, true, MODE, XCD>;
    case 37: return k_srb<0, MODE, XCD>;
    case 38: return k_srb<1, MODE, XCD>;
    case 39: return k_srb<2, MODE, XCD>;
    case 40: return k_srb<3, MODE, XCD>;
    case 41: return k_srb<4, MODE, XCD>;
    case 42: return k_srb<5, MODE, XCD>;
    default: break;
  }
  if (p >= 8 && p <= 11) return k_sem<false, false, MODE, false, XCD>;
  if (p >= 12 && p <= 15) return k_sem<true, false, MODE, false, XCD>;
  if (p >= 24 && p <= 27) return k_sem<false, true, MODE, false, XCD>;
  if (p >= 28 && p <= 31) return k_sem<true, true, MODE, false, XCD>;
  if (p >= 16 && p <= 19) return k_sem<false, false, MODE, true, XCD>;
  if (p >= 32 && p <= 35) return k_sem<true, false, MODE, true, XCD>;
  return nullptr;
}

uint32_t sem_count_of(int p) {
  static const uint32_t kN[4] = {1, 2, 10, 120};  // spinSem1/2/10/120
  if (p >= 8 && p <= 35 && !(p >= 20 && p <= 23)) return kN[(p - 8) % 4];
  return 0;
}

// The no-allocation form's buffers: the caller's device storage, and the
// per-device scratch array (see hs_syncprim_run_noalloc).
struct NoAlloc {
  int32_t *d_storage;
  uint32_t *scratch;
  size_t scratch_words;
};

int run_once(hs_ctx *c, const hs_syncprim_cfg *cfg, int32_t *storage, long n,
             hs_syncprim_result *res, bool xcd_mode, bool *topology_fail,
             const NoAlloc *na = nullptr);

}  // namespace

// The no-allocation entry point's state: lock blocks, queue ring, barrier
// words, flags, control words and placement map of one run, carved from
// this array.  Each GPU the library runs on holds its own copy (4 MiB) for
// the life of the process; hs_open looks up its device address once per
// context, with that context's device current.
constexpr size_t kScratchWords = (size_t)1 << 20;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace hs {
int scratch_init(hs_ctx *c) {
  void *ptr = nullptr;
  HS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  c->scratch = static_cast<uint32_t *>(ptr);
  c->scratch_words = kScratchWords;
  return HS_OK;
}
}  // namespace hs

extern "C" int hs_syncprim_run_noalloc(hs_ctx *c, const hs_syncprim_cfg *cfg,
                                       int32_t *d_storage, long n, hs_syncprim_result *res) {
  if (!c || !cfg || !d_storage || !res) return HS_EINVAL;
  if (!c->scratch) return HS_EHIP;
  const NoAlloc na{d_storage, c->scratch, c->scratch_words};
  const int p = cfg->prim, ncu = c->num_cu;
  const int per = cfg->num_wgs > 0 ? (cfg->num_wgs + std::min(cfg->num_wgs, ncu) - 1) /
                                         std::min(cfg->num_wgs, ncu)
                                   : 1;
  const bool xcd_form = (p >= 0 && p < 4) || ((p == 37 || p == 38) && per > 4);
  const bool srb_flat = p >= 37 && p <= 42 && !xcd_form;
  const int nx = num_xcds(c);
  if (cfg->sync_scope == 0 && (is_mutex(p) || is_sem(p) || (srb_flat && nx == 1)))
    return run_once(c, cfg, nullptr, n, res, true, nullptr, &na);
  const bool xcd = cfg->sync_scope == 0 && (xcd_form || srb_flat) && nx > 1;
  if (!xcd) return run_once(c, cfg, nullptr, n, res, false, nullptr, &na);
  // The placement check cannot fall back here: the storage has no saved
  // copy to restart from.  sync_scope_used 2 reports it.
  bool topo = false;
  const int rc = run_once(c, cfg, nullptr, n, res, true, &topo, &na);
  if (rc == HS_ETIMEOUT && topo) res->sync_scope_used = 2;
  return rc;
}

extern "C" int hs_syncprim_run(hs_ctx *c, const hs_syncprim_cfg *cfg,
                               int32_t *storage, long n, hs_syncprim_result *res) {
  if (!c || !cfg || !storage || !res) return HS_EINVAL;
  const int p = cfg->prim;
  // XCD-aware form: the tree barriers, and the SRB family where it is a tree
  // (more than 4 WGs per group); flat barriers stay release/acquire.
  const int ncu = c->num_cu;
  const int per = cfg->num_wgs > 0 ? (cfg->num_wgs + std::min(cfg->num_wgs, ncu) - 1) /
                                         std::min(cfg->num_wgs, ncu)
                                   : 1;
  const bool xcd_form = (p >= 0 && p < 4) || ((p == 37 || p == 38) && per > 4);
  // mutexes and semaphores: the fence-free (agent-scope data) form needs no
  // placement, so it runs directly; so do the flat sense-reversing barriers
  // on a one-XCD device.  With several XCDs their data walk keeps stores in
  // L2 and checks the placement like the trees.
  const bool srb_flat = p >= 37 && p <= 42 && !xcd_form;
  const int nx = num_xcds(c);
  if (cfg->sync_scope == 0 && (is_mutex(p) || is_sem(p) || (srb_flat && nx == 1)))
    return run_once(c, cfg, storage, n, res, true, nullptr);
  const bool xcd = cfg->sync_scope == 0 && (xcd_form || srb_flat) && nx > 1;
  if (!xcd) return run_once(c, cfg, storage, n, res, false, nullptr);
  std::vector<int32_t> init(storage, storage + (n > 0 ? n : 0));
  bool topo = false;
  int rc = run_once(c, cfg, storage, n, res, true, &topo);
  if (rc == HS_ETIMEOUT && topo) {  // placement was not round-robin: strict form
    std::copy(init.begin(), init.end(), storage);
    rc = run_once(c, cfg, storage, n, res, false, nullptr);
  }
  return rc;
}

namespace {

int run_once(hs_ctx *c, const hs_syncprim_cfg *cfg, int32_t *storage, long n,
             hs_syncprim_result *res, bool xcd_mode, bool *topology_fail, const NoAlloc *na) {
  std::memset(res, 0, sizeof(*res));
  const int p = cfg->prim;
  if (hs_prim_name(p) == nullptr) return HS_EINVAL;
  if (!hs_prim_supported(p)) return HS_EUNSUPPORTED;
  const int L = cfg->num_ldst, nw = cfg->num_wgs, iters = cfg->num_cs_iters;
  const int wg = cfg->wg_size ? cfg->wg_size : 64;
  const int reps = cfg->reps ? cfg->reps : 10;
  if (L < 0 || nw < 1 || nw > c->max_wgs || iters < 0 || reps < 1) return HS_EINVAL;
  if (wg % 64 || wg > c->max_threads || wg > 1024) return HS_EINVAL;
  if (cfg->mode != 0 && cfg->mode != 1) return HS_EINVAL;
  const int ncu = c->num_cu;
  const int nab = std::min(nw, ncu);
  const int per = (nw + nab - 1) / nab;
  const bool uniq = is_uniq(p);
  if (is_barrier(p) && nw > ncu && nw % ncu) return HS_EINVAL;  // RD:80
  if (is_sem(p) && uniq && nw % nab) return HS_EINVAL;  // each group needs its writer
  if (n != hs_storage_words(p, nw, ncu, wg, L)) return HS_EINVAL;

  KernelFn kfn = xcd_mode ? (cfg->mode ? pick_kernel<1, true>(p) : pick_kernel<0, true>(p))
                          : (cfg->mode ? pick_kernel<1, false>(p) : pick_kernel<0, false>(p));
  const int nxcd = xcd_mode ? num_xcds(c) : 1;
  res->sync_scope_used = xcd_mode ? 0 : 1;
  if (!kfn) return HS_EUNSUPPORTED;
  HS_HIP(hipSetDevice(c->device));
  if (is_barrier(p) || is_sem(p)) {  // forward progress needs co-residency
    int fit = 0;
    int rc = coresident_wgs(c, (const void *)kfn, wg, &fit);
    if (rc) return rc;
    if (fit < nw) return HS_EOCCUPANCY;
  }

  // ---- state, sized from the workload -------------------------------------
  const int nlocks = uniq ? ncu : 1;
  const bool sleep_mx = p == 6 || p == 22;
  long ring_len = 0;
  int ring_stride = 0;
  if (sleep_mx) {
    long tickets = uniq ? (long)((nw + ncu - 1) / ncu) : (long)nw;
    tickets *= (long)iters * reps;
    ring_len = tickets + 1;
    // each slot on its own 128-byte line while the rings stay under 128 MiB
    // (the no-allocation form: while they take at most a quarter of its
    // scratch)
    const long line_limit = na ? (long)(na->scratch_words / 4 / HS_LINE_WORDS) : (1L << 20);
    ring_stride = ((long)nlocks * ring_len <= line_limit) ? HS_LINE_WORDS : 1;
  }
  DevArena arena;
  size_t carved = 0;  // words of the scratch handed out (no-allocation form)
  // Device buffers: allocated, or carved from the scratch at 256-byte
  // boundaries (HS_ESCRATCH once it is used up).
  auto take = [&](auto **out, size_t count) -> int {
    if (!na) return arena.alloc(out, count);
    const size_t words = (count * sizeof(**out) + 3) / 4;
    if (carved + words > na->scratch_words) return HS_ESCRATCH;
    *out = reinterpret_cast<std::remove_reference_t<decltype(*out)>>(na->scratch + carved);
    carved += (words + 63) & ~(size_t)63;
    return HS_OK;
  };
  // Fills on the stream: memsets, or a fill kernel in the no-allocation form
  // (which callers may run where memset nodes are unwelcome).
  auto fill = [&](uint32_t *p, size_t words, uint32_t v) -> int {
    if (!na) {
      HS_HIP(hipMemsetAsync(p, v == 0 ? 0 : 0xff, words * 4, c->stream));
      return HS_OK;
    }
    const int blocks = (int)std::min<size_t>((words + 255) / 256, 1024);
    k_fill<<<blocks > 0 ? blocks : 1, 256, 0, c->stream>>>(p, words, v);
    HS_HIP(hipGetLastError());
    return HS_OK;
  };
  int32_t *d_storage = na ? na->d_storage : nullptr;
  uint32_t *d_ctl, *d_locks, *d_ring = nullptr, *d_bar, *d_flags;
  int rc;
  if (!na && (rc = arena.alloc(&d_storage, n))) return rc;
  if ((rc = take(&d_ctl, 64))) return rc;
  if ((rc = take(&d_locks, (size_t)nlocks * kLockWords))) return rc;
  if (sleep_mx && (rc = take(&d_ring, (size_t)nlocks * ring_len * ring_stride))) return rc;
  // barrier words, kBarWords blocks: [global][XCD-aware sub-counters:
  // nxcd x sub_max][groups: nab][release words: nxcd][episode count line]
  int sub_max = 0, n_sub = 0;  // XCD-aware tree: sub-counters per XCD, in all
  const bool lf = p == 2 || p == 3;
  const bool a2a = xcd_mode && lf && lf_xcd_form(nab, per, nxcd, wg);
  const int sub = a2a ? (nab / nxcd + (nab % nxcd ? 1 : 0)) * per : sub_group_size(nw, lf);
  if (xcd_mode)
    for (int x = 0; x < nxcd; ++x) {
      const int nwx = (nab / nxcd + (x < nab % nxcd ? 1 : 0)) * per;  // WGs on XCD x
      sub_max = std::max(sub_max, (nwx + sub - 1) / sub);
      n_sub += (nwx + sub - 1) / sub;
    }
  const size_t o_sub = kBarWords, o_grp = o_sub + (size_t)kBarWords * nxcd * sub_max;
  const size_t o_rel = o_grp + (size_t)kBarWords * nab, o_ep = o_rel + (size_t)kBarWords * nxcd;
  const size_t bar_words = o_ep + HS_LINE_WORDS;
  if ((rc = take(&d_bar, bar_words))) return rc;
  // strict: packed flag arrays; XCD: one line per flag (kFlagStride)
  // flags: strict [gin: nab][gout: nab][lin, lout: nab x per], packed;
  // XCD-aware [gout: nxcd][din: nxcd x sub_max][sin: nxcd x sub_max x
  // sub][lin, lout: nab x per], one line per flag
  const size_t n_sub_slots = (size_t)nxcd * sub_max;
  const size_t x_flags = nxcd + n_sub_slots * (1 + sub) + 2 * (size_t)nab * per;
  const size_t flag_words = (xcd_mode ? x_flags * kFlagStride
                                      : 2 * (size_t)nab + 2 * (size_t)nab * per) +
                            4 * HS_LINE_WORDS;
  if ((rc = take(&d_flags, flag_words))) return rc;
  uint32_t *d_xmap, *d_flat;
  if ((rc = take(&d_xmap, (size_t)reps * nxcd))) return rc;
  if ((rc = take(&d_flat, 2 * (size_t)kBarWords))) return rc;
  // the scratch keeps the last call's contents: every word used is cleared
  if ((rc = fill(d_flat, 2 * kBarWords, 0))) return rc;
  if ((rc = fill(d_xmap, (size_t)reps * nxcd, 0xffffffffu))) return rc;
  if ((rc = fill(d_ctl, 64, 0))) return rc;
  if ((rc = fill(d_locks, (size_t)nlocks * kLockWords, 0))) return rc;
  if ((rc = fill(d_bar, bar_words, 0))) return rc;
  if ((rc = fill(d_flags, flag_words, 0))) return rc;
  const uint32_t sem_max = sem_count_of(p) ? sem_count_of(p) + 1 : 0;  // count+1, HSS:15
  if (sem_max || d_ring) {
    long work = std::max<long>(nlocks, nlocks * ring_len);
    int blocks = (int)std::min<long>((work + 255) / 256, 4096);
    k_init_locks<<<blocks, 256, 0, c->stream>>>(d_locks, nlocks, sem_max, d_ring,
                                                ring_len, ring_stride);
    HS_HIP(hipGetLastError());
  }
  if (!na)
    HS_HIP(hipMemcpyAsync(d_storage, storage, n * sizeof(int32_t), hipMemcpyHostToDevice,
                          c->stream));
  HS_HIP(hipStreamSynchronize(c->stream));

  SPArgs a;
  std::memset(&a, 0, sizeof(a));
  a.storage = d_storage;
  a.ctl = d_ctl;
  a.locks = d_locks;
  a.ring = d_ring;
  a.ring_len = ring_len;
  a.ring_stride = ring_stride;
  uint32_t *episodes = d_bar + o_ep;
  // strict: global block, nab group blocks;
  // XCD: global, nxcd XCD blocks, nab group blocks, nxcd release words
  a.tb = TreeBar{d_bar, d_bar + o_grp, episodes, nab, per};
  a.xt = XcdTree{d_bar, d_bar + o_sub, d_bar + o_grp, d_bar + o_rel, episodes, nab, per, nxcd,
                 sub_max, n_sub, sub};
  uint32_t *fl = d_flags + 4 * HS_LINE_WORDS;
  a.fb = FlagBar{fl, fl + nab, fl + 2 * nab, fl + 2 * nab + (size_t)nab * per, episodes,
                 nab, per};
  {
    const size_t S = kFlagStride, o_din = nxcd, o_sin = o_din + n_sub_slots;
    const size_t o_lin = o_sin + n_sub_slots * sub, o_lout = o_lin + (size_t)nab * per;
    a.xf = XcdFlags{fl, fl + o_din * S, fl + o_sin * S, fl + o_lin * S, fl + o_lout * S,
                    episodes, nab, per, nxcd, sub_max, sub, a2a ? 1 : 0};
  }
  a.flat = d_flat;
  a.nxcd = nxcd;
  a.L = L;
  a.inc_bar = cfg->mode == 1 && inc_band_barrier(wg, L);
  a.iters = iters;
  a.num_cu = ncu;
  a.nab = nab;
  a.per_cu = per;
  a.bo.min_q = cfg->backoff_min > 0 ? cfg->backoff_min : 1;
  a.bo.max_q = cfg->backoff_max > 0 ? cfg->backoff_max : 1023;
  // Default: one sleep quantum between polls, none for the XCD-aware tree
  // barriers, whose waits mostly poll lines kept in their XCD's L2 (1024 x
  // 256, numLdSt 10: lfTreeBarrUniq 3.91 -> 3.85 us per episode, the
  // atomic tree 3.29 -> 3.28).
  const bool tree = p <= 3;
  a.bo.poll_q = cfg->poll_sleep < 0 ? 0
                : cfg->poll_sleep > 0 ? cfg->poll_sleep
                : (tree && xcd_mode) ? 0 : 1;
  a.ticks = timeout_ticks(c, cfg->timeout_s);
  const bool lx = p == 1 || p == 3 || p == 38 || p == 40 || p == 42;
  const bool srb = p >= 37 && p <= 42;
  const bool srb_tree = (p == 37 || p == 38) && per > 4;
  const uint32_t local_per_iter = (lx && per > 1) ? 2 : 1;

  // ---- region of interest (main.hip.cpp:2354-2511) -----------------------
  if ((rc = load_kernel((const void *)kfn))) return rc;  // not inside the ROI
  if (na) {  // no events: the host clock around the launches and their syncs
    HS_HIP(hipStreamSynchronize(c->stream));
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) {
      a.xcd_map = d_xmap + (size_t)r * nxcd;
      a.e_global0 = (uint32_t)r * iters;
      a.e_local0 = (uint32_t)r * iters * local_per_iter;
      hipLaunchKernelGGL(kfn, dim3(nw), dim3(wg), 0, c->stream, a);
      HS_HIP(hipGetLastError());
      HS_HIP(hipStreamSynchronize(c->stream));
    }
    res->roi_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    res->kernel_ms = res->roi_ms;
  }
  EventSet ev;  // [0] ROI start, [1] ROI end, then a pair per launch
  if (!na && (rc = ev.create(2 + 2 * (size_t)reps))) return rc;
  const hipEvent_t ev0 = na ? nullptr : ev[0], ev1 = na ? nullptr : ev[1];
  auto kev = [&](int i) { return ev[2 + i]; };
  int launch_rc = HS_OK;
  if (!na) {
  HS_HIP(hipDeviceSynchronize());
  HS_HIP(hipEventRecord(ev0, c->stream));
  for (int r = 0; r < reps; ++r) {
    a.xcd_map = d_xmap + (size_t)r * nxcd;
    a.e_global0 = (uint32_t)r * iters;
    a.e_local0 = (uint32_t)r * iters * local_per_iter;
    HS_HIP(hipEventRecord(kev(2 * r), c->stream));
    hipLaunchKernelGGL(kfn, dim3(nw), dim3(wg), 0, c->stream, a);
    if (hipGetLastError() != hipSuccess) { launch_rc = HS_EHIP; break; }
    HS_HIP(hipEventRecord(kev(2 * r + 1), c->stream));
    HS_HIP(hipStreamSynchronize(c->stream));
  }
  HS_HIP(hipEventRecord(ev1, c->stream));
  HS_HIP(hipEventSynchronize(ev1));
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  res->roi_ms = ms;
  double kms = 0;
  for (int r = 0; r < reps && launch_rc == HS_OK; ++r) {
    float k = 0;
    HS_HIP(hipEventElapsedTime(&k, kev(2 * r), kev(2 * r + 1)));
    kms += k;
  }
  res->kernel_ms = kms;
  }
  if (launch_rc) return launch_rc;

  // ---- results and idle-state invariants ----------------------------------
  if (!na) HS_HIP(hipMemcpy(storage, d_storage, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  uint32_t ctlw[64];
  HS_HIP(hipMemcpy(ctlw, d_ctl, sizeof ctlw, hipMemcpyDeviceToHost));
  const uint32_t abort_w = ctlw[0];
  if (topology_fail) *topology_fail = ctlw[32] != 0;
  std::vector<uint32_t> locks((size_t)nlocks * kLockWords), bar(bar_words), flags(flag_words),
      flat(2 * kBarWords);
  HS_HIP(hipMemcpy(locks.data(), d_locks, locks.size() * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(flat.data(), d_flat, flat.size() * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(bar.data(), d_bar, bar.size() * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(flags.data(), d_flags, flags.size() * 4, hipMemcpyDeviceToHost));
  const uint64_t total_ep = (uint64_t)iters * reps;
  bool ok = true;
  if (is_mutex(p) || is_sem(p)) {
    for (int l = 0; l < nlocks; ++l) {
      const uint32_t *lk = &locks[(size_t)l * kLockWords];
      res->cs_entries += lk[kLkCs];
      res->rd_entries += lk[kLkRd];
      // per-lock acquisitions: WGs mapped to this lock x iters x reps
      long wgs_here = uniq ? (nw / ncu + (l < nw % ncu ? 1 : 0)) : nw;
      if (is_sem(p) && uniq) wgs_here = (l < nab) ? per : 0;
      const uint64_t acq = (uint64_t)wgs_here * iters * reps;
      if (p == 4 || p == 5 || p == 20 || p == 21) ok = ok && lk[kLkHead] == 0;
      if (p == 7 || p == 23)
        ok = ok && lk[kLkHead] == (uint32_t)acq && lk[kLkTail] == (uint32_t)acq;
      if (sleep_mx) {
        ok = ok && lk[kLkTail] == (uint32_t)acq;
        if (acq + 1 <= (uint64_t)ring_len) {
          std::vector<uint32_t> ring(ring_len * ring_stride);
          HS_HIP(hipMemcpy(ring.data(), d_ring + (size_t)l * ring_len * ring_stride,
                           ring.size() * 4, hipMemcpyDeviceToHost));
          for (long k = 0; k < ring_len; ++k)
            ok = ok && ring[k * ring_stride] == ((uint64_t)k == acq ? 1u : 0xffffffffu);
        } else {
          ok = false;
        }
      }
      // count back to max, no writer waiting; a semaphore nobody entered
      // keeps its initial word (the writer goes first, HSS:23)
      if (is_sem(p) && (!uniq || l < nab))
        ok = ok && lk[kLkHead] == (acq ? sem_max : sem_max | kSemWaitBit);
      if (is_prio(p)) ok = ok && lk[kLkPrio] == 0;  // no post in flight
    }
    const uint32_t *lk0 = &locks[0];
    const bool semw = is_sem(p);
    const uint32_t w[8] = {lk0[kLkHead], semw ? (lk0[kLkHead] & 0xffffu) : 0u,
                           semw ? (lk0[kLkHead] >> 16) : 0u, lk0[kLkMax],
                           lk0[kLkTail], lk0[kLkCs], lk0[kLkRd], (uint32_t)nlocks};
    std::memcpy(res->lock_words, w, sizeof(w));
  } else {
    const uint32_t ep = bar[o_ep];
    res->episodes = ep;
    ok = ok && ep == (uint32_t)total_ep;
    const uint32_t te = (uint32_t)total_ep, tl = (uint32_t)(total_ep * local_per_iter);
    if (srb && !srb_tree) {
      // flat barriers: the arrival count wraps to 0 at every episode's last
      // arrival (atomicInc), and the generation counts the episodes
      auto flat_ok = [&](int k) {
        return flat[(size_t)k * kBarWords] == 0 && flat[(size_t)k * kBarWords + 32] == te;
      };
      ok = ok && flat_ok(0);
      if ((p == 40 || p == 42) && per > 1) ok = ok && flat_ok(1);
      if (p == 38 && per > 1) {  // per-group local barrier
        // counters back at 0 (reset each episode, or wrapped by atomicInc)
        const size_t g0 = o_grp;
        for (int gi = 0; gi < nab; ++gi)
          ok = ok && bar[g0 + kBarWords * gi] == 0 && bar[g0 + kBarWords * gi + 32] == te;
      }
      const uint32_t w[8] = {flat[0], flat[1], flat[32], flat[kBarWords + 32], ep, 0, 0, 0};
      std::memcpy(res->lock_words, w, sizeof(w));
    } else if (p <= 1 || srb) {
      const size_t g0 = o_grp;
      if (xcd_mode) {
        // atomicInc counters wrap to 0 at each episode's last arrival
        const int nx = std::min(nab, nxcd);
        ok = ok && bar[0] == 0 && bar[32] == te;
        for (int x = 0; x < nx; ++x) {
          const int nwx = (nab / nxcd + (x < nab % nxcd ? 1 : 0)) * per;
          for (int k = 0; k < sub_max; ++k) {
            const int ck = std::max(0, std::min(sub, nwx - k * sub));
            ok = ok && (ck == 0 || bar[o_sub + (size_t)kBarWords * (x * sub_max + k)] == 0);
          }
          ok = ok && bar[o_rel + (size_t)kBarWords * x] == te;  // release word
        }
        if (per > 1)  // groups: local-exchange arrivals only
          for (int gi = 0; gi < nab; ++gi)
            ok = ok && bar[g0 + kBarWords * gi] == 0 &&
                 bar[g0 + kBarWords * gi + 32] == (lx ? te : 0u);
      } else {
        ok = ok && bar[0] == 0 && bar[32] == te;
        if (per > 1)
          for (int gi = 0; gi < nab; ++gi)
            ok = ok && bar[g0 + kBarWords * gi] == 0 && bar[g0 + kBarWords * gi + 32] == tl;
      }
      const uint32_t w[8] = {bar[0], bar[32], bar[g0], bar[g0 + 32], ep, 0, 0, 0};
      std::memcpy(res->lock_words, w, sizeof(w));
    } else {
      const uint32_t *f = &flags[4 * HS_LINE_WORDS];
      const size_t S = xcd_mode ? kFlagStride : 1;  // words between flags
      const uint32_t *lin, *lout;
      uint32_t lin_want = tl, lout_want = tl;
      uint32_t w1 = 0;
      if (xcd_mode) {
        // release words gout[x]; arrival flags at the sub and device levels
        // carry episodes; group flags carry local epochs (local exchange only)
        const uint32_t *gout = f, *din = f + nxcd * S, *sinf = din + n_sub_slots * S;
        const int nx = std::min(nab, nxcd);
        for (int x = 0; x < nx; ++x) {
          ok = ok && gout[x * S] == te;
          const int nwx = (nab / nxcd + (x < nab % nxcd ? 1 : 0)) * per;
          for (int k = 0; k * sub < nwx; ++k) {
            const size_t sid = (size_t)x * sub_max + k;
            if (sid != 0) ok = ok && din[sid * S] == te;
            const int ck = std::min(sub, nwx - k * sub);
            for (int m = 1; m < ck; ++m) ok = ok && sinf[(sid * sub + m) * S] == te;
          }
        }
        lin = sinf + n_sub_slots * sub * S;
        lin_want = lx ? tl - 1 : 0;  // the group barrier is all-to-all: no out-flags
        lout_want = 0;
        w1 = n_sub_slots > 1 ? din[S] : 0;
      } else {
        const uint32_t *gin = f, *gout = f + nab;
        for (int gi = 1; gi < nab; ++gi) ok = ok && gin[gi] == te && gout[gi] == te;
        lin = gout + nab;
        w1 = f[1 + nab];
      }
      lout = lin + (size_t)nab * per * S;
      if (per > 1)
        for (int gi = 0; gi < nab; ++gi)
          for (int m = 1; m < per; ++m)
            ok = ok && lin[((size_t)gi * per + m) * S] == lin_want &&
                 lout[((size_t)gi * per + m) * S] == lout_want;
      const uint32_t w[8] = {f[S], w1, per > 1 ? lin[S] : 0, per > 1 ? lout[S] : 0, ep, 0, 0, 0};
      std::memcpy(res->lock_words, w, sizeof(w));
    }
  }
  res->state_ok = ok ? 1 : 0;
  if (abort_w) return HS_ETIMEOUT;
  return HS_OK;
}

}  // namespace
