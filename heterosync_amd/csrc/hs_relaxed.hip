// hs_relaxed.hip -- relaxed-atomics kernels: seqlock, flags, reference
// counter and split counter, rewritten for gfx950 wave64.
//
//   seqlock     hip/relaxedAtomics/uvm/seqlocks/seqlocks_kernel.h:23-240,
//               seqlocks_hipAtomics/seqlocks_kernel.h:23-134
//   flags       cuda/relaxedAtomics/uvm/flags/flags_kernel.cu:13-242
//   refCounter  cuda/relaxedAtomics/uvm/refCounter/refCounter_kernel.cu:4-55
//   splitCounter cuda/relaxedAtomics/uvm/splitCounter/splitCounter_kernel.cu:10-103
//
// Every sync word is an agent-scope atomic; data words touched by several
// workgroups (seqlock data) are atomics too, so the kernels are race-free
// under the HIP memory model.  Waits are bounded by the watchdog Guard.
#include <algorithm>
#include <cstring>
#include <vector>

#include "hs_host.h"
#include "../../include/hs_detail/hs_syncprims_dev.h"

using namespace hs;

namespace {

// Per-WG event counter in LDS, flushed with one atomic at the end.
__device__ __forceinline__ void flush_count(uint64_t *dst, uint32_t *s_cnt) {
  __syncthreads();
  if (is_master() && *s_cnt)
    __hip_atomic_fetch_add(dst, (uint64_t)*s_cnt, __ATOMIC_RELAXED, HS_AGENT);
}

// ---------------------------------------------------------------------------
// seqlock.  WG b is a writer iff b % writer_period == 0 (reference: 16,
// HQ:206), lock index b % group_size (HQ:201).  Each lock owns a pair of
// 64-word data rows (lane = data word), so a reader's snapshot can be
// checked: with a stable even sequence S the two rows must both hold S-2
// (the writer that moved S-2 -> S stored its starting value) or 0 for S=0.
//   variant 0: acquire/release scoped atomics, readers load (HQA:23-109)
//   variant 1: relaxed atomics + agent fences, readers exchange 0 into the
//              data as the reference's DeNovo-style "exch as load" (HQ:23-95)
// STRICT keeps the reference's acquire/release orderings (each an agent-
// scope cache write-back or invalidate, serialised per XCD).  Without it the
// orderings come from completion alone: every access to the sequence word
// and the rows is an agent-scope atomic (L1 bypassed, stores written
// through), so a load performed after s_waitcnt has drained the wave's
// earlier accesses sees them in order, and a store issued after the drain
// becomes visible after them.  Readers: sequence load, row loads, drain,
// sequence load.  Writers: CAS odd, row stores, drain, sequence store.
struct SeqArgs {
  uint32_t *seq;    // n_seq words, each on its own line (stride 32)
  int32_t *data;    // n_seq * 2 * 64 words
  int32_t *out;     // num_wgs * wg
  uint32_t *ctl;    // abort
  uint64_t *stats;  // [0] torn snapshots, [1] atomic ops, [2] reader retries
  int group_size, writer_period, iters;
  uint64_t ticks;
};

template <int VARIANT, bool STRICT>
__global__ void __launch_bounds__(1024) k_seqlock(SeqArgs a) {
  __shared__ uint32_t s_seq0, s_torn, s_ops;
  __shared__ int s_state;
  __shared__ int32_t s_snap[64];
  Guard g;
  g.init(a.ctl, a.ticks);
  const int b = blockIdx.x, lane = threadIdx.x & 63;
  const bool wave0 = threadIdx.x < 64;
  const int lock = b % a.group_size;
  uint32_t *seq = a.seq + (long)lock * HS_LINE_WORDS;
  int32_t *d0 = a.data + (long)lock * 128, *d1 = d0 + 64;
  const bool writer = (b % a.writer_period) == 0;
  if (is_master()) {
    s_torn = 0;
    s_ops = 0;
  }
  if (wave0) s_snap[lane] = 0;
  __syncthreads();
  bool alive = true;
  for (int it = 0; it < a.iters && alive; ++it) {
    if (writer) {
      if (is_master()) {  // HQ:109-130: wait for an even sequence, CAS it odd
        uint32_t s0 = (STRICT && VARIANT == 0) ? ld_acq(seq) : ld_rlx(seq);
        uint32_t ops = 1, bo = 1;
        int st = 1;
        while ((s0 & 1u) ||
               (STRICT ? cas_acq(seq, s0, s0 + 1) : cas_rlx(seq, s0, s0 + 1)) != s0) {
          ++ops;
          if (!g.ok()) { st = -1; break; }
          sleep_quanta(bo);
          bo = ((bo << 1) + 1) & 127u;  // MAX_BACKOFF 128 (HQ:16)
          s0 = (STRICT && VARIANT == 0) ? ld_acq(seq) : ld_rlx(seq);
          ++ops;
        }
        s_seq0 = s0;
        s_state = st;
        s_ops += ops + 2;
      }
      __syncthreads();
      const uint32_t s0 = s_seq0;
      alive = s_state > 0;
      if (alive && wave0) {  // the 64-word rows belong to the first wave
        if (STRICT) {
          if (VARIANT == 0) fence_rel(); else fence_acq();
        }
        __hip_atomic_store(d0 + lane, (int32_t)s0, __ATOMIC_RELAXED, HS_AGENT);
        __hip_atomic_store(d1 + lane, (int32_t)s0, __ATOMIC_RELAXED, HS_AGENT);
        if (!STRICT) wait_mem();  // the rows are visible before the sequence moves on
      }
      __syncthreads();
      if (alive && is_master()) {  // HQ:133-135
        if (STRICT) st_rel(seq, s0 + 2); else st_rlx(seq, s0 + 2);
      }
    } else {
      for (;;) {  // HQ:23-56: retry until the sequence is even and unchanged
        if (is_master()) {  // the value is stored, so the load has completed
          s_seq0 = !STRICT         ? ld_rlx(seq)
                   : VARIANT == 0 ? ld_acq(seq)
                                  : (fence_acq(), ld_rlx(seq));
          s_ops += 2;
        }
        __syncthreads();
        const uint32_t sa = s_seq0;
        int32_t r1 = 0, r2 = 0;
        if (wave0) {
          if (VARIANT == 0) {
            r1 = __hip_atomic_load(d0 + lane, __ATOMIC_RELAXED, HS_AGENT);
            r2 = __hip_atomic_load(d1 + lane, __ATOMIC_RELAXED, HS_AGENT);
          } else {
            r1 = __hip_atomic_exchange(d0 + lane, 0, __ATOMIC_RELAXED, HS_AGENT);
            r2 = __hip_atomic_exchange(d1 + lane, 0, __ATOMIC_RELAXED, HS_AGENT);
          }
        }
        if (is_master()) {
          // the row loads complete before the second read
          if (STRICT) fence_acq(); else wait_mem();
          const uint32_t sb = ld_rlx(seq);
          s_state = (sa == sb && !(sa & 1u)) ? 1 : (g.ok() ? 0 : -1);
        }
        __syncthreads();
        const int st = s_state;
        if (st == 1 && wave0) {
          s_snap[lane] = r1 + r2;
          const int32_t want = sa == 0 ? 0 : (int32_t)(sa - 2);
          if (VARIANT == 0 && (r1 != want || r2 != want)) atomicAdd(&s_torn, 1u);
        }
        __syncthreads();  // every wave has read s_state/s_seq0
        if (st != 0) {
          alive = st > 0;
          break;
        }
      }
      if (alive) a.out[(long)b * blockDim.x + threadIdx.x] = s_snap[lane];  // HQ:56
    }
  }
  __syncthreads();
  if (is_master()) {
    if (s_torn) __hip_atomic_fetch_add(a.stats + 0, (uint64_t)s_torn, __ATOMIC_RELAXED, HS_AGENT);
    __hip_atomic_fetch_add(a.stats + 1, (uint64_t)s_ops, __ATOMIC_RELAXED, HS_AGENT);
  }
}

// ---------------------------------------------------------------------------
// flags (CFK:13-242).  The last WG is the main WG: LDS work, stop += 1,
// flag tree barrier, then it reads `dirty` and writes its output rows if
// set.  Workers run numIters iterations (and keep polling until stop is
// seen), bumping `dirty` when their random value is <= cutoff; then they
// join the barrier.  Only it < numIters iterations can set dirty, so the
// final count is exact: sum over workers and it < numIters of
// [rand[b + it*numTBs] <= cutoff].
// The reference's `dirty` is one word that every dirty iteration bumps
// (CFK:138) and the main TB reads after the barrier (CFK:73); on MI355X RMWs to one address serialise at about 11 ns, so
// that word alone bounds the run.  Here every dirty iteration still makes
// one relaxed atomic add, to its own WG's line of `dirty`, and the main WG
// sums the lines after the barrier (one per lane) before it tests the flag.
struct FlagsArgs {
  uint32_t *stop;   // own line
  uint32_t *dirty;  // one line per WG
  uint32_t *dirty_total;  // the main WG's sum, for the host
  uint32_t *out;
  const float *rnd;
  uint32_t *ctl;
  uint64_t *stats;
  FlagBar fb;
  int accs, repeats, iters;
  float cutoff;
  uint64_t ticks;
};

__global__ void __launch_bounds__(1024) k_flags(FlagsArgs a) {
  __shared__ uint32_t lds[1024];
  __shared__ int s_stop, s_ok, s_flag;
  __shared__ uint32_t s_ops, s_dirty;
  Guard g;
  g.init(a.ctl, a.ticks);
  const int b = blockIdx.x, grid = gridDim.x;
  const bool main_wg = b == grid - 1;
  const int grp = b % a.fb.nab, q = b / a.fb.nab;
  lds[threadIdx.x] = main_wg ? 0 : threadIdx.x;
  if (is_master()) {
    s_ok = 1;
    s_ops = 0;
  }
  __syncthreads();
  if (main_wg) {  // CFK:28-60
    for (int i = 0; i < a.repeats; ++i) {
      for (int j = 0; j < a.accs; ++j) lds[threadIdx.x] += a.accs;
      __syncthreads();
    }
    if (is_master()) {
      add_rel(a.stop, 1);
      ++s_ops;
    }
  } else {  // CFK:88-176
    uint32_t bo = 2;
    if (is_master()) {
      s_stop = ld_rlx(a.stop) != 0;
      ++s_ops;
    }
    const int dummy = a.accs / 4;
    // Neither load of an iteration waits in line: the random value is read
    // one iteration ahead (it is input), and the stop flag's load is issued
    // before the iteration's work and consumed after it.
    float r_next = a.iters > 0 ? a.rnd[b] : 1.0f;
    for (int it = 0;; ++it) {
      __syncthreads();
      const bool go = (!s_stop || it < a.iters) && s_ok;
      __syncthreads();  // all waves have read s_stop/s_ok
      if (!go) break;
      const float r = r_next;
      if (it + 1 < a.iters) r_next = a.rnd[(long)b + (long)(it + 1) * grid];
      uint32_t stop_v = 0;
      if (is_master()) stop_v = ld_rlx(a.stop);
      if (it < a.iters && r <= a.cutoff) {
        for (int i = 0; i < dummy; ++i) lds[threadIdx.x]++;
        __syncthreads();
        if (is_master()) {
          add_rlx(a.dirty + (long)b * HS_LINE_WORDS, 1);
          ++s_ops;
        }
      } else if (is_master()) {
        bo = (bo * 2) & 255u;  // MAX_BACKOFF_EXP 256 (CFK:6)
        sleep_quanta(bo >> 3);
      }
      if (is_master()) {
        // an opaque use here, so the compiler does not fold the test (and
        // the wait for the load) into the branch that issued it
        asm volatile("" : "+v"(stop_v));
        s_stop = stop_v != 0;
        ++s_ops;
        if (!s_stop && it >= a.iters && !g.ok()) s_ok = 0;
      }
    }
  }
  __syncthreads();
  bool ok = s_ok;
  if (ok) ok = flag_tree_barrier(a.fb, grp, q, 1, 1, false, 1, g, &s_flag);
  if (main_wg && ok) {  // CFK:62-77
    if (is_master()) {
      fence_acq();
      s_dirty = 0;
    }
    __syncthreads();
    uint32_t part = 0;
    for (int k = threadIdx.x; k < grid; k += blockDim.x)
      part += ld_rlx(a.dirty + (long)k * HS_LINE_WORDS);
    if (part) atomicAdd(&s_dirty, part);
    __syncthreads();
    if (is_master()) {
      st_rlx(a.dirty_total, s_dirty);
      s_ops += (uint32_t)grid;
    }
    if (s_dirty) a.out[(long)b * blockDim.x + threadIdx.x] = lds[threadIdx.x];
  }
  __syncthreads();
  if (is_master())
    __hip_atomic_fetch_add(a.stats + 1, (uint64_t)s_ops, __ATOMIC_RELAXED, HS_AGENT);
}

// ---------------------------------------------------------------------------
// refCounter (CRK:4-55).  Thread t of WG b owns counter
// (b % numGroups) * wg + t in both arrays, shared with the same lane of the
// `sharers` WGs of its group.  Per repeat: +1 on both, `spacing` LDS
// steps, then CUDA-semantics atomicDec(c, 1e9) on both; the thread whose
// decrement saw old <= 1 marks its del entry.  Releases precede the
// decrements and an acquire follows a final one (refcount discipline).
struct RefArgs {
  uint32_t *c0, *c1, *del0, *del1;
  uint64_t *stats;
  int num_groups, repeats, spacing;
};

// STRICT keeps the agent-scope release fence before the decrements and the
// acquire fence after a final one.  Without it the counters, the only
// memory the benchmark shares, are agent-scope atomics: draining the wave's
// increments (s_waitcnt) orders them before the decrements, and the del
// store depends on the decrement's returned value, so no cache maintenance
// (one write-back per wave per repeat, serialised per XCD) is needed.
template <bool STRICT>
__global__ void __launch_bounds__(1024) k_refcounter(RefArgs a) {
  __shared__ volatile int lds[1024];
  const int b = blockIdx.x, t = threadIdx.x;
  const long me = (long)b * blockDim.x + t;
  const long ci = (long)(b % a.num_groups) * blockDim.x + t;
  uint32_t *p0 = a.c0 + ci, *p1 = a.c1 + ci;
  lds[t] = 0;
  __syncthreads();
  for (int i = 0; i < a.repeats; ++i) {
    add_rlx(p0, 1);
    add_rlx(p1, 1);
    for (int j = 0; j < a.spacing; ++j) {
      lds[t] += j;
      __syncthreads();
    }
    if (STRICT) fence_rel(); else wait_mem();
    const uint32_t o0 = __builtin_amdgcn_atomic_dec32(p0, 1000000000u, __ATOMIC_RELAXED, "agent");
    const uint32_t o1 = __builtin_amdgcn_atomic_dec32(p1, 1000000000u, __ATOMIC_RELAXED, "agent");
    if (STRICT && (o0 <= 1 || o1 <= 1)) fence_acq();
    if (o0 <= 1) a.del0[me] = 1;
    if (o1 <= 1) a.del1[me] = 1;
  }
}

// ---------------------------------------------------------------------------
// splitCounter (CSK:10-103), wave-granular: even waves of a WG are writers
// (one relaxed atomicAdd(+1) per iteration on their own counter), odd waves
// are readers.  The reference's read is WG-collective: the reader threads
// of a TB share one sweep and all output its sum (CSK:33-76, 95).  Here the
// first reader wave makes the WG's read every `read_every` iterations: it
// loads the WG's writer counters, reduces them across the wave and
// publishes the WG's partial sum, then reads all partials (a strided sweep
// shared by the lanes) for the approximate global total, which every
// reader thread outputs at the end.  Counters are exact at the end (iters
// each); reader totals lie in [0, total].
struct SplitArgs {
  uint32_t *counters;  // num_wgs * wg/2
  uint32_t *partial;   // num_wgs
  uint32_t *out;       // num_wgs * wg
  uint64_t *stats;
  int iters, read_every;
};

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of p[lane], p[lane + 64], ... below n, by relaxed loads issued eight at
// a time before any is added: the words are independent and the sum is an
// approximate snapshot anyway, so the loads need not wait for each other (a
// load-add loop waits out one memory round trip per word).
__device__ __forceinline__ uint32_t sweep_sum(const uint32_t *p, int n, int lane) {
  uint32_t s = 0;
  for (int k = lane; k < n; k += 8 * 64) {
    uint32_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = k + u * 64 < n ? ld_rlx(p + k + u * 64) : 0u;
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  return s;
}

__global__ void __launch_bounds__(1024) k_splitcounter(SplitArgs a) {
  const int b = blockIdx.x, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int half = blockDim.x >> 1;  // writer threads per WG
  uint32_t *mine = a.counters + (long)b * half;
  const int nw = gridDim.x;
  __shared__ uint32_t s_total;
  uint64_t ops = 0;
  if ((wave & 1) == 0) {
    uint32_t *c = mine + (wave >> 1) * 64 + lane;
    for (int i = 0; i < a.iters; ++i) add_rlx(c, 1u);
    ops = a.iters;
  } else if (wave == 1) {  // the WG's read, made once (CSK:33-76 is WG-collective)
    uint32_t total = 0;
    for (int i = a.read_every; i <= a.iters; i += a.read_every) {
      const uint32_t s = wave_sum(sweep_sum(mine, half, lane));
      if (lane == 0) st_rlx(a.partial + b, s);
      total = wave_sum(sweep_sum(a.partial, nw, lane));
      ops += (half + 63) / 64 + (nw + 63) / 64 + 1;
    }
    if (lane == 0) s_total = total;
  }
  __syncthreads();
  if (wave & 1) a.out[(long)b * blockDim.x + threadIdx.x] = s_total;  // every reader thread
  if (lane == 0)
    __hip_atomic_fetch_add(a.stats + 1, ops * 64, __ATOMIC_RELAXED, HS_AGENT);
}

// ---------------------------------------------------------------------------
struct Timer {
  hipEvent_t e0{}, e1{};
  Timer() {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
  }
  ~Timer() {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
};

int read_stats(uint64_t *d_stats, uint32_t *d_ctl, hs_relaxed_result *res,
               uint64_t *torn) {
  uint64_t st[4];
  uint32_t ab = 0;
  HS_HIP(hipMemcpy(st, d_stats, sizeof st, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(&ab, d_ctl, 4, hipMemcpyDeviceToHost));
  if (torn) *torn = st[0];
  res->atomic_ops = st[1];
  res->aborted = ab != 0;
  return ab ? HS_ETIMEOUT : HS_OK;
}

}  // namespace

extern "C" int hs_seqlock_run(hs_ctx *c, int num_wgs, int wg, int group_size,
                              int writer_period, int iters, int variant,
                              uint32_t *seq_out, int n_seq, int32_t *out_arr,
                              uint64_t *torn, hs_relaxed_result *res) {
  if (!c || !seq_out || !out_arr || !res || num_wgs < 1 || wg < 64 || wg % 64 ||
      wg > 1024 || group_size < 1 || writer_period < 1 || iters < 0 ||
      n_seq < group_size || variant < 0 || variant > 3)
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  HS_HIP(hipSetDevice(c->device));
  // bit 0: the reference variant; bit 1: its strict (fenced) orderings
  static void (*const kKernels[4])(SeqArgs) = {k_seqlock<0, false>, k_seqlock<1, false>,
                                               k_seqlock<0, true>, k_seqlock<1, true>};
  void (*kfn)(SeqArgs) = kKernels[variant];
  int fit = 0, rc;
  // writers and readers of one lock wait on each other
  if ((rc = coresident_wgs(c, (const void *)kfn, wg, &fit))) return rc;
  if (fit < num_wgs) return HS_EOCCUPANCY;
  DevArena ar;
  uint32_t *d_seq, *d_ctl;
  int32_t *d_data, *d_out;
  uint64_t *d_stats;
  if ((rc = ar.alloc(&d_seq, (size_t)n_seq * HS_LINE_WORDS))) return rc;
  if ((rc = ar.alloc(&d_data, (size_t)n_seq * 128))) return rc;
  if ((rc = ar.alloc(&d_out, (size_t)num_wgs * wg))) return rc;
  if ((rc = ar.alloc(&d_ctl, 64))) return rc;
  if ((rc = ar.alloc(&d_stats, 4))) return rc;
  HS_HIP(hipMemset(d_seq, 0, (size_t)n_seq * HS_LINE_WORDS * 4));
  HS_HIP(hipMemset(d_data, 0, (size_t)n_seq * 128 * 4));
  HS_HIP(hipMemset(d_out, 0xff, (size_t)num_wgs * wg * 4));  // -1 (seqlocks.cpp:56)
  HS_HIP(hipMemset(d_ctl, 0, 256));
  HS_HIP(hipMemset(d_stats, 0, 32));
  SeqArgs a{d_seq, d_data, d_out, d_ctl, d_stats, group_size, writer_period, iters,
            timeout_ticks(c, 0)};
  Timer t;
  HS_HIP(hipDeviceSynchronize());
  HS_HIP(hipEventRecord(t.e0, c->stream));
  hipLaunchKernelGGL(kfn, dim3(num_wgs), dim3(wg), 0, c->stream, a);
  HS_HIP(hipGetLastError());
  HS_HIP(hipEventRecord(t.e1, c->stream));
  HS_HIP(hipEventSynchronize(t.e1));
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, t.e0, t.e1));
  res->kernel_ms = ms;
  std::vector<uint32_t> seq((size_t)n_seq * HS_LINE_WORDS);
  HS_HIP(hipMemcpy(seq.data(), d_seq, seq.size() * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < n_seq; ++i) seq_out[i] = seq[(size_t)i * HS_LINE_WORDS];
  HS_HIP(hipMemcpy(out_arr, d_out, (size_t)num_wgs * wg * 4, hipMemcpyDeviceToHost));
  uint64_t tr = 0;
  rc = read_stats(d_stats, d_ctl, res, &tr);
  if (torn) *torn = tr;
  return rc;
}

extern "C" int hs_flags_run(hs_ctx *c, int num_tbs, int tb, int accs, int repeats,
                            float cutoff, int iters, const float *rnd, uint32_t *out,
                            uint32_t *dirty, hs_relaxed_result *res) {
  if (!c || !rnd || !out || !dirty || !res || num_tbs < 2 || tb < 64 || tb % 64 ||
      tb > 1024 || accs < 0 || repeats < 0 || iters < 0 || !(cutoff >= 0.f && cutoff <= 1.f))
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  const int ncu = c->num_cu;
  const int nab = std::min(num_tbs, ncu);
  const int per = (num_tbs + nab - 1) / nab;
  if (num_tbs > ncu && num_tbs % ncu) return HS_EINVAL;  // CFK:213-221
  HS_HIP(hipSetDevice(c->device));
  int fit = 0, rc;
  if ((rc = coresident_wgs(c, (const void *)k_flags, tb, &fit))) return rc;
  if (fit < num_tbs) return HS_EOCCUPANCY;
  DevArena ar;
  uint32_t *d_sync, *d_out, *d_ctl, *d_flags;
  float *d_rnd;
  uint64_t *d_stats;
  const size_t nrand = (size_t)num_tbs * iters;
  const size_t flag_words = 2 * (size_t)nab + 2 * (size_t)nab * per + HS_LINE_WORDS;
  if ((rc = ar.alloc(&d_sync, (4 + (size_t)num_tbs) * HS_LINE_WORDS))) return rc;
  if ((rc = ar.alloc(&d_out, (size_t)num_tbs * tb))) return rc;
  if ((rc = ar.alloc(&d_ctl, 64))) return rc;
  if ((rc = ar.alloc(&d_flags, flag_words))) return rc;
  if ((rc = ar.alloc(&d_rnd, nrand))) return rc;
  if ((rc = ar.alloc(&d_stats, 4))) return rc;
  HS_HIP(hipMemset(d_sync, 0, (4 + (size_t)num_tbs) * HS_LINE_WORDS * 4));
  HS_HIP(hipMemset(d_out, 0, (size_t)num_tbs * tb * 4));
  HS_HIP(hipMemset(d_ctl, 0, 256));
  HS_HIP(hipMemset(d_flags, 0, flag_words * 4));
  HS_HIP(hipMemset(d_stats, 0, 32));
  if (nrand) HS_HIP(hipMemcpy(d_rnd, rnd, nrand * 4, hipMemcpyHostToDevice));
  FlagsArgs a;
  a.stop = d_sync;
  a.dirty_total = d_sync + 2 * HS_LINE_WORDS;
  a.dirty = d_sync + 4 * HS_LINE_WORDS;
  a.out = d_out;
  a.rnd = d_rnd;
  a.ctl = d_ctl;
  a.stats = d_stats;
  uint32_t *f = d_flags + HS_LINE_WORDS;
  a.fb = FlagBar{f, f + nab, f + 2 * nab, f + 2 * nab + (size_t)nab * per,
                 d_sync + 3 * HS_LINE_WORDS, nab, per};
  a.accs = accs;
  a.repeats = repeats;
  a.iters = iters;
  a.cutoff = cutoff;
  a.ticks = timeout_ticks(c, 0);
  Timer t;
  HS_HIP(hipDeviceSynchronize());
  HS_HIP(hipEventRecord(t.e0, c->stream));
  hipLaunchKernelGGL(k_flags, dim3(num_tbs), dim3(tb), 0, c->stream, a);
  HS_HIP(hipGetLastError());
  HS_HIP(hipEventRecord(t.e1, c->stream));
  HS_HIP(hipEventSynchronize(t.e1));
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, t.e0, t.e1));
  res->kernel_ms = ms;
  HS_HIP(hipMemcpy(out, d_out, (size_t)num_tbs * tb * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(dirty, a.dirty_total, 4, hipMemcpyDeviceToHost));
  return read_stats(d_stats, d_ctl, res, nullptr);
}

extern "C" int hs_refcounter_run(hs_ctx *c, int sharers, int num_tbs, int tb, int repeats,
                                 int spacing, int strict, uint32_t *c0, uint32_t *c1,
                                 uint32_t *del0, uint32_t *del1, hs_relaxed_result *res) {
  if (!c || !c0 || !c1 || !del0 || !del1 || !res || sharers < 1 || num_tbs < sharers ||
      num_tbs % sharers || tb < 1 || tb > 1024 || repeats < 0)
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  HS_HIP(hipSetDevice(c->device));
  const long nthr = (long)num_tbs * tb, ncount = nthr / sharers;
  DevArena ar;
  uint32_t *d_c, *d_del;
  uint64_t *d_stats;
  int rc;
  if ((rc = ar.alloc(&d_c, 2 * (size_t)ncount))) return rc;
  if ((rc = ar.alloc(&d_del, 2 * (size_t)nthr))) return rc;
  if ((rc = ar.alloc(&d_stats, 4))) return rc;
  HS_HIP(hipMemset(d_c, 0, 2 * (size_t)ncount * 4));
  HS_HIP(hipMemset(d_del, 0, 2 * (size_t)nthr * 4));
  RefArgs a{d_c, d_c + ncount, d_del, d_del + nthr, d_stats, num_tbs / sharers, repeats,
            spacing < 0 ? 2 * repeats : spacing};
  if ((rc = load_kernel(strict ? (const void *)k_refcounter<true>
                               : (const void *)k_refcounter<false>)))
    return rc;  // not inside the timed launch
  Timer t;
  HS_HIP(hipDeviceSynchronize());
  HS_HIP(hipEventRecord(t.e0, c->stream));
  if (strict)
    hipLaunchKernelGGL(k_refcounter<true>, dim3(num_tbs), dim3(tb), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_refcounter<false>, dim3(num_tbs), dim3(tb), 0, c->stream, a);
  HS_HIP(hipGetLastError());
  HS_HIP(hipEventRecord(t.e1, c->stream));
  HS_HIP(hipEventSynchronize(t.e1));
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, t.e0, t.e1));
  res->kernel_ms = ms;
  HS_HIP(hipMemcpy(c0, d_c, ncount * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(c1, d_c + ncount, ncount * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(del0, d_del, nthr * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(del1, d_del + nthr, nthr * 4, hipMemcpyDeviceToHost));
  res->atomic_ops = 4ull * (uint64_t)nthr * (uint64_t)repeats;
  return HS_OK;
}

extern "C" int hs_splitcounter_run(hs_ctx *c, int num_tbs, int tb, int iters, int read_every,
                                   uint32_t *counters, uint32_t *out, hs_relaxed_result *res) {
  if (!c || !counters || !out || !res || num_tbs < 1 || tb < 128 || tb % 128 || tb > 1024 ||
      iters < 0 || read_every < 0)
    return HS_EINVAL;
  std::memset(res, 0, sizeof *res);
  HS_HIP(hipSetDevice(c->device));
  const long nthr = (long)num_tbs * tb;
  DevArena ar;
  uint32_t *d_cnt, *d_part, *d_out;
  uint64_t *d_stats;
  int rc;
  if ((rc = ar.alloc(&d_cnt, nthr / 2))) return rc;
  if ((rc = ar.alloc(&d_part, num_tbs))) return rc;
  if ((rc = ar.alloc(&d_out, nthr))) return rc;
  if ((rc = ar.alloc(&d_stats, 4))) return rc;
  HS_HIP(hipMemset(d_cnt, 0, nthr / 2 * 4));
  HS_HIP(hipMemset(d_part, 0, num_tbs * 4));
  HS_HIP(hipMemset(d_out, 0, nthr * 4));
  HS_HIP(hipMemset(d_stats, 0, 32));
  // every WG reads at least once: the period is capped at the run's length
  const int longest = iters > 0 ? iters : 1;
  if (read_every == 0) read_every = longest < 64 ? longest : 64;
  if (read_every > longest) read_every = longest;
  SplitArgs a{d_cnt, d_part, d_out, d_stats, iters, read_every};
  if ((rc = load_kernel((const void *)k_splitcounter))) return rc;  // not inside the timed launch
  Timer t;
  HS_HIP(hipDeviceSynchronize());
  HS_HIP(hipEventRecord(t.e0, c->stream));
  hipLaunchKernelGGL(k_splitcounter, dim3(num_tbs), dim3(tb), 0, c->stream, a);
  HS_HIP(hipGetLastError());
  HS_HIP(hipEventRecord(t.e1, c->stream));
  HS_HIP(hipEventSynchronize(t.e1));
  float ms = 0;
  HS_HIP(hipEventElapsedTime(&ms, t.e0, t.e1));
  res->kernel_ms = ms;
  HS_HIP(hipMemcpy(counters, d_cnt, nthr / 2 * 4, hipMemcpyDeviceToHost));
  HS_HIP(hipMemcpy(out, d_out, nthr * 4, hipMemcpyDeviceToHost));
  uint64_t st[4];
  HS_HIP(hipMemcpy(st, d_stats, sizeof st, hipMemcpyDeviceToHost));
  res->atomic_ops = st[1];
  return HS_OK;
}
