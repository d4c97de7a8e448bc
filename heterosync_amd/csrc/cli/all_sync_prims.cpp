// all_sync_prims.cpp -- drop-in replacement for the reference's
// `allSyncPrims-1kernel <syncPrim> <numLdSt> <numWGs> <numCSIters>`
// (hip/syncPrims/uvm/main.hip.cpp:1929-2715).  Same positional arguments,
// same exit codes (0 / -1) and the stdout lines consumers parse:
//   "elapsed time: %f ms, average: %f ms"   (main.hip.cpp:2510, parsed by
//                                            syncPrims_results.sh:35)
//   "PASSED!"                               (main.hip.cpp:2702)
//   stderr "ERROR: %s has %u output errors" (main.hip.cpp:2698)
// Optional knobs go AFTER the four positional arguments so existing sweep
// scripts keep working: --wg=N --mode=copy|inc --reps=N --backoff=MIN:MAX
// --poll=N --timeout=S --json.
//
// The self-check is the library's hs_syncprim_expected: the closed form of
// the reference's host golden (main.hip.cpp:2524-2682).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/heterosync.h"

namespace {

void usage() {
  std::fprintf(stderr,
               "./allSyncPrims-1kernel <syncPrim> <numLdSt> <numWGs> <numCSIters> [options]\n"
               "where:\n"
               "\t<syncPrim>: atomicTreeBarrUniq, atomicTreeBarrUniqLocalExch, lfTreeBarrUniq,\n"
               "\t\tlfTreeBarrUniqLocalExch, spinMutex, spinMutexEBO, sleepMutex, faMutex,\n"
               "\t\tspinSem{1,2,10,120}, spinSemEBO{1,2,10,120}, spinMutexUniq,\n"
               "\t\tspinMutexEBOUniq, sleepMutexUniq, faMutexUniq, spinSemUniq{1,2,10,120},\n"
               "\t\tspinSemEBOUniq{1,2,10,120}\n"
               "\t<numLdSt>: the # of LDs and STs to do for each thread in the critical section.\n"
               "\t<numWGs>: the # of WGs to execute (want to be divisible by the number of CUs).\n"
               "\t<numCSIters>: number of iterations of the critical section.\n"
               "options: --wg=N (threads/WG, default 64) --mode=copy|inc --reps=N (default 10)\n"
               "         --backoff=MIN:MAX --poll=N --timeout=SECONDS --json\n"
               "         --strict (agent-scope release/acquire in every WG)\n");
}

// The line the reference prints before a run (main.hip.cpp:2177-2340),
// including its semaphore counts: ids 11, 19 and 27 print 2 for their
// 120-count semaphores there, and so do they here.
std::string banner(int p, int iters) {
  static const char *kBarrier[] = {"atomic_tree_barrier", "atomic_tree_barrier_localExch",
                                   "fbs_tree_barrier", "fbs_tree_barrier_localExch"};
  static const char *kMutex[] = {"spin_lock_mutex", "ebo_mutex", "sleeping_mutex",
                                 "fetchadd_mutex"};
  static const char *kSrb[] = {"atomic_tree_barrier_SRB", "atomic_tree_barrier_localExch_SRB",
                               "atomic_tree_barrier_NaiveSRB",
                               "atomic_tree_barrier_localExch_NaiveSRB",
                               "atomic_tree_barrier_NaiveAllSRB",
                               "atomic_tree_barrier_localExch_NaiveAllSRB"};
  static const int kShown[4] = {1, 2, 10, 120};
  char buf[96];
  if (p < 4) {
    std::snprintf(buf, sizeof buf, "%s_%03d", kBarrier[p], iters);
  } else if (p <= 7 || (p >= 20 && p <= 23)) {
    std::snprintf(buf, sizeof buf, "%s%s_%03d", kMutex[(p - 4) % 16], p >= 20 ? "_uniq" : "",
                  iters);
  } else if (p >= 37) {
    std::snprintf(buf, sizeof buf, "%s_%03d", kSrb[p - 37], iters);
  } else {
    const char *kind = p <= 11 ? "spin_lock_sem" : p <= 15 ? "ebo_sem"
                     : p <= 19 ? "prior_spin_lock_sem" : p <= 27 ? "spin_lock_sem_uniq"
                     : p <= 31 ? "ebo_sem_uniq" : "prior_ebo_sem";
    const int shown = (p == 11 || p == 19 || p == 27) ? 2 : kShown[(p - 8) % 4];
    std::snprintf(buf, sizeof buf, "%s_%03d_%03d", kind, shown, iters);
  }
  return buf;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 5) {
    usage();
    return -1;
  }
  int wg = 64, reps = 10, mode = 0, bmin = 0, bmax = 0, poll = 0, strict = 0;
  double timeout = 0;
  bool json = false;
  for (int i = 5; i < argc; ++i) {
    const char *a = argv[i];
    if (!std::strncmp(a, "--wg=", 5)) wg = std::atoi(a + 5);
    else if (!std::strncmp(a, "--reps=", 7)) reps = std::atoi(a + 7);
    else if (!std::strcmp(a, "--mode=inc")) mode = 1;
    else if (!std::strcmp(a, "--mode=copy")) mode = 0;
    else if (!std::strncmp(a, "--backoff=", 10)) std::sscanf(a + 10, "%d:%d", &bmin, &bmax);
    else if (!std::strncmp(a, "--poll=", 7)) poll = std::atoi(a + 7);
    else if (!std::strncmp(a, "--timeout=", 10)) timeout = std::atof(a + 10);
    else if (!std::strcmp(a, "--json")) json = true;
    else if (!std::strcmp(a, "--strict")) strict = 1;
    else {
      usage();
      return -1;
    }
  }

  hs_ctx *ctx = nullptr;
  int rc = hs_open(0, &ctx);
  if (rc) {
    std::fprintf(stderr, "There is no device supporting HIP (%s)\n", hs_strerror(rc));
    return -1;
  }
  int ncu = 0, max_thr = 0, max_wgs = 0;
  char arch[64];
  hs_device_info(ctx, &ncu, &max_thr, &max_wgs, arch, sizeof arch);
  std::printf("GPU: %s\n", arch);
  std::printf("# CU: %d, Max Thrs/WG: %d, Max WG/CU: %d, Max # WG: %d\n", ncu, max_thr,
              max_thr / 64, max_wgs);

  const char *prim_str = argv[1];
  const int L = std::atoi(argv[2]);
  const int nw = std::atoi(argv[3]);
  const int iters = std::atoi(argv[4]);
  if (nw > max_wgs || nw < 1) {  // main.hip.cpp:2028
    std::fprintf(stderr, "ERROR: numWGs (%d) must be in [1, %d]\n", nw, max_wgs);
    hs_close(ctx);
    return -1;
  }
  const int p = hs_prim_id(prim_str);
  if (p < 0) {
    std::fprintf(stderr, "ERROR: Unknown synchronization primitive: %s\n", prim_str);
    hs_close(ctx);
    return -1;
  }
  if (!hs_prim_supported(p)) {
    std::fprintf(stderr, "ERROR: %s (#%d) is not implemented by this library\n", prim_str, p);
    hs_close(ctx);
    return -1;
  }
  const long n = hs_storage_words(p, nw, ncu, wg, L);
  if (n <= 0) {
    std::fprintf(stderr, "ERROR: bad sizes\n");
    hs_close(ctx);
    return -1;
  }
  const long uniq_per_wg = (long)wg * (L + 1);
  std::printf("# WGs: %d, # Ld/St: %d, # Locs Mult: %ld, # Uniq Locs/WG: %ld, # Storage Locs: %ld\n",
              nw, L, n / uniq_per_wg, uniq_per_wg, n);
  std::vector<int32_t> storage(n);
  for (long i = 0; i < n; ++i) storage[i] = (int32_t)i;

  std::printf("%s\n", banner(p, iters).c_str());
  std::fflush(stdout);

  hs_syncprim_cfg cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.prim = p;
  cfg.num_ldst = L;
  cfg.num_wgs = nw;
  cfg.num_cs_iters = iters;
  cfg.wg_size = wg;
  cfg.reps = reps;
  cfg.mode = mode;
  cfg.backoff_min = bmin;
  cfg.backoff_max = bmax;
  cfg.poll_sleep = poll;
  cfg.timeout_s = timeout;
  cfg.sync_scope = strict;
  hs_syncprim_result res;
  rc = hs_syncprim_run(ctx, &cfg, storage.data(), n, &res);
  if (rc) {
    std::fprintf(stderr, "ERROR hs_syncprim_run - %s\n", hs_strerror(rc));
    hs_close(ctx);
    return -1;
  }
  std::printf("elapsed time: %f ms, average: %f ms\n", res.roi_ms, res.roi_ms / reps);
  std::fflush(stdout);

  std::vector<int32_t> golden(n);
  if (hs_syncprim_expected(p, L, nw, iters, reps, wg, ncu, mode, golden.data(), n) != HS_OK) {
    std::fprintf(stderr, "ERROR: no self-check for %s\n", prim_str);
    hs_close(ctx);
    return -1;
  }
  unsigned errors = 0;
  std::printf("Comparing GPU results to golden results:\n");
  for (long i = 0; i < n; ++i) {
    if (storage[i] != golden[i]) {
      if (errors < 16)
        std::fprintf(stderr, "\tERROR: storage[%ld] = %d, golden[%ld] = %d\n", i,
                     storage[i], i, golden[i]);
      ++errors;
    }
  }
  if (!res.state_ok) {
    std::fprintf(stderr, "\tERROR: lock/barrier state not idle after the run\n");
    ++errors;
  }
  if (errors) {
    std::fprintf(stderr, "ERROR: %s has %u output errors\n", prim_str, errors);
    hs_close(ctx);
    return -1;
  }
  std::printf("PASSED!\n");
  if (json) {
    const double secs = res.roi_ms / 1e3;
    const double acq = (double)(res.cs_entries + res.rd_entries);
    std::printf("{\"prim\": \"%s\", \"roi_ms\": %.4f, \"kernel_ms\": %.4f, "
                "\"cs_entries\": %lld, \"rd_entries\": %lld, \"episodes\": %lld, "
                "\"acquire_release_per_s\": %.1f, \"episodes_per_s\": %.1f, \"sync_scope\": %d}\n",
                prim_str, res.roi_ms, res.kernel_ms, (long long)res.cs_entries,
                (long long)res.rd_entries, (long long)res.episodes,
                secs > 0 ? acq / secs : 0.0, secs > 0 ? res.episodes / secs : 0.0,
                res.sync_scope_used);
  }
  hs_close(ctx);
  return 0;
}
