// split_counter.cpp -- drop-in for `./splitCounter <numTBs> <tbSize>`
// (cuda/relaxedAtomics/uvm/splitCounter/splitCounter.cu).  Checks every
// counter == iters*INC_VAL (splitCounter.cu:52-60) and that every reader's
// approximate total lies in [0, total].  Options: --iters=N (reference 1),
// --read-every=N (writer iterations between a WG's reads; 0 = min(64, N)).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../../include/heterosync.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "./splitCounter <numTBs> <tbSize> [--iters=N] [--read-every=N]\n");
    return -1;
  }
  const int ntb = std::atoi(argv[1]), tb = std::atoi(argv[2]);
  int iters = 1, read_every = 0;
  for (int i = 3; i < argc; ++i) {
    if (!std::strncmp(argv[i], "--iters=", 8)) iters = std::atoi(argv[i] + 8);
    else if (!std::strncmp(argv[i], "--read-every=", 13)) read_every = std::atoi(argv[i] + 13);
  }
  hs_ctx *ctx;
  if (hs_open(0, &ctx)) { std::fprintf(stderr, "no HIP device\n"); return -1; }
  const size_t nthr = (size_t)ntb * tb;
  std::vector<uint32_t> cnt(nthr / 2), out(nthr);
  hs_relaxed_result r;
  std::printf("Launching kernel - 1 runs with %d TBs and %d threads/TB\n", ntb, tb);
  int rc = hs_splitcounter_run(ctx, ntb, tb, iters, read_every, cnt.data(), out.data(), &r);
  if (rc) { std::fprintf(stderr, "ERROR hs_splitcounter_run - %s\n", hs_strerror(rc)); return -1; }
  bool pass = true;
  for (size_t i = 0; i < cnt.size(); ++i)
    if (cnt[i] != (uint32_t)iters) { std::fprintf(stderr, "ERROR: counters[%zu] = %u\n", i, cnt[i]); pass = false; break; }
  const uint64_t total = (uint64_t)iters * cnt.size();
  for (size_t i = 0; i < out.size(); ++i)
    if (out[i] > total) { std::fprintf(stderr, "ERROR: reader %zu saw %u > %llu\n", i, out[i], (unsigned long long)total); pass = false; break; }
  std::printf("kernel time: %f ms, atomic ops: %llu\n", r.kernel_ms, (unsigned long long)r.atomic_ops);
  std::printf(pass ? "PASSED\n" : "FAILED\n");
  hs_close(ctx);
  return pass ? 0 : -1;
}
