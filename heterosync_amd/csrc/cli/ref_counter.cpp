// ref_counter.cpp -- drop-in for `./refCounter <numSharersPerGroup> <numTBs>
// <tbSize> <numRepeats>` (cuda/relaxedAtomics/uvm/refCounter/refCounter.cu).
// Checks refCounter.cu:85-126 (all counters 0, some del set) and the
// stronger per-counter rule: every counter had a final decrement, so one of
// its sharers marked del.  Options: --spacing=N (default 2*numRepeats),
// --strict (the reference's fences around the decrements).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../../include/heterosync.h"

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "./refCounter <numSharersPerGroup> <numTBs> <tbSize> <numRepeats> "
                         "[--spacing=N] [--strict]\n");
    return -1;
  }
  const int sh = std::atoi(argv[1]), ntb = std::atoi(argv[2]), tb = std::atoi(argv[3]);
  const int reps = std::atoi(argv[4]);
  int spacing = -1, strict = 0;
  for (int i = 5; i < argc; ++i) {
    if (!std::strncmp(argv[i], "--spacing=", 10)) spacing = std::atoi(argv[i] + 10);
    else if (!std::strcmp(argv[i], "--strict")) strict = 1;
  }
  hs_ctx *ctx;
  if (hs_open(0, &ctx)) { std::fprintf(stderr, "no HIP device\n"); return -1; }
  const size_t nthr = (size_t)ntb * tb, nc = nthr / sh;
  const int groups = ntb / sh;
  std::printf("# Thr: %zu, # TB: %d, # Counters: %zu, # Sharers: %d, # Groups: %d, # Counters/Group: %d\n",
              nthr, ntb, nc, sh, groups, tb);
  std::vector<uint32_t> c0(nc), c1(nc), d0(nthr), d1(nthr);
  hs_relaxed_result r;
  int rc = hs_refcounter_run(ctx, sh, ntb, tb, reps, spacing, strict, c0.data(), c1.data(),
                             d0.data(), d1.data(), &r);
  if (rc) { std::fprintf(stderr, "ERROR hs_refcounter_run - %s\n", hs_strerror(rc)); return -1; }
  bool pass = true;
  for (size_t i = 0; i < nc; ++i)
    if (c0[i] || c1[i]) { pass = false; std::fprintf(stderr, "ERROR: counter %zu: %u %u\n", i, c0[i], c1[i]); break; }
  for (size_t ci = 0; ci < nc && pass; ++ci) {
    const int g = (int)(ci / tb), t = (int)(ci % tb);
    bool s0 = false, s1 = false;
    for (int b = g; b < ntb; b += groups) {
      s0 |= d0[(size_t)b * tb + t] != 0;
      s1 |= d1[(size_t)b * tb + t] != 0;
    }
    if (reps > 0 && (!s0 || !s1)) { pass = false; std::fprintf(stderr, "ERROR: counter %zu never released\n", ci); }
  }
  std::printf("kernel time: %f ms, atomic ops: %llu\n", r.kernel_ms, (unsigned long long)r.atomic_ops);
  std::fprintf(stderr, pass ? "TEST PASSED!\n" : "TEST FAILED!\n");
  hs_close(ctx);
  return pass ? 0 : -1;
}
