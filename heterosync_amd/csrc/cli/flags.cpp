// flags.cpp -- drop-in for `./flags <numTBs> <tbSize> <numAccsPerThr>
// <numRepeats> <cutoffVal> <numIters> <pageAlign>`
// (cuda/relaxedAtomics/uvm/flags/flags.cu:19-176).  Random input as the
// reference: srand(2018), rand()/UINT32_MAX (flags.cu:92-101), one value per
// worker and iteration (rand[b + it*numTBs]).  Checks the output rule of
// flags.cu:130-160 plus the exact dirty count.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../../include/heterosync.h"

int main(int argc, char **argv) {
  if (argc != 8) {
    std::fprintf(stderr, "./flags <numTBs> <tbSize> <numAccsPerThr> <numRepeats> <cutoffVal> "
                         "<numIters> <pageAlign>\n");
    return -1;
  }
  const int ntb = std::atoi(argv[1]), tb = std::atoi(argv[2]), accs = std::atoi(argv[3]);
  const int reps = std::atoi(argv[4]), iters = std::atoi(argv[6]);
  const float cut = (float)std::atof(argv[5]);
  hs_ctx *ctx;
  if (hs_open(0, &ctx)) { std::fprintf(stderr, "no HIP device\n"); return -1; }
  std::vector<float> rnd((size_t)ntb * iters);
  srand(2018);
  for (auto &v : rnd) v = (float)rand() / (float)(unsigned int)-1;
  uint64_t want_dirty = 0;
  for (int b = 0; b < ntb - 1; ++b)
    for (int it = 0; it < iters; ++it) want_dirty += rnd[(size_t)b + (size_t)it * ntb] <= cut;
  std::vector<uint32_t> out((size_t)ntb * tb);
  uint32_t dirty = 0;
  hs_relaxed_result r;
  std::printf("Launching kernel - 1 runs with %d TBs and %d threads/TB, %d iterations of worker TB loop\n",
              ntb, tb, iters);
  int rc = hs_flags_run(ctx, ntb, tb, accs, reps, cut, iters, rnd.data(), out.data(), &dirty, &r);
  if (rc) { std::fprintf(stderr, "ERROR hs_flags_run - %s\n", hs_strerror(rc)); return -1; }
  std::printf("isDirty - %u\n", dirty);
  bool pass = dirty == want_dirty;
  if (!pass) std::fprintf(stderr, "ERROR: dirty = %u, expected %llu\n", dirty, (unsigned long long)want_dirty);
  const uint32_t exp_main = (uint32_t)reps * accs * accs;
  for (size_t i = 0; i < out.size(); ++i) {
    const bool main_part = i >= (size_t)tb * (ntb - 1);
    const uint32_t want = (dirty && main_part) ? exp_main : 0;
    if (out[i] != want) {
      if (pass) std::fprintf(stderr, "ERROR: outArr[%zu] = %u, expected %u\n", i, out[i], want);
      pass = false;
    }
  }
  std::printf("kernel time: %f ms\n", r.kernel_ms);
  std::printf("Test %s!\n", pass ? "PASSED" : "FAILED");
  hs_close(ctx);
  return pass ? 0 : -1;
}
