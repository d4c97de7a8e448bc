// seqlocks.cpp -- drop-in for `./seqlocks <numWGs> <wgSize> <groupSize_seqlock>
// <useTFs>` (hip/relaxedAtomics/uvm/seqlocks/seqlocks.cpp:6-112).  useTFs=1
// selects the relaxed+fence variant, 0 the scoped acquire/release one
// (seqlocks_hipAtomics).  Options after the positional arguments:
// --iters=N (reference NUM_ITERS 10), --writer-period=N (reference 16),
// --strict (the reference's fences and acquire/release orderings instead of
// drained agent-scope accesses).
// Checks: every sequence number even (seqlocks.cpp:80-88), each equals
// 2*iters*#writers of that lock, and no torn reader snapshot.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../../include/heterosync.h"

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "./seqlocks <numWGs> <wgSize> <groupSize_seqlock> <useTFs> "
                         "[--iters=N] [--writer-period=N] [--strict]\n");
    return -1;
  }
  const int nw = std::atoi(argv[1]), wg = std::atoi(argv[2]), gs = std::atoi(argv[3]);
  const int tfs = std::atoi(argv[4]) == 1;
  int iters = 10, period = 16, strict = 0;
  for (int i = 5; i < argc; ++i) {
    if (!std::strncmp(argv[i], "--iters=", 8)) iters = std::atoi(argv[i] + 8);
    else if (!std::strncmp(argv[i], "--writer-period=", 16)) period = std::atoi(argv[i] + 16);
    else if (!std::strcmp(argv[i], "--strict")) strict = 1;
  }
  hs_ctx *ctx;
  if (hs_open(0, &ctx)) { std::fprintf(stderr, "no HIP device\n"); return -1; }
  const int nseq = gs > nw / (gs > 0 ? gs : 1) ? gs : nw / gs;
  std::vector<uint32_t> seq(nseq);
  std::vector<int32_t> out((size_t)nw * wg);
  uint64_t torn = 0;
  hs_relaxed_result r;
  std::printf("Launching kernel - 1 runs with %d WGs and %d threads/WG\n", nw, wg);
  int rc = hs_seqlock_run(ctx, nw, wg, gs, period, iters, (tfs ? 1 : 0) | (strict ? 2 : 0),
                          seq.data(), nseq, out.data(), &torn, &r);
  if (rc) { std::fprintf(stderr, "ERROR hs_seqlock_run - %s\n", hs_strerror(rc)); return -1; }
  int errors = 0;
  for (int l = 0; l < gs; ++l) {
    uint32_t want = 0;
    for (int b = 0; b < nw; ++b) if (b % period == 0 && b % gs == l) want += 2u * iters;
    if (seq[l] % 2 != 0) { std::fprintf(stderr, "ERROR: seqlock should be a multiple of 2!\n"); ++errors; }
    if (seq[l] != want) { std::fprintf(stderr, "ERROR: seqlock[%d] = %u, expected %u\n", l, seq[l], want); ++errors; }
  }
  if (torn) { std::fprintf(stderr, "ERROR: %llu torn reader snapshots\n", (unsigned long long)torn); ++errors; }
  std::printf("kernel time: %f ms, atomic ops: %llu\n", r.kernel_ms, (unsigned long long)r.atomic_ops);
  std::printf(errors ? "FAILED\n" : "PASSED\n");
  hs_close(ctx);
  return errors ? -1 : 0;
}
