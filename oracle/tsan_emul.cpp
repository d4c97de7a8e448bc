// tsan_emul.cpp -- the std::thread + std::atomic emulation (cpu_emul.cpp)
// under ThreadSanitizer.  Test infrastructure: every primitive the
// emulation covers runs at small sizes, in both critical-section modes, and
// its storage must equal the closed form (hso_golden_closed).  Built by
// `make -C oracle tsan` with -fsanitize=thread; a data race in the
// emulation's lock, semaphore or barrier protocols makes TSan report and
// exit nonzero.
#include <cstdio>
#include <cstring>
#include <vector>

#include "hs_oracle.h"

int main() {
  int runs = 0, bad = 0;
  for (int prim = 0; prim <= 42; ++prim) {
    if (prim == 36) continue;  // reserved in the reference's table
    for (int mode = 0; mode <= 1; ++mode) {
      const int nw = 8, ncu = 4, L = 3, iters = 3, reps = 2, wg = 64;
      const long n = hso_num_storage_locs(prim, nw, ncu, wg, L);
      if (n <= 0) continue;
      std::vector<int32_t> want(n), got(n, 0);
      if (hso_golden_closed(prim, L, nw, iters, reps, wg, ncu, mode, want.data(), n) != 0)
        continue;
      int64_t cs = 0;
      const double t = hso_cpu_run(prim, L, nw, iters, reps, wg, ncu, mode, got.data(), n, &cs);
      ++runs;
      if (t < 0 || std::memcmp(got.data(), want.data(), n * 4) != 0) {
        std::fprintf(stderr, "prim %d mode %d: storage differs from the closed form\n", prim,
                     mode);
        ++bad;
      }
      if (prim <= 3) {  // the threaded barrier pool (the CPU baseline)
        const long n16 = hso_num_storage_locs(prim, 16, ncu, wg, L);
        std::vector<int32_t> pool(n16, 0), w16(n16);
        const double tb = hso_cpu_barrier_run(prim, L, 16, iters, reps, wg, ncu, mode, 3,
                                              pool.data(), n16);
        hso_golden_closed(prim, L, 16, iters, reps, wg, ncu, mode, w16.data(), n16);
        ++runs;
        if (tb < 0 || std::memcmp(pool.data(), w16.data(), n16 * 4) != 0) {
          std::fprintf(stderr, "barrier pool prim %d mode %d differs\n", prim, mode);
          ++bad;
        }
      }
    }
  }
  std::printf("tsan_emul: %d runs, %d mismatches\n", runs, bad);
  return bad ? 1 : 0;
}
