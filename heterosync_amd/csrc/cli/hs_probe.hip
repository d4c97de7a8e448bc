// hs_probe.hip -- MI355X synchronization cost probe: the measured
// denominators of the sync-path rooflines.
//   1. workgroup -> XCD placement (XCC_ID hardware register)
//   2. agent-scope release / acquire cost vs. number of concurrent WGs
//   3. one-way flag hand-off latency (same XCD / across XCDs)
//   4. device-scope atomic RMW throughput: one address, one per XCD, one per
//      WG, one per lane, eight per lane in turn (the uncontended peak)
//   5. visibility: which store/load flavours observe each other across WGs
//      on the same XCD and across XCDs
//   6. a line's return to an XCD whose L2 an earlier store left it in
// Every wait is bounded; prints one line per measurement.
// `hs_probe floor <WGs> <lanes> <numLdSt>` prints only the terms of the
// synchronization floors (one JSON line): the grid barrier's data pass at
// that footprint, dependent RMW / load / store round trips, the flag
// hand-off on one XCD and across XCDs, and one address's RMW spacing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "../../../include/hs_detail/hs_syncprims_dev.h"

using namespace hs;

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                           \
    }                                                                         \
  } while (0)

namespace {

__global__ void k_xcc(uint32_t *out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

// mode 0: release fence; 1: acquire fence; 2: both; 3: both after dirtying 11 KB
__global__ void k_fence(int mode, int iters, int32_t *data, uint32_t *sink) {
  int32_t *mine = data + (long)blockIdx.x * 2816;
  for (int i = 0; i < iters; ++i) {
    if (mode == 3) {
      for (int k = threadIdx.x; k < 2816; k += blockDim.x) mine[k] += 1;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (mode != 1) fence_rel();
      if (mode != 0) fence_acq();
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sink[blockIdx.x] = mine[0];
}

// One-way hand-off between WGs a and b: ping-pong on one flag, polled with
// agent-scope loads.  plain = 0: agent-scope stores (written through);
// plain = 1: plain stores, which stay dirty in the storing XCD's L2 (only
// for two WGs on one XCD: the form the barriers use for flags that never
// leave an XCD).
__global__ void k_pingpong(int a, int b, int rounds, uint32_t *flag, uint32_t *res,
                           int plain) {
  if (threadIdx.x != 0) return;
  const int me = blockIdx.x;
  if (me != a && me != b) return;
  const uint32_t parity = me == a ? 0 : 1;
  uint32_t polls = 0;
  for (int r = 0; r < rounds; ++r) {
    const uint32_t want = 2 * r + parity;
    while (ld_rlx(flag) != want)
      if (++polls > (1u << 23)) {
        res[2] = 1;
        return;
      }
    if (plain) st_xcd(flag, want + 1);
    else st_rlx(flag, want + 1);
  }
}

// 0: one address; 1: one per XCD; 2: one per WG; 3: one per lane; 4: eight
// per lane in turn (consecutive adds of a lane hit different addresses, so
// no address sees back-to-back RMWs from one lane)
__global__ void k_atomic(int mode, int iters, uint32_t *ctr) {
  if (mode < 3 && threadIdx.x != 0) return;
  uint32_t *p = ctr;
  if (mode == 1) p = ctr + (blockIdx.x % 8) * 64;
  if (mode == 2) p = ctr + blockIdx.x * 64;
  if (mode >= 3) p = ctr + blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 4) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (int i = 0; i < iters; ++i) add_rlx(p + (i & 7) * stride, 1);
    return;
  }
  for (int i = 0; i < iters; ++i) add_rlx(p, 1);
}

// writer WG a: stores (wmode 0 plain, 1 plain + release fence, 2 agent-scope
// relaxed atomic stores), waits for completion, raises a flag; reader WG b:
// loads (rmode 0 plain, 1 agent-scope relaxed atomic, 2 acquire fence +
// plain) and counts values that are not the current round.
__global__ void k_vis(int a, int b, int rounds, int wmode, int rmode, uint32_t *flag,
                      int32_t *data, uint32_t *res) {
  const int me = blockIdx.x;
  if (me != a && me != b) return;
  for (int r = 1; r <= rounds; ++r) {
    if (me == a) {
      if (threadIdx.x == 0) {
        uint32_t n = 0;
        while (ld_rlx(flag + 32) != (uint32_t)(r - 1))
          if (++n > (1u << 22)) {
            res[3] = 1;
            break;
          }
      }
      __syncthreads();
      if (wmode == 2)
        __hip_atomic_store(data + threadIdx.x * 16, r, __ATOMIC_RELAXED, HS_AGENT);
      else
        data[threadIdx.x * 16] = r;
      wait_mem();
      __syncthreads();
      if (threadIdx.x == 0) {
        if (wmode == 1) fence_rel();
        st_rlx(flag, r);
      }
    } else {
      if (threadIdx.x == 0) {
        uint32_t n = 0;
        while (ld_rlx(flag) != (uint32_t)r)
          if (++n > (1u << 22)) {
            res[3] = 1;
            break;
          }
      }
      __syncthreads();
      if (rmode == 2) fence_acq();
      const int32_t v = rmode == 1 ? __hip_atomic_load(data + threadIdx.x * 16, __ATOMIC_RELAXED, HS_AGENT)
                                   : data[threadIdx.x * 16];
      if (v != r) atomicAdd(res + 1, 1u);
      __syncthreads();
      if (threadIdx.x == 0) st_rlx(flag + 32, r);
    }
  }
}

// A line's return to an XCD.  Per round, four steps handed on by a flag:
//   0  WG a  (XCD x) stores with flavour f0 (0 plain, 2 agent-scope)
//   1  WG a2 (XCD x) agent-loads it, then stores with flavour f1
//   2  WG b  (XCD y) agent-loads it, then agent-stores
//   3  WG a  (XCD x) agent-loads it
// res[1..3] count the stale values read at steps 1..3, res[4] a timeout.
// Step 3 reads stale exactly when XCD x's L2 still returns a line that an
// earlier store on x left there after y's store reached memory.
__global__ void k_return(int a, int a2, int b, int rounds, int f0, int f1, uint32_t *flag,
                         int32_t *data, uint32_t *res) {
  const int me = blockIdx.x;
  if (me != a && me != a2 && me != b) return;
  int32_t *p = data + threadIdx.x * 16;
  bool dead = false;  // master: a wait timed out, stop waiting
  for (int r = 1; r <= rounds; ++r) {
    for (int step = 0; step < 4; ++step) {
      const int who = step == 1 ? a2 : step == 2 ? b : a;
      if (who != me) continue;
      const uint32_t s = 4u * (uint32_t)(r - 1) + (uint32_t)step;
      if (threadIdx.x == 0 && !dead) {
        uint32_t n = 0;
        while (ld_rlx(flag) != s)
          if (++n > (1u << 22)) {
            res[4] = 1;
            dead = true;
            break;
          }
      }
      __syncthreads();
      if (step > 0) {
        const int32_t v = __hip_atomic_load(p, __ATOMIC_RELAXED, HS_AGENT);
        if (v != 10 * r + step) atomicAdd(res + step, 1u);
      }
      const int fl = step == 0 ? f0 : step == 1 ? f1 : 2;
      if (step < 3) {
        if (fl == 2)
          __hip_atomic_store(p, 10 * r + step + 1, __ATOMIC_RELAXED, HS_AGENT);
        else
          *p = 10 * r + step + 1;
      }
      wait_mem();
      __syncthreads();
      if (threadIdx.x == 0) st_rlx(flag, s + 1);
    }
  }
}

// The data pass of a grid-barrier episode with nothing else: every WG moves
// its own region (blockDim chains of L+1 16-word lines) `passes` times with
// the barrier kernels' own walk (16-byte accesses, agent-scope loads, stores
// kept in this XCD's L2, lines 1..L incremented in place), each pass drained
// and closed by a WG barrier.  No WG waits for another, so the time per pass
// is what the memory system needs to move one episode's data: the data term
// of the barrier's episode floor.
__global__ void k_l2pass(int32_t *data, int L, int passes) {
  for (int i = 0; i < passes; ++i) {
    wg_move<1, true>(data, (long)blockIdx.x * blockDim.x, L, true);
    wait_mem();
    __syncthreads();
  }
}

// Dependent chains of one lane of WG `who`, `n` steps each; out[0] keeps the
// values live.
//   0: agent-scope RMW (fetch_add) whose address depends on the last result
//   1: agent-scope load whose address depends on the last loaded value
//   2: agent-scope store followed by a drain of the memory counter
__global__ void k_chain(int who, int mode, int n, uint32_t *p, uint32_t *out) {
  if (threadIdx.x != 0 || (int)blockIdx.x != who) return;
  uint32_t v = 0;
  for (int i = 0; i < n; ++i) {
    if (mode == 0) {
      v = add_rlx(p + (v >> 30), 1u);  // v < 2^30: always p[0], but not known to the compiler
    } else if (mode == 1) {
      v = ld_rlx(p + 64 * ((v + i) & 7));  // p[64 k] holds 0
    } else {
      st_rlx(p + 64 * (i & 7) + 32, (uint32_t)i);
      wait_mem();
    }
  }
  out[0] = v;
}

// Agent-scope stores of the chase's links (chase[line * 32] = next[line])
// by a grid spread over every XCD: each line is then one another XCD has
// just written through, 7 times in 8, as a lock's next holder finds the
// critical section its predecessor wrote.
__global__ void k_chase_write(const uint32_t *next, uint32_t nl, uint32_t *chase) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x)
    st_rlx(chase + (size_t)i * 32, next[i]);
}

// Dependent agent-scope loads of one lane of WG `who` chasing a cyclic
// permutation of 128-byte lines (chase[line * 32] holds the next line): n
// steps, each a line no other step of the launch touches.
__global__ void k_chase(int who, int n, uint32_t start, const uint32_t *chase, uint32_t *out) {
  if (threadIdx.x != 0 || (int)blockIdx.x != who) return;
  uint32_t v = start;
  for (int i = 0; i < n; ++i) v = ld_rlx(chase + (size_t)v * 32);
  out[0] = v;
}

// Plain streaming copy of known size: the calibration reference for the
// FETCH_SIZE / WRITE_SIZE hardware counters.
__global__ void k_copy(const int4 *__restrict__ src, int4 *__restrict__ dst, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

float time_ms(const std::function<void()> &f) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  f();
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0));
  f();
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return ms;
}

}  // namespace

// hs_probe floor <WGs> <lanes> <numLdSt>: the measured terms of the
// synchronization floors, one JSON line.
int floor_terms(int nwg, int lanes, int L) {
  if (nwg < 1 || lanes < 64 || lanes % 64 || lanes > 1024 || L < 1) return 1;
  int32_t *data;
  uint32_t *buf, *res;
  const long words = (long)nwg * lanes * (L + 1);
  CHECK(hipMalloc(&data, words * 4));
  CHECK(hipMemset(data, 0, words * 4));
  CHECK(hipMalloc(&buf, 1 << 16));
  CHECK(hipMalloc(&res, 4096));
  const int passes = 200;
  const float pass_ms = time_ms([&] { k_l2pass<<<nwg, lanes>>>(data, L, passes); });
  // Latencies depend on where the two ends sit: XCDs on the same I/O die
  // hand a line over faster than XCDs on different dies, and a memory-side
  // RMW's round trip depends on the issuing XCD and the counter's channel.
  // WG b of a small launch runs on XCD (b + s) mod 8, s set by the launch,
  // so each term is measured from all 8 XCDs (chains: WG 0..7 in turn) or
  // to all 7 other XCDs (hops: WG 0 with WG 1..7), and reported as the
  // mean; the release of a grid barrier must reach every XCD, so the
  // slowest of the 7 hops is reported as well (hop_far).
  double chain_us[3] = {0, 0, 0};
  for (int mode = 0; mode < 3; ++mode) {
    const int n = 2000;
    for (int who = 0; who < 8; ++who) {
      CHECK(hipMemset(buf, 0, 1 << 16));
      const float ms = time_ms([&] { k_chain<<<8, 64>>>(who, mode, n, buf, res); });
      chain_us[mode] += ms * 1000 / n / 8;
    }
  }
  // loads of lines another XCD has just written through: a chase over
  // 64 MiB (512 Ki lines of 128 B, more than the L2s hold) in a random
  // cyclic order, its links rewritten by agent-scope stores from every XCD
  // right before each timed launch
  double remote_us = 0;
  {
    const uint32_t nl = 1u << 19;
    std::vector<uint32_t> order(nl), next(nl);
    for (uint32_t i = 0; i < nl; ++i) order[i] = i;
    uint64_t x = 0x9e3779b97f4a7c15ull;
    for (uint32_t i = nl - 1; i > 0; --i) {  // Fisher-Yates with a xorshift generator
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      std::swap(order[i], order[x % (i + 1)]);
    }
    for (uint32_t i = 0; i < nl; ++i) next[order[i]] = order[(i + 1) % nl];
    uint32_t *chase, *d_next;
    CHECK(hipMalloc(&chase, (size_t)nl * 32 * 4));
    CHECK(hipMalloc(&d_next, (size_t)nl * 4));
    CHECK(hipMemcpy(d_next, next.data(), (size_t)nl * 4, hipMemcpyHostToDevice));
    k_chase_write<<<1024, 256>>>(d_next, nl, chase);
    // each launch starts at its own point of the cycle, 4096 steps apart,
    // so no launch revisits a line an earlier one loaded
    const int n = 2000;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    k_chase<<<8, 64>>>(0, n, order[(size_t)16 * 4096], chase, res);  // warm-up launch
    CHECK(hipDeviceSynchronize());
    for (int who = 0; who < 8; ++who) {
      k_chase_write<<<1024, 256>>>(d_next, nl, chase);
      CHECK(hipDeviceSynchronize());
      CHECK(hipEventRecord(e0));
      k_chase<<<8, 64>>>(who, n, order[(size_t)who * 4096], chase, res);
      CHECK(hipEventRecord(e1));
      CHECK(hipEventSynchronize(e1));
      float ms = 0;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      remote_us += ms * 1000 / n / 8;
    }
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    CHECK(hipFree(chase));
    CHECK(hipFree(d_next));
  }
  auto hop = [&](int a, int b, double *us, int plain) {
    const int rounds = 2000;
    CHECK(hipMemset(res, 0, 64));
    const float ms = time_ms([&] {
      CHECK(hipMemset(buf, 0, 256));
      k_pingpong<<<16, 64>>>(a, b, rounds, buf, res, plain);
    });
    uint32_t hr[4];
    CHECK(hipMemcpy(hr, res, 16, hipMemcpyDeviceToHost));
    if (hr[2]) return false;
    *us = ms * 1000 / (2 * rounds);
    return true;
  };
  double hop_same = 0, hop_same_l2 = 0, hop_other = 0, hop_far = 0, us = 0;
  for (int j = 0; j < 8; ++j) {  // WG j and WG j + 8 share an XCD
    if (!hop(j, j + 8, &us, 0)) return 2;
    hop_same += us / 8;
    if (!hop(j, j + 8, &us, 1)) return 2;
    hop_same_l2 += us / 8;
  }
  for (int j = 1; j < 8; ++j) {  // WG 0 and WG j do not
    if (!hop(0, j, &us, 0)) return 2;
    hop_other += us / 7;
    hop_far = std::max(hop_far, us);
  }
  CHECK(hipMemset(buf, 0, 1 << 16));
  const int it = 1000;
  const float one_ms = time_ms([&] { k_atomic<<<1024, 256>>>(0, it, buf); });
  const double bytes = (double)nwg * lanes * L * 8;  // 4 B load + 4 B store per line
  std::printf("{\"wgs\": %d, \"lanes\": %d, \"num_ldst\": %d, \"l2_pass_us\": %.4f, "
              "\"l2_pass_gbs\": %.1f, \"rmw_rt_us\": %.4f, \"load_rt_us\": %.4f, "
              "\"store_rt_us\": %.4f, \"remote_load_rt_us\": %.4f, \"hop_same_xcd_us\": %.4f, "
              "\"hop_same_xcd_l2_us\": %.4f, "
              "\"hop_other_xcd_us\": %.4f, \"hop_far_xcd_us\": %.4f, "
              "\"same_address_rmw_ns\": %.3f}\n",
              nwg, lanes, L, pass_ms * 1000 / passes, bytes * passes / (pass_ms * 1e6),
              chain_us[0], chain_us[1], chain_us[2], remote_us, hop_same, hop_same_l2, hop_other,
              hop_far,
              one_ms * 1e6 / ((double)it * 1024));
  CHECK(hipFree(data));
  CHECK(hipFree(buf));
  CHECK(hipFree(res));
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && std::string(argv[1]) == "floor")  // hs_probe floor <WGs> <lanes> <numLdSt>
    return floor_terms(argc >= 3 ? std::atoi(argv[2]) : 1024, argc >= 4 ? std::atoi(argv[3]) : 256,
                       argc >= 5 ? std::atoi(argv[4]) : 10);
  if (argc >= 2 && std::string(argv[1]) == "copy") {  // hs_probe copy <MiB>
    const long mib = argc >= 3 ? std::atol(argv[2]) : 256;
    const long n = mib * (1L << 20) / 16;
    int4 *src, *dst;
    CHECK(hipMalloc(&src, n * 16));
    CHECK(hipMalloc(&dst, n * 16));
    CHECK(hipMemset(src, 1, n * 16));
    CHECK(hipDeviceSynchronize());
    const float ms = time_ms([&] { k_copy<<<4096, 256>>>(src, dst, n); });
    std::printf("copy %ld MiB: %.3f ms, %.1f GB/s (read+write)\n", mib, ms,
                2.0 * n * 16 / (ms * 1e6));
    return 0;
  }
  uint32_t *buf, *res;
  int32_t *data;
  CHECK(hipMalloc(&buf, 1 << 24));
  CHECK(hipMalloc(&res, 4096));
  CHECK(hipMalloc(&data, 4096L * 2816 * 4));
  CHECK(hipMemset(data, 0, 4096L * 2816 * 4));
  {
    CHECK(hipMemset(buf, 0xff, 4096 * 4));
    k_xcc<<<2048, 64>>>(buf);
    CHECK(hipDeviceSynchronize());
    std::vector<uint32_t> h(2048);
    CHECK(hipMemcpy(h.data(), buf, 2048 * 4, hipMemcpyDeviceToHost));
    int rr = 0;
    for (int i = 0; i < 2048; ++i) rr += h[i] == (h[0] + i) % 8;
    std::printf("placement: WG b on XCD (b + %u) mod 8 for %d/2048 WGs\n", h[0], rr);
  }
  for (int mode = 0; mode <= 3; ++mode)
    for (int nwg : {1, 8, 256, 1024}) {
      const int it = 200;
      const float ms = time_ms([&] { k_fence<<<nwg, 256>>>(mode, it, data, buf); });
      static const char *kName[] = {"release", "acquire", "release+acquire",
                                    "release+acquire, 11 KB dirty"};
      std::printf("fence %-30s %4d WGs: %8.3f us per WG-iteration\n", kName[mode], nwg,
                  ms * 1000 / it);
    }
  for (auto pr : std::vector<std::pair<int, int>>{{0, 8}, {0, 1}, {0, 4}}) {
    const int rounds = 2000;
    CHECK(hipMemset(res, 0, 64));
    const float ms = time_ms([&] {
      CHECK(hipMemset(buf, 0, 256));
      k_pingpong<<<16, 64>>>(pr.first, pr.second, rounds, buf, res, 0);
    });
    uint32_t hr[4];
    CHECK(hipMemcpy(hr, res, 16, hipMemcpyDeviceToHost));
    std::printf("hand-off WG%d<->WG%d (%s): %.3f us one way%s\n", pr.first, pr.second,
                pr.second == 8 ? "same XCD" : "other XCD", ms * 1000 / (2 * rounds),
                hr[2] ? " (TIMEOUT)" : "");
  }
  for (int mode = 0; mode < 5; ++mode)
    for (int nwg : {256, 1024}) {  // mode 4 spans 8 x 1024 x 256 words = 8 MiB of buf
      const int it = mode >= 3 ? 100 : 1000;
      CHECK(hipMemset(buf, 0, 1 << 24));
      const float ms = time_ms([&] { k_atomic<<<nwg, 256>>>(mode, it, buf); });
      const double ops = (double)it * nwg * (mode >= 3 ? 256 : 1);
      static const char *kName[] = {"one address", "one per XCD", "one per WG", "one per lane",
                                    "8 per lane"};
      std::printf("atomic add %-12s %4d WGs: %9.3f G ops/s\n", kName[mode], nwg, ops / ms / 1e6);
    }
  for (int wmode = 0; wmode < 3; ++wmode)
    for (int rmode = 0; rmode < 3; ++rmode)
      for (int pr = 0; pr < 2; ++pr) {
        const int a = 0, b = pr ? 1 : 8, rounds = 2000;
        CHECK(hipMemset(buf, 0, 4096));
        CHECK(hipMemset(res, 0, 64));
        CHECK(hipMemset(data, 0, 1 << 20));
        k_vis<<<16, 256>>>(a, b, rounds, wmode, rmode, buf, data, res);
        CHECK(hipDeviceSynchronize());
        uint32_t hr[4];
        CHECK(hipMemcpy(hr, res, 16, hipMemcpyDeviceToHost));
        static const char *kW[] = {"plain st", "plain st + release", "agent st"};
        static const char *kR[] = {"plain ld", "agent ld", "acquire + plain ld"};
        std::printf("visibility %-18s -> %-18s %s: stale %u / %d%s\n", kW[wmode], kR[rmode],
                    pr ? "other XCD" : "same XCD ", hr[1], rounds * 256,
                    hr[3] ? " (TIMEOUT)" : "");
      }
  {  // a line's return to an XCD: WGs 0 and 8 share an XCD, WG 1 is on another
    static const int kF[3][2] = {{0, 0}, {0, 2}, {2, 2}};
    static const char *kN[] = {"plain", "", "agent"};
    for (int m = 0; m < 3; ++m) {
      const int rounds = 2000;
      CHECK(hipMemset(buf, 0, 4096));
      CHECK(hipMemset(res, 0, 64));
      CHECK(hipMemset(data, 0, 1 << 20));
      k_return<<<16, 256>>>(0, 8, 1, rounds, kF[m][0], kF[m][1], buf, data, res);
      CHECK(hipDeviceSynchronize());
      uint32_t hr[5];
      CHECK(hipMemcpy(hr, res, 20, hipMemcpyDeviceToHost));
      std::printf("return x:%s st, x:agent ld + %s st, y:agent ld + agent st, x:agent ld: "
                  "stale %u / %u / %u of %d%s\n",
                  kN[kF[m][0]], kN[kF[m][1]], hr[1], hr[2], hr[3], rounds * 256,
                  hr[4] ? " (TIMEOUT)" : "");
    }
  }
  std::printf("done\n");
  return 0;
}
