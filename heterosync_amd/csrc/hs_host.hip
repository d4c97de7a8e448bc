// hs_host.hip -- context, device query, error strings and the primitive
// table of the C ABI.
#include <cstring>

#include "hs_host.h"

namespace {

// main.hip.cpp:2036-2089 (ids 36 and 43-46 are reserved there).
struct PrimName {
  const char *name;
  int id;
};
const PrimName kPrims[] = {
    {"atomicTreeBarrUniq", 0},
    {"atomicTreeBarrUniqLocalExch", 1},
    {"lfTreeBarrUniq", 2},
    {"lfTreeBarrUniqLocalExch", 3},
    {"spinMutex", 4},
    {"spinMutexEBO", 5},
    {"sleepMutex", 6},
    {"faMutex", 7},
    {"spinSem1", 8},
    {"spinSem2", 9},
    {"spinSem10", 10},
    {"spinSem120", 11},
    {"spinSemEBO1", 12},
    {"spinSemEBO2", 13},
    {"spinSemEBO10", 14},
    {"spinSemEBO120", 15},
    {"PRspinSem1", 16},
    {"PRspinSem2", 17},
    {"PRspinSem10", 18},
    {"PRspinSem120", 19},
    {"spinMutexUniq", 20},
    {"spinMutexEBOUniq", 21},
    {"sleepMutexUniq", 22},
    {"faMutexUniq", 23},
    {"spinSemUniq1", 24},
    {"spinSemUniq2", 25},
    {"spinSemUniq10", 26},
    {"spinSemUniq120", 27},
    {"spinSemEBOUniq1", 28},
    {"spinSemEBOUniq2", 29},
    {"spinSemEBOUniq10", 30},
    {"spinSemEBOUniq120", 31},
    {"PRspinSemEBO1", 32},
    {"PRspinSemEBO2", 33},
    {"PRspinSemEBO10", 34},
    {"PRspinSemEBO120", 35},
    {"atomicTreeBarrUniqSRB", 37},
    {"atomicTreeBarrUniqLocalExchSRB", 38},
    {"atomicTreeBarrUniqNaiveSRB", 39},
    {"atomicTreeBarrUniqLocalExchNaiveSRB", 40},
    {"atomicTreeBarrUniqNaiveAllSRB", 41},
    {"atomicTreeBarrUniqLocalExchNaiveAllSRB", 42},
};

}  // namespace

namespace hs {

int coresident_wgs(const hs_ctx *c, const void *kernel, int wg, int *out) {
  int per_cu = 0;
  HS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, wg, 0));
  *out = per_cu * c->num_cu;
  return HS_OK;
}

int load_kernel(const void *kernel) {
  int per_cu = 0;  // the occupancy query needs the loaded code object
  HS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0));
  return HS_OK;
}

}  // namespace hs

extern "C" {

const char *hs_strerror(int code) {
  switch (code) {
    case HS_OK: return "ok";
    case HS_EINVAL: return "invalid argument or unsupported shape";
    case HS_EHIP: return "HIP runtime error";
    case HS_ENOMEM: return "device allocation failed";
    case HS_EOCCUPANCY: return "workgroups cannot all be co-resident";
    case HS_ETIMEOUT: return "device watchdog fired (wait exceeded timeout)";
    case HS_ENODEV: return "no HIP device";
    case HS_EUNSUPPORTED: return "primitive not implemented by this library";
    case HS_ESCRATCH: return "workload needs more than the no-allocation scratch (4 MiB)";
    default: return "unknown error";
  }
}

int hs_open(int device, hs_ctx **out) {
  if (!out) return HS_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return HS_ENODEV;
  }
  if (device < 0 || device >= n) return HS_EINVAL;
  HS_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  HS_HIP(hipGetDeviceProperties(&prop, device));
  int wall_khz = 0;
  HS_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device));
  hs_ctx *c = new hs_ctx();
  c->device = device;
  c->num_cu = prop.multiProcessorCount;
  c->max_threads = prop.maxThreadsPerBlock;
  c->max_wgs = (prop.maxThreadsPerBlock / 64) * prop.multiProcessorCount;
  c->wall_hz = (wall_khz > 0 ? wall_khz : 100000) * 1000.0;
  std::strncpy(c->arch, prop.gcnArchName, sizeof(c->arch) - 1);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return HS_EHIP;
  }
  if (hs::scratch_init(c) != HS_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return HS_EHIP;
  }
  *out = c;
  return HS_OK;
}

void hs_close(hs_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->uts_cache_free) c->uts_cache_free(c->uts_cache);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int hs_device_info(hs_ctx *c, int *num_cu, int *max_threads, int *max_wgs,
                   char *arch, int arch_len) {
  if (!c) return HS_EINVAL;
  if (num_cu) *num_cu = c->num_cu;
  if (max_threads) *max_threads = c->max_threads;
  if (max_wgs) *max_wgs = c->max_wgs;
  if (arch && arch_len > 0) {
    std::strncpy(arch, c->arch, arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return HS_OK;
}

int hs_prim_id(const char *name) {
  if (!name) return -1;
  for (const auto &p : kPrims)
    if (std::strcmp(name, p.name) == 0) return p.id;
  return -1;
}

const char *hs_prim_name(int id) {
  for (const auto &p : kPrims)
    if (p.id == id) return p.name;
  return nullptr;
}

int hs_prim_supported(int id) {
  return (id >= 0 && id <= 35) || (id >= 37 && id <= 42);
}

long hs_storage_words(int p, int num_wgs, int num_cu, int wg, int L) {
  if (num_wgs <= 0 || num_cu <= 0 || wg <= 0 || L < 0) return HS_EINVAL;
  long mult;
  if (p < 4 || p >= 37 || (p >= 24 && p <= 35))
    mult = num_wgs;                                    // barriers, uniq sems
  else if (p >= 4 && p <= 7)
    mult = 1;                                          // mutexes
  else if (p >= 8 && p <= 19)
    mult = (num_wgs + num_cu - 1) / num_cu;            // semaphores
  else if (p >= 20 && p <= 23)
    mult = num_wgs < num_cu ? num_wgs : num_cu;        // uniq mutexes
  else
    return HS_EINVAL;
  return mult * (long)wg * (long)(L + 1);
}

// The reference binary's own self-check (main.hip.cpp:2524-2702) in closed
// form.  The kernels' region walk fixes how many critical-section passes k
// touch each WG-sized region; a chain of L+1 16-word lines that k passes
// touched ends as "line j = original line max(0, j-k)" (copy) or
// "original + k on lines 1..L" (inc).
int hs_syncprim_expected(int p, int L, int nw, int iters, int reps, int wg, int ncu, int mode,
                         int32_t *out, long n_words) {
  if (!out || nw <= 0 || ncu <= 0 || wg <= 0 || L < 0 || iters < 0 || reps <= 0 ||
      (mode != 0 && mode != 1) || !hs_prim_supported(p))
    return HS_EINVAL;
  const long n = hs_storage_words(p, nw, ncu, wg, L);
  if (n < 0 || n != n_words) return HS_EINVAL;
  const long ir = (long)iters * reps;
  const int nab = nw < ncu ? nw : ncu;
  const int per = (nw + nab - 1) / nab;
  long writers = 0;  // semaphores: one writer per group (main.hip.cpp:883)
  for (int b = 0; b < nw; ++b) writers += (b / nab == (b % nab) % per);
  std::vector<long> k;
  if (p < 4 || p >= 37) {  // barriers: every region, twice with local exchange
    const bool lx = p == 1 || p == 3 || p == 38 || p == 40 || p == 42;
    k.assign(nw, ir * ((lx && per > 1) ? 2 : 1));
  } else if (p <= 7) {
    k.assign(1, (long)nw * ir);
  } else if (p <= 19) {  // global semaphores: regions 0..perCU-1
    k.assign(per, writers * ir);
  } else if (p >= 32) {  // the global EBO priority kernel over num_wgs regions
    k.assign(nw, 0);
    for (int r = 0; r < per && r < nw; ++r) k[r] = writers * ir;
  } else if (p <= 23) {  // uniq mutexes: one region per group
    k.assign(nab, 0);
    for (int b = 0; b < nw; ++b) k[b % ncu] += ir;
  } else {  // uniq semaphores: each writer walks its group's regions
    k.assign(nw, 0);
    for (int b = 0; b < nw; ++b) {
      if (b / nab != (b % nab) % per) continue;
      int cur = b;
      for (int j = 0; j < per; ++j) {
        k[cur] += ir;
        cur = (cur + nab) % nw;
      }
    }
  }
  const long region = (long)wg * (L + 1);
  if ((long)k.size() * region != n) return HS_EINVAL;
  const long chain = 16L * (L + 1);
  for (long i = 0; i < n; ++i) {
    const long kk = k[i / region];
    const long c0 = (i / chain) * chain, j = (i - c0) / 16, w = (i - c0) % 16;
    out[i] = (int32_t)(mode == 0 ? c0 + (j - kk > 0 ? j - kk : 0) * 16 + w
                                 : i + (j >= 1 ? kk : 0));
  }
  return HS_OK;
}

}  // extern "C"
