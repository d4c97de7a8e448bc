// hs_host.h -- host-side internals shared by the library's translation
// units: the context, error plumbing and small device-buffer helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/heterosync.h"

struct hs_ctx {
  int device = 0;
  int num_cu = 0;
  int max_threads = 0;
  int max_wgs = 0;
  double wall_hz = 0;  // wall_clock64() ticks per second
  char arch[64] = {0};
  hipStream_t stream = nullptr;
  // the no-allocation entry point's per-device scratch array (device
  // address from hipGetSymbolAddress) and its size in words
  uint32_t *scratch = nullptr;
  size_t scratch_words = 0;
  // engine kept between UTS calls, released by hs_close
  void *uts_cache = nullptr;
  void (*uts_cache_free)(void *) = nullptr;
};

namespace hs {

#define HS_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      std::fprintf(stderr, "HIP error %s at %s:%d: %s\n", #call, __FILE__, \
                   __LINE__, hipGetErrorString(e_));                      \
      return HS_EHIP;                                                     \
    }                                                                     \
  } while (0)

// Owns a set of device allocations; frees them on scope exit.
class DevArena {
 public:
  ~DevArena() {
    for (void *p : ptrs_) (void)hipFree(p);
  }
  template <typename T>
  int alloc(T **out, size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return HS_ENOMEM;
    }
    ptrs_.push_back(p);
    *out = static_cast<T *>(p);
    return HS_OK;
  }

 private:
  std::vector<void *> ptrs_;
};

// Owns a set of HIP events; destroys them on every return path.
class EventSet {
 public:
  ~EventSet() {
    for (hipEvent_t e : ev_)
      if (e) (void)hipEventDestroy(e);
  }
  int create(size_t n) {
    ev_.assign(n, nullptr);
    for (hipEvent_t &e : ev_)
      if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
        (void)hipGetLastError();
        return HS_EHIP;
      }
    return HS_OK;
  }
  hipEvent_t operator[](size_t i) const { return ev_[i]; }

 private:
  std::vector<hipEvent_t> ev_;
};

inline uint64_t timeout_ticks(const hs_ctx *c, double s) {
  if (s <= 0) s = 20.0;
  return (uint64_t)(s * c->wall_hz);
}

// XCDs of the device: gfx950 (MI355X) / gfx942 (MI300X) have 32 CUs per XCD.
inline int num_xcds(const hs_ctx *c) {
  if (std::strncmp(c->arch, "gfx950", 6) && std::strncmp(c->arch, "gfx942", 6)) return 1;
  return c->num_cu >= 64 && c->num_cu % 32 == 0 ? c->num_cu / 32 : 1;
}

// Looks up the device address of the no-allocation scratch array for the
// context's device (current when called).
int scratch_init(hs_ctx *c);

// Max co-resident WGs of `kernel` at `wg` threads per WG on this device.
int coresident_wgs(const hs_ctx *c, const void *kernel, int wg, int *out);

// Makes HIP load `kernel`'s code object now.  HIP loads a module on the
// first use of one of its kernels; a timed region that begins with that
// first launch would include the load (about 3 ms measured on MI355X).
int load_kernel(const void *kernel);

}  // namespace hs
