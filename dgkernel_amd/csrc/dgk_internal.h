// Internal definitions shared by the libdgk translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/dgk.h"

namespace dgk {

void set_error(const char* fmt, ...);
void clear_error();
// Resolves this device's address of the allocation-free partition table;
// called once per device handle at creation, with the device current.
int noalloc_init(dgk_device dev);

}  // namespace dgk

struct dgk_device_s {
  int id = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  int lds_bytes = 0;
  int arch_major = 0, arch_minor = 0;
  char name[256] = {0};
  bool capturing = false;
  // Device copy of a host K array for launches over more than 256
  // partitions (grown on demand, freed with the device).
  void* kcache = nullptr;
  size_t kcache_bytes = 0;
  // This device's address of the allocation-free partition table (set at
  // device creation, see deriv.hip).
  int64_t* noalloc_k = nullptr;
};

#define DGK_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      dgk::set_error("%s failed at %s:%d: %s", #call, __FILE__, __LINE__,               \
                     hipGetErrorString(e_));                                            \
      return DGK_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define DGK_REQUIRE(cond, ...)        \
  do {                                \
    if (!(cond)) {                    \
      dgk::set_error(__VA_ARGS__);    \
      return DGK_ERR_INVALID;         \
    }                                 \
  } while (0)

#define DGK_LAUNCH_CHECK()                                                            \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess) {                                                           \
      dgk::set_error("kernel launch failed at %s:%d: %s", __FILE__, __LINE__,         \
                     hipGetErrorString(e_));                                          \
      return DGK_ERR_HIP;                                                             \
    }                                                                                 \
  } while (0)

namespace dgk {

inline int np_of(int N) { return (N + 1) * (N + 2) * (N + 3) / 6; }
inline int nfp_of(int N) { return (N + 1) * (N + 2) / 2; }

// Element tiles over partitions: tile t covers elements [first, first+count)
// of partition part (local numbering). Built on the host from K.
struct TileList {
  std::vector<int32_t> part, first, count;
};
TileList make_tiles(int npart, const int64_t* K, int per_tile);

}  // namespace dgk
