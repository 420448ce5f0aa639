// FaceConnector exchange on one device.
//
// The reference builds pick/place lists per (source, destination) partition
// pair (utils/face_connector.go:143-187) and exchanges through them as
//   Pbuf[dst][place[dst][src][i]] = u[src][pick[src][dst][i]]
// (utils/face_connector_test.go:104-156), src == dst included. With every
// partition resident in one pooled allocation this is one gather-scatter
// launch over all pairs: the pair of each entry is found by binary search on
// the segment starts (a few scalar loads that stay in cache), the copy is a
// plain 8-byte move, so the result is bit-exact by construction.
//
// When the place lists give every P-buffer slot of every destination exactly
// once -- which the FaceConnector's do, one slot per face point -- the plan
// stores the entries in slot order instead (destination-major, then slot).
// The launch then writes P with unit stride and gathers the reads, which are
// neighbour traces and mostly hit in cache; any other lists keep the
// pair-ordered form.
#include <vector>

#include "dgk_internal.h"

struct dgk_face_plan_s {
  dgk_device dev = nullptr;
  int npart = 0;
  int nseg = 0;
  int64_t total = 0;
  // slot-ordered form (dense): entry i fills slot i - dst_start[d] of partition d
  bool dense = false;
  int64_t* dst_start = nullptr;        // [npart+1] device
  int32_t* src_of = nullptr;           // [total] device
  // pair-ordered form (dense == false), and the source index of the dense form
  int64_t* seg_pick_start = nullptr;   // [nseg+1] device
  int64_t* seg_place_start = nullptr;  // [nseg] device
  int32_t* seg_src = nullptr;          // [nseg] device
  int32_t* seg_dst = nullptr;          // [nseg] device
  int32_t* pick = nullptr;             // [total] device
  int32_t* place = nullptr;            // [total] device
};

namespace {

__global__ void __launch_bounds__(256)
face_exchange_kernel(int nseg, int64_t total, const int64_t* __restrict__ seg_pick_start,
                     const int64_t* __restrict__ seg_place_start,
                     const int32_t* __restrict__ seg_src, const int32_t* __restrict__ seg_dst,
                     const int32_t* __restrict__ pick, const int32_t* __restrict__ place,
                     const double* __restrict__ u, const int64_t* __restrict__ u_off,
                     double* __restrict__ P, const int64_t* __restrict__ P_off) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    int lo = 0, hi = nseg;  // last s with seg_pick_start[s] <= i
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (seg_pick_start[mid] <= i) lo = mid; else hi = mid;
    }
    int64_t k = i - seg_pick_start[lo];
    int src = seg_src[lo], dst = seg_dst[lo];
    P[P_off[dst] + place[seg_place_start[lo] + k]] = u[u_off[src] + pick[i]];
  }
}

// Slot-ordered form: unit-stride writes of P, gathered reads of u.
__global__ void __launch_bounds__(256)
face_exchange_dense_kernel(int npart, int64_t total, const int64_t* __restrict__ dst_start,
                           const int32_t* __restrict__ src_of, const int32_t* __restrict__ pick,
                           const double* __restrict__ u, const int64_t* __restrict__ u_off,
                           double* __restrict__ P, const int64_t* __restrict__ P_off) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < total; i += stride) {
    int lo = 0, hi = npart;  // last d with dst_start[d] <= i
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (dst_start[mid] <= i) lo = mid; else hi = mid;
    }
    __builtin_nontemporal_store(u[u_off[src_of[i]] + pick[i]], P + P_off[lo] + (i - dst_start[lo]));
  }
}

}  // namespace

extern "C" int dgk_face_plan_create(dgk_device dev, int npart, const int64_t* counts_host,
                                    const int32_t* pick_host, const int32_t* place_host,
                                    dgk_face_plan* out) {
  dgk::clear_error();
  DGK_REQUIRE(dev && counts_host && out && npart > 0, "dgk_face_plan_create: bad argument");
  *out = nullptr;
  // Segments = non-empty (src, dst) pairs in src-major order; place lists are
  // stored dst-major, so each pair's place start is found by a second scan.
  std::vector<int64_t> pick_start, place_start;
  std::vector<int32_t> src, dst;
  std::vector<int64_t> place_pos((size_t)npart * npart, 0);
  int64_t acc = 0;
  for (int d = 0; d < npart; ++d)
    for (int s = 0; s < npart; ++s) {
      place_pos[(size_t)s * npart + d] = acc;
      acc += counts_host[(size_t)s * npart + d];
    }
  int64_t total = 0;
  for (int s = 0; s < npart; ++s)
    for (int d = 0; d < npart; ++d) {
      int64_t c = counts_host[(size_t)s * npart + d];
      DGK_REQUIRE(c >= 0, "dgk_face_plan_create: negative count");
      if (c == 0) continue;
      pick_start.push_back(total);
      place_start.push_back(place_pos[(size_t)s * npart + d]);
      src.push_back(s);
      dst.push_back(d);
      total += c;
    }
  DGK_REQUIRE(total == acc, "dgk_face_plan_create: inconsistent counts");
  DGK_REQUIRE(total == 0 || (pick_host && place_host), "dgk_face_plan_create: NULL lists");
  pick_start.push_back(total);
  dgk_face_plan p = new dgk_face_plan_s();
  p->dev = dev;
  p->npart = npart;
  p->nseg = (int)src.size();
  p->total = total;
  // Slot order: destination d receives sum_s counts[s][d] values; if its place
  // entries cover slots 0 .. that count - 1 exactly once, entry (s, d, k) moves
  // to position dst_start[d] + place.
  std::vector<int64_t> dst_start((size_t)npart + 1, 0);
  for (int d = 0; d < npart; ++d) {
    int64_t n = 0;
    for (int s2 = 0; s2 < npart; ++s2) n += counts_host[(size_t)s2 * npart + d];
    dst_start[d + 1] = dst_start[d] + n;
  }
  std::vector<int32_t> src_sorted, pick_sorted;
  bool dense = total > 0;
  if (dense) {
    src_sorted.assign((size_t)total, -1);
    pick_sorted.resize((size_t)total);
    for (size_t g = 0; g < src.size() && dense; ++g) {
      const int s2 = src[g], d = dst[g];
      const int64_t n = pick_start[g + 1] - pick_start[g], slots = dst_start[d + 1] - dst_start[d];
      for (int64_t k = 0; k < n; ++k) {
        const int64_t pl = place_host[place_start[g] + k];
        if (pl < 0 || pl >= slots || src_sorted[(size_t)(dst_start[d] + pl)] >= 0) {
          dense = false;
          break;
        }
        src_sorted[(size_t)(dst_start[d] + pl)] = s2;
        pick_sorted[(size_t)(dst_start[d] + pl)] = pick_host[pick_start[g] + k];
      }
    }
  }
  p->dense = dense;
  auto up = [&](void** dptr, const void* h, size_t bytes) -> int {
    return dgk_malloc(dev, bytes, h, dptr);
  };
  int rc = DGK_OK;
  if (dense) {
    rc |= up((void**)&p->dst_start, dst_start.data(), dst_start.size() * 8);
    rc |= up((void**)&p->src_of, src_sorted.data(), (size_t)total * 4);
    rc |= up((void**)&p->pick, pick_sorted.data(), (size_t)total * 4);
  } else {
    rc |= up((void**)&p->seg_pick_start, pick_start.data(), pick_start.size() * 8);
    if (p->nseg) {
      rc |= up((void**)&p->seg_place_start, place_start.data(), place_start.size() * 8);
      rc |= up((void**)&p->seg_src, src.data(), src.size() * 4);
      rc |= up((void**)&p->seg_dst, dst.data(), dst.size() * 4);
    }
    if (total) {
      rc |= up((void**)&p->pick, pick_host, (size_t)total * 4);
      rc |= up((void**)&p->place, place_host, (size_t)total * 4);
    }
  }
  if (rc != DGK_OK) {
    dgk_face_plan_destroy(p);
    return DGK_ERR_HIP;
  }
  *out = p;
  return DGK_OK;
}

extern "C" int dgk_face_plan_destroy(dgk_face_plan p) {
  if (!p) return DGK_OK;
  dgk_free(p->dev, p->dst_start);
  dgk_free(p->dev, p->src_of);
  dgk_free(p->dev, p->seg_pick_start);
  dgk_free(p->dev, p->seg_place_start);
  dgk_free(p->dev, p->seg_src);
  dgk_free(p->dev, p->seg_dst);
  dgk_free(p->dev, p->pick);
  dgk_free(p->dev, p->place);
  delete p;
  return DGK_OK;
}

extern "C" int dgk_face_exchange(dgk_face_plan p, const double* u_dev, const int64_t* u_off_dev,
                                 double* P_dev, const int64_t* P_off_dev) {
  dgk::clear_error();
  DGK_REQUIRE(p && u_dev && u_off_dev && P_dev && P_off_dev, "dgk_face_exchange: NULL argument");
  if (p->total == 0) return DGK_OK;
  int64_t blocks = (p->total + 255) / 256;
  int64_t cap = (int64_t)p->dev->cu_count * 16;
  if (blocks > cap) blocks = cap;
  if (p->dense)
    hipLaunchKernelGGL(face_exchange_dense_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       p->dev->stream, p->npart, p->total, p->dst_start, p->src_of, p->pick, u_dev,
                       u_off_dev, P_dev, P_off_dev);
  else
    hipLaunchKernelGGL(face_exchange_kernel, dim3((unsigned)blocks), dim3(256), 0, p->dev->stream,
                       p->nseg, p->total, p->seg_pick_start, p->seg_place_start, p->seg_src,
                       p->seg_dst, p->pick, p->place, u_dev, u_off_dev, P_dev, P_off_dev);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}
