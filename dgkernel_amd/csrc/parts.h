// Partition tiling shared by the per-partition kernels (derivative, MATMUL,
// pointwise statements of source-built kernels).
//
// A launch covers every partition at once. Up to kMaxPartsByValue partitions
// the tile table travels by value in the kernel arguments: tile t of a flat
// list covers elements [first, first+count) of one partition (no empty
// blocks for uneven partitions). Above that the grid is two-dimensional:
// blockIdx.y is the partition, blockIdx.x the tile within it, and K[part] is
// read from a device array (the runner's K array, int_t wide, or a per-device
// copy of the host K the library keeps).
#pragma once

#include "dgk_internal.h"

namespace dgk {

constexpr int kMaxPartsByValue = 256;

struct PartTiles {
  int npart;
  int mode2d;          // 1: part = blockIdx.y, K from kdev
  const void* kdev;    // device K (int32 or int64 per ksize) when mode2d
  int ksize;
  int start[kMaxPartsByValue + 1];  // prefix sum of tiles (flat mode)
  int K[kMaxPartsByValue];          // elements per partition (flat mode)
};

// Tile tx of the flat list (flat mode), or tile tx of partition blockIdx.y
// (2-D mode); false if it has no elements.
__device__ __forceinline__ bool find_tile_at(const PartTiles& pt, int tx, int per_tile, int& part,
                                             int& first, int& count) {
  if (pt.mode2d) {
    part = blockIdx.y;
    first = tx * per_tile;
    long long k = pt.ksize == 8 ? (long long)((const int64_t*)pt.kdev)[part]
                                : (long long)((const int32_t*)pt.kdev)[part];
    long long rem = k - first;
    count = rem < per_tile ? (int)rem : per_tile;
    return count > 0;
  }
  const int tile = tx;
  int lo = 0, hi = pt.npart;  // last p with start[p] <= tile
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pt.start[mid] <= tile) lo = mid; else hi = mid;
  }
  part = lo;
  first = (tile - pt.start[lo]) * per_tile;
  int rem = pt.K[lo] - first;
  count = rem < per_tile ? rem : per_tile;
  return count > 0;
}

// Tile of this block; false if the block has no elements (2-D mode only).
__device__ __forceinline__ bool find_tile(const PartTiles& pt, int per_tile, int& part, int& first,
                                          int& count) {
  if (pt.mode2d) {
    part = blockIdx.y;
    first = blockIdx.x * per_tile;
    long long k = pt.ksize == 8 ? (long long)((const int64_t*)pt.kdev)[part]
                                : (long long)((const int32_t*)pt.kdev)[part];
    long long rem = k - first;
    count = rem < per_tile ? (int)rem : per_tile;
    return count > 0;
  }
  const int tile = blockIdx.x;
  int lo = 0, hi = pt.npart;  // last p with start[p] <= tile
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pt.start[mid] <= tile) lo = mid; else hi = mid;
  }
  part = lo;
  first = (tile - pt.start[lo]) * per_tile;
  int rem = pt.K[lo] - first;
  count = rem < per_tile ? rem : per_tile;
  return count > 0;
}

// Host: tile table and grid for K (host values). kdev/ksize: a device copy of
// K if the caller has one (the runner's K array); else, above
// kMaxPartsByValue partitions, K is copied into a per-device buffer
// (synchronous; that call is not graph-capturable). grid.x == 0: no work.
// force2d: use the two-dimensional grid (K from kdev, which must be given)
// whatever the partition count.
int build_part_tiles(dgk_device dev, int npart, const int64_t* K, int per_tile, const void* kdev,
                     int ksize, PartTiles& pt, dim3& grid, bool force2d = false);

}  // namespace dgk
