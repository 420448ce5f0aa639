// Physical derivative and MATMUL kernels for gfx950 over the runner's pooled
// partition layout.
//
// Reference semantics: the OKL kernel "differentiate"
// (tetnudg/tet_nudg_mesh_physical_derivative_test.go:850-892) computes
//   Ur = Dr u, Us = Ds u, Ut = Dt u            (MATMUL_Dr/Ds/Dt macros,
//                                               runner/builder/builder.go:355-386)
//   du/dx = rx Ur + sx Us + tx Ut, likewise y, z (per node).
// The reference runs one thread per element with three global temporaries
// and stride-Np loads. Here one block owns a tile of elements of one
// partition: the tile's u is staged through LDS with contiguous (coalesced)
// loads, Dr/Ds/Dt sit in LDS, each thread produces one node of one element
// for a group of elements (register blocking over EPG elements reuses each
// D row), and the per-node metrics and outputs stream with unit stride. The
// temporaries Ur/Us/Ut are written only when the caller names them (a
// source-built "differentiate" kernel does, the direct entry point does not).
//
// The operators are read through strides, D(n, j) = D[n*rs + j*cs], so the
// runner's column-major device matrices (runner/runner.go:395-400) are used
// in place.
#include "parts.h"

namespace dgk {

int build_part_tiles(dgk_device dev, int npart, const int64_t* K, int per_tile, const void* kdev,
                     int ksize, PartTiles& pt, dim3& grid, bool force2d) {
  DGK_REQUIRE(npart >= 1, "partition count %d < 1", npart);
  DGK_REQUIRE(per_tile >= 1, "tile size %d < 1", per_tile);
  pt.npart = npart;
  pt.mode2d = 0;
  pt.kdev = nullptr;
  pt.ksize = 8;
  int64_t kmax = 0, tiles = 0;
  for (int p = 0; p < npart; ++p) {
    DGK_REQUIRE(K[p] >= 0 && K[p] < (1LL << 31), "K[%d] = %lld out of range", p, (long long)K[p]);
    kmax = std::max<int64_t>(kmax, K[p]);
    tiles += (K[p] + per_tile - 1) / per_tile;
  }
  DGK_REQUIRE(!force2d || kdev, "two-dimensional tiling needs a device K");
  if (npart <= kMaxPartsByValue && !force2d) {
    DGK_REQUIRE(tiles < (1LL << 31), "too many tiles (%lld)", (long long)tiles);
    int t = 0;
    for (int p = 0; p < npart; ++p) {
      pt.start[p] = t;
      pt.K[p] = (int)K[p];
      t += (int)((K[p] + per_tile - 1) / per_tile);
    }
    pt.start[npart] = t;
    grid = dim3((unsigned)t, 1, 1);
    return DGK_OK;
  }
  DGK_REQUIRE(npart <= 65535, "partition count %d above 65535", npart);
  pt.mode2d = 1;
  if (kdev) {
    DGK_REQUIRE(ksize == 4 || ksize == 8, "K element size %d", ksize);
    pt.kdev = kdev;
    pt.ksize = ksize;
  } else {
    size_t bytes = (size_t)npart * 8;
    if (dev->kcache_bytes < bytes) {
      if (dev->kcache) DGK_HIP(hipFree(dev->kcache));
      dev->kcache = nullptr;
      dev->kcache_bytes = 0;
      DGK_HIP(hipMalloc(&dev->kcache, bytes));
      dev->kcache_bytes = bytes;
    }
    DGK_HIP(hipMemcpyAsync(dev->kcache, K, bytes, hipMemcpyHostToDevice, dev->stream));
    DGK_HIP(hipStreamSynchronize(dev->stream));
    pt.kdev = dev->kcache;
    pt.ksize = 8;
  }
  grid = dim3((unsigned)((kmax + per_tile - 1) / per_tile), (unsigned)npart, 1);
  return DGK_OK;
}

}  // namespace dgk

namespace {

using dgk::PartTiles;
using dgk::find_tile;

template <typename T>
__device__ __forceinline__ T load_as(const void* p, int64_t i) {
  return static_cast<const T*>(p)[i];
}

// Arrays of one physical-derivative launch. Each array has its own
// partition offsets (value offsets, OFF wide), as the runner passes them.
struct NodalArgs {
  const double* D[3];  // Dr, Ds, Dt; D_q(n, j) = D[q][n*drs + j*dcs]
  int drs, dcs;
  const double* u;
  const void* u_off;
  const void* m[9];  // rx ry rz sx sy sz tx ty tz
  const void* m_off[9];
  void* o[3];  // du/dx, du/dy, du/dz
  const void* o_off[3];
  double* t[3];  // Ur, Us, Ut (optional)
  const void* t_off[3];
};

// A block's arrays based at its partition: every offset is read once per
// block (the stores could alias the offset tables, so offsets read per node
// would be read again after every store).
template <typename MT, typename OT>
struct NodalBase {
  const MT* m[9];
  OT* o[3];
  double* t[3];
};

template <typename MT, typename OT, typename OFF>
__device__ __forceinline__ NodalBase<MT, OT> nodal_base(const NodalArgs& a, int part) {
  NodalBase<MT, OT> b;
#pragma unroll
  for (int q = 0; q < 9; ++q) b.m[q] = (const MT*)a.m[q] + (int64_t)((const OFF*)a.m_off[q])[part];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    b.o[q] = (OT*)a.o[q] + (int64_t)((const OFF*)a.o_off[q])[part];
    b.t[q] = a.t[0] ? a.t[q] + (int64_t)((const OFF*)a.t_off[q])[part] : nullptr;
  }
  return b;
}

template <typename MT, typename OT>
__device__ __forceinline__ void load_metrics(const NodalBase<MT, OT>& b, int64_t li, double (&m)[9]) {
#pragma unroll
  for (int q = 0; q < 9; ++q) m[q] = (double)b.m[q][li];
}

// The metrics are loaded (load_metrics) before any output of the block's
// thread is stored: the stores could alias the metric arrays, so loads
// placed after a store would wait for it.
template <typename MT, typename OT>
__device__ __forceinline__ void store_node(const NodalBase<MT, OT>& b, int64_t li,
                                           const double (&m)[9], double ar, double as, double at) {
  b.o[0][li] = (OT)(m[0] * ar + m[3] * as + m[6] * at);
  b.o[1][li] = (OT)(m[1] * ar + m[4] * as + m[7] * at);
  b.o[2][li] = (OT)(m[2] * ar + m[5] * as + m[8] * at);
  if (b.t[0]) {
    b.t[0][li] = ar;
    b.t[1][li] = as;
    b.t[2][li] = at;
  }
}

// One block: tiles of G*EPG elements of one partition, tiles tx = blockIdx.x,
// blockIdx.x + gridDim.x, ... below ntx (the flat list, or the tiles of
// partition blockIdx.y), so Dr/Ds/Dt are staged in LDS once per block;
// threads = NP * G (padded to a wave multiple), thread (node n, group g)
// handles node n of elements g, g+G, ..., g+(EPG-1)G.
template <int NP, int G, int EPG, typename MT, typename OT, typename OFF>
__global__ void __launch_bounds__(256)
deriv_nodal_kernel(PartTiles pt, NodalArgs a, int ntx) {
  constexpr int E = G * EPG;
  __shared__ double sD[3 * NP * NP];
  __shared__ double sU[E * NP];
  const int tid = threadIdx.x;
  for (int i = tid; i < 3 * NP * NP; i += blockDim.x) {
    int q = i / (NP * NP), r = i - q * NP * NP;
    int n = r / NP, j = r - n * NP;
    sD[i] = a.D[q][n * a.drs + j * a.dcs];
  }
  const int n = tid % NP, g = tid / NP;
  const double* dr = sD + n * NP;
  const double* ds = sD + NP * NP + n * NP;
  const double* dt = sD + 2 * NP * NP + n * NP;
  for (int tx = blockIdx.x; tx < ntx; tx += gridDim.x) {
    int part, first, count;
    if (!find_tile_at(pt, tx, E, part, first, count)) continue;
    const int64_t ub = (int64_t)((const OFF*)a.u_off)[part] + (int64_t)first * NP;
    const NodalBase<MT, OT> base = nodal_base<MT, OT, OFF>(a, part);
    const int nval = count * NP;
    __syncthreads();  // the previous tile's readers of sU are done
    for (int i = tid; i < nval; i += blockDim.x) sU[i] = a.u[ub + i];
    __syncthreads();
    if (g >= G) continue;
    double ar[EPG], as[EPG], at[EPG];
#pragma unroll
    for (int q = 0; q < EPG; ++q) ar[q] = as[q] = at[q] = 0.0;
#pragma unroll 5
    for (int j = 0; j < NP; ++j) {
      double x = dr[j], y = ds[j], z = dt[j];
#pragma unroll
      for (int q = 0; q < EPG; ++q) {
        double uj = sU[(g + q * G) * NP + j];
        ar[q] = fma(x, uj, ar[q]);
        as[q] = fma(y, uj, as[q]);
        at[q] = fma(z, uj, at[q]);
      }
    }
    double mv[EPG][9];
#pragma unroll
    for (int q = 0; q < EPG; ++q) {
      const int e = g + q * G;
      load_metrics(base, (int64_t)(first + (e < count ? e : 0)) * NP + n, mv[q]);
    }
#pragma unroll
    for (int q = 0; q < EPG; ++q) {
      int e = g + q * G;
      if (e >= count) continue;
      store_node(base, (int64_t)(first + e) * NP + n, mv[q], ar[q], as[q], at[q]);
    }
  }
}

// Large NP (P >= 5): D rows are read from global memory (L1/L2 resident),
// u is staged in LDS.
template <int NP, int G, typename MT, typename OT, typename OFF>
__global__ void __launch_bounds__(256)
deriv_nodal_big_kernel(PartTiles pt, NodalArgs a) {
  __shared__ double sU[G * NP];
  int part, first, count;
  if (!find_tile(pt, G, part, first, count)) return;
  const int tid = threadIdx.x;
  const int64_t ub = (int64_t)((const OFF*)a.u_off)[part] + (int64_t)first * NP;
  const NodalBase<MT, OT> base = nodal_base<MT, OT, OFF>(a, part);
  for (int i = tid; i < count * NP; i += blockDim.x) sU[i] = a.u[ub + i];
  __syncthreads();
  for (int w = tid; w < count * NP; w += blockDim.x) {
    int e = w / NP, n = w - e * NP;
    double ar = 0, as = 0, at = 0;
    for (int j = 0; j < NP; ++j) {
      double uj = sU[e * NP + j];
      int di = n * a.drs + j * a.dcs;
      ar = fma(a.D[0][di], uj, ar);
      as = fma(a.D[1][di], uj, as);
      at = fma(a.D[2][di], uj, at);
    }
    double mv[9];
    load_metrics(base, (int64_t)first * NP + w, mv);
    store_node(base, (int64_t)first * NP + w, mv, ar, as, at);
  }
}

// Elements per thread group and groups per block for each order: threads =
// NP*G <= 256, LDS = 3 NP^2 + G*EPG*NP doubles.
template <int N> struct NodalShape;
template <> struct NodalShape<1> { static constexpr int G = 64, EPG = 1; };
template <> struct NodalShape<2> { static constexpr int G = 25, EPG = 2; };
template <> struct NodalShape<3> { static constexpr int G = 12, EPG = 4; };
template <> struct NodalShape<4> { static constexpr int G = 7, EPG = 4; };

template <int N, typename MT, typename OT, typename OFF>
int launch_nodal(dgk_device dev, int npart, const int64_t* K, const void* kdev, int ksize,
                 bool force2d, const NodalArgs& a) {
  constexpr int NP = (N + 1) * (N + 2) * (N + 3) / 6;
  PartTiles pt;
  dim3 grid;
  if constexpr (N <= 4) {
    constexpr int G = NodalShape<N>::G, EPG = NodalShape<N>::EPG;
    int rc = dgk::build_part_tiles(dev, npart, K, G * EPG, kdev, ksize, pt, grid, force2d);
    if (rc) return rc;
    if (grid.x == 0) return DGK_OK;
    int threads = ((NP * G + 63) / 64) * 64;
    // a few blocks per CU walk the tiles (the operators are staged once per
    // block); in the 2-D form, per partition
    const int ntx = (int)grid.x;
    const int cap = pt.mode2d ? std::max(1, 4 * dev->cu_count / (int)grid.y) : 4 * dev->cu_count;
    if ((int)grid.x > cap) grid.x = cap;
    hipLaunchKernelGGL((deriv_nodal_kernel<NP, G, EPG, MT, OT, OFF>), grid, dim3(threads), 0,
                       dev->stream, pt, a, ntx);
  } else {
    constexpr int G = 16;
    int rc = dgk::build_part_tiles(dev, npart, K, G, kdev, ksize, pt, grid, force2d);
    if (rc) return rc;
    if (grid.x == 0) return DGK_OK;
    hipLaunchKernelGGL((deriv_nodal_big_kernel<NP, G, MT, OT, OFF>), grid, dim3(256), 0, dev->stream,
                       pt, a);
  }
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <typename MT, typename OT, typename OFF>
int dispatch_nodal_N(int N, dgk_device dev, int npart, const int64_t* K, const void* kdev,
                     int ksize, bool force2d, const NodalArgs& a) {
  switch (N) {
#define DGK_CASE(n) \
  case n:           \
    return launch_nodal<n, MT, OT, OFF>(dev, npart, K, kdev, ksize, force2d, a);
    DGK_CASE(1) DGK_CASE(2) DGK_CASE(3) DGK_CASE(4) DGK_CASE(5) DGK_CASE(6)
#undef DGK_CASE
    default:
      dgk::set_error("physical derivative: order N=%d not built (1..6)", N);
      return DGK_ERR_UNSUPPORTED;
  }
}

// MATMUL over partitions, thread per (element, row):
//   OUT[e*rows + i] (= | +=) sum_j A(i,j) IN[e*cols + j],  A(i,j) = A[i*rs + j*cs].
// The sum runs in the accumulator type AT of the runner's macro ("double sum"
// or "float sum", runner/builder/builder.go:355-410): each step adds the
// product (in the common type of A and IN) to sum and rounds to AT.
template <typename MT, typename IT, typename OT, typename AT, typename OFF>
__global__ void __launch_bounds__(256)
matmul_kernel(PartTiles pt, int per_tile, int rows, int cols, const MT* __restrict__ A, int rs,
              int cs, const IT* __restrict__ in, const OFF* __restrict__ in_off, OT* __restrict__ out,
              const OFF* __restrict__ out_off, int add) {
  int part, first, count;
  if (!find_tile(pt, per_tile, part, first, count)) return;
  const int64_t ib = (int64_t)in_off[part] + (int64_t)first * cols;
  const int64_t ob = (int64_t)out_off[part] + (int64_t)first * rows;
  for (int w = threadIdx.x; w < count * rows; w += blockDim.x) {
    int e = w / rows, i = w - e * rows;
    AT s = AT(0);
    for (int j = 0; j < cols; ++j) s = (AT)(s + A[(int64_t)i * rs + (int64_t)j * cs] * in[ib + (int64_t)e * cols + j]);
    if (add) out[ob + w] = (OT)(out[ob + w] + s); else out[ob + w] = (OT)s;
  }
}

template <typename MT, typename IT, typename OT, typename AT>
int launch_matmul_t(dgk_device dev, const PartTiles& pt, dim3 grid, int per_tile, int rows, int cols,
                    const void* A, int rs, int cs, const void* in, const void* in_off, void* out,
                    const void* out_off, int add, int off_size) {
  if (off_size == 8)
    hipLaunchKernelGGL((matmul_kernel<MT, IT, OT, AT, int64_t>), grid, dim3(256), 0, dev->stream, pt,
                       per_tile, rows, cols, (const MT*)A, rs, cs, (const IT*)in,
                       (const int64_t*)in_off, (OT*)out, (const int64_t*)out_off, add);
  else
    hipLaunchKernelGGL((matmul_kernel<MT, IT, OT, AT, int32_t>), grid, dim3(256), 0, dev->stream, pt,
                       per_tile, rows, cols, (const MT*)A, rs, cs, (const IT*)in,
                       (const int32_t*)in_off, (OT*)out, (const int32_t*)out_off, add);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

}  // namespace

namespace dgk {

// Physical derivative over partitions. D rows/cols through (drs, dcs);
// per-array offsets; temps t[3] (and t_off) may be NULL.
int phys_deriv_nodal_full(dgk_device dev, int N, int npart, const int64_t* K, const void* kdev,
                          int ksize, const double* const* D, int drs, int dcs, const double* u,
                          const void* u_off, const void* const* metrics, const void* const* m_off,
                          int mtype, void* const* outs, const void* const* o_off, int otype,
                          double* const* temps, const void* const* t_off, int off_size,
                          bool force2d) {
  DGK_REQUIRE(dev && K && D && u && u_off && metrics && m_off && outs && o_off,
              "physical derivative: NULL argument");
  NodalArgs a{};
  for (int q = 0; q < 3; ++q) {
    DGK_REQUIRE(D[q], "physical derivative: operator %d is NULL", q);
    DGK_REQUIRE(outs[q] && o_off[q], "physical derivative: output %d or its offsets is NULL", q);
    a.D[q] = D[q];
    a.o[q] = outs[q];
    a.o_off[q] = o_off[q];
    a.t[q] = temps ? temps[q] : nullptr;
    a.t_off[q] = t_off ? t_off[q] : nullptr;
    DGK_REQUIRE(!temps || (temps[q] && t_off && t_off[q]),
                "physical derivative: temporary %d or its offsets is NULL", q);
  }
  for (int i = 0; i < 9; ++i) {
    DGK_REQUIRE(metrics[i] && m_off[i], "physical derivative: metric %d or its offsets is NULL", i);
    a.m[i] = metrics[i];
    a.m_off[i] = m_off[i];
  }
  a.drs = drs;
  a.dcs = dcs;
  a.u = u;
  a.u_off = u_off;
  DGK_REQUIRE(mtype == DGK_F64 || mtype == DGK_F32, "physical derivative: bad metric type");
  DGK_REQUIRE(otype == DGK_F64 || otype == DGK_F32, "physical derivative: bad output type");
  DGK_REQUIRE(off_size == 8 || off_size == 4, "physical derivative: offset size %d", off_size);
#define DGK_GO(MT, OT, OFF) \
  return dispatch_nodal_N<MT, OT, OFF>(N, dev, npart, K, kdev, ksize, force2d, a)
  if (off_size == 8) {
    if (mtype == DGK_F64 && otype == DGK_F64) DGK_GO(double, double, int64_t);
    if (mtype == DGK_F32 && otype == DGK_F64) DGK_GO(float, double, int64_t);
    if (mtype == DGK_F64 && otype == DGK_F32) DGK_GO(double, float, int64_t);
    DGK_GO(float, float, int64_t);
  } else {
    if (mtype == DGK_F64 && otype == DGK_F64) DGK_GO(double, double, int32_t);
    if (mtype == DGK_F32 && otype == DGK_F64) DGK_GO(float, double, int32_t);
    if (mtype == DGK_F64 && otype == DGK_F32) DGK_GO(double, float, int32_t);
    DGK_GO(float, float, int32_t);
  }
#undef DGK_GO
}

// Typed MATMUL: mtype/itype/otype/atype are DGK_F64 or DGK_F32.
int matmul_typed(dgk_device dev, int npart, const int64_t* K, const void* kdev, int ksize, int rows,
                 int cols, const void* A, int rs, int cs, int mtype, const void* in,
                 const void* in_off, int itype, void* out, const void* out_off, int otype,
                 int atype, int add, int off_size) {
  DGK_REQUIRE(dev && K && A && in && in_off && out && out_off, "matmul: NULL argument");
  DGK_REQUIRE(rows > 0 && cols > 0, "matmul: bad shape %dx%d", rows, cols);
  DGK_REQUIRE(off_size == 8 || off_size == 4, "matmul: offset size %d", off_size);
  int per_tile = rows >= 256 ? 1 : 256 / rows;
  PartTiles pt;
  dim3 grid;
  int rc = build_part_tiles(dev, npart, K, per_tile, kdev, ksize, pt, grid);
  if (rc) return rc;
  if (grid.x == 0) return DGK_OK;
  const int code = (mtype == DGK_F32) * 8 + (itype == DGK_F32) * 4 + (otype == DGK_F32) * 2 +
                   (atype == DGK_F32);
  switch (code) {
#define DGK_MM(c, MT, IT, OT, AT)                                                               \
  case c:                                                                                       \
    return launch_matmul_t<MT, IT, OT, AT>(dev, pt, grid, per_tile, rows, cols, A, rs, cs, in,  \
                                           in_off, out, out_off, add, off_size);
    DGK_MM(0, double, double, double, double)
    DGK_MM(1, double, double, double, float)
    DGK_MM(2, double, double, float, double)
    DGK_MM(3, double, double, float, float)
    DGK_MM(4, double, float, double, double)
    DGK_MM(5, double, float, double, float)
    DGK_MM(6, double, float, float, double)
    DGK_MM(7, double, float, float, float)
    DGK_MM(8, float, double, double, double)
    DGK_MM(9, float, double, double, float)
    DGK_MM(10, float, double, float, double)
    DGK_MM(11, float, double, float, float)
    DGK_MM(12, float, float, double, double)
    DGK_MM(13, float, float, double, float)
    DGK_MM(14, float, float, float, double)
    DGK_MM(15, float, float, float, float)
#undef DGK_MM
  }
  dgk::set_error("matmul: bad type code");
  return DGK_ERR_INVALID;
}

}  // namespace dgk

extern "C" int dgk_phys_deriv_nodal(dgk_device dev, int N, int npart, const int64_t* K_host,
                                    const double* Dr_dev, const double* Ds_dev,
                                    const double* Dt_dev, const double* u_dev,
                                    const int64_t* u_off_dev, const void* const* metrics_dev,
                                    int mtype, const int64_t* m_off_dev, void* dudx_dev,
                                    void* dudy_dev, void* dudz_dev, int otype,
                                    const int64_t* o_off_dev) {
  dgk::clear_error();
  DGK_REQUIRE(metrics_dev, "dgk_phys_deriv_nodal: NULL metrics");
  const int np = dgk::np_of(N);
  const double* D[3] = {Dr_dev, Ds_dev, Dt_dev};
  const void* moff[9];
  for (int i = 0; i < 9; ++i) moff[i] = m_off_dev;
  void* outs[3] = {dudx_dev, dudy_dev, dudz_dev};
  const void* ooff[3] = {o_off_dev, o_off_dev, o_off_dev};
  return dgk::phys_deriv_nodal_full(dev, N, npart, K_host, nullptr, 8, D, np, 1, u_dev, u_off_dev,
                                    metrics_dev, moff, mtype, outs, ooff, otype, nullptr, nullptr, 8,
                                    false);
}

// ---------------------------------------------------------------------------
// Allocation-free physical derivative.
//
// dgk_phys_deriv_nodal keeps the partition sizes of a launch over more than
// 256 partitions in a per-device buffer it allocates on first need. The
// variant below keeps them in `scratch`, a file-scope table of
// kNoAllocMaxParts int64 values (~512 KiB) that every GPU the library runs on
// holds for the life of the process. Host code never passes the array's name
// to a kernel: it uses the address hipGetSymbolAddress returns for each
// device (dgk::noalloc_init at device creation, kept in dev->noalloc_k). The
// host K is staged into the table by small kernels that carry 256 values each
// in their arguments, so the call makes no allocation, no copy from host
// memory and no synchronisation (graph-capturable). Every call on one GPU
// shares the table and overwrites its first npart entries; calls that overlap
// on one GPU (two streams) race.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocMaxParts = 65535;  // grid.y limit of the 2-D launch
constexpr int kStageChunk = 256;

static __device__ __align__(256) int64_t scratch[kNoAllocMaxParts];

struct KChunk {
  int64_t k[kStageChunk];
};

__global__ void stage_k_kernel(int64_t* __restrict__ dst, int first, int n, KChunk c) {
  const int i = threadIdx.x;
  if (i < n) dst[first + i] = c.k[i];
}

}  // namespace

int dgk::noalloc_init(dgk_device dev) {
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    dgk::set_error("hipGetSymbolAddress(partition table) failed on device %d: %s", dev->id,
                   hipGetErrorString(e));
    return DGK_ERR_HIP;
  }
  dev->noalloc_k = (int64_t*)p;
  return DGK_OK;
}

extern "C" int dgk_phys_deriv_nodal_noalloc(dgk_device dev, int N, int npart,
                                            const int64_t* K_host, const double* Dr_dev,
                                            const double* Ds_dev, const double* Dt_dev,
                                            const double* u_dev, const int64_t* u_off_dev,
                                            const void* const* metrics_dev, int mtype,
                                            const int64_t* m_off_dev, void* dudx_dev,
                                            void* dudy_dev, void* dudz_dev, int otype,
                                            const int64_t* o_off_dev) {
  dgk::clear_error();
  DGK_REQUIRE(dev && K_host && metrics_dev, "dgk_phys_deriv_nodal_noalloc: NULL argument");
  DGK_REQUIRE(dev->noalloc_k, "dgk_phys_deriv_nodal_noalloc: partition table not resolved");
  DGK_REQUIRE(npart >= 1 && npart <= kNoAllocMaxParts,
              "dgk_phys_deriv_nodal_noalloc: %d partitions outside 1..%d (the table's size)", npart,
              kNoAllocMaxParts);
  for (int p = 0; p < npart; ++p)
    DGK_REQUIRE(K_host[p] >= 0 && K_host[p] < (1LL << 31),
                "dgk_phys_deriv_nodal_noalloc: K[%d] = %lld out of range", p, (long long)K_host[p]);
  for (int first = 0; first < npart; first += kStageChunk) {
    KChunk c;
    const int n = std::min(kStageChunk, npart - first);
    for (int i = 0; i < n; ++i) c.k[i] = K_host[first + i];
    hipLaunchKernelGGL(stage_k_kernel, dim3(1), dim3(kStageChunk), 0, dev->stream, dev->noalloc_k,
                       first, n, c);
    DGK_LAUNCH_CHECK();
  }
  const int np = dgk::np_of(N);
  const double* D[3] = {Dr_dev, Ds_dev, Dt_dev};
  const void* moff[9];
  for (int i = 0; i < 9; ++i) moff[i] = m_off_dev;
  void* outs[3] = {dudx_dev, dudy_dev, dudz_dev};
  const void* ooff[3] = {o_off_dev, o_off_dev, o_off_dev};
  return dgk::phys_deriv_nodal_full(dev, N, npart, K_host, dev->noalloc_k, 8, D, np, 1, u_dev,
                                    u_off_dev, metrics_dev, moff, mtype, outs, ooff, otype, nullptr,
                                    nullptr, 8, true);
}

extern "C" int dgk_matmul(dgk_device dev, int rows, int cols, const double* A_colmajor_dev,
                          int npart, const int64_t* K_host, const double* in_dev,
                          const int64_t* in_off_dev, double* out_dev, const int64_t* out_off_dev,
                          int add) {
  dgk::clear_error();
  return dgk::matmul_typed(dev, npart, K_host, nullptr, 8, rows, cols, A_colmajor_dev, 1, rows,
                           DGK_F64, in_dev, in_off_dev, DGK_F64, out_dev, out_off_dev, DGK_F64,
                           DGK_F64, add, 8);
}
