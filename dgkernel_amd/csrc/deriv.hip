// Physical derivative and MATMUL kernels for gfx950.
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
// for a group of elements (register blocking over GROUP elements reuses each
// D row), and the per-node metrics and outputs stream with unit stride.
// Nothing goes back to HBM except the three outputs.
#include "dgk_internal.h"

namespace {

constexpr int kMaxPartsByValue = 256;

struct PartTiles {
  int npart;
  int tiles_per_part_start[kMaxPartsByValue + 1];  // prefix sum of tiles
  int K[kMaxPartsByValue];                          // elements per partition
};

__device__ __forceinline__ void find_tile(const PartTiles& pt, int tile, int per_tile, int& part,
                                          int& first, int& count) {
  int lo = 0, hi = pt.npart;  // last p with start[p] <= tile
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pt.tiles_per_part_start[mid] <= tile) lo = mid; else hi = mid;
  }
  part = lo;
  first = (tile - pt.tiles_per_part_start[lo]) * per_tile;
  int rem = pt.K[lo] - first;
  count = rem < per_tile ? rem : per_tile;
}

int build_part_tiles(int npart, const int64_t* K, int per_tile, PartTiles& pt, int& ntiles) {
  DGK_REQUIRE(npart >= 1 && npart <= kMaxPartsByValue,
              "partition count %d outside 1..%d", npart, kMaxPartsByValue);
  pt.npart = npart;
  int t = 0;
  for (int p = 0; p < npart; ++p) {
    DGK_REQUIRE(K[p] >= 0 && K[p] < (1LL << 31), "K[%d] = %lld out of range", p, (long long)K[p]);
    pt.tiles_per_part_start[p] = t;
    pt.K[p] = (int)K[p];
    t += (int)((K[p] + per_tile - 1) / per_tile);
  }
  pt.tiles_per_part_start[npart] = t;
  ntiles = t;
  return DGK_OK;
}

template <typename T>
__device__ __forceinline__ T load_as(const void* p, int64_t i) {
  return static_cast<const T*>(p)[i];
}

struct Metrics9 {
  const void* m[9];  // rx ry rz sx sy sz tx ty tz
};

// One block: GROUP*EPG elements of one partition (EPG = elements per thread
// group); threads = NP * GROUPS_PAD, thread (node n, group g) handles node n
// of elements g, g+G, ..., g+(EPG-1)G.
template <int NP, int G, int EPG, typename MT, typename OT, typename OFF>
__global__ void __launch_bounds__(256)
deriv_nodal_kernel(PartTiles pt, const double* __restrict__ Dr, const double* __restrict__ Ds,
                   const double* __restrict__ Dt, const double* __restrict__ u,
                   const OFF* __restrict__ u_off, Metrics9 met, const OFF* __restrict__ m_off,
                   OT* __restrict__ ox, OT* __restrict__ oy, OT* __restrict__ oz,
                   const OFF* __restrict__ o_off) {
  constexpr int E = G * EPG;
  __shared__ double sD[3 * NP * NP];
  __shared__ double sU[E * NP];
  int part, first, count;
  find_tile(pt, blockIdx.x, E, part, first, count);
  const int tid = threadIdx.x;
  for (int i = tid; i < 3 * NP * NP; i += blockDim.x) {
    int m = i / (NP * NP), r = i - m * NP * NP;
    sD[i] = (m == 0 ? Dr : (m == 1 ? Ds : Dt))[r];
  }
  const int64_t ub = (int64_t)u_off[part] + (int64_t)first * NP;
  const int nval = count * NP;
  for (int i = tid; i < nval; i += blockDim.x) sU[i] = u[ub + i];
  __syncthreads();
  const int n = tid % NP, g = tid / NP;
  if (g >= G) return;
  double ar[EPG], as[EPG], at[EPG];
#pragma unroll
  for (int q = 0; q < EPG; ++q) ar[q] = as[q] = at[q] = 0.0;
  const double* dr = sD + n * NP;
  const double* ds = sD + NP * NP + n * NP;
  const double* dt = sD + 2 * NP * NP + n * NP;
#pragma unroll 5
  for (int j = 0; j < NP; ++j) {
    double a = dr[j], b = ds[j], c = dt[j];
#pragma unroll
    for (int q = 0; q < EPG; ++q) {
      double uj = sU[(g + q * G) * NP + j];
      ar[q] = fma(a, uj, ar[q]);
      as[q] = fma(b, uj, as[q]);
      at[q] = fma(c, uj, at[q]);
    }
  }
  const int64_t mb = (int64_t)m_off[part] + (int64_t)first * NP;
  const int64_t obase = (int64_t)o_off[part] + (int64_t)first * NP;
#pragma unroll
  for (int q = 0; q < EPG; ++q) {
    int e = g + q * G;
    if (e >= count) continue;
    int64_t mi = mb + e * NP + n, oi = obase + e * NP + n;
    double rx = (double)load_as<MT>(met.m[0], mi), ry = (double)load_as<MT>(met.m[1], mi),
           rz = (double)load_as<MT>(met.m[2], mi), sx = (double)load_as<MT>(met.m[3], mi),
           sy = (double)load_as<MT>(met.m[4], mi), sz = (double)load_as<MT>(met.m[5], mi),
           tx = (double)load_as<MT>(met.m[6], mi), ty = (double)load_as<MT>(met.m[7], mi),
           tz = (double)load_as<MT>(met.m[8], mi);
    ox[oi] = (OT)(rx * ar[q] + sx * as[q] + tx * at[q]);
    oy[oi] = (OT)(ry * ar[q] + sy * as[q] + ty * at[q]);
    oz[oi] = (OT)(rz * ar[q] + sz * as[q] + tz * at[q]);
  }
}

// Large NP (P >= 5): D rows are read from global memory (L1/L2 resident),
// u is staged in LDS.
template <int NP, int G, typename MT, typename OT, typename OFF>
__global__ void __launch_bounds__(256)
deriv_nodal_big_kernel(PartTiles pt, const double* __restrict__ Dr, const double* __restrict__ Ds,
                       const double* __restrict__ Dt, const double* __restrict__ u,
                       const OFF* __restrict__ u_off, Metrics9 met, const OFF* __restrict__ m_off,
                       OT* __restrict__ ox, OT* __restrict__ oy, OT* __restrict__ oz,
                       const OFF* __restrict__ o_off) {
  constexpr int E = G;
  __shared__ double sU[E * NP];
  int part, first, count;
  find_tile(pt, blockIdx.x, E, part, first, count);
  const int tid = threadIdx.x;
  const int64_t ub = (int64_t)u_off[part] + (int64_t)first * NP;
  for (int i = tid; i < count * NP; i += blockDim.x) sU[i] = u[ub + i];
  __syncthreads();
  for (int w = tid; w < count * NP; w += blockDim.x) {
    int e = w / NP, n = w - e * NP;
    double ar = 0, as = 0, at = 0;
    for (int j = 0; j < NP; ++j) {
      double uj = sU[e * NP + j];
      ar = fma(Dr[n * NP + j], uj, ar);
      as = fma(Ds[n * NP + j], uj, as);
      at = fma(Dt[n * NP + j], uj, at);
    }
    int64_t mi = (int64_t)m_off[part] + (int64_t)first * NP + w;
    int64_t oi = (int64_t)o_off[part] + (int64_t)first * NP + w;
    double rx = (double)load_as<MT>(met.m[0], mi), ry = (double)load_as<MT>(met.m[1], mi),
           rz = (double)load_as<MT>(met.m[2], mi), sx = (double)load_as<MT>(met.m[3], mi),
           sy = (double)load_as<MT>(met.m[4], mi), sz = (double)load_as<MT>(met.m[5], mi),
           tx = (double)load_as<MT>(met.m[6], mi), ty = (double)load_as<MT>(met.m[7], mi),
           tz = (double)load_as<MT>(met.m[8], mi);
    ox[oi] = (OT)(rx * ar + sx * as + tx * at);
    oy[oi] = (OT)(ry * ar + sy * as + ty * at);
    oz[oi] = (OT)(rz * ar + sz * as + tz * at);
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
int launch_nodal(dgk_device dev, int npart, const int64_t* K, const double* Dr, const double* Ds,
                 const double* Dt, const double* u, const OFF* u_off, const void* const* metrics,
                 const OFF* m_off, void* dx, void* dy, void* dz, const OFF* o_off) {
  constexpr int NP = (N + 1) * (N + 2) * (N + 3) / 6;
  Metrics9 met;
  for (int i = 0; i < 9; ++i) met.m[i] = metrics[i];
  PartTiles pt;
  int ntiles = 0;
  if constexpr (N <= 4) {
    constexpr int G = NodalShape<N>::G, EPG = NodalShape<N>::EPG;
    int rc = build_part_tiles(npart, K, G * EPG, pt, ntiles);
    if (rc) return rc;
    if (ntiles == 0) return DGK_OK;
    int threads = ((NP * G + 63) / 64) * 64;
    hipLaunchKernelGGL((deriv_nodal_kernel<NP, G, EPG, MT, OT, OFF>), dim3(ntiles), dim3(threads), 0,
                       dev->stream, pt, Dr, Ds, Dt, u, u_off, met, m_off, (OT*)dx, (OT*)dy, (OT*)dz,
                       o_off);
  } else {
    constexpr int G = 16;
    int rc = build_part_tiles(npart, K, G, pt, ntiles);
    if (rc) return rc;
    if (ntiles == 0) return DGK_OK;
    hipLaunchKernelGGL((deriv_nodal_big_kernel<NP, G, MT, OT, OFF>), dim3(ntiles), dim3(256), 0,
                       dev->stream, pt, Dr, Ds, Dt, u, u_off, met, m_off, (OT*)dx, (OT*)dy, (OT*)dz,
                       o_off);
  }
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <typename MT, typename OT, typename OFF>
int dispatch_nodal_N(int N, dgk_device dev, int npart, const int64_t* K, const double* Dr,
                     const double* Ds, const double* Dt, const double* u, const OFF* u_off,
                     const void* const* metrics, const OFF* m_off, void* dx, void* dy, void* dz,
                     const OFF* o_off) {
  switch (N) {
#define DGK_CASE(n)                                                                             \
  case n:                                                                                       \
    return launch_nodal<n, MT, OT, OFF>(dev, npart, K, Dr, Ds, Dt, u, u_off, metrics, m_off, dx, \
                                        dy, dz, o_off);
    DGK_CASE(1) DGK_CASE(2) DGK_CASE(3) DGK_CASE(4) DGK_CASE(5) DGK_CASE(6)
#undef DGK_CASE
    default:
      dgk::set_error("physical derivative: order N=%d not built (1..6)", N);
      return DGK_ERR_UNSUPPORTED;
  }
}

// Generic MATMUL over partitions: thread per (element, row).
template <typename OFF>
__global__ void __launch_bounds__(256)
matmul_kernel(PartTiles pt, int per_tile, int rows, int cols, const double* __restrict__ A,
              const double* __restrict__ in, const OFF* __restrict__ in_off,
              double* __restrict__ out, const OFF* __restrict__ out_off, int add) {
  int part, first, count;
  find_tile(pt, blockIdx.x, per_tile, part, first, count);
  int64_t ib = (int64_t)in_off[part] + (int64_t)first * cols;
  int64_t ob = (int64_t)out_off[part] + (int64_t)first * rows;
  for (int w = threadIdx.x; w < count * rows; w += blockDim.x) {
    int e = w / rows, i = w - e * rows;
    double s = 0.0;
    for (int j = 0; j < cols; ++j) s = fma(A[(int64_t)j * rows + i], in[ib + (int64_t)e * cols + j], s);
    if (add) out[ob + w] += s; else out[ob + w] = s;
  }
}

// OUT[e*rows + i] = s[e*rows + i] * sum_j A(i,j) IN[e*cols + j], A column-major
// (the runner's device matrix layout); s may be NULL (plain MATMUL).
template <typename OFF, typename ST>
__global__ void __launch_bounds__(256)
scaled_matmul_kernel(PartTiles pt, int per_tile, int rows, int cols, const double* __restrict__ A,
                     const double* __restrict__ in, const OFF* __restrict__ in_off,
                     const ST* __restrict__ sc, const OFF* __restrict__ sc_off,
                     double* __restrict__ out, const OFF* __restrict__ out_off) {
  int part, first, count;
  find_tile(pt, blockIdx.x, per_tile, part, first, count);
  int64_t ib = (int64_t)in_off[part] + (int64_t)first * cols;
  int64_t ob = (int64_t)out_off[part] + (int64_t)first * rows;
  int64_t sb = sc ? (int64_t)sc_off[part] + (int64_t)first * rows : 0;
  for (int w = threadIdx.x; w < count * rows; w += blockDim.x) {
    int e = w / rows, i = w - e * rows;
    double s = 0.0;
    for (int j = 0; j < cols; ++j) s = fma(A[(int64_t)j * rows + i], in[ib + (int64_t)e * cols + j], s);
    if (sc) s *= (double)sc[sb + w];
    out[ob + w] = s;
  }
}

__global__ void transpose_kernel(int rows, int cols, const double* __restrict__ in_colmajor,
                                 double* __restrict__ out_rowmajor) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= rows * cols) return;
  int i = w / cols, j = w - i * cols;
  out_rowmajor[w] = in_colmajor[(int64_t)j * rows + i];
}

}  // namespace

namespace dgk {

int phys_deriv_nodal_any(dgk_device dev, int N, int npart, const int64_t* K, const double* Dr,
                         const double* Ds, const double* Dt, const double* u, const void* u_off,
                         const void* const* metrics, int mtype, const void* m_off, void* dx,
                         void* dy, void* dz, int otype, const void* o_off, int off_size) {
  DGK_REQUIRE(dev && K && Dr && Ds && Dt && u && u_off && metrics && m_off && dx && dy && dz &&
                  o_off,
              "physical derivative: NULL argument");
  for (int i = 0; i < 9; ++i) DGK_REQUIRE(metrics[i], "physical derivative: metric %d is NULL", i);
  DGK_REQUIRE(mtype == DGK_F64 || mtype == DGK_F32, "physical derivative: bad metric type");
  DGK_REQUIRE(otype == DGK_F64 || otype == DGK_F32, "physical derivative: bad output type");
#define DGK_GO(MT, OT, OFF)                                                                   \
  return dispatch_nodal_N<MT, OT, OFF>(N, dev, npart, K, Dr, Ds, Dt, u, (const OFF*)u_off,    \
                                       metrics, (const OFF*)m_off, dx, dy, dz, (const OFF*)o_off)
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

int matmul_any(dgk_device dev, int rows, int cols, const double* A, int npart, const int64_t* K,
               const double* in, const void* in_off, double* out, const void* out_off, int add,
               int off_size) {
  DGK_REQUIRE(dev && A && K && in && in_off && out && out_off, "matmul: NULL argument");
  DGK_REQUIRE(rows > 0 && cols > 0, "matmul: bad shape %dx%d", rows, cols);
  int per_tile = rows >= 256 ? 1 : 256 / rows;
  PartTiles pt;
  int ntiles = 0;
  int rc = build_part_tiles(npart, K, per_tile, pt, ntiles);
  if (rc) return rc;
  if (!ntiles) return DGK_OK;
  if (off_size == 8)
    hipLaunchKernelGGL(matmul_kernel<int64_t>, dim3(ntiles), dim3(256), 0, dev->stream, pt, per_tile,
                       rows, cols, A, in, (const int64_t*)in_off, out, (const int64_t*)out_off, add);
  else
    hipLaunchKernelGGL(matmul_kernel<int32_t>, dim3(ntiles), dim3(256), 0, dev->stream, pt, per_tile,
                       rows, cols, A, in, (const int32_t*)in_off, out, (const int32_t*)out_off, add);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

int scaled_matmul_any(dgk_device dev, int rows, int cols, const double* A_colmajor, int npart,
                      const int64_t* K, const double* in, const void* in_off, const void* sc,
                      int sc_type, const void* sc_off, double* out, const void* out_off,
                      int off_size) {
  DGK_REQUIRE(dev && A_colmajor && K && in && in_off && out && out_off, "matmul: NULL argument");
  int per_tile = rows >= 256 ? 1 : 256 / rows;
  PartTiles pt;
  int ntiles = 0;
  int rc = build_part_tiles(npart, K, per_tile, pt, ntiles);
  if (rc) return rc;
  if (!ntiles) return DGK_OK;
#define DGK_GO(OFF, ST)                                                                          \
  hipLaunchKernelGGL((scaled_matmul_kernel<OFF, ST>), dim3(ntiles), dim3(256), 0, dev->stream, pt, \
                     per_tile, rows, cols, A_colmajor, in, (const OFF*)in_off, (const ST*)sc,      \
                     (const OFF*)sc_off, out, (const OFF*)out_off)
  if (off_size == 8) {
    if (sc_type == DGK_F32) DGK_GO(int64_t, float); else DGK_GO(int64_t, double);
  } else {
    if (sc_type == DGK_F32) DGK_GO(int32_t, float); else DGK_GO(int32_t, double);
  }
#undef DGK_GO
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

int transpose_colmajor(dgk_device dev, int rows, int cols, const double* in, double* out) {
  int n = rows * cols;
  hipLaunchKernelGGL(transpose_kernel, dim3((n + 255) / 256), dim3(256), 0, dev->stream, rows, cols,
                     in, out);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
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
  return dgk::phys_deriv_nodal_any(dev, N, npart, K_host, Dr_dev, Ds_dev, Dt_dev, u_dev, u_off_dev,
                                   metrics_dev, mtype, m_off_dev, dudx_dev, dudy_dev, dudz_dev,
                                   otype, o_off_dev, 8);
}

extern "C" int dgk_matmul(dgk_device dev, int rows, int cols, const double* A_colmajor_dev,
                          int npart, const int64_t* K_host, const double* in_dev,
                          const int64_t* in_off_dev, double* out_dev, const int64_t* out_off_dev,
                          int add) {
  dgk::clear_error();
  return dgk::matmul_any(dev, rows, cols, A_colmajor_dev, npart, K_host, in_dev, in_off_dev, out_dev,
                         out_off_dev, add, 8);
}
