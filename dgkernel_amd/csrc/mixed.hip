// Physical derivative on affine tetrahedra with matrix cores, in FP64, FP32
// and split BF16x3.
//
// For an affine element the metrics are constant, so
//   du/dx = rx (Dr u) + sx (Ds u) + tx (Dt u)   (likewise y, z),
// the reference's calcPhysicalDerivative
// (tetnudg/tet_nudg_mesh_physical_derivative_test.go:344-367) with one
// metric value per element instead of one per node.
//
// Per wave, 16 elements are the N dimension of the MFMA. The stacked operator
// [Dr; Ds; Dt] (blocks at row offsets 0, SO, 2 SO) is the A operand, u the B
// operand, read straight from HBM in the lane order the MFMA wants. SO is
// chosen so that the lane holding row i of Dr u also holds rows i of Ds u
// and Dt u (FP64 result rows are interleaved by 4, FP32 rows are blocked by
// 4 within 16-row tiles), so the chain rule happens in registers.
//
//   F64:    v_mfma_f64_16x16x4_f64, operator fragments streamed from L2.
//   F32:    v_mfma_f32_16x16x4_f32 on FP32-rounded operator and u.
//   BF16X3: v_mfma_f32_16x16x32_bf16 on hi/lo splits, hi*hi + hi*lo + lo*hi,
//           FP32 accumulate (Henry, Tang & Heinecke, ARITH 2019).
// F32 and BF16X3 keep the operator fragments in LDS for the block's life.
#include <cmath>
#include <cstring>
#include <vector>

#include "dgk_internal.h"

namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kF64 = 0, kF32 = 1, kBF = 2;

template <int N, int PREC>
struct DShape {
  static constexpr int NP = (N + 1) * (N + 2) * (N + 3) / 6;
  static constexpr int SO = PREC == kF64 ? (NP + 3) / 4 * 4 : (NP + 15) / 16 * 16;
  static constexpr int MT = (2 * SO + NP + 15) / 16;
  static constexpr int KSTEP = PREC == kBF ? 32 : 4;
  static constexpr int KS = (NP + KSTEP - 1) / KSTEP;
  static constexpr int NFRAG = MT * KS * 64;  // fragments (per part for BF16)
};

int shape_query(int N, int prec, int& NP, int& SO, int& MT, int& KSTEP, int& KS) {
  NP = dgk::np_of(N);
  SO = prec == kF64 ? (NP + 3) / 4 * 4 : (NP + 15) / 16 * 16;
  MT = (2 * SO + NP + 15) / 16;
  KSTEP = prec == kBF ? 32 : 4;
  KS = (NP + KSTEP - 1) / KSTEP;
  return DGK_OK;
}

// Node groups: the rows are processed in NG passes over K, each pass
// accumulating only the 16-row tiles that hold its nodes' Dr u, Ds u, Dt u,
// which bounds the live accumulators (the B operand is re-read from cache).
template <int N, int PREC, int NG>
struct Groups {
  using SH = DShape<N, PREC>;
  static constexpr int QN = PREC == kF64 ? SH::SO / 4 : SH::SO / 16;  // node slots per lane
  static constexpr int QG = (QN + NG - 1) / NG;
  // Does group G use 16-row tile tt?
  static constexpr bool needs(int G, int tt) {
    for (int q = G * QG; q < (G + 1) * QG && q < QN; ++q)
      for (int b = 0; b < 3; ++b) {
        const int row0 = PREC == kF64 ? b * SH::SO + 4 * q : b * SH::SO + 16 * q;
        if (row0 / 16 == tt) return true;
      }
    return false;
  }
};

template <int N, int PREC, int NG, int G>
__device__ __forceinline__ void deriv_group(const unsigned char* sA, const void* frag, const double* ue,
                                            bool valid, int lane, int g, const double* m64,
                                            void* ox, void* oy, void* oz, int64_t ge) {
  using SH = DShape<N, PREC>;
  using GR = Groups<N, PREC, NG>;
  constexpr int NP = SH::NP, SO = SH::SO, MT = SH::MT, KS = SH::KS, NFRAG = SH::NFRAG;
  constexpr int QN = GR::QN, QG = GR::QG;
  if constexpr (PREC == kF64) {
    const double* A = (const double*)frag;
    doublex4 acc[MT];
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) acc[tt] = (doublex4){0.0, 0.0, 0.0, 0.0};
#pragma unroll 3
    for (int s = 0; s < KS; ++s) {
      const int k = 4 * s + g;
      const double b = (valid && k < NP) ? ue[k] : 0.0;
#pragma unroll
      for (int tt = 0; tt < MT; ++tt)
        if (GR::needs(G, tt))
          acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(tt * KS + s) * 64 + lane], b, acc[tt],
                                                          0, 0, 0);
    }
    if (valid) {
      double* px = (double*)ox + ge * NP;
      double* py = (double*)oy + ge * NP;
      double* pz = (double*)oz + ge * NP;
#pragma unroll
      for (int q = G * QG; q < (G + 1) * QG && q < QN; ++q) {
        const int i = g + 4 * q;
        const double ur = acc[(4 * q) / 16][((4 * q) % 16) / 4];
        const double us = acc[(SO + 4 * q) / 16][((SO + 4 * q) % 16) / 4];
        const double ut = acc[(2 * SO + 4 * q) / 16][((2 * SO + 4 * q) % 16) / 4];
        if (i < NP) {
          px[i] = m64[0] * ur + m64[3] * us + m64[6] * ut;
          py[i] = m64[1] * ur + m64[4] * us + m64[7] * ut;
          pz[i] = m64[2] * ur + m64[5] * us + m64[8] * ut;
        }
      }
    }
  } else {
    floatx4 acc[MT];
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) acc[tt] = (floatx4){0.f, 0.f, 0.f, 0.f};
    if constexpr (PREC == kF32) {
      const float* A = (const float*)sA;
#pragma unroll 3
      for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        const float b = (valid && k < NP) ? (float)ue[k] : 0.f;
#pragma unroll
        for (int tt = 0; tt < MT; ++tt)
          if (GR::needs(G, tt))
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[(tt * KS + s) * 64 + lane], b, acc[tt],
                                                            0, 0, 0);
      }
    } else {
      const bf16x8* Ahi = (const bf16x8*)sA;
      const bf16x8* Alo = Ahi + NFRAG;
#pragma unroll 1
      for (int s = 0; s < KS; ++s) {
        bf16x8 bhi, blo;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = 32 * s + 8 * g + e;
          const double x = (valid && k < NP) ? ue[k] : 0.0;
          const __bf16 h = (__bf16)(float)x;
          bhi[e] = h;
          blo[e] = (__bf16)(float)(x - (double)(float)h);
        }
#pragma unroll
        for (int tt = 0; tt < MT; ++tt)
          if (GR::needs(G, tt)) {
            const bf16x8 ah = Ahi[(tt * KS + s) * 64 + lane];
            const bf16x8 al = Alo[(tt * KS + s) * 64 + lane];
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bhi, acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, blo, acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bhi, acc[tt], 0, 0, 0);
          }
      }
    }
    if (valid) {
      float m[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) m[q] = (float)m64[q];
      float* px = (float*)ox + ge * NP;
      float* py = (float*)oy + ge * NP;
      float* pz = (float*)oz + ge * NP;
#pragma unroll
      for (int tq = G * QG; tq < (G + 1) * QG && tq < QN; ++tq) {
        floatx4 vx, vy, vz;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float ur = acc[tq][r], us = acc[SO / 16 + tq][r], ut = acc[2 * SO / 16 + tq][r];
          vx[r] = m[0] * ur + m[3] * us + m[6] * ut;
          vy[r] = m[1] * ur + m[4] * us + m[7] * ut;
          vz[r] = m[2] * ur + m[5] * us + m[8] * ut;
        }
        const int i0 = 16 * tq + 4 * g;
        if constexpr (NP % 4 == 0) {
          if (i0 < NP) {
            *(floatx4*)(px + i0) = vx;
            *(floatx4*)(py + i0) = vy;
            *(floatx4*)(pz + i0) = vz;
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (i0 + r < NP) {
              px[i0 + r] = vx[r];
              py[i0 + r] = vy[r];
              pz[i0 + r] = vz[r];
            }
        }
      }
    }
  }
}

template <int N, int PREC, int NG, int G>
__device__ __forceinline__ void deriv_groups(const unsigned char* sA, const void* frag,
                                             const double* ue, bool valid, int lane, int g,
                                             const double* m64, void* ox, void* oy, void* oz,
                                             int64_t ge) {
  if constexpr (G < NG) {
    deriv_group<N, PREC, NG, G>(sA, frag, ue, valid, lane, g, m64, ox, oy, oz, ge);
    deriv_groups<N, PREC, NG, G + 1>(sA, frag, ue, valid, lane, g, m64, ox, oy, oz, ge);
  }
}

template <int N, int PREC, int WPB, int NG>
__global__ void __launch_bounds__(64 * WPB)
deriv_affine_kernel(int64_t K, const void* __restrict__ frag, const double* __restrict__ u,
                    const double* __restrict__ geo, int64_t gs, void* __restrict__ ox,
                    void* __restrict__ oy, void* __restrict__ oz) {
  using SH = DShape<N, PREC>;
  constexpr int NP = SH::NP, NFRAG = SH::NFRAG;
  constexpr int NT = 64 * WPB;
  // LDS copy of the fragments (F32: floats; BF16: hi then lo, 8 bf16 each).
  constexpr int LDS_BYTES = PREC == kF32 ? NFRAG * 4 : (PREC == kBF ? NFRAG * 32 : 16);
  __shared__ __attribute__((aligned(16))) unsigned char sA[LDS_BYTES];
  if constexpr (PREC != kF64) {
    const uint4* src = (const uint4*)frag;
    uint4* dst = (uint4*)sA;
    for (int i = threadIdx.x; i < LDS_BYTES / 16; i += NT) dst[i] = src[i];
    __syncthreads();
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int64_t ntiles = (K + 15) / 16;
  const int64_t NW = (int64_t)gridDim.x * WPB;
  for (int64_t t = (int64_t)blockIdx.x * WPB + w; t < ntiles; t += NW) {
    const int64_t e0 = t * 16;
    const bool valid = e0 + c < K;
    const int64_t ge = valid ? e0 + c : 0;
    double m64[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) m64[q] = valid ? geo[q * gs + ge] : 0.0;
    deriv_groups<N, PREC, NG, 0>(sA, frag, u + ge * NP, valid, lane, g, m64, ox, oy, oz, ge);
  }
}

}  // namespace

struct dgk_deriv_plan_s {
  dgk_device dev = nullptr;
  int N = 0, prec = 0;
  void* frag = nullptr;
};

namespace {

// bf16 round-to-nearest-even of a float, as raw bits (host side).
uint16_t bf16_bits(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  if ((x & 0x7f800000u) == 0x7f800000u) return (uint16_t)(x >> 16);
  uint32_t lsb = (x >> 16) & 1u;
  x += 0x7fffu + lsb;
  return (uint16_t)(x >> 16);
}

float bf16_to_float(uint16_t b) {
  uint32_t x = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

template <int N, int PREC, int WPB, int NG>
int launch_affine_t(dgk_deriv_plan pl, int64_t K, const double* u, const double* geo, int64_t gs,
                    void* ox, void* oy, void* oz) {
  int64_t ntiles = (K + 15) / 16;
  int64_t blocks = (ntiles + WPB - 1) / WPB;
  int bpc = PREC == kF64 ? 4 : 1;
  int64_t cap = (int64_t)pl->dev->cu_count * bpc;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) return DGK_OK;
  hipLaunchKernelGGL((deriv_affine_kernel<N, PREC, WPB, NG>), dim3((unsigned)blocks),
                     dim3(64 * WPB), 0, pl->dev->stream, K, pl->frag, u, geo, gs, ox, oy, oz);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <int N>
int launch_affine_n(dgk_deriv_plan pl, int64_t K, const double* u, const double* geo, int64_t gs,
                    void* ox, void* oy, void* oz) {
  constexpr int NG = N >= 5 ? 2 : 1;
  switch (pl->prec) {
    case kF64: return launch_affine_t<N, kF64, 4, NG>(pl, K, u, geo, gs, ox, oy, oz);
    case kF32: return launch_affine_t<N, kF32, 8, 1>(pl, K, u, geo, gs, ox, oy, oz);
    default: return launch_affine_t<N, kBF, 8, NG>(pl, K, u, geo, gs, ox, oy, oz);
  }
}

}  // namespace

extern "C" int dgk_deriv_plan_create(dgk_device dev, int N, const double* Dr, const double* Ds,
                                     const double* Dt, int prec, dgk_deriv_plan* out) {
  dgk::clear_error();
  DGK_REQUIRE(dev && Dr && Ds && Dt && out, "dgk_deriv_plan_create: NULL argument");
  DGK_REQUIRE(N >= 1 && N <= 6, "dgk_deriv_plan_create: N=%d outside 1..6", N);
  DGK_REQUIRE(prec == kF64 || prec == kF32 || prec == kBF, "dgk_deriv_plan_create: prec %d", prec);
  *out = nullptr;
  int NP, SO, MT, KSTEP, KS;
  shape_query(N, prec, NP, SO, MT, KSTEP, KS);
  const double* D[3] = {Dr, Ds, Dt};
  auto val = [&](int row, int k) -> double {
    int blk = row / SO, i = row % SO;
    if (blk >= 3 || i >= NP || k >= NP) return 0.0;
    return D[blk][i * NP + k];
  };
  const size_t nfrag = (size_t)MT * KS * 64;
  std::vector<unsigned char> buf;
  if (prec == kF64 || prec == kF32) {
    buf.resize(nfrag * (prec == kF64 ? 8 : 4));
    for (int tt = 0; tt < MT; ++tt)
      for (int s = 0; s < KS; ++s)
        for (int l = 0; l < 64; ++l) {
          double v = val(16 * tt + (l & 15), 4 * s + (l >> 4));
          size_t q = ((size_t)tt * KS + s) * 64 + l;
          if (prec == kF64) ((double*)buf.data())[q] = v;
          else ((float*)buf.data())[q] = (float)v;
        }
  } else {
    buf.resize(nfrag * 32);
    uint16_t* hi = (uint16_t*)buf.data();
    uint16_t* lo = hi + nfrag * 8;
    for (int tt = 0; tt < MT; ++tt)
      for (int s = 0; s < KS; ++s)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e) {
            double v = val(16 * tt + (l & 15), 32 * s + 8 * (l >> 4) + e);
            uint16_t h = bf16_bits((float)v);
            size_t q = (((size_t)tt * KS + s) * 64 + l) * 8 + e;
            hi[q] = h;
            lo[q] = bf16_bits((float)(v - (double)bf16_to_float(h)));
          }
  }
  dgk_deriv_plan p = new dgk_deriv_plan_s();
  p->dev = dev;
  p->N = N;
  p->prec = prec;
  int rc = dgk_malloc(dev, buf.size(), buf.data(), &p->frag);
  if (rc != DGK_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return DGK_OK;
}

extern "C" int dgk_deriv_plan_destroy(dgk_deriv_plan p) {
  if (!p) return DGK_OK;
  dgk_free(p->dev, p->frag);
  delete p;
  return DGK_OK;
}

extern "C" int dgk_phys_deriv_affine(dgk_deriv_plan p, int64_t K, const double* u_dev,
                                     const double* geo_dev, int64_t geo_stride, void* dudx_dev,
                                     void* dudy_dev, void* dudz_dev) {
  dgk::clear_error();
  DGK_REQUIRE(p && u_dev && geo_dev && dudx_dev && dudy_dev && dudz_dev,
              "dgk_phys_deriv_affine: NULL argument");
  DGK_REQUIRE(K >= 0 && geo_stride >= K, "dgk_phys_deriv_affine: K=%lld stride=%lld", (long long)K,
              (long long)geo_stride);
  if (K == 0) return DGK_OK;
  switch (p->N) {
#define DGK_CASE(n) \
  case n: return launch_affine_n<n>(p, K, u_dev, geo_dev, geo_stride, dudx_dev, dudy_dev, dudz_dev);
    DGK_CASE(1) DGK_CASE(2) DGK_CASE(3) DGK_CASE(4) DGK_CASE(5) DGK_CASE(6)
#undef DGK_CASE
  }
  dgk::set_error("dgk_phys_deriv_affine: N=%d", p->N);
  return DGK_ERR_UNSUPPORTED;
}
