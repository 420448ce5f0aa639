// Split Burgers stage for orders above 4 (where the fused stage's operators
// no longer fit a CU) and for contractions below FP64:
//   1. volume term  rhs = -div f(u)           deriv_affine_kernel<.., DIV>
//      ([Dr; Ds; Dt] f on matrix cores, metrics folded in FP64);
//   2. surface term and RK update, this file: per element the Lax-Friedrichs
//      flux jump g on the 4 Nfp face points (FP64, neighbour trace gathered
//      through the orientation-code table), LIFT g on matrix cores, then
//        out = a0 u0 + a1 (u + dt (rhs + LIFT g))   (mode 0)
//        out = rhs + LIFT g                          (mode 1).
// The flux is the same strong-form LF flux as the fused stage (burgers.hip):
//   g = Fscale (n.f(u-) - n.f*),  n.f* = (n.f(u-) + n.f(u+))/2
//       - alpha |n.(1,1,1)| (u+ - u-)/2,  alpha = max(|u-|, |u+|).
//
// Surface kernel layout: each wave takes 16-element tiles on its own (the
// MFMA N dimension); the LIFT fragments sit in LDS for the block. Lane
// (c, g) computes the B operand for element c directly: the flux values at
// face points k = 4 s + g (FP64/FP32) or 32 s + 8 g + e (bf16), so the flux
// never leaves registers. State stays FP64; the product runs in the plan's
// precision with FP32 accumulation below FP64.
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dgmesh_internal.h"

namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kF64 = 0, kF32 = 1, kBF3 = 2, kBF6 = 3;
constexpr bool is_bf(int prec) { return prec == kBF3 || prec == kBF6; }
constexpr int parts_of(int prec) { return prec == kBF3 ? 2 : (prec == kBF6 ? 3 : 1); }

template <int N, int PREC>
struct SCfg {
  static constexpr int NP = dgk::Tet<N>::Np, NFP = dgk::Tet<N>::Nfp, NF = 4 * NFP;
  static constexpr int MTL = (NP + 15) / 16;  // 16-row tiles of LIFT
  static constexpr int KSTEP = is_bf(PREC) ? 32 : 4;
  static constexpr int KSL = (NF + KSTEP - 1) / KSTEP;
  static constexpr int PARTS = parts_of(PREC);
  static constexpr size_t NFRAG = (size_t)MTL * KSL * 64;  // fragments per part
  static constexpr size_t FRAG_BYTES =
      PREC == kF64 ? NFRAG * 8 : (PREC == kF32 ? NFRAG * 4 : NFRAG * 16 * PARTS);
  static constexpr int W = 8;  // waves per block
};

struct SurfArgs {
  const dgk::Tile* tiles;
  int ntiles;
  const void* frag;           // LIFT fragments, [part][MTL][KSL][64]
  const uint8_t* face_nodes;  // [4][24][Nfp]
  const double* bgeo;         // [11][K]
  int64_t K;
  const int32_t* nbr;         // [4][K] value offset of each neighbour's node 0
  const int32_t* codes;       // [K] four 5-bit orientation codes
  const double* u;
  const double* u0;
  const double* rhs;
  double* out;
  double a0, a1;
  const double* dt;
  unsigned long long* maxabs;
  int mode;
};

template <int PREC>
using AccT = typename std::conditional<PREC == kF64, doublex4, floatx4>::type;

template <int N, int PREC>
__global__ void __launch_bounds__((64 * SCfg<N, PREC>::W)) surface_update_kernel(SurfArgs a) {
  using C = SCfg<N, PREC>;
  constexpr int NP = C::NP, NFP = C::NFP, NF = C::NF, MTL = C::MTL, KSL = C::KSL, W = C::W;
  constexpr int PARTS = C::PARTS;
  constexpr dgk::FmaskTable<N> FM{};
  __shared__ __attribute__((aligned(16))) unsigned char sA[C::FRAG_BYTES];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ int16_t sFM[NF];
  {
    const uint4* src = (const uint4*)a.frag;
    uint4* dst = (uint4*)sA;
    for (int i = threadIdx.x; i < (int)(C::FRAG_BYTES / 16); i += 64 * W) dst[i] = src[i];
    for (int i = threadIdx.x; i < 4 * 24 * NFP; i += 64 * W) sFN[i] = a.face_nodes[i];
    for (int i = threadIdx.x; i < NF; i += 64 * W) sFM[i] = (int16_t)FM.v[i / NFP][i % NFP];
  }
  __syncthreads();
  const double dt = a.mode == 0 ? *a.dt : 0.0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = lane & 15, g = lane >> 4;
  const int64_t K = a.K;
  double lmax = 0.0;
  for (int t = blockIdx.x * W + w; t < a.ntiles; t += gridDim.x * W) {
    const dgk::Tile tile = a.tiles[t];
    const bool valid = c < tile.count;
    const int64_t ge = tile.gbase + (valid ? c : 0);
    const int64_t vb = tile.vbase + (int64_t)(valid ? c : 0) * NP;
    const int packed = a.codes[ge];
    int32_t nb[4];
    double sn[4], fs[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      nb[f] = a.nbr[f * K + ge];
      sn[f] = a.bgeo[(3 + f) * K + ge];
      fs[f] = a.bgeo[(7 + f) * K + ge];
    }
    // LF flux jump at face point k of element c (FP64).
    auto flux = [&](int k) -> double {
      if (!valid || k >= NF) return 0.0;
      const int f = k / NFP, fp = k - f * NFP;
      const int code = (packed >> (5 * f)) & 31;
      const int32_t nbf = f == 0 ? nb[0] : f == 1 ? nb[1] : f == 2 ? nb[2] : nb[3];
      const double snf = f == 0 ? sn[0] : f == 1 ? sn[1] : f == 2 ? sn[2] : sn[3];
      const double fsf = f == 0 ? fs[0] : f == 1 ? fs[1] : f == 2 ? fs[2] : fs[3];
      const double uM = a.u[vb + sFM[k]];
      const double uP = a.u[nbf + sFN[(f * 24 + code) * NFP + fp]];
      const double fM = 0.5 * uM * uM, fP = 0.5 * uP * uP;
      const double al = fmax(fabs(uM), fabs(uP));
      return fsf * (0.5 * (fM - fP) * snf + 0.5 * al * fabs(snf) * (uP - uM));
    };
    AccT<PREC> acc[MTL];
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[tt][r] = 0;
    // Not unrolled: the flux gathers of every step would otherwise be
    // hoisted together and exhaust the registers.
#pragma unroll 1
    for (int s = 0; s < KSL; ++s) {
      if constexpr (PREC == kF64 || PREC == kF32) {
        const double b = flux(4 * s + g);
#pragma unroll
        for (int tt = 0; tt < MTL; ++tt) {
          const size_t q = ((size_t)tt * KSL + s) * 64 + lane;
          if constexpr (PREC == kF64)
            acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(((const double*)sA)[q], b, acc[tt], 0, 0, 0);
          else
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(((const float*)sA)[q], (float)b, acc[tt], 0,
                                                           0, 0);
        }
      } else {
        bf16x8 bp[PARTS];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const double x = flux(32 * s + 8 * g + e);
          const __bf16 h = (__bf16)(float)x;
          const double r1 = x - (double)(float)h;
          const __bf16 m = (__bf16)(float)r1;
          bp[0][e] = h;
          bp[1][e] = m;
          if constexpr (PARTS == 3) bp[2][e] = (__bf16)(float)(r1 - (double)(float)m);
        }
#pragma unroll
        for (int tt = 0; tt < MTL; ++tt) {
          const size_t q = ((size_t)tt * KSL + s) * 64 + lane;
          const bf16x8* F = (const bf16x8*)sA;
          if constexpr (PREC == kBF3) {
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[0], acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[1], acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[0], acc[tt], 0, 0, 0);
          } else {
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[2 * C::NFRAG + q], bp[0], acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[2], acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[1], acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[0], acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[1], acc[tt], 0, 0, 0);
            acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[0], acc[tt], 0, 0, 0);
          }
        }
      }
    }
    // Epilogue. FP64 D lanes hold rows g + 4r of a tile, FP32 D lanes rows 4g + r.
    if (valid) {
#pragma unroll
      for (int tt = 0; tt < MTL; ++tt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * tt + (PREC == kF64 ? g + 4 * r : 4 * g + r);
          if (i < NP) {
            const double l = a.rhs[vb + i] + (double)acc[tt][r];
            double o;
            if (a.mode == 0) {
              o = a.a1 * fma(dt, l, a.u[vb + i]);
              if (a.a0 != 0.0) o = fma(a.a0, a.u0[vb + i], o);
              lmax = fmax(lmax, fabs(o));
            } else {
              o = l;
            }
            a.out[vb + i] = o;
          }
        }
      }
    }
  }
  if (a.maxabs) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o, 64));
    if (lane == 0 && lmax > 0.0)
      atomicMax(a.maxabs, (unsigned long long)__double_as_longlong(lmax));
  }
}

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

// LIFT fragments in the MFMA A-operand lane order of the precision.
std::vector<unsigned char> lift_fragments(const std::vector<double>& LIFT, int N, int prec) {
  const int NP = dgk::np_of(N), NF = 4 * dgk::nfp_of(N);
  const int MTL = (NP + 15) / 16, KSTEP = is_bf(prec) ? 32 : 4;
  const int KSL = (NF + KSTEP - 1) / KSTEP, PARTS = parts_of(prec);
  const size_t nfrag = (size_t)MTL * KSL * 64;
  auto val = [&](int i, int k) { return (i < NP && k < NF) ? LIFT[(size_t)i * NF + k] : 0.0; };
  std::vector<unsigned char> buf;
  if (!is_bf(prec)) {
    buf.resize(nfrag * (prec == kF64 ? 8 : 4));
    for (int tt = 0; tt < MTL; ++tt)
      for (int s = 0; s < KSL; ++s)
        for (int l = 0; l < 64; ++l) {
          const double v = val(16 * tt + (l & 15), 4 * s + (l >> 4));
          const size_t q = ((size_t)tt * KSL + s) * 64 + l;
          if (prec == kF64) ((double*)buf.data())[q] = v;
          else ((float*)buf.data())[q] = (float)v;
        }
  } else {
    buf.resize(nfrag * 16 * PARTS);
    uint16_t* P0 = (uint16_t*)buf.data();
    for (int tt = 0; tt < MTL; ++tt)
      for (int s = 0; s < KSL; ++s)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e) {
            const double v = val(16 * tt + (l & 15), 32 * s + 8 * (l >> 4) + e);
            const size_t q = (((size_t)tt * KSL + s) * 64 + l) * 8 + e;
            const uint16_t h = bf16_bits((float)v);
            const double r1 = v - (double)bf16_to_float(h);
            const uint16_t m = bf16_bits((float)r1);
            P0[q] = h;
            P0[nfrag * 8 + q] = m;
            if (PARTS == 3) P0[2 * nfrag * 8 + q] = bf16_bits((float)(r1 - (double)bf16_to_float(m)));
          }
  }
  return buf;
}

template <int N, int PREC>
int surface_blocks(dgk_dgmesh m) {
  int per_cu = 0;
  DGK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, surface_update_kernel<N, PREC>,
                                                       64 * SCfg<N, PREC>::W, 0));
  if (per_cu < 1) per_cu = 1;
  m->split_blocks = m->dev->cu_count * per_cu;
  return DGK_OK;
}

template <int N, int PREC>
int surface_launch(dgk_dgmesh m, const SurfArgs& a) {
  const int W = SCfg<N, PREC>::W;
  int grid = (a.ntiles + W - 1) / W;
  if (grid > m->split_blocks) grid = m->split_blocks;
  if (grid < 1) return DGK_OK;
  hipLaunchKernelGGL((surface_update_kernel<N, PREC>), dim3(grid), dim3(64 * W), 0, m->dev->stream,
                     a);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <int N>
int surface_n(dgk_dgmesh m, const SurfArgs* a) {
  switch (m->prec) {
    case kF64: return a ? surface_launch<N, kF64>(m, *a) : surface_blocks<N, kF64>(m);
    case kF32: return a ? surface_launch<N, kF32>(m, *a) : surface_blocks<N, kF32>(m);
    case kBF3: return a ? surface_launch<N, kBF3>(m, *a) : surface_blocks<N, kBF3>(m);
    default: return a ? surface_launch<N, kBF6>(m, *a) : surface_blocks<N, kBF6>(m);
  }
}

// a == nullptr: size the persistent grid.
int surface(dgk_dgmesh m, const SurfArgs* a) {
  switch (m->N) {
    case 1: return surface_n<1>(m, a);
    case 2: return surface_n<2>(m, a);
    case 3: return surface_n<3>(m, a);
    case 4: return surface_n<4>(m, a);
    case 5: return surface_n<5>(m, a);
    case 6: return surface_n<6>(m, a);
  }
  dgk::set_error("split stage: N=%d outside 1..6", m->N);
  return DGK_ERR_UNSUPPORTED;
}

}  // namespace

int dgk::split_prepare(dgk_dgmesh m) {
  if (m->plan && m->rhs && m->lift_frag) return DGK_OK;
  dgk::split_release(m);
  int rc = dgk_deriv_plan_create(m->dev, m->N, m->Dr_h.data(), m->Ds_h.data(), m->Dt_h.data(),
                                 m->r_h.data(), m->s_h.data(), m->t_h.data(), m->prec, &m->plan);
  if (rc != DGK_OK) return rc;
  const size_t state_bytes = (size_t)(m->state_off.empty() ? 0 : m->state_off.back()) * 8 + 16;
  rc = dgk_malloc(m->dev, state_bytes, nullptr, (void**)&m->rhs);
  if (rc != DGK_OK) return rc;
  const std::vector<unsigned char> frag = lift_fragments(m->LIFT_h, m->N, m->prec);
  rc = dgk_malloc(m->dev, frag.size(), frag.data(), &m->lift_frag);
  if (rc != DGK_OK) return rc;
  return surface(m, nullptr);
}

void dgk::split_release(dgk_dgmesh m) {
  if (m->plan) dgk_deriv_plan_destroy(m->plan);
  if (m->rhs) dgk_free(m->dev, m->rhs);
  if (m->lift_frag) dgk_free(m->dev, m->lift_frag);
  m->plan = nullptr;
  m->rhs = nullptr;
  m->lift_frag = nullptr;
}

int dgk::split_volume(dgk_dgmesh m, const double* u) {
  int rc = split_prepare(m);
  if (rc != DGK_OK) return rc;
  return dgk::deriv_plan_divergence(m->plan, m->tiles64, m->ntiles64, u, m->geo, m->K, m->rhs);
}

int dgk::split_stage(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                     double a1, const double* dt, unsigned long long* maxabs, int mode) {
  int rc = split_volume(m, u);
  if (rc != DGK_OK) return rc;
  return split_surface(m, u, u0, out, a0, a1, dt, maxabs, mode);
}

int dgk::split_surface(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                       double a1, const double* dt, unsigned long long* maxabs, int mode) {
  SurfArgs a;
  a.tiles = m->tiles64;
  a.ntiles = m->ntiles64;
  a.frag = m->lift_frag;
  a.face_nodes = m->face_nodes;
  a.bgeo = m->bgeo;
  a.K = m->K;
  a.nbr = m->nbr;
  a.codes = m->codes;
  a.u = u;
  a.u0 = u0;
  a.rhs = m->rhs;
  a.out = out;
  a.a0 = a0;
  a.a1 = a1;
  a.dt = dt;
  a.maxabs = maxabs;
  a.mode = mode;
  return surface(m, &a);
}
