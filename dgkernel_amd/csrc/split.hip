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
// (c, g) owns face g of element c: it gathers that face's own and
// neighbour traces (all loads in flight at once) and computes the LF jumps
// of the face's points directly as its B operand. The LIFT columns are
// taken face by face: MFMA step s multiplies point s of every face
// (FP64/FP32, k = 4) or points 8s..8s+7 (bf16, k = 32), so no value moves
// between lanes. State stays FP64; the product runs in the plan's precision
// with FP32 accumulation below FP64.
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
  // MFMA steps over the face points: one point per face (k = 4) or eight
  // points per face (bf16, k = 32, padded past Nfp with zero columns).
  static constexpr int KSL = is_bf(PREC) ? (NFP + 7) / 8 : NFP;
  static constexpr int NPT = is_bf(PREC) ? 8 * KSL : NFP;  // padded points per face
  static constexpr int PARTS = parts_of(PREC);
  static constexpr size_t NFRAG = (size_t)MTL * KSL * 64;  // fragments per part
  static constexpr size_t FRAG_BYTES =
      PREC == kF64 ? NFRAG * 8 : (PREC == kF32 ? NFRAG * 4 : NFRAG * 16 * PARTS);
  // Waves per block: three per SIMD to cover the latency of the trace
  // gathers. More waves only add contention for the L1's outstanding misses
  // (four per SIMD made the BF16x3 kernel ~8 % slower at P=6), fewer leave
  // latency uncovered. BF16x3 runs them as two blocks of six (its fragments
  // leave room in LDS for two), which measured 0.6 % faster per step at P=6.
  static constexpr int W = PREC == kBF3 ? 6 : 12;
  // Per-wave LDS buffer: a quarter tile (4 elements, row stride NPH) for the
  // own traces, reused as the 16 x 17 update buffer.
  static constexpr int NPH = NP | 1;
  static constexpr int BUF = 4 * NPH > 16 * 17 ? 4 * NPH : 16 * 17;
  static constexpr int NVH = (4 * NP + 63) / 64;  // quarter-tile values per lane
  // Own traces staged through LDS (unit-stride loads) or gathered per lane.
  static constexpr bool STAGE = PREC == kF32;
};

struct SurfArgs {
  const dgk::Tile* tiles;
  int ntiles;
  const void* frag;           // LIFT fragments, [part][MTL][KSL][64]
  const uint8_t* face_nodes;  // [4][24][Nfp]
  const double* bgeo;         // [19][K]; rows 0..2 the folded metrics
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

// Row of LIFT g held in accumulator slot r of lane group g: FP64 D lanes hold
// rows g + 4r of a 16-row tile, FP32 D lanes rows 4g + r.
template <int PREC>
__device__ inline int acc_row(int g, int r) {
  return PREC == kF64 ? g + 4 * r : 4 * g + r;
}

// LDS writes of one lane made visible to the other lanes of the same wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per tile, each wave: lane (c, g) gathers the neighbour trace of face g
// (all loads in flight together), reads its own trace from the tile staged
// in LDS a quarter at a time with unit-stride loads, and forms the LF jumps as its
// B operand; then, per 16-row tile of LIFT: the rhs, u, u0 values this lane
// will update are loaded, the product runs on matrix cores, and the result
// goes through the wave's 16 x 17 LDS buffer so that each lane updates and
// stores rows of 16 consecutive values.
template <int N, int PREC>
__global__ void __launch_bounds__((64 * SCfg<N, PREC>::W)) surface_update_kernel(SurfArgs a) {
  using C = SCfg<N, PREC>;
  constexpr int NP = C::NP, NFP = C::NFP, NF = C::NF, MTL = C::MTL, KSL = C::KSL, W = C::W;
  constexpr int NPT = C::NPT, PARTS = C::PARTS, NPH = C::NPH, NVH = C::NVH;
  constexpr bool STAGE = C::STAGE;
  constexpr dgk::FmaskTable<N> FM{};
  __shared__ __attribute__((aligned(16))) unsigned char sA[C::FRAG_BYTES];
  __shared__ double sTall[W][C::BUF];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFM[NF];
  {
    const uint4* src = (const uint4*)a.frag;
    uint4* dst = (uint4*)sA;
    for (int i = threadIdx.x; i < (int)(C::FRAG_BYTES / 16); i += 64 * W) dst[i] = src[i];
    for (int i = threadIdx.x; i < 4 * 24 * NFP; i += 64 * W) sFN[i] = a.face_nodes[i];
    for (int i = threadIdx.x; i < NF; i += 64 * W) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
  }
  __syncthreads();
  const double dt = a.mode == 0 ? *a.dt : 0.0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = lane & 15, g = lane >> 4;  // element of the tile, face
  double* sT = sTall[w];
  const int64_t K = a.K;
  double lmax = 0.0;
  // Tile order: with a grid that is a multiple of 8, block b runs on XCD
  // b % 8 (round-robin dispatch), and each XCD walks one contiguous eighth of
  // the tile list, so neighbour traces mostly hit that XCD's L2; otherwise a
  // plain grid stride.
  int t0 = blockIdx.x * W + w, t1 = a.ntiles, ts = gridDim.x * W;
  if (gridDim.x % 8 == 0) {
    const int x = blockIdx.x % 8;
    t0 = (int)((int64_t)a.ntiles * x / 8) + (blockIdx.x / 8) * W + w;
    t1 = (int)((int64_t)a.ntiles * (x + 1) / 8);
    ts = (gridDim.x / 8) * W;
  }
  for (int t = t0; t < t1; t += ts) {
    const dgk::Tile tile = a.tiles[t];
    const int cnt = tile.count;
    const int ce = c < cnt ? c : cnt - 1;  // clamped: every load stays in the tile
    const int64_t ge = tile.gbase + ce;
    const double* ut = a.u + tile.vbase;
    const int code = (a.codes[ge] >> (5 * g)) & 31;
    const int32_t nb = a.nbr[(int64_t)g * K + ge];
    double fa, fb;
    dgk::face_factors(g, a.bgeo[ge], a.bgeo[K + ge], a.bgeo[2 * K + ge], fa, fb);
    // Traces of face g, all loads in flight together.
    const uint8_t* fn = sFN + (g * 24 + code) * NFP;
    const uint8_t* fm = sFM + g * NFP;
    double uP[NFP], uM[NFP];
    // Own traces through LDS, a quarter tile at a time: unit-stride loads of
    // elements 4 h .. 4 h + 3, then the lanes of those elements read theirs.
#pragma unroll 1
    for (int h = 0; h < (STAGE ? 4 : 0); ++h) {
      const int lim = min(cnt, 4 * h + 4) * NP - 4 * h * NP;
      wave_sync();  // previous readers of the buffer are done
#pragma unroll
      for (int j = 0; j < NVH; ++j) {
        const int v = lane + 64 * j;
        if (v < lim) {
          const int e = v / NP;
          sT[e * NPH + (v - e * NP)] = ut[4 * h * NP + v];
        }
      }
      wave_sync();
      if ((ce >> 2) == h) {
        const double* se = sT + (ce - 4 * h) * NPH;
#pragma unroll
        for (int q = 0; q < NFP; ++q) uM[q] = se[fm[q]];
      }
    }
#pragma unroll
    for (int q = 0; q < NFP; ++q) {
      uP[q] = a.u[nb + fn[q]];
      if constexpr (!STAGE) uM[q] = ut[ce * NP + fm[q]];
    }
    // Fscale (n.f(u-) - n.f*), f(u) = u^2/2 (1,1,1), factored; zero past Nfp.
    auto jump = [&](int q) -> double {
      if (q >= NFP) return 0.0;
      const double al = fmax(fabs(uM[q]), fabs(uP[q]));
      return (uM[q] - uP[q]) * fma(fa, uM[q] + uP[q], -fb * al);
    };
    // B operand of each step in the product's precision: the jumps (k = 4)
    // or their bf16 parts (k = 32), split as soon as each is formed.
    double gj[is_bf(PREC) ? 1 : NPT];
    bf16x8 bp[is_bf(PREC) ? KSL : 1][PARTS];
    if constexpr (is_bf(PREC)) {
#pragma unroll
      for (int s = 0; s < KSL; ++s)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const double x = jump(8 * s + e);
          if constexpr (PARTS == 2) {
            // two parts carry ~16 bits: the FP32 residual is exact enough
            const float xf = (float)x;
            const __bf16 h = (__bf16)xf;
            bp[s][0][e] = h;
            bp[s][1][e] = (__bf16)(xf - (float)h);
          } else {
            const __bf16 h = (__bf16)(float)x;
            const double r1 = x - (double)(float)h;
            const __bf16 m = (__bf16)(float)r1;
            bp[s][0][e] = h;
            bp[s][1][e] = m;
            bp[s][2][e] = (__bf16)(float)(r1 - (double)(float)m);
          }
        }
    } else {
#pragma unroll
      for (int q = 0; q < NPT; ++q) gj[q] = jump(q);
    }
    const double* rt = a.rhs + tile.vbase;
    const double* u0t = a.a0 != 0.0 ? a.u0 + tile.vbase : nullptr;
    double* ot = a.out + tile.vbase;
    // out from preloaded rhs, u, u0 values
    auto finish = [&](int64_t idx, double lg, double rv, double uv, double u0v) {
      const double l = rv + lg;
      double o;
      if (a.mode == 0) {
        o = a.a1 * fma(dt, l, uv);
        if (a.a0 != 0.0) o = fma(a.a0, u0v, o);
        lmax = fmax(lmax, fabs(o));
      } else {
        o = l;
      }
      __builtin_nontemporal_store(o, ot + idx);
    };
    // LIFT g, row tile by row tile (not unrolled: the preloads of every row
    // tile would otherwise be hoisted together).
#pragma unroll 1
    for (int tt = 0; tt < MTL; ++tt) {
      // The rows this lane updates after the product (element 4 j + lane / 16,
      // row 16 tt + lane % 16), loaded ahead of the product.
      double pr[4], pu[4], p0[4];
      int64_t pidx[4];
      bool pok[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = 4 * j + (lane >> 4), i = 16 * tt + (lane & 15);
        pok[j] = e < cnt && i < NP;
        pidx[j] = pok[j] ? (int64_t)e * NP + i : 0;
        pr[j] = rt[pidx[j]];
        pu[j] = a.mode == 0 ? ut[pidx[j]] : 0.0;
        p0[j] = a.a0 != 0.0 ? __builtin_nontemporal_load(u0t + pidx[j]) : 0.0;
      }
      AccT<PREC> acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = 0;
      __builtin_amdgcn_s_setprio(1);  // matrix phase first (wave priority)
#pragma unroll
      for (int s = 0; s < KSL; ++s) {
        const size_t q = ((size_t)tt * KSL + s) * 64 + lane;
        if constexpr (PREC == kF64) {
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(((const double*)sA)[q], gj[s], acc, 0, 0, 0);
        } else if constexpr (PREC == kF32) {
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(((const float*)sA)[q], (float)gj[s], acc, 0, 0,
                                                     0);
        } else {
          const bf16x8* F = (const bf16x8*)sA;
          if constexpr (PREC == kBF3) {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[s][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[s][1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[s][0], acc, 0, 0, 0);
          } else {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[2 * C::NFRAG + q], bp[s][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[s][2], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[s][1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[s][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[s][1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[s][0], acc, 0, 0, 0);
          }
        }
      }
      __builtin_amdgcn_s_setprio(0);
      // Through the wave's 16 x 17 buffer to rows of 16 consecutive values.
      wave_sync();  // previous readers of the buffer are done
#pragma unroll
      for (int r = 0; r < 4; ++r) sT[c * 17 + acc_row<PREC>(g, r)] = (double)acc[r];
      wave_sync();
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (pok[j]) finish(pidx[j], sT[(4 * j + (lane >> 4)) * 17 + (lane & 15)], pr[j], pu[j], p0[j]);
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

// LIFT fragments in the MFMA A-operand lane order of the precision, columns
// taken face by face: lane (row, g) of step s holds LIFT[row][g Nfp + s]
// (k = 4) or LIFT[row][g Nfp + 8 s + e], e = 0..7 (bf16, zero past Nfp).
std::vector<unsigned char> lift_fragments(const std::vector<double>& LIFT, int N, int prec) {
  const int NP = dgk::np_of(N), NFP = dgk::nfp_of(N), NF = 4 * NFP;
  const int MTL = (NP + 15) / 16;
  const int KSL = is_bf(prec) ? (NFP + 7) / 8 : NFP, PARTS = parts_of(prec);
  const size_t nfrag = (size_t)MTL * KSL * 64;
  auto val = [&](int i, int g, int q) {
    return (i < NP && q < NFP) ? LIFT[(size_t)i * NF + g * NFP + q] : 0.0;
  };
  std::vector<unsigned char> buf;
  if (!is_bf(prec)) {
    buf.resize(nfrag * (prec == kF64 ? 8 : 4));
    for (int tt = 0; tt < MTL; ++tt)
      for (int s = 0; s < KSL; ++s)
        for (int l = 0; l < 64; ++l) {
          const double v = val(16 * tt + (l & 15), l >> 4, s);
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
            const double v = val(16 * tt + (l & 15), l >> 4, 8 * s + e);
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
  switch (m->lift_prec) {
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
  const std::vector<unsigned char> frag = lift_fragments(m->LIFT_h, m->N, m->lift_prec);
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
