// Physical derivative on affine tetrahedra with matrix cores, in FP64, FP32
// and split BF16 (three- or six-product) arithmetic.
//
// For an affine element the metrics are constant, so
//   du/dx = rx (Dr u) + sx (Ds u) + tx (Dt u)   (likewise y, z),
// the reference's calcPhysicalDerivative
// (tetnudg/tet_nudg_mesh_physical_derivative_test.go:344-367) with one
// metric value per element instead of one per node.
//
// Work decomposition. The stacked operator A = [Dr; Ds; Dt] (3 Np rows,
// padded to MT tiles of 16 rows) is the MFMA A operand; 16 elements of u are
// the B operand (N dimension). A block of G waves shares one 16-element
// tile at a time; wave w owns row tiles [w TPW, (w+1) TPW) and keeps their
// operator fragments in registers for the whole launch, so the operator is
// read from memory once per wave, not once per tile. Per element tile:
//   1. u (16 Np contiguous doubles) and the 9 metrics are staged in LDS with
//      coalesced 16-byte loads; the next tile's loads are issued before the
//      matrix phase;
//   2. low-precision modes subtract each element's linear interpolant through
//      its four vertex nodes (w = u - L u, in FP64, a group of memory threads
//      per element); D L u is the constant (u_v1 - u_v0, u_v2 - u_v0,
//      u_v3 - u_v0)/2 in (r, s, t) and is added back in the epilogue.
//      Rounding errors then scale with |u - L u| = O(h^2 |u''|) instead of
//      |u|;
//   3. every wave runs its MFMAs with B read from LDS, and parks the results
//      in LDS;
//   4. the epilogue applies the chain rule and writes the outputs with
//      contiguous stores: in FP64 for the FP64 plan and the BF16x6
//      divergence, in FP32 for the other plans below FP64 (FP32 outputs,
//      the divergence of the split Burgers stage included).
//
//   F64:     v_mfma_f64_4x4x4_4b_f64 (four per 16x16x4 step, the operator as
//            the B operand).
//   F32:     v_mfma_f32_16x16x4_f32 on FP32-rounded operator and w.
//   BF16X3:  v_mfma_f32_16x16x32_bf16 on two-part bf16 splits,
//            hi*hi + hi*lo + lo*hi (Henry, Tang & Heinecke, ARITH 2019);
//            ~16 significant bits per operand.
//   BF16X6:  three-part splits (of the FP32 value of w), hi*hi + hi*mid +
//            mid*hi + mid*mid + hi*lo + lo*hi; ~24 bits, FP32-level accuracy.
//   I8X6:    exact-integer split (Ozaki, Ogita & Oishi, Numer. Algorithms
//            2012; Ootomo, Ozaki & Yokota, IJHPCA 2024, on integer matrix
//            units): each operator row and each element's w is scaled by a
//            power of two to |x| <= 1 and rounded to a 46-bit integer, which
//            is written as six signed base-256 digits (int8). The digit pairs
//            of weight a + b >= 5 (21 of 36) run as v_mfma_i32_16x16x64_i8
//            with exact int32 accumulation, one accumulator per weight; the
//            weights are combined and scaled back in FP64. Error ~2^-46 of the
//            row and element magnitudes: FP64-level results (~3e-12 rel.).
// BF16 and FP32 modes accumulate in FP32 and write FP32 outputs; I8X6 writes
// FP64.
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dgmesh_internal.h"

namespace {

typedef double doublex2 __attribute__((ext_vector_type(2)));
typedef unsigned int uintx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx4 __attribute__((ext_vector_type(4)));

constexpr int kF64 = 0, kF32 = 1, kBF3 = 2, kBF6 = 3, kI8 = 4;
// Exact-integer split: digits per operand, and the lowest pair weight kept.
constexpr int kDig = 6, kThr = 5;
constexpr int kDiag = 2 * kDig - 1 - kThr;  // accumulators (pair weights kThr..2 kDig-2)
constexpr int kBits = 8 * kDig - 2;          // integer bits of a scaled operand
// Integer plan K steps: 16x16x64 steps, and a final 16x16x32 step when the
// rest of the Np columns fits in 32 (same issue cost, half the registers).
constexpr int i8_k64(int np) { return np / 64 + (np % 64 > 32 ? 1 : 0); }
constexpr int i8_t32(int np) { return (np % 64 != 0 && np % 64 <= 32) ? 1 : 0; }

constexpr int parts_of(int prec) {
  return prec == kBF3 ? 2 : (prec == kBF6 ? 3 : (prec == kI8 ? kDig : 1));
}
constexpr bool is_bf(int prec) { return prec == kBF3 || prec == kBF6; }
constexpr int freg_of(int prec) { return prec == kF64 ? 2 : (prec == kF32 ? 1 : 4); }

template <int N, int PREC>
struct ACfg {
  static constexpr int NP = (N + 1) * (N + 2) * (N + 3) / 6;
  static constexpr int ROWS = 3 * NP;
  static constexpr int MT = (ROWS + 15) / 16;
  static constexpr int KSTEP = is_bf(PREC) ? 32 : (PREC == kI8 ? 64 : 4);
  static constexpr int K64 = PREC == kI8 ? i8_k64(NP) : 0;  // integer plan: x64 steps
  static constexpr int T32 = PREC == kI8 ? i8_t32(NP) : 0;  // integer plan: x32 tail step
  static constexpr int KS = PREC == kI8 ? K64 + T32 : (NP + KSTEP - 1) / KSTEP;
  static constexpr int KPAD = 64 * K64 + 32 * T32;           // integer plan: padded K
  static constexpr int PARTS = parts_of(PREC);
  static constexpr int FRAG_REGS = MT * KS * PARTS * freg_of(PREC);
  // G matrix waves hold the operator; G memory waves stage, convert and
  // write back. Four matrix waves (one per SIMD) keep long MFMA chains; the
  // low-precision kernels are held to 128 registers (WPE = 4 waves per
  // SIMD) so that two blocks share a CU and one block's matrix phase
  // overlaps the other's memory phase. FP64 needs ~170 fragment and
  // accumulator registers per matrix wave and runs one block per CU.
  // The integer plan keeps six operator digits in registers (72 at P=6
  // with two row tiles per wave), so eight matrix waves share them.
  static constexpr int G =
      MT < 4 ? MT
             : ((PREC == kF64 && MT >= 8) || (PREC == kI8 && MT >= 8) ||
                        (FRAG_REGS / 4 > 100 && PREC != kF64 && PREC != kI8)
                    ? 8
                    : 4);
  static constexpr int WPE = PREC == kF64 && G < 8 ? 2 : 4;
  static constexpr int TPW = (MT + G - 1) / G;
  static constexpr int NTM = 64 * G;   // memory threads
  static constexpr int NT = 64 * G + NTM;  // threads per block
  static constexpr int NPU = NP | 1;                     // LDS row stride of u, in doubles
  static constexpr int NPF = NP | 1;                     // LDS row stride of FP32 w, in floats
  static constexpr int KPB = KS * 32 + 8;                // LDS row stride of bf16 w parts
  static constexpr int KPI = KPAD + 16;                  // LDS row stride of int8 w digits, bytes
  static constexpr int ROWSP = MT * 16 + 1;              // LDS row stride of results
  static constexpr int NV2 = 8 * NP;                     // double2 per 16-element tile
  static constexpr int NL = (NV2 + NTM - 1) / NTM;       // double2 loads per memory thread
  static constexpr int NIT = (16 * NP + NTM - 1) / NTM;  // values per memory thread
  static constexpr size_t NFRAG = (size_t)MT * KS * 64;  // fragments per part
};

// Host-side copy of the fragment shape (same formulas as ACfg).
struct HostCfg {
  int NP, MT, KSTEP, KS, PARTS;
};
HostCfg host_cfg(int N, int prec) {
  HostCfg c;
  c.NP = dgk::np_of(N);
  c.MT = (3 * c.NP + 15) / 16;
  c.KSTEP = is_bf(prec) ? 32 : (prec == kI8 ? 64 : 4);
  c.KS = prec == kI8 ? i8_k64(c.NP) + i8_t32(c.NP) : (c.NP + c.KSTEP - 1) / c.KSTEP;
  c.PARTS = parts_of(prec);
  return c;
}

template <int PREC>
struct Frag;
template <>
struct Frag<kF64> {
  double v;
};
template <>
struct Frag<kF32> {
  float v;
};
template <>
struct Frag<kBF3> {
  bf16x8 v[2];
};
template <>
struct Frag<kBF6> {
  bf16x8 v[3];
};
template <>
struct Frag<kI8> {
  intx4 v[kDig];  // 16x16x64 step: digit a has weight 256^a
  long t[kDig];   // 16x16x32 tail step
};

template <int PREC>
using AccT = typename std::conditional<PREC == kF64, doublex4, floatx4>::type;
template <int PREC>
using ResT = typename std::conditional<PREC == kF64 || PREC == kI8, double, float>::type;

struct AffineArgs {
  int64_t K;                // dense mode: elements; tile mode: unused
  const dgk::Tile* tiles;   // tile mode: 16-element tiles in the runner layout (else dense)
  int64_t ntiles;
  const void* frag;    // operator fragments, [part][MT][KS][64]
  const double* lam;   // [3][NP] barycentric weights l1, l2, l3 of each node (mixed modes)
  const int* rowexp;   // I8X6: power-of-two scale of each operator row [MT*16]
  const double* rowcorr;  // I8X6: [MT*16][4] D applied to 1, l1, l2, l3 minus its exact value
  int vtx[4];          // indices of the vertex nodes (mixed modes)
  const double* u;
  const double* geo;
  int64_t gs;
  void* ox;
  void* oy;
  void* oz;
};

// DIV = false: du/dx, du/dy, du/dz into ox, oy, oz (FP64 or FP32).
// DIV = true: the Burgers volume term -div f(u), f(u) = u^2/2 (1,1,1), into
// ox (FP32 for the plans with an FP32 epilogue, FP32 and BF16x3; else FP64):
// the staged values are f, and the epilogue folds the metrics to
// -( (rx+ry+rz) Dr f + (sx+sy+sz) Ds f + (tx+ty+tz) Dt f ).
template <int N, int PREC, bool DIV>
__global__ void __launch_bounds__((ACfg<N, PREC>::NT))
__attribute__((amdgpu_waves_per_eu(ACfg<N, PREC>::WPE)))
deriv_affine_kernel(AffineArgs a) {
  using C = ACfg<N, PREC>;
  constexpr int NP = C::NP, MT = C::MT, KS = C::KS, TPW = C::TPW, G = C::G, NTM = C::NTM;
  constexpr int NPU = C::NPU, NPF = C::NPF, KPB = C::KPB, KPI = C::KPI, ROWSP = C::ROWSP;
  constexpr int NV2 = C::NV2, NL = C::NL, NIT = C::NIT, PARTS = C::PARTS;
  constexpr bool MIXED = PREC != kF64;
  using R = ResT<PREC>;
  // Software pipeline over the block's tiles n = 0, 1, ... (tile index
  // blockIdx.x + n gridDim.x). Between two barriers, phase n runs
  //   memory waves: stage(n + LEAD)  prefetched registers -> sU, sM
  //                 lin(n + 1)       (below FP64) sU -> w = u - Lu in the
  //                                  matrix precision; slopes of Lu -> sM
  //                 epi(n - 1)       sR, sM -> outputs
  //   matrix waves: mma(n)           w (or sU) -> sR
  // so the matrix cores of one tile overlap the memory and vector work of
  // its neighbours. Buffers are indexed by tile parity (sU, w, sR) or by
  // n mod 4 (sM), which keeps every live tile in its own buffer.
  constexpr int LEAD = MIXED ? 2 : 1;
  __shared__ double sU[2][16 * NPU];
  __shared__ float sWf[PREC == kF32 ? 2 : 1][PREC == kF32 ? 16 * NPF : 1];
  __shared__ __attribute__((aligned(16))) __bf16 sWb[is_bf(PREC) ? 2 : 1][PARTS]
                                                    [is_bf(PREC) ? 16 * KPB : 1];
  __shared__ R sR[2][16 * ROWSP];
  // I8X6: the digits of w, [parity][digit][element][KPI bytes], and each
  // element's power-of-two scale
  __shared__ __attribute__((aligned(16))) int8_t sWi[PREC == kI8 ? 2 : 1][PREC == kI8 ? kDig : 1]
                                                     [PREC == kI8 ? 16 * KPI : 1];
  __shared__ int sF[2][16];
  // I8X6: each element's vertex values (a0, a1 - a0, a2 - a0, a3 - a0), and
  // the per-row correction that applies the stored operator to the linear
  // part (the exact slopes are added in the epilogue)
  __shared__ __attribute__((aligned(16))) double sA[PREC == kI8 ? 2 : 1][PREC == kI8 ? 64 : 1];
  __shared__ __attribute__((aligned(16))) double sDel[PREC == kI8 ? MT * 64 : 1];
  __shared__ int sRx[PREC == kI8 ? MT * 16 : 1];  // I8X6 row scales
  // I8X6: the operator's least significant digit (weight 1), which enters a
  // single pair (with w's top digit), sits in LDS; the matrix waves keep the
  // other five digits in registers.
  __shared__ intx4 sOpx[PREC == kI8 ? MT * (C::K64 > 0 ? C::K64 : 1) * 64 : 1];
  __shared__ long sOpt[PREC == kI8 && C::T32 ? MT * 64 : 1];
  // Per element: 9 metrics rx..tz, then (below FP64) half the linear part's
  // r, s, t slopes, padded to 12 doubles so the epilogue reads it with six
  // 16-byte loads.
  __shared__ __attribute__((aligned(16))) double sM[4][16 * 12];
  // Below FP64 the same per-element values in FP32, for an FP32 epilogue: its
  // rounding adds a few FP32 ulps, below the products' own error -- except for
  // the BF16x6 divergence (the split Burgers stage), whose ~2^-24 products
  // keep the FP64 epilogue.
  constexpr bool EPI32 = PREC == kF32 || PREC == kBF3 || (PREC == kBF6 && !DIV);
  __shared__ __attribute__((aligned(16))) float sMf[EPI32 ? 4 : 1][EPI32 ? 16 * 12 : 1];
  __shared__ double sLam[MIXED ? 3 * NP : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const bool is_mma = w < G;
  const int mt = tid - 64 * G;  // memory-thread index (valid when !is_mma)
  const int c = lane & 15, g = lane >> 4;

  // All setup loads complete here, so that the wait counters inside the
  // tile loop only ever cover the tile prefetch.
  __builtin_amdgcn_s_waitcnt(0);

  const int64_t K = a.K;
  const int64_t ntiles = a.tiles ? a.ntiles : (K + 15) / 16;
  const int64_t Gd = gridDim.x;
  const int64_t b0 = blockIdx.x;
  const int64_t nb = b0 < ntiles ? (ntiles - b0 + Gd - 1) / Gd : 0;  // tiles of this block
  constexpr int PD = 2;  // tiles whose loads are in flight per block
  if (nb == 0) return;

  // Tile t: value offset of its first element, first element (metrics),
  // element count.
  struct TileRef {
    int64_t vbase, gbase;
    int count;
  };
  auto tile_ref = [&](int64_t t) -> TileRef {
    if (a.tiles) {
      const dgk::Tile tl = a.tiles[t];
      return {tl.vbase, (int64_t)tl.gbase, (int)tl.count};
    }
    const int64_t left = K - t * 16;
    return {t * 16 * NP, t * 16, (int)(left < 16 ? left : 16)};
  };

  // Tile loads through a buffer resource based at the tile: the per-thread
  // offsets are loop-invariant registers, and reads past the tile's values
  // return zero. A pair that straddles the end is fixed up when staged.
  auto prefetch = [&](int64_t n, doublex2 (&P)[NL], double& PM) {
    const TileRef tr = tile_ref(b0 + n * Gd);
    if (mt < 9 * 16) {
      const int e = mt & 15;
      PM = a.geo[(mt >> 4) * a.gs + tr.gbase + (e < tr.count ? e : 0)];
    }
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.u + tr.vbase), 0, tr.count * NP * 8, 0x00020000);
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int v = mt + q * NTM;
      P[q] = __builtin_bit_cast(doublex2, __builtin_amdgcn_raw_buffer_load_b128(rs, 16 * v, 0, 0));
    }
  };

  // stage(n): registers of slot n % PD -> sU[n % 2] (u, or f = u^2/2 for
  // DIV), sM[n % 4]; refills the slot.
  auto stage = [&](int64_t n, int bu, int bm, doublex2 (&P)[NL], double& PM) {
    const TileRef tr = tile_ref(b0 + n * Gd);
    const int lim = tr.count * NP;  // values of this tile
    double* U = sU[bu];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int v = mt + q * NTM;
      if (v < NV2) {
        const int i0 = 2 * v, i1 = 2 * v + 1;
        double x0 = P[q].x, x1 = P[q].y;
        if (i1 >= lim) {
          x0 = i0 < lim ? a.u[tr.vbase + i0] : 0.0;
          x1 = 0.0;
        }
        if constexpr (DIV) {
          x0 = 0.5 * x0 * x0;
          x1 = 0.5 * x1 * x1;
        }
        U[(i0 / NP) * NPU + i0 % NP] = x0;
        U[(i1 / NP) * NPU + i1 % NP] = x1;
      }
    }
    if constexpr (NTM >= 9 * 16) {
      if (mt < 9 * 16) {
        sM[bm][(mt & 15) * 12 + (mt >> 4)] = PM;
        if constexpr (EPI32) sMf[bm][(mt & 15) * 12 + (mt >> 4)] = (float)PM;
      }
    } else {
      for (int v = mt; v < 9 * 16; v += NTM) {
        const int q = v >> 4, e = v & 15;
        const double mv = e < tr.count ? a.geo[q * a.gs + tr.gbase + e] : 0.0;
        sM[bm][e * 12 + q] = mv;
        if constexpr (EPI32) sMf[bm][e * 12 + q] = (float)mv;
      }
    }
    if (n + PD < nb) prefetch(n + PD, P, PM);
  };

  // lin: w = u minus the linear interpolant through the vertex nodes, in the
  // matrix precision (FP32, or bf16 parts with zero padding up to KS*32).
  auto lin = [&](int bu, int bl) {
    static_assert(NTM % 16 == 0, "memory threads split evenly over 16 elements");
    if constexpr (MIXED) {
      // Memory thread (e, t) handles values t + TPE j (FP32) or value pairs
      // (bf16 parts, stored as one 32-bit word each) of element e, so that the
      // vertex values are read once per element.
      constexpr int TPE = NTM / 16;
      const int e = mt / TPE, t = mt % TPE;
      const double* ue = sU[bu] + e * NPU;
      const double a0 = ue[a.vtx[0]];
      const double a1 = ue[a.vtx[1]] - a0, a2 = ue[a.vtx[2]] - a0, a3 = ue[a.vtx[3]] - a0;
      if (t == 0) {
        sM[bl][e * 12 + 9] = 0.5 * a1;
        sM[bl][e * 12 + 10] = 0.5 * a2;
        sM[bl][e * 12 + 11] = 0.5 * a3;
        if constexpr (EPI32) {
          sMf[bl][e * 12 + 9] = (float)(0.5 * a1);
          sMf[bl][e * 12 + 10] = (float)(0.5 * a2);
          sMf[bl][e * 12 + 11] = (float)(0.5 * a3);
        }
      }
      auto rem = [&](int i) -> double {
        return i < NP ? ue[i] - (a0 + a1 * sLam[i] + a2 * sLam[NP + i] + a3 * sLam[2 * NP + i])
                      : 0.0;
      };
      if constexpr (PREC == kF32) {
        constexpr int NJ = (NP + TPE - 1) / TPE;
  #pragma unroll 1
        for (int j = 0; j < NJ; ++j) {
          const int i = t + TPE * j;
          if (NP % TPE == 0 || i < NP) sWf[bu][e * NPF + i] = (float)rem(i);
        }
      } else if constexpr (PREC == kI8) {
        // Thread t of the element's TPE threads takes groups of four values
        // t, t + TPE, ...; the element's scale 2^E (|w| < 2^E) comes from a
        // max over those threads (consecutive lanes).
        constexpr int NG = C::KPAD / 4, GPT = (NG + TPE - 1) / TPE;
        static_assert(64 % TPE == 0, "an element's threads sit in one wave");
        double x[GPT][4];
        double mx = 0.0;
#pragma unroll
        for (int j = 0; j < GPT; ++j)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const int i = 4 * (t + TPE * j) + b;
            x[j][b] = rem(i);
            mx = fmax(mx, fabs(x[j][b]));
          }
#pragma unroll
        for (int o = TPE / 2; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
        int E = 0;
        if (mx > 0.0) (void)frexp(mx, &E);
        if (t == 0) {
          sF[bu][e] = E;
          sA[bu][4 * e] = a0;
          sA[bu][4 * e + 1] = a1;
          sA[bu][4 * e + 2] = a2;
          sA[bu][4 * e + 3] = a3;
        }
        // y = round(w 2^(kBits - E)), |y| <= 2^46. The bytes of y + B, B =
        // 0x808080808080, each minus 128 (XOR 0x80), are the balanced
        // base-256 digits of y; split at 2^24 into exact uint32 halves, and
        // transpose the bytes of four values into one word per digit.
        const double B = 141289400074368.0;
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
          const int gi = t + TPE * j;
          if (NG % TPE != 0 && gi >= NG) continue;
          uint32_t L[4], H[4];
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const double z = rint(ldexp(x[j][b], kBits - E)) + B;
            const double zh = floor(z * 0x1p-24);
            L[b] = (uint32_t)(z - zh * 0x1p24) ^ 0x808080u;
            H[b] = (uint32_t)zh ^ 0x808080u;
          }
#pragma unroll
          for (int d = 0; d < kDig; ++d) {
            const uint32_t* S = d < 3 ? L : H;
            const uint32_t k = (uint32_t)(d % 3);
            const uint32_t lo = __builtin_amdgcn_perm(S[1], S[0], 0x0c0c0000u | ((4u + k) << 8) | k);
            const uint32_t hi = __builtin_amdgcn_perm(S[3], S[2], ((4u + k) << 24) | (k << 16) | 0x0c0cu);
            *(uint32_t*)(sWi[bu][d] + e * KPI + 4 * gi) = lo | hi;
          }
        }
      } else {
        // bf16 rows are zero-padded up to KS*32 values
        constexpr int NPAIR = KS * 16, NJ = (NPAIR + TPE - 1) / TPE;
        auto bits = [](__bf16 h) -> uint32_t { return (uint32_t)__builtin_bit_cast(uint16_t, h); };
  #pragma unroll 1
        for (int j = 0; j < NJ; ++j) {
          const int pr = t + TPE * j;
          if (NPAIR % TPE == 0 || pr < NPAIR) {
            const int i0 = 2 * pr;
            const double x0 = rem(i0), x1 = rem(i0 + 1);
            uint32_t* w0 = (uint32_t*)(sWb[bu][0] + e * KPB + i0);
            uint32_t* w1 = (uint32_t*)(sWb[bu][1] + e * KPB + i0);
            if constexpr (PREC == kBF3) {
              // two parts carry ~16 bits: the FP32 residual is exact enough
              const float f0 = (float)x0, f1 = (float)x1;
              const __bf16 h0 = (__bf16)f0, h1 = (__bf16)f1;
              *w0 = bits(h0) | (bits(h1) << 16);
              *w1 = bits((__bf16)(f0 - (float)h0)) | (bits((__bf16)(f1 - (float)h1)) << 16);
            } else {
              // three parts of the FP32 value hold it to ~2^-24: each
              // difference below is exact in FP32
              const float f0 = (float)x0, f1 = (float)x1;
              const __bf16 h0 = (__bf16)f0, h1 = (__bf16)f1;
              const float r0 = f0 - (float)h0, r1 = f1 - (float)h1;
              const __bf16 m0 = (__bf16)r0, m1 = (__bf16)r1;
              uint32_t* w2 = (uint32_t*)(sWb[bu][2] + e * KPB + i0);
              *w0 = bits(h0) | (bits(h1) << 16);
              *w1 = bits(m0) | (bits(m1) << 16);
              *w2 = bits((__bf16)(r0 - (float)m0)) | (bits((__bf16)(r1 - (float)m1)) << 16);
            }
          }
        }
      }
    }
  };

  // mma: this wave's row tiles of [Dr; Ds; Dt] times the tile -> sR[n % 2].
  auto mma = [&](int bu, const Frag<PREC> (&fa)[TPW][KS]) {
    AccT<PREC> acc[TPW];
#pragma unroll
    for (int j = 0; j < TPW; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[j][r] = 0;
    if constexpr (PREC == kF64) {
      // FP64 on v_mfma_f64_4x4x4_4b, which issues about twice the FP64 rate
      // of v_mfma_f64_16x16x4 on gfx950 (72 against 36-47 TFLOP/s in a
      // 1-2 wave per SIMD loop). Its operand layout (checked on gfx950):
      // block j of four is lanes 16 k + 4 j + i, with A[i][k] and B[k][i] at
      // that lane and D[i][n] at lane 16 i + 4 j + n. The operator fragment
      // (lane l: row 16 tt + l%16, column 4 s + l/16) is the B operand
      // unchanged, block j taking operator rows 4 j..4 j+3; instruction q
      // takes elements 4 q..4 q+3 as A (lane l: element 4 q + l%4, node
      // 4 s + l/16), and leaves row 16 tt + l%16 of element 4 q + l/16.
      double a4[TPW][4];
#pragma unroll
      for (int j = 0; j < TPW; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) a4[j][q] = 0.0;
      const double* ua = sU[bu] + (lane & 3) * NPU + (lane >> 4);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const bool kin = 4 * s + (lane >> 4) < NP;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double a = kin ? ua[4 * q * NPU + 4 * s] : 0.0;
#pragma unroll
          for (int j = 0; j < TPW; ++j)
            a4[j][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, fa[j][s].v, a4[j][q], 0, 0, 0);
        }
      }
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        const int tt = w * TPW + j;
        if (tt < MT) {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            sR[bu][(4 * q + (lane >> 4)) * ROWSP + 16 * tt + (lane & 15)] = a4[j][q];
        }
      }
      return;
    } else if constexpr (PREC == kF32) {
      using BT = float;
      const BT* ub = sWf[bu] + c * NPF;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + g;
        const BT b = k < NP ? ub[k] : (BT)0;
#pragma unroll
        for (int j = 0; j < TPW; ++j)
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[j][s].v, b, acc[j], 0, 0, 0);
      }
    } else if constexpr (PREC == kI8) {
      // Per row tile: kDiag int32 accumulators, one per pair weight; B digits
      // are read from LDS one at a time, so only one is live.
      constexpr int K64 = C::K64;
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        intx4 ac[kDiag];
#pragma unroll
        for (int d = 0; d < kDiag; ++d) ac[d] = intx4{0, 0, 0, 0};
        // One B digit ahead is in flight; the scheduling barriers keep the
        // compiler from hoisting more (registers: the operator digits fill
        // most of the budget).
#pragma unroll
        for (int s = 0; s < K64; ++s) {
          const int8_t* wb = sWi[bu][0] + c * KPI + 64 * s + 16 * g;
          intx4 bn = *(const intx4*)wb;
#pragma unroll
          for (int q = 0; q < kDig; ++q) {
            const intx4 b = bn;
            if (q + 1 < kDig) bn = *(const intx4*)(wb + (q + 1) * 16 * KPI);
#pragma unroll
            for (int p = kDig - 1; p >= 0; --p)
              if (p + q >= kThr) {
                const intx4 af = p > 0 ? fa[j][s].v[p] : sOpx[((w * TPW + j) * K64 + s) * 64 + lane];
                ac[p + q - kThr] =
                    __builtin_amdgcn_mfma_i32_16x16x64_i8(af, b, ac[p + q - kThr], 0, 0, 0);
              }
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        if constexpr (C::T32) {
          const int8_t* wb = sWi[bu][0] + c * KPI + 64 * K64 + 8 * g;
          long bn = *(const long*)wb;
#pragma unroll
          for (int q = 0; q < kDig; ++q) {
            const long b = bn;
            if (q + 1 < kDig) bn = *(const long*)(wb + (q + 1) * 16 * KPI);
#pragma unroll
            for (int p = kDig - 1; p >= 0; --p)
              if (p + q >= kThr) {
                const long af = p > 0 ? fa[j][K64].t[p] : sOpt[(w * TPW + j) * 64 + lane];
                ac[p + q - kThr] =
                    __builtin_amdgcn_mfma_i32_16x16x32_i8(af, b, ac[p + q - kThr], 0, 0, 0);
              }
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        const int tt = w * TPW + j;
        if (tt < MT) {
          const int E = sF[bu][c] + 8 * kThr - 2 * kBits;
          const doublex2 av01 = *(const doublex2*)(sA[bu] + 4 * c);
          const doublex2 av23 = *(const doublex2*)(sA[bu] + 4 * c + 2);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 16 * tt + 4 * g + r;
            double y = (double)ac[kDiag - 1][r];
#pragma unroll
            for (int d = kDiag - 2; d >= 0; --d) y = fma(y, 256.0, (double)ac[d][r]);
            const doublex2 d01 = *(const doublex2*)(sDel + 4 * row);
            const doublex2 d23 = *(const doublex2*)(sDel + 4 * row + 2);
            const double corr = av01.x * d01.x + av01.y * d01.y + av23.x * d23.x + av23.y * d23.y;
            sR[bu][c * ROWSP + row] = ldexp(y, sRx[row] + E) + corr;
          }
        }
      }
      return;
    } else {
      bf16x8 bb[KS][PARTS];
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int p = 0; p < PARTS; ++p)
          bb[s][p] = *(const bf16x8*)(sWb[bu][p] + c * KPB + 32 * s + 8 * g);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const bf16x8* b = bb[s];
#pragma unroll
        for (int j = 0; j < TPW; ++j) {
          if constexpr (PREC == kBF3) {
            // parts: [0] = hi, [1] = lo; smallest products first
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[1], b[0], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[0], b[1], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[0], b[0], acc[j], 0, 0, 0);
          } else {
            // parts: [0] = hi, [1] = mid, [2] = lo; smallest products first
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[2], b[0], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[0], b[2], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[1], b[1], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[1], b[0], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[0], b[1], acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[j][s].v[0], b[0], acc[j], 0, 0, 0);
          }
        }
      }
    }
    // Results to LDS, [element][row]; FP32 D lanes hold rows 4g + r.
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int tt = w * TPW + j;
      if (tt < MT) {
#pragma unroll
        for (int r = 0; r < 4; ++r) sR[bu][c * ROWSP + 16 * tt + 4 * g + r] = acc[j][r];
      }
    }
  };

  // epi(n): chain rule in FP64 and contiguous stores.
  auto epi = [&](int64_t n, int br, int bm) {
    const TileRef tr = tile_ref(b0 + n * Gd);
    const int lim = tr.count * NP;
#pragma unroll
    for (int q = 0; q < NIT; ++q) {
      const int v = mt + q * NTM;
      if (EPI32 && v < lim) {
        const int e = v / NP, i = v - e * NP;
        const R* re = sR[br] + e * ROWSP;
        const float4* M4 = (const float4*)(sMf[bm] + e * 12);
        // m0 = (rx, ry, rz, sx), m1 = (sy, sz, tx, ty), m2 = (tz, slopes r, s, t)
        const float4 m0 = M4[0], m1 = M4[1], m2 = M4[2];
        const float ur = (float)re[i] + m2.y, us = (float)re[NP + i] + m2.z,
                    ut = (float)re[2 * NP + i] + m2.w;
        if constexpr (DIV) {
          // an FP32 result: stored as FP32, which halves the volume term's
          // traffic here and in the surface kernel that reads it
          const float ga = m0.x + m0.y + m0.z, gb = m0.w + m1.x + m1.y, gc = m1.z + m1.w + m2.x;
          __builtin_nontemporal_store(-fmaf(gc, ut, fmaf(gb, us, ga * ur)),
                                      (float*)a.ox + tr.vbase + v);
        } else {
          __builtin_nontemporal_store(fmaf(m1.z, ut, fmaf(m0.w, us, m0.x * ur)),
                                      (float*)a.ox + tr.vbase + v);
          __builtin_nontemporal_store(fmaf(m1.w, ut, fmaf(m1.x, us, m0.y * ur)),
                                      (float*)a.oy + tr.vbase + v);
          __builtin_nontemporal_store(fmaf(m2.x, ut, fmaf(m1.y, us, m0.z * ur)),
                                      (float*)a.oz + tr.vbase + v);
        }
      } else if (v < lim) {
        const int e = v / NP, i = v - e * NP;
        const R* re = sR[br] + e * ROWSP;
        const doublex2* M2 = (const doublex2*)(sM[bm] + e * 12);
        const doublex2 m01 = M2[0], m23 = M2[1], m45 = M2[2], m67 = M2[3], m8l = M2[4];
        double ur = (double)re[i], us = (double)re[NP + i], ut = (double)re[2 * NP + i];
        if constexpr (MIXED) {
          const doublex2 l23 = M2[5];
          ur += m8l.y;
          us += l23.x;
          ut += l23.y;
        }
        if constexpr (DIV) {
          const double ga = m01.x + m01.y + m23.x, gb = m23.y + m45.x + m45.y,
                       gc = m67.x + m67.y + m8l.x;
          __builtin_nontemporal_store(-(ga * ur + gb * us + gc * ut), (double*)a.ox + tr.vbase + v);
        } else {
          const double dx = m01.x * ur + m23.y * us + m67.x * ut;
          const double dy = m01.y * ur + m45.x * us + m67.y * ut;
          const double dz = m23.x * ur + m45.y * us + m8l.x * ut;
          if constexpr (PREC == kF64 || PREC == kI8) {
            __builtin_nontemporal_store(dx, (double*)a.ox + tr.vbase + v);
            __builtin_nontemporal_store(dy, (double*)a.oy + tr.vbase + v);
            __builtin_nontemporal_store(dz, (double*)a.oz + tr.vbase + v);
          } else {
            __builtin_nontemporal_store((float)dx, (float*)a.ox + tr.vbase + v);
            __builtin_nontemporal_store((float)dy, (float*)a.oy + tr.vbase + v);
            __builtin_nontemporal_store((float)dz, (float*)a.oz + tr.vbase + v);
          }
        }
      }
    }
  };

  // The two roles run separate loops with the same barrier sequence, so that
  // the operator registers of the matrix waves and the prefetch registers of
  // the memory waves are never live together.
  if (is_mma) {
    // Matrix waves issue first when both roles are ready on a SIMD (wave
    // priority, not a queue priority): 0.3-1 % faster across the plans.
    __builtin_amdgcn_s_setprio(2);
    // Operator fragments of this wave's row tiles, resident for the launch.
    Frag<PREC> fa[TPW][KS];
#pragma unroll
    for (int j = 0; j < TPW; ++j) {
      const int tt = w * TPW + j;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const size_t q = ((size_t)tt * KS + s) * 64 + lane;
        if constexpr (PREC == kF64) {
          fa[j][s].v = tt < MT ? ((const double*)a.frag)[q] : 0.0;
        } else if constexpr (PREC == kF32) {
          fa[j][s].v = tt < MT ? ((const float*)a.frag)[q] : 0.f;
        } else if constexpr (PREC == kI8) {
#pragma unroll
          for (int p = 1; p < kDig; ++p) {
            const intx4* F = (const intx4*)a.frag + (size_t)p * C::NFRAG;
            if (s < C::K64) fa[j][s].v[p] = tt < MT ? F[q] : intx4{0, 0, 0, 0};
            else fa[j][s].t[p] = tt < MT ? ((const long*)(F + q))[0] : 0L;
          }
        } else {
#pragma unroll
          for (int p = 0; p < PARTS; ++p) {
            const bf16x8* F = (const bf16x8*)a.frag + (size_t)p * C::NFRAG;
            bf16x8 z;
#pragma unroll
            for (int e = 0; e < 8; ++e) z[e] = (__bf16)0.0f;
            fa[j][s].v[p] = tt < MT ? F[q] : z;
          }
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if constexpr (MIXED) __syncthreads();
    for (int64_t n0 = 0; n0 <= nb; n0 += 2) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int64_t n = n0 + d;
        if (n > nb) break;
        if (n < nb) mma(d & 1, fa);
        __syncthreads();
      }
    }
  } else {
    if constexpr (MIXED) {
      for (int v = mt; v < 3 * NP; v += NTM) sLam[v] = a.lam[v];
    }
    if constexpr (PREC == kI8) {
      for (int v = mt; v < MT * 64; v += NTM) sDel[v] = a.rowcorr[v];
      for (int v = mt; v < MT * 16; v += NTM) sRx[v] = a.rowexp[v];
      // digit 0 slots [tile][kstep][lane] (16 bytes; the tail step uses 8)
      const intx4* F0 = (const intx4*)a.frag;
      for (int v = mt; v < MT * KS * 64; v += NTM) {
        const int tt = v / (KS * 64), ks = (v / 64) % KS, l = v % 64;
        if (ks < C::K64) sOpx[(tt * C::K64 + ks) * 64 + l] = F0[v];
        else if constexpr (C::T32) sOpt[tt * 64 + l] = ((const long*)(F0 + v))[0];
      }
    }
    doublex2 pre[PD][NL];
    double pm[PD];  // metrics: memory thread v < 144 holds row v / 16 of element v % 16
#pragma unroll
    for (int d = 0; d < PD; ++d)
      if (d < nb) prefetch(d, pre[d], pm[d]);
    // prologue: tiles 0 .. LEAD-1 staged, tile 0 linearised below FP64
#pragma unroll
    for (int d = 0; d < LEAD; ++d)
      if (d < nb) stage(d, d & 1, d, pre[d & 1], pm[d & 1]);
    __syncthreads();
    if constexpr (MIXED) {
      lin(0, 0);
      __syncthreads();
    }
    // Phase n; unrolled by 2 so that the prefetch slot and the parity-indexed
    // buffers are compile-time constants (PD = 2).
    for (int64_t n0 = 0; n0 <= nb; n0 += 2) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int64_t n = n0 + d;
        if (n > nb) break;
        const int n4 = (int)(n & 3);
        if (n + LEAD < nb)
          stage(n + LEAD, (d + LEAD) & 1, (n4 + LEAD) & 3, pre[(d + LEAD) & 1], pm[(d + LEAD) & 1]);
        if constexpr (MIXED) {
          if (n + 1 < nb) lin((d + 1) & 1, (n4 + 1) & 3);
        }
        if (n >= 1) epi(n - 1, (d + 1) & 1, (n4 + 3) & 3);
        __syncthreads();
      }
    }
  }
}

}  // namespace

struct dgk_deriv_plan_s {
  dgk_device dev = nullptr;
  int N = 0, prec = 0;
  void* frag = nullptr;
  double* lam = nullptr;
  int* rowexp = nullptr;  // I8X6 row scales
  double* rowcorr = nullptr;  // I8X6 [MT*16][4] operator on the linear part minus exact
  int vtx[4] = {0, 0, 0, 0};
  int blocks = 0;  // persistent grid size
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

// Persistent grid: blocks resident per CU (the same for both modes) times
// the CU count.
template <int N, int PREC>
int occupancy_blocks(dgk_deriv_plan pl) {
  int per_cu = 0;
  DGK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, deriv_affine_kernel<N, PREC, false>, ACfg<N, PREC>::NT, 0));
  if (per_cu < 1) per_cu = 1;
  pl->blocks = pl->dev->cu_count * per_cu;
  return DGK_OK;
}

template <int N, int PREC, bool DIV>
int launch_affine_t(dgk_deriv_plan pl, const AffineArgs& a) {
  const int64_t ntiles = a.tiles ? a.ntiles : (a.K + 15) / 16;
  const int64_t blocks = pl->blocks < ntiles ? pl->blocks : ntiles;
  if (blocks < 1) return DGK_OK;
  hipLaunchKernelGGL((deriv_affine_kernel<N, PREC, DIV>), dim3((unsigned)blocks),
                     dim3(ACfg<N, PREC>::NT), 0, pl->dev->stream, a);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

// what: 0 size the persistent grid, 1 gradient launch, 2 divergence launch.
template <int N, int PREC>
int dispatch_p(dgk_deriv_plan pl, const AffineArgs* a, int what) {
  if (what == 0) return occupancy_blocks<N, PREC>(pl);
  return what == 1 ? launch_affine_t<N, PREC, false>(pl, *a) : launch_affine_t<N, PREC, true>(pl, *a);
}

template <int N>
int dispatch_n(dgk_deriv_plan pl, const AffineArgs* a, int what) {
  switch (pl->prec) {
    case kF64: return dispatch_p<N, kF64>(pl, a, what);
    case kF32: return dispatch_p<N, kF32>(pl, a, what);
    case kBF3: return dispatch_p<N, kBF3>(pl, a, what);
    case kI8: return dispatch_p<N, kI8>(pl, a, what);
    default: return dispatch_p<N, kBF6>(pl, a, what);
  }
}

int dispatch(dgk_deriv_plan pl, const AffineArgs* a, int what) {
  switch (pl->N) {
    case 1: return dispatch_n<1>(pl, a, what);
    case 2: return dispatch_n<2>(pl, a, what);
    case 3: return dispatch_n<3>(pl, a, what);
    case 4: return dispatch_n<4>(pl, a, what);
    case 5: return dispatch_n<5>(pl, a, what);
    case 6: return dispatch_n<6>(pl, a, what);
  }
  dgk::set_error("dgk_phys_deriv_affine: N=%d", pl->N);
  return DGK_ERR_UNSUPPORTED;
}

}  // namespace

extern "C" int dgk_deriv_plan_create(dgk_device dev, int N, const double* Dr, const double* Ds,
                                     const double* Dt, const double* r, const double* s,
                                     const double* t, int prec, dgk_deriv_plan* out) {
  dgk::clear_error();
  DGK_REQUIRE(dev && Dr && Ds && Dt && out, "dgk_deriv_plan_create: NULL argument");
  DGK_REQUIRE(N >= 1 && N <= 6, "dgk_deriv_plan_create: N=%d outside 1..6", N);
  DGK_REQUIRE(prec >= kF64 && prec <= kI8, "dgk_deriv_plan_create: precision %d", prec);
  DGK_REQUIRE(prec == kF64 || (r && s && t),
              "dgk_deriv_plan_create: node coordinates r, s, t are required below FP64");
  *out = nullptr;
  const HostCfg hc = host_cfg(N, prec);
  const int NP = hc.NP;
  const double* D[3] = {Dr, Ds, Dt};
  auto val = [&](int row, int k) -> double {
    const int blk = row / NP, i = row % NP;
    if (blk >= 3 || k >= NP) return 0.0;
    return D[blk][i * NP + k];
  };
  const size_t nfrag = (size_t)hc.MT * hc.KS * 64;
  std::vector<unsigned char> buf;
  std::vector<int> rowexp;
  if (prec == kI8) {
    // Row r scaled by 2^-E_r (|A_rk| < 2^E_r), rounded to kBits bits, written
    // as kDig balanced base-256 digits. Digit a of (row, k) sits in the
    // 16-byte slot [a][tile][kstep][lane], lane = row % 16 + 16 g: byte
    // k % 16 with g = (k % 64) / 16 for a 16x16x64 step, byte k % 8 with
    // g = (k % 32) / 8 for the 16x16x32 tail (the same K placement the
    // kernel uses for w).
    const int K64 = i8_k64(NP);
    rowexp.assign((size_t)hc.MT * 16, 0);
    buf.assign(nfrag * 16 * kDig, 0);
    for (int row = 0; row < hc.MT * 16; ++row) {
      double mx = 0.0;
      for (int k = 0; k < NP; ++k) mx = std::max(mx, std::fabs(val(row, k)));
      int E = 0;
      if (mx > 0.0) (void)std::frexp(mx, &E);
      rowexp[row] = E;
      const int tt = row / 16, lr = row % 16;
      for (int k = 0; k < NP; ++k) {
        long long X = std::llrint(std::ldexp(val(row, k), kBits - E));
        const int ks = k < 64 * K64 ? k / 64 : K64;
        const int l = ks < K64 ? lr + 16 * ((k % 64) / 16) : lr + 16 * ((k - 64 * K64) / 8);
        const int b = ks < K64 ? k % 16 : (k - 64 * K64) % 8;
        for (int d = 0; d < kDig; ++d) {
          long long v = X & 255, cy = v >> 7;
          const int dg = d < kDig - 1 ? (int)(v - (cy << 8)) : (int)X;
          if (d < kDig - 1) X = (X >> 8) + cy;
          DGK_REQUIRE(dg >= -128 && dg <= 127, "dgk_deriv_plan_create: digit out of range");
          const size_t q = (((size_t)d * hc.MT + tt) * hc.KS + ks) * 64 + l;
          buf[q * 16 + b] = (unsigned char)(signed char)dg;
        }
      }
    }
  } else if (prec == kF64 || prec == kF32) {
    buf.resize(nfrag * (prec == kF64 ? 8 : 4));
    for (int tt = 0; tt < hc.MT; ++tt)
      for (int ks = 0; ks < hc.KS; ++ks)
        for (int l = 0; l < 64; ++l) {
          const double v = val(16 * tt + (l & 15), 4 * ks + (l >> 4));
          const size_t q = ((size_t)tt * hc.KS + ks) * 64 + l;
          if (prec == kF64) ((double*)buf.data())[q] = v;
          else ((float*)buf.data())[q] = (float)v;
        }
  } else {
    buf.resize(nfrag * 16 * hc.PARTS);
    uint16_t* P0 = (uint16_t*)buf.data();
    for (int tt = 0; tt < hc.MT; ++tt)
      for (int ks = 0; ks < hc.KS; ++ks)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e) {
            const double v = val(16 * tt + (l & 15), 32 * ks + 8 * (l >> 4) + e);
            const size_t q = (((size_t)tt * hc.KS + ks) * 64 + l) * 8 + e;
            const uint16_t h = bf16_bits((float)v);
            const double r1 = v - (double)bf16_to_float(h);
            const uint16_t m = bf16_bits((float)r1);
            P0[q] = h;
            P0[nfrag * 8 + q] = m;
            if (hc.PARTS == 3) {
              const double r2 = r1 - (double)bf16_to_float(m);
              P0[2 * nfrag * 8 + q] = bf16_bits((float)r2);
            }
          }
  }
  // Vertex nodes and barycentric weights (reference tet (-1,-1,-1), (1,-1,-1),
  // (-1,1,-1), (-1,-1,1)) for the linear-part subtraction.
  std::vector<double> lam;
  int vtx[4] = {0, 0, 0, 0};
  if (prec != kF64) {
    const double V[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    for (int v = 0; v < 4; ++v) {
      int found = -1;
      for (int i = 0; i < NP; ++i)
        if (std::fabs(r[i] - V[v][0]) < 1e-10 && std::fabs(s[i] - V[v][1]) < 1e-10 &&
            std::fabs(t[i] - V[v][2]) < 1e-10)
          found = i;
      DGK_REQUIRE(found >= 0, "dgk_deriv_plan_create: vertex %d of the reference tet is not a node",
                  v);
      vtx[v] = found;
    }
    lam.resize(3 * NP);
    for (int i = 0; i < NP; ++i) {
      lam[i] = 0.5 * (1.0 + r[i]);
      lam[NP + i] = 0.5 * (1.0 + s[i]);
      lam[2 * NP + i] = 0.5 * (1.0 + t[i]);
    }
  }
  // I8X6: Lu = a0 + a1 l1 + a2 l2 + a3 l3 enters as its exact slopes (a1, a2,
  // a3)/2 in the epilogue; the stored operator gives D Lu, which differs by
  // its round-off (D 1 is ~1e-11, not 0, at P=6). The row corrections
  // [D 1, D l1 - e1, D l2 - e2, D l3 - e3] restore the stored operator's
  // action, so the plan agrees with FP64 to its own rounding.
  std::vector<double> rowcorr;
  if (prec == kI8) {
    rowcorr.assign((size_t)hc.MT * 64, 0.0);
    for (int row = 0; row < 3 * NP; ++row) {
      const int blk = row / NP;
      long double c[4] = {0, 0, 0, 0};
      for (int k = 0; k < NP; ++k) {
        const long double v = val(row, k);
        c[0] += v;
        c[1] += v * lam[k];
        c[2] += v * lam[NP + k];
        c[3] += v * lam[2 * NP + k];
      }
      c[1 + blk] -= 0.5L;
      for (int q = 0; q < 4; ++q) rowcorr[(size_t)row * 4 + q] = (double)c[q];
    }
  }
  dgk_deriv_plan p = new dgk_deriv_plan_s();
  p->dev = dev;
  p->N = N;
  p->prec = prec;
  std::memcpy(p->vtx, vtx, sizeof vtx);
  int rc = dgk_malloc(dev, buf.size(), buf.data(), &p->frag);
  if (rc == DGK_OK && !lam.empty())
    rc = dgk_malloc(dev, lam.size() * sizeof(double), lam.data(), (void**)&p->lam);
  if (rc == DGK_OK && !rowexp.empty())
    rc = dgk_malloc(dev, rowexp.size() * sizeof(int), rowexp.data(), (void**)&p->rowexp);
  if (rc == DGK_OK && !rowcorr.empty())
    rc = dgk_malloc(dev, rowcorr.size() * sizeof(double), rowcorr.data(), (void**)&p->rowcorr);
  if (rc == DGK_OK) rc = dispatch(p, nullptr, 0);
  if (rc != DGK_OK) {
    if (p->frag) dgk_free(dev, p->frag);
    if (p->lam) dgk_free(dev, p->lam);
    if (p->rowexp) dgk_free(dev, p->rowexp);
    if (p->rowcorr) dgk_free(dev, p->rowcorr);
    delete p;
    return rc;
  }
  *out = p;
  return DGK_OK;
}

extern "C" int dgk_deriv_plan_destroy(dgk_deriv_plan p) {
  if (!p) return DGK_OK;
  dgk_free(p->dev, p->frag);
  if (p->lam) dgk_free(p->dev, p->lam);
  if (p->rowexp) dgk_free(p->dev, p->rowexp);
  if (p->rowcorr) dgk_free(p->dev, p->rowcorr);
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
  DGK_REQUIRE(((uintptr_t)u_dev & 15) == 0, "dgk_phys_deriv_affine: u must be 16-byte aligned");
  if (K == 0) return DGK_OK;
  AffineArgs a;
  a.K = K;
  a.tiles = nullptr;
  a.ntiles = 0;
  a.frag = p->frag;
  a.lam = p->lam;
  a.rowexp = p->rowexp;
  a.rowcorr = p->rowcorr;
  std::memcpy(a.vtx, p->vtx, sizeof a.vtx);
  a.u = u_dev;
  a.geo = geo_dev;
  a.gs = geo_stride;
  a.ox = dudx_dev;
  a.oy = dudy_dev;
  a.oz = dudz_dev;
  return dispatch(p, &a, 1);
}

bool dgk::deriv_plan_divergence_f32(dgk_deriv_plan p) {
  return p->prec == kF32 || p->prec == kBF3;
}

int dgk::deriv_plan_divergence(dgk_deriv_plan p, const dgk::Tile* tiles, int ntiles, const double* u,
                               const double* geo, int64_t gs, void* out) {
  if (ntiles == 0) return DGK_OK;
  AffineArgs a;
  a.K = 0;
  a.tiles = tiles;
  a.ntiles = ntiles;
  a.frag = p->frag;
  a.lam = p->lam;
  a.rowexp = p->rowexp;
  a.rowcorr = p->rowcorr;
  std::memcpy(a.vtx, p->vtx, sizeof a.vtx);
  a.u = u;
  a.geo = geo;
  a.gs = gs;
  a.ox = out;
  a.oy = out;
  a.oz = out;
  return dispatch(p, &a, 2);
}
