// Shock capturing for the tetrahedral Burgers stage on gfx950: the
// Persson-Peraire modal-decay sensor, artificial viscosity div(eps grad u) in
// BR1 form, and a learned sensor (MLP over element modal coefficients) with
// its training kernels.
//
// None of this exists in the reference (it has no shock capturing); the
// learned-viscosity workload defines it. Written from:
//  * Persson & Peraire, AIAA 2006-112: S = sum_{deg m = N} uh_m^2 /
//    sum_m uh_m^2 over the orthonormal modal coefficients uh = V^-1 u,
//    s = log10 S, eps = eps0 * ramp(s; s0, kappa) with the sine ramp, and
//    eps0 = ceps h lambda / N, h = (6 vol)^(1/3), lambda = sqrt(3) max|u|
//    (the largest |f'(u).n| for f(u) = u^2/2 (1,1,1));
//  * BR1 (Bassi & Rebay 1997; Hesthaven & Warburton 2008 §7.1-7.2), strong
//    form with central fluxes and a jump penalty:
//      q   = grad u + LIFT(Fscale n (u+ - u-)/2),
//      L_v = div(eps q) - LIFT(Fscale (n.(s- - s+)/2 + tau (u- - u+))),
//      s = eps q, tau = tau0 max(eps-, eps+).
// A viscous stage (Burgers + L_v + SSP-RK3 update) is two launches: a
// volume kernel (BR1 gradient on matrix cores, kept in registers; the
// divergence of eps q minus that of u^2/2 (1,1,1); the face traces eps n.q)
// and a surface kernel (LF and viscous flux jumps, one face per lane; LIFT on
// matrix cores; the RK update).
//
// Kernel shapes. The Persson-Peraire sensor takes one element per lane, 64
// elements (four 16-element tiles) per single-wave block; the tile's nodal
// values are staged through LDS with unit-stride loads, and V^-1 is read
// with wave-uniform addresses (scalar loads, SGPR operands of v_fma_f64).
// The stage kernels and the learned sensor work on 16-element tiles per wave
// on matrix cores. The learned sensor: modal transform on split BF16 MFMA
// (three products), both hidden layers on BF16 MFMA (16x16x32), FP32
// accumulation. Hidden activations never leave registers between layers:
// the k order of each layer's product is permuted so that the FP32 result
// layout of one layer (lane (c, g) holds rows 4g + r of each 16-row tile) is
// directly the BF16 B operand of the next (lane (c, g) holds k = 8g + j of a
// 32-wide k step).
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dgmesh_internal.h"

namespace {

using dgk::FmaskTable;
using dgk::Tet;
using dgk::Tile;

typedef const __attribute__((address_space(4))) double* cdouble_ptr;
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

constexpr int kH = 64;  // hidden width of the sensor network

// Modal degree of each column of Vandermonde3D (i, j, k nested, degree
// i + j + k; element/library/gonudg/vandermonde3d.go:15-41).
template <int N>
struct ModeDeg {
  int d[Tet<N>::Np];
  constexpr ModeDeg() : d{} {
    int sk = 0;
    for (int i = 0; i <= N; ++i)
      for (int j = 0; j <= N - i; ++j)
        for (int k = 0; k <= N - i - j; ++k) d[sk++] = i + j + k;
  }
};

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double pp_fraction(double S, double s0, double kappa) {
  if (!(S > 0.0)) return 0.0;
  const double sv = log10(S);
  if (sv < s0 - kappa) return 0.0;
  if (sv > s0 + kappa) return 1.0;
  return 0.5 * (1.0 + sin(M_PI * (sv - s0) / (2.0 * kappa)));
}

struct AvArgs {
  const Tile* tiles;  // 16-element tiles, full tiles first
  int ntiles;
  int64_t K;
  const double* u;
  const double* geo;          // [26][K]
  const double* Dr;
  const double* Ds;
  const double* Dt;
  const double* LIFT;
  const double* Vinv;         // Np x Np row-major (modal transform)
  const uint8_t* face_nodes;  // [4][24][Nfp]
  const int32_t* nbr;         // [4][K] neighbour value offsets
  const int32_t* nbe;         // [4][K] neighbour element indices
  const int32_t* codes;       // [K]
  double* vol;                // state layout: volume terms of the stage RHS
  double* sig;                // [K][4][Nfp]: eps n.q at each own face point
  double* eps;                // [K]
  const double* eps_in;
  double* rhs;                // state layout, += L_v
  double* frac;               // [K] or NULL
  float* frac32;              // [K] or NULL
  float* feats;               // [K][Np] or NULL
  double* maxeps;             // atomic max of eps, or NULL
  double s0, kappa, ceps, tau0;
};

// The wave's four tiles: lane l belongs to tile 4 grp + l / 16, element
// l % 16 of it (count 0 past the list).
__device__ __forceinline__ Tile group_tile(const Tile* tiles, int ntiles, int grp, int lane) {
  Tile t;
  const int ti = 4 * grp + (lane >> 4);
  if (ti < ntiles) {
    t = tiles[ti];
  } else {
    t.vbase = 0;
    t.gbase = 0;
    t.count = 0;
  }
  return t;
}

// A tile record loaded by every lane, in scalar registers (it is wave-uniform).
__device__ __forceinline__ Tile tile_uniform(const Tile& v) {
  Tile u;
  u.vbase = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v.vbase >> 32)) << 32) |
                      (uint32_t)__builtin_amdgcn_readfirstlane((int)v.vbase));
  u.gbase = __builtin_amdgcn_readfirstlane(v.gbase);
  u.count = __builtin_amdgcn_readfirstlane(v.count);
  return u;
}

// Stage the nodal values of the wave's four tiles of src into LDS rows of S
// doubles (row 16 q + e for element e of tile q), unit-stride loads.
template <int NP, int S>
__device__ __forceinline__ void stage4(const double* __restrict__ src, const Tile& my, int lane,
                                       double* sRow) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long long vb = __shfl((long long)my.vbase, 16 * q, 64);
    const int cnt = __shfl(my.count, 16 * q, 64);
    for (int i = lane; i < cnt * NP; i += 64) {
      const int e = i / NP, n = i - e * NP;
      sRow[(16 * q + e) * S + n] = src[vb + i];
    }
  }
}

// dst[tile values] (=|+=) LDS rows, unit-stride stores.
template <int NP, int S, bool ADD>
__device__ __forceinline__ void unstage4(double* __restrict__ dst, const Tile& my, int lane,
                                         const double* sRow) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long long vb = __shfl((long long)my.vbase, 16 * q, 64);
    const int cnt = __shfl(my.count, 16 * q, 64);
    for (int i = lane; i < cnt * NP; i += 64) {
      const int e = i / NP, n = i - e * NP;
      const double v = sRow[(16 * q + e) * S + n];
      if (ADD) dst[vb + i] += v;
      else dst[vb + i] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Matrix-core kernels of the viscous stage (FP64, v_mfma_f64_16x16x4_f64).
// Each wave takes 16-element tiles (the MFMA N dimension); the operator
// fragments sit in LDS for the block. FP64 MFMA layout (checked on gfx950):
// lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[l/16 + 4r][l%16], r = 0..3, so with blocks stacked at multiples of 4
// rows the lane (c, g) holding row i = g + 4q of one block holds row i of
// every block, for element c.
//
// The BR1 gradient is q_d = grad_d u + sum_f (Fscale_f n_d,f / 2) T_f with
// [Dr; Ds; Dt] u (stacked rows, K = Np) and T_f = LIFT_f (u+ - u-) per face
// (face-contiguous k with each face padded to a multiple of 4 points, one
// accumulator set per face); the divergence is one product [Dr | Ds | Dt]
// [S_r; S_s; S_t]; the surface term one LIFT product over face-per-lane k.
// ---------------------------------------------------------------------------
typedef double double4v __attribute__((ext_vector_type(4)));

// One k = 4 step of a 16-row operator tile, frag = the tile's 64-entry
// fragment table in the 16x16x4 FP64 A-operand lane order. A tile whose only
// rows are its first four (the last row tile at Np = 35, 20, 4, 84) runs as
// one v_mfma_f64_4x4x4_4b, which issues about eight times faster than the
// 16x16x4 on gfx950. Its block j is lanes 16 k + 4 j + i (A[i][k], B[k][i];
// D[i][n] at lane 16 i + 4 j + n, checked on gfx950): B is the 16x16x4 B
// operand unchanged, lane l reads A row l%4, column l/16 from lane
// 16 (l/16) + l%4 of the table, and the result lands in register 0's place
// (row l/16, element l%16); registers 1-3 stay zero.
__device__ __forceinline__ void mfma_step(double4v& acc, const double* frag, int lane, double b,
                                          bool four_rows) {
  if (four_rows)
    acc[0] = __builtin_amdgcn_mfma_f64_4x4x4f64(frag[16 * (lane >> 4) + (lane & 3)], b, acc[0], 0,
                                                0, 0);
  else
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[lane], b, acc, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Persson-Peraire sensor on FP64 matrix cores. Per 16-element tile each wave
// forms the modal coefficients uh = V^-1 u with v_mfma_f64_16x16x4_f64 (the
// tile's elements are the N dimension; the V^-1 fragments sit in LDS for the
// block). Lane (c, g) then holds modes 16 tt + g + 4 r of element c, so the
// mode energies and max|u| of an element reduce over its four lanes. The
// nodal values and volume of the next tile are loaded during the product.
// ---------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(256) pp_sensor_kernel(AvArgs a) {
  constexpr int NP = Tet<N>::Np, MT = (NP + 15) / 16, KS = (NP + 3) / 4;
  constexpr ModeDeg<N> MD{};
  __shared__ __attribute__((aligned(16))) double sV[MT * KS * 64];
  for (int q = threadIdx.x; q < MT * KS * 64; q += 256) {
    const int l = q & 63, ts = q >> 6, tt = ts / KS, s = ts % KS;
    const int row = 16 * tt + (l & 15), k = 4 * s + (l >> 4);
    sV[q] = row < NP && k < NP ? a.Vinv[row * NP + k] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, g = lane >> 4;
  // modes of top degree among the rows this lane holds (bit 4 tt + r)
  uint32_t topm = 0;
#pragma unroll
  for (int tt = 0; tt < MT; ++tt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * tt + g + 4 * r;
      if (row < NP && MD.d[row] == N) topm |= 1u << (4 * tt + r);
    }
  // The nodal values of tile d with unit-stride loads (staged in LDS for the
  // B operands) and the volume factor of element c, a tile ahead.
  constexpr int PV = (16 * NP + 63) / 64, S = NP + 1;
  __shared__ double sUall[4][16 * S];
  double* sU = sUall[w];
  double pre[PV], jn = 0.0;
  auto issue = [&](const Tile& d) {
#pragma unroll
    for (int j = 0; j < PV; ++j) pre[j] = a.u[d.vbase + min(lane + 64 * j, d.count * NP - 1)];
    jn = a.geo[9 * a.K + d.gbase + (c < d.count ? c : d.count - 1)];
  };
  const int t0 = blockIdx.x * 4 + w, ts = gridDim.x * 4, t1 = a.ntiles;
  double emax = 0.0;
  Tile tile_n{}, tile_raw{};
  if (t0 < t1) {
    tile_n = tile_uniform(a.tiles[t0]);
    issue(tile_n);
    if (t0 + ts < t1) tile_raw = a.tiles[t0 + ts];
  }
  for (int t = t0; t < t1; t += ts) {
    const Tile tile = tile_n;
    const double jv = jn;
    wave_fence();  // previous readers of the tile buffer are done
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      if (i < 16 * NP) {
        const int e = i / NP, nn = i - e * NP;
        sU[e * S + nn] = i < tile.count * NP ? pre[j] : 0.0;
      }
    }
    if (t + ts < t1) {
      tile_n = tile_uniform(tile_raw);
      issue(tile_n);
      if (t + 2 * ts < t1) tile_raw = a.tiles[t + 2 * ts];
    }
    wave_fence();
    // B operand of step s: node 4 s + g of element c (zero past Np)
    double b[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) b[s] = 4 * s + g < NP ? sU[c * S + 4 * s + g] : 0.0;
    double4v acc[MT];
#pragma unroll
    for (int tt = 0; tt < MT; ++tt) acc[tt] = (double4v){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int tt = 0; tt < MT; ++tt)
        mfma_step(acc[tt], sV + (tt * KS + s) * 64, lane, b[s], NP - 16 * tt <= 4);
    double umax = 0.0, tot = 0.0, top = 0.0;
#pragma unroll
    for (int s = 0; s < KS; ++s) umax = fmax(umax, fabs(b[s]));
#pragma unroll
    for (int tt = 0; tt < MT; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double e2 = acc[tt][r] * acc[tt][r];  // rows past Np are zero
        tot += e2;
        if (topm >> (4 * tt + r) & 1u) top += e2;
      }
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) {
      umax = fmax(umax, __shfl_xor(umax, o, 64));
      tot += __shfl_xor(tot, o, 64);
      top += __shfl_xor(top, o, 64);
    }
    const bool own = c < tile.count;
    const int64_t gi = (int64_t)tile.gbase + c;
    if (own && g == 0) {
      const double Sv = tot > 0.0 ? top / tot : 0.0;
      const double fr = pp_fraction(Sv, a.s0, a.kappa);
      const double h = cbrt(6.0 * (4.0 / 3.0) * jv);
      const double e = a.ceps * h * sqrt(3.0) * umax / N * fr;
      a.eps[gi] = e;
      emax = fmax(emax, e);
      if (a.frac) a.frac[gi] = fr;
      if (a.frac32) a.frac32[gi] = (float)fr;
    }
    if (a.feats && own) {
      const double inv = tot > 0.0 ? 1.0 / sqrt(tot) : 0.0;
#pragma unroll
      for (int tt = 0; tt < MT; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * tt + g + 4 * r;
          if (row < NP) a.feats[gi * NP + row] = (float)(acc[tt][r] * inv);
        }
    }
  }
  if (a.maxeps) {
    emax = wmax(emax);
    if (lane == 0 && emax > 0.0)
      atomicMax((unsigned long long*)a.maxeps, (unsigned long long)__double_as_longlong(emax));
  }
}

template <int N>
struct AvShape {
  static constexpr int NP = Tet<N>::Np, NFP = Tet<N>::Nfp, NF = 4 * NFP;
  static constexpr int NPR = (NP + 3) / 4 * 4;      // block row/column offset
  static constexpr int QN = NPR / 4;                // rows per lane
  static constexpr int MTG = (3 * NPR + 15) / 16;   // row tiles of [Dr; Ds; Dt]
  static constexpr int KSG = NPR / 4;               // k steps over Np
  static constexpr int MTL = (NP + 15) / 16;        // row tiles of Np rows
  static constexpr int SPF = (NFP + 3) / 4;         // k steps per (padded) face
  static constexpr int KSD = 3 * NPR / 4;           // k steps of [Dr | Ds | Dt]
  static constexpr int KSL = NFP;                   // k steps, face-per-lane LIFT
  static constexpr int S = NP + 1;                  // LDS row stride
  static constexpr int NA_G = (MTG * KSG + MTL * 4 * SPF) * 64;
};

// Own volume node of face point fp of face f, fp = f4 + g (clamped).
template <int N>
__device__ __forceinline__ int face_node(int f, int f4, int g) {
  constexpr FmaskTable<N> FM{};
  constexpr int NFP = Tet<N>::Nfp;
  const int n0 = FM.v[f][min(f4 + 0, NFP - 1)], n1 = FM.v[f][min(f4 + 1, NFP - 1)];
  const int n2 = FM.v[f][min(f4 + 2, NFP - 1)], n3 = FM.v[f][min(f4 + 3, NFP - 1)];
  return g == 0 ? n0 : g == 1 ? n1 : g == 2 ? n2 : n3;
}

// MODE 0: out = a0 u0 + a1 (u + dt R); 1: out = a1 (u + dt R); 2: out = R,
// with R = L_v(u) + (BURG ? L(u) : 0).
struct StageAv {
  const double* u0;
  double* out;
  double a0, a1;
  const double* dt;
  double* maxabs;
  const double* bgeo;  // [19][K]
};

// ---------------------------------------------------------------------------
// Viscous stage in two launches (the gradient never goes to memory):
// av_vol_kernel: per 16-element tile, the BR1 gradient q on matrix cores
//   (kept in registers), then
//     vol = [Dr | Ds | Dt] [S_r; S_s; S_t],  S_b = eps (b.q) - g_b u^2/2 (BURG)
//   (the FP64 result layout of q -- lane (c, g) holds nodes g + 4 i -- is
//   exactly the B layout of this product, so q feeds it from registers),
//   and the face traces sig[f][fp] = eps n_f . q at own face points (q goes
//   through LDS one component at a time).
// av_surf_kernel: lanes along the (element, point) pairs of one face at a
//   time gather the own and neighbour traces of u and sig (two faces in
//   flight; the next tile's first face during this tile's products) and form
//     -Fscale ((sig- + sig+)/2 + tau (u- - u+))  [+ LF jump (BURG)]
//   (sig+ is the neighbour's own eps n'.q with n' = -n, so the central flux
//   n.(eps q)- - n.(eps q)+ is sig- + sig+; a boundary face, its own
//   neighbour, has zero jump) into LDS, from which each lane reads its B
//   operands; then LIFT on matrix cores per 16-row tile (columns face by
//   face) and the RK update through a 16 x 17 LDS buffer with row-contiguous
//   accesses.
// ---------------------------------------------------------------------------
template <int N, int WPB, bool BURG>
__global__ void __launch_bounds__(64 * WPB, WPB >= 8 ? 2 : 1)
    av_vol_kernel(AvArgs a, const double* __restrict__ bgeo) {
  using SH = AvShape<N>;
  constexpr int NP = SH::NP, NFP = SH::NFP, NF = SH::NF, NPR = SH::NPR, QN = SH::QN;
  constexpr int MTG = SH::MTG, KSG = SH::KSG, MTL = SH::MTL, SPF = SH::SPF, S = SH::S;
  constexpr int KSD = SH::KSD;
  constexpr int NT = 64 * WPB;
  constexpr int NAD = MTL * KSD * 64;  // [Dr | Ds | Dt] fragments
  constexpr FmaskTable<N> FM{};
  __shared__ __attribute__((aligned(16))) double sA[SH::NA_G + NAD];
  __shared__ __attribute__((aligned(16))) double sUall[WPB][16 * S];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFM[NF];
  for (int i = threadIdx.x; i < 4 * 24 * NFP; i += NT) sFN[i] = a.face_nodes[i];
  for (int i = threadIdx.x; i < NF; i += NT) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
  for (int q = threadIdx.x; q < SH::NA_G + NAD; q += NT) {
    const int l = q & 63, ts = q >> 6, cl = l & 15, gl = l >> 4;
    double v = 0.0;
    if (ts < MTG * KSG) {  // [Dr; Ds; Dt] for the gradient
      const int tt = ts / KSG, s = ts % KSG;
      const int row = 16 * tt + cl, k = 4 * s + gl, blk = row / NPR, i = row % NPR;
      if (k < NP && blk < 3 && i < NP) v = (blk == 0 ? a.Dr : blk == 1 ? a.Ds : a.Dt)[i * NP + k];
    } else if (ts < MTG * KSG + MTL * 4 * SPF) {  // LIFT per face, padded faces
      const int r2 = ts - MTG * KSG, tt = r2 / (4 * SPF), s = r2 % (4 * SPF);
      const int row = 16 * tt + cl, f = s / SPF, fp = 4 * (s % SPF) + gl;
      if (row < NP && fp < NFP) v = a.LIFT[row * NF + f * NFP + fp];
    } else {  // [Dr | Ds | Dt] for the divergence
      const int r3 = ts - MTG * KSG - MTL * 4 * SPF, tt = r3 / KSD, s = r3 % KSD;
      const int row = 16 * tt + cl, k = 4 * s + gl, blk = k / NPR, j = k % NPR;
      if (row < NP && j < NP) v = (blk == 0 ? a.Dr : blk == 1 ? a.Ds : a.Dt)[row * NP + j];
    }
    sA[q] = v;
  }
  __syncthreads();
  const double* sAv = sA;
  const double* sAl = sA + MTG * KSG * 64;
  const double* sAd = sA + SH::NA_G;
  // wave index as a scalar, so that tile indices and records stay in SGPRs
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, g = lane >> 4;
  double* sU = sUall[w];
  const int64_t K = a.K;
  const int nw = gridDim.x * WPB;
  for (int t = blockIdx.x * WPB + w; t < a.ntiles; t += nw) {
    const Tile tl = tile_uniform(a.tiles[t]);
    const int cnt = tl.count;
    wave_fence();
    for (int i = lane; i < cnt * NP; i += 64) {
      const int e = i / NP, n = i - e * NP;
      sU[e * S + n] = a.u[tl.vbase + i];
    }
    wave_fence();
    const int ec = min(c, cnt - 1);
    const int64_t ge = (int64_t)tl.gbase + ec;
    const double* su = sU + ec * S;
    const int packed = a.codes[ge];
    int32_t nb[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) nb[f] = a.nbr[f * K + ge];
    // Every other load of the tile is issued here, before the products that
    // cover its latency: the neighbour traces of the gradient's face jumps
    // (they need the neighbour offsets), the metrics, the face normals and
    // Fscale, eps and the folded metrics.
    double uPv[4][SPF];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int code = (packed >> (5 * f)) & 31;
#pragma unroll
      for (int sp = 0; sp < SPF; ++sp) {
        const int fpc = min(4 * sp + g, NFP - 1);
        uPv[f][sp] = a.u[nb[f] + sFN[(f * 24 + code) * NFP + fpc]];
      }
    }
    double mt[9], nrm[4][3], hfs[4];
#pragma unroll
    for (int q = 0; q < 9; ++q) mt[q] = a.geo[q * K + ge];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      hfs[f] = 0.5 * a.geo[(13 + 4 * f) * K + ge];
#pragma unroll
      for (int d = 0; d < 3; ++d) nrm[f][d] = a.geo[(10 + 4 * f + d) * K + ge];
    }
    const double ee = a.eps_in[ge];
    double gh[3] = {0.0, 0.0, 0.0};
    if (BURG) {
#pragma unroll
      for (int d = 0; d < 3; ++d) gh[d] = 0.5 * bgeo[d * K + ge];
    }
    // Gradient: [Dr; Ds; Dt] u and LIFT_f (u+ - u-) per face.
    double4v accV[MTG];
#pragma unroll
    for (int tt = 0; tt < MTG; ++tt) accV[tt] = (double4v){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KSG; ++s) {
      const int k = 4 * s + g;
      const double b = k < NP ? su[k] : 0.0;
#pragma unroll
      for (int tt = 0; tt < MTG; ++tt)
        accV[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sAv[(tt * KSG + s) * 64 + lane], b,
                                                         accV[tt], 0, 0, 0);
    }
    double4v accF[4][MTL];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
#pragma unroll
      for (int tt = 0; tt < MTL; ++tt) accF[f][tt] = (double4v){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int sp = 0; sp < SPF; ++sp) {
        const int fp = 4 * sp + g;
        const double uM = su[sFM[f * NFP + min(fp, NFP - 1)]];
        const double b = fp < NFP ? uPv[f][sp] - uM : 0.0;
#pragma unroll
        for (int tt = 0; tt < MTL; ++tt)
          mfma_step(accF[f][tt], sAl + (tt * 4 * SPF + f * SPF + sp) * 64, lane, b,
                    NP - 16 * tt <= 4);
      }
    }
    double qv[3][QN];
    {
      double cf[4][3];
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int d = 0; d < 3; ++d) cf[f][d] = hfs[f] * nrm[f][d];
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const double ur = accV[(4 * q) / 16][((4 * q) % 16) / 4];
        const double us = accV[(NPR + 4 * q) / 16][((NPR + 4 * q) % 16) / 4];
        const double ut = accV[(2 * NPR + 4 * q) / 16][((2 * NPR + 4 * q) % 16) / 4];
        double T[4];
#pragma unroll
        for (int f = 0; f < 4; ++f)
          T[f] = (4 * q) / 16 < MTL ? accF[f][((4 * q) / 16) % MTL][((4 * q) % 16) / 4] : 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double v = mt[d] * ur + mt[3 + d] * us + mt[6 + d] * ut;
#pragma unroll
          for (int f = 0; f < 4; ++f) v = fma(cf[f][d], T[f], v);
          qv[d][q] = v;
        }
      }
    }
    // Volume: [Dr | Ds | Dt] [S_r; S_s; S_t], B from registers.
    double4v accD[MTL];
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt) accD[tt] = (double4v){0.0, 0.0, 0.0, 0.0};
    {
#pragma unroll
      for (int s = 0; s < KSD; ++s) {
        const int blk = (4 * s) / NPR;
        const int sj = s - blk * (NPR / 4), j = 4 * sj + g;
        double b = 0.0;
        if (j < NP) {
          b = ee * (mt[3 * blk] * qv[0][sj] + mt[3 * blk + 1] * qv[1][sj] +
                    mt[3 * blk + 2] * qv[2][sj]);
          if (BURG) b = fma(-gh[blk] * su[j], su[j], b);
        }
#pragma unroll
        for (int tt = 0; tt < MTL; ++tt)
          mfma_step(accD[tt], sAd + (tt * KSD + s) * 64, lane, b, NP - 16 * tt <= 4);
      }
    }
    // Face traces sig = eps n_g . q at the points of face g of element c.
    double sg[NFP];
#pragma unroll
    for (int fp = 0; fp < NFP; ++fp) sg[fp] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      wave_fence();  // previous readers of sU are done
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const int i = g + 4 * q;
        if (i < NP) sU[c * S + i] = qv[d][q];
      }
      wave_fence();
      const double nd = g == 0 ? nrm[0][d] : g == 1 ? nrm[1][d] : g == 2 ? nrm[2][d] : nrm[3][d];
#pragma unroll
      for (int fp = 0; fp < NFP; ++fp) sg[fp] = fma(nd, su[sFM[g * NFP + fp]], sg[fp]);
    }
    if (c < cnt) {
      double* so = a.sig + (ge * 4 + g) * NFP;
#pragma unroll
      for (int fp = 0; fp < NFP; ++fp) so[fp] = ee * sg[fp];
    }
    // vol through sU, unit-stride stores.
    wave_fence();
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * tt + g + 4 * r;
        if (i < NP) sU[c * S + i] = accD[tt][r];
      }
    wave_fence();
    for (int i = lane; i < cnt * NP; i += 64) {
      const int e = i / NP, n = i - e * NP;
      a.vol[tl.vbase + i] = sU[e * S + n];
    }
  }
}

// Per-wave face record of the surface kernel: value offset of the
// neighbour's node 0, row of the orientation tables, rows of the own and
// neighbour face traces in sig, the boundary flag, Fscale, the penalty and
// the Burgers flux factors.
struct AvFaceRec {
  int32_t nb, fnrow, mrow, prow, bnd, pad;
  double fs, tau, fa, fb;
};

template <int N, int WPB, int MODE, bool BURG>
__global__ void __launch_bounds__(64 * WPB, 2) av_surf_kernel(AvArgs a, StageAv o) {
  using SH = AvShape<N>;
  constexpr int NP = SH::NP, NFP = SH::NFP, MTL = SH::MTL, KSF = SH::SPF, NPT = 4 * KSF;
  constexpr int NG = (16 * NFP + 63) / 64;  // gather instructions per face
  constexpr int NT = 64 * WPB;
  constexpr FmaskTable<N> FM{};
  constexpr int BUF = NPT * 17 > 16 * 17 ? NPT * 17 : 16 * 17;
  __shared__ __attribute__((aligned(16))) double sA[MTL * 4 * KSF * 64];
  __shared__ double sBuf[WPB][BUF];
  __shared__ AvFaceRec sRall[WPB][2][64];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFP[4 * 24 * NFP];  // neighbour's face point of each (face, code, point)
  __shared__ uint8_t sFM[4 * NFP];
  for (int i = threadIdx.x; i < 4 * NFP; i += NT) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
  for (int i = threadIdx.x; i < 4 * 24 * NFP; i += NT) {
    const int nn = a.face_nodes[i];
    sFN[i] = (uint8_t)nn;
    const int f2 = ((i / NFP) % 24) / 6;
    int fp2 = 0;
    for (int q = 0; q < NFP; ++q)
      if (FM.v[f2][q] == nn) fp2 = q;
    sFP[i] = (uint8_t)fp2;
  }
  // LIFT fragments, columns face by face: lane (row, k) of step s of face f
  // holds LIFT[row][f Nfp + 4 s + k] (zero past Nfp).
  for (int q = threadIdx.x; q < MTL * 4 * KSF * 64; q += NT) {
    const int l = q & 63, ts = q >> 6, cl = l & 15, kl = l >> 4;
    const int tt = ts / (4 * KSF), f = (ts / KSF) % 4, s = ts % KSF;
    const int row = 16 * tt + cl, fp = 4 * s + kl;
    sA[q] = row < NP && fp < NFP ? a.LIFT[row * 4 * NFP + f * NFP + fp] : 0.0;
  }
  __syncthreads();
  // wave index as a scalar, so that tile indices and records stay in SGPRs
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, g = lane >> 4;  // element of the tile, face (or k group)
  double* sB = sBuf[w];
  const int64_t K = a.K;
  double lmax = 0.0;
  // Gather pair of instruction j: element pc[j] of the tile, point pq[j].
  int pc[NG], pq[NG];
  bool pv[NG];
#pragma unroll
  for (int j = 0; j < NG; ++j) {
    const int p = lane + 64 * j;
    pv[j] = p < 16 * NFP;
    const int pp = pv[j] ? p : 16 * NFP - 1;
    pc[j] = pp / NFP;
    pq[j] = pp - pc[j] * NFP;
  }
  // XCD-contiguous tile ranges (see the split surface kernel).
  int t0 = blockIdx.x * WPB + w, t1 = a.ntiles, ts = gridDim.x * WPB;
  if (gridDim.x % 8 == 0) {
    const int x = blockIdx.x % 8;
    t0 = (int)((int64_t)a.ntiles * x / 8) + (blockIdx.x / 8) * WPB + w;
    t1 = (int)((int64_t)a.ntiles * (x + 1) / 8);
    ts = (gridDim.x / 8) * WPB;
  }
  // Tile records are read one tile ahead and made wave-uniform only when
  // used, so waiting for one never drains the loads issued after it.
  // Face-record inputs of lane (c, g) for a tile (element clamped to the tile).
  int32_t rnb = 0, rne = 0, rcode = 0, rge = 0;
  double ree = 0.0, rep = 0.0, rfs = 0.0, rga = 0.0, rgb = 0.0, rgc = 0.0;
  auto rec_load = [&](const Tile& tl) {
    const int ce = c < tl.count ? c : tl.count - 1;
    const int64_t ge = tl.gbase + ce;
    rge = (int32_t)ge;
    rnb = a.nbr[(int64_t)g * K + ge];
    rne = a.nbe[(int64_t)g * K + ge];
    rcode = a.codes[ge];
    ree = a.eps_in[ge];
    rfs = a.geo[(13 + 4 * g) * K + ge];
    if (BURG) {
      rga = o.bgeo[ge];
      rgb = o.bgeo[K + ge];
      rgc = o.bgeo[2 * K + ge];
    }
  };
  auto rec_store = [&](AvFaceRec* sR) {
    const int code = (rcode >> (5 * g)) & 31, f2 = code / 6;
    AvFaceRec r;
    r.nb = rnb;
    r.fnrow = (g * 24 + code) * NFP;
    r.mrow = rge * 4 + g;
    r.prow = rne * 4 + f2;
    r.bnd = rne == rge && f2 == g;  // boundary face: its own neighbour
    r.pad = 0;
    r.fs = rfs;
    r.tau = a.tau0 * fmax(ree, rep);
    r.fa = 0.0;
    r.fb = 0.0;
    if (BURG) dgk::face_factors(g, rga, rgb, rgc, r.fa, r.fb);
    sR[g * 16 + c] = r;
    wave_fence();
  };
  // Own and neighbour traces of u and sig of face f at the pairs of this
  // lane: two faces in flight, face f in slot f % 2.
  double uM[2][NG], uP[2][NG], sM[2][NG], sP[2][NG];
  auto gather = [&](const Tile& tl, const AvFaceRec* sR, int f) {
    const double* ut = a.u + tl.vbase;
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      const int ce = pc[j] < tl.count ? pc[j] : tl.count - 1;
      const AvFaceRec& r = sR[f * 16 + pc[j]];
      uP[f & 1][j] = a.u[r.nb + sFN[r.fnrow + pq[j]]];
      uM[f & 1][j] = ut[ce * NP + sFM[f * NFP + pq[j]]];
      sM[f & 1][j] = a.sig[(int64_t)r.mrow * NFP + pq[j]];
      sP[f & 1][j] = a.sig[(int64_t)r.prow * NFP + sFP[r.fnrow + pq[j]]];
    }
  };
  // MODE 2 (R) is the update with a0 = 0, a1 = dt = 1 and u = 0
  const double dt = MODE == 2 ? 1.0 : *o.dt;
  const double a0 = MODE == 0 ? o.a0 : 0.0, a1 = MODE == 2 ? 1.0 : o.a1;
  Tile tile_n{}, tile_raw{};
  int rb = 0;  // record buffer of the current tile
  if (t0 < t1) {
    tile_n = tile_uniform(a.tiles[t0]);
    rec_load(tile_n);
    rep = a.eps_in[rne];
    rec_store(sRall[w][0]);
    gather(tile_n, sRall[w][0], 0);
    if (t0 + ts < t1) tile_raw = a.tiles[t0 + ts];
  }
  for (int t = t0; t < t1; t += ts, rb ^= 1) {
    const Tile tile = tile_n;
    const bool more = t + ts < t1;
    if (more) {
      tile_n = tile_uniform(tile_raw);
      rec_load(tile_n);
      if (t + 2 * ts < t1) tile_raw = a.tiles[t + 2 * ts];
    }
    const int cnt = tile.count;
    const AvFaceRec* sR = sRall[w][rb];
    gather(tile, sR, 1);
    // The tile's state values through buffer resources: loads past the
    // tile read zero and stores past it are dropped, so partial tiles take
    // no branches.
    const int nbytes = cnt * NP * 8;
    const __amdgpu_buffer_rsrc_t rsV = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.vol + tile.vbase), 0, nbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rsU = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.u + tile.vbase), 0, MODE == 2 ? 0 : nbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(MODE == 0 ? o.u0 + tile.vbase : a.vol), 0, MODE == 0 ? nbytes : 0, 0x00020000);
    const __amdgpu_buffer_rsrc_t rsO = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(o.out + tile.vbase), 0, nbytes, 0x00020000);
    // Byte offset of the value lane updates for row tile tt, slot j (element
    // 4 j + lane / 16, row 16 tt + lane % 16); past the tile for rows >= Np.
    auto upd_off = [&](int tt, int j) -> int {
      const int e = 4 * j + (lane >> 4), i = 16 * tt + (lane & 15);
      return i < NP ? (e * NP + i) * 8 : 0x40000000;
    };
    double pv_[2][4], pu[2][4], p0[2][4];
    auto preload = [&](int tt) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int off = upd_off(tt, j);
        pv_[tt & 1][j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsV, off, 0, 0));
        pu[tt & 1][j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsU, off, 0, 0));
        p0[tt & 1][j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs0, off, 0, 2));
      }
    };
    // B operands of each face's steps: lane (c, g) takes point 4 s + g of
    // element c.
    double b[4][KSF];
    if constexpr (NPT > NFP) {
      // columns past Nfp of the jump buffer are zero (the update buffer
      // overwrote them)
      wave_fence();
      for (int i = lane; i < 16 * (NPT - NFP); i += 64) sB[(NFP + i / 16) * 17 + (i & 15)] = 0.0;
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (f > 0 || NPT == NFP) wave_fence();  // previous readers of the buffer are done
      // -Fscale ((sig- + sig+)/2 + tau (u- - u+))  [+ Fscale (n.f(u-) - n.f*)]
#pragma unroll
      for (int j = 0; j < NG; ++j) {
        const AvFaceRec& r = sR[f * 16 + pc[j]];
        const double m = uM[f & 1][j], p = uP[f & 1][j];
        const double sm = sM[f & 1][j], sn = r.bnd ? -sm : sP[f & 1][j];
        double x = -r.fs * (0.5 * (sm + sn) + r.tau * (m - p));
        if (BURG) x = fma(m - p, fma(r.fa, m + p, -r.fb * fmax(fabs(m), fabs(p))), x);
        if (pv[j]) sB[pq[j] * 17 + pc[j]] = x;
      }
      // the next tile's neighbour eps (its index arrived with the records)
      if (f == 0 && more) rep = a.eps_in[rne];
      if (f < 2) gather(tile, sR, f + 2);
      else if (f == 2) preload(0);
      wave_fence();
#pragma unroll
      for (int s = 0; s < KSF; ++s) b[f][s] = sB[(4 * s + g) * 17 + c];
    }
    // The next tile: face records, then its first face in flight.
    if (more) {
      AvFaceRec* sRn = sRall[w][rb ^ 1];
      rec_store(sRn);
      gather(tile_n, sRn, 0);
    }
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt) {
      if (tt + 1 < MTL) preload(tt + 1);
      double4v acc = (double4v){0.0, 0.0, 0.0, 0.0};
      __builtin_amdgcn_s_setprio(1);  // matrix phase first (wave priority)
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int s = 0; s < KSF; ++s)
          mfma_step(acc, sA + ((tt * 4 + f) * KSF + s) * 64, lane, b[f][s], NP - 16 * tt <= 4);
      __builtin_amdgcn_s_setprio(0);
      // Through the wave's 16 x 17 buffer to rows of 16 consecutive values.
      wave_fence();  // previous readers of the buffer are done
#pragma unroll
      for (int r = 0; r < 4; ++r) sB[c * 17 + g + 4 * r] = acc[r];
      wave_fence();
      // out = a0 u0 + a1 (u + dt (vol + LIFT g))
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double l = pv_[tt & 1][j] + sB[(4 * j + (lane >> 4)) * 17 + (lane & 15)];
        const double v = fma(a0, p0[tt & 1][j], a1 * fma(dt, l, pu[tt & 1][j]));
        const bool ok = 4 * j + (lane >> 4) < cnt && 16 * tt + (lane & 15) < NP;
        lmax = fmax(lmax, ok ? fabs(v) : 0.0);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rsO, upd_off(tt, j), 0, 0);
      }
    }
  }
  if (MODE != 2 && o.maxabs) {
    lmax = wmax(lmax);
    if (lane == 0 && lmax > 0.0)
      atomicMax((unsigned long long*)o.maxabs, (unsigned long long)__double_as_longlong(lmax));
  }
}

// ---------------------------------------------------------------------------
// av_surf_hi_kernel: the surface half of the viscous stage on the block
// layout of the split stage's P = 5, 6 surface kernel (surf_hi.hip): MT
// matrix waves (one 16-row tile of LIFT each, FP64 fragments in registers)
// and 4 memory waves (4 elements of every 16-element tile each), one tile per
// phase between block barriers; in phase n
//   matrix waves: tile n's vol and u0 rows loaded during the product
//                 G^T LIFT^T (v_mfma_f64_16x16x4_f64, result rows l/16 + 4 r),
//                 then out = a0 u0 + a1 (u + dt (vol + LIFT g)) with 16
//                 consecutive rows per store;
//   memory waves: the face records of tile n+3 into LDS, the loads of tile
//                 n+4's records and of tile n+2 (u, the neighbour traces of
//                 u, the own and neighbour traces of sig, the neighbours' eps,
//                 gathered through the records stored a phase before), and
//                 tile n+1's u into LDS and its jumps
//                   -Fscale ((sig- + sig+)/2 + tau (u- - u+))  [+ LF jump (BURG)]
//                 (tau = tau0 max(eps-, eps+); a boundary face, its own
//                 neighbour, has sig+ = -sig- and u+ = u-) into G.
// MODE 0: a0 u0 + ...; 1: a0 = 0; 2: out = vol + LIFT g.
// ---------------------------------------------------------------------------
template <int N>
struct AvHiCfg {
  static constexpr int NP = Tet<N>::Np, NFP = Tet<N>::Nfp, NF = 4 * NFP;
  static constexpr int MT = (NP + 15) / 16;  // row tiles of LIFT = matrix waves
  static constexpr int KF = NFP;             // k steps of 4 over the 4 Nfp columns
  static constexpr int MW = 4;               // memory waves
  static constexpr int EPW = 16 / MW;        // elements per memory wave
  static constexpr int NT = 64 * (MT + MW);
  static constexpr int NPU = NP | 1;         // LDS row stride of u
  static constexpr int GRS = NF + 1;         // LDS row stride of the jumps (+ a dump column), odd
  static constexpr int UPL = (EPW * NP + 63) / 64;   // u values per lane
  static constexpr int GPL = (EPW * NF + 63) / 64;   // trace pairs per lane
};

// Face record of (face f, element j) of a memory wave's elements: value
// offset of the neighbour's node 0, row of the orientation table, rows of
// the own and neighbour sig traces ([K][4] face rows of Nfp values), the
// neighbour element, boundary flag, own eps, Fscale, LF factors, and the
// neighbour's eps (loaded with the tile's traces, written when the jumps are
// formed).
struct AvHiRec {
  int32_t nb, fnrow, smrow, sprow, ne, bnd;
  double eo, fs, fa, fb, ep;
};

template <int N, int MODE, bool BURG>
__global__ void __launch_bounds__(AvHiCfg<N>::NT)
    av_surf_hi_kernel(AvArgs a, StageAv o, const Tile* __restrict__ tiles) {
  using C = AvHiCfg<N>;
  constexpr int NP = C::NP, NFP = C::NFP, NF = C::NF, MT = C::MT, KF = C::KF, MW = C::MW;
  constexpr int EPW = C::EPW, NPU = C::NPU, GRS = C::GRS, UPL = C::UPL, GPL = C::GPL;
  constexpr FmaskTable<N> FM{};
  static_assert(16 * NPU < 4096 && GRS < 512, "packed pair words");
  __shared__ double sG[2][16 * GRS];
  __shared__ double sU[2][16 * NPU + 1];
  __shared__ int64_t sTv[2];
  __shared__ int sTc[2];
  __shared__ AvHiRec sRec[MW][4][4 * EPW];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFP[4 * 24 * NFP];  // neighbour's face point of each (face, code, point)
  __shared__ uint8_t sFM[NF];
  for (int i = threadIdx.x; i < NF; i += C::NT) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
  for (int i = threadIdx.x; i < 4 * 24 * NFP; i += C::NT) {
    const int nn = a.face_nodes[i];
    sFN[i] = (uint8_t)nn;
    const int f2 = ((i / NFP) % 24) / 6;
    int fp2 = 0;
    for (int qq = 0; qq < NFP; ++qq)
      if (FM.v[f2][qq] == nn) fp2 = qq;
    sFP[i] = (uint8_t)fp2;
  }
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, g = lane >> 4;
  const int64_t K = a.K;
  // Tiles of this block: with a grid that is a multiple of 8 (block b on XCD
  // b % 8), each XCD walks one contiguous eighth of the tile list.
  int t0 = blockIdx.x, t1 = a.ntiles, ts = gridDim.x;
  if (gridDim.x % 8 == 0) {
    const int x = blockIdx.x % 8;
    t0 = (int)((int64_t)a.ntiles * x / 8) + blockIdx.x / 8;
    t1 = (int)((int64_t)a.ntiles * (x + 1) / 8);
    ts = gridDim.x / 8;
  }
  const int nb = t0 < t1 ? (t1 - t0 + ts - 1) / ts : 0;
  auto tile_of = [&](int n) -> Tile { return tile_uniform(tiles[t0 + n * ts]); };
  // Lane-derived values are rederived in each step (an empty asm hides the
  // lane index) rather than held across the tile loop.
  auto lane_now = [&]() {
    int l = lane;
    asm volatile("" : "+v"(l));
    return l;
  };
  __syncthreads();

  if (w < MT) {
    // ---------------- matrix wave: LIFT rows 16 w .. 16 w + 15 ----------------
    // fragment s: LIFT[16 w + l % 16][4 s + l / 16] in lane l
    double lf[KF];
#pragma unroll
    for (int s = 0; s < KF; ++s) {
      const int row = 16 * w + c;
      lf[s] = row < NP ? a.LIFT[row * NF + 4 * s + g] : 0.0;
    }
    const double dt = MODE == 2 ? 1.0 : *o.dt;
    const double a0 = MODE == 0 ? o.a0 : 0.0, a1 = MODE == 2 ? 1.0 : o.a1;
    // The fragments and dt are complete before the loop: without this wait
    // the compiler counts them as pending at the loop head.
    __builtin_amdgcn_s_waitcnt(0);
    double lmax = 0.0;
    __syncthreads();  // phase -1: the memory waves prepare tile 0
    for (int n = 0; n < nb; ++n) {
      const int bu = n & 1;
      const int64_t tv = sTv[bu];
      const int64_t vbase = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(tv >> 32)) << 32) |
                                      (uint32_t)__builtin_amdgcn_readfirstlane((int)tv));
      const int count = __builtin_amdgcn_readfirstlane(sTc[bu]);
      const int row = 16 * w + c;
      const int rr = row < NP ? row : NP - 1;
      // vol and u0 of the rows this lane updates (element g + 4 r, row
      // 16 w + c), loaded during the product through buffer resources sized
      // to the tile (loads past it read zero)
      const int nbytes = count * NP * 8;
      const __amdgpu_buffer_rsrc_t rsV =
          __builtin_amdgcn_make_buffer_rsrc((void*)(a.vol + vbase), 0, nbytes, 0x00020000);
      const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(MODE == 0 ? o.u0 + vbase : a.vol), 0, MODE == 0 ? nbytes : 0, 0x00020000);
      double rv[4], u0v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int off = ((g + 4 * r) * NP + rr) * 8;
        rv[r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsV, off, 0, 0));
        u0v[r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs0, off, 0, 2));
      }
      // G^T LIFT^T: register r of lane l holds row 16 w + l % 16 of element
      // l / 16 + 4 r
      double4v acc = (double4v){0.0, 0.0, 0.0, 0.0};
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int s = 0; s < KF; ++s)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sG[bu][c * GRS + 4 * s + g], lf[s], acc, 0, 0, 0);
      __builtin_amdgcn_s_setprio(0);
      // Stores through a buffer resource sized to the tile: the ones past it
      // (elements >= count, rows >= Np) are dropped.
      const __amdgpu_buffer_rsrc_t rsO =
          __builtin_amdgcn_make_buffer_rsrc((void*)(o.out + vbase), 0, nbytes, 0x00020000);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = g + 4 * r;
        const double uv = MODE == 2 ? 0.0 : sU[bu][e * NPU + rr];
        const double v = fma(a0, u0v[r], a1 * fma(dt, rv[r] + acc[r], uv));
        const bool ok = row < NP && e < count;
        lmax = fmax(lmax, ok ? fabs(v) : 0.0);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), rsO,
                                              row < NP ? (e * NP + row) * 8 : 0x40000000, 0, 0);
      }
      __syncthreads();  // end of phase n
    }
    if (MODE != 2 && o.maxabs) {
      lmax = wmax(lmax);
      if (lane == 0 && lmax > 0.0)
        atomicMax((unsigned long long*)o.maxabs, (unsigned long long)__double_as_longlong(lmax));
    }
    return;
  }

  // ---------------- memory wave q: elements EPW q .. EPW q + EPW - 1 ----------------
  const int q = w - MT;
  const int e0 = EPW * q;
  // Face records of this wave's elements for tile n go to slot n % 4 of
  // sRec[q], written three phases ahead from loads issued a phase before:
  // lane s < 4 EPW takes face s / EPW of element e0 + s % EPW (clamped to the
  // tile). Every lane loads (lane l the record of lane l % (4 EPW)), and
  // tiles past the block's last read its last, so that every phase issues
  // the same loads.
  static_assert((4 * EPW & (4 * EPW - 1)) == 0, "records per wave: a power of two");
  int32_t rnb = 0, rne = 0, rcode = 0, rge = 0;
  double reo = 0.0, rfs = 0.0, rga = 0.0, rgb = 0.0, rgc = 0.0;
  auto rec_load = [&](int n) {
    const int ln = lane_now() & (4 * EPW - 1), rf = ln / EPW, rj = ln % EPW;
    const Tile tl = tile_of(n < nb ? n : nb - 1);
    const int e = e0 + rj < tl.count ? e0 + rj : tl.count - 1;
    const int64_t ge = tl.gbase + e;
    rge = (int32_t)ge;
    rnb = a.nbr[(int64_t)rf * K + ge];
    rne = a.nbe[(int64_t)rf * K + ge];
    rcode = a.codes[ge];
    reo = a.eps_in[ge];
    rfs = a.geo[(13 + 4 * rf) * K + ge];
    if (BURG) {
      rga = o.bgeo[ge];
      rgb = o.bgeo[K + ge];
      rgc = o.bgeo[2 * K + ge];
    }
  };
  auto rec_store = [&](int n) {
    if (n >= nb) return;
    const int ln = lane_now(), rf = ln / EPW;
    if (ln < 4 * EPW) {
      const int code = (rcode >> (5 * rf)) & 31, f2 = code / 6;
      AvHiRec r;
      r.nb = rnb;
      r.fnrow = (rf * 24 + code) * NFP;
      r.smrow = rge * 4 + rf;
      r.sprow = rne * 4 + f2;
      r.ne = rne;
      r.bnd = rne == rge && f2 == rf;
      r.eo = reo;
      r.fs = rfs;
      // Fscale n.(1,1,1) of face rf: -gc, -gb, ga + gb + gc, -ga, as
      // coefficients in {-1, 0, 1} of (ga, gb, gc) rather than a branch on
      // the lane's face
      const double c0 = (double)(rf == 2) - (double)(rf == 3);
      const double c1 = (double)(rf == 2) - (double)(rf == 1);
      const double c2 = (double)(rf == 2) - (double)(rf == 0);
      const double sn = BURG ? fma(c0, rga, fma(c1, rgb, c2 * rgc)) : 0.0;
      r.fa = 0.25 * sn;
      r.fb = 0.5 * fabs(sn);
      sRec[q][n & 3][ln] = r;
    }
  };
  // Per lane and trace pair k (pair p = lane + 64 k: slot p / NFP = f EPW + j,
  // point qq = p % NFP), fixed for the launch: index of the own trace in a
  // tile's u (12 bits), column of the jump in G (9 bits; the dump column past
  // the pairs), record slot (5 bits), element (2 bits), point (5 bits).
  uint32_t fw[GPL];
#pragma unroll
  for (int k = 0; k < GPL; ++k) {
    const int p = lane + 64 * k;
    const bool ok = EPW * NF % 64 == 0 || p < EPW * NF;
    const int pp = ok ? p : EPW * NF - 1;
    const int sl = pp / NFP, qq = pp % NFP, f = sl / EPW, j = sl % EPW;
    const int iu = (e0 + j) * NPU + sFM[f * NFP + qq];
    const int col = ok ? f * NFP + qq : NF;
    fw[k] = (uint32_t)iu | ((uint32_t)col << 12) | ((uint32_t)sl << 21) | ((uint32_t)j << 26);
  }
  // u values (pair p = lane + 64 k: element e0 + p / NP, node p % NP), per
  // trace pair u+, sig- and sig+, and eps+ of lane l's record (l < 4 EPW), by
  // slot (compile-time constants, so the arrays stay in registers)
  double U[2][UPL], P[2][GPL], SM[2][GPL], SP[2][GPL], EP[2];
  int64_t TV[2] = {0, 0};
  int TC[2] = {1, 1};
  auto issue = [&](int n, auto SL) {
    constexpr int slot = decltype(SL)::value;
    const int nn = n < nb ? n : nb - 1;
    const Tile tl = tile_of(nn);
    const AvHiRec* R = sRec[q][nn & 3];
    const int lane = lane_now();
#pragma unroll
    for (int k = 0; k < UPL; ++k) {
      const int p = lane + 64 * k;
      const int e = e0 + p / NP;
      const int64_t i = (int64_t)(e < tl.count ? e : tl.count - 1) * NP + p % NP;
      U[slot][k] = a.u[tl.vbase + i];
    }
#pragma unroll
    for (int k = 0; k < GPL; ++k) {
      const int p = lane + 64 * k;
      const int pp = p < EPW * NF ? p : EPW * NF - 1;
      const int qq = pp % NFP;
      const AvHiRec& r = R[pp / NFP];
      P[slot][k] = a.u[r.nb + sFN[r.fnrow + qq]];
      SM[slot][k] = a.sig[(int64_t)r.smrow * NFP + qq];
      SP[slot][k] = a.sig[(int64_t)r.sprow * NFP + sFP[r.fnrow + qq]];
    }
    EP[slot] = a.eps_in[R[lane & (4 * EPW - 1)].ne];
    TV[slot] = tl.vbase;
    TC[slot] = tl.count;
  };
  const double tau0 = a.tau0;
  auto prepare = [&](int n, auto SL) {
    constexpr int slot = decltype(SL)::value;
    const int lane = lane_now();
    const int bu = n & 1;
    const AvHiRec* R = sRec[q][n & 3];
    double* su = sU[bu];
#pragma unroll
    for (int k = 0; k < UPL; ++k) {
      const int p = lane + 64 * k;
      const int o2 = p < EPW * NP ? (e0 + p / NP) * NPU + p % NP : 16 * NPU;
      su[o2] = U[slot][k];
    }
    if (q == 0 && lane == 0) {
      sTv[bu] = TV[slot];
      sTc[bu] = TC[slot];
    }
    if (lane < 4 * EPW) sRec[q][n & 3][lane].ep = EP[slot];
    wave_fence();  // the wave's u of the tile and eps+ of its records are in LDS
    double* G = sG[bu];
#pragma unroll
    for (int k = 0; k < GPL; ++k) {
      const uint32_t fk = fw[k];
      const int col = (fk >> 12) & 511, j = fk >> 26;
      const AvHiRec& r = R[(fk >> 21) & 31];
      const double m = su[fk & 4095], pv = P[slot][k];
      const double sm = SM[slot][k], sn = r.bnd ? -sm : SP[slot][k];
      const double tau = tau0 * fmax(r.eo, r.ep);
      double x = -r.fs * (0.5 * (sm + sn) + tau * (m - pv));
      if (BURG) x = fma(m - pv, fma(r.fa, m + pv, -r.fb * fmax(fabs(m), fabs(pv))), x);
      G[(e0 + j) * GRS + col] = x;
    }
  };
  // Prologue (phase -1 of the matrix waves): records of tiles 0..2 stored and
  // tile 3's loaded, loads of tiles 0 and 1, tile 0 prepared.
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  if (nb > 0) {
#pragma unroll 1
    for (int n = 0; n < 3; ++n) {
      rec_load(n);
      rec_store(n);
    }
    rec_load(3);
    wave_fence();
    issue(0, S0{});
    issue(1, S1{});
    prepare(0, S0{});
  }
  __syncthreads();  // end of phase -1
  // Phase n: the records of tile n+3 stored, the loads of tile n+4's records
  // and of tile n+2 into slot n & 1, then tile n+1 prepared from slot
  // (n+1) & 1, loaded a phase ago.
  auto phase = [&](int n, auto A1, auto A2) {
    rec_store(n + 3);
    rec_load(n + 4);
    wave_fence();
    issue(n + 2, A2);
    if (n + 1 < nb) prepare(n + 1, A1);
    __syncthreads();  // end of phase n
  };
  for (int n = 0; n < nb; n += 2) {
    phase(n, S1{}, S0{});
    if (n + 1 < nb) phase(n + 1, S0{}, S1{});
  }
}

// dt = min(cfl hmin / max|u|, cfl_v hmin^2 / max eps), clipped to t_end;
// scal = [max|u|, dt, t, max eps]; max|u| and max eps are cleared for the
// next step's reductions.
__global__ void av_dt_kernel(double* scal, double cfl, double cfl_v, double hmin, double t_end) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const double m = scal[0], e = scal[3], t = scal[2];
    double dt = cfl * hmin / m;
    if (e > 0.0) dt = fmin(dt, cfl_v * hmin * hmin / e);
    if (t_end > 0.0) dt = fmin(dt, fmax(t_end - t, 0.0));
    scal[1] = dt;
    scal[2] = t + dt;
    scal[0] = 0.0;
    scal[3] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Sensor network: y = sigmoid(w3 . tanh(W2 tanh(W1 x + b1) + b2) + b3),
// x = modal coefficients / their Euclidean norm, eps = eps0 y.
// Parameters (FP32 master copy): W1[64][nin], b1[64], W2[64][64], b2[64],
// w3[64], b3.
// ---------------------------------------------------------------------------
struct MlpLayout {
  int nin, oW1, ob1, oW2, ob2, ow3, ob3, total;
};

__host__ __device__ inline MlpLayout mlp_layout(int nin) {
  MlpLayout L;
  L.nin = nin;
  L.oW1 = 0;
  L.ob1 = kH * nin;
  L.oW2 = L.ob1 + kH;
  L.ob2 = L.oW2 + kH * kH;
  L.ow3 = L.ob2 + kH;
  L.ob3 = L.ow3 + kH;
  L.total = L.ob3 + 1;
  return L;
}

// Feature index of element j of the 8-wide BF16 k slice held by lane group g
// in k step s: the rows a 16x16 FP32 result places in that lane.
__host__ __device__ inline int perm_k(int s, int g, int j) {
  return 16 * (2 * s + (j >> 2)) + 4 * g + (j & 3);
}

// Per-lane network state: BF16 fragments of W1 and W2 in the permuted k
// order, and the biases / output weights of the rows this lane's results
// hold (rows 16 mt + 4 g + r).
struct MlpRegs {
  bf16x8 A1[4][2], A2[4][2];
  float b1[4][4], b2[4][4], w3[4][4];
  float b3;
};

__device__ __forceinline__ void mlp_load(const float* __restrict__ P, const MlpLayout& L, int lane,
                                         MlpRegs& R) {
  const int row = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = perm_k(s, g, j);
        const float w1 = k < L.nin ? P[L.oW1 + (16 * mt + row) * L.nin + k] : 0.f;
        R.A1[mt][s][j] = (__bf16)w1;
        R.A2[mt][s][j] = (__bf16)P[L.oW2 + (16 * mt + row) * kH + k];
      }
    }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int h = 16 * mt + 4 * g + r;
      R.b1[mt][r] = P[L.ob1 + h];
      R.b2[mt][r] = P[L.ob2 + h];
      R.w3[mt][r] = P[L.ow3 + h];
    }
  R.b3 = P[L.ob3];
}

// tanh(x) = 1 - 2 / (1 + e^(2x)) on the hardware exponential and
// reciprocal: ~1e-7 relative, far below the BF16 rounding of this path.
__device__ __forceinline__ float tanh_fast(float x) {
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __expf(2.f * x));
}

// Forward pass of 16 inputs (columns); X[s] is lane (c, g)'s BF16 slice of
// input c in k step s. Returns y for column c (all four lane groups).
// The biases enter as the accumulators' initial values.
__device__ __forceinline__ float mlp_forward_tile(const MlpRegs& R, const bf16x8 (&X)[2]) {
  floatx4 z[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    z[mt] = (floatx4){R.b1[mt][0], R.b1[mt][1], R.b1[mt][2], R.b1[mt][3]};
#pragma unroll
    for (int s = 0; s < 2; ++s) z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.A1[mt][s], X[s], z[mt], 0, 0, 0);
  }
  bf16x8 H[2];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int mt = 2 * s + (j >> 2), r = j & 3;
      H[s][j] = (__bf16)tanh_fast(z[mt][r]);
    }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    z[mt] = (floatx4){R.b2[mt][0], R.b2[mt][1], R.b2[mt][2], R.b2[mt][3]};
#pragma unroll
    for (int s = 0; s < 2; ++s) z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.A2[mt][s], H[s], z[mt], 0, 0, 0);
  }
  float part = 0.f;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) part = fmaf(R.w3[mt][r], tanh_fast(z[mt][r]), part);
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  return __builtin_amdgcn_rcpf(1.f + __expf(-(part + R.b3)));
}

// Network on given features x[n][nin] (FP32) -> y[n]; 16 inputs per wave step.
__global__ void __launch_bounds__(256) mlp_features_kernel(const float* __restrict__ P, int nin,
                                                           const float* __restrict__ x, int64_t n,
                                                           float* __restrict__ y) {
  const MlpLayout L = mlp_layout(nin);
  const int lane = threadIdx.x & 63;
  MlpRegs R;
  mlp_load(P, L, lane, R);
  const int c = lane & 15, g = lane >> 4;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t t = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t * 16 < n; t += nw) {
    const int64_t e = t * 16 + c;
    const int64_t ec = e < n ? e : n - 1;
    bf16x8 X[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = perm_k(s, g, j);
        X[s][j] = (__bf16)(k < nin ? x[ec * nin + k] : 0.f);
      }
    const float yv = mlp_forward_tile(R, X);
    if (g == 0 && e < n) y[e] = yv;
  }
}

// Two-part BF16 split of an FP32 value: x ~ hi + lo, ~16 significant bits.
__device__ __forceinline__ __bf16 bf16_hi(float x) { return (__bf16)x; }
__device__ __forceinline__ __bf16 bf16_lo(float x) { return (__bf16)(x - (float)(__bf16)x); }

// Learned sensor on the state: per 16-element tile, uh = V^-1 u on BF16
// matrix cores in the split form hi*hi + hi*lo + lo*hi (Henry, Tang &
// Heinecke, ARITH 2019; ~16 significant bits, ample for an energy fraction
// resolved to a few digits), x = uh / |uh|, y = network(x) on BF16 MFMA,
// eps = eps0 y.
template <int N>
__global__ void __launch_bounds__(256) mlp_sensor_kernel(AvArgs a, const float* __restrict__ P) {
  constexpr int NP = Tet<N>::Np, S = NP + 1;
  constexpr int MTM = (NP + 15) / 16, KSM = (NP + 31) / 32;
  constexpr int PV = (16 * NP + 63) / 64;  // tile values per lane
  static_assert(NP <= 48, "modal rows must fit three 16-row tiles");
  __shared__ double sUall[4][16 * S + 1];  // + a slot for the stores past the tile
  // V^-1 fragments (BF16 A operand of 16x16x32, hi and lo parts), shared by
  // the block: entry [part][mt][s][lane] holds A[16 mt + row][32 s + 8 g + j].
  __shared__ bf16x8 sV[2][MTM][KSM][64];
  // wave index as a scalar, so that tile indices and records stay in SGPRs
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, g = lane >> 4;
  for (int q = threadIdx.x; q < MTM * KSM * 64; q += 256) {
    const int l = q & 63, ms = q >> 6, mt = ms / KSM, s = ms % KSM;
    bf16x8 h, lo;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = 16 * mt + (l & 15), k = 32 * s + 8 * (l >> 4) + j;
      const float v = (i < NP && k < NP) ? (float)a.Vinv[i * NP + k] : 0.f;
      h[j] = bf16_hi(v);
      lo[j] = bf16_lo(v);
    }
    sV[0][mt][s][l] = h;
    sV[1][mt][s][l] = lo;
  }
  double* sU = sUall[w];
  const MlpLayout L = mlp_layout(NP);
  MlpRegs R;
  mlp_load(P, L, lane, R);
  __syncthreads();
  double emax = 0.0;
  // The wave's tiles: each next tile's values and volume factors are in
  // flight during the current one, its record a tile earlier.
  const int t0 = blockIdx.x * 4 + w, ts = gridDim.x * 4, t1 = a.ntiles;
  Tile tl{}, tile_raw{};
  double pre[PV], jn = 0.0;
  auto issue = [&](const Tile& d) {
#pragma unroll
    for (int j = 0; j < PV; ++j) pre[j] = a.u[d.vbase + min(lane + 64 * j, d.count * NP - 1)];
    jn = a.geo[9 * a.K + d.gbase + (c < d.count ? c : d.count - 1)];
  };
  if (t0 < t1) {
    tl = tile_uniform(a.tiles[t0]);
    issue(tl);
    if (t0 + ts < t1) tile_raw = a.tiles[t0 + ts];
  }
  for (int t = t0; t < t1; t += ts) {
    const Tile cur = tl;
    const double jv = jn;
    wave_fence();
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      const int e = i / NP, nn = i - e * NP;
      sU[i < cur.count * NP ? e * S + nn : 16 * S] = pre[j];
    }
    if (t + ts < t1) {
      tl = tile_uniform(tile_raw);
      issue(tl);
      if (t + 2 * ts < t1) tile_raw = a.tiles[t + 2 * ts];
    }
    wave_fence();
    const int cc = c < cur.count ? c : 0;
    floatx4 acc[MTM];
#pragma unroll
    for (int mt = 0; mt < MTM; ++mt) acc[mt] = (floatx4){0.f, 0.f, 0.f, 0.f};
    float umax = 0.f;
#pragma unroll
    for (int s = 0; s < KSM; ++s) {
      bf16x8 bh, bl;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = 32 * s + 8 * g + j;
        const float uv = k < NP ? (float)sU[cc * S + k] : 0.f;
        umax = fmaxf(umax, fabsf(uv));
        bh[j] = bf16_hi(uv);
        bl[j] = bf16_lo(uv);
      }
#pragma unroll
      for (int mt = 0; mt < MTM; ++mt) {
        const bf16x8 vh = sV[0][mt][s][lane], vl = sV[1][mt][s][lane];
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl, bh, acc[mt], 0, 0, 0);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh, bl, acc[mt], 0, 0, 0);
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh, bh, acc[mt], 0, 0, 0);
      }
    }
    // lane (c, g) holds uh rows 16 mt + 4 g + r of element c
    float tot = 0.f;
#pragma unroll
    for (int mt = 0; mt < MTM; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) tot = fmaf(acc[mt][r], acc[mt][r], tot);
    tot += __shfl_xor(tot, 16, 64);
    tot += __shfl_xor(tot, 32, 64);
    umax = fmaxf(umax, __shfl_xor(umax, 16, 64));
    umax = fmaxf(umax, __shfl_xor(umax, 32, 64));
    const float inv = tot > 0.f ? 1.f / sqrtf(tot) : 0.f;
    bf16x8 X[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int mt = 2 * s + (j >> 2), r = j & 3;
        X[s][j] = (__bf16)(mt < MTM ? acc[mt < MTM ? mt : 0][r] * inv : 0.f);
      }
    const float y = mlp_forward_tile(R, X);
    if (g == 0 && c < cur.count) {
      const int64_t ge = (int64_t)cur.gbase + c;
      // h = (6 vol)^(1/3), vol = 4/3 J; FP32 (the network output it scales
      // carries BF16 rounding)
      const float h = 2.f * cbrtf((float)jv);
      const double e = (double)((float)a.ceps * h * 1.7320508f * umax / N * y);
      a.eps[ge] = e;
      if (a.frac) a.frac[ge] = (double)y;
      emax = fmax(emax, e);
    }
  }
  if (a.maxeps) {
    emax = wmax(emax);
    if (lane == 0 && emax > 0.0)
      atomicMax((unsigned long long*)a.maxeps, (unsigned long long)__double_as_longlong(emax));
  }
}

// Sample index i of a training batch takes data row hash(seed, i) % ndata.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Gradient of the sensor network on BF16 matrix cores (FP32 accumulation).
// Each wave takes 32-row chunks (two 16-row tiles). The forward pass is the
// inference path (x, W1, W2 and h1 rounded to BF16; tanh and the sigmoid in
// FP32). Backward, with y - t of each row:
//   dy = 2 (y - t) / B y (1 - y),   dz2 = dy w3 (1 - h2^2),
//   dH1 = W2^T dz2 on matrix cores (dz2 and W2 in BF16; W2^T's k order is
//         permuted like W2's, so dz2's result layout is the B operand),
//   dz1 = dH1 (1 - h1^2),
//   dW2 = dz2^T h1 and dW1 = dz1^T x on matrix cores with the chunk's 32 rows
//         as the k dimension (dz2, h1, dz1 and x transposed through LDS in
//         BF16),
//   b1, b2, w3, b3 and the loss summed in FP32 on the vector units.
// Each wave writes its sums to its own slot of the partial buffer, and
// mlp_reduce_kernel adds the slots in a fixed order: no atomics, so a run
// is reproducible.
// ---------------------------------------------------------------------------
constexpr int kGradWaves = 4;        // waves per block of the gradient kernel
constexpr int kGradMaxWaves = 1024;  // partial slots

struct GradArgs {
  const float* P;   // FP32 master parameters
  int nin;
  const float* x;   // [ndata][nin]
  const float* t;   // [ndata]
  int64_t ndata;
  int64_t rows;     // rows of this gradient (the batch)
  uint64_t seed;    // 0: rows 0 .. rows-1 in order; else drawn with replacement
  float* part;      // [waves][stride]: the parameters' sums, then the loss
  int stride;
};

__global__ void __launch_bounds__(64 * kGradWaves, 1) mlp_grad_kernel(GradArgs a) {
  constexpr int W = kGradWaves, RS = 40;  // transpose row stride (BF16): 80 B
  const MlpLayout L = mlp_layout(a.nin);
  const int nin = a.nin;
  // A-operand fragments (16x16x32 BF16): W1 and W2 in the permuted k order
  // (forward), W2^T in the permuted k order (backward).
  __shared__ bf16x8 sA1[4][2][64], sA2[4][2][64], sAT[4][2][64];
  __shared__ float sB1[kH], sB2[kH], sW3[kH];
  // Per wave: [unit or feature][row of the chunk] in BF16.
  __shared__ __attribute__((aligned(16))) __bf16 sT2[W][kH * RS], sTh[W][kH * RS],
      sT1[W][kH * RS], sTx[W][kH * RS];
  for (int q = threadIdx.x; q < 4 * 2 * 64; q += 64 * W) {
    const int l = q & 63, ms = q >> 6, mt = ms >> 1, s = ms & 1, row = l & 15, g = l >> 4;
    bf16x8 f1, f2, ft;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = perm_k(s, g, j), u = 16 * mt + row;
      f1[j] = (__bf16)(k < nin ? a.P[L.oW1 + u * nin + k] : 0.f);
      f2[j] = (__bf16)a.P[L.oW2 + u * kH + k];
      ft[j] = (__bf16)a.P[L.oW2 + k * kH + u];
    }
    sA1[mt][s][l] = f1;
    sA2[mt][s][l] = f2;
    sAT[mt][s][l] = ft;
  }
  for (int q = threadIdx.x; q < kH; q += 64 * W) {
    sB1[q] = a.P[L.ob1 + q];
    sB2[q] = a.P[L.ob2 + q];
    sW3[q] = a.P[L.ow3 + q];
  }
  __syncthreads();
  const float b3 = a.P[L.ob3];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, g = lane >> 4;
  __bf16* T2 = sT2[w];
  __bf16* Th = sTh[w];
  __bf16* T1 = sT1[w];
  __bf16* Tx = sTx[w];
  const int nft = (nin + 15) / 16;  // feature tiles of dW1
  const float inv_b = 1.f / (float)a.rows;
  floatx4 gW2[4][4], gW1[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) gW2[i][j] = gW1[i][j] = (floatx4){0.f, 0.f, 0.f, 0.f};
  float gb1[4][4], gb2[4][4], gw3[4][4], gb3 = 0.f, lsum = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) gb1[i][r] = gb2[i][r] = gw3[i][r] = 0.f;
  const int64_t nchunk = (a.rows + 31) / 32;
  const int wave = blockIdx.x * W + w, nwave = gridDim.x * W;
  for (int64_t ch = wave; ch < nchunk; ch += nwave) {
    wave_fence();  // previous readers of the transposes are done
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const int64_t rr = ch * 32 + 16 * tt + c;  // batch position of this lane's row
      const bool valid = rr < a.rows;
      const int64_t idx = !valid ? 0
                          : (a.seed ? (int64_t)(mix64(a.seed + 0x9e3779b97f4a7c15ULL *
                                                                   (uint64_t)(rr + 1)) %
                                                (uint64_t)a.ndata)
                                    : rr);
      const float* xr = a.x + idx * nin;
      bf16x8 X[2];
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = perm_k(s, g, j);
          X[s][j] = (__bf16)(k < nin ? xr[k] : 0.f);
          Tx[k * RS + 16 * tt + c] = X[s][j];
        }
      const float tv = a.t[idx];
      // forward
      floatx4 z[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        z[mt] = (floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s)
          z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sA1[mt][s][lane], X[s], z[mt], 0, 0, 0);
      }
      float h1[4][4];
      bf16x8 H[2];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * mt + 4 * g + r;
          h1[mt][r] = tanh_fast(z[mt][r] + sB1[u]);
          const __bf16 hb = (__bf16)h1[mt][r];
          H[mt >> 1][4 * (mt & 1) + r] = hb;
          Th[u * RS + 16 * tt + c] = hb;
        }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        z[mt] = (floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s)
          z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sA2[mt][s][lane], H[s], z[mt], 0, 0, 0);
      }
      float h2[4][4], part = 0.f;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * mt + 4 * g + r;
          h2[mt][r] = tanh_fast(z[mt][r] + sB2[u]);
          part = fmaf(sW3[u], h2[mt][r], part);
        }
      part += __shfl_xor(part, 16, 64);
      part += __shfl_xor(part, 32, 64);
      const float y = __builtin_amdgcn_rcpf(1.f + __expf(-(part + b3)));
      const float err = valid ? y - tv : 0.f;
      const float dy = 2.f * err * inv_b * y * (1.f - y);
      if (g == 0) {
        lsum = fmaf(err * err, inv_b, lsum);
        gb3 += dy;
      }
      // backward
      bf16x8 D2[2];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * mt + 4 * g + r;
          const float d = dy * sW3[u] * (1.f - h2[mt][r] * h2[mt][r]);
          gb2[mt][r] += d;
          gw3[mt][r] = fmaf(dy, h2[mt][r], gw3[mt][r]);
          const __bf16 db = (__bf16)d;
          D2[mt >> 1][4 * (mt & 1) + r] = db;
          T2[u * RS + 16 * tt + c] = db;
        }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        z[mt] = (floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s)
          z[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sAT[mt][s][lane], D2[s], z[mt], 0, 0, 0);
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * mt + 4 * g + r;
          const float d = z[mt][r] * (1.f - h1[mt][r] * h1[mt][r]);
          gb1[mt][r] += d;
          T1[u * RS + 16 * tt + c] = (__bf16)d;
        }
    }
    wave_fence();
    // weight gradients: the chunk's 32 rows are the k dimension; lane (m, k)
    // reads rows 8 k .. 8 k + 7 of unit (or feature) 16 tile + m
    bf16x8 at2[4], at1[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      at2[mt] = *(const bf16x8*)(T2 + (16 * mt + c) * RS + 8 * g);
      at1[mt] = *(const bf16x8*)(T1 + (16 * mt + c) * RS + 8 * g);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const bf16x8 bh = *(const bf16x8*)(Th + (16 * nt + c) * RS + 8 * g);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        gW2[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at2[mt], bh, gW2[mt][nt], 0, 0, 0);
      if (nt < nft) {
        const bf16x8 bx = *(const bf16x8*)(Tx + (16 * nt + c) * RS + 8 * g);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
          gW1[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(at1[mt], bx, gW1[mt][nt], 0, 0, 0);
      }
    }
  }
  // The wave's sums to its slot: D lanes (n, g) hold rows 16 mt + 4 g + r,
  // column 16 nt + n; the vector sums reduce over the 16 rows of a lane
  // group.
  float* out = a.part + (int64_t)wave * a.stride;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int u = 16 * mt + 4 * g + r;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        out[L.oW2 + u * kH + 16 * nt + c] = gW2[mt][nt][r];
        const int f = 16 * nt + c;
        if (nt < nft && f < nin) out[L.oW1 + u * nin + f] = gW1[mt][nt][r];
      }
      float s1 = gb1[mt][r], s2 = gb2[mt][r], s3 = gw3[mt][r];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
        s3 += __shfl_xor(s3, o, 64);
      }
      if (c == 0) {
        out[L.ob1 + u] = s1;
        out[L.ob2 + u] = s2;
        out[L.ow3 + u] = s3;
      }
    }
  gb3 = wave_sum(gb3);
  lsum = wave_sum(lsum);
  if (lane == 0) {
    out[L.ob3] = gb3;
    out[L.total] = lsum;
  }
}

// Sum of the nw partial slots in slot order. mode 0: grad[p] = sum (and the
// loss into *loss); mode 1: Adam step on parameter p (beta1, beta2, eps,
// bias corrections c1, c2), the loss added to *loss.
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float* __restrict__ part, int nw,
                                                         int stride, int total, int mode,
                                                         float* __restrict__ grad,
                                                         float* __restrict__ P,
                                                         float* __restrict__ m,
                                                         float* __restrict__ v, float lr, float b1,
                                                         float b2, float eps, float c1, float c2,
                                                         float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > total) return;
  // eight slots' loads in flight, added in slot order
  float s = 0.f;
  int k = 0;
  for (; k + 8 <= nw; k += 8) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = part[(int64_t)(k + i) * stride + p];
#pragma unroll
    for (int i = 0; i < 8; ++i) s += v[i];
  }
  for (; k < nw; ++k) s += part[(int64_t)k * stride + p];
  if (p == total) {
    if (loss) *loss = mode == 0 ? s : *loss + s;
    return;
  }
  if (mode == 0) {
    grad[p] = s;
    return;
  }
  const float mi = b1 * m[p] + (1.f - b1) * s;
  const float vi = b2 * v[p] + (1.f - b2) * s * s;
  m[p] = mi;
  v[p] = vi;
  P[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}


}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
struct dgk_av_s {
  dgk_dgmesh m = nullptr;
  double* Vinv = nullptr;
  double* eps = nullptr;
  double* vol = nullptr;  // volume terms of the viscous stage RHS (state layout)
  double* sig = nullptr;  // [K][4][Nfp] viscous face traces
  double* zero_eps = nullptr;
  dgk_av_params p{};
};

struct dgk_mlp_s {
  dgk_device dev = nullptr;
  int nin = 0, total = 0;
  long step = 0;
  float* P = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  float* loss = nullptr;
  float* part = nullptr;   // per-wave partial sums of the gradient kernel
};

namespace {

AvArgs av_args(dgk_av a) {
  dgk_dgmesh m = a->m;
  AvArgs r{};
  r.tiles = m->tiles64;
  r.ntiles = m->ntiles64;
  r.K = m->K;
  r.geo = m->geo;
  r.Dr = m->Dr;
  r.Ds = m->Ds;
  r.Dt = m->Dt;
  r.LIFT = m->LIFT;
  r.Vinv = a->Vinv;
  r.face_nodes = m->face_nodes;
  r.nbr = m->nbr;
  r.nbe = m->nbe;
  r.codes = m->codes;
  r.eps = a->eps;
  r.eps_in = a->eps;
  r.vol = a->vol;
  r.sig = a->sig;
  r.s0 = a->p.s0;
  r.kappa = a->p.kappa;
  r.ceps = a->p.ceps;
  r.tau0 = a->p.tau0;
  return r;
}

int groups_grid(dgk_dgmesh m, int per_cu) {
  const int groups = (m->ntiles64 + 3) / 4;
  return std::max(1, std::min(groups, m->dev->cu_count * per_cu));
}

#define DGK_AV_LAUNCH(kernel, grid, block, shmem, ...)                               \
  switch (m->N) {                                                                    \
    case 1: hipLaunchKernelGGL(kernel<1>, grid, block, shmem, m->dev->stream, __VA_ARGS__); break; \
    case 2: hipLaunchKernelGGL(kernel<2>, grid, block, shmem, m->dev->stream, __VA_ARGS__); break; \
    case 3: hipLaunchKernelGGL(kernel<3>, grid, block, shmem, m->dev->stream, __VA_ARGS__); break; \
    case 4: hipLaunchKernelGGL(kernel<4>, grid, block, shmem, m->dev->stream, __VA_ARGS__); break; \
    default:                                                                         \
      dgk::set_error("shock capturing is built for N <= 4 (N=%d)", m->N);            \
      return DGK_ERR_UNSUPPORTED;                                                    \
  }

// Waves per block of the volume kernel: at N = 4 its operator fragments take
// 98 KiB of LDS, so one block of 8 waves (two per SIMD) shares them; lower
// orders use 4 waves and two blocks per CU.
template <int N>
constexpr int av_wpb() { return N >= 4 ? 8 : 4; }
// Waves per block of the surface kernel: one block of eight per CU, two per
// SIMD (the kernel keeps its own loads in flight a face and a row tile
// ahead, which takes the registers of two waves per SIMD).
constexpr int kSurfW = 8;

int av_blocks(dgk_dgmesh m) {
  const int wpb = m->N >= 4 ? 8 : 4;
  return std::max(1, std::min((m->ntiles64 + wpb - 1) / wpb, m->dev->cu_count * (m->N >= 4 ? 1 : 2)));
}

template <int N, bool BURG>
int av_vol_launch(dgk_dgmesh m, const AvArgs& r) {
  hipLaunchKernelGGL((av_vol_kernel<N, av_wpb<N>(), BURG>), dim3(av_blocks(m)),
                     dim3(64 * av_wpb<N>()), 0, m->dev->stream, r, (const double*)m->bgeo);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

// Persistent grid of the surface kernel: resident blocks per CU times the
// CU count (a multiple of 8, for the XCD-contiguous tile order).
template <int N, int MODE, bool BURG>
int av_surf_launch(dgk_dgmesh m, const AvArgs& r, const StageAv& o) {
  // Resident blocks per CU of this instantiation (thread-safe one-time
  // initialisation; the same for every gfx950 device).
  static const int per_cu = [] {
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, av_surf_kernel<N, kSurfW, MODE, BURG>,
                                                     64 * kSurfW, 0) != hipSuccess)
      v = 1;
    return v < 1 ? 1 : v;
  }();
  const int grid = std::max(1, std::min((m->ntiles64 + kSurfW - 1) / kSurfW, m->dev->cu_count * per_cu));
  hipLaunchKernelGGL((av_surf_kernel<N, kSurfW, MODE, BURG>), dim3(grid), dim3(64 * kSurfW), 0,
                     m->dev->stream, r, o);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

// The surface half on the memory-wave / matrix-wave block layout: one block
// of MT + 4 waves per resident slot, grid a multiple of 8 while the tiles
// allow.
template <int N, int MODE, bool BURG>
int av_surf_hi_launch(dgk_dgmesh m, const AvArgs& r, const StageAv& o) {
  static const int per_cu = [] {
    int v = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, av_surf_hi_kernel<N, MODE, BURG>,
                                                     AvHiCfg<N>::NT, 0) != hipSuccess)
      v = 1;
    return v < 1 ? 1 : v;
  }();
  const int grid = std::max(1, std::min(r.ntiles, m->dev->cu_count * per_cu));
  hipLaunchKernelGGL((av_surf_hi_kernel<N, MODE, BURG>), dim3(grid), dim3(AvHiCfg<N>::NT), 0,
                     m->dev->stream, r, o, r.tiles);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

// Surface kernel of the viscous stage: the memory-wave form at P = 4 (the
// one-wave-per-tile form below it).
template <int N, int MODE, bool BURG>
int av_surf_any(dgk_dgmesh m, const AvArgs& r, const StageAv& o) {
  if constexpr (N == 4) return av_surf_hi_launch<N, MODE, BURG>(m, r, o);
  return av_surf_launch<N, MODE, BURG>(m, r, o);
}

template <int N>
int av_run_n(dgk_dgmesh m, const AvArgs& r, const StageAv& o, int mode, bool burg) {
  int rc = burg ? av_vol_launch<N, true>(m, r) : av_vol_launch<N, false>(m, r);
  if (rc != DGK_OK) return rc;
  if (mode == 2)
    return burg ? av_surf_any<N, 2, true>(m, r, o) : av_surf_any<N, 2, false>(m, r, o);
  if (mode == 0) return av_surf_any<N, 0, true>(m, r, o);
  return av_surf_any<N, 1, true>(m, r, o);
}

// Volume kernel, then the surface kernel: mode 0/1 stage (a0 != 0 / a0 == 0),
// mode 2 rhs.
int av_run(dgk_av a, const double* u, const double* u0, double* out, double a0, double a1,
           const double* dt, double* maxabs, int mode, bool burg) {
  dgk_dgmesh m = a->m;
  if (m->ntiles64 == 0) return DGK_OK;
  AvArgs r = av_args(a);
  r.u = u;
  StageAv o;
  o.u0 = u0;
  o.out = out;
  o.a0 = a0;
  o.a1 = a1;
  o.dt = dt;
  o.maxabs = maxabs;
  o.bgeo = m->bgeo;
  switch (m->N) {
    case 1: return av_run_n<1>(m, r, o, mode, burg);
    case 2: return av_run_n<2>(m, r, o, mode, burg);
    case 3: return av_run_n<3>(m, r, o, mode, burg);
    case 4: return av_run_n<4>(m, r, o, mode, burg);
  }
  dgk::set_error("shock capturing is built for N <= 4 (N=%d)", m->N);
  return DGK_ERR_UNSUPPORTED;
}

}  // namespace

extern "C" int dgk_av_create(dgk_dgmesh m, const double* Vinv_host, const dgk_av_params* p,
                             dgk_av* out) {
  dgk::clear_error();
  DGK_REQUIRE(m && Vinv_host && p && out, "dgk_av_create: NULL argument");
  DGK_REQUIRE(m->N >= 1 && m->N <= 4, "dgk_av_create: shock capturing is built for N <= 4");
  DGK_REQUIRE(p->kappa > 0.0 && p->ceps >= 0.0 && p->tau0 >= 0.0,
              "dgk_av_create: need kappa > 0, ceps >= 0, tau0 >= 0");
  *out = nullptr;
  dgk_av a = new dgk_av_s();
  a->m = m;
  a->p = *p;
  const size_t nstate = (size_t)m->state_off.back() * 8 + 16;
  int rc = DGK_OK;
  rc |= dgk_malloc(m->dev, (size_t)m->Np * m->Np * 8, Vinv_host, (void**)&a->Vinv);
  rc |= dgk_malloc(m->dev, (size_t)m->K * 8 + 8, nullptr, (void**)&a->eps);
  rc |= dgk_malloc(m->dev, nstate, nullptr, (void**)&a->vol);
  rc |= dgk_malloc(m->dev, (size_t)m->K * 4 * m->Nfp * 8 + 8, nullptr, (void**)&a->sig);
  if (rc == DGK_OK) rc = dgk_fill_bytes(m->dev, a->eps, 0, (size_t)m->K * 8 + 8);
  if (rc != DGK_OK) {
    dgk_av_destroy(a);
    return rc;
  }
  *out = a;
  return DGK_OK;
}

extern "C" int dgk_av_destroy(dgk_av a) {
  if (!a) return DGK_OK;
  dgk_device dev = a->m->dev;
  dgk_free(dev, a->Vinv);
  dgk_free(dev, a->eps);
  dgk_free(dev, a->vol);
  dgk_free(dev, a->sig);
  delete a;
  return DGK_OK;
}

extern "C" int dgk_av_eps_dev(dgk_av a, double** eps_dev) {
  DGK_REQUIRE(a && eps_dev, "dgk_av_eps_dev: NULL argument");
  *eps_dev = a->eps;
  return DGK_OK;
}

extern "C" int dgk_av_sensor_pp(dgk_av a, const double* u_dev, double* maxeps_dev,
                                double* frac_dev, float* frac32_dev, float* feats_dev) {
  dgk::clear_error();
  DGK_REQUIRE(a && u_dev, "dgk_av_sensor_pp: NULL argument");
  dgk_dgmesh m = a->m;
  if (m->ntiles64 == 0) return DGK_OK;
  AvArgs r = av_args(a);
  r.u = u_dev;
  r.maxeps = maxeps_dev;
  r.frac = frac_dev;
  r.frac32 = frac32_dev;
  r.feats = feats_dev;
  // four waves (tiles) per block, two blocks per CU
  const dim3 grid(std::max(1, std::min((m->ntiles64 + 3) / 4, 2 * m->dev->cu_count))), block(256);
  DGK_AV_LAUNCH(pp_sensor_kernel, grid, block, 0, r);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

extern "C" int dgk_av_sensor_mlp(dgk_av a, dgk_mlp net, const double* u_dev, double* maxeps_dev,
                                 double* frac_dev) {
  dgk::clear_error();
  DGK_REQUIRE(a && net && u_dev, "dgk_av_sensor_mlp: NULL argument");
  dgk_dgmesh m = a->m;
  DGK_REQUIRE(net->nin == m->Np, "dgk_av_sensor_mlp: network input width %d != Np %d", net->nin,
              m->Np);
  if (m->ntiles64 == 0) return DGK_OK;
  AvArgs r = av_args(a);
  r.u = u_dev;
  r.maxeps = maxeps_dev;
  r.frac = frac_dev;
  const int grid = std::max(1, std::min((m->ntiles64 + 3) / 4, m->dev->cu_count * 4));
  DGK_AV_LAUNCH(mlp_sensor_kernel, dim3(grid), dim3(256), 0, r, (const float*)net->P);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

extern "C" int dgk_av_rhs(dgk_av a, const double* u_dev, double* out_dev) {
  dgk::clear_error();
  DGK_REQUIRE(a && u_dev && out_dev, "dgk_av_rhs: NULL argument");
  DGK_REQUIRE(out_dev != u_dev, "dgk_av_rhs: out must not alias u");
  return av_run(a, u_dev, u_dev, out_dev, 0.0, 1.0, nullptr, nullptr, 2, false);
}

extern "C" int dgk_av_stage(dgk_av a, const double* u_dev, const double* u0_dev, double* out_dev,
                            double a0, double a1, const double* dt_dev, double* maxabs_dev) {
  dgk::clear_error();
  DGK_REQUIRE(a && u_dev && out_dev && dt_dev, "dgk_av_stage: NULL argument");
  DGK_REQUIRE(out_dev != u_dev, "dgk_av_stage: out must not alias u");
  DGK_REQUIRE(a0 == 0.0 || u0_dev, "dgk_av_stage: u0 is NULL with a0 != 0");
  return av_run(a, u_dev, u0_dev ? u0_dev : u_dev, out_dev, a0, a1, dt_dev, maxabs_dev,
                a0 != 0.0 ? 0 : 1, true);
}

extern "C" int dgk_av_rhs_total(dgk_av a, const double* u_dev, double* out_dev) {
  dgk::clear_error();
  DGK_REQUIRE(a && u_dev && out_dev, "dgk_av_rhs_total: NULL argument");
  DGK_REQUIRE(out_dev != u_dev, "dgk_av_rhs_total: out must not alias u");
  return av_run(a, u_dev, u_dev, out_dev, 0.0, 1.0, nullptr, nullptr, 2, true);
}

extern "C" int dgk_av_cfl_dt(dgk_device dev, double* scal_dev, double cfl, double cfl_v,
                             double hmin, double t_end) {
  dgk::clear_error();
  DGK_REQUIRE(dev && scal_dev, "dgk_av_cfl_dt: NULL argument");
  hipLaunchKernelGGL(av_dt_kernel, dim3(1), dim3(64), 0, dev->stream, scal_dev, cfl, cfl_v, hmin,
                     t_end);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

// ------------------------------- network ----------------------------------
extern "C" int dgk_mlp_create(dgk_device dev, int nin, const float* params_host, dgk_mlp* out) {
  dgk::clear_error();
  DGK_REQUIRE(dev && params_host && out, "dgk_mlp_create: NULL argument");
  // the training kernel's LDS (64 samples x (nin + 3 x 64) floats) stays within 64 KiB
  DGK_REQUIRE(nin >= 1 && nin <= 56, "dgk_mlp_create: input width %d outside 1..56", nin);
  *out = nullptr;
  dgk_mlp n = new dgk_mlp_s();
  n->dev = dev;
  n->nin = nin;
  n->total = mlp_layout(nin).total;
  const size_t bytes = (size_t)n->total * 4;
  int rc = DGK_OK;
  rc |= dgk_malloc(dev, bytes, params_host, (void**)&n->P);
  rc |= dgk_malloc(dev, bytes, nullptr, (void**)&n->m);
  rc |= dgk_malloc(dev, bytes, nullptr, (void**)&n->v);
  rc |= dgk_malloc(dev, 16, nullptr, (void**)&n->loss);
  rc |= dgk_malloc(dev, (size_t)kGradMaxWaves * (n->total + 1) * 4, nullptr, (void**)&n->part);
  if (rc == DGK_OK) rc = dgk_fill_bytes(dev, n->m, 0, bytes);
  if (rc == DGK_OK) rc = dgk_fill_bytes(dev, n->v, 0, bytes);
  if (rc == DGK_OK) rc = dgk_fill_bytes(dev, n->loss, 0, 16);
  if (rc != DGK_OK) {
    dgk_mlp_destroy(n);
    return rc;
  }
  *out = n;
  return DGK_OK;
}

extern "C" int dgk_mlp_destroy(dgk_mlp n) {
  if (!n) return DGK_OK;
  dgk_free(n->dev, n->P);
  dgk_free(n->dev, n->m);
  dgk_free(n->dev, n->v);
  dgk_free(n->dev, n->loss);
  dgk_free(n->dev, n->part);
  delete n;
  return DGK_OK;
}

extern "C" int dgk_mlp_num_params(dgk_mlp n, int* count) {
  DGK_REQUIRE(n && count, "dgk_mlp_num_params: NULL argument");
  *count = n->total;
  return DGK_OK;
}

extern "C" int dgk_mlp_get(dgk_mlp n, float* params_host) {
  DGK_REQUIRE(n && params_host, "dgk_mlp_get: NULL argument");
  return dgk_copy_d2h(n->dev, params_host, n->P, (size_t)n->total * 4, 0);
}

extern "C" int dgk_mlp_set(dgk_mlp n, const float* params_host) {
  DGK_REQUIRE(n && params_host, "dgk_mlp_set: NULL argument");
  n->step = 0;
  const size_t bytes = (size_t)n->total * 4;
  int rc = dgk_copy_h2d(n->dev, n->P, params_host, bytes, 0);
  if (rc == DGK_OK) rc = dgk_fill_bytes(n->dev, n->m, 0, bytes);
  if (rc == DGK_OK) rc = dgk_fill_bytes(n->dev, n->v, 0, bytes);
  return rc;
}

extern "C" int dgk_mlp_forward(dgk_mlp n, const float* x_dev, int64_t count, float* y_dev) {
  dgk::clear_error();
  DGK_REQUIRE(n && x_dev && y_dev, "dgk_mlp_forward: NULL argument");
  DGK_REQUIRE(count >= 0, "dgk_mlp_forward: negative count");
  if (count == 0) return DGK_OK;
  const int64_t tiles = (count + 15) / 16;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, n->dev->cu_count * 4));
  hipLaunchKernelGGL(mlp_features_kernel, dim3(grid), dim3(256), 0, n->dev->stream, n->P, n->nin,
                     x_dev, count, y_dev);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

// Gradient of the mean squared error over all `count` rows into grad_dev
// (FP32 [num_params]); loss_dev (may be NULL) receives the loss.
namespace {

// The gradient kernel over `rows` rows (drawn from the ndata rows when seed
// is nonzero) into the per-wave slots; returns the number of slots used.
int mlp_grad_launch(dgk_mlp n, const float* x, const float* t, int64_t ndata, int64_t rows,
                    uint64_t seed) {
  const int64_t chunks = (rows + 31) / 32;
  const int waves = (int)std::min<int64_t>(chunks, kGradMaxWaves);
  const int blocks = (waves + kGradWaves - 1) / kGradWaves;
  GradArgs g;
  g.P = n->P;
  g.nin = n->nin;
  g.x = x;
  g.t = t;
  g.ndata = ndata;
  g.rows = rows;
  g.seed = seed;
  g.part = n->part;
  g.stride = n->total + 1;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64 * kGradWaves), 0, n->dev->stream, g);
  return blocks * kGradWaves;
}

}  // namespace

// Gradient of the mean squared error over all `count` rows into grad_dev
// (FP32 [num_params]); loss_dev (may be NULL) receives the loss.
extern "C" int dgk_mlp_grad(dgk_mlp n, const float* x_dev, const float* t_dev, int64_t count,
                            float* grad_dev, float* loss_dev) {
  dgk::clear_error();
  DGK_REQUIRE(n && x_dev && t_dev && grad_dev, "dgk_mlp_grad: NULL argument");
  DGK_REQUIRE(count >= 1 && count <= (1 << 30), "dgk_mlp_grad: count %lld", (long long)count);
  const int nw = mlp_grad_launch(n, x_dev, t_dev, count, count, 0);
  DGK_LAUNCH_CHECK();
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((n->total + 1 + 255) / 256), dim3(256), 0,
                     n->dev->stream, n->part, nw, n->total + 1, n->total, 0, grad_dev, nullptr,
                     nullptr, nullptr, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, loss_dev ? loss_dev : n->loss);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

// One Adam step on a batch sampled (with replacement) from the ndata rows by
// a hash of (seed, step); loss of the batch added to *loss_acc_dev if not
// NULL.
extern "C" int dgk_mlp_train_step(dgk_mlp n, const float* x_dev, const float* t_dev, int64_t ndata,
                                  int batch, uint64_t seed, float lr, float* loss_acc_dev) {
  dgk::clear_error();
  DGK_REQUIRE(n && x_dev && t_dev, "dgk_mlp_train_step: NULL argument");
  DGK_REQUIRE(ndata >= 1 && batch >= 1, "dgk_mlp_train_step: empty data or batch");
  n->step += 1;
  const uint64_t s = (seed | 1ULL) * 0x2545F4914F6CDD1DULL + (uint64_t)n->step;
  const int nw = mlp_grad_launch(n, x_dev, t_dev, ndata, batch, s ? s : 1ULL);
  DGK_LAUNCH_CHECK();
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((n->total + 1 + 255) / 256), dim3(256), 0,
                     n->dev->stream, n->part, nw, n->total + 1, n->total, 1, nullptr, n->P, n->m,
                     n->v, lr, b1, b2, 1e-8f, c1, c2, loss_acc_dev ? loss_acc_dev : n->loss);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}
