// Tetrahedral DG Burgers stage on gfx950: device mesh, FP64 affine geometry,
// the fused strong-form RHS + SSP-RK3 stage, max|u| and the CFL step.
//
// The reference has only a volume-only, non-conservative Burgers RHS
// (examples/3d_scalar_burgers_equation.go:295-340) and sketches LF + LIFT in
// documentation/DGKernel/example_application.md:296-353. This file
// implements the strong form of Hesthaven & Warburton (2008) §2.4/§6:
//   L(u) = -div f(u) + LIFT (Fscale (n.f(u-) - n.f*)),  f(u) = u^2/2 (1,1,1),
//   n.f* = (n.f(u-) + n.f(u+))/2 - alpha |n.(1,1,1)| (u+ - u-)/2,
//   alpha = max(|u-|, |u+|),
// and one SSP-RK3 (Shu & Osher 1988) stage out = a0 u0 + a1 (u + dt L(u)).
//
// Design (one element per thread, one tile of TB elements per block):
//  * the tile's nodal values are contiguous in the runner layout
//    (X_offsets[p] + l*Np), so they are staged through LDS with unit-stride
//    loads and each thread then holds its element's Np values in VGPRs;
//  * Dr, Ds, Dt and LIFT are read with wave-uniform addresses (scalar loads,
//    SGPR operands of v_fma_f64), so the O(Np^2) contractions use no LDS;
//  * affine geometry is folded per element at setup into 11 FP64 numbers
//    (rx+ry+rz, sx+sy+sz, tx+ty+tz, and per face n.(1,1,1) and Fscale);
//  * the neighbour trace u+ is gathered straight from the neighbour's nodes
//    (value offset + orientation code -> node table in LDS); no trace arrays
//    are written between stages;
//  * the RK combination is written back through LDS with unit stride;
//  * one persistent launch walks a flat list of tiles covering every
//    partition, so uneven partition sizes cost no padding.
#include <algorithm>
#include <cmath>

#include "dgmesh_internal.h"

namespace {

using dgk::FmaskTable;
using dgk::kBGeoRows;
using dgk::kGeoRows;
using dgk::Tet;
using dgk::Tile;


std::vector<int> host_fmask(int N) {
  std::vector<int> v;
  int Nfp = dgk::nfp_of(N);
  v.assign(4 * Nfp, 0);
  int c[4] = {0, 0, 0, 0}, sk = 0;
  for (int n = 0; n <= N; ++n)
    for (int m = 0; m <= N - n; ++m)
      for (int l = 0; l <= N - n - m; ++l) {
        if (n == 0) v[0 * Nfp + c[0]++] = sk;
        if (m == 0) v[1 * Nfp + c[1]++] = sk;
        if (l + m + n == N) v[2 * Nfp + c[2]++] = sk;
        if (l == 0) v[3 * Nfp + c[3]++] = sk;
        ++sk;
      }
  return v;
}



// FP64 affine geometry per element from its four vertices: the
// GeometricFactors3D / Normals3D / ComputeFscale formulas
// (element/library/gonudg/normals3d.go:11-207) for a map whose Jacobian is
// constant: xr = (vb - va)/2, xs = (vc - va)/2, xt = (vd - va)/2.
__global__ void geometry_kernel(int64_t K, const double* __restrict__ vtx,
                                double* __restrict__ geo, double* __restrict__ bgeo,
                                int* __restrict__ bad) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const double* v = vtx + k * 12;
  double xr = 0.5 * (v[3] - v[0]), yr = 0.5 * (v[4] - v[1]), zr = 0.5 * (v[5] - v[2]);
  double xs = 0.5 * (v[6] - v[0]), ys = 0.5 * (v[7] - v[1]), zs = 0.5 * (v[8] - v[2]);
  double xt = 0.5 * (v[9] - v[0]), yt = 0.5 * (v[10] - v[1]), zt = 0.5 * (v[11] - v[2]);
  double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
  if (!(J > 0.0)) atomicAdd(bad, 1);
  double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
  double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
  double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
  double m[9] = {rx, ry, rz, sx, sy, sz, tx, ty, tz};
  for (int q = 0; q < 9; ++q) geo[q * K + k] = m[q];
  geo[9 * K + k] = J;
  double nv[4][3] = {{-tx, -ty, -tz}, {-sx, -sy, -sz}, {rx + sx + tx, ry + sy + ty, rz + sz + tz},
                     {-rx, -ry, -rz}};
  bgeo[0 * K + k] = rx + ry + rz;
  bgeo[1 * K + k] = sx + sy + sz;
  bgeo[2 * K + k] = tx + ty + tz;
  for (int f = 0; f < 4; ++f) {
    double mag = sqrt(nv[f][0] * nv[f][0] + nv[f][1] * nv[f][1] + nv[f][2] * nv[f][2]);
    double nx = nv[f][0] / mag, ny = nv[f][1] / mag, nz = nv[f][2] / mag;
    double sJ = mag * J;
    double fs = sJ / J;
    geo[(10 + 4 * f) * K + k] = nx;
    geo[(11 + 4 * f) * K + k] = ny;
    geo[(12 + 4 * f) * K + k] = nz;
    geo[(13 + 4 * f) * K + k] = fs;
    bgeo[(3 + f) * K + k] = nx + ny + nz;
    bgeo[(7 + f) * K + k] = fs;
    // LF jump factored as (u- - u+) (A (u- + u+) - B alpha):
    bgeo[(11 + f) * K + k] = 0.25 * fs * (nx + ny + nz);
    bgeo[(15 + f) * K + k] = 0.5 * fs * fabs(nx + ny + nz);
  }
}

struct StageParams {
  const double* Dr;
  const double* Ds;
  const double* Dt;
  const double* LIFT;
  const uint8_t* face_nodes;  // [4][24][Nfp]
  const double* bgeo;         // [11][K]
  const int32_t* nbr;         // [4][K] value offset of each neighbour's node 0
  const int32_t* codes;       // [K] four 5-bit orientation codes
  const Tile* tiles;
  int ntiles;
  int chunked;                // 1: each block takes a contiguous run of tiles
  int64_t K;
  const double* u;
  const double* u0;
  double* out;
  double a0, a1;
  const double* dt;
  unsigned long long* maxabs;
  int mode;                   // 0 stage, 1 rhs only (out = L(u)), 2 trace u+ (out = uP)
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Pointer into the constant address space: uniform loads through it become
// scalar loads (SGPR operands) even after the kernel has stored to memory.
typedef const __attribute__((address_space(4))) double* cdouble_ptr;

template <int N, int TB>
__global__ void __launch_bounds__(TB) burgers_stage_kernel(StageParams p) {
  constexpr int NP = Tet<N>::Np, NFP = Tet<N>::Nfp, NF = 4 * NFP;
  constexpr int S = NP + 1;  // padded LDS row: odd stride in doubles spreads banks
  constexpr FmaskTable<N> FM{};
  constexpr int NOP = 3 * NP * NP + NP * NF;  // Dr, Ds, Dt, LIFT
  __shared__ __attribute__((aligned(16))) double sOp[NOP];
  __shared__ __attribute__((aligned(16))) double sU[TB * S];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  // The operators stay in LDS for the block's lifetime; every read below has
  // a wave-uniform address, so it is a broadcast (one LDS access per wave).
  for (int i = threadIdx.x; i < 4 * 24 * NFP; i += TB) sFN[i] = p.face_nodes[i];
  for (int i = threadIdx.x; i < NP * NP; i += TB) {
    sOp[i] = p.Dr[i];
    sOp[NP * NP + i] = p.Ds[i];
    sOp[2 * NP * NP + i] = p.Dt[i];
  }
  for (int i = threadIdx.x; i < NP * NF; i += TB) sOp[3 * NP * NP + i] = p.LIFT[i];
  const double dt = p.mode == 0 ? *p.dt : 0.0;
  const int64_t K = p.K;
  double lmax = 0.0;
  int t0, t1, ts;
  if (p.chunked) {
    // XCD-aware: consecutive hardware blocks land on different XCDs
    // (round robin over 8), so renumber blocks so that each XCD owns a
    // contiguous range of tiles and its L2 sees the neighbour data of
    // nearby tiles.
    const int G = gridDim.x;
    const int b = blockIdx.x;
    const int vb = (G % 8 == 0) ? (b % 8) * (G / 8) + b / 8 : b;
    t0 = (int)((int64_t)p.ntiles * vb / G);
    t1 = (int)((int64_t)p.ntiles * (vb + 1) / G);
    ts = 1;
  } else {
    t0 = blockIdx.x;
    t1 = p.ntiles;
    ts = gridDim.x;
  }
  for (int t = t0; t < t1; t += ts) {
    const Tile tile = p.tiles[t];
    const int cnt = tile.count;
    __syncthreads();
    for (int i = threadIdx.x; i < cnt * NP; i += TB) {
      int e = i / NP, n = i - e * NP;
      sU[e * S + n] = p.u[tile.vbase + i];
    }
    __syncthreads();
    const int e = threadIdx.x;
    if (e < cnt) {
      const int64_t g = tile.gbase + e;
      const double* su = sU + e * S;
      const int packed = p.codes[g];
      int32_t nb[4];
#pragma unroll
      for (int f = 0; f < 4; ++f) nb[f] = p.nbr[f * K + g];
      if (p.mode == 2) {
        // Trace of the neighbour, as the stage sees it.
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const int code = (packed >> (5 * f)) & 31;
#pragma unroll
          for (int fp = 0; fp < NFP; ++fp)
            p.out[g * NF + f * NFP + fp] = p.u[nb[f] + sFN[(f * 24 + code) * NFP + fp]];
        }
      } else {
        // Operators through scalar loads (wave-uniform addresses). The zero
        // offset depends on this tile's record, so the loads stay inside the
        // tile loop instead of being hoisted into registers for all tiles.
        const int64_t z = (int64_t)(__builtin_amdgcn_readfirstlane(cnt) >> 30);
        const cdouble_ptr Dr = (cdouble_ptr)(p.Dr + z);
        const cdouble_ptr Ds = (cdouble_ptr)(p.Ds + z);
        const cdouble_ptr Dt = (cdouble_ptr)(p.Dt + z);
        const cdouble_ptr LIFT = (cdouble_ptr)(p.LIFT + z);
        double v[NP];
        {
          const double ga = p.bgeo[0 * K + g], gb = p.bgeo[1 * K + g], gc = p.bgeo[2 * K + g];
          double F[NP];
#pragma unroll
          for (int n = 0; n < NP; ++n) F[n] = 0.5 * su[n] * su[n];
#pragma unroll
          for (int i = 0; i < NP; ++i) {
            double r = 0.0, s = 0.0, q = 0.0;
#pragma unroll
            for (int j = 0; j < NP; ++j) {
              r = fma(Dr[i * NP + j], F[j], r);
              s = fma(Ds[i * NP + j], F[j], s);
              q = fma(Dt[i * NP + j], F[j], q);
            }
            v[i] = -(ga * r + gb * s + gc * q);
          }
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const int code = (packed >> (5 * f)) & 31;
          const double sn = p.bgeo[(3 + f) * K + g];
          const double fs = p.bgeo[(7 + f) * K + g];
          const double asn = fabs(sn);
          double gf[NFP];
#pragma unroll
          for (int fp = 0; fp < NFP; ++fp) {
            const double uM = su[FM.v[f][fp]];
            const double uP = p.u[nb[f] + sFN[(f * 24 + code) * NFP + fp]];
            const double fM = 0.5 * uM * uM, fP = 0.5 * uP * uP;
            const double al = fmax(fabs(uM), fabs(uP));
            gf[fp] = fs * (0.5 * (fM - fP) * sn + 0.5 * al * asn * (uP - uM));
          }
#pragma unroll
          for (int i = 0; i < NP; ++i) {
            double acc = v[i];
#pragma unroll
            for (int fp = 0; fp < NFP; ++fp) acc = fma(LIFT[i * NF + f * NFP + fp], gf[fp], acc);
            v[i] = acc;
          }
        }
        // Only this thread reads its own row of sU (uM above), so the row is
        // overwritten in place with u + dt*L(u) (or L(u) alone).
        double* sv = sU + e * S;
        if (p.mode == 0) {
#pragma unroll
          for (int n = 0; n < NP; ++n) sv[n] = fma(dt, v[n], sv[n]);
        } else {
#pragma unroll
          for (int n = 0; n < NP; ++n) sv[n] = v[n];
        }
      }
    }
    if (p.mode == 2) continue;
    __syncthreads();
    for (int i = threadIdx.x; i < cnt * NP; i += TB) {
      int ee = i / NP, n = i - ee * NP;
      double o = sU[ee * S + n];
      if (p.mode == 0) {
        o = p.a1 * o;
        if (p.a0 != 0.0) o = fma(p.a0, p.u0[tile.vbase + i], o);
      }
      p.out[tile.vbase + i] = o;
      lmax = fmax(lmax, fabs(o));
    }
  }
  if (p.maxabs) {
    lmax = wave_max(lmax);
    if ((threadIdx.x & 63) == 0 && lmax > 0.0)
      atomicMax(p.maxabs, (unsigned long long)__double_as_longlong(lmax));
  }
}

// ---------------------------------------------------------------------------
// MFMA variant of the stage (FP64 matrix cores, v_mfma_f64_16x16x4_f64).
//
// A block of 4 waves owns a tile of 64 elements, 16 per wave (the MFMA N
// dimension). Per wave, in FP64 throughout:
//   C_vol  = [Dr; Ds; Dt] * F,  F = u^2/2       (3 NPR x 16, K = Np)
//   C_lift = LIFT * G,          G = flux jumps  (Np x 16,    K = 4 Nfp)
// where the three derivative blocks are stacked at row offsets 0, NPR, 2 NPR
// with NPR = Np rounded up to a multiple of 4.
//
// Operand layout of the 16x16x4 FP64 MFMA (wave64), checked on gfx950: lane
// l provides A[l%16][l/16] and B[l/16][l%16] of the 16x4 / 4x16 operands and
// holds D[l/16 + 4r][l%16], r = 0..3, of the 16x16 result (rows interleaved
// by 4). With blocks stacked at multiples of 4 rows, the lane that holds row
// i of Dr F also holds rows i of Ds F, Dt F and LIFT G for the same element,
// so the chain rule with the folded affine metrics and the RK combination
// happen in registers.
//
// The operator fragments are written to LDS once per block, in per-lane
// order, and read back with one conflict-free ds_read_b64 per MFMA. The next
// tile's nodal values and the current tile's u0 are loaded while the
// current tile computes.
// ---------------------------------------------------------------------------
typedef double double4_t __attribute__((ext_vector_type(4)));

template <int N>
struct MfmaShape {
  static constexpr int NP = Tet<N>::Np, NFP = Tet<N>::Nfp, NF = 4 * NFP;
  static constexpr int NPR = (NP + 3) / 4 * 4;       // row offset between Dr, Ds, Dt blocks
  static constexpr int KSV = (NP + 3) / 4;           // K steps of the volume product
  static constexpr int MTV = (3 * NPR + 15) / 16;    // 16-row tiles of the stacked operator
  static constexpr int MTL = (NP + 15) / 16;         // 16-row tiles of LIFT
  static constexpr int KSL = NF / 4;                 // K steps of the lift product
  static constexpr int QN = NPR / 4;                 // result rows per lane
  static constexpr int S = NP + 1;                   // LDS row stride of the u tile
  static constexpr int NA = (MTV * KSV + MTL * KSL) * 64;  // operator fragments (doubles)
};

// LDS writes of one lane made visible to the other lanes of the same wave.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One k = 4 step of a 16-row operator tile whose only rows are its first
// four (the last LIFT tile at N = 1, 3, 4) as one v_mfma_f64_4x4x4_4b
// instead of a v_mfma_f64_16x16x4_f64, which on gfx950 issues about eight
// times slower. The 4x4x4 operand layout (checked on gfx950): block j of the
// four is lanes 16 k + 4 j + i, with A[i][k] and B[k][i] at that lane and
// D[i][n] at lane 16 i + 4 j + n. With block j as elements 4 j..4 j+3, B is
// the 16x16x4 B operand unchanged, lane l reads A row l%4, column l/16 from
// the tile's 16x16x4 fragment table, and the result lands where register 0
// of the 16x16x4 result would (row l/16, element l%16); registers 1-3, rows
// past the operator, stay zero.
__device__ __forceinline__ void mfma_f64_rows4(double4_t& acc, const double* frag, int lane,
                                               double b) {
  acc[0] = __builtin_amdgcn_mfma_f64_4x4x4f64(frag[16 * (lane >> 4) + (lane & 3)], b, acc[0], 0,
                                              0, 0);
}

// The same for a tile whose rows end inside it: one 4x4x4 per group of four
// rows that holds operator rows (ng of them, a constant once the callers'
// loops are unrolled); group r reads A row 4 r + l%4 from lane
// 16 (l/16) + 4 r + l%4 and fills register r's place.
__device__ __forceinline__ void mfma_f64_row_groups(double4_t& acc, const double* frag, int lane,
                                                    double b, int ng) {
  const double* f = frag + 16 * (lane >> 4) + (lane & 3);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (r < ng) acc[r] = __builtin_amdgcn_mfma_f64_4x4x4f64(f[4 * r], b, acc[r], 0, 0, 0);
}

// Each wave works through its own sequence of 16-element tiles; the waves of a
// block only share the operator fragments (written once, one block barrier).
template <int N, int WPB>
__global__ void __launch_bounds__(64 * WPB) burgers_stage_mfma_kernel(StageParams p) {
  using SH = MfmaShape<N>;
  constexpr int NP = SH::NP, NFP = SH::NFP, NF = SH::NF, S = SH::S, NPR = SH::NPR;
  constexpr int KSV = SH::KSV, MTV = SH::MTV, MTL = SH::MTL, KSL = SH::KSL, QN = SH::QN;
  constexpr int NT = 64 * WPB;               // threads per block
  constexpr int E = 16;                      // elements per wave tile
  constexpr int PV = (E * NP + 63) / 64;     // tile values per lane
  constexpr FmaskTable<N> FM{};
  __shared__ __attribute__((aligned(16))) double sA[SH::NA];
  __shared__ __attribute__((aligned(16))) double sUall[WPB][E * S];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFM[NF];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, gq = lane >> 4;
  for (int i = threadIdx.x; i < 4 * 24 * NFP; i += NT) sFN[i] = p.face_nodes[i];
  for (int i = threadIdx.x; i < NF; i += NT) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
  // Operator fragments: entry (tile, kstep, lane) in lane order.
  for (int q = threadIdx.x; q < SH::NA; q += NT) {
    const int l = q & 63, ts_ = q >> 6, cl = l & 15, gl = l >> 4;
    double v = 0.0;
    if (ts_ < MTV * KSV) {
      const int tt = ts_ / KSV, s = ts_ % KSV;
      const int row = 16 * tt + cl, k = 4 * s + gl, blk = row / NPR, i = row % NPR;
      if (k < NP && blk < 3 && i < NP) v = (blk == 0 ? p.Dr : blk == 1 ? p.Ds : p.Dt)[i * NP + k];
    } else {
      const int r2 = ts_ - MTV * KSV, tt = r2 / KSL, s = r2 % KSL;
      const int row = 16 * tt + cl, k = 4 * s + gl;
      if (row < NP) v = p.LIFT[row * NF + k];
    }
    sA[q] = v;
  }
  __syncthreads();
  const double dt = p.mode == 0 ? *p.dt : 0.0;
  const int64_t K = p.K;
  double lmax = 0.0;
  // Wave-level persistent schedule.
  const int NW = gridDim.x * WPB;
  const int gw = blockIdx.x * WPB + w;
  const int vw = (p.chunked && NW % 8 == 0) ? (gw % 8) * (NW / 8) + gw / 8 : gw;
  const int t0 = p.chunked ? (int)((int64_t)p.ntiles * vw / NW) : gw;
  const int t1 = p.chunked ? (int)((int64_t)p.ntiles * (vw + 1) / NW) : p.ntiles;
  const int ts = p.chunked ? 1 : NW;
  const double* sAv = sA;
  const double* sAl = sA + MTV * KSV * 64;
  double* sU = sUall[w];

  Tile tile;
  tile.vbase = 0; tile.gbase = 0; tile.count = 0;
  if (t0 < t1) tile = p.tiles[t0];
  double pre[PV];
#pragma unroll
  for (int j = 0; j < PV; ++j) {
    const int i = lane + 64 * j;
    pre[j] = i < tile.count * NP ? p.u[tile.vbase + i] : 0.0;
  }
  for (int t = t0; t < t1; t += ts) {
    const Tile cur = tile;
    const int cnt = cur.count;
    wave_sync();  // this wave's previous readers of sU are done
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      if (i < cnt * NP) {
        const int e = i / NP, n = i - e * NP;
        sU[e * S + n] = pre[j];
      }
    }
    // u0 of this tile and the next tile's u: in flight while this tile computes.
    double u0v[PV];
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      u0v[j] = (p.mode == 0 && p.a0 != 0.0 && i < cnt * NP)
                   ? __builtin_nontemporal_load(p.u0 + cur.vbase + i)
                   : 0.0;
    }
    {
      const int tn = t + ts;
      if (tn < t1) tile = p.tiles[tn]; else tile.count = 0;
#pragma unroll
      for (int j = 0; j < PV; ++j) {
        const int i = lane + 64 * j;
        pre[j] = i < tile.count * NP ? p.u[tile.vbase + i] : 0.0;
      }
    }
    const bool valid = c < cnt;
    const int64_t g = cur.gbase + (valid ? c : 0);
    const int packed = valid ? p.codes[g] : 0;
    int32_t nb[4];
    double sn[4], fs[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      nb[f] = valid ? p.nbr[f * K + g] : 0;
      sn[f] = valid ? p.bgeo[(3 + f) * K + g] : 0.0;
      fs[f] = valid ? p.bgeo[(7 + f) * K + g] : 0.0;
    }
    const double ga = valid ? p.bgeo[0 * K + g] : 0.0;
    const double gb = valid ? p.bgeo[1 * K + g] : 0.0;
    const double gc = valid ? p.bgeo[2 * K + g] : 0.0;
    double uP[KSL];
#pragma unroll
    for (int s = 0; s < KSL; ++s) {
      const int m = gq + 4 * s, f = m / NFP, fp = m - f * NFP;
      const int code = (packed >> (5 * f)) & 31;
      const int32_t base = f == 0 ? nb[0] : f == 1 ? nb[1] : f == 2 ? nb[2] : nb[3];
      uP[s] = valid ? p.u[base + sFN[(f * 24 + code) * NFP + fp]] : 0.0;
    }
    wave_sync();  // sU holds this tile
    const double* su = sU + c * S;
    double4_t accV[MTV];
#pragma unroll
    for (int tt = 0; tt < MTV; ++tt) accV[tt] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KSV; ++s) {
      const int k = 4 * s + gq;
      double Fk = 0.0;
      if (valid && k < NP) {
        const double uk = su[k];
        Fk = 0.5 * uk * uk;
      }
#pragma unroll
      for (int tt = 0; tt < MTV; ++tt)
        accV[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sAv[(tt * KSV + s) * 64 + lane], Fk,
                                                         accV[tt], 0, 0, 0);
    }
    double4_t accL[MTL];
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt) accL[tt] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KSL; ++s) {
      const int m = gq + 4 * s, f = m / NFP;
      const double snf = f == 0 ? sn[0] : f == 1 ? sn[1] : f == 2 ? sn[2] : sn[3];
      const double fsf = f == 0 ? fs[0] : f == 1 ? fs[1] : f == 2 ? fs[2] : fs[3];
      double gm = 0.0;
      if (valid) {
        const double uM = su[sFM[m]], up = uP[s];
        const double fM = 0.5 * uM * uM, fP = 0.5 * up * up;
        const double al = fmax(fabs(uM), fabs(up));
        gm = fsf * (0.5 * (fM - fP) * snf + 0.5 * al * fabs(snf) * (up - uM));
      }
#pragma unroll
      for (int tt = 0; tt < MTL; ++tt) {
        if (NP - 16 * tt <= 4)
          mfma_f64_rows4(accL[tt], sAl + (tt * KSL + s) * 64, lane, gm);
        else
          accL[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sAl[(tt * KSL + s) * 64 + lane], gm,
                                                           accL[tt], 0, 0, 0);
      }
    }
    // Rows i = gq + 4q of this lane's element: Dr F at row i, Ds F at NPR + i,
    // Dt F at 2 NPR + i, LIFT G at i; tile and slot are compile-time.
    double vout[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const double ur = accV[(4 * q) / 16][((4 * q) % 16) / 4];
      const double us = accV[(NPR + 4 * q) / 16][((NPR + 4 * q) % 16) / 4];
      const double ut = accV[(2 * NPR + 4 * q) / 16][((2 * NPR + 4 * q) % 16) / 4];
      const double lf = (4 * q) / 16 < MTL ? accL[((4 * q) / 16) % MTL][((4 * q) % 16) / 4] : 0.0;
      const double rhs = -(ga * ur + gb * us + gc * ut) + lf;
      const int i = gq + 4 * q;
      vout[q] = (i < NP && p.mode == 0) ? fma(dt, rhs, su[i < NP ? i : 0]) : rhs;
    }
    wave_sync();  // all reads of sU as u are done
    if (valid) {
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const int i = gq + 4 * q;
        if (i < NP) sU[c * S + i] = vout[q];
      }
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      if (i < cnt * NP) {
        const int e = i / NP, n = i - e * NP;
        double o = sU[e * S + n];
        if (p.mode == 0) o = fma(p.a0, u0v[j], p.a1 * o);
        __builtin_nontemporal_store(o, p.out + cur.vbase + i);
        lmax = fmax(lmax, fabs(o));
      }
    }
  }
  if (p.maxabs) {
    lmax = wave_max(lmax);
    if (lane == 0 && lmax > 0.0) atomicMax(p.maxabs, (unsigned long long)__double_as_longlong(lmax));
  }
}

// ---------------------------------------------------------------------------
// Pipelined MFMA stage: the same stage as the kernel above, scheduled for
// memory-level parallelism and laid out so that each lane serves one face.
//
//  * LIFT k ordering: k = 4 fp + f, i.e. the 4 lanes of a column group gq
//    hold face f = gq, and k step s covers face point s of all four faces.
//    Each lane therefore needs the normal, Fscale, neighbour offset and
//    orientation code of a single face, fixed for the whole tile: no
//    per-step selects and no face padding (4 Nfp / 4 = Nfp k steps).
//  * Every load is unconditional: addresses of lanes past a partial tile are
//    clamped into the tile, so the loop body has no divergent branches and
//    the compiler can count outstanding loads exactly. Only the stores of a
//    partial tile are predicated; full tiles (all but at most one per
//    partition) run a loop without any predicate.
//  * Software pipeline per wave: tile records are loaded two tiles ahead;
//    the next tile's nodal values and neighbour records are issued after the
//    current tile's gathers and geometry, so they are in flight while the
//    current tile computes; u0 is issued after the volume product.
//  * The -1/2 of -div(u^2/2 (1,1,1)) is folded into the stacked derivative
//    operator, and the Lax-Friedrichs jump is factored as
//    (u- - u+) (A (u- + u+) - B alpha), A = Fscale n.(1,1,1)/4,
//    B = Fscale |n.(1,1,1)|/2, precomputed per face at setup.
// ---------------------------------------------------------------------------
template <int N>
struct PipeShape {
  static constexpr int NP = Tet<N>::Np, NFP = Tet<N>::Nfp, NF = 4 * NFP;
  static constexpr int KSL = NFP;                    // k steps of the lift product
  static constexpr int NPR = (NP + 3) / 4 * 4;       // row offset between Dr, Ds, Dt blocks
  static constexpr int KSV = (NP + 3) / 4;           // k steps of the volume product
  static constexpr int MTV = (3 * NPR + 15) / 16;    // 16-row tiles of the stacked operator
  static constexpr int MTL = (NP + 15) / 16;         // 16-row tiles of LIFT
  static constexpr int QN = NPR / 4;                 // result rows per lane
  static constexpr int S = NP + 1;                   // LDS row stride of the u tile
  static constexpr int E = 16;                       // elements per wave tile
  static constexpr int PV = (E * NP + 63) / 64;      // tile values per lane
  static constexpr int NA = (MTV * KSV + MTL * KSL) * 64;  // operator fragments (doubles)
};

template <int PV>
struct TileIn {
  double u[PV];    // nodal values of the tile, value lane + 64 j
  int32_t nb;      // neighbour value offset across face lane / 16 of element lane % 16
  int32_t packed;  // the element's four orientation codes
};

// A tile record as loaded by every lane (vector load, one address).
__device__ __forceinline__ int4 load_tile_raw(const Tile* tiles, int t) {
  return *reinterpret_cast<const int4*>(tiles + t);
}

// The record in scalar registers (it is wave-uniform).
__device__ __forceinline__ Tile tile_uniform(int4 v) {
  Tile d;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane(v.x);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane(v.y);
  d.vbase = (int64_t)(((uint64_t)hi << 32) | lo);
  d.gbase = __builtin_amdgcn_readfirstlane(v.z);
  d.count = __builtin_amdgcn_readfirstlane(v.w);
  return d;
}

template <int N>
__device__ __forceinline__ void load_tile_in(const StageParams& p, const Tile& d, int lane,
                                             TileIn<PipeShape<N>::PV>& in) {
  using SH = PipeShape<N>;
  const int last = d.count * SH::NP - 1;
#pragma unroll
  for (int j = 0; j < SH::PV; ++j) in.u[j] = p.u[d.vbase + min(lane + 64 * j, last)];
  const int64_t g = (int64_t)d.gbase + min(lane & 15, d.count - 1);
  in.packed = p.codes[g];
  in.nb = p.nbr[(lane >> 4) * p.K + g];
}

// Packed volume-node indices of face f's points (6 bits each, five per
// word): word w holds points 5w .. 5w+4.
template <int N>
struct FacePack {
  static constexpr int NFP = Tet<N>::Nfp, NW = (NFP + 4) / 5;
  uint32_t w[4][NW];
  constexpr FacePack() : w{} {
    constexpr FmaskTable<N> FM{};
    for (int f = 0; f < 4; ++f)
      for (int q = 0; q < NFP; ++q) w[f][q / 5] |= (uint32_t)FM.v[f][q] << (6 * (q % 5));
  }
};

// MODE 0: out = a0 u0 + a1 (u + dt L(u)); 1: out = a1 (u + dt L(u)); 2: out = L(u).
// REGA: the operator fragments are held in registers for the whole sweep
// (40 doubles per lane at N = 3) instead of being read from LDS per MFMA.
template <int N, int MODE, bool FULL, bool REGA>
__device__ __forceinline__ void pipe_tiles(const StageParams& p, const double* sA, double* sU,
                                           const uint8_t* sFN, int t0, int t1, int ts,
                                           double dt, double& lmax) {
  using SH = PipeShape<N>;
  constexpr int NP = SH::NP, NFP = SH::NFP, S = SH::S, NPR = SH::NPR;
  constexpr int KSV = SH::KSV, MTV = SH::MTV, MTL = SH::MTL, KSL = SH::KSL, QN = SH::QN;
  constexpr int E = SH::E, PV = SH::PV;
  constexpr FacePack<N> FPK{};
  constexpr int NW = FacePack<N>::NW;
  if (t0 >= t1) return;
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, gq = lane >> 4;  // element column, face of this lane
  const int64_t K = p.K;
  constexpr int NFR = REGA ? MTV * KSV + MTL * KSL : 1;
  // LIFT fragments of a last tile with four rows or fewer feed a 4x4x4
  // product (mfma_f64_rows4), which reads lane 16 (l/16) + l%4 of the table.
  constexpr bool L4 = NP - 16 * (MTL - 1) <= 4;
  double ra[NFR];
  if (REGA) {
#pragma unroll
    for (int q = 0; q < NFR; ++q)
      ra[q] = sA[q * 64 + (L4 && q >= MTV * KSV + (MTL - 1) * KSL ? 16 * (lane >> 4) + (lane & 3)
                                                                   : lane)];
  }
  const double* sAv = sA;
  const double* sAl = sA + MTV * KSV * 64;
  const double* su = sU + c * S;
  // Own volume nodes of face gq, packed (lane-dependent, loop-invariant).
  uint32_t own[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w)
    own[w] = gq == 0 ? FPK.w[0][w] : gq == 1 ? FPK.w[1][w] : gq == 2 ? FPK.w[2][w] : FPK.w[3][w];
  const uint8_t* sfn_face = sFN + gq * 24 * NFP;

  Tile dcur = tile_uniform(load_tile_raw(p.tiles, t0));
  Tile dnxt = tile_uniform(load_tile_raw(p.tiles, min(t0 + ts, t1 - 1)));
  TileIn<PV> in;
  load_tile_in<N>(p, dcur, lane, in);
  for (int t = t0; t < t1; t += ts) {
    const int4 d2 = load_tile_raw(p.tiles, min(t + 2 * ts, t1 - 1));
    const int cnt = FULL ? E : dcur.count;
    // Stage this tile's values (rows of S doubles; lanes past the tile write
    // the pad slot of the last row).
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      const int e = i / NP, n = i - e * NP;
      const bool ok = (E * NP % 64 == 0 || i < E * NP) && (FULL || i < cnt * NP);
      sU[ok ? e * S + n : E * S - 1] = in.u[j];
    }
    // Neighbour trace across face gq: neighbour node of each face point.
    double uP[KSL];
    {
      const int code = (in.packed >> (5 * gq)) & 31;
      const uint8_t* fn = sfn_face + code * NFP;
      int idx[KSL];
#pragma unroll
      for (int s = 0; s < KSL; ++s) idx[s] = fn[s];
#pragma unroll
      for (int s = 0; s < KSL; ++s) uP[s] = p.u[in.nb + idx[s]];
    }
    __builtin_amdgcn_sched_barrier(0);
    // Geometry of element c (clamped into the tile): the folded metrics; the
    // flux factors of face gq follow from them (face_factors).
    const int64_t g = (int64_t)dcur.gbase + (FULL ? c : min(c, cnt - 1));
    const double ga = p.bgeo[0 * K + g], gb = p.bgeo[1 * K + g], gc = p.bgeo[2 * K + g];
    double fa, fb;
    dgk::face_factors(gq, ga, gb, gc, fa, fb);
    // Next tile, in flight while this one computes.
    TileIn<PV> nxt;
    load_tile_in<N>(p, dnxt, lane, nxt);
    __builtin_amdgcn_sched_barrier(0);
    wave_sync();  // sU holds this tile
    __builtin_amdgcn_s_setprio(1);  // matrix phase first (wave priority)
    // Volume: [Dr; Ds; Dt] F, folded at once with the affine metrics.
    double vout[QN];
    {
      double4_t accV[MTV];
#pragma unroll
      for (int tt = 0; tt < MTV; ++tt) accV[tt] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KSV; ++s) {
        const int k = 4 * s + gq;
        const double uk = su[k < NP ? k : NP - 1];
        const double Fk = k < NP ? uk * uk : 0.0;  // the operator carries -1/2
#pragma unroll
        for (int tt = 0; tt < MTV; ++tt)
          accV[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              REGA ? ra[(tt * KSV + s) % NFR] : sAv[(tt * KSV + s) * 64 + lane], Fk,
                                                           accV[tt], 0, 0, 0);
      }
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const double ur = accV[(4 * q) / 16][((4 * q) % 16) / 4];
        const double us = accV[(NPR + 4 * q) / 16][((NPR + 4 * q) % 16) / 4];
        const double ut = accV[(2 * NPR + 4 * q) / 16][((2 * NPR + 4 * q) % 16) / 4];
        vout[q] = fma(gc, ut, fma(gb, us, ga * ur));
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    double u0v[PV];
    if (MODE == 0) {
#pragma unroll
      for (int j = 0; j < PV; ++j)
        u0v[j] = __builtin_nontemporal_load(p.u0 + dcur.vbase + min(lane + 64 * j, cnt * NP - 1));
    }
    __builtin_amdgcn_sched_barrier(0);
    // Surface: Lax-Friedrichs jumps at point s of face gq, then LIFT G.
    double4_t accL[MTL];
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt) accL[tt] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KSL; ++s) {
      const int nm = (own[s / 5] >> (6 * (s % 5))) & 63;
      const double uM = su[nm], up = uP[s];
      const double al = fmax(fabs(uM), fabs(up));
      // Fscale (n.f(u-) - n.f*) with f(u) = u^2/2 (1,1,1), factored
      const double gm = (uM - up) * fma(fa, uM + up, -fb * al);
#pragma unroll
      for (int tt = 0; tt < MTL; ++tt) {
        if (L4 && tt == MTL - 1) {
          if (REGA)
            accL[tt][0] = __builtin_amdgcn_mfma_f64_4x4x4f64(ra[(MTV * KSV + tt * KSL + s) % NFR],
                                                             gm, accL[tt][0], 0, 0, 0);
          else
            mfma_f64_rows4(accL[tt], sAl + (tt * KSL + s) * 64, lane, gm);
        } else {
          accL[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              REGA ? ra[(MTV * KSV + tt * KSL + s) % NFR] : sAl[(tt * KSL + s) * 64 + lane], gm,
              accL[tt], 0, 0, 0);
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const double lf = (4 * q) / 16 < MTL ? accL[((4 * q) / 16) % MTL][((4 * q) % 16) / 4] : 0.0;
      const double rhs = vout[q] + lf;
      const int i = gq + 4 * q;
      vout[q] = MODE == 2 ? rhs : fma(dt, rhs, su[i < NP ? i : NP - 1]);
    }
    wave_sync();  // all reads of sU as u are done
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int i = gq + 4 * q;
      const bool ok = (NPR == NP || i < NP) && (FULL || c < cnt);
      sU[ok ? c * S + i : E * S - 1] = vout[q];
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      const int e = i / NP, n = i - e * NP;
      const bool inb = E * NP % 64 == 0 || i < E * NP;
      double o = sU[inb ? e * S + n : E * S - 1];
      if (MODE == 0) o = fma(p.a0, u0v[j], p.a1 * o);
      if (MODE == 1) o = p.a1 * o;
      if (FULL ? inb : i < cnt * NP) {
        __builtin_nontemporal_store(o, p.out + dcur.vbase + i);
        lmax = fmax(lmax, fabs(o));
      }
    }
    dcur = dnxt;
    dnxt = tile_uniform(d2);
#pragma unroll
    for (int j = 0; j < PV; ++j) in.u[j] = nxt.u[j];
    in.packed = nxt.packed;
    in.nb = nxt.nb;
  }
}

// Deep variant of the pipeline (two waves per SIMD, twice the registers):
// every load of tile t+1 -- nodal values, the neighbour gathers, geometry and
// u0 -- is issued at the start of tile t, with the neighbour records of tile
// t+2 and the tile record of t+3 behind them, so the whole compute of tile t
// covers their latency. Same arithmetic as pipe_tiles.
template <int N, int MODE>
struct DeepIn {
  double u[PipeShape<N>::PV];
  double uP[PipeShape<N>::KSL];
  double u0[MODE == 0 ? PipeShape<N>::PV : 1];
  double ga, gb, gc;
};

// Issues the loads of tile d. The neighbour traces are gathered along the
// face points: the tile's 64 face slots (slot 16 f + c: face f of element
// c) times Nfp points are Nfp loads per lane, pair lane + 64 j at load j, so
// one load touches the nodes of a few neighbours instead of one per lane.
// The slots' neighbour offsets and orientation rows go through the wave's
// LDS table first; the values go back to one face per lane through LDS
// when the tile is computed.
template <int N, int MODE>
__device__ __forceinline__ void deep_issue(const StageParams& p, const Tile& d, int lane,
                                           int32_t nb, int32_t packed, const uint8_t* sFN,
                                           int2* sMeta, DeepIn<N, MODE>& x) {
  using SH = PipeShape<N>;
  constexpr int NP = SH::NP, NFP = SH::NFP, KSL = SH::KSL, PV = SH::PV;
  const int c = lane & 15, gq = lane >> 4;
  const int last = d.count * NP - 1;
#pragma unroll
  for (int j = 0; j < PV; ++j) x.u[j] = p.u[d.vbase + min(lane + 64 * j, last)];
  wave_sync();  // previous readers of the table are done
  sMeta[gq * 16 + c] = make_int2(nb, (gq * 24 + ((packed >> (5 * gq)) & 31)) * NFP);
  wave_sync();
#pragma unroll
  for (int j = 0; j < KSL; ++j) {
    const int pr = lane + 64 * j, fs = pr / NFP, q = pr - fs * NFP;
    const int2 m = sMeta[fs];
    x.uP[j] = p.u[m.x + sFN[m.y + q]];
  }
  const int64_t K = p.K;
  const int64_t g = (int64_t)d.gbase + min(c, d.count - 1);
  x.ga = p.bgeo[0 * K + g];
  x.gb = p.bgeo[1 * K + g];
  x.gc = p.bgeo[2 * K + g];
  if (MODE == 0) {
#pragma unroll
    for (int j = 0; j < PV; ++j)
      x.u0[j] = __builtin_nontemporal_load(p.u0 + d.vbase + min(lane + 64 * j, last));
  }
}

__device__ __forceinline__ void deep_meta(const StageParams& p, const Tile& d, int lane,
                                          int32_t& nb, int32_t& packed) {
  const int64_t g = (int64_t)d.gbase + min(lane & 15, d.count - 1);
  packed = p.codes[g];
  nb = p.nbr[(lane >> 4) * p.K + g];
}

// DEPTH 1: the loads of tile t+1 are in flight during tile t. DEPTH 2: those
// of t+1 and t+2 (stage 1, which reads no u0, has the registers for it).
template <int N, int MODE, bool FULL, int DEPTH = 1>
__device__ __forceinline__ void deep_tiles(const StageParams& p, const double* sA, double* sU,
                                           const uint8_t* sFN, double* sTr, int2* sMeta, int t0,
                                           int t1, int ts, double dt, double& lmax) {
  using SH = PipeShape<N>;
  constexpr int NP = SH::NP, NFP = SH::NFP, S = SH::S, NPR = SH::NPR;
  constexpr int KSV = SH::KSV, MTV = SH::MTV, MTL = SH::MTL, KSL = SH::KSL, QN = SH::QN;
  constexpr int E = SH::E, PV = SH::PV;
  constexpr FacePack<N> FPK{};
  constexpr int NW = FacePack<N>::NW;
  if (t0 >= t1) return;
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, gq = lane >> 4;
  const double* sAv = sA;
  const double* sAl = sA + MTV * KSV * 64;
  const double* su = sU + c * S;
  uint32_t own[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w)
    own[w] = gq == 0 ? FPK.w[0][w] : gq == 1 ? FPK.w[1][w] : gq == 2 ? FPK.w[2][w] : FPK.w[3][w];

  static_assert(DEPTH == 1 || DEPTH == 2, "prefetch depth 1 or 2");
  Tile dcur = tile_uniform(load_tile_raw(p.tiles, t0));
  Tile dn1 = tile_uniform(load_tile_raw(p.tiles, min(t0 + ts, t1 - 1)));
  Tile dn2 = tile_uniform(load_tile_raw(p.tiles, min(t0 + 2 * ts, t1 - 1)));
  Tile dn3 = DEPTH == 2 ? tile_uniform(load_tile_raw(p.tiles, min(t0 + 3 * ts, t1 - 1))) : dn2;
  int32_t nb0, pk0, nb1, pk1;
  deep_meta(p, dcur, lane, nb0, pk0);
  DeepIn<N, MODE> cur, nx1;
  deep_issue<N, MODE>(p, dcur, lane, nb0, pk0, sFN, sMeta, cur);
  deep_meta(p, dn1, lane, nb1, pk1);
  if constexpr (DEPTH == 2) {
    // tile t+1 in flight from the start; the meta of t+2 ready
    deep_issue<N, MODE>(p, dn1, lane, nb1, pk1, sFN, sMeta, nx1);
    deep_meta(p, dn2, lane, nb1, pk1);
  }
  for (int t = t0; t < t1; t += ts) {
    const int4 d3 = load_tile_raw(p.tiles, min(t + (DEPTH + 2) * ts, t1 - 1));
    const int cnt = FULL ? E : dcur.count;
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      const int e = i / NP, n = i - e * NP;
      const bool ok = (E * NP % 64 == 0 || i < E * NP) && (FULL || i < cnt * NP);
      sU[ok ? e * S + n : E * S - 1] = cur.u[j];
    }
    // neighbour traces, pair lane + 64 j (read back one face per lane below)
#pragma unroll
    for (int j = 0; j < KSL; ++j) sTr[lane + 64 * j] = cur.uP[j];
    __builtin_amdgcn_sched_barrier(0);
    // Loads of tile t+DEPTH (its meta was read a tile ago), meta of t+DEPTH+1.
    DeepIn<N, MODE> nx;
    deep_issue<N, MODE>(p, DEPTH == 2 ? dn2 : dn1, lane, nb1, pk1, sFN, sMeta, nx);
    int32_t nb2, pk2;
    deep_meta(p, DEPTH == 2 ? dn3 : dn2, lane, nb2, pk2);
    __builtin_amdgcn_sched_barrier(0);
    wave_sync();  // sU and sTr hold this tile
    double up[KSL];
#pragma unroll
    for (int s = 0; s < KSL; ++s) up[s] = sTr[(gq * 16 + c) * NFP + s];
    // A wave in its matrix phase issues first (wave priority): the other wave
    // of the SIMD does its loads and VALU work in the MFMA shadow.
    __builtin_amdgcn_s_setprio(1);
    double vout[QN];
    {
      double4_t accV[MTV];
#pragma unroll
      for (int tt = 0; tt < MTV; ++tt) accV[tt] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KSV; ++s) {
        const int k = 4 * s + gq;
        const double uk = su[k < NP ? k : NP - 1];
        const double Fk = k < NP ? uk * uk : 0.0;  // the operator carries -1/2
#pragma unroll
        for (int tt = 0; tt < MTV; ++tt) {
          if (3 * NPR - 16 * tt < 16)  // a last row tile with padding rows
            mfma_f64_row_groups(accV[tt], sAv + (tt * KSV + s) * 64, lane, Fk,
                                (3 * NPR - 16 * tt + 3) / 4);
          else
            accV[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sAv[(tt * KSV + s) * 64 + lane], Fk,
                                                             accV[tt], 0, 0, 0);
        }
      }
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const double ur = accV[(4 * q) / 16][((4 * q) % 16) / 4];
        const double us = accV[(NPR + 4 * q) / 16][((NPR + 4 * q) % 16) / 4];
        const double ut = accV[(2 * NPR + 4 * q) / 16][((2 * NPR + 4 * q) % 16) / 4];
        vout[q] = fma(cur.gc, ut, fma(cur.gb, us, cur.ga * ur));
      }
    }
    double4_t accL[MTL];
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt) accL[tt] = (double4_t){0.0, 0.0, 0.0, 0.0};
    double fa, fb;
    dgk::face_factors(gq, cur.ga, cur.gb, cur.gc, fa, fb);
#pragma unroll
    for (int s = 0; s < KSL; ++s) {
      const int nm = (own[s / 5] >> (6 * (s % 5))) & 63;
      const double uM = su[nm];
      const double al = fmax(fabs(uM), fabs(up[s]));
      const double gm = (uM - up[s]) * fma(fa, uM + up[s], -fb * al);
#pragma unroll
      for (int tt = 0; tt < MTL; ++tt) {
        if (NP - 16 * tt <= 4)
          mfma_f64_rows4(accL[tt], sAl + (tt * KSL + s) * 64, lane, gm);
        else
          accL[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(sAl[(tt * KSL + s) * 64 + lane], gm,
                                                           accL[tt], 0, 0, 0);
      }
    }
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const double lf = (4 * q) / 16 < MTL ? accL[((4 * q) / 16) % MTL][((4 * q) % 16) / 4] : 0.0;
      const double rhs = vout[q] + lf;
      const int i = gq + 4 * q;
      vout[q] = MODE == 2 ? rhs : fma(dt, rhs, su[i < NP ? i : NP - 1]);
    }
    wave_sync();  // all reads of sU as u are done
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const int i = gq + 4 * q;
      const bool ok = (NPR == NP || i < NP) && (FULL || c < cnt);
      sU[ok ? c * S + i : E * S - 1] = vout[q];
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < PV; ++j) {
      const int i = lane + 64 * j;
      const int e = i / NP, n = i - e * NP;
      const bool inb = E * NP % 64 == 0 || i < E * NP;
      double o = sU[inb ? e * S + n : E * S - 1];
      if (MODE == 0) o = fma(p.a0, cur.u0[j], p.a1 * o);
      if (MODE == 1) o = p.a1 * o;
      if (FULL ? inb : i < cnt * NP) {
        __builtin_nontemporal_store(o, p.out + dcur.vbase + i);
        lmax = fmax(lmax, fabs(o));
      }
    }
    if constexpr (DEPTH == 2) {
      cur = nx1;
      nx1 = nx;
    } else {
      cur = nx;
    }
    nb1 = nb2;
    pk1 = pk2;
    dcur = dn1;
    dn1 = dn2;
    if constexpr (DEPTH == 2) {
      dn2 = dn3;
      dn3 = tile_uniform(d3);
    } else {
      dn2 = tile_uniform(d3);
    }
  }
}

// Tiles [0, nfull) hold 16 elements each; [nfull, ntiles) are the partial
// last tiles of partitions.
// KIND 0: pipe_tiles (3 waves per SIMD); 1: deep_tiles (2 waves per SIMD);
// 2: pipe_tiles with register-resident operator fragments (2 waves per SIMD).
template <int N, int WPB, int MODE, int KIND>
__global__ void __launch_bounds__(64 * WPB, KIND == 0 ? 3 : 2)
    burgers_stage_pipe_kernel(StageParams p, int nfull) {
  using SH = PipeShape<N>;
  constexpr int NP = SH::NP, NFP = SH::NFP, NF = SH::NF, NPR = SH::NPR;
  constexpr int KSV = SH::KSV, MTV = SH::MTV, KSL = SH::KSL;
  constexpr int NT = 64 * WPB;
  __shared__ __attribute__((aligned(16))) double sA[SH::NA];
  __shared__ __attribute__((aligned(16))) double sUall[WPB][SH::E * SH::S];
  // deep pipeline: neighbour traces (64 face slots x Nfp) and the face
  // slots' neighbour offsets and orientation rows, per wave
  __shared__ __attribute__((aligned(16))) double sTrall[KIND == 1 ? WPB : 1][KIND == 1 ? 64 * NFP : 1];
  __shared__ int2 sMetaall[KIND == 1 ? WPB : 1][64];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  for (int i = threadIdx.x; i < 4 * 24 * NFP; i += NT) sFN[i] = p.face_nodes[i];
  // Operator fragments, entry (tile, kstep, lane) in lane order. Volume: the
  // stacked -1/2 [Dr; Ds; Dt]. LIFT: column k = 4 fp + f of k step fp is
  // face f, point fp.
  for (int q = threadIdx.x; q < SH::NA; q += NT) {
    const int l = q & 63, ts_ = q >> 6, cl = l & 15, gl = l >> 4;
    double v = 0.0;
    if (ts_ < MTV * KSV) {
      const int tt = ts_ / KSV, s = ts_ % KSV;
      const int row = 16 * tt + cl, k = 4 * s + gl, blk = row / NPR, i = row % NPR;
      if (k < NP && blk < 3 && i < NP)
        v = -0.5 * (blk == 0 ? p.Dr : blk == 1 ? p.Ds : p.Dt)[i * NP + k];  // exact scaling
    } else {
      const int r2 = ts_ - MTV * KSV, tt = r2 / KSL, s = r2 % KSL;
      const int row = 16 * tt + cl;
      if (row < NP) v = p.LIFT[row * NF + gl * NFP + s];
    }
    sA[q] = v;
  }
  __syncthreads();
  const double dt = MODE == 2 ? 0.0 : *p.dt;
  double lmax = 0.0;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int NW = gridDim.x * WPB;
  const int gw = blockIdx.x * WPB + w;
  double* sU = sUall[w];
  int f0 = gw, f1 = nfull, fs = NW;
  if (p.chunked && gridDim.x % 8 == 0) {
    // Blocks are dealt round robin over the 8 XCDs (block b on XCD b % 8).
    // Each XCD takes a contiguous eighth of the full tiles and its waves
    // sweep it together, so tiles processed at the same time on one XCD are
    // neighbours in the tile list and their face neighbours (a few tiles to
    // a few dozen tiles away) are read from that XCD's L2.
    const int x = blockIdx.x % 8;
    const int lw = (blockIdx.x / 8) * WPB + w;
    f0 = (int)((int64_t)nfull * x / 8) + lw;
    f1 = (int)((int64_t)nfull * (x + 1) / 8);
    fs = NW / 8;
  }
  if (KIND == 1) {
    double* sTr = sTrall[KIND == 1 ? w : 0];
    int2* sMeta = sMetaall[KIND == 1 ? w : 0];
    // One tile of loads in flight in every mode: with the gathers routed
    // through LDS, stage 1 ran 3.5 % faster than with two (673.6 against
    // 697.7 us at P=3, 5.3 M tets), which the DEPTH 2 form gave before.
    deep_tiles<N, MODE, true, 1>(p, sA, sU, sFN, sTr, sMeta, f0, f1, fs, dt, lmax);
    deep_tiles<N, MODE, false>(p, sA, sU, sFN, sTr, sMeta, nfull + gw, p.ntiles, NW, dt, lmax);
  } else {
    pipe_tiles<N, MODE, true, KIND == 2>(p, sA, sU, sFN, f0, f1, fs, dt, lmax);
    pipe_tiles<N, MODE, false, false>(p, sA, sU, sFN, nfull + gw, p.ntiles, NW, dt, lmax);
  }
  if (MODE != 2 && p.maxabs) {
    lmax = wave_max(lmax);
    if ((threadIdx.x & 63) == 0 && lmax > 0.0)
      atomicMax(p.maxabs, (unsigned long long)__double_as_longlong(lmax));
  }
}

__global__ void max_abs_kernel(const Tile* __restrict__ tiles, int ntiles, int np,
                               const double* __restrict__ u, unsigned long long* maxabs) {
  double m = 0.0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    Tile tl = tiles[t];
    for (int i = threadIdx.x; i < tl.count * np; i += blockDim.x) m = fmax(m, fabs(u[tl.vbase + i]));
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(maxabs, (unsigned long long)__double_as_longlong(m));
}

__global__ void cfl_dt_kernel(double* maxabs, double cfl, double hmin, double* dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double m = *maxabs;
    *dt = cfl * hmin / m;
    *maxabs = 0.0;
  }
}

// Minimum distance between distinct nodes of an element (node coordinates
// by the affine map of startup3d.go:126-148), reduced over all elements.
template <int NP>
__global__ void min_node_distance_kernel(int64_t K, const double* __restrict__ vtx,
                                         const double* __restrict__ rst,
                                         unsigned long long* out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double best = INFINITY;
  if (k < K) {
    const double* v = vtx + k * 12;
    double X[NP], Y[NP], Z[NP];
    for (int i = 0; i < NP; ++i) {
      double r = rst[i], s = rst[NP + i], t = rst[2 * NP + i];
      double la = -(1.0 + r + s + t), lb = 1.0 + r, lc = 1.0 + s, ld = 1.0 + t;
      X[i] = 0.5 * (la * v[0] + lb * v[3] + lc * v[6] + ld * v[9]);
      Y[i] = 0.5 * (la * v[1] + lb * v[4] + lc * v[7] + ld * v[10]);
      Z[i] = 0.5 * (la * v[2] + lb * v[5] + lc * v[8] + ld * v[11]);
    }
    for (int i = 0; i < NP; ++i)
      for (int j = i + 1; j < NP; ++j) {
        double dx = X[j] - X[i], dy = Y[j] - Y[i], dz = Z[j] - Z[i];
        double len = sqrt(dx * dx + dy * dy + dz * dz);
        if (len > 1e-12 && len < best) best = len;
      }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) best = fmin(best, __shfl_xor(best, o, 64));
  if ((threadIdx.x & 63) == 0 && best < INFINITY)
    atomicMin(out, (unsigned long long)__double_as_longlong(best));
}

}  // namespace


namespace {

// Tiles of per_tile elements over all partitions, full tiles first (in
// partition order), then the partial last tiles of partitions.
int build_tile_list(dgk_dgmesh m, int per_tile, Tile** out, int* nout, int* nfull) {
  std::vector<Tile> full, part;
  int64_t g = 0;
  for (int p = 0; p < m->npart; ++p) {
    for (int64_t f = 0; f < m->Kp[p]; f += per_tile) {
      Tile t;
      t.vbase = m->state_off[p] + f * m->Np;
      t.gbase = (int32_t)(g + f);
      t.count = (int32_t)std::min<int64_t>(per_tile, m->Kp[p] - f);
      (t.count == per_tile ? full : part).push_back(t);
    }
    g += m->Kp[p];
  }
  if (m->tile_order == 1 && !m->ekey.empty()) {
    // Full tiles of all partitions along the Morton curve of the centroid of
    // each tile's middle element: tiles that run at the same time on one XCD
    // then cover a compact region, across partition boundaries.
    std::stable_sort(full.begin(), full.end(), [&](const Tile& x, const Tile& y) {
      return m->ekey[x.gbase + x.count / 2] < m->ekey[y.gbase + y.count / 2];
    });
  }
  if (nfull) *nfull = (int)full.size();
  full.insert(full.end(), part.begin(), part.end());
  if (*out) dgk_free(m->dev, *out);
  *out = nullptr;
  *nout = (int)full.size();
  if (full.empty()) return DGK_OK;
  return dgk_malloc(m->dev, full.size() * sizeof(Tile), full.data(), (void**)out);
}

int build_tiles(dgk_dgmesh m) {
  int rc = build_tile_list(m, m->tb, &m->tiles, &m->ntiles, nullptr);
  if (rc == DGK_OK) rc = build_tile_list(m, 16, &m->tiles64, &m->ntiles64, &m->nfull64);
  return rc;
}

template <int N, int TB>
int launch_stage_t(dgk_dgmesh m, const StageParams& p) {
  int grid = m->ntiles;
  int bpc = m->blocks_per_cu > 0 ? m->blocks_per_cu : (TB == 64 ? 8 : (TB == 128 ? 4 : 2));
  int cap = m->dev->cu_count * bpc;
  if (grid > cap) grid = cap;
  if (grid < 1) return DGK_OK;
  if (p.chunked && grid >= 8) grid -= grid % 8;
  hipLaunchKernelGGL((burgers_stage_kernel<N, TB>), dim3(grid), dim3(TB), 0, m->dev->stream, p);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <int N>
int launch_stage_n(dgk_dgmesh m, const StageParams& p) {
  switch (m->tb) {
    case 64: return launch_stage_t<N, 64>(m, p);
    case 128: return launch_stage_t<N, 128>(m, p);
    case 256: return launch_stage_t<N, 256>(m, p);
    default:
      dgk::set_error("burgers stage: unsupported tile size %d", m->tb);
      return DGK_ERR_INVALID;
  }
}

template <int N, int WPB>
int launch_stage_mfma_w(dgk_dgmesh m, const StageParams& p) {
  int grid = (p.ntiles + WPB - 1) / WPB;
  int bpc = m->blocks_per_cu > 0 ? m->blocks_per_cu : (WPB == 4 ? 3 : WPB == 2 ? 6 : 8);
  int cap = m->dev->cu_count * bpc;
  if (grid > cap) grid = cap;
  if (grid < 1) return DGK_OK;
  if (p.chunked && grid >= 8) grid -= grid % 8;
  hipLaunchKernelGGL((burgers_stage_mfma_kernel<N, WPB>), dim3(grid), dim3(64 * WPB), 0,
                     m->dev->stream, p);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <int N>
int launch_stage_mfma(dgk_dgmesh m, const StageParams& p) {
  switch (m->tbm) {
    case 16: return launch_stage_mfma_w<N, 1>(m, p);
    case 32: return launch_stage_mfma_w<N, 2>(m, p);
    default: return launch_stage_mfma_w<N, 4>(m, p);
  }
}

template <int N, int WPB, int MODE, int KIND>
int launch_stage_pipe_wm(dgk_dgmesh m, const StageParams& p) {
  int grid = (p.ntiles + WPB - 1) / WPB;
  const int waves_per_simd = KIND == 0 ? 3 : 2;
  int bpc = m->blocks_per_cu > 0 ? m->blocks_per_cu : (4 * waves_per_simd) / WPB;
  int cap = m->dev->cu_count * bpc;
  if (grid > cap) grid = cap;
  if (grid < 1) return DGK_OK;
  if (p.chunked && grid >= 8) grid -= grid % 8;
  hipLaunchKernelGGL((burgers_stage_pipe_kernel<N, WPB, MODE, KIND>), dim3(grid), dim3(64 * WPB),
                     0, m->dev->stream, p, m->nfull64);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <int N, int WPB, int KIND>
int launch_stage_pipe_w(dgk_dgmesh m, const StageParams& p) {
  if (p.mode == 1) return launch_stage_pipe_wm<N, WPB, 2, KIND>(m, p);
  if (p.a0 != 0.0) return launch_stage_pipe_wm<N, WPB, 0, KIND>(m, p);
  return launch_stage_pipe_wm<N, WPB, 1, KIND>(m, p);
}

template <int N, int KIND>
int launch_stage_pipe(dgk_dgmesh m, const StageParams& p) {
  switch (m->tbm) {
    case 32: return launch_stage_pipe_w<N, 2, KIND>(m, p);
    default: return launch_stage_pipe_w<N, 4, KIND>(m, p);
  }
}

bool use_mfma(dgk_dgmesh m, int mode) {
  if (mode == 2 || m->N > 4) return false;
  return m->variant != 1 && m->variant != 3;
}

bool use_split(dgk_dgmesh m, int mode) {
  return mode != 2 && (m->N > 4 || m->prec != 0 || m->variant == 3);
}

int launch_stage(dgk_dgmesh m, StageParams p) {
  if (use_split(m, p.mode))
    return dgk::split_stage(m, p.u, p.u0, p.out, p.a0, p.a1, p.dt, p.maxabs, p.mode);
  p.Dr = m->Dr; p.Ds = m->Ds; p.Dt = m->Dt; p.LIFT = m->LIFT;
  p.face_nodes = m->face_nodes;
  p.bgeo = m->bgeo;
  p.nbr = m->nbr;
  p.codes = m->codes;
  p.tiles = m->tiles;
  p.ntiles = m->ntiles;
  p.chunked = m->chunked;
  p.K = m->K;
  if (use_mfma(m, p.mode)) {
    p.tiles = m->tiles64;
    p.ntiles = m->ntiles64;
    if (m->variant == 4) {
      switch (m->N) {
        case 1: return launch_stage_mfma<1>(m, p);
        case 2: return launch_stage_mfma<2>(m, p);
        case 3: return launch_stage_mfma<3>(m, p);
        case 4: return launch_stage_mfma<4>(m, p);
      }
    }
    if (m->variant == 5) {
      switch (m->N) {
        case 1: return launch_stage_pipe<1, 1>(m, p);
        case 2: return launch_stage_pipe<2, 1>(m, p);
        case 3: return launch_stage_pipe<3, 1>(m, p);
      }
    }
    if (m->variant == 6) {
      switch (m->N) {
        case 1: return launch_stage_pipe<1, 2>(m, p);
        case 2: return launch_stage_pipe<2, 2>(m, p);
        case 3: return launch_stage_pipe<3, 2>(m, p);
      }
    }
    switch (m->N) {
      case 1: return launch_stage_pipe<1, 0>(m, p);
      case 2: return launch_stage_pipe<2, 0>(m, p);
      case 3: return launch_stage_pipe<3, 0>(m, p);
      case 4: return launch_stage_pipe<4, 0>(m, p);
    }
  }
  switch (m->N) {
    case 1: return launch_stage_n<1>(m, p);
    case 2: return launch_stage_n<2>(m, p);
    case 3: return launch_stage_n<3>(m, p);
    case 4: return launch_stage_n<4>(m, p);
    default:
      dgk::set_error("burgers stage: the fused kernel and the trace are built for N <= 4 (N=%d)",
                     m->N);
      return DGK_ERR_UNSUPPORTED;
  }
}

}  // namespace

extern "C" int dgk_dgmesh_create(dgk_device dev, const dgk_dgmesh_desc* d, dgk_dgmesh* out) {
  dgk::clear_error();
  DGK_REQUIRE(dev && d && out, "dgk_dgmesh_create: NULL argument");
  *out = nullptr;
  DGK_REQUIRE(d->N >= 1 && d->N <= 6, "dgk_dgmesh_create: N=%d outside 1..6", d->N);
  DGK_REQUIRE(d->npart >= 1, "dgk_dgmesh_create: npart must be >= 1");
  DGK_REQUIRE(d->K && d->state_off && d->Dr && d->Ds && d->Dt && d->LIFT && d->Fmask &&
                  d->face_nodes && d->EToE && d->codes && d->vtx && d->r && d->s && d->t,
              "dgk_dgmesh_create: NULL field in descriptor");
  const int N = d->N, Np = dgk::np_of(N), Nfp = dgk::nfp_of(N);
  std::vector<int> fm = host_fmask(N);
  for (int i = 0; i < 4 * Nfp; ++i)
    DGK_REQUIRE(fm[i] == d->Fmask[i],
                "dgk_dgmesh_create: Fmask differs from the NUDG lattice ordering at %d", i);
  dgk_dgmesh m = new dgk_dgmesh_s();
  m->dev = dev;
  m->N = N; m->Np = Np; m->Nfp = Nfp; m->npart = d->npart;
  m->Kp.assign(d->K, d->K + d->npart);
  m->state_off.assign(d->state_off, d->state_off + d->npart + 1);
  std::vector<int64_t> gstart(d->npart + 1, 0);
  for (int p = 0; p < d->npart; ++p) {
    if (m->Kp[p] < 0) { delete m; dgk::set_error("negative K"); return DGK_ERR_INVALID; }
    gstart[p + 1] = gstart[p] + m->Kp[p];
  }
  const int64_t K = gstart[d->npart];
  m->K = K;
  if (K >= (1LL << 31) || m->state_off[d->npart] >= (1LL << 31)) {
    delete m;
    dgk::set_error("dgk_dgmesh_create: mesh too large for 32-bit neighbour offsets");
    return DGK_ERR_UNSUPPORTED;
  }
  // Neighbour value offsets and packed codes, SoA.
  std::vector<int32_t> nbr((size_t)4 * K), codes((size_t)K), nbe((size_t)4 * K);
  int p2 = 0;
  for (int64_t g = 0; g < K; ++g) {
    int packed = 0;
    for (int f = 0; f < 4; ++f) {
      int64_t g2 = d->EToE[g * 4 + f];
      int c = d->codes[g * 4 + f];
      if (g2 < 0 || g2 >= K || c < 0 || c >= 24) {
        delete m;
        dgk::set_error("dgk_dgmesh_create: bad neighbour/code at element %lld face %d", (long long)g, f);
        return DGK_ERR_INVALID;
      }
      // partition of g2 (partition-major numbering)
      if (!(g2 >= gstart[p2] && g2 < gstart[p2 + 1])) {
        p2 = (int)(std::upper_bound(gstart.begin(), gstart.end(), g2) - gstart.begin()) - 1;
      }
      nbr[(size_t)f * K + g] = (int32_t)(m->state_off[p2] + (g2 - gstart[p2]) * Np);
      nbe[(size_t)f * K + g] = (int32_t)g2;
      packed |= c << (5 * f);
    }
    codes[g] = packed;
  }
  // Morton keys of the element centroids (10 bits per axis over the mesh's
  // bounding box), for the spatial tile order.
  {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    std::vector<float> cen((size_t)3 * K);
    for (int64_t g = 0; g < K; ++g)
      for (int a = 0; a < 3; ++a) {
        const double* v = d->vtx + g * 12;
        const double c = 0.25 * (v[a] + v[3 + a] + v[6 + a] + v[9 + a]);
        cen[(size_t)3 * g + a] = (float)c;
        lo[a] = std::min(lo[a], c);
        hi[a] = std::max(hi[a], c);
      }
    auto spread = [](uint32_t x) {  // bit i of x to bit 3 i
      x &= 0x3ff;
      x = (x | (x << 16)) & 0x30000ff;
      x = (x | (x << 8)) & 0x300f00f;
      x = (x | (x << 4)) & 0x30c30c3;
      x = (x | (x << 2)) & 0x9249249;
      return x;
    };
    m->ekey.resize((size_t)K);
    for (int64_t g = 0; g < K; ++g) {
      uint32_t q[3];
      for (int a = 0; a < 3; ++a) {
        const double span = hi[a] - lo[a];
        const double t = span > 0 ? (cen[(size_t)3 * g + a] - lo[a]) / span : 0.0;
        q[a] = (uint32_t)std::min(1023.0, std::max(0.0, t * 1023.0));
      }
      m->ekey[(size_t)g] = spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2);
    }
  }
  std::vector<uint8_t> fn((size_t)4 * 24 * Nfp);
  for (size_t i = 0; i < fn.size(); ++i) {
    int v = d->face_nodes[i];
    if (v < 0 || v >= Np) {
      delete m;
      dgk::set_error("dgk_dgmesh_create: face node table entry %d out of range", v);
      return DGK_ERR_INVALID;
    }
    fn[i] = (uint8_t)v;
  }
  std::vector<double> rst(3 * Np);
  for (int i = 0; i < Np; ++i) { rst[i] = d->r[i]; rst[Np + i] = d->s[i]; rst[2 * Np + i] = d->t[i]; }
  m->Dr_h.assign(d->Dr, d->Dr + (size_t)Np * Np);
  m->Ds_h.assign(d->Ds, d->Ds + (size_t)Np * Np);
  m->Dt_h.assign(d->Dt, d->Dt + (size_t)Np * Np);
  m->LIFT_h.assign(d->LIFT, d->LIFT + (size_t)Np * 4 * Nfp);
  m->r_h.assign(d->r, d->r + Np);
  m->s_h.assign(d->s, d->s + Np);
  m->t_h.assign(d->t, d->t + Np);
  int rc = DGK_OK;
  rc |= dgk_malloc(dev, (size_t)Np * Np * 8, d->Dr, (void**)&m->Dr);
  rc |= dgk_malloc(dev, (size_t)Np * Np * 8, d->Ds, (void**)&m->Ds);
  rc |= dgk_malloc(dev, (size_t)Np * Np * 8, d->Dt, (void**)&m->Dt);
  rc |= dgk_malloc(dev, (size_t)Np * 4 * Nfp * 8, d->LIFT, (void**)&m->LIFT);
  rc |= dgk_malloc(dev, rst.size() * 8, rst.data(), (void**)&m->rst);
  rc |= dgk_malloc(dev, fn.size(), fn.data(), (void**)&m->face_nodes);
  rc |= dgk_malloc(dev, (size_t)K * 12 * 8 + 8, d->vtx, (void**)&m->vtx);
  rc |= dgk_malloc(dev, (size_t)K * kGeoRows * 8 + 8, nullptr, (void**)&m->geo);
  rc |= dgk_malloc(dev, (size_t)K * kBGeoRows * 8 + 8, nullptr, (void**)&m->bgeo);
  rc |= dgk_malloc(dev, nbr.size() * 4 + 4, nbr.data(), (void**)&m->nbr);
  rc |= dgk_malloc(dev, codes.size() * 4 + 4, codes.data(), (void**)&m->codes);
  rc |= dgk_malloc(dev, nbe.size() * 4 + 4, nbe.data(), (void**)&m->nbe);
  rc |= dgk_malloc(dev, 64, nullptr, (void**)&m->scratch);
  if (rc == DGK_OK) rc = build_tiles(m);
  if (rc != DGK_OK) { dgk_dgmesh_destroy(m); return DGK_ERR_HIP; }
  // Geometry on the device (FP64), with a count of non-positive Jacobians.
  int* bad = (int*)m->scratch;
  rc = dgk_fill_bytes(dev, m->scratch, 0, 64);
  if (rc == DGK_OK && K > 0) {
    hipLaunchKernelGGL(geometry_kernel, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, dev->stream,
                       K, m->vtx, m->geo, m->bgeo, bad);
    if (hipGetLastError() != hipSuccess) rc = DGK_ERR_HIP;
  }
  int nbad = 0;
  if (rc == DGK_OK) rc = dgk_copy_d2h(dev, &nbad, bad, sizeof(int), 0);
  if (rc != DGK_OK) { dgk_dgmesh_destroy(m); return rc; }
  if (nbad) {
    dgk_dgmesh_destroy(m);
    dgk::set_error("dgk_dgmesh_create: %d elements with non-positive Jacobian", nbad);
    return DGK_ERR_NUMERIC;
  }
  if (use_split(m, 0)) {
    rc = dgk::split_prepare(m);
    if (rc != DGK_OK) {
      dgk_dgmesh_destroy(m);
      return rc;
    }
  }
  *out = m;
  return DGK_OK;
}

extern "C" int dgk_dgmesh_destroy(dgk_dgmesh m) {
  if (!m) return DGK_OK;
  dgk_device dev = m->dev;
  dgk::split_release(m);
  void* ptrs[] = {m->Dr, m->Ds, m->Dt, m->LIFT, m->rst, m->face_nodes, m->vtx, m->geo,
                  m->bgeo, m->nbr, m->codes, m->nbe, m->tiles, m->tiles64, m->scratch};
  for (void* q : ptrs) dgk_free(dev, q);
  delete m;
  return DGK_OK;
}

extern "C" int dgk_dgmesh_geometry(dgk_dgmesh m, double* geo_host) {
  DGK_REQUIRE(m && geo_host, "dgk_dgmesh_geometry: NULL argument");
  return dgk_copy_d2h(m->dev, geo_host, m->geo, (size_t)m->K * kGeoRows * 8, 0);
}

extern "C" int dgk_dgmesh_geometry_dev(dgk_dgmesh m, const double** geo_dev) {
  DGK_REQUIRE(m && geo_dev, "dgk_dgmesh_geometry_dev: NULL argument");
  *geo_dev = m->geo;
  return DGK_OK;
}

extern "C" int dgk_dgmesh_min_node_distance(dgk_dgmesh m, double* hmin) {
  dgk::clear_error();
  DGK_REQUIRE(m && hmin, "dgk_dgmesh_min_node_distance: NULL argument");
  unsigned long long init = (unsigned long long)0x7FF0000000000000ULL;  // +inf
  int rc = dgk_copy_h2d(m->dev, m->scratch, &init, 8, 0);
  if (rc) return rc;
  dim3 grid((unsigned)((m->K + 255) / 256)), block(256);
  switch (m->N) {
#define DGK_CASE(n)                                                                              \
  case n:                                                                                        \
    hipLaunchKernelGGL(min_node_distance_kernel<(n + 1) * (n + 2) * (n + 3) / 6>, grid, block, 0, \
                       m->dev->stream, m->K, m->vtx, m->rst, m->scratch);                        \
    break;
    DGK_CASE(1) DGK_CASE(2) DGK_CASE(3) DGK_CASE(4) DGK_CASE(5) DGK_CASE(6)
#undef DGK_CASE
    default: return DGK_ERR_UNSUPPORTED;
  }
  DGK_LAUNCH_CHECK();
  return dgk_copy_d2h(m->dev, hmin, m->scratch, 8, 0);
}

extern "C" int dgk_dgmesh_trace_plus(dgk_dgmesh m, const double* u_dev, double* uP_dev) {
  dgk::clear_error();
  DGK_REQUIRE(m && u_dev && uP_dev, "dgk_dgmesh_trace_plus: NULL argument");
  StageParams p{};
  p.u = u_dev;
  p.u0 = u_dev;
  p.out = uP_dev;
  p.mode = 2;
  return launch_stage(m, p);
}

extern "C" int dgk_burgers_stage(dgk_dgmesh m, const double* u_dev, const double* u0_dev,
                                 double* out_dev, double a0, double a1, const double* dt_dev,
                                 double* maxabs_dev) {
  dgk::clear_error();
  DGK_REQUIRE(m && u_dev && out_dev && dt_dev, "dgk_burgers_stage: NULL argument");
  DGK_REQUIRE(out_dev != u_dev, "dgk_burgers_stage: out must not alias u");
  DGK_REQUIRE(a0 == 0.0 || u0_dev, "dgk_burgers_stage: u0 is NULL with a0 != 0");
  StageParams p{};
  p.u = u_dev;
  p.u0 = u0_dev ? u0_dev : u_dev;
  p.out = out_dev;
  p.a0 = a0;
  p.a1 = a1;
  p.dt = dt_dev;
  p.maxabs = (unsigned long long*)maxabs_dev;
  p.mode = 0;
  return launch_stage(m, p);
}

extern "C" int dgk_burgers_rhs(dgk_dgmesh m, const double* u_dev, double* rhs_dev) {
  dgk::clear_error();
  DGK_REQUIRE(m && u_dev && rhs_dev, "dgk_burgers_rhs: NULL argument");
  DGK_REQUIRE(rhs_dev != u_dev, "dgk_burgers_rhs: rhs must not alias u");
  StageParams p{};
  p.u = u_dev;
  p.u0 = u_dev;
  p.out = rhs_dev;
  p.mode = 1;
  return launch_stage(m, p);
}

extern "C" int dgk_max_abs(dgk_dgmesh m, const double* u_dev, double* maxabs_dev) {
  dgk::clear_error();
  DGK_REQUIRE(m && u_dev && maxabs_dev, "dgk_max_abs: NULL argument");
  int grid = std::min(m->ntiles, m->dev->cu_count * 8);
  if (grid < 1) return DGK_OK;
  hipLaunchKernelGGL(max_abs_kernel, dim3(grid), dim3(256), 0, m->dev->stream, m->tiles, m->ntiles,
                     m->Np, u_dev, (unsigned long long*)maxabs_dev);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

extern "C" int dgk_cfl_dt(dgk_device dev, double* maxabs_dev, double cfl, double hmin,
                          double* dt_dev) {
  dgk::clear_error();
  DGK_REQUIRE(dev && maxabs_dev && dt_dev, "dgk_cfl_dt: NULL argument");
  hipLaunchKernelGGL(cfl_dt_kernel, dim3(1), dim3(64), 0, dev->stream, maxabs_dev, cfl, hmin, dt_dev);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

extern "C" int dgk_burgers_set_order(dgk_dgmesh m, int xcd_chunked) {
  dgk::clear_error();
  DGK_REQUIRE(m, "dgk_burgers_set_order: NULL mesh");
  m->chunked = xcd_chunked ? 1 : 0;
  return DGK_OK;
}

extern "C" int dgk_dgmesh_set_tile_order(dgk_dgmesh m, int order) {
  dgk::clear_error();
  DGK_REQUIRE(m, "dgk_dgmesh_set_tile_order: NULL mesh");
  DGK_REQUIRE(order == 0 || order == 1, "dgk_dgmesh_set_tile_order: order %d", order);
  if (order == m->tile_order) return DGK_OK;
  m->tile_order = order;
  return build_tiles(m);
}

extern "C" int dgk_burgers_set_variant(dgk_dgmesh m, int variant) {
  dgk::clear_error();
  DGK_REQUIRE(m, "dgk_burgers_set_variant: NULL mesh");
  DGK_REQUIRE(variant >= 0 && variant <= 6, "dgk_burgers_set_variant: variant %d", variant);
  DGK_REQUIRE((variant != 2 && variant != 4) || m->N <= 4,
              "dgk_burgers_set_variant: MFMA stage built for N <= 4");
  DGK_REQUIRE((variant != 5 && variant != 6) || m->N <= 3,
              "dgk_burgers_set_variant: variants 5, 6 of the MFMA stage are built for N <= 3");
  m->variant = variant;
  return use_split(m, 0) ? dgk::split_prepare(m) : DGK_OK;
}

extern "C" int dgk_burgers_set_tile(dgk_dgmesh m, int elems_per_block, int blocks_per_cu) {
  dgk::clear_error();
  DGK_REQUIRE(m, "dgk_burgers_set_tile: NULL mesh");
  DGK_REQUIRE(blocks_per_cu >= 0 && blocks_per_cu <= 64, "dgk_burgers_set_tile: blocks_per_cu %d",
              blocks_per_cu);
  m->blocks_per_cu = blocks_per_cu;
  if (use_mfma(m, 0)) {
    int e = elems_per_block == 0 ? 64 : elems_per_block;
    DGK_REQUIRE(e == 16 || e == 32 || e == 64, "dgk_burgers_set_tile: MFMA tile %d (16, 32, 64)", e);
    m->tbm = e;
    return DGK_OK;
  }
  int tb = elems_per_block == 0 ? 64 : elems_per_block;
  DGK_REQUIRE(tb == 64 || tb == 128 || tb == 256, "dgk_burgers_set_tile: elems_per_block %d", tb);
  if (tb != m->tb) {
    m->tb = tb;
    return build_tiles(m);
  }
  return DGK_OK;
}

extern "C" int dgk_burgers_set_precision(dgk_dgmesh m, int prec) {
  dgk::clear_error();
  DGK_REQUIRE(m, "dgk_burgers_set_precision: NULL mesh");
  const int vol = prec & ~DGK_PREC_LIFT_F64;
  DGK_REQUIRE(vol >= 0 && vol <= DGK_PREC_I8X6, "dgk_burgers_set_precision: precision %d", prec);
  const int lift = ((prec & DGK_PREC_LIFT_F64) || vol == DGK_PREC_I8X6) ? 0 : vol;
  if (vol != m->prec || lift != m->lift_prec) {
    dgk::split_release(m);
    m->prec = vol;
    m->lift_prec = lift;
  }
  return use_split(m, 0) ? dgk::split_prepare(m) : DGK_OK;
}
