// Fused Burgers stage at P = 5, 6 in split BF16 (three products): the volume
// term, the Lax-Friedrichs flux, LIFT and the RK update in one launch, in
// place of the split stage's two launches (volume term through a derivative
// plan into a state-sized buffer, then flux + LIFT + update), which read u
// three times and the volume term twice. The stage is the one the other
// stage kernels compute (strong-form LF + LIFT after Hesthaven & Warburton
// 2008 §6, against the reference's volume-only RHS in
// examples/3d_scalar_burgers_equation.go:295-340 and its LF + LIFT sketch in
// documentation/DGKernel/example_application.md:296-353).
//
// Per 16-element tile the stage is four products sharing one B,
//   rhs = ga (-Dr W) + gb (-Ds W) + gc (-Dt W) + LIFT G + corr,
// W = [w] (Np x 16) and G = [g] (4 Nfp x 16) one column per element, with
// ga = rx+ry+rz, gb = sx+sy+sz, gc = tx+ty+tz the folded metrics of the
// affine element, w = f - Lf for f = u^2/2 minus its linear interpolant
// through the vertex nodes (rounding then scales with |f - Lf|; D Lf is the
// constant (a1, a2, a3)/2 of the interpolant's slopes and enters as the
// per-element correction corr = -(ga a1 + gb a2 + gc a3)/2), and g the LF
// jumps
//   g = (u- - u+) (A (u- + u+) - B max(|u-|, |u+|)),
// A = Fscale n.(1,1,1)/4, B = Fscale |n.(1,1,1)|/2 -- the same strong-form
// flux as the fused P <= 4 stage (burgers.hip) and the split stage
// (split.hip). Then out = a0 u0 + a1 (u + dt rhs). The metrics scale the
// products' FP32 results in FP64, so B carries w once.
//
// Block: MT matrix waves (one per 16-row tile of [-Dr | -Ds | -Dt | LIFT],
// its BF16 hi/lo fragments held in registers for the launch) and MW memory
// waves. The block walks its tiles in phases separated by block barriers;
// in phase n
//   matrix waves: B of tile n from LDS -> v_mfma_f32_16x16x32_bf16 (A_lo B_hi
//                 + A_hi B_lo + A_hi B_hi, computed as its transpose so that
//                 a lane holds consecutive rows of one element) -> RK update
//                 of their rows (u, u0, the element's metrics and the tile
//                 record from LDS: no global loads);
//   memory waves: the face records of tile n+3 (loaded in phase n-1) into
//                 LDS; the loads of tile n+4's records and of tile n+2 (u,
//                 u0 and the neighbour traces, gathered through the records
//                 stored in phase n-1); tile n+1's u and u0 (loaded in phase
//                 n-1) into LDS, w and the jumps (own traces from that LDS
//                 tile) as BF16 hi/lo into B of tile n+1.
// Memory wave q owns elements 4q..4q+3 of every tile, so it only reads LDS it
// wrote itself between block barriers.
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dgmesh_internal.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

template <int N>
struct FCfg {
  static constexpr int NP = dgk::Tet<N>::Np, NFP = dgk::Tet<N>::Nfp;
  static constexpr int MT = (NP + 15) / 16;   // row tiles of A = matrix waves
  static constexpr int KV = (NP + 31) / 32;   // k = 32 steps per volume block
  static constexpr int KL = (NFP + 31) / 32;  // k = 32 steps per face
  static constexpr int KS = 3 * KV + 4 * KL;  // k steps of A
  static constexpr int KB = KV + 4 * KL;      // k steps of B (w once, then the jumps)
  static constexpr int MW = 4;                // memory waves
  static constexpr int EPW = 16 / MW;         // elements per memory wave
  static constexpr int NT = 64 * (MT + MW);
  static constexpr int NPU = NP | 1;          // LDS row stride of u, doubles
  static constexpr int UPL = (EPW * NP + 63) / 64;        // u values per lane
  static constexpr int GPL = (EPW * 4 * NFP + 63) / 64;  // trace pairs per lane
  static constexpr size_t NFRAG = (size_t)MT * KS * 64;   // fragments per part
  // vertex nodes of the reference tet in the nodal order (checked on the host)
  static constexpr int VTX[4] = {0, N, (N + 1) * (N + 2) / 2 - 1, NP - 1};
  __device__ static int vtx(int v) {
    return v == 0 ? VTX[0] : v == 1 ? VTX[1] : v == 2 ? VTX[2] : VTX[3];
  }
};

// Face record of (face f, element j) of a memory wave's elements: value
// offset of the neighbour's node 0, row of the orientation table, flux
// factors.
struct FRec {
  int32_t nb, fnrow;
  float fa, fb;
};

struct FusedArgs {
  const dgk::Tile* tiles;
  int ntiles;
  const bf16x8* frag;         // [part][MT][KS][64]
  const uint8_t* face_nodes;  // [4][24][Nfp]
  const double* bgeo;         // [19][K]; rows 0..2 the folded metrics
  int64_t K;
  const int32_t* nbr;         // [4][K]
  const int32_t* codes;       // [K]
  const float* lam;           // [6][Np] barycentric weights of the nodes, FP32 and remainders
  const double* u;
  const double* u0;
  double* out;
  double a0, a1;
  const double* dt;
  unsigned long long* maxabs;
  int mode;                   // 0: RK stage, 1: out = rhs
};

__device__ __forceinline__ uint16_t bf16_of(float x) {
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}

// (The tile list is a separate const __restrict__ argument: the compiler may
// then read its wave-uniform entries through the scalar cache, whose wait
// counter does not hold back on the vector loads in flight.)
template <int N, int NPART>
__global__ void __launch_bounds__(FCfg<N>::NT)
    fused_bf16_kernel(FusedArgs a, const dgk::Tile* __restrict__ tiles) {
  static_assert(NPART == 2 || NPART == 3, "BF16x3 (two parts) or BF16x6 (three parts)");
  using C = FCfg<N>;
  // FP64 interpolant for BF16x6, whose products carry less error than its
  // FP32 rounding would add
  using LT = typename std::conditional<NPART == 3, double, float>::type;
  constexpr int NP = C::NP, NFP = C::NFP, MT = C::MT, KV = C::KV, KL = C::KL, KS = C::KS;
  constexpr int KB = C::KB, MW = C::MW, EPW = C::EPW, NPU = C::NPU, UPL = C::UPL, GPL = C::GPL;
  constexpr dgk::FmaskTable<N> FM{};
  // B of two tiles: [buffer][part][element][k] BF16 parts, k = 32 s + the
  // step's column: w at the nodes (32 KV columns), then the jumps face by
  // face (32 KL columns each). Rows are padded by 8 columns, so that the 16
  // elements of a 16-byte read fall in distinct banks; the padding also takes
  // the writes of lanes past the tile's values.
  constexpr int KBR = 32 * KB + 8, DUMP = 32 * KB;
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][NPART][16][KBR];
  // u of two tiles (+ a dump slot), and their tile records, staged by the
  // memory waves for the matrix waves' update
  __shared__ double sU[2][16 * NPU + 1];
  __shared__ int64_t sTv[2];
  __shared__ int sTc[2];
  // per element of two tiles: ga, gb, gc and the linear part's correction
  __shared__ __attribute__((aligned(16))) double sE[2][16][4];
  // per memory wave: face records and folded metrics of its elements for
  // four tiles (slot = tile % 4)
  __shared__ FRec sRec[MW][4][4 * EPW];
  __shared__ double sGm[MW][4][EPW][4];
  __shared__ __attribute__((aligned(16))) float sLam[NP][4];
  // BF16x6: the FP32 remainders of the weights
  __shared__ __attribute__((aligned(16))) float sLamLo[NPART == 3 ? NP : 1][4];
  // per memory wave and element: f(v0) and the slopes f(v1..3) - f(v0) of
  // the linear interpolant of f = u^2/2
  __shared__ __attribute__((aligned(16))) LT sV[MW][EPW][4];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFM[4 * NFP];
  // LIFT fragments of the matrix waves (their Dr, Ds, Dt fragments are held
  // in registers): [row tile][face step][part][lane]
  __shared__ bf16x8 sFL[MT][4 * KL][NPART][64];
  {
    uint4* z = (uint4*)&sB[0][0][0][0];
    for (int i = threadIdx.x; i < (int)(sizeof(sB) / 16); i += C::NT) z[i] = uint4{0, 0, 0, 0};
    for (int i = threadIdx.x; i < 4 * 24 * NFP; i += C::NT) sFN[i] = a.face_nodes[i];
    for (int i = threadIdx.x; i < 4 * NFP; i += C::NT) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
    for (int i = threadIdx.x; i < 4 * NP; i += C::NT) {
      sLam[i / 4][i % 4] = i % 4 < 3 ? a.lam[(i % 4) * NP + i / 4] : 0.f;
      if (NPART == 3) sLamLo[i / 4][i % 4] = i % 4 < 3 ? a.lam[(3 + i % 4) * NP + i / 4] : 0.f;
    }
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
  auto tile_of = [&](int n) -> dgk::Tile {
    const dgk::Tile tr = tiles[t0 + n * ts];
    dgk::Tile u;
    u.vbase = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(tr.vbase >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)tr.vbase));
    u.gbase = __builtin_amdgcn_readfirstlane(tr.gbase);
    u.count = __builtin_amdgcn_readfirstlane(tr.count);
    return u;
  };
  // Lane-derived values are rederived in each step (an empty asm hides the
  // lane index) rather than held across the tile loop, where they spilled.
  auto lane_now = [&]() {
    int l = lane;
    asm volatile("" : "+v"(l));
    return l;
  };
  auto wsync = []() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // the value of lane 4 (l / 4) in lane l (a quad broadcast), and of lane s
  // in every lane, for FP32 and FP64 values
  auto quad_first = [](LT x) -> LT {
    if constexpr (sizeof(LT) == 8) {
      const uint64_t b = __builtin_bit_cast(uint64_t, x);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, 0x00, 0xf, 0xf, false);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), 0x00, 0xf, 0xf, false);
      return __builtin_bit_cast(LT, ((uint64_t)hi << 32) | lo);
    } else {
      return __builtin_bit_cast(LT, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x00, 0xf, 0xf, false));
    }
  };

  // The BF16 parts of an FP32 value into B (part q 16 q rows further): two
  // parts hold ~16 bits, three the whole FP32 value.
  auto put_parts = [](uint16_t* b, float v) {
    const __bf16 h = (__bf16)v;
    const float r = v - (float)h;
    b[0] = __builtin_bit_cast(uint16_t, h);
    if constexpr (NPART == 3) {
      const __bf16 m = (__bf16)r;
      b[16 * KBR] = __builtin_bit_cast(uint16_t, m);
      b[32 * KBR] = bf16_of(r - (float)m);
    } else {
      b[16 * KBR] = bf16_of(r);
    }
  };
  // Face records of memory wave q's elements for tile n go to slot n % 4 of
  // sRec[q] (their folded metrics to sGm[q]), written by that wave three
  // phases ahead from loads it issued a phase before: lane s < 4 EPW takes
  // face s / EPW of element EPW q + s % EPW (clamped to the tile).
  int32_t rnb = 0, rcode = 0;
  double rga = 0.0, rgb = 0.0, rgc = 0.0;
  // (Every lane loads, lane l taking the record of lane l % (4 EPW), and
  // tiles past the block's last read its last: the loads are the same in
  // every phase, so the waits for them count exactly.)
  static_assert((4 * EPW & (4 * EPW - 1)) == 0, "records per wave: a power of two");
  auto rec_load = [&](int qq, int n) {
    const int ln = lane_now() & (4 * EPW - 1), rf = ln / EPW, rj = ln % EPW;
    const dgk::Tile tl = tile_of(n < nb ? n : nb - 1);
    const int e = EPW * qq + rj < tl.count ? EPW * qq + rj : tl.count - 1;
    const int64_t ge = tl.gbase + e;
    rnb = a.nbr[(int64_t)rf * K + ge];
    rcode = a.codes[ge];
    rga = a.bgeo[ge];
    rgb = a.bgeo[K + ge];
    rgc = a.bgeo[2 * K + ge];
  };
  auto rec_store = [&](int qq, int n) {
    if (n >= nb) return;
    const int ln = lane_now(), rf = ln / EPW, rj = ln % EPW;
    if (ln < 4 * EPW) {
      FRec r;
      r.nb = rnb;
      r.fnrow = (rf * 24 + ((rcode >> (5 * rf)) & 31)) * NFP;
      // Fscale n.(1,1,1) of face rf: -gc, -gb, ga + gb + gc, -ga (the face
      // factors of dgk::face_factors), as coefficients in {-1, 0, 1} of
      // (ga, gb, gc) rather than a branch on the lane's face
      const double c0 = (double)(rf == 2) - (double)(rf == 3);
      const double c1 = (double)(rf == 2) - (double)(rf == 1);
      const double c2 = (double)(rf == 2) - (double)(rf == 0);
      const double sn = fma(c0, rga, fma(c1, rgb, c2 * rgc));
      r.fa = (float)(0.25 * sn);
      r.fb = (float)(0.5 * fabs(sn));
      sRec[qq][n & 3][ln] = r;
      if (rf == 0) {
        sGm[qq][n & 3][rj][0] = rga;
        sGm[qq][n & 3][rj][1] = rgb;
        sGm[qq][n & 3][rj][2] = rgc;
      }
    }
  };
  __syncthreads();

  if (w < MT) {
    // ---------------- matrix wave: rows 16 w .. 16 w + 15 ----------------
    bf16x8 fa[NPART][3 * KV];
#pragma unroll
    for (int q = 0; q < NPART; ++q) {
#pragma unroll
      for (int s = 0; s < 3 * KV; ++s) fa[q][s] = a.frag[q * C::NFRAG + ((size_t)w * KS + s) * 64 + lane];
#pragma unroll
      for (int t = 0; t < 4 * KL; ++t)
        sFL[w][t][q][lane] = a.frag[q * C::NFRAG + ((size_t)w * KS + 3 * KV + t) * 64 + lane];
    }
    const double dt = a.mode == 0 ? *a.dt : 1.0;
    const double a0 = a.mode == 0 ? a.a0 : 0.0, a1 = a.mode == 0 ? a.a1 : 1.0;
    // The fragments and dt are complete before the loop: without this wait
    // the compiler counts them as pending at the loop head and holds the
    // products behind the stores of the previous tile.
    __builtin_amdgcn_s_waitcnt(0);
    double lmax = 0.0;
    __syncthreads();  // phase -1: the memory waves prepare tile 0
    for (int n = 0; n < nb; ++n) {
      const int bu = n & 1;
      const int64_t tv = sTv[bu];
      const int64_t vbase = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(tv >> 32)) << 32) |
                                      (uint32_t)__builtin_amdgcn_readfirstlane((int)tv));
      const int count = __builtin_amdgcn_readfirstlane(sTc[bu]);
      // u0 of the rows this lane updates (below), loaded during the products
      // (BF16x6: during the LIFT products only, so that it is not held across
      // the volume products, where its registers would spill); loads past
      // the tile read zero
      const int row = 16 * w + c;
      const int rr = row < NP ? row : NP - 1;
      double u0v[4] = {0.0, 0.0, 0.0, 0.0};
      auto load_u0 = [&]() {
        if (a0 != 0.0) {
          const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
              (void*)(a.u0 + vbase), 0, count * NP * 8, 0x00020000);
#pragma unroll
          for (int r = 0; r < 4; ++r)
            u0v[r] = __builtin_bit_cast(
                double, __builtin_amdgcn_raw_buffer_load_b64(rs0, ((4 * g + r) * NP + rr) * 8, 0, 2));
        }
      };
      if (NPART == 2) load_u0();
      // The products run transposed, B^T A^T: the tile's B as the first
      // operand (its registers are B's as read) and the fragments as the
      // second, so that lane l, register r holds row 16 w + l % 16 of element
      // 4 (l / 16) + r and each store below writes 16 consecutive rows of an
      // element. -Dr w, -Ds w, -Dt w and LIFT g accumulate apart; the
      // element's metrics scale them in the update.
      // The split products, smallest first: BF16x3 b0 a1 + b1 a0 + b0 a0,
      // BF16x6 b0 a2 + b2 a0 + b1 a1 + b0 a1 + b1 a0 + b0 a0 (part 0 the
      // leading one).
      auto prod = [](floatx4 acc, const bf16x8* b, const bf16x8* op) {
        if constexpr (NPART == 3) {
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[0], op[2], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[2], op[0], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[1], op[1], acc, 0, 0, 0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[0], op[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[1], op[0], acc, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[0], op[0], acc, 0, 0, 0);
      };
      floatx4 aR = {0.f, 0.f, 0.f, 0.f}, aS = aR, aT = aR, aL = aR;
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int s = 0; s < KV; ++s) {
        bf16x8 b[NPART], oR[NPART], oS[NPART], oT[NPART];
#pragma unroll
        for (int q = 0; q < NPART; ++q) {
          b[q] = *(const bf16x8*)&sB[bu][q][c][32 * s + 8 * g];
          oR[q] = fa[q][s];
          oS[q] = fa[q][KV + s];
          oT[q] = fa[q][2 * KV + s];
        }
        aR = prod(aR, b, oR);
        aS = prod(aS, b, oS);
        aT = prod(aT, b, oT);
      }
      if (NPART == 3) load_u0();
#pragma unroll
      for (int t = 0; t < 4 * KL; ++t) {
        bf16x8 b[NPART], oL[NPART];
#pragma unroll
        for (int q = 0; q < NPART; ++q) {
          b[q] = *(const bf16x8*)&sB[bu][q][c][32 * (KV + t) + 8 * g];
          oL[q] = sFL[w][t][q][lane];
        }
        aL = prod(aL, b, oL);
      }
      __builtin_amdgcn_s_setprio(0);
      // Stores through a buffer resource sized to the tile: the ones past it
      // (elements >= count, rows >= Np) are dropped, so the partial tile takes
      // no branches.
      const __amdgpu_buffer_rsrc_t rsO = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(a.out + vbase), 0, count * NP * 8, 0x00020000);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = 4 * g + r;
        const double* E = sE[bu][e];
        const double rhs = fma(E[0], (double)aR[r],
                               fma(E[1], (double)aS[r], fma(E[2], (double)aT[r], (double)aL[r] + E[3])));
        const double uv = a.mode == 0 ? sU[bu][e * NPU + rr] : 0.0;
        const double o = fma(a0, u0v[r], a1 * fma(dt, rhs, uv));
        const bool ok = row < NP && e < count;
        lmax = fmax(lmax, ok ? fabs(o) : 0.0);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, o), rsO,
                                              row < NP ? (e * NP + row) * 8 : 0x40000000, 0, 2);
      }
      __syncthreads();  // end of phase n
    }
    if (a.maxabs) {
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o, 64));
      if (lane == 0 && lmax > 0.0)
        atomicMax(a.maxabs, (unsigned long long)__double_as_longlong(lmax));
    }
    return;
  }

  // ---------------- memory wave q: elements EPW q .. EPW q + EPW - 1 ----------------
  const int q = w - MT;
  const int e0 = EPW * q;
  // Per lane and trace pair k (pair p = lane + 64 k: slot p / NFP = f EPW + j,
  // point p % NFP), fixed for the launch: index of the own trace in a tile's
  // u (12 bits), column of the jump in B (9 bits; the padding past the pairs),
  // record slot (5 bits), element (2 bits).
  uint32_t fw[GPL];
#pragma unroll
  for (int k = 0; k < GPL; ++k) {
    const int p = lane + 64 * k;
    const bool ok = EPW * 4 * NFP % 64 == 0 || p < EPW * 4 * NFP;
    const int pp = ok ? p : EPW * 4 * NFP - 1;
    const int sl = pp / NFP, qq = pp % NFP, f = sl / EPW, j = sl % EPW;
    const int iu = (e0 + j) * NPU + sFM[f * NFP + qq];
    const int col = ok ? 32 * (KV + f * KL) + qq : DUMP + (lane & 7);
    fw[k] = (uint32_t)iu | ((uint32_t)col << 12) | ((uint32_t)sl << 21) | ((uint32_t)j << 26);
  }
  // Per lane and value k (p = lane + 64 k: element j = p / NP, node
  // i = p % NP), fixed for the launch: the barycentric weights of node i, and
  // a word with the value's column in B (12 bits; the padding past the
  // values), j (bits 12..13) and its index in a tile's u (bits 16..27; the
  // dump slot past the values).
  // (BF16x6 reads the weights with their remainders from LDS instead: its
  // memory waves have no registers to spare.)
  constexpr int UPLW = NPART == 3 ? 1 : UPL;
  float LX[UPLW], LY[UPLW], LZ[UPLW];
  uint32_t vw[UPL];
#pragma unroll
  for (int k = 0; k < UPL; ++k) {
    const int p = lane + 64 * k;
    const bool ok = EPW * NP % 64 == 0 || p < EPW * NP;
    const int pc = ok ? p : EPW * NP - 1;
    const int j = pc / NP, i = pc % NP;
    if (NPART == 2) {
      LX[k % UPLW] = sLam[i][0];
      LY[k % UPLW] = sLam[i][1];
      LZ[k % UPLW] = sLam[i][2];
    }
    const int bo = (e0 + j) * KBR + (ok ? i : DUMP + (lane & 7));
    const int so = ok ? (e0 + j) * NPU + i : 16 * NPU;
    vw[k] = (uint32_t)bo | ((uint32_t)j << 12) | ((uint32_t)so << 16);
  }
  // Lane 4 j + v (< 16) reads u at vertex v of element j from a tile's u.
  const int vxo = (e0 + ((lane >> 2) & 3)) * NPU + C::vtx(lane & 3);
  // u values (pair p = lane + 64 k: element e0 + p / NP, node p % NP) and
  // neighbour traces (pair p as above), by slot (compile-time constants, so
  // the arrays stay in registers)
  double U[2][UPL], P[2][GPL];
  int64_t TV[2] = {0, 0};  // tile records by slot
  int TC[2] = {1, 1};
  // Loads of tile n (clamped to the block's last tile, so that every phase
  // issues the same loads) into a slot.
  auto issue = [&](int n, auto SL) {
    constexpr int slot = decltype(SL)::value;
    const int nn = n < nb ? n : nb - 1;
    const dgk::Tile tl = tile_of(nn);
    const FRec* R = sRec[q][nn & 3];
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
      const int pp = p < EPW * 4 * NFP ? p : EPW * 4 * NFP - 1;
      const FRec& r = R[pp / NFP];
      P[slot][k] = a.u[r.nb + sFN[r.fnrow + pp % NFP]];
    }
    TV[slot] = tl.vbase;
    TC[slot] = tl.count;
  };
  // Tile n's values (loaded into slot) -> LDS u, the per-element metrics and
  // correction, and B.
  auto prepare = [&](int n, auto SL) {
    constexpr int slot = decltype(SL)::value;
    const int lane = lane_now();
    const int bu = n & 1;
    const FRec* R = sRec[q][n & 3];
    double* su = sU[bu];
#pragma unroll
    for (int k = 0; k < UPL; ++k) su[vw[k] >> 16] = U[slot][k];
    if (q == 0 && lane == 0) {
      sTv[bu] = TV[slot];
      sTc[bu] = TC[slot];
    }
    wsync();  // the wave's u of the tile is in LDS
    // The jumps' own traces and the vertex values are read here; the jumps
    // use theirs after the volume columns, whose arithmetic covers the
    // latency.
    double MV[GPL];
#pragma unroll
    for (int k = 0; k < GPL; ++k) MV[k] = su[fw[k] & 4095];
    const double xv = su[vxo];
    const double* G = sGm[q][n & 3][lane & 3];
    const double ga = G[0], gb = G[1], gc = G[2];
    // f = u^2/2 at vertex v of element j on lane 4 j + v and its difference
    // from f at vertex 0 (a quad broadcast) into the wave's table: f(v0) and
    // the slopes f(v1..3) - f(v0) of the linear interpolant
    const LT fv = (LT)(0.5 * xv * xv);
    const LT dv = fv - quad_first(fv);
    if (lane < 4 * EPW) sV[q][lane >> 2][lane & 3] = (lane & 3) == 0 ? fv : dv;
    wsync();
    // per element (lane j < EPW): the metrics and the correction
    // -(ga a1 + gb a2 + gc a3)/2 for the interpolant's slopes
    if (lane < EPW) {
      const LT* V = sV[q][lane];
      double* E = sE[bu][e0 + lane];
      E[0] = ga;
      E[1] = gb;
      E[2] = gc;
      E[3] = -0.5 * (ga * (double)V[1] + gb * (double)V[2] + gc * (double)V[3]);
    }
    // volume columns: w = f - Lf at node i of element j (BF16x3: the
    // interpolant in FP32, its rounding entering w as the FP32 weights' does;
    // BF16x6: in FP64 with the weights' FP32 remainders)
    uint16_t* bB = &sB[bu][0][0][0];
#pragma unroll
    for (int k = 0; k < UPL; ++k) {
      const uint32_t vk = vw[k];
      const int j = (vk >> 12) & 3;
      const LT* V = sV[q][j];
      const LT f0 = V[0], d1 = V[1], d2 = V[2], d3 = V[3];
      LT lf;
      if constexpr (NPART == 3) {
        const int i = (int)(vk & 4095) - (e0 + j) * KBR;  // past NP: the padding
        const float4 hi = *(const float4*)sLam[i < NP ? i : 0];
        const float4 lo = *(const float4*)sLamLo[i < NP ? i : 0];
        lf = fma(d3, (double)hi.z + (double)lo.z,
                 fma(d2, (double)hi.y + (double)lo.y, fma(d1, (double)hi.x + (double)lo.x, f0)));
      } else {
        lf = fmaf(d3, LZ[k % UPLW], fmaf(d2, LY[k % UPLW], fmaf(d1, LX[k % UPLW], f0)));
      }
      const double x = U[slot][k];
      put_parts(bB + (vk & 4095), (float)(0.5 * x * x - (double)lf));
    }
    // jump columns: the jump at point qq of face f of element j
#pragma unroll
    for (int k = 0; k < GPL; ++k) {
      const uint32_t fk = fw[k];
      const int col = (fk >> 12) & 511, j = fk >> 26;
      const FRec& r = R[(fk >> 21) & 31];
      const double m = MV[k], pv = P[slot][k];
      const double x =
          (m - pv) * fma((double)r.fa, m + pv, -(double)r.fb * fmax(fabs(m), fabs(pv)));
      put_parts(&sB[bu][0][e0 + j][col], (float)x);
    }
  };

  // Prologue (phase -1 of the matrix waves): records of tiles 0..2 stored
  // and tile 3's loaded, loads of tiles 0 and 1, tile 0 prepared.
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  if (nb > 0) {
#pragma unroll 1
    for (int n = 0; n < 3; ++n) {
      rec_load(q, n);
      rec_store(q, n);
    }
    rec_load(q, 3);
    wsync();
    issue(0, S0{});
    issue(1, S1{});
    prepare(0, S0{});
  }
  __syncthreads();  // end of phase -1
  // Phase n (the matrix waves on tile n): the records of tile n+3 (loaded a
  // phase ago) stored, the loads of tile n+4's records and of tile n+2 (its
  // records are in LDS since phase n-1) into slot n & 1, then tile n+1
  // prepared from slot (n+1) & 1, loaded a phase ago: each load is consumed
  // a phase after it is issued, and the ones it waits for are the oldest in
  // flight.
  auto phase = [&](int n, auto A1, auto A2) {
    rec_store(q, n + 3);
    rec_load(q, n + 4);
    wsync();
    issue(n + 2, A2);
    if (n + 1 < nb) prepare(n + 1, A1);
    __syncthreads();  // end of phase n
  };
  for (int n = 0; n < nb; n += 2) {
    phase(n, S1{}, S0{});
    if (n + 1 < nb) phase(n + 1, S0{}, S1{});
  }
}

// bf16 round-to-nearest-even of a float, as raw bits (host side).
uint16_t bf16_bits_h(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  if ((x & 0x7f800000u) == 0x7f800000u) return (uint16_t)(x >> 16);
  uint32_t lsb = (x >> 16) & 1u;
  x += 0x7fffu + lsb;
  return (uint16_t)(x >> 16);
}

float bf16_float_h(uint16_t b) {
  uint32_t x = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

template <int N, int NPART>
int fused_launch(dgk_dgmesh m, const FusedArgs& a) {
  if (a.ntiles == 0) return DGK_OK;
  int grid = m->dev->cu_count;
  if (grid > a.ntiles) grid = a.ntiles;
  hipLaunchKernelGGL((fused_bf16_kernel<N, NPART>), dim3(grid), dim3(FCfg<N>::NT), 0, m->dev->stream,
                     a, a.tiles);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

}  // namespace

bool dgk::fused_applies(dgk_dgmesh m) {
  return (m->N == 5 || m->N == 6) && (m->prec == DGK_PREC_BF16X3 || m->prec == DGK_PREC_BF16X6) &&
         m->lift_prec == m->prec && m->variant != 3;
}

int dgk::fused_prepare(dgk_dgmesh m) {
  const int npart = m->prec == DGK_PREC_BF16X6 ? 3 : 2;
  if (m->fused_frag && m->fused_npart == npart) return DGK_OK;
  fused_release(m);
  const int N = m->N, NP = dgk::np_of(N), NFP = dgk::nfp_of(N), NF = 4 * NFP;
  const int MT = (NP + 15) / 16, KV = (NP + 31) / 32, KL = (NFP + 31) / 32, KS = 3 * KV + 4 * KL;
  const size_t nfrag = (size_t)MT * KS * 64;
  // A = [-Dr | -Ds | -Dt | LIFT] in the 16x16x32 BF16 A-operand lane order:
  // lane l of step s of row tile t holds A[16 t + l % 16][32 s' + 8 (l / 16) + e],
  // e = 0..7, within the block (volume block s / KV, or face (s - 3 KV) / KL);
  // part q is the BF16 value of what parts 0..q-1 leave of the FP64 entry.
  std::vector<uint16_t> buf((size_t)npart * nfrag * 8, 0);
  const std::vector<double>* D[3] = {&m->Dr_h, &m->Ds_h, &m->Dt_h};
  for (int t = 0; t < MT; ++t)
    for (int s = 0; s < KS; ++s)
      for (int l = 0; l < 64; ++l)
        for (int e = 0; e < 8; ++e) {
          const int row = 16 * t + (l & 15);
          double v = 0.0;
          if (row < NP) {
            if (s < 3 * KV) {
              const int k = 32 * (s % KV) + 8 * (l >> 4) + e;
              if (k < NP) v = -(*D[s / KV])[(size_t)row * NP + k];
            } else {
              const int f = (s - 3 * KV) / KL;
              const int qq = 32 * ((s - 3 * KV) % KL) + 8 * (l >> 4) + e;
              if (qq < NFP) v = m->LIFT_h[(size_t)row * NF + f * NFP + qq];
            }
          }
          const size_t q = (((size_t)t * KS + s) * 64 + l) * 8 + e;
          double rest = v;
          for (int part = 0; part < npart; ++part) {
            const uint16_t h = bf16_bits_h((float)rest);
            buf[part * nfrag * 8 + q] = h;
            rest -= (double)bf16_float_h(h);
          }
        }
  // the kernel takes the vertices of the reference tet at nodes 0, N,
  // (N+1)(N+2)/2 - 1 and Np - 1; barycentric weights of the nodes
  const double V[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
  const int vtx[4] = {0, N, (N + 1) * (N + 2) / 2 - 1, NP - 1};
  for (int v = 0; v < 4; ++v) {
    const int i = vtx[v];
    DGK_REQUIRE(std::fabs(m->r_h[i] - V[v][0]) < 1e-10 && std::fabs(m->s_h[i] - V[v][1]) < 1e-10 &&
                    std::fabs(m->t_h[i] - V[v][2]) < 1e-10,
                "fused stage: node %d is not vertex %d of the reference tet", i, v);
  }
  // [0..2][Np] the FP32 weights, [3..5][Np] their FP32 remainders
  std::vector<float> lam(6 * NP);
  for (int i = 0; i < NP; ++i) {
    const double l[3] = {0.5 * (1.0 + m->r_h[i]), 0.5 * (1.0 + m->s_h[i]), 0.5 * (1.0 + m->t_h[i])};
    for (int d = 0; d < 3; ++d) {
      lam[d * NP + i] = (float)l[d];
      lam[(3 + d) * NP + i] = (float)(l[d] - (double)(float)l[d]);
    }
  }
  int rc = dgk_malloc(m->dev, buf.size() * 2, buf.data(), &m->fused_frag);
  if (rc == DGK_OK) rc = dgk_malloc(m->dev, lam.size() * 4, lam.data(), (void**)&m->fused_lam);
  if (rc == DGK_OK) m->fused_npart = npart;
  return rc;
}

void dgk::fused_release(dgk_dgmesh m) {
  if (m->fused_frag) dgk_free(m->dev, m->fused_frag);
  if (m->fused_lam) dgk_free(m->dev, m->fused_lam);
  m->fused_frag = nullptr;
  m->fused_lam = nullptr;
  m->fused_npart = 0;
}

int dgk::fused_stage(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                     double a1, const double* dt, unsigned long long* maxabs, int mode) {
  int rc = fused_prepare(m);
  if (rc != DGK_OK) return rc;
  FusedArgs a;
  a.tiles = m->tiles64;
  a.ntiles = m->ntiles64;
  a.frag = (const bf16x8*)m->fused_frag;
  a.face_nodes = m->face_nodes;
  a.bgeo = m->bgeo;
  a.K = m->K;
  a.nbr = m->nbr;
  a.codes = m->codes;
  a.lam = m->fused_lam;
  a.u = u;
  a.u0 = u0;
  a.out = out;
  a.a0 = a0;
  a.a1 = a1;
  a.dt = dt;
  a.maxabs = maxabs;
  a.mode = mode;
  const bool six = m->fused_npart == 3;
  switch (m->N) {
    case 5: return six ? fused_launch<5, 3>(m, a) : fused_launch<5, 2>(m, a);
    case 6: return six ? fused_launch<6, 3>(m, a) : fused_launch<6, 2>(m, a);
  }
  dgk::set_error("fused stage: N=%d outside 5..6", m->N);
  return DGK_ERR_UNSUPPORTED;
}
