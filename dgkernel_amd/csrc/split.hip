// Split Burgers stage for orders above 4 (where the fused stage's operators
// no longer fit a CU) and for contractions below FP64:
//   1. volume term  rhs = -div f(u)           deriv_affine_kernel<.., DIV>
//      ([Dr; Ds; Dt] f on matrix cores, metrics folded in FP64, or in FP32
//      for the FP32 and BF16x3 plans, whose rhs is then stored as FP32);
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
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

constexpr int kF64 = 0, kF32 = 1, kBF3 = 2, kBF6 = 3;
constexpr bool is_bf(int prec) { return prec == kBF3 || prec == kBF6; }
constexpr int parts_of(int prec) { return prec == kBF3 ? 2 : (prec == kBF6 ? 3 : 1); }

template <int N, int PREC>
struct SCfg {
  static constexpr int NP = dgk::Tet<N>::Np, NFP = dgk::Tet<N>::Nfp, NF = 4 * NFP;
  static constexpr int MTL = (NP + 15) / 16;  // 16-row tiles of LIFT
  // LIFT columns face by face: the MFMA steps of face f take its points four
  // at a time (FP64/FP32, k = 4) or 32 at a time (bf16, k = 32), with zero
  // columns past Nfp.
  static constexpr int KB = is_bf(PREC) ? 32 : 4;
  static constexpr int KSF = (NFP + KB - 1) / KB;  // steps per face
  static constexpr int NPT = KB * KSF;             // padded points per face
  static constexpr int PARTS = parts_of(PREC);
  static constexpr size_t NFRAG = (size_t)MTL * 4 * KSF * 64;  // fragments per part
  static constexpr size_t FRAG_BYTES =
      PREC == kF64 ? NFRAG * 8 : (PREC == kF32 ? NFRAG * 4 : NFRAG * 16 * PARTS);
  // Waves per block: two per SIMD, in one block per CU, or for BF16x3 in
  // two blocks of four (its fragments leave room in LDS for two). The
  // kernel keeps its own loads in flight two phases ahead; a third wave per
  // SIMD would not fit the registers that takes (it spilled, and ran 20-30 %
  // slower).
  static constexpr int W = PREC == kBF3 ? 4 : 8;
  // Trace gathers of one face: lanes along its 16 x Nfp (element, point)
  // pairs, point fastest, so one load instruction touches the nodes of two
  // or three neighbours rather than one per lane.
  static constexpr int NG = (16 * NFP + 63) / 64;
  // The jumps of one face in LDS, [point][element] with row stride 17 (the
  // point-major writes and element-major reads are both conflict-free), in
  // FP64 where the product needs more than FP32 carries; the same per-wave
  // buffer serves as the 16 x 17 update buffer.
  using JT = typename std::conditional<PREC == kF64 || PREC == kBF6, double, float>::type;
  static constexpr size_t BUF_BYTES =
      NPT * 17 * sizeof(JT) > 16 * 17 * 8 ? NPT * 17 * sizeof(JT) : 16 * 17 * 8;
  static constexpr int BUF = (int)(BUF_BYTES / 8);
};

// Per-wave face record for the gathers: value offset of the neighbour's
// node 0, row of the orientation table, flux factors.
struct FaceRec {
  int32_t nb, fnrow;
  double fa, fb;
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
  const void* rhs;            // volume term, FP64 or (R32) FP32
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

// Per tile, each wave:
//   1. lanes along each face's (element, point) pairs hold the own and
//      neighbour traces (face 0 was gathered during the previous tile; the
//      others are gathered two faces ahead of their use) and
//      form the LF jumps face by face into the LDS buffer, from which each
//      lane reads its B operands of the face's MFMA steps into registers;
//   2. the next tile's face records (neighbour offset, orientation row, flux
//      factors of face g of element c at lane (c, g)) go to the other record
//      buffer, and its face 0 is gathered;
//   3. per 16-row tile of LIFT: the products over the four faces, then the
//      result goes through the wave's 16 x 17 buffer so that each lane
//      updates and stores rows of 16 consecutive values. The rhs, u, u0
//      values of row tile tt + 1 are loaded before the products of row
//      tile tt (those of row tile 0 during the jumps).
// Wait counts are in order, so every prefetch is consumed after the loads
// issued before it, never before.
template <int N, int PREC, bool R32>
__global__ void __launch_bounds__((64 * SCfg<N, PREC>::W), 2) surface_update_kernel(SurfArgs a) {
  using C = SCfg<N, PREC>;
  using JT = typename C::JT;
  constexpr int NP = C::NP, NFP = C::NFP, MTL = C::MTL, KSF = C::KSF, W = C::W;
  constexpr int NPT = C::NPT, PARTS = C::PARTS, NG = C::NG;
  constexpr dgk::FmaskTable<N> FM{};
  __shared__ __attribute__((aligned(16))) unsigned char sA[C::FRAG_BYTES];
  __shared__ __attribute__((aligned(16))) double sTall[W][C::BUF];
  __shared__ FaceRec sRall[W][2][64];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFM[4 * NFP];
  {
    const uint4* src = (const uint4*)a.frag;
    uint4* dst = (uint4*)sA;
    for (int i = threadIdx.x; i < (int)(C::FRAG_BYTES / 16); i += 64 * W) dst[i] = src[i];
    for (int i = threadIdx.x; i < 4 * 24 * NFP; i += 64 * W) sFN[i] = a.face_nodes[i];
    for (int i = threadIdx.x; i < 4 * NFP; i += 64 * W) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
  }
  __syncthreads();
  // wave index as a scalar, so that tile indices and records stay in SGPRs
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = lane & 15, g = lane >> 4;  // element of the tile, face (or k group)
  double* sT = sTall[w];
  JT* sJ = (JT*)sT;
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
  // Tile records are read one tile ahead and made wave-uniform only when
  // used, so waiting for one never drains the loads issued after it.
  auto uniform = [](const dgk::Tile& v) -> dgk::Tile {
    dgk::Tile u;
    u.vbase = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v.vbase >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)v.vbase));
    u.gbase = __builtin_amdgcn_readfirstlane(v.gbase);
    u.count = __builtin_amdgcn_readfirstlane(v.count);
    return u;
  };
  // Face-record inputs of lane (c, g) for a tile (element clamped to the
  // tile, so every load stays in it).
  int32_t rnb = 0, rcode = 0;
  double rga = 0.0, rgb = 0.0, rgc = 0.0;
  auto rec_load = [&](const dgk::Tile& tl) {
    const int ce = c < tl.count ? c : tl.count - 1;
    const int64_t ge = tl.gbase + ce;
    int gg = g;
    asm volatile("" : "+v"(gg));  // (the row base is derived here, not kept live)
    rnb = a.nbr[(int64_t)gg * K + ge];
    rcode = a.codes[ge];
    rga = a.bgeo[ge];
    rgb = a.bgeo[K + ge];
    rgc = a.bgeo[2 * K + ge];
  };
  auto rec_store = [&](FaceRec* sR) {
    FaceRec r;
    r.nb = rnb;
    r.fnrow = (g * 24 + ((rcode >> (5 * g)) & 31)) * NFP;
    dgk::face_factors(g, rga, rgb, rgc, r.fa, r.fb);
    sR[g * 16 + c] = r;
    wave_sync();
  };
  // Own and neighbour traces of face f of a tile at the pairs of this lane:
  // two faces in flight, face f in slot f % 2.
  double uM[2][NG], uP[2][NG];
  // (The pair indices pass through an empty asm, so that the compiler
  // derives the LDS addresses from them here instead of keeping one address
  // register per face and pair live across the tile loop.)
  auto gather = [&](const dgk::Tile& tl, const FaceRec* sR, int f) {
    const double* ut = a.u + tl.vbase;
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      int jc = pc[j], jq = pq[j];
      asm volatile("" : "+v"(jc), "+v"(jq));
      const int ce = jc < tl.count ? jc : tl.count - 1;
      const FaceRec& r = sR[f * 16 + jc];
      uP[f & 1][j] = a.u[r.nb + sFN[r.fnrow + jq]];
      uM[f & 1][j] = ut[ce * NP + sFM[f * NFP + jq]];
    }
  };
  // mode 1 (rhs + LIFT g) is the update with a0 = 0, a1 = dt = 1 and u = 0
  const double dt = a.mode == 0 ? *a.dt : 1.0;
  const double a0 = a.mode == 0 ? a.a0 : 0.0, a1 = a.mode == 0 ? a.a1 : 1.0;
  dgk::Tile tile_n{}, tile_raw{};
  int rb = 0;  // record buffer of the current tile
  if (t0 < t1) {
    tile_n = uniform(a.tiles[t0]);
    rec_load(tile_n);
    rec_store(sRall[w][0]);
    gather(tile_n, sRall[w][0], 0);
    if (t0 + ts < t1) tile_raw = a.tiles[t0 + ts];
  }
  for (int t = t0; t < t1; t += ts, rb ^= 1) {
    const dgk::Tile tile = tile_n;
    const bool more = t + ts < t1;
    if (more) {
      tile_n = uniform(tile_raw);
      rec_load(tile_n);
      if (t + 2 * ts < t1) tile_raw = a.tiles[t + 2 * ts];
    }
    const int cnt = tile.count;
    const FaceRec* sR = sRall[w][rb];
    gather(tile, sR, 1);
    // The tile's state values through buffer resources: loads past the
    // tile read zero and stores past it are dropped, so partial tiles take
    // no branches.
    const int nbytes = cnt * NP * 8;
    const __amdgpu_buffer_rsrc_t rsR = __builtin_amdgcn_make_buffer_rsrc(
        (void*)((const char*)a.rhs + tile.vbase * (R32 ? 4 : 8)), 0, R32 ? nbytes / 2 : nbytes,
        0x00020000);
    const __amdgpu_buffer_rsrc_t rsU = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.u + tile.vbase), 0, a.mode == 0 ? nbytes : 0, 0x00020000);
    const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a0 != 0.0 ? a.u0 + tile.vbase : a.u), 0, a0 != 0.0 ? nbytes : 0, 0x00020000);
    const __amdgpu_buffer_rsrc_t rsO = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.out + tile.vbase), 0, nbytes, 0x00020000);
    // Byte offset of the value lane updates for row tile tt, slot j (element
    // 4 j + lane / 16, row 16 tt + lane % 16); past the tile for rows >= Np.
    auto upd_off = [&](int tt, int j) -> int {
      const int e = 4 * j + (lane >> 4), i = 16 * tt + (lane & 15);
      return i < NP ? (e * NP + i) * 8 : 0x40000000;
    };
    typename std::conditional<R32, float, double>::type pr[2][4];
    double pu[2][4], p0[2][4];
    auto preload = [&](int tt) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int off = upd_off(tt, j);
        if constexpr (R32)
          pr[tt & 1][j] =
              __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsR, off >> 1, 0, 0));
        else
          pr[tt & 1][j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsR, off, 0, 0));
        pu[tt & 1][j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsU, off, 0, 0));
        p0[tt & 1][j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs0, off, 0, 2));
      }
    };
    // B operands of each face's steps: lane (c, g) takes points 4 s + g
    // (k = 4) or 32 s + 8 g + e (bf16) of element c.
    typename std::conditional<PREC == kF64, double, float>::type b[4][is_bf(PREC) ? 1 : KSF];
    bf16x8 bp[4][is_bf(PREC) ? KSF : 1][PARTS];
    if constexpr (NPT > NFP) {
      // columns past Nfp of the jump buffer are zero (the update buffer
      // overwrote them)
      wave_sync();
      for (int i = lane; i < 16 * (NPT - NFP); i += 64) sJ[(NFP + i / 16) * 17 + (i & 15)] = (JT)0;
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      // Fscale (n.f(u-) - n.f*), f(u) = u^2/2 (1,1,1), factored.
      if (f > 0 || NPT == NFP) wave_sync();  // previous readers of the buffer are done
#pragma unroll
      for (int j = 0; j < NG; ++j) {
        const FaceRec& r = sR[f * 16 + pc[j]];
        const double m = uM[f & 1][j], p = uP[f & 1][j];
        const double x = (m - p) * fma(r.fa, m + p, -r.fb * fmax(fabs(m), fabs(p)));
        if (pv[j]) sJ[pq[j] * 17 + pc[j]] = (JT)x;
      }
      if (f < 2) gather(tile, sR, f + 2);
      else if (f == 2) preload(0);
      wave_sync();
      if constexpr (!is_bf(PREC)) {
#pragma unroll
        for (int s = 0; s < KSF; ++s) b[f][s] = sJ[(4 * s + g) * 17 + c];
      } else {
#pragma unroll
        for (int s = 0; s < KSF; ++s)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const JT x = sJ[(32 * s + 8 * g + e) * 17 + c];
            if constexpr (PARTS == 2) {
              // two parts carry ~16 bits: the FP32 value is exact enough
              const __bf16 h = (__bf16)x;
              bp[f][s][0][e] = h;
              bp[f][s][1][e] = (__bf16)(x - (float)h);
            } else {
              const __bf16 h = (__bf16)(float)x;
              const double r1 = x - (double)(float)h;
              const __bf16 m = (__bf16)(float)r1;
              bp[f][s][0][e] = h;
              bp[f][s][1][e] = m;
              bp[f][s][2][e] = (__bf16)(float)(r1 - (double)(float)m);
            }
          }
      }
    }
    // The next tile: face records, then its first face in flight.
    if (more) {
      FaceRec* sRn = sRall[w][rb ^ 1];
      rec_store(sRn);
      gather(tile_n, sRn, 0);
    }
#pragma unroll
    for (int tt = 0; tt < MTL; ++tt) {
      if (tt + 1 < MTL) preload(tt + 1);
      AccT<PREC> acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = 0;
      __builtin_amdgcn_s_setprio(1);  // matrix phase first (wave priority)
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int s = 0; s < KSF; ++s) {
          const size_t q = (((size_t)tt * 4 + f) * KSF + s) * 64 + lane;
          if constexpr (PREC == kF64) {
            if (NP - 16 * tt <= 4) {
              // a last row tile of four rows or fewer: one v_mfma_f64_4x4x4_4b
              // (about eight times cheaper to issue on gfx950). Its block j is
              // lanes 16 k + 4 j + i (A[i][k], B[k][i]; D[i][n] at lane
              // 16 i + 4 j + n): B unchanged, A row l%4 from lane
              // 16 (l/16) + l%4 of the table, result in register 0's place
              int lp = lane;
              asm volatile("" : "+v"(lp));  // (derived here, not held across the tile)
              acc[0] = __builtin_amdgcn_mfma_f64_4x4x4f64(
                  ((const double*)sA)[q - lane + 16 * (lp >> 4) + (lp & 3)], b[f][s], acc[0], 0, 0,
                  0);
            } else {
              acc = __builtin_amdgcn_mfma_f64_16x16x4f64(((const double*)sA)[q], b[f][s], acc, 0, 0,
                                                         0);
            }
          } else if constexpr (PREC == kF32) {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(((const float*)sA)[q], b[f][s], acc, 0, 0, 0);
          } else {
            const bf16x8* F = (const bf16x8*)sA;
            if constexpr (PREC == kBF3) {
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[f][s][0], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[f][s][1], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[f][s][0], acc, 0, 0, 0);
            } else {
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[2 * C::NFRAG + q], bp[f][s][0], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[f][s][2], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[f][s][1], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[C::NFRAG + q], bp[f][s][0], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[f][s][1], acc, 0, 0, 0);
              acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(F[q], bp[f][s][0], acc, 0, 0, 0);
            }
          }
        }
      __builtin_amdgcn_s_setprio(0);
      // Through the wave's 16 x 17 buffer to rows of 16 consecutive values.
      wave_sync();  // previous readers of the buffer are done
#pragma unroll
      for (int r = 0; r < 4; ++r) sT[c * 17 + acc_row<PREC>(g, r)] = (double)acc[r];
      wave_sync();
      // out = a0 u0 + a1 (u + dt (rhs + LIFT g))
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double l = (double)pr[tt & 1][j] + sT[(4 * j + (lane >> 4)) * 17 + (lane & 15)];
        const double o = fma(a0, p0[tt & 1][j], a1 * fma(dt, l, pu[tt & 1][j]));
        const bool ok = 4 * j + (lane >> 4) < cnt && 16 * tt + (lane & 15) < NP;
        lmax = fmax(lmax, ok ? fabs(o) : 0.0);
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, o), rsO, upd_off(tt, j), 0, 2);
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

// LIFT fragments in the MFMA A-operand lane order of the precision, columns
// taken face by face: lane (row, k) of step s of face f holds
// LIFT[row][f Nfp + 4 s + k] (k = 4) or LIFT[row][f Nfp + 32 s + 8 k + e],
// e = 0..7 (bf16); zero past Nfp.
std::vector<unsigned char> lift_fragments(const std::vector<double>& LIFT, int N, int prec) {
  const int NP = dgk::np_of(N), NFP = dgk::nfp_of(N), NF = 4 * NFP;
  const int MTL = (NP + 15) / 16;
  const int KB = is_bf(prec) ? 32 : 4, KSF = (NFP + KB - 1) / KB, PARTS = parts_of(prec);
  const size_t nfrag = (size_t)MTL * 4 * KSF * 64;
  auto val = [&](int i, int f, int q) {
    return (i < NP && q < NFP) ? LIFT[(size_t)i * NF + f * NFP + q] : 0.0;
  };
  std::vector<unsigned char> buf;
  if (!is_bf(prec)) {
    buf.resize(nfrag * (prec == kF64 ? 8 : 4));
    for (int tt = 0; tt < MTL; ++tt)
      for (int f = 0; f < 4; ++f)
        for (int s = 0; s < KSF; ++s)
          for (int l = 0; l < 64; ++l) {
            const double v = val(16 * tt + (l & 15), f, 4 * s + (l >> 4));
            const size_t q = (((size_t)tt * 4 + f) * KSF + s) * 64 + l;
            if (prec == kF64) ((double*)buf.data())[q] = v;
            else ((float*)buf.data())[q] = (float)v;
          }
  } else {
    buf.resize(nfrag * 16 * PARTS);
    uint16_t* P0 = (uint16_t*)buf.data();
    for (int tt = 0; tt < MTL; ++tt)
      for (int f = 0; f < 4; ++f)
        for (int s = 0; s < KSF; ++s)
          for (int l = 0; l < 64; ++l)
            for (int e = 0; e < 8; ++e) {
              const double v = val(16 * tt + (l & 15), f, 32 * s + 8 * (l >> 4) + e);
              const size_t q = ((((size_t)tt * 4 + f) * KSF + s) * 64 + l) * 8 + e;
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

template <int N, int PREC, bool R32>
int surface_blocks(dgk_dgmesh m) {
  int per_cu = 0;
  DGK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, surface_update_kernel<N, PREC, R32>,
                                                       64 * SCfg<N, PREC>::W, 0));
  if (per_cu < 1) per_cu = 1;
  m->split_blocks = m->dev->cu_count * per_cu;
  return DGK_OK;
}

template <int N, int PREC, bool R32>
int surface_launch(dgk_dgmesh m, const SurfArgs& a) {
  const int W = SCfg<N, PREC>::W;
  int grid = (a.ntiles + W - 1) / W;
  if (grid > m->split_blocks) grid = m->split_blocks;
  if (grid < 1) return DGK_OK;
  hipLaunchKernelGGL((surface_update_kernel<N, PREC, R32>), dim3(grid), dim3(64 * W), 0,
                     m->dev->stream, a);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

template <int N, int PREC>
int surface_w(dgk_dgmesh m, const SurfArgs* a) {
  if (m->rhs32) return a ? surface_launch<N, PREC, true>(m, *a) : surface_blocks<N, PREC, true>(m);
  return a ? surface_launch<N, PREC, false>(m, *a) : surface_blocks<N, PREC, false>(m);
}

template <int N>
int surface_n(dgk_dgmesh m, const SurfArgs* a) {
  switch (m->lift_prec) {
    case kF64: return surface_w<N, kF64>(m, a);
    case kF32: return surface_w<N, kF32>(m, a);
    case kBF3: return surface_w<N, kBF3>(m, a);
    default: return surface_w<N, kBF6>(m, a);
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
  if (dgk::fused_applies(m)) return dgk::fused_prepare(m);
  if (m->plan && m->rhs && m->lift_frag) return DGK_OK;
  dgk::split_release(m);
  int rc = dgk_deriv_plan_create(m->dev, m->N, m->Dr_h.data(), m->Ds_h.data(), m->Dt_h.data(),
                                 m->r_h.data(), m->s_h.data(), m->t_h.data(), m->prec, &m->plan);
  if (rc != DGK_OK) return rc;
  m->rhs32 = dgk::deriv_plan_divergence_f32(m->plan) ? 1 : 0;
  const size_t state_bytes = (size_t)(m->state_off.empty() ? 0 : m->state_off.back()) * 8 + 16;
  rc = dgk_malloc(m->dev, state_bytes, nullptr, (void**)&m->rhs);
  if (rc != DGK_OK) return rc;
  const std::vector<unsigned char> frag = lift_fragments(m->LIFT_h, m->N, m->lift_prec);
  rc = dgk_malloc(m->dev, frag.size(), frag.data(), &m->lift_frag);
  if (rc != DGK_OK) return rc;
  if (dgk::surf_hi_applies(m)) return dgk::surf_hi_prepare(m);
  return surface(m, nullptr);
}

void dgk::split_release(dgk_dgmesh m) {
  dgk::fused_release(m);
  dgk::surf_hi_release(m);
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
  if (dgk::fused_applies(m)) return dgk::fused_stage(m, u, u0, out, a0, a1, dt, maxabs, mode);
  int rc = split_volume(m, u);
  if (rc != DGK_OK) return rc;
  return split_surface(m, u, u0, out, a0, a1, dt, maxabs, mode);
}

int dgk::split_surface(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                       double a1, const double* dt, unsigned long long* maxabs, int mode) {
  if (dgk::surf_hi_applies(m)) return dgk::surf_hi_surface(m, u, u0, out, a0, a1, dt, maxabs, mode);
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
