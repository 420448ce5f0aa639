// Surface kernel of the split stage at P = 5, 6 with LIFT in FP64 (the FP64,
// I8X6 and LIFT-in-FP64 settings) or FP32 (the FP32 setting): the
// Lax-Friedrichs jumps, LIFT on FP64 or FP32 matrix cores and the RK update,
//   out = a0 u0 + a1 (u + dt (rhs + LIFT g)),   (mode 0)
//   out = rhs + LIFT g                         (mode 1),
// rhs the volume term the derivative plan wrote (FP64, or FP32 for the plans
// with an FP32 epilogue). Same flux as the other stage kernels:
//   g = (u- - u+) (A (u- + u+) - B max(|u-|, |u+|)),
// A = Fscale n.(1,1,1)/4, B = Fscale |n.(1,1,1)|/2.
//
// Block: MT matrix waves (one 16-row tile of LIFT each, its fragments in
// registers) and 4 memory waves (4 elements of every 16-element tile each),
// one tile per phase between block barriers; in phase n
//   matrix waves: tile n's rhs and u0 rows loaded during the product
//                 G^T LIFT^T (v_mfma_f64_16x16x4_f64 or v_mfma_f32_16x16x4_f32,
//                 G the tile's jumps from LDS, transposed so that 16 lanes hold 16 consecutive rows
//                 of an element) -> RK update, 16 consecutive rows per store;
//   memory waves: the face records of tile n+3 (loaded in phase n-1) into
//                 LDS; the loads of tile n+4's records and of tile n+2 (u and
//                 the neighbour traces, gathered through the records stored in
//                 phase n-1); tile n+1's u (loaded in phase n-1) into LDS and
//                 its jumps (own traces from that LDS tile) into G.
// The fused BF16 stage (fused_hi.hip) walks its tiles the same way; here the
// memory waves have no volume columns to form, and u is loaded once by this
// kernel (own traces and update) where the one-wave-per-tile surface kernel
// of split.hip reads it twice. With FP32 LIFT the jumps are formed in FP64 and
// rounded to FP32 for the product, whose FP32 sums are added to the volume
// term in FP64 (the rounding of the split.hip kernel's FP32 form).
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dgmesh_internal.h"

namespace {

typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

template <int N>
struct SCfgHi {
  static constexpr int NP = dgk::Tet<N>::Np, NFP = dgk::Tet<N>::Nfp;
  static constexpr int MT = (NP + 15) / 16;  // row tiles of LIFT = matrix waves
  static constexpr int KF = NFP;             // k = 4 steps over the 4 Nfp columns
  static constexpr int MW = 4;               // memory waves
  static constexpr int EPW = 16 / MW;        // elements per memory wave
  static constexpr int NT = 64 * (MT + MW);
  static constexpr int NPU = NP | 1;         // LDS row stride of u, doubles
  static constexpr int GRS = 4 * NFP + 1;    // LDS row stride of the jumps (+ a dump column), odd
  static constexpr int UPL = (EPW * NP + 63) / 64;        // u values per lane
  static constexpr int GPL = (EPW * 4 * NFP + 63) / 64;  // trace pairs per lane
};

// Face record of (face f, element j) of a memory wave's elements: value
// offset of the neighbour's node 0, row of the orientation table, flux
// factors.
struct FRecHi {
  int32_t nb, fnrow;
  double fa, fb;
};

struct SurfHiArgs {
  int ntiles;
  const void* frag;           // [MT][KF][64] LIFT fragments, FP64 or FP32 (LT)
  const uint8_t* face_nodes;  // [4][24][Nfp]
  const double* bgeo;         // [19][K]; rows 0..2 the folded metrics
  int64_t K;
  const int32_t* nbr;         // [4][K]
  const int32_t* codes;       // [K]
  const double* u;
  const double* u0;
  const void* rhs;            // volume term, FP64 or FP32 (R32)
  double* out;
  double a0, a1;
  const double* dt;
  unsigned long long* maxabs;
  int mode;                   // 0: RK stage, 1: out = rhs + LIFT g
};

// (The tile list is a separate const __restrict__ argument: the compiler may
// then read its wave-uniform entries through the scalar cache.)
// LT: type of LIFT and of the jumps in the product (double or float).
template <int N, bool R32, typename LT>
__global__ void __launch_bounds__(SCfgHi<N>::NT)
    surf_hi_kernel(SurfHiArgs a, const dgk::Tile* __restrict__ tiles) {
  constexpr bool L64 = std::is_same<LT, double>::value;
  using AccT = typename std::conditional<L64, doublex4, floatx4>::type;
  using C = SCfgHi<N>;
  constexpr int NP = C::NP, NFP = C::NFP, MT = C::MT, KF = C::KF, MW = C::MW, EPW = C::EPW;
  constexpr int NPU = C::NPU, GRS = C::GRS, UPL = C::UPL, GPL = C::GPL;
  constexpr dgk::FmaskTable<N> FM{};
  // jumps of two tiles, [element][face f, point qq] (+ a dump column); u of
  // two tiles (+ a dump slot) and their tile records, for the update
  __shared__ LT sG[2][16 * GRS];
  __shared__ double sU[2][16 * NPU + 1];
  __shared__ int64_t sTv[2];
  __shared__ int sTc[2];
  // per memory wave: face records of its elements for four tiles (slot =
  // tile % 4)
  __shared__ FRecHi sRec[MW][4][4 * EPW];
  __shared__ uint8_t sFN[4 * 24 * NFP];
  __shared__ uint8_t sFM[4 * NFP];
  for (int i = threadIdx.x; i < 4 * 24 * NFP; i += C::NT) sFN[i] = a.face_nodes[i];
  for (int i = threadIdx.x; i < 4 * NFP; i += C::NT) sFM[i] = (uint8_t)FM.v[i / NFP][i % NFP];
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
  // lane index) rather than held across the tile loop.
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
  __syncthreads();

  if (w < MT) {
    // ---------------- matrix wave: LIFT rows 16 w .. 16 w + 15 ----------------
    // fragment s: LIFT[16 w + l % 16][4 s + l / 16] in lane l
    LT lf[KF];
#pragma unroll
    for (int s = 0; s < KF; ++s) lf[s] = ((const LT*)a.frag)[((size_t)w * KF + s) * 64 + lane];
    const double dt = a.mode == 0 ? *a.dt : 1.0;
    const double a0 = a.mode == 0 ? a.a0 : 0.0, a1 = a.mode == 0 ? a.a1 : 1.0;
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
      // The product's register r of lane l holds row 16 w + l % 16 of element
      // g + 4 r (FP64: result rows interleave as l / 16 + 4 r) or 4 g + r
      // (FP32).
      auto elem = [&](int r) { return L64 ? g + 4 * r : 4 * g + r; };
      // rhs and u0 of the rows this lane updates, loaded during the product
      // through buffer resources sized to the tile (loads past it read zero)
      const int nbytes = count * NP * 8;
      const __amdgpu_buffer_rsrc_t rsR = __builtin_amdgcn_make_buffer_rsrc(
          (void*)((const char*)a.rhs + vbase * (R32 ? 4 : 8)), 0, R32 ? nbytes / 2 : nbytes, 0x00020000);
      const bool with_u0 = a0 != 0.0;
      const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(with_u0 ? a.u0 + vbase : a.u), 0, with_u0 ? nbytes : 0, 0x00020000);
      double rv[4], u0v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int off = elem(r) * NP + rr;
        if constexpr (R32)
          rv[r] = (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsR, off * 4, 0, 0));
        else
          rv[r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsR, off * 8, 0, 0));
        u0v[r] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs0, off * 8, 0, 2));
      }
      // G^T LIFT^T
      AccT acc = {};
      __builtin_amdgcn_s_setprio(1);
      if constexpr (L64) {
#pragma unroll
        for (int s = 0; s < KF; ++s)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(sG[bu][c * GRS + 4 * s + g], lf[s], acc, 0, 0, 0);
      } else {
#pragma unroll
        for (int s = 0; s < KF; ++s)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sG[bu][c * GRS + 4 * s + g], lf[s], acc, 0, 0, 0);
      }
      __builtin_amdgcn_s_setprio(0);
      // Stores through a buffer resource sized to the tile: the ones past it
      // (elements >= count, rows >= Np) are dropped, so the partial tile takes
      // no branches.
      const __amdgpu_buffer_rsrc_t rsO =
          __builtin_amdgcn_make_buffer_rsrc((void*)(a.out + vbase), 0, nbytes, 0x00020000);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = elem(r);
        const double uv = a.mode == 0 ? sU[bu][e * NPU + rr] : 0.0;
        const double o = fma(a0, u0v[r], a1 * fma(dt, rv[r] + (double)acc[r], uv));
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
  // Face records of this wave's elements for tile n go to slot n % 4 of
  // sRec[q], written three phases ahead from loads issued a phase before:
  // lane s < 4 EPW takes face s / EPW of element e0 + s % EPW (clamped to the
  // tile). Every lane loads, lane l taking the record of lane l % (4 EPW), and
  // tiles past the block's last read its last: the loads are the same in
  // every phase, so the waits for them count exactly.
  static_assert((4 * EPW & (4 * EPW - 1)) == 0, "records per wave: a power of two");
  int32_t rnb = 0, rcode = 0;
  double rga = 0.0, rgb = 0.0, rgc = 0.0;
  auto rec_load = [&](int n) {
    const int ln = lane_now() & (4 * EPW - 1), rf = ln / EPW, rj = ln % EPW;
    const dgk::Tile tl = tile_of(n < nb ? n : nb - 1);
    const int e = e0 + rj < tl.count ? e0 + rj : tl.count - 1;
    const int64_t ge = tl.gbase + e;
    rnb = a.nbr[(int64_t)rf * K + ge];
    rcode = a.codes[ge];
    rga = a.bgeo[ge];
    rgb = a.bgeo[K + ge];
    rgc = a.bgeo[2 * K + ge];
  };
  auto rec_store = [&](int n) {
    if (n >= nb) return;
    const int ln = lane_now(), rf = ln / EPW;
    if (ln < 4 * EPW) {
      FRecHi r;
      r.nb = rnb;
      r.fnrow = (rf * 24 + ((rcode >> (5 * rf)) & 31)) * NFP;
      // Fscale n.(1,1,1) of face rf: -gc, -gb, ga + gb + gc, -ga (the face
      // factors of dgk::face_factors), as coefficients in {-1, 0, 1} of
      // (ga, gb, gc) rather than a branch on the lane's face
      const double c0 = (double)(rf == 2) - (double)(rf == 3);
      const double c1 = (double)(rf == 2) - (double)(rf == 1);
      const double c2 = (double)(rf == 2) - (double)(rf == 0);
      const double sn = fma(c0, rga, fma(c1, rgb, c2 * rgc));
      r.fa = 0.25 * sn;
      r.fb = 0.5 * fabs(sn);
      sRec[q][n & 3][ln] = r;
    }
  };
  // Per lane and trace pair k (pair p = lane + 64 k: slot p / NFP = f EPW + j,
  // point qq = p % NFP), fixed for the launch: index of the own trace in a
  // tile's u (12 bits), column of the jump in G (9 bits; the dump column past
  // the pairs), record slot (5 bits), element (2 bits).
  uint32_t fw[GPL];
#pragma unroll
  for (int k = 0; k < GPL; ++k) {
    const int p = lane + 64 * k;
    const bool ok = EPW * 4 * NFP % 64 == 0 || p < EPW * 4 * NFP;
    const int pp = ok ? p : EPW * 4 * NFP - 1;
    const int sl = pp / NFP, qq = pp % NFP, f = sl / EPW, j = sl % EPW;
    const int iu = (e0 + j) * NPU + sFM[f * NFP + qq];
    const int col = ok ? f * NFP + qq : 4 * NFP;
    fw[k] = (uint32_t)iu | ((uint32_t)col << 12) | ((uint32_t)sl << 21) | ((uint32_t)j << 26);
  }
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
    const FRecHi* R = sRec[q][nn & 3];
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
      const FRecHi& r = R[pp / NFP];
      P[slot][k] = a.u[r.nb + sFN[r.fnrow + pp % NFP]];
    }
    TV[slot] = tl.vbase;
    TC[slot] = tl.count;
  };
  // Tile n's values (loaded into slot) -> LDS u and the jumps.
  auto prepare = [&](int n, auto SL) {
    constexpr int slot = decltype(SL)::value;
    const int lane = lane_now();
    const int bu = n & 1;
    const FRecHi* R = sRec[q][n & 3];
    double* su = sU[bu];
#pragma unroll
    for (int k = 0; k < UPL; ++k) {
      const int p = lane + 64 * k;
      const int o = p < EPW * NP ? (e0 + p / NP) * NPU + p % NP : 16 * NPU;
      su[o] = U[slot][k];
    }
    if (q == 0 && lane == 0) {
      sTv[bu] = TV[slot];
      sTc[bu] = TC[slot];
    }
    wsync();  // the wave's u of the tile is in LDS
    LT* G = sG[bu];
#pragma unroll
    for (int k = 0; k < GPL; ++k) {
      const uint32_t fk = fw[k];
      const int col = (fk >> 12) & 511, j = fk >> 26;
      const FRecHi& r = R[(fk >> 21) & 31];
      const double m = su[fk & 4095], pv = P[slot][k];
      G[(e0 + j) * GRS + col] = (LT)((m - pv) * fma(r.fa, m + pv, -r.fb * fmax(fabs(m), fabs(pv))));
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
    wsync();
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

template <int N, bool R32, typename LT>
int surf_hi_launch(dgk_dgmesh m, const SurfHiArgs& a) {
  if (a.ntiles == 0) return DGK_OK;
  int grid = m->dev->cu_count;
  if (grid > a.ntiles) grid = a.ntiles;
  hipLaunchKernelGGL((surf_hi_kernel<N, R32, LT>), dim3(grid), dim3(SCfgHi<N>::NT), 0, m->dev->stream, a,
                     m->tiles64);
  DGK_LAUNCH_CHECK();
  return DGK_OK;
}

}  // namespace

bool dgk::surf_hi_applies(dgk_dgmesh m) {
  return (m->N == 5 || m->N == 6) && (m->lift_prec == DGK_PREC_F64 || m->lift_prec == DGK_PREC_F32);
}

int dgk::surf_hi_prepare(dgk_dgmesh m) {
  if (m->surf_hi_frag) return DGK_OK;
  const int N = m->N, NP = dgk::np_of(N), NFP = dgk::nfp_of(N), NF = 4 * NFP;
  const int MT = (NP + 15) / 16, KF = NFP;
  const bool l64 = m->lift_prec == DGK_PREC_F64;
  // fragment s of row tile t: lane l holds LIFT[16 t + l % 16][4 s + l / 16]
  // (rounded to FP32 for the FP32 form)
  std::vector<double> frag((size_t)MT * KF * 64, 0.0);
  for (int t = 0; t < MT; ++t)
    for (int s = 0; s < KF; ++s)
      for (int l = 0; l < 64; ++l) {
        const int row = 16 * t + (l & 15), k = 4 * s + (l >> 4);
        if (row < NP && k < NF) frag[((size_t)t * KF + s) * 64 + l] = m->LIFT_h[(size_t)row * NF + k];
      }
  m->surf_hi_l64 = l64 ? 1 : 0;
  if (l64) return dgk_malloc(m->dev, frag.size() * 8, frag.data(), &m->surf_hi_frag);
  const std::vector<float> f32(frag.begin(), frag.end());
  return dgk_malloc(m->dev, f32.size() * 4, f32.data(), &m->surf_hi_frag);
}

void dgk::surf_hi_release(dgk_dgmesh m) {
  if (m->surf_hi_frag) dgk_free(m->dev, m->surf_hi_frag);
  m->surf_hi_frag = nullptr;
}

int dgk::surf_hi_surface(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                         double a1, const double* dt, unsigned long long* maxabs, int mode) {
  int rc = surf_hi_prepare(m);
  if (rc != DGK_OK) return rc;
  const bool l64 = m->lift_prec == DGK_PREC_F64;
  DGK_REQUIRE(m->surf_hi_l64 == (l64 ? 1 : 0), "split surface: LIFT fragments prepared for another precision");
  SurfHiArgs a;
  a.ntiles = m->ntiles64;
  a.frag = m->surf_hi_frag;
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
  const bool r32 = m->rhs32 != 0;
  switch (m->N) {
    case 5:
      if (l64) return r32 ? surf_hi_launch<5, true, double>(m, a) : surf_hi_launch<5, false, double>(m, a);
      return r32 ? surf_hi_launch<5, true, float>(m, a) : surf_hi_launch<5, false, float>(m, a);
    case 6:
      if (l64) return r32 ? surf_hi_launch<6, true, double>(m, a) : surf_hi_launch<6, false, double>(m, a);
      return r32 ? surf_hi_launch<6, true, float>(m, a) : surf_hi_launch<6, false, float>(m, a);
  }
  dgk::set_error("split surface: N=%d outside 5..6", m->N);
  return DGK_ERR_UNSUPPORTED;
}
