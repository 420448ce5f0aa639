// Device-mesh internals shared by the fused stage (burgers.hip) and the
// split stage (split.hip): element shape constants, face masks, element
// tiles and the dgk_dgmesh_s state.
#pragma once

#include <vector>

#include "dgk_internal.h"

namespace dgk {

template <int N>
struct Tet {
  static constexpr int Np = (N + 1) * (N + 2) * (N + 3) / 6;
  static constexpr int Nfp = (N + 1) * (N + 2) / 2;
};

// Face masks of the NUDG node ordering (t=-1, s=-1, r+s+t=-1, r=-1), derived
// from the lattice indices of EquiNodes3D (element/library/gonudg/nodes3d.go:
// 191-201); the warp & blend map keeps face nodes on their faces. Checked
// against the caller's Fmask in dgk_dgmesh_create.
template <int N>
struct FmaskTable {
  int v[4][Tet<N>::Nfp];
  constexpr FmaskTable() : v{} {
    int c[4] = {0, 0, 0, 0};
    int sk = 0;
    for (int n = 0; n <= N; ++n)
      for (int m = 0; m <= N - n; ++m)
        for (int l = 0; l <= N - n - m; ++l) {
          if (n == 0) v[0][c[0]++] = sk;
          if (m == 0) v[1][c[1]++] = sk;
          if (l + m + n == N) v[2][c[2]++] = sk;
          if (l == 0) v[3][c[3]++] = sk;
          ++sk;
        }
  }
};

struct Tile {
  int64_t vbase;  // value offset of the tile's first element in state arrays
  int32_t gbase;  // dense global index of the tile's first element
  int32_t count;  // elements in the tile
};

// Flux factors of face f, A = Fscale n.(1,1,1)/4 and B = Fscale |n.(1,1,1)|/2,
// from the folded metrics alone: for an affine element Fscale_f n_f is the
// unnormalised face normal, -grad t, -grad s, grad(r+s+t), -grad r for
// faces 0..3 (the Normals3D construction, element/library/gonudg/
// normals3d.go:82-185), so Fscale_f n_f.(1,1,1) is -gc, -gb, ga+gb+gc, -ga
// with ga = rx+ry+rz, gb = sx+sy+sz, gc = tx+ty+tz.
__device__ inline void face_factors(int f, double ga, double gb, double gc, double& A,
                                             double& B) {
  const double sn = f == 0 ? -gc : f == 1 ? -gb : f == 2 ? ga + gb + gc : -ga;
  A = 0.25 * sn;
  B = 0.5 * fabs(sn);
}

constexpr int kGeoRows = 26;   // rx..tz, J, 4 x (nx, ny, nz, Fscale)
// rx+ry+rz, sx+sy+sz, tx+ty+tz, 4 x n.(1,1,1), 4 x Fscale,
// 4 x Fscale n.(1,1,1)/4, 4 x Fscale |n.(1,1,1)|/2
constexpr int kBGeoRows = 19;

}  // namespace dgk

struct dgk_dgmesh_s {
  dgk_device dev = nullptr;
  int N = 0, Np = 0, Nfp = 0, npart = 0;
  int64_t K = 0;
  std::vector<int64_t> Kp, state_off;
  double *Dr = nullptr, *Ds = nullptr, *Dt = nullptr, *LIFT = nullptr, *rst = nullptr;
  uint8_t* face_nodes = nullptr;
  double *vtx = nullptr, *geo = nullptr, *bgeo = nullptr;
  int32_t *nbr = nullptr, *codes = nullptr;
  int32_t* nbe = nullptr;          // [4][K] dense global element index of each face's neighbour
  dgk::Tile* tiles = nullptr;      // tiles of tb elements (VALU kernel)
  int ntiles = 0;
  dgk::Tile* tiles64 = nullptr;    // tiles of 16 elements (MFMA kernel, one per wave step)
  int ntiles64 = 0;
  int nfull64 = 0;                 // tiles64[0, nfull64) hold 16 elements each
  int tbm = 64;               // MFMA block: 16 elements per wave, 1/2/4 waves per block
  int tb = 64, blocks_per_cu = 0, chunked = 0;
  int variant = 0;            // 0 auto, 1 VALU, 2 MFMA
  int tile_order = 0;         // 0 partition order, 1 Morton order of tile centroids
  std::vector<uint32_t> ekey; // per element: 30-bit Morton key of its centroid
  unsigned long long* scratch = nullptr;  // reduction slot for synchronous helpers
  // Host copies of the reference operators (row-major) and nodes.
  std::vector<double> Dr_h, Ds_h, Dt_h, LIFT_h, r_h, s_h, t_h;
  // Split stage (orders above 4, or contractions below FP64): volume term
  // through a derivative plan into rhs, then flux + LIFT + RK update.
  int prec = 0;                          // DGK_PREC_* of the D contractions
  int lift_prec = 0;                     // DGK_PREC_* of the LIFT contraction
  dgk_deriv_plan plan = nullptr;         // built on first use for (N, prec)
  void* rhs = nullptr;                   // state-sized volume term (FP64, or FP32
                                         // where the plan's epilogue is FP32)
  int rhs32 = 0;                         // rhs holds FP32 values
  void* lift_frag = nullptr;             // LIFT fragments for lift_prec
  int split_blocks = 0;                  // persistent grid of the surface kernel
  // Fused BF16x3 stage at P = 5, 6 (fused_hi.hip)
  void* fused_frag = nullptr;            // [-Dr | -Ds | -Dt | LIFT] BF16 part fragments
  float* fused_lam = nullptr;            // [6][Np] barycentric weights of the nodes (FP32, remainders)
  int fused_npart = 0;                   // BF16 parts of the fragments built (2: BF16x3, 3: BF16x6)
  // Split stage's FP64/FP32-LIFT surface kernel at P = 5, 6 (surf_hi.hip)
  void* surf_hi_frag = nullptr;          // [MT][Nfp][64] LIFT fragments, FP64 or FP32
  int surf_hi_l64 = 0;                   // surf_hi_frag holds FP64 values
};

namespace dgk {

// Split stage for orders above 4 or contractions below FP64 (split.hip):
// mode 0: out = a0 u0 + a1 (u + dt L(u)) (max |out| into maxabs if given);
// mode 1: out = L(u).
int split_stage(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0, double a1,
                const double* dt, unsigned long long* maxabs, int mode);
// The two halves of split_stage: the volume term into m->rhs, then flux,
// LIFT and the update (rhs may be modified in between, e.g. by viscosity).
int split_volume(dgk_dgmesh m, const double* u);
int split_surface(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                  double a1, const double* dt, unsigned long long* maxabs, int mode);
// Builds the split stage's plan, FP64 volume buffer and LIFT fragments for
// (N, prec) if not there yet (allocates: call outside graph capture).
int split_prepare(dgk_dgmesh m);
void split_release(dgk_dgmesh m);

// Split stage at P = 5, 6 with LIFT in FP64 or FP32: jumps, LIFT and update
// in surf_hi.hip (surf_hi_surface) in place of split.hip's surface kernel.
bool surf_hi_applies(dgk_dgmesh m);
int surf_hi_prepare(dgk_dgmesh m);
void surf_hi_release(dgk_dgmesh m);
int surf_hi_surface(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                    double a1, const double* dt, unsigned long long* maxabs, int mode);
// Fused stage for P = 5, 6 with the D and LIFT contractions both in BF16x3
// or both in BF16x6 (fused_hi.hip): volume term, flux, LIFT and RK update in
// one launch; used by split_stage in place of its two launches unless
// variant 3 asks for them.
bool fused_applies(dgk_dgmesh m);
int fused_prepare(dgk_dgmesh m);
void fused_release(dgk_dgmesh m);
int fused_stage(dgk_dgmesh m, const double* u, const double* u0, double* out, double a0,
                double a1, const double* dt, unsigned long long* maxabs, int mode);

// Volume term of the Burgers RHS through a derivative plan (mixed.hip):
// out = -(div f(u)), f(u) = u^2/2 (1,1,1), for the elements of the tiles
// (16 elements each, runner layout), geo = the 26-row geometry with stride
// gs. The output is FP32 where deriv_plan_divergence_f32 says so (plans
// whose epilogue computes in FP32), else FP64.
int deriv_plan_divergence(dgk_deriv_plan plan, const Tile* tiles, int ntiles, const double* u,
                          const double* geo, int64_t gs, void* out);
bool deriv_plan_divergence_f32(dgk_deriv_plan plan);

}  // namespace dgk
