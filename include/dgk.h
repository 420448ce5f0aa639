/*
 * dgk.h -- C ABI of libdgk, the MI355X (gfx950) native backend for the
 * tetrahedral DG hot path of DGKernel.
 *
 * Two layers:
 *
 *  1. The device/memory/kernel/graph layer that replaces the gocca API the
 *     reference's runner calls through cgo (citations are the reference call
 *     sites each function replaces). Kernels are built from the same OKL
 *     text the runner hands to gocca, read by a front end that maps its
 *     statements onto hand-written gfx950 kernels (no JIT compiler).
 *
 *  2. Hot-path entry points, callable directly (and from inside a graph
 *     capture): physical derivative, FaceConnector exchange, fused
 *     Lax-Friedrichs/LIFT/SSP-RK3 stage, CFL step size, Persson-Peraire
 *     sensor and the learned-viscosity MLP.
 *
 * Conventions: every function returns an int status (DGK_OK == 0). On
 * failure dgk_last_error() returns a message for the calling thread. No
 * function aborts the process. Pointers named *_dev are device pointers;
 * all other pointers are host memory. Entry points enqueue work on the
 * device's stream and return without synchronizing unless stated; host
 * buffers passed to copies may be reused as soon as the call returns.
 *
 * Partitioned arrays follow the runner's pooled layout
 * (runner/builder/builder.go:113-159): one allocation X_global per array,
 * partition p's values start at value offset X_offsets[p], element l of
 * partition p has its Np nodal values at X_global[X_offsets[p] + l*Np + n]
 * (element-major, runner/kernel_copy.go:86-110).
 */
#ifndef DGK_H
#define DGK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DGK_OK = 0,
  DGK_ERR_INVALID = 1,     /* bad argument, shape or state */
  DGK_ERR_HIP = 2,         /* a HIP runtime call failed */
  DGK_ERR_NOT_FOUND = 3,   /* unknown kernel name / signature in the registry */
  DGK_ERR_UNSUPPORTED = 4, /* valid request this build does not implement */
  DGK_ERR_NUMERIC = 5      /* setup found a non-positive Jacobian or similar */
};

typedef struct dgk_device_s* dgk_device;
typedef struct dgk_kernel_s* dgk_kernel;
typedef struct dgk_graph_s* dgk_graph;
typedef struct dgk_dgmesh_s* dgk_dgmesh;
typedef struct dgk_timer_s* dgk_timer;

/* Last error message of the calling thread ("" if none). */
const char* dgk_last_error(void);
/* Library version string. */
const char* dgk_version(void);

/* ---------------------------------------------------------------------- */
/* Device: replaces gocca.NewDevice (utils/device_helpers.go:9-39),        */
/* OCCADevice.Finish (runner/kernel_execution.go:61), .Mode, .Free.        */
/* props: JSON such as {"mode": "HIP", "device_id": 0}. Modes "HIP" and    */
/* "CUDA" (accepted for existing callers) select the GPU; other modes      */
/* return DGK_ERR_UNSUPPORTED.                                              */
/* ---------------------------------------------------------------------- */
int dgk_device_create(const char* props, dgk_device* out);
int dgk_device_destroy(dgk_device dev);
int dgk_device_sync(dgk_device dev);
const char* dgk_device_mode(dgk_device dev);
/* name: buffer of at least 256 bytes. cu_count/lds_bytes may be NULL. */
int dgk_device_info(dgk_device dev, char* name, int* cu_count, int* lds_bytes, int* arch_major,
                    int* arch_minor);

/* ---------------------------------------------------------------------- */
/* Memory: replaces OCCADevice.Malloc (runner/runner.go:83,196,210) and    */
/* OCCAMemory.CopyFrom/CopyTo[WithOffset] (runner/kernel_copy.go:30,98,    */
/* 331; runner/memory_operations.go:13-418). Copies are synchronous with   */
/* respect to the host buffer (cgo may not retain Go pointers).            */
/* ---------------------------------------------------------------------- */
int dgk_malloc(dgk_device dev, size_t bytes, const void* init_host, void** out_dev);
int dgk_free(dgk_device dev, void* ptr_dev);
int dgk_copy_h2d(dgk_device dev, void* dst_dev, const void* src_host, size_t bytes,
                 size_t dst_offset_bytes);
int dgk_copy_d2h(dgk_device dev, void* dst_host, const void* src_dev, size_t bytes,
                 size_t src_offset_bytes);
int dgk_copy_d2d(dgk_device dev, void* dst_dev, const void* src_dev, size_t bytes);
/* Fill with a byte value by a small kernel (safe under graph capture). */
int dgk_fill_bytes(dgk_device dev, void* dst_dev, int value, size_t bytes);

/* ---------------------------------------------------------------------- */
/* Source-built kernels: replace OCCADevice.BuildKernelFromString          */
/* (runner/runner.go:88-123, called with the runner's preamble + "\n" +    */
/* the caller's kernel source, the kernel name, and OCCA props) and        */
/* OCCAKernel.RunWithArgs (runner/kernel_execution.go:57).                 */
/* There is no OKL compiler: the text is read by a front end that knows    */
/* the runner's preamble (runner/builder/builder.go:172-410: int_t, NPART, */
/* KpartMax, static matrices, X_PART and MATMUL_X / MATMUL_ADD_X macros)   */
/* and these statements in the @outer partition loop: X_PART aliases,      */
/* MATMUL macros, element-node and flat-node pointwise loops, and the      */
/* element matvec loop (dgk_kernel_list prints the forms). Each statement  */
/* runs as one launch over all partitions on hand-written kernels; the     */
/* reference's "differentiate" body (three MATMULs + chain rule,           */
/* tetnudg/tet_nudg_mesh_physical_derivative_test.go:850-892, :240-278)    */
/* runs as one launch of the physical-derivative kernel. Other kernels     */
/* return DGK_ERR_UNSUPPORTED with the offending source line.              */
/* ---------------------------------------------------------------------- */
enum { DGK_ARG_MEM = 0, DGK_ARG_F64 = 1, DGK_ARG_F32 = 2, DGK_ARG_I64 = 3, DGK_ARG_I32 = 4 };

typedef struct {
  int kind;
  int pad;
  union {
    const void* mem;
    double f64;
    float f32;
    int64_t i64;
    int32_t i32;
  } v;
} dgk_arg;

/* The statement forms the front end knows, newline-separated. */
int dgk_kernel_list(char* buf, size_t size);
/* Read source (the full text gocca would compile) and prepare kernel name.
 * Static matrices are copied to the device here (allocates: call outside
 * graph capture). */
int dgk_kernel_build_from_source(dgk_device dev, const char* source, const char* name,
                                 dgk_kernel* out);
/* Parse only (no device): the statement program the library would run,
 * one line per launch, into buf. */
int dgk_kernel_describe(const char* source, const char* name, char* buf, size_t size);
/* Enqueue the kernel's launches. args: one per signature parameter, in
 * signature order (runner/kernel_execution.go:150-227): device memory for
 * pointers (non-NULL), DGK_ARG_F64/F32/I64/I32 matching each scalar's type.
 * host_K: the partition sizes (NPART int_t values) if the caller has them;
 * NULL reads them from the K argument (a synchronous copy, as the runner's
 * launch is synchronous anyway). */
int dgk_kernel_launch(dgk_kernel k, const void* host_K, int nargs, const dgk_arg* args);
int dgk_kernel_free(dgk_kernel k);

/* ---------------------------------------------------------------------- */
/* Graphs and timing                                                       */
/* ---------------------------------------------------------------------- */
int dgk_graph_begin(dgk_device dev);
int dgk_graph_end(dgk_device dev, dgk_graph* out);
int dgk_graph_launch(dgk_graph g);
int dgk_graph_destroy(dgk_graph g);
int dgk_timer_create(dgk_device dev, dgk_timer* out);
int dgk_timer_start(dgk_timer t);
int dgk_timer_stop(dgk_timer t);
/* Synchronizes on the stop event; milliseconds between start and stop. */
int dgk_timer_elapsed_ms(dgk_timer t, double* ms);
int dgk_timer_destroy(dgk_timer t);

/* ---------------------------------------------------------------------- */
/* Hot path: physical derivative                                            */
/* Replaces the OKL kernel "differentiate" of                              */
/* tetnudg/tet_nudg_mesh_physical_derivative_test.go:850-892 (FP64) and    */
/* :240-278 (FP32 metrics/outputs) with the host result of                 */
/* calcPhysicalDerivative (:344-367):                                       */
/*   Ur = Dr u, Us = Ds u, Ut = Dt u;                                       */
/*   du/dx = rx Ur + sx Us + tx Ut (likewise y, z).                         */
/* ---------------------------------------------------------------------- */
enum { DGK_F64 = 0, DGK_F32 = 1 };

/* Per-node metrics (the reference's API layout: 13 Np*K arrays).
 * Dr/Ds/Dt_dev: Np x Np row-major FP64 on the device.
 * metrics_dev[9]: rx, ry, rz, sx, sy, sz, tx, ty, tz global pointers.
 * Offsets arrays (value offsets, int64, device, npart+1 entries) for u,
 * metrics and outputs. mtype/otype: DGK_F64 or DGK_F32 for the metric and
 * output arrays (u is always FP64). */
int dgk_phys_deriv_nodal(dgk_device dev, int N, int npart, const int64_t* K_host,
                         const double* Dr_dev, const double* Ds_dev, const double* Dt_dev,
                         const double* u_dev, const int64_t* u_off_dev,
                         const void* const* metrics_dev, int mtype,
                         const int64_t* m_off_dev, void* dudx_dev, void* dudy_dev, void* dudz_dev,
                         int otype, const int64_t* o_off_dev);

/* Allocation-free form of dgk_phys_deriv_nodal, with the same arguments and
 * results. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops). dgk_phys_deriv_nodal keeps the partition sizes of a launch over more
 * than 256 partitions in a per-device buffer it allocates (and fills with a
 * synchronous copy) on first need; this form stages them into a fixed table
 * the library holds on every GPU it runs on (65535 partitions at most, ~512
 * KiB) with small kernels, so it makes no allocation, no host copy and no
 * synchronisation. All calls on one GPU share that table: calls that overlap
 * on one GPU (different streams) race. npart above 65535 is rejected. */
int dgk_phys_deriv_nodal_noalloc(dgk_device dev, int N, int npart, const int64_t* K_host,
                                 const double* Dr_dev, const double* Ds_dev, const double* Dt_dev,
                                 const double* u_dev, const int64_t* u_off_dev,
                                 const void* const* metrics_dev, int mtype,
                                 const int64_t* m_off_dev, void* dudx_dev, void* dudy_dev,
                                 void* dudz_dev, int otype, const int64_t* o_off_dev);

/* Affine elements (native layout): the operator is prepared once per order
 * and precision (MFMA fragments), then applied to K elements:
 *   u_dev: dense element-major [k*Np + n], FP64, 16-byte aligned;
 *   geo_dev: SoA rows rx ry rz sx sy sz tx ty tz at geo_dev[q*geo_stride + k]
 *            (e.g. the first 9 rows of dgk_dgmesh_geometry_dev, stride K);
 *   outputs dense element-major, FP64 for DGK_PREC_F64, FP32 otherwise.
 * DGK_PREC_F64: FP64 MFMA (v_mfma_f64_16x16x4), FP64 everywhere.
 * Below FP64 each element's linear interpolant through its vertex nodes is
 * removed in FP64 before rounding and its derivative added back, so rounding
 * scales with |u - Lu| rather than |u|; r, s, t (the reference nodes) locate
 * the vertex nodes and are required.
 * DGK_PREC_F32: operator and remainder rounded to FP32, FP32 MFMA.
 * DGK_PREC_BF16X3: two-part bf16 splits, hi*hi + hi*lo + lo*hi
 *   (Henry, Tang & Heinecke, ARITH 2019), ~16 bits per operand.
 * DGK_PREC_BF16X6: three-part bf16 splits of the FP32 value, six products,
 *   ~24 bits.
 * These accumulate in FP32 and write FP32 outputs; the slopes of Lu and the
 * metrics are applied in FP32 (a few FP32 ulps, below the products' own
 * error).
 * DGK_PREC_I8X6: exact-integer split on the INT8 matrix cores (Ozaki scheme):
 *   operator rows and each element's remainder scaled by powers of two and
 *   written as six signed 8-bit digits (46 bits); the 21 digit pairs of
 *   highest weight accumulate exactly in int32 and are combined in FP64.
 *   FP64 outputs, ~3e-12 relative to FP64. */
enum {
  DGK_PREC_F64 = 0,
  DGK_PREC_F32 = 1,
  DGK_PREC_BF16X3 = 2,
  DGK_PREC_BF16X6 = 3,
  DGK_PREC_I8X6 = 4
};
typedef struct dgk_deriv_plan_s* dgk_deriv_plan;
int dgk_deriv_plan_create(dgk_device dev, int N, const double* Dr_host, const double* Ds_host,
                          const double* Dt_host, const double* r_host, const double* s_host,
                          const double* t_host, int prec, dgk_deriv_plan* out);
int dgk_deriv_plan_destroy(dgk_deriv_plan plan);
int dgk_phys_deriv_affine(dgk_deriv_plan plan, int64_t K, const double* u_dev,
                          const double* geo_dev, int64_t geo_stride, void* dudx_dev,
                          void* dudy_dev, void* dudz_dev);

/* MATMUL_<X>(IN, OUT, K[part]) / MATMUL_ADD_<X> of
 * runner/builder/builder.go:355-410 over all partitions in one launch:
 * OUT[e*rows + i] (=|+=) sum_j A(i,j) IN[e*cols + j]. A_dev is the
 * column-major device matrix of runner/runner.go:395-400 (A[j*rows + i]). */
int dgk_matmul(dgk_device dev, int rows, int cols, const double* A_colmajor_dev, int npart,
               const int64_t* K_host, const double* in_dev, const int64_t* in_off_dev,
               double* out_dev, const int64_t* out_off_dev, int add);

/* ---------------------------------------------------------------------- */
/* Hot path: FaceConnector exchange (utils/face_connector.go:143-187 index */
/* semantics; runtime exchange of utils/face_connector_test.go:104-156):   */
/*   P_global[P_off[dst] + place[j]] = u_global[u_off[src] + pick[i]]       */
/* for every (src, dst) pair, including src == dst, in one launch.         */
/* The caller passes the concatenated lists: pick lists src-major         */
/* (pick[src][dst] for src, then dst), place lists dst-major (place[dst]   */
/* [src]); counts[src*npart + dst] gives each pair's length. If the place  */
/* lists fill every slot of every destination once (the FaceConnector's    */
/* do), the plan stores the entries in destination-slot order and the      */
/* launch writes P with unit stride; other lists are kept pair by pair.    */
/* ---------------------------------------------------------------------- */
typedef struct dgk_face_plan_s* dgk_face_plan;
int dgk_face_plan_create(dgk_device dev, int npart, const int64_t* counts_host,
                         const int32_t* pick_host, const int32_t* place_host,
                         dgk_face_plan* out);
int dgk_face_plan_destroy(dgk_face_plan plan);
int dgk_face_exchange(dgk_face_plan plan, const double* u_dev, const int64_t* u_off_dev,
                      double* P_dev, const int64_t* P_off_dev);

/* ---------------------------------------------------------------------- */
/* DG mesh on the device: partitions, neighbours, orientation codes and    */
/* FP64 affine geometry computed on the GPU from vertex coordinates.       */
/* ---------------------------------------------------------------------- */
typedef struct {
  int N;                       /* polynomial order */
  int npart;                   /* partitions */
  const int64_t* K;            /* [npart] elements per partition */
  const int64_t* state_off;    /* [npart+1] value offsets of state arrays (runner layout) */
  const double* Dr;            /* Np x Np row-major */
  const double* Ds;
  const double* Dt;
  const double* LIFT;          /* Np x 4Nfp row-major */
  const int* Fmask;            /* [4][Nfp] */
  const double* r;             /* [Np] reference node coordinates */
  const double* s;
  const double* t;
  const int* face_nodes;       /* [4][24][Nfp]: neighbour volume node of own face node, per code */
  const int64_t* EToE;         /* [Ktot][4] global element of each face's neighbour;  */
                               /* global elements are numbered partition-major        */
  const int* codes;            /* [Ktot][4] orientation code 6*f2 + sigma */
  const double* vtx;           /* [Ktot][4][3] vertex coordinates (affine map corners) */
} dgk_dgmesh_desc;

int dgk_dgmesh_create(dgk_device dev, const dgk_dgmesh_desc* desc, dgk_dgmesh* out);
int dgk_dgmesh_destroy(dgk_dgmesh m);
/* Device geometry, SoA [q][Ktot]: q = 0..8 rx ry rz sx sy sz tx ty tz,
 * 9 J, then for faces f = 0..3: 10+4f nx, 11+4f ny, 12+4f nz, 13+4f Fscale
 * (26 rows). Copied to host. */
int dgk_dgmesh_geometry(dgk_dgmesh m, double* geo_host);
/* Device pointer to the SoA geometry (26 x Ktot doubles). */
int dgk_dgmesh_geometry_dev(dgk_dgmesh m, const double** geo_dev);
/* Minimum distance between distinct nodes of one element over the mesh,
 * FP64 on the device (examples/3d_scalar_burgers_equation.go:697-717).
 * Synchronous. */
int dgk_dgmesh_min_node_distance(dgk_dgmesh m, double* hmin);
/* Writes uP[k][f][fp] (dense global element order) as the fused stage
 * kernel sees it: the neighbour trace gathered through EToE + codes. */
int dgk_dgmesh_trace_plus(dgk_dgmesh m, const double* u_dev, double* uP_dev);

/* ---------------------------------------------------------------------- */
/* Hot path: fused Burgers stage (strong-form DG, Lax-Friedrichs flux,     */
/* LIFT, SSP-RK3 combination) on all partitions in one persistent launch:  */
/*   out = a0*u0 + a1*(u + dt*L(u)),                                        */
/*   L(u) = -div(u^2/2 (1,1,1)) + LIFT (Fscale (n.f(u-) - n.f*)),           */
/*   n.f* = (n.f(u-) + n.f(u+))/2 - alpha |n.(1,1,1)| (u+ - u-)/2,          */
/*   alpha = max(|u-|, |u+|).                                               */
/* dt is read from device memory (dt_dev). If maxabs_dev is not NULL the   */
/* kernel also folds max|out| into *maxabs_dev (atomic, bitwise max of the */
/* non-negative double). out may alias u0 but not u.                        */
/* ---------------------------------------------------------------------- */
int dgk_burgers_stage(dgk_dgmesh m, const double* u_dev, const double* u0_dev, double* out_dev,
                      double a0, double a1, const double* dt_dev, double* maxabs_dev);
/* L(u) alone into rhs_dev (runner layout). N = 1..6. */
int dgk_burgers_rhs(dgk_dgmesh m, const double* u_dev, double* rhs_dev);
/* max|u| over all elements into *maxabs_dev (which must hold 0 or a
 * previous max). */
int dgk_max_abs(dgk_dgmesh m, const double* u_dev, double* maxabs_dev);
/* dt = cfl * hmin / *maxabs_dev into *dt_dev (FP64), then *maxabs_dev = 0. */
int dgk_cfl_dt(dgk_device dev, double* maxabs_dev, double cfl, double hmin, double* dt_dev);
/* Tile shape of the selected stage kernel: elements per block (MFMA kernel:
 * 16, 32 or 64, i.e. 1, 2 or 4 waves of 16 elements; VALU kernel: 64, 128
 * or 256) and blocks per CU of the persistent grid. 0 restores the tuned
 * default. Set the variant first. */
int dgk_burgers_set_tile(dgk_dgmesh m, int elems_per_block, int blocks_per_cu);
/* Stage kernel: 0 automatic (FP64 MFMA for N <= 4, split for N = 5, 6,
 * the fused stage for N = 5, 6 with D and LIFT both in DGK_PREC_BF16X3 or
 * both in DGK_PREC_BF16X6),
 * 1 VALU (one element per thread, FP64 FMA with scalar operator operands),
 * 2 FP64 MFMA, 3 split (volume kernel + surface kernel, see
 * dgk_burgers_set_precision; also for N = 5, 6 in BF16x3 and BF16x6). */
int dgk_burgers_set_variant(dgk_dgmesh m, int variant);
/* Tile order of the persistent grid: 1 gives each XCD a contiguous range of
 * tiles (blocks renumbered around the round-robin XCD dispatch), 0 (default)
 * walks the tile list with a grid stride. */
int dgk_burgers_set_order(dgk_dgmesh m, int xcd_chunked);
/* Order of the 16-element tiles that every persistent launch on the mesh
 * walks (stage, split stage, viscosity, sensors): 0 (default) partition
 * by partition, 1 the full tiles of all partitions along the Morton curve
 * of their centroids (partial tiles last). A scheduling choice only:
 * results are identical. No counterpart in the reference, whose kernels
 * loop over partitions (runner/builder/builder.go:172-277). Call outside
 * graph capture (it reallocates the tile list). */
int dgk_dgmesh_set_tile_order(dgk_dgmesh m, int order);
/* Precision of the Dr/Ds/Dt and LIFT contractions in the stage (DGK_PREC_*,
 * default DGK_PREC_F64); OR-ing DGK_PREC_LIFT_F64 keeps the LIFT product in
 * FP64 while Dr/Ds/Dt run in the given precision (DGK_PREC_I8X6 always
 * keeps LIFT in FP64). The state, geometry, flux,
 * RK combination and dt stay FP64. Below FP64, and for N = 5, 6 at any
 * precision, the stage runs split: a volume kernel (-div f(u) through a
 * derivative plan, with the linear-part removal of dgk_phys_deriv_affine)
 * into a mesh-owned FP64 buffer, then a surface kernel (LF flux, LIFT on
 * matrix cores, RK update). The fused kernels (N <= 4, FP64) are used
 * otherwise. At N = 5, 6 with D and LIFT both in DGK_PREC_BF16X3 (or both in
 * DGK_PREC_BF16X6) one fused launch computes the stage (w = u^2/2 minus its
 * vertex interpolant once per element column, Dr/Ds/Dt and LIFT products on
 * BF16 matrix cores scaled by the element's metrics in FP64, LF flux and RK
 * update), unless variant 3 asks for the split stage. */
enum { DGK_PREC_LIFT_F64 = 16 };
int dgk_burgers_set_precision(dgk_dgmesh m, int prec);

/* ---------------------------------------------------------------------- */
/* Shock capturing (no counterpart in the reference, which has no shock    */
/* capturing; the BASELINE learned-viscosity workload defines it):         */
/*   u_t + div(u^2/2 (1,1,1)) = div(eps grad u),  eps constant per element. */
/* Sensor (Persson & Peraire, AIAA 2006-112), on uh = V^-1 u:              */
/*   S = sum_{deg m = N} uh_m^2 / sum uh_m^2, s = log10 S,                  */
/*   eps = eps0 ramp(s), ramp = 0 below s0 - kappa, 1 above s0 + kappa,     */
/*   (1 + sin(pi (s - s0) / (2 kappa)))/2 between;                          */
/*   eps0 = ceps h sqrt(3) max|u| / N, h = (6 vol)^(1/3).                   */
/* Viscous operator, BR1 strong form with a jump penalty:                   */
/*   q = grad u + LIFT(Fscale n (u+ - u-)/2),                               */
/*   L_v = div(eps q) - LIFT(Fscale (n.(eps q)- - n.(eps q)+)/2             */
/*                           + Fscale tau (u- - u+)), tau = tau0 max(eps-,eps+). */
/* The viscous stage uses the split stage path (volume kernel, viscous      */
/* kernels, surface kernel). Built for N = 1..4. eps lives in the context   */
/* (dense global element order, FP64).                                      */
/* ---------------------------------------------------------------------- */
typedef struct dgk_av_s* dgk_av;
typedef struct dgk_mlp_s* dgk_mlp;
typedef struct {
  double s0;     /* sensor threshold (log10), e.g. -4 log10(N) */
  double kappa;  /* ramp half width (log10) */
  double ceps;   /* eps0 = ceps h lambda / N */
  double tau0;   /* penalty factor */
} dgk_av_params;

/* Vinv_host: the Np x Np inverse Vandermonde matrix (row-major) of the
 * orthonormal basis in Vandermonde3D's mode order. Allocates (call outside
 * graph capture). */
int dgk_av_create(dgk_dgmesh m, const double* Vinv_host, const dgk_av_params* p, dgk_av* out);
int dgk_av_destroy(dgk_av a);
/* Device pointer of eps [Ktot] (FP64). */
int dgk_av_eps_dev(dgk_av a, double** eps_dev);
/* Persson-Peraire sensor of u (runner layout) into eps, FP64. maxeps_dev:
 * atomic max of eps (must hold 0 or a previous max), or NULL. Optional
 * outputs (NULL to skip): frac = eps/eps0 (FP64 or FP32, [Ktot]) and the
 * normalised modal coefficients uh/|uh| (FP32, [Ktot][Np]). */
int dgk_av_sensor_pp(dgk_av a, const double* u_dev, double* maxeps_dev, double* frac_dev,
                     float* frac32_dev, float* feats_dev);
/* Learned sensor: eps = eps0 * net(uh/|uh|), modal transform on BF16 matrix
 * cores in split form (hi*hi + hi*lo + lo*hi, ~16 bits), network on BF16
 * matrix cores with FP32 accumulation. The network's
 * input width must be Np. frac_dev (FP64, may be NULL) receives net(.). */
int dgk_av_sensor_mlp(dgk_av a, dgk_mlp net, const double* u_dev, double* maxeps_dev,
                      double* frac_dev);
/* L_v(u; eps) alone into out_dev (runner layout). */
int dgk_av_rhs(dgk_av a, const double* u_dev, double* out_dev);
/* L(u) + L_v(u; eps) into out_dev (runner layout). */
int dgk_av_rhs_total(dgk_av a, const double* u_dev, double* out_dev);
/* One SSP-RK3 stage with viscosity: out = a0 u0 + a1 (u + dt (L + L_v)),
 * max|out| folded into *maxabs_dev if not NULL. out may alias u0, not u. */
int dgk_av_stage(dgk_av a, const double* u_dev, const double* u0_dev, double* out_dev, double a0,
                 double a1, const double* dt_dev, double* maxabs_dev);
/* scal_dev = [max|u|, dt, t, max eps] (FP64, device):
 *   dt = min(cfl hmin / max|u|, cfl_v hmin^2 / max eps) (second term when
 *   max eps > 0), clipped so that t does not pass t_end (t_end > 0);
 *   t += dt; max|u| and max eps are cleared. */
int dgk_av_cfl_dt(dgk_device dev, double* scal_dev, double cfl, double cfl_v, double hmin,
                  double t_end);

/* Sensor network nin -> 64 -> 64 -> 1 (nin <= 56), tanh hidden, sigmoid output.
 * Parameters (FP32, flat): W1[64][nin], b1[64], W2[64][64], b2[64],
 * w3[64], b3. The FP32 copy is the training master; inference rounds W1,
 * W2, the inputs and the first hidden layer to BF16. */
int dgk_mlp_create(dgk_device dev, int nin, const float* params_host, dgk_mlp* out);
int dgk_mlp_destroy(dgk_mlp n);
int dgk_mlp_num_params(dgk_mlp n, int* count);
int dgk_mlp_get(dgk_mlp n, float* params_host);
/* Replaces the parameters and resets the optimiser state. */
int dgk_mlp_set(dgk_mlp n, const float* params_host);
/* y[i] = net(x[i]) for x [count][nin] (FP32, device) on BF16 matrix cores. */
int dgk_mlp_forward(dgk_mlp n, const float* x_dev, int64_t count, float* y_dev);
/* Gradient of mean_i (net(x_i) - t_i)^2 over all rows into grad_dev
 * [num_params] (FP32), the loss into *loss_dev if not NULL. Computed on BF16
 * matrix cores with FP32 accumulation: the forward pass rounds as
 * dgk_mlp_forward, the backward products take BF16 operands (dz2 and W2;
 * dz2, h1, dz1 and x), the bias sums are FP32. Per-wave partial sums are
 * added in a fixed order, so repeated calls give identical results. */
int dgk_mlp_grad(dgk_mlp n, const float* x_dev, const float* t_dev, int64_t count,
                 float* grad_dev, float* loss_dev);
/* One Adam step (beta 0.9, 0.999) with learning rate lr on a batch drawn
 * with replacement from the ndata rows by a hash of (seed, step count), the
 * gradient as dgk_mlp_grad's; the batch loss is added to *loss_acc_dev if
 * not NULL. Two launches, no allocation. */
int dgk_mlp_train_step(dgk_mlp n, const float* x_dev, const float* t_dev, int64_t ndata,
                       int batch, uint64_t seed, float lr, float* loss_acc_dev);

/* ---------------------------------------------------------------------- */
/* Element-graph partitioning (host only: no device is needed or touched). */
/* Replaces metis.PartGraphKwayWeighted as MeshPartitioner.Partition calls */
/* it (mesh/partitioner/mesh_partitioner.go:87-142; graph from             */
/* buildMetisGraph, :145-184).                                              */
/* ---------------------------------------------------------------------- */
/* Multilevel k-way partition of the CSR graph (xadj [nv+1], adjncy        */
/* [xadj[nv]], symmetric); vwgt [nv] and adjwgt [xadj[nv]] may be NULL      */
/* (weights 1). part[v] in [0, nparts); every part weighs at most           */
/* max(ubfactor * total / nparts, ceil(total / nparts)) unless the vertex   */
/* weights make that impossible. objective 0 minimises the edge cut (sum of */
/* adjwgt over edges between parts, each once), 1 the communication volume */
/* (sum over vertices of the number of other parts among the neighbours);  */
/* *objval (may be NULL) receives that value. Deterministic for a seed.     */
int dgk_partition_kway(int64_t nv, const int64_t* xadj, const int64_t* adjncy,
                       const int64_t* vwgt, const int64_t* adjwgt, int nparts, double ubfactor,
                       int objective, uint64_t seed, int32_t* part, int64_t* objval);

#ifdef __cplusplus
}
#endif

#endif /* DGK_H */
