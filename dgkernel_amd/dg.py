"""Host driver for the device DG path: setup of a partitioned tet mesh, the
fused Burgers SSP-RK3 stepper, and the physical-derivative / FaceConnector
entry points, all through libdgk's C ABI.

Partitioned state arrays use the runner's pooled layout
(runner/builder/builder.go:113-159): partitions are contiguous ranges of
global elements, each starting at an aligned value offset.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .mesh import TetMesh, connect, face_codes, face_node_tables, vmap_pm
from .nudg import RefTet, build_ref_tet

WARP_ALIGN = 128  # builder.WarpAlign (runner/builder/builder.go:25-30)


def aligned_offsets(K: np.ndarray, values_per_elem: int, value_size: int = 8,
                    alignment: int = WARP_ALIGN):
    """CalculateAlignedOffsetsAndSize; runner/builder/builder.go:113-159.

    Returns (offsets[npart+1] in values, total bytes).
    """
    offs = np.zeros(len(K) + 1, dtype=np.int64)
    cur = 0
    a = max(1, int(alignment))
    for i, k in enumerate(K):
        if cur % a:
            cur = (cur + a - 1) // a * a
        offs[i] = cur // value_size
        cur += int(k) * values_per_elem * value_size
    if cur % a:
        cur = (cur + a - 1) // a * a
    offs[-1] = cur // value_size
    return offs, int(offs[-1]) * value_size


@dataclass
class DGSetup:
    """Host-side setup of a mesh at order N with contiguous partitions."""

    ref: RefTet
    mesh: TetMesh
    Kp: np.ndarray  # elements per partition
    EToE: np.ndarray
    EToF: np.ndarray
    codes: np.ndarray
    state_off: np.ndarray  # value offsets of Np-per-element arrays
    state_bytes: int
    pooled_index: np.ndarray  # dense node index -> pooled value index

    @property
    def K(self) -> int:
        return self.mesh.K

    @property
    def npart(self) -> int:
        return int(self.Kp.size)

    def vmaps(self):
        return vmap_pm(self.ref, self.EToE, self.codes)

    def to_pooled(self, dense: np.ndarray) -> np.ndarray:
        out = np.zeros(self.state_bytes // 8, dtype=np.float64)
        out[self.pooled_index] = dense.reshape(-1)
        return out

    def from_pooled(self, pooled: np.ndarray) -> np.ndarray:
        return pooled[self.pooled_index]

    def element_vertices(self) -> np.ndarray:
        m = self.mesh
        v = np.stack([m.VX[m.EToV], m.VY[m.EToV], m.VZ[m.EToV]], axis=2)  # [K,4,3]
        return np.ascontiguousarray(v.reshape(-1), dtype=np.float64)

    def node_coords(self):
        """Physical node coordinates [K, Np] (affine map, startup3d.go:126-148)."""
        r, s, t = self.ref.r, self.ref.s, self.ref.t
        m = self.mesh
        out = []
        for V in (m.VX, m.VY, m.VZ):
            va, vb, vc, vd = (V[m.EToV[:, q]][:, None] for q in range(4))
            out.append(0.5 * (-(1.0 + r + s + t) * va + (1.0 + r) * vb + (1.0 + s) * vc
                              + (1.0 + t) * vd))
        return out


def make_setup(mesh: TetMesh, N: int, Kp=None, ref: RefTet | None = None,
               alignment: int = WARP_ALIGN) -> DGSetup:
    ref = ref or build_ref_tet(N)
    K = mesh.K
    Kp = np.array([K] if Kp is None else Kp, dtype=np.int64)
    if int(Kp.sum()) != K:
        raise ValueError("partition sizes do not sum to K")
    EToE, EToF = connect(mesh)
    codes = face_codes(mesh, EToE, EToF)
    offs, nbytes = aligned_offsets(Kp, ref.Np, 8, alignment)
    starts = np.concatenate(([0], np.cumsum(Kp)))
    pidx = np.empty(K * ref.Np, dtype=np.int64)
    for p in range(Kp.size):
        g0, g1 = starts[p], starts[p + 1]
        base = offs[p]
        pidx[g0 * ref.Np:g1 * ref.Np] = base + np.arange((g1 - g0) * ref.Np)
    return DGSetup(ref, mesh, Kp, EToE, EToF, codes, offs, nbytes, pidx)


class DGMesh:
    """Device mesh (dgk_dgmesh): operators, neighbours, codes, FP64 geometry."""

    def __init__(self, dev: _lib.Device, st: DGSetup):
        self.dev, self.st = dev, st
        ref = st.ref
        self._keep = dict(
            K=np.ascontiguousarray(st.Kp, dtype=np.int64),
            off=np.ascontiguousarray(st.state_off, dtype=np.int64),
            Dr=np.ascontiguousarray(ref.Dr), Ds=np.ascontiguousarray(ref.Ds),
            Dt=np.ascontiguousarray(ref.Dt), LIFT=np.ascontiguousarray(ref.LIFT),
            Fmask=np.ascontiguousarray(ref.Fmask, dtype=np.int32),
            r=np.ascontiguousarray(ref.r), s=np.ascontiguousarray(ref.s),
            t=np.ascontiguousarray(ref.t),
            fn=np.ascontiguousarray(face_node_tables(ref), dtype=np.int32),
            EToE=np.ascontiguousarray(st.EToE, dtype=np.int64),
            codes=np.ascontiguousarray(st.codes, dtype=np.int32),
            vtx=st.element_vertices())
        k = self._keep
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
        d = _lib.DGMeshDesc(ref.N, st.npart, P(k["K"], C.c_int64), P(k["off"], C.c_int64),
                            P(k["Dr"], C.c_double), P(k["Ds"], C.c_double), P(k["Dt"], C.c_double),
                            P(k["LIFT"], C.c_double), P(k["Fmask"], C.c_int), P(k["r"], C.c_double),
                            P(k["s"], C.c_double), P(k["t"], C.c_double), P(k["fn"], C.c_int),
                            P(k["EToE"], C.c_int64), P(k["codes"], C.c_int), P(k["vtx"], C.c_double))
        h = C.c_void_p()
        _lib.call("dgk_dgmesh_create", dev.h, C.byref(d), C.byref(h))
        self.h = h

    def geometry(self) -> np.ndarray:
        """[26, K]: rx ry rz sx sy sz tx ty tz J, then per face nx ny nz Fscale."""
        out = np.empty((26, self.st.K))
        _lib.call("dgk_dgmesh_geometry", self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def geometry_dev_ptr(self):
        p = C.c_void_p()
        _lib.call("dgk_dgmesh_geometry_dev", self.h, C.byref(p))
        return p

    def min_node_distance(self) -> float:
        v = C.c_double()
        _lib.call("dgk_dgmesh_min_node_distance", self.h, C.byref(v))
        return v.value

    def set_tile(self, elems_per_block: int, blocks_per_cu: int = 0) -> None:
        _lib.call("dgk_burgers_set_tile", self.h, elems_per_block, blocks_per_cu)

    def set_precision(self, prec: int) -> None:
        """DGK_PREC_* of the D and LIFT contractions of the stage (state stays FP64);
        | DGK_PREC_LIFT_F64 keeps the LIFT product in FP64."""
        _lib.call("dgk_burgers_set_precision", self.h, prec)

    def set_variant(self, variant: int) -> None:
        """0 automatic, 1 VALU stage kernel, 2 FP64 MFMA stage kernel, 3 split stage."""
        _lib.call("dgk_burgers_set_variant", self.h, variant)

    def set_order(self, xcd_chunked: bool) -> None:
        _lib.call("dgk_burgers_set_order", self.h, 1 if xcd_chunked else 0)

    def set_tile_order(self, spatial: bool) -> None:
        """Walk the 16-element tiles partition by partition (False) or, across
        partitions, along the Morton curve of their centroids (True)."""
        _lib.call("dgk_dgmesh_set_tile_order", self.h, 1 if spatial else 0)

    def trace_plus(self, u: _lib.DeviceBuffer, out: _lib.DeviceBuffer) -> None:
        _lib.call("dgk_dgmesh_trace_plus", self.h, u.ptr, out.ptr)

    def __del__(self):
        try:
            if _lib.alive(self.dev):
                _lib.load().dgk_dgmesh_destroy(self.h)
        except Exception:
            pass


class BurgersSolver:
    """SSP-RK3 time stepping of u_t + div(u^2/2 (1,1,1)) = 0 on the device.

    Every step: dt = cfl*hmin/max|u| (FP64 on the device), then three fused
    stage launches. ``step_graph`` captures one step in a HIP graph.
    """

    def __init__(self, dev: _lib.Device, st: DGSetup, cfl: float = 0.3,
                 hmin: float | None = None, tuned: bool = True):
        self.dev, self.st, self.cfl = dev, st, cfl
        self.mesh = DGMesh(dev, st)
        # launch shape from the autotuner's table (dgkernel_amd/tuned_stage.json), if any
        self.tuned = None
        if tuned:
            from . import autotune
            self.tuned = autotune.apply(self.mesh)
        n = st.state_bytes // 8
        self.u = dev.zeros(n)
        self.u1 = dev.zeros(n)
        self.u2 = dev.zeros(n)
        self.scal = dev.zeros(4)  # [maxabs, dt, spare, spare]
        self.hmin = self.mesh.min_node_distance() if hmin is None else hmin
        self._graph = None

    @property
    def maxabs_ptr(self):
        return self.scal.ptr

    @property
    def dt_ptr(self):
        return C.c_void_p(self.scal.ptr.value + 8)

    def set_state(self, u_dense: np.ndarray) -> None:
        self.u.copy_from(self.st.to_pooled(u_dense))
        self.scal.fill0()
        _lib.call("dgk_max_abs", self.mesh.h, self.u.ptr, self.maxabs_ptr)

    def state(self) -> np.ndarray:
        return self.st.from_pooled(self.u.to_host()).reshape(self.st.K, self.st.ref.Np)

    def dt(self) -> float:
        return float(self.scal.to_host()[1])

    def _enqueue_step(self) -> None:
        m = self.mesh.h
        _lib.call("dgk_cfl_dt", self.dev.h, self.maxabs_ptr, self.cfl, self.hmin, self.dt_ptr)
        _lib.call("dgk_burgers_stage", m, self.u.ptr, None, self.u1.ptr, 0.0, 1.0, self.dt_ptr, None)
        _lib.call("dgk_burgers_stage", m, self.u1.ptr, self.u.ptr, self.u2.ptr, 0.75, 0.25,
                  self.dt_ptr, None)
        _lib.call("dgk_burgers_stage", m, self.u2.ptr, self.u.ptr, self.u.ptr, 1.0 / 3.0, 2.0 / 3.0,
                  self.dt_ptr, self.maxabs_ptr)

    def step(self, n: int = 1, graph: bool = False) -> None:
        if graph:
            if self._graph is None:
                self._graph = _lib.Graph(self.dev, self._enqueue_step)
            for _ in range(n):
                self._graph.launch()
        else:
            for _ in range(n):
                self._enqueue_step()

    def rhs(self, u_dense: np.ndarray) -> np.ndarray:
        """L(u) for a dense [K,Np] field (testing aid)."""
        self.u1.copy_from(self.st.to_pooled(u_dense))
        _lib.call("dgk_burgers_rhs", self.mesh.h, self.u1.ptr, self.u2.ptr)
        return self.st.from_pooled(self.u2.to_host()).reshape(self.st.K, self.st.ref.Np)


class DerivPlan:
    """Affine physical derivative on matrix cores (dgk_deriv_plan).

    ``prec`` is one of ``_lib.DGK_PREC_F64``, ``DGK_PREC_F32``,
    ``DGK_PREC_BF16X3``, ``DGK_PREC_BF16X6``, ``DGK_PREC_I8X6``; outputs are
    FP64 for F64 and I8X6 and FP32 otherwise. ``D`` is (Dr, Ds, Dt), Np x Np row-major; ``rst`` the
    reference node coordinates (required below FP64).
    """

    def __init__(self, dev: _lib.Device, N: int, D, prec: int = _lib.DGK_PREC_F64, rst=None):
        self.dev, self.N, self.prec = dev, N, prec
        D = [np.ascontiguousarray(m, dtype=np.float64) for m in D]
        R = [None] * 3 if rst is None else [np.ascontiguousarray(v, dtype=np.float64) for v in rst]
        h = C.c_void_p()
        P = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        _lib.call("dgk_deriv_plan_create", dev.h, N, *(P(m) for m in D), *(P(v) for v in R), prec,
                  C.byref(h))
        self.h = h
        self.out_dtype = (np.float64 if prec in (_lib.DGK_PREC_F64, _lib.DGK_PREC_I8X6)
                          else np.float32)

    @classmethod
    def from_ref(cls, dev: _lib.Device, ref: RefTet, prec: int = _lib.DGK_PREC_F64):
        return cls(dev, ref.N, (ref.Dr, ref.Ds, ref.Dt), prec, (ref.r, ref.s, ref.t))

    def apply(self, K: int, u, geo_ptr, geo_stride: int, outs) -> None:
        """Enqueue d/dx, d/dy, d/dz of dense [K, Np] ``u`` (device buffer) into ``outs``."""
        _lib.call("dgk_phys_deriv_affine", self.h, K, u.ptr, geo_ptr, geo_stride,
                  outs[0].ptr, outs[1].ptr, outs[2].ptr)

    def free(self) -> None:
        if self.h:
            _lib.call("dgk_deriv_plan_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib.alive(self.dev):
                _lib.load().dgk_deriv_plan_destroy(self.h)
        except Exception:
            pass
