"""ctypes binding of the CPU oracle (oracle/liboracle.so, built from
oracle/dgk_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker (or the timed CPU
baseline). The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_long)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: build with `python -m dgkernel_amd.build --only oracle`")
        lib = C.CDLL(str(LIB_PATH))
        lib.orc_setup.restype = C.c_void_p
        lib.orc_setup.argtypes = [C.c_int, C.c_int, _D, _D, _D, C.c_long, _I, _I, _D, _I]
        lib.orc_free.argtypes = [C.c_void_p]
        lib.orc_get_field.argtypes = [C.c_void_p, C.c_int, _D]
        lib.orc_get_imap.argtypes = [C.c_void_p, C.c_int, _I]
        lib.orc_burgers_rhs.argtypes = [C.c_void_p, _D, _D]
        lib.orc_min_node_distance.argtypes = [C.c_void_p]
        lib.orc_min_node_distance.restype = C.c_double
        lib.orc_ssprk3.argtypes = [C.c_void_p, _D, C.c_int, C.c_double, C.c_double, C.c_double, _D]
        lib.orc_burgers_rhs_partitions.argtypes = [C.c_void_p, _D, _D, _L, C.c_int]
        lib.orc_ssprk3_partitions.argtypes = [C.c_void_p, _D, C.c_int, C.c_double, C.c_double,
                                              _L, C.c_int, _D]
        lib.orc_jacobi_gl.argtypes = [C.c_double, C.c_double, C.c_int, _D]
        lib.orc_jacobi_gq.argtypes = [C.c_double, C.c_double, C.c_int, _D, _D]
        lib.orc_jacobi_p.argtypes = [_D, C.c_int, C.c_double, C.c_double, C.c_int, _D]
        lib.orc_grad_jacobi_p.argtypes = [_D, C.c_int, C.c_double, C.c_double, C.c_int, _D]
        lib.orc_simplex3dp.argtypes = [_D, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int, _D, _D, _D, _D]
        lib.orc_nodes3d.argtypes = [C.c_int, _D, _D, _D]
        lib.orc_element.argtypes = [C.c_int] + [_D] * 10 + [_I]
        lib.orc_phys_deriv.argtypes = [C.c_int, C.c_long, _D, _D, _D, _D, C.POINTER(_D), _D, _D, _D]
        lib.orc_matmul.argtypes = [C.c_int, C.c_int, _D, C.c_long, _D, _D, C.c_int]
        lib.orc_aligned_offsets.argtypes = [C.c_int, _I, C.c_long, C.c_int, C.c_int, _L]
        lib.orc_aligned_offsets.restype = C.c_long
        lib.orc_face_connector_count.argtypes = [C.c_long, _I]
        lib.orc_face_connector.argtypes = [C.c_long, C.c_int, C.c_int, C.c_int, _I, _I, _L, _I, _I]
        lib.orc_pp_sensor.argtypes = [C.c_void_p, _D, C.c_double, C.c_double, C.c_double, _D, _D,
                                      _D, C.POINTER(C.c_float)]
        lib.orc_av_rhs.argtypes = [C.c_void_p, _D, _D, C.c_double, _D]
        lib.orc_ssprk3_av.argtypes = [C.c_void_p, _D, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _D, _D]
        lib.orc_ssprk3_av.restype = C.c_int
        lib.orc_num_threads.restype = C.c_int
        lib.orc_set_num_threads.argtypes = [C.c_int]
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def set_threads(n: int) -> None:
    load().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(load().orc_num_threads())


def jacobi_gl(alpha: float, beta: float, N: int) -> np.ndarray:
    x = np.zeros(N + 1)
    load().orc_jacobi_gl(alpha, beta, N, _d(x))
    return x


def jacobi_gq(alpha: float, beta: float, N: int):
    x, w = np.zeros(N + 1), np.zeros(N + 1)
    load().orc_jacobi_gq(alpha, beta, N, _d(x), _d(w))
    return x, w


def jacobi_p(x, alpha, beta, N, grad=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    fn = load().orc_grad_jacobi_p if grad else load().orc_jacobi_p
    fn(_d(x), x.size, alpha, beta, N, _d(out))
    return out


def simplex3dp(r, s, t, i, j, k):
    r, s, t = (np.ascontiguousarray(v, dtype=np.float64) for v in (r, s, t))
    P, dr, ds, dt = (np.zeros_like(r) for _ in range(4))
    load().orc_simplex3dp(_d(r), _d(s), _d(t), r.size, i, j, k, _d(P), _d(dr), _d(ds), _d(dt))
    return P, dr, ds, dt


def nodes3d(N: int):
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    X, Y, Z = np.zeros(Np), np.zeros(Np), np.zeros(Np)
    load().orc_nodes3d(N, _d(X), _d(Y), _d(Z))
    return X, Y, Z


def element(N: int) -> dict:
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    Nfp = (N + 1) * (N + 2) // 2
    out = {k: np.zeros(Np) for k in ("r", "s", "t")}
    for k in ("V", "Vinv", "Mass", "Dr", "Ds", "Dt"):
        out[k] = np.zeros((Np, Np))
    out["LIFT"] = np.zeros((Np, 4 * Nfp))
    fm = np.zeros(4 * Nfp, dtype=np.int32)
    rc = load().orc_element(N, *[_d(out[k]) for k in ("r", "s", "t", "V", "Vinv", "Mass", "Dr",
                                                      "Ds", "Dt", "LIFT")], _i(fm))
    if rc:
        raise RuntimeError(f"oracle element construction failed at N={N}")
    out["Fmask"] = fm.reshape(4, Nfp)
    out["Np"], out["Nfp"], out["N"] = Np, Nfp, N
    return out


_FIELDS = {"X": 0, "Y": 1, "Z": 2, "J": 3, "rx": 4, "ry": 5, "rz": 6, "sx": 7, "sy": 8, "sz": 9,
           "tx": 10, "ty": 11, "tz": 12, "nx": 20, "ny": 21, "nz": 22, "sJ": 23, "Fscale": 24}
_IMAPS = {"EToE": 0, "EToF": 1, "vmapM": 2, "vmapP": 3, "mapP": 4}


class Setup:
    """StartUp3D on a mesh (the reference's NUDGTet for that mesh)."""

    def __init__(self, N: int, VX, VY, VZ, EToV, EToVc=None, period=None):
        self.N = N
        self.Np = (N + 1) * (N + 2) * (N + 3) // 6
        self.Nfp = (N + 1) * (N + 2) // 2
        self._VX, self._VY, self._VZ = (np.ascontiguousarray(v, dtype=np.float64) for v in (VX, VY, VZ))
        self._EToV = np.ascontiguousarray(EToV, dtype=np.int32)
        self.K = self._EToV.shape[0]
        self._EToVc = None if EToVc is None else np.ascontiguousarray(EToVc, dtype=np.int32)
        self._per = None if period is None else np.ascontiguousarray(period, dtype=np.float64)
        st = C.c_int(0)
        self.h = load().orc_setup(N, self._VX.size, _d(self._VX), _d(self._VY), _d(self._VZ), self.K,
                                  _i(self._EToV), _i(self._EToVc) if self._EToVc is not None else None,
                                  _d(self._per) if self._per is not None else None, C.byref(st))
        self.status = st.value
        if not self.h:
            raise RuntimeError("oracle setup failed")

    @classmethod
    def from_mesh(cls, N: int, mesh):
        per = mesh.period if any(p > 0 for p in mesh.period) else None
        return cls(N, mesh.VX, mesh.VY, mesh.VZ, mesh.EToV, mesh.EToVc, per)

    def field(self, name: str) -> np.ndarray:
        n = self.Np if _FIELDS[name] < 20 else 4 * self.Nfp
        out = np.zeros(self.K * n)
        load().orc_get_field(self.h, _FIELDS[name], _d(out))
        return out.reshape(self.K, n)

    def imap(self, name: str) -> np.ndarray:
        n = 4 if _IMAPS[name] < 2 else 4 * self.Nfp
        out = np.zeros(self.K * n, dtype=np.int32)
        load().orc_get_imap(self.h, _IMAPS[name], _i(out))
        return out.reshape(self.K, n)

    def burgers_rhs(self, u: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        load().orc_burgers_rhs(self.h, _d(u), _d(out))
        return out

    def min_node_distance(self) -> float:
        return float(load().orc_min_node_distance(self.h))

    def pp_sensor(self, u: np.ndarray, s0: float, kappa: float, ceps: float):
        """Persson-Peraire sensor: (eps, frac, eps0) per element and the
        normalised modal coefficients [K, Np] (float32)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        eps, frac, e0 = np.zeros(self.K), np.zeros(self.K), np.zeros(self.K)
        feats = np.zeros((self.K, self.Np), dtype=np.float32)
        load().orc_pp_sensor(self.h, _d(u), s0, kappa, ceps, _d(eps), _d(frac), _d(e0),
                             feats.ctypes.data_as(C.POINTER(C.c_float)))
        return eps, frac, e0, feats

    def av_rhs(self, u: np.ndarray, eps: np.ndarray, tau0: float) -> np.ndarray:
        """Viscous operator L_v(u; eps) (BR1 + jump penalty)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        out = np.zeros_like(u)
        load().orc_av_rhs(self.h, _d(u), _d(eps), tau0, _d(out))
        return out

    def ssprk3_av(self, u: np.ndarray, t_end: float, cfl: float, cfl_v: float, hmin: float,
                  s0: float, kappa: float, ceps: float, tau0: float, mode: int = 1,
                  max_steps: int = 100000):
        """SSP-RK3 with Persson-Peraire viscosity to t_end; (u, dts, t)."""
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        dts = np.zeros(max_steps)
        t = C.c_double()
        n = load().orc_ssprk3_av(self.h, _d(u), max_steps, t_end, cfl, cfl_v, hmin, mode, s0,
                                 kappa, ceps, tau0, _d(dts), C.byref(t))
        return u, dts[:n], t.value

    def ssprk3(self, u: np.ndarray, nsteps: int, cfl: float = 0.3, hmin: float | None = None,
               fixed_dt: float = 0.0):
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        dts = np.zeros(max(nsteps, 1))
        if hmin is None:
            hmin = self.min_node_distance()
        load().orc_ssprk3(self.h, _d(u), nsteps, cfl, hmin, fixed_dt, _d(dts))
        return u, dts[:nsteps]

    @staticmethod
    def _poff(Kp) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(np.asarray(Kp, dtype=np.int64))]).astype(np.int64)

    def burgers_rhs_partitions(self, u: np.ndarray, Kp) -> np.ndarray:
        """burgers_rhs with the reference's OCCA-OpenMP schedule (parallel over
        the contiguous partitions of sizes Kp, serial element loop inside)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        poff = self._poff(Kp)
        load().orc_burgers_rhs_partitions(self.h, _d(u), _d(out),
                                          poff.ctypes.data_as(C.POINTER(C.c_long)), len(Kp))
        return out

    def ssprk3_partitions(self, u: np.ndarray, nsteps: int, cfl: float, hmin: float, Kp):
        """ssprk3 with the partition-parallel schedule."""
        u = np.array(u, dtype=np.float64, copy=True, order="C")
        dts = np.zeros(max(nsteps, 1))
        poff = self._poff(Kp)
        load().orc_ssprk3_partitions(self.h, _d(u), nsteps, cfl, hmin,
                                     poff.ctypes.data_as(C.POINTER(C.c_long)), len(Kp), _d(dts))
        return u, dts[:nsteps]

    def __del__(self):
        try:
            if self.h:
                load().orc_free(self.h)
        except Exception:
            pass


def phys_deriv(Dr, Ds, Dt, U, metrics):
    """calcPhysicalDerivative; U and metrics element-major [K, Np]."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    K, Np = U.shape
    ms = [np.ascontiguousarray(m, dtype=np.float64) for m in metrics]
    arr = (_D * 9)(*[_d(m) for m in ms])
    Dr, Ds, Dt = (np.ascontiguousarray(m, dtype=np.float64) for m in (Dr, Ds, Dt))
    out = [np.zeros_like(U) for _ in range(3)]
    load().orc_phys_deriv(Np, K, _d(Dr), _d(Ds), _d(Dt), _d(U), arr, *[_d(o) for o in out])
    return out


def matmul(A, IN, add_to=None):
    """MATMUL_<X> (add_to None) or MATMUL_ADD_<X>; IN [K, cols] -> [K, rows]."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    IN = np.ascontiguousarray(IN, dtype=np.float64)
    rows, cols = A.shape
    K = IN.shape[0]
    out = np.zeros((K, rows)) if add_to is None else np.array(add_to, dtype=np.float64, copy=True)
    load().orc_matmul(rows, cols, _d(A), K, _d(IN), _d(out), 0 if add_to is None else 1)
    return out


def aligned_offsets(K, size_bytes: int, data_type: int, alignment: int):
    K = np.ascontiguousarray(K, dtype=np.int32)
    offs = np.zeros(K.size + 1, dtype=np.int64)
    total = load().orc_aligned_offsets(K.size, _i(K), size_bytes, data_type, alignment,
                                       offs.ctypes.data_as(_L))
    return offs, int(total)


def face_connector(K: int, Nfaces: int, Nfp: int, Np: int, VmapP, EToP):
    """Returns (num_partitions, counts[src,dst], pick dict, place dict)."""
    VmapP = np.ascontiguousarray(VmapP, dtype=np.int32)
    EToP = np.ascontiguousarray(EToP, dtype=np.int32)
    NP = load().orc_face_connector_count(K, _i(EToP))
    total = K * Nfaces * Nfp
    counts = np.zeros(NP * NP, dtype=np.int64)
    pick = np.zeros(total, dtype=np.int32)
    place = np.zeros(total, dtype=np.int32)
    rc = load().orc_face_connector(K, Nfaces, Nfp, Np, _i(VmapP), _i(EToP),
                                   counts.ctypes.data_as(_L), _i(pick), _i(place))
    if rc:
        raise ValueError("invalid dimensions")
    counts = counts.reshape(NP, NP)
    pk, pl = {}, {}
    o = 0
    for s in range(NP):
        for d in range(NP):
            c = int(counts[s, d])
            pk[(s, d)] = pick[o:o + c].astype(np.int64)
            o += c
    o = 0
    for d in range(NP):
        for s in range(NP):
            c = int(counts[s, d])
            pl[(d, s)] = place[o:o + c].astype(np.int64)
            o += c
    return NP, counts, pk, pl
